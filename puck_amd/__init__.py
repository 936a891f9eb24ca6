"""puck-mi355x: CKY parsing on MI355X behind the C ABI in include/puck.h.

Python mirror of the reference's parse-path API (ctypes over libpuck.so):
  Grammar.load_text   ~ SimpleGrammar.parseBerkeleyText + CLParserData.make (CLParser.scala:1026,1264)
  CLParser.parse      ~ CLParser.parse (CLParser.scala:50-53)
  CLParser.partitions ~ CLParser.partitions (CLParser.scala:55-58)
  CLParser.inside_outside ~ CLParser.insideOutside (CLParser.scala:60-63)

There is no fallback: if libpuck.so is missing or cannot be loaded, every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpuck.so")
DATA_DIR = os.path.join(HERE, "data", "textGrammars")

VITERBI = 0
REAL = 1
LOGSUM = 2

_lib = None


class PuckError(RuntimeError):
    pass


class TreeNode(ctypes.Structure):
    _fields_ = [("top_sym", ctypes.c_int32), ("bot_sym", ctypes.c_int32),
                ("width", ctypes.c_int32), ("score", ctypes.c_float)]


class MbrNode(ctypes.Structure):
    _fields_ = [("top_coarse", ctypes.c_int32), ("bot_coarse", ctypes.c_int32),
                ("begin", ctypes.c_int32), ("end", ctypes.c_int32), ("score", ctypes.c_float)]


class GrammarInfo(ctypes.Structure):
    _fields_ = [("num_nonterms", ctypes.c_int32), ("num_terms", ctypes.c_int32),
                ("num_coarse", ctypes.c_int32), ("cell_size", ctypes.c_int32),
                ("mask_words", ctypes.c_int32), ("root", ctypes.c_int32),
                ("num_binary", ctypes.c_int32 * 4), ("num_unary_n", ctypes.c_int32),
                ("num_unary_t", ctypes.c_int32), ("num_words", ctypes.c_int32),
                ("semiring", ctypes.c_int32), ("num_blocks", ctypes.c_int32 * 4),
                ("rules_dropped", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("rule_kernel_ms", ctypes.c_double), ("reduce_ms", ctypes.c_double),
                ("other_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("rule_apps", ctypes.c_double), ("rule_launches", ctypes.c_int64),
                ("device_ms", ctypes.c_double)]


class LaunchConfig(ctypes.Structure):
    _fields_ = [("persistent", ctypes.c_int32), ("piece_target", ctypes.c_int32),
                ("queue_wgs_per_cu", ctypes.c_int32), ("reduce_targets", ctypes.c_int32)]


FLAG_ONE_ITEM_PER_LANE = 1
FLAG_PERSISTENT = 2
FLAG_GRAPHS = 4
FLAG_NO_PLAN_AHEAD = 8  # accepted and ignored: every parser works on one HIP stream
FLAG_OUTSIDE_FINAL = 16
FLAG_DEVICE_SCHEDULE = 32
FLAG_SLOT_PARTIALS = 64


class BatchConfig(ctypes.Structure):
    _fields_ = [("streams", ctypes.c_int32), ("bucket_cells", ctypes.c_int64), ("persistent", ctypes.c_int32),
                ("graphs", ctypes.c_int32), ("piece_target", ctypes.c_int32), ("items_per_lane", ctypes.c_int32),
                ("outside_final", ctypes.c_int32), ("device_schedule", ctypes.c_int32),
                ("reduce_targets", ctypes.c_int32), ("slot_partials", ctypes.c_int32)]


class BatchStats(ctypes.Structure):
    _fields_ = [("device_ms", ctypes.c_double), ("buckets", ctypes.c_int64), ("graph_captures", ctypes.c_int64),
                ("graph_replays", ctypes.c_int64)]

# Exported symbols and their signatures: (name, restype, argtypes).
_P = ctypes.c_void_p
_I32P = ctypes.POINTER(ctypes.c_int32)
_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)
_U32P = ctypes.POINTER(ctypes.c_uint32)
SIGNATURES = [
    ("puck_last_error", ctypes.c_char_p, []),
    ("puck_abi_version", ctypes.c_int32, []),
    ("puck_grammar_load_text", ctypes.c_int32, [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(_P)]),
    ("puck_grammar_free", None, [_P]),
    ("puck_grammar_reproject", ctypes.c_int32, [_P, _P, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    ("puck_grammar_save_grz", ctypes.c_int32, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.c_char_p]),
    ("puck_grammar_load_grz", ctypes.c_int32, [ctypes.c_char_p, ctypes.POINTER(_P), ctypes.c_int32]),
    ("puck_grammar_info", ctypes.c_int32, [_P, ctypes.POINTER(GrammarInfo)]),
    ("puck_grammar_word_id", ctypes.c_int32, [_P, ctypes.c_char_p]),
    ("puck_grammar_rules", ctypes.c_int32, [_P, ctypes.c_int32, _I32P, _F32P, _F32P, ctypes.c_int32]),
    ("puck_grammar_symbols", ctypes.c_int32, [_P, ctypes.c_int32, _I32P, _I32P, ctypes.c_int32]),
    ("puck_grammar_coarse_name", ctypes.c_int32, [_P, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32]),
    ("puck_grammar_blocks", ctypes.c_int32, [_P, ctypes.c_int32, ctypes.c_int32, _I32P]),
    ("puck_grammar_lexicon", ctypes.c_int32, [_P, ctypes.c_int32, _I32P, _F64P, ctypes.c_int32]),
    ("puck_synth_wsj6", ctypes.c_int32, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64]),
    ("puck_sample_sentences", ctypes.c_int32, [_P, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, ctypes.c_int32, _I32P, ctypes.c_int64, _I32P]),
    ("puck_parser_create", ctypes.c_int32, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_P)]),
    ("puck_parser_free", None, [_P]),
    ("puck_parser_configure", ctypes.c_int32, [_P, ctypes.POINTER(LaunchConfig)]),
    ("puck_batch_create", ctypes.c_int32, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                           ctypes.POINTER(BatchConfig), ctypes.POINTER(_P)]),
    ("puck_batch_free", None, [_P]),
    ("puck_batch_parse_viterbi", ctypes.c_int32, [_P, _I32P, _I32P, ctypes.c_int32, ctypes.POINTER(TreeNode),
                                                  ctypes.c_int64]),
    ("puck_batch_parse_mbr", ctypes.c_int32, [_P, _I32P, _I32P, ctypes.c_int32, ctypes.POINTER(MbrNode),
                                              ctypes.c_int64]),
    ("puck_batch_stats", ctypes.c_int32, [_P, ctypes.POINTER(BatchStats)]),
    ("puck_parse_viterbi", ctypes.c_int32, [_P, _I32P, _I32P, ctypes.c_int32, ctypes.POINTER(TreeNode), ctypes.c_int64]),
    ("puck_parse_mbr", ctypes.c_int32, [_P, _I32P, _I32P, ctypes.c_int32, ctypes.POINTER(MbrNode), ctypes.c_int64]),
    ("puck_parse_viterbi_noalloc", ctypes.c_int32, [_P, _I32P, _I32P, ctypes.c_int32, ctypes.POINTER(TreeNode),
                                                    ctypes.c_int64]),
    ("puck_partitions", ctypes.c_int32, [_P, _I32P, _I32P, ctypes.c_int32, _F32P]),
    ("puck_inside_outside", ctypes.c_int32, [_P, _I32P, _I32P, ctypes.c_int32]),
    ("puck_masks", ctypes.c_int32, [_P, _I32P, _I32P, ctypes.c_int32, _U32P, ctypes.c_int64]),
    ("puck_debug_charts", ctypes.c_int32, [_P, _I32P, _I32P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           _F32P, _F32P, _F32P, _F32P, _U32P, _F64P, ctypes.c_int64]),
    ("puck_parser_stats", ctypes.c_int32, [_P, ctypes.POINTER(Stats)]),
    ("puck_parser_launches", ctypes.c_int64, [_P, _F32P, _F64P, _I32P, ctypes.c_int64]),
    ("puck_server_create", ctypes.c_int32, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_P)]),
    ("puck_server_free", None, [_P]),
    ("puck_server_submit", ctypes.c_int32, [_P, _I32P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    ("puck_server_result", ctypes.c_int32, [_P, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64]),
    ("puck_server_flush", ctypes.c_int32, [_P]),
    ("puck_server_stats", ctypes.c_int32, [_P, ctypes.c_int32, ctypes.c_void_p]),
]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_int64), ("batches", ctypes.c_int64), ("mean_batch", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p90_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double)]


def lib():
    """Loads libpuck.so (raises PuckError if it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PuckError(f"{LIB_PATH} is not built; run `python -m puck_amd.build`")
        l = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def _check(status):
    if status < 0:
        raise PuckError(lib().puck_last_error().decode() or f"status {status}")
    return status


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_I32P)


def grammar_path(name):
    """Path prefix of a bundled text grammar, e.g. grammar_path('wsj_1')."""
    return os.path.join(DATA_DIR, name + ".gr")


class Grammar:
    """A compiled text grammar (CLParserData.make over one Berkeley text grammar)."""

    def __init__(self, handle, prefix, semiring):
        self._h = handle
        self.prefix = prefix
        self.semiring = semiring
        info = GrammarInfo()
        _check(lib().puck_grammar_info(self._h, ctypes.byref(info)))
        self.info = info

    @classmethod
    def load_text(cls, prefix, semiring=VITERBI):
        h = _P()
        _check(lib().puck_grammar_load_text(prefix.encode(), semiring, ctypes.byref(h)))
        return cls(h, prefix, semiring)

    @classmethod
    def load_grz(cls, path):
        """All layers of a grammar.grz container (CLParserData.readSequence), coarse first."""
        n = _check(lib().puck_grammar_load_grz(path.encode(), None, 0))
        hs = (_P * n)()
        _check(lib().puck_grammar_load_grz(path.encode(), hs, n))
        out = []
        for h in hs:
            info = GrammarInfo()
            _check(lib().puck_grammar_info(h, ctypes.byref(info)))
            out.append(cls(_P(h), f"{path}#{len(out)}", info.semiring))
        return out

    @staticmethod
    def reproject(coarse, fine, hierarchy=None, coarse_level=-1, fine_level=-1):
        """CLParser.reprojectGrammar (the --reproject option): new (coarse, fine) layers whose coarse
        symbols are the coarse grammar's refined symbols, fine symbols projected through the
        hierarchy (default: fine prefix + ".hierarchy"; levels default to the last number in each
        prefix)."""
        hc, hf = _P(), _P()
        _check(lib().puck_grammar_reproject(coarse._h, fine._h, hierarchy.encode() if hierarchy else None,
                                            coarse_level, fine_level, ctypes.byref(hc), ctypes.byref(hf)))
        return Grammar(hc, coarse.prefix, coarse.semiring), Grammar(hf, fine.prefix, fine.semiring)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.puck_grammar_free(self._h)
            self._h = None

    def word_id(self, word):
        return lib().puck_grammar_word_id(self._h, word.encode())

    def rules(self, family):
        """(idx[n,3], w[n], q[n]) for family 0..3 binary NN/NT/TN/TT, 4 unary N, 5 unary T."""
        n = _check(lib().puck_grammar_rules(self._h, family, None, None, None, 0))
        idx = np.zeros((max(n, 1), 3), np.int32)
        w = np.zeros(max(n, 1), np.float32)
        q = np.zeros(max(n, 1), np.float32)
        _check(lib().puck_grammar_rules(self._h, family, idx.ctypes.data_as(_I32P), w.ctypes.data_as(_F32P),
                                        q.ctypes.data_as(_F32P), n))
        return idx[:n], w[:n], q[:n]

    def symbols(self, terminal):
        n = _check(lib().puck_grammar_symbols(self._h, int(terminal), None, None, 0))
        c = np.zeros(max(n, 1), np.int32)
        s = np.zeros(max(n, 1), np.int32)
        _check(lib().puck_grammar_symbols(self._h, int(terminal), c.ctypes.data_as(_I32P), s.ctypes.data_as(_I32P), n))
        return c[:n], s[:n]

    def coarse_names(self):
        """Coarse symbol names (.numstates order; a reprojected layer's are coarse substates)."""
        out = []
        buf = ctypes.create_string_buffer(256)
        for c in range(self.info.num_coarse):
            _check(lib().puck_grammar_coarse_name(self._h, c, buf, 256))
            out.append(buf.value.decode())
        return out

    def blocks(self, which, family):
        boc = np.zeros(self.info.num_coarse, np.int32)
        nb = _check(lib().puck_grammar_blocks(self._h, which, family, boc.ctypes.data_as(_I32P)))
        return nb, boc

    def lexicon(self, word):
        n = _check(lib().puck_grammar_lexicon(self._h, word, None, None, 0))
        t = np.zeros(max(n, 1), np.int32)
        lp = np.zeros(max(n, 1), np.float64)
        _check(lib().puck_grammar_lexicon(self._h, word, t.ctypes.data_as(_I32P), lp.ctypes.data_as(_F64P), n))
        return t[:n], lp[:n]

    def sample(self, seed, n, mode=0, min_len=5, max_len=20):
        """Seeded PCFG sentences: returns (words int32[total], offsets int32[n+1])."""
        cap = n * max(max_len, 80) + 16
        words = np.zeros(cap, np.int32)
        offsets = np.zeros(n + 1, np.int32)
        _check(lib().puck_sample_sentences(self._h, seed, n, mode, min_len, max_len, words.ctypes.data_as(_I32P),
                                           cap, offsets.ctypes.data_as(_I32P)))
        return words[:offsets[-1]].copy(), offsets


def save_grz(path, layers):
    """Writes the compiled layers (coarse first) into one grammar.grz (CLParserData.writeSequence)."""
    arr = (_P * len(layers))(*[g._h for g in layers])
    _check(lib().puck_grammar_save_grz(arr, len(layers), path.encode()))


def synth_wsj6(out_prefix, seed=6, data_dir=DATA_DIR):
    """Writes the synthetic wsj_6-shape grammar to out_prefix.* and returns its summary line."""
    _check(lib().puck_synth_wsj6(data_dir.encode(), out_prefix.encode(), seed))
    return lib().puck_last_error().decode()


def synth_wsj6_cached(seed=6, cache_dir=None):
    """Prefix of the synthetic wsj_6-shape grammar, generated once per seed into cache_dir."""
    cache_dir = cache_dir or os.environ.get("PUCK_CACHE", "/tmp/puck_cache")
    os.makedirs(cache_dir, exist_ok=True)
    prefix = os.path.join(cache_dir, f"wsj6syn_s{seed}.gr")
    done = prefix + ".done"
    # Concurrent processes (pytest workers) serialise on a lock file, so no
    # reader sees a grammar another process is still writing.
    import fcntl
    with open(prefix + ".lock", "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        if not (os.path.exists(done) and os.path.exists(prefix + ".hierarchy")):
            synth_wsj6(prefix, seed)
            with open(done, "w") as f:
                f.write("ok\n")
    return prefix


def _parse_with(obj, sentences, vit_fn, mbr_fn):
    w, wp, o, op, n = CLParser._pack(sentences)
    lens = np.diff(o)
    total = int(np.sum(2 * lens - 1))
    if obj.layers[-1].semiring == VITERBI:
        nodes = (TreeNode * max(total, 1))()
        _check(getattr(lib(), vit_fn)(obj._h, wp, op, n, nodes, total))
        arr = np.ctypeslib.as_array(nodes).view(np.int32).reshape(-1, 4)[:total]
    else:
        nodes = (MbrNode * max(total, 1))()
        _check(getattr(lib(), mbr_fn)(obj._h, wp, op, n, nodes, total))
        arr = np.ctypeslib.as_array(nodes).view(np.int32).reshape(-1, 5)[:total]
    out, s = [], 0
    for L in lens:
        out.append(arr[s:s + 2 * L - 1].copy())
        s += 2 * L - 1
    return out


class CLParser:
    """Coarse-to-fine CKY parser over grammar layers on one GPU (CLParser.scala:46-952)."""

    def __init__(self, layers, device=0, mem_bytes=16 << 30, items_per_lane=2, persistent=False, graphs=False,
                 outside_final=False, device_schedule=False, slot_partials=False):
        """outside_final: parse() also runs a Viterbi final layer's outside pass before the decode,
        as CLParser.insideOutside does (the trees are unchanged). device_schedule: each pass runs as
        one persistent launch over a device queue of all its spans (PUCK_FLAG_DEVICE_SCHEDULE).
        slot_partials: Viterbi layers write one partial row per rule slot instead of folding them
        per target (PUCK_FLAG_SLOT_PARTIALS; same results)."""
        self.layers = list(layers)
        arr = (_P * len(self.layers))(*[g._h for g in self.layers])
        h = _P()
        flags = ((FLAG_ONE_ITEM_PER_LANE if items_per_lane == 1 else 0) | (FLAG_PERSISTENT if persistent else 0)
                 | (FLAG_GRAPHS if graphs else 0) | (FLAG_OUTSIDE_FINAL if outside_final else 0)
                 | (FLAG_DEVICE_SCHEDULE if device_schedule else 0) | (FLAG_SLOT_PARTIALS if slot_partials else 0))
        _check(lib().puck_parser_create(arr, len(self.layers), device, mem_bytes, flags, ctypes.byref(h)))
        self._h = h

    def configure(self, persistent, piece_target=512, queue_wgs_per_cu=1, reduce_targets=4):
        """Launch shape of the rule and reduce kernels (see puck_parser_configure in include/puck.h)."""
        cfg = LaunchConfig(int(bool(persistent)), int(piece_target), int(queue_wgs_per_cu), int(reduce_targets))
        _check(lib().puck_parser_configure(self._h, ctypes.byref(cfg)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.puck_parser_free(self._h)
            self._h = None

    @staticmethod
    def _pack(sentences):
        if isinstance(sentences, tuple):
            words, offsets = sentences
        else:
            offsets = np.zeros(len(sentences) + 1, np.int32)
            offsets[1:] = np.cumsum([len(s) for s in sentences])
            words = np.concatenate([np.asarray(s, np.int32) for s in sentences]) if len(sentences) else np.zeros(0, np.int32)
        w, wp = _i32(words)
        o, op = _i32(offsets)
        return w, wp, o, op, len(o) - 1

    def parse(self, sentences):
        """Viterbi trees (list of TreeNode arrays) or MBR trees for a Real final layer."""
        return _parse_with(self, sentences, "puck_parse_viterbi", "puck_parse_mbr")

    def partitions(self, sentences):
        w, wp, o, op, n = self._pack(sentences)
        out = np.zeros(max(n, 1), np.float32)
        _check(lib().puck_partitions(self._h, wp, op, n, out.ctypes.data_as(_F32P)))
        return out[:n]

    def inside_outside(self, sentences):
        w, wp, o, op, n = self._pack(sentences)
        _check(lib().puck_inside_outside(self._h, wp, op, n))

    def masks(self, sentences):
        w, wp, o, op, n = self._pack(sentences)
        lens = np.diff(o)
        cells = int(np.sum(lens * (lens + 1)))
        mw = self.layers[0].info.mask_words
        out = np.zeros(max(cells * mw, 1), np.uint32)
        _check(lib().puck_masks(self._h, wp, op, n, out.ctypes.data_as(_U32P), out.size))
        return out[:cells * mw].reshape(cells, mw)

    def debug_charts(self, sentences, layer, with_outside=True):
        w, wp, o, op, n = self._pack(sentences)
        lens = np.diff(o)
        cells = int(np.sum(lens * (lens + 1)))
        cs = self.layers[layer].info.cell_size
        mw = self.layers[0].info.mask_words
        ins = np.zeros(cells * cs, np.float32)
        outs = np.zeros(cells * cs, np.float32)
        isc = np.zeros(cells, np.float32)
        osc = np.zeros(cells, np.float32)
        m = np.zeros(cells * mw, np.uint32)
        ps = np.zeros(n, np.float64)
        _check(lib().puck_debug_charts(self._h, wp, op, n, layer, int(with_outside), ins.ctypes.data_as(_F32P),
                                       outs.ctypes.data_as(_F32P), isc.ctypes.data_as(_F32P),
                                       osc.ctypes.data_as(_F32P), m.ctypes.data_as(_U32P),
                                       ps.ctypes.data_as(_F64P), ins.size))
        return dict(inside=ins.reshape(cells, cs), outside=outs.reshape(cells, cs), in_scale=isc,
                    out_scale=osc, masks=m.reshape(cells, mw), partition_scales=ps)

    def stats(self):
        s = Stats()
        _check(lib().puck_parser_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def launches(self):
        """Per rule-kernel launch of the last call: (ms float32[], rule_apps float64[], family int32[]);
        family = family id + 16 x layer index."""
        n = lib().puck_parser_launches(self._h, None, None, None, 0)
        ms = np.zeros(max(n, 1), np.float32)
        apps = np.zeros(max(n, 1), np.float64)
        fam = np.zeros(max(n, 1), np.int32)
        lib().puck_parser_launches(self._h, ms.ctypes.data_as(_F32P), apps.ctypes.data_as(_F64P),
                                   fam.ctypes.data_as(_I32P), n)
        return ms[:n], apps[:n], fam[:n]


class BatchParser:
    """Batched engine (puck_batch_* in include/puck.h): sentences sorted by length into buckets of
    `bucket_cells` chart cells that run concurrently on `streams` HIP streams, with persistent rule
    launches and per-shape HIP graphs. Same results as CLParser.parse."""

    def __init__(self, layers, device=0, mem_bytes=16 << 30, streams=3, bucket_cells=1 << 20, persistent=True,
                 graphs=True, piece_target=512, items_per_lane=2, outside_final=False, device_schedule=False,
                 reduce_targets=4, slot_partials=False):
        self.layers = list(layers)
        arr = (_P * len(self.layers))(*[g._h for g in self.layers])
        self.config = BatchConfig(int(streams), int(bucket_cells), int(bool(persistent)), int(bool(graphs)),
                                  int(piece_target), int(items_per_lane), int(bool(outside_final)),
                                  int(bool(device_schedule)), int(reduce_targets), int(bool(slot_partials)))
        h = _P()
        _check(lib().puck_batch_create(arr, len(self.layers), device, mem_bytes, ctypes.byref(self.config),
                                       ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.puck_batch_free(self._h)
            self._h = None

    def parse(self, sentences):
        return _parse_with(self, sentences, "puck_batch_parse_viterbi", "puck_batch_parse_mbr")

    def stats(self):
        s = BatchStats()
        _check(lib().puck_batch_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in BatchStats._fields_}


class Server:
    """Continuous-batching parse server (replaces BatchFunctionAnnotatorService,
    AnnotatorService.scala:29-78): submit sentences from any thread, collect results by ticket."""

    def __init__(self, layers, device=0, mem_bytes=16 << 30, max_inflight=2, max_batch=20000, items_per_lane=2,
                 persistent=True):
        self.layers = list(layers)
        arr = (_P * len(self.layers))(*[g._h for g in self.layers])
        h = _P()
        flags = (FLAG_ONE_ITEM_PER_LANE if items_per_lane == 1 else 0) | (FLAG_PERSISTENT if persistent else 0)
        _check(lib().puck_server_create(arr, len(self.layers), device, mem_bytes, max_inflight, max_batch, flags,
                                        ctypes.byref(h)))
        self._h = h
        self._mbr = self.layers[-1].semiring == REAL

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.puck_server_free(self._h)
            self._h = None

    def submit(self, words):
        w = np.ascontiguousarray(words, np.int32)
        t = ctypes.c_int64()
        _check(lib().puck_server_submit(self._h, w.ctypes.data_as(_I32P), len(w), ctypes.byref(t)))
        return t.value

    def result(self, ticket, n_words, wait=True):
        """Tree of a finished request (2n-1 nodes), or None if still pending and wait is False."""
        k = 2 * n_words - 1
        if self._mbr:
            buf = (MbrNode * k)()
        else:
            buf = (TreeNode * k)()
        rc = _check(lib().puck_server_result(self._h, ticket, int(wait), ctypes.cast(buf, ctypes.c_void_p), k))
        if rc == 0:
            return None
        return np.ctypeslib.as_array(buf).view(np.int32).reshape(k, -1).copy()

    def flush(self):
        _check(lib().puck_server_flush(self._h))

    def stats(self, reset=False):
        s = ServerStats()
        _check(lib().puck_server_stats(self._h, int(reset), ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}
