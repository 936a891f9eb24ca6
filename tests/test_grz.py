"""grammar.grz container (CPU): a zip with the reference's entry names whose binary payloads
load back into identical compiled tables; corrupted or truncated files are rejected."""
import zipfile

import numpy as np
import pytest


@pytest.fixture(scope="module")
def grz(tmp_path_factory, puck, wsj0, wsj1):
    path = str(tmp_path_factory.mktemp("grz") / "grammar.grz")
    puck.save_grz(path, [wsj0, wsj1])
    return path


# Entry names the reference writes per layer (CLParser.scala:1249-1260 and the write methods it
# calls: CLInside.scala:16-23, CLOutside.scala:27-38, CLParserUtils.scala:14-16,
# CLMaskKernels.scala:23-25, CLMBRKernels.scala:20-22, CLViterbi.scala:73-80,
# CLScalingKernels.scala:19-20; a kernel is `<name>.program` + `<name>.name`, ZipUtil.scala:86-89).
REFERENCE_LAYER_ENTRIES = (
    ["grammar", "structure", "semiring"]
    + [f"{b}/numKernels" for b in ("insideNT", "insideNN", "insideTN", "insideTT", "insideNU", "insideTU",
                                   "outside_L_NN", "outside_R_NN", "outside_L_NT", "outside_R_NT", "outside_L_TN",
                                   "outside_R_TN", "outside_L_TT", "outside_R_TT", "outsideNU", "outsideTU")]
    + [f"{b}/{e}" for b in ("insideNN", "outside_L_NN") for e in ("directWrite", "extra", "enqueuer/ints",
                                                                 "enqueuer/enqueue.name", "0/bits", "0/rules")]
    + [f"{k}.{x}" for k in ("setRootScoresKernel", "getRootScoresKernel", "computeMasksKernel", "computeMBRKernel",
                            "computeViterbiKernel", "getScalingConstantsKernel") for x in ("program", "name")]
    + ["MasksInts", "MBRInts", "ViterbiWGSize", "ViterbiParents", "ViterbiLeft", "ViterbiRight", "ViterbiScores",
       "ViterbiOffsets"])


def test_zip_layout_and_entry_names(grz, puck, wsj0, wsj1):
    with zipfile.ZipFile(grz) as z:
        names = set(z.namelist())
        assert z.namelist()[0] == "parserNames"
        # CLParserData.writeSequence (CLParser.scala:1307-1310): "parser-i" lines.
        assert z.read("parserNames").decode().split("\n") == ["parser-0", "parser-1"]
        for i, g in enumerate((wsj0, wsj1)):
            for e in REFERENCE_LAYER_ENTRIES:
                assert f"parser-{i}/{e}" in names, e
            # one rule-kernel entry per CoarseParent block of the inside families
            for f, fam in enumerate(("NN", "NT", "TN", "TT")):
                assert z.read(f"parser-{i}/inside{fam}/numKernels") == str(g.info.num_blocks[f]).encode()
            n_nn = np.frombuffer(z.read(f"parser-{i}/insideNN/0/rules"), np.int32)
            assert n_nn.size > 0 and np.all(np.diff(n_nn) > 0)
            off = np.frombuffer(z.read(f"parser-{i}/ViterbiOffsets"), np.int32)
            assert list(off) == [k * g.info.num_nonterms for k in range(6)]
        assert z.testzip() is None  # CRCs agree
        assert z.read("parser-1/semiring") == b"Viterbi"
        assert z.read("parser-0/computeViterbiKernel.name") == b"k_viterbi"


def test_block_rule_lists_partition_each_family(grz, wsj1):
    """The per-block rule lists of a bundle partition the family's rules."""
    with zipfile.ZipFile(grz) as z:
        for fam in ("NN", "NT", "TN", "TT"):
            nb = int(z.read(f"parser-1/inside{fam}/numKernels"))
            rules = np.concatenate([np.frombuffer(z.read(f"parser-1/inside{fam}/{j}/rules"), np.int32)
                                    for j in range(nb)])
            n = len(wsj1.rules(("NN", "NT", "TN", "TT").index(fam))[0])
            assert sorted(rules.tolist()) == list(range(n))


def test_tables_round_trip(grz, puck, wsj0, wsj1):
    back = puck.Grammar.load_grz(grz)
    assert len(back) == 2
    for a, b in zip([wsj0, wsj1], back):
        for f in ("num_nonterms", "num_terms", "num_coarse", "cell_size", "mask_words", "root",
                  "num_unary_n", "num_unary_t", "num_words", "semiring", "rules_dropped"):
            assert getattr(a.info, f) == getattr(b.info, f), f
        assert list(a.info.num_binary) == list(b.info.num_binary)
        assert list(a.info.num_blocks) == list(b.info.num_blocks)
        for fam in range(6):
            for x, y in zip(a.rules(fam), b.rules(fam)):
                assert np.array_equal(x, y)
        for t in (0, 1):
            for x, y in zip(a.symbols(t), b.symbols(t)):
                assert np.array_equal(x, y)
        for which in range(3):
            for fam in range(4):
                na, ba = a.blocks(which, fam)
                nb, bb = b.blocks(which, fam)
                assert na == nb and np.array_equal(ba, bb)
        for wid in (0, 7, 1234, a.info.num_words - 1):
            for x, y in zip(a.lexicon(wid), b.lexicon(wid)):
                assert np.array_equal(x, y)
        assert a.word_id("the") == b.word_id("the")


def test_real_semiring_layer(tmp_path, puck, wsj0, wsj1_real):
    path = str(tmp_path / "real.grz")
    puck.save_grz(path, [wsj0, wsj1_real])
    back = puck.Grammar.load_grz(path)
    assert back[0].info.semiring == puck.VITERBI and back[1].info.semiring == puck.REAL
    _, _, qa = wsj1_real.rules(0)
    _, _, qb = back[1].rules(0)
    assert np.array_equal(qa, qb)


def test_logsum_layer(tmp_path, puck, wsj0):
    ls = puck.Grammar.load_text(puck.grammar_path("wsj_1"), puck.LOGSUM)
    path = str(tmp_path / "ls.grz")
    puck.save_grz(path, [wsj0, ls])
    with zipfile.ZipFile(path) as z:
        assert z.read("parser-1/semiring") == b"LogSum"
    back = puck.Grammar.load_grz(path)
    assert back[1].info.semiring == puck.LOGSUM
    for x, y in zip(ls.rules(0), back[1].rules(0)):
        assert np.array_equal(x, y)


def test_corrupted_and_truncated_files_are_rejected(tmp_path, grz, puck):
    data = bytearray(open(grz, "rb").read())
    bad = tmp_path / "bad.grz"
    data[200] ^= 0xFF  # inside the first payload
    bad.write_bytes(bytes(data))
    with pytest.raises(puck.PuckError, match="CRC|bad grz"):
        puck.Grammar.load_grz(str(bad))
    short = tmp_path / "short.grz"
    short.write_bytes(open(grz, "rb").read()[:1000])
    with pytest.raises(puck.PuckError, match="bad grz"):
        puck.Grammar.load_grz(str(short))
    with pytest.raises(puck.PuckError, match="cannot open"):
        puck.Grammar.load_grz(str(tmp_path / "missing.grz"))
