// extern "C" boundary over the grammar compiler, input generators and the CKY engine.
#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/puck.h"
#include "batch.h"
#include "engine.h"
#include "grz.h"
#include "server.h"
#include "synth.h"

struct puck_grammar {
  std::shared_ptr<puck::Grammar> g;
};

struct puck_parser {
  std::unique_ptr<puck::Parser> p;
  mutable std::mutex mu;  // serialises calls on one handle (the reference's synchronized methods)
};

struct puck_batch {
  std::unique_ptr<puck::BatchParser> b;
  std::mutex mu;
};

struct puck_server {
  std::unique_ptr<puck::Server> s;
  bool mbr = false;
};

namespace {

thread_local std::string g_last_error;

int32_t set_error(int32_t code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

template <class F>
int32_t guarded(F&& f) {
  try {
    g_last_error.clear();
    return f();
  } catch (const std::bad_alloc&) {
    return set_error(PUCK_ERR_NOMEM, "allocation failed");
  } catch (const std::runtime_error& e) {
    const std::string m = e.what();
    if (m.rfind("HIP error", 0) == 0) {
      if (m.find("out of memory") != std::string::npos) return set_error(PUCK_ERR_NOMEM, m);
      return set_error(PUCK_ERR_DEVICE, m);
    }
    if (m.rfind("cannot open", 0) == 0 || m.rfind("bad ", 0) == 0 || m.rfind("missing", 0) == 0)
      return set_error(PUCK_ERR_IO, m);
    return set_error(PUCK_ERR_ARG, m);
  } catch (const std::exception& e) {
    return set_error(PUCK_ERR_INTERNAL, e.what());
  }
}

int64_t nodes_needed(const int32_t* offsets, int32_t n) {
  int64_t t = 0;
  for (int32_t i = 0; i < n; ++i) t += 2 * (int64_t)(offsets[i + 1] - offsets[i]) - 1;
  return t;
}

}  // namespace

extern "C" {

const char* puck_last_error(void) { return g_last_error.c_str(); }
int32_t puck_abi_version(void) { return 3; }

int32_t puck_grammar_load_text(const char* prefix, int32_t semiring, puck_grammar_t** out) {
  return guarded([&]() -> int32_t {
    if (!prefix || !out) return set_error(PUCK_ERR_ARG, "null argument");
    if (semiring != PUCK_VITERBI && semiring != PUCK_REAL && semiring != PUCK_LOGSUM)
      return set_error(PUCK_ERR_ARG, "unknown semiring");
    auto* h = new puck_grammar;
    try {
      h->g = std::make_shared<puck::Grammar>(puck::load_text_grammar(prefix, (puck::Semiring)semiring));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return PUCK_OK;
  });
}

void puck_grammar_free(puck_grammar_t* g) { delete g; }

int32_t puck_grammar_reproject(const puck_grammar_t* coarse, const puck_grammar_t* fine, const char* hierarchy_path,
                               int32_t coarse_level, int32_t fine_level, puck_grammar_t** coarse_out,
                               puck_grammar_t** fine_out) {
  return guarded([&]() -> int32_t {
    if (!coarse || !fine || !coarse_out || !fine_out) return set_error(PUCK_ERR_ARG, "null argument");
    auto c = std::make_shared<puck::Grammar>(*coarse->g);
    auto f = std::make_shared<puck::Grammar>(*fine->g);
    const std::string path = hierarchy_path ? std::string(hierarchy_path) : f->prefix + ".hierarchy";
    if (coarse_level < 0) coarse_level = puck::grammar_level_from_name(c->prefix);
    if (fine_level < 0) fine_level = puck::grammar_level_from_name(f->prefix);
    puck::reproject_grammars(*c, *f, path, coarse_level, fine_level);
    auto* hc = new puck_grammar;
    hc->g = c;
    auto* hf = new puck_grammar;
    hf->g = f;
    *coarse_out = hc;
    *fine_out = hf;
    return PUCK_OK;
  });
}

int32_t puck_grammar_save_grz(const puck_grammar_t* const* layers, int32_t n_layers, const char* path) {
  return guarded([&]() -> int32_t {
    if (!layers || n_layers < 1 || !path) return set_error(PUCK_ERR_ARG, "bad argument");
    std::vector<const puck::Grammar*> gs;
    for (int32_t i = 0; i < n_layers; ++i) {
      if (!layers[i]) return set_error(PUCK_ERR_ARG, "null grammar layer");
      gs.push_back(layers[i]->g.get());
    }
    puck::save_grz(path, gs);
    return PUCK_OK;
  });
}

int32_t puck_grammar_load_grz(const char* path, puck_grammar_t** out, int32_t cap) {
  return guarded([&]() -> int32_t {
    if (!path) return set_error(PUCK_ERR_ARG, "null argument");
    auto gs = puck::load_grz(path);
    if ((int32_t)gs.size() > cap) return (int32_t)gs.size();
    if (!out) return set_error(PUCK_ERR_ARG, "null output array");
    for (size_t i = 0; i < gs.size(); ++i) {
      auto* h = new puck_grammar;
      h->g = gs[i];
      out[i] = h;
    }
    return (int32_t)gs.size();
  });
}

int32_t puck_grammar_info(const puck_grammar_t* h, puck_grammar_info_t* out) {
  return guarded([&]() -> int32_t {
    if (!h || !out) return set_error(PUCK_ERR_ARG, "null argument");
    const puck::Grammar& g = *h->g;
    std::memset(out, 0, sizeof(*out));
    out->num_nonterms = g.num_nonterms();
    out->num_terms = g.num_terms();
    out->num_coarse = (int32_t)g.coarse_names.size();
    out->cell_size = g.cell_size;
    out->mask_words = g.mask_words;
    out->root = g.root;
    for (int f = 0; f < 4; ++f) {
      out->num_binary[f] = (int32_t)g.binaries[f].size();
      out->num_blocks[f] = g.inside_blocks[f].num_blocks;
    }
    out->num_unary_n = (int32_t)g.unary_n.size();
    out->num_unary_t = (int32_t)g.unary_t.size();
    out->num_words = (int32_t)g.words.size();
    out->semiring = g.semiring;
    out->rules_dropped = g.num_rules_dropped;
    return PUCK_OK;
  });
}

int32_t puck_grammar_word_id(const puck_grammar_t* h, const char* word) {
  if (!h || !word) return -1;
  auto it = h->g->word_ids.find(word);
  return it == h->g->word_ids.end() ? -1 : it->second;
}

int32_t puck_grammar_rules(const puck_grammar_t* h, int32_t family, int32_t* idx, float* w, float* q,
                           int32_t cap) {
  return guarded([&]() -> int32_t {
    if (!h) return set_error(PUCK_ERR_ARG, "null grammar");
    const puck::Grammar& g = *h->g;
    int32_t n = 0;
    if (family >= 0 && family < 4) {
      const auto& v = g.binaries[family];
      for (const auto& r : v) {
        if (n < cap) {
          if (idx) { idx[3 * n] = r.p; idx[3 * n + 1] = r.l; idx[3 * n + 2] = r.r; }
          if (w) w[n] = r.w;
          if (q) q[n] = r.q;
        }
        ++n;
      }
    } else if (family == 4 || family == 5) {
      const auto& v = family == 4 ? g.unary_n : g.unary_t;
      for (const auto& r : v) {
        if (n < cap) {
          if (idx) { idx[3 * n] = r.p; idx[3 * n + 1] = r.c; idx[3 * n + 2] = -1; }
          if (w) w[n] = r.w;
          if (q) q[n] = r.q;
        }
        ++n;
      }
    } else {
      return set_error(PUCK_ERR_ARG, "family must be in [0, 5]");
    }
    return n;
  });
}

int32_t puck_grammar_symbols(const puck_grammar_t* h, int32_t terminal, int32_t* coarse, int32_t* sub,
                             int32_t cap) {
  return guarded([&]() -> int32_t {
    if (!h) return set_error(PUCK_ERR_ARG, "null grammar");
    const auto& v = terminal ? h->g->terms : h->g->nonterms;
    for (int32_t i = 0; i < (int32_t)v.size() && i < cap; ++i) {
      if (coarse) coarse[i] = v[i].coarse;
      if (sub) sub[i] = v[i].sub;
    }
    return (int32_t)v.size();
  });
}

int32_t puck_grammar_coarse_name(const puck_grammar_t* h, int32_t coarse, char* buf, int32_t cap) {
  return guarded([&]() -> int32_t {
    if (!h || coarse < 0 || coarse >= (int32_t)h->g->coarse_names.size()) return set_error(PUCK_ERR_ARG, "bad argument");
    const std::string& s = h->g->coarse_names[coarse];
    if (buf && cap > 0) {
      const size_t n = std::min<size_t>(s.size(), (size_t)cap - 1);
      std::memcpy(buf, s.data(), n);
      buf[n] = 0;
    }
    return (int32_t)s.size();
  });
}

int32_t puck_grammar_blocks(const puck_grammar_t* h, int32_t which, int32_t family, int32_t* boc) {
  return guarded([&]() -> int32_t {
    if (!h || family < 0 || family > 3 || which < 0 || which > 2) return set_error(PUCK_ERR_ARG, "bad argument");
    const puck::Grammar& g = *h->g;
    const puck::BlockPartition& bp =
        which == 0 ? g.inside_blocks[family] : which == 1 ? g.outside_l_blocks[family] : g.outside_r_blocks[family];
    if (boc) std::memcpy(boc, bp.block_of_coarse.data(), bp.block_of_coarse.size() * 4);
    return bp.num_blocks;
  });
}

int32_t puck_grammar_lexicon(const puck_grammar_t* h, int32_t word, int32_t* tags, double* logp, int32_t cap) {
  return guarded([&]() -> int32_t {
    if (!h) return set_error(PUCK_ERR_ARG, "null grammar");
    const puck::Grammar& g = *h->g;
    if (word < 0 || word + 1 >= (int32_t)g.lex_offsets.size()) return set_error(PUCK_ERR_ARG, "word id out of range");
    int32_t n = 0;
    for (int e = g.lex_offsets[word]; e < g.lex_offsets[word + 1]; ++e, ++n)
      if (n < cap) {
        if (tags) tags[n] = g.lex_tags[e];
        if (logp) logp[n] = g.lex_logp[e];
      }
    return n;
  });
}

int32_t puck_synth_wsj6(const char* data_dir, const char* out_prefix, uint64_t seed) {
  return guarded([&]() -> int32_t {
    if (!data_dir || !out_prefix) return set_error(PUCK_ERR_ARG, "null argument");
    g_last_error = puck::synth_wsj6(data_dir, out_prefix, seed);
    return PUCK_OK;
  });
}

int32_t puck_sample_sentences(const puck_grammar_t* h, uint64_t seed, int32_t n, int32_t mode,
                              int32_t min_len, int32_t max_len, int32_t* words_out, int64_t words_cap,
                              int32_t* offsets_out) {
  return guarded([&]() -> int32_t {
    if (!h || !words_out || !offsets_out || n < 0) return set_error(PUCK_ERR_ARG, "bad argument");
    std::vector<int32_t> w, o;
    if (!puck::sample_sentences(*h->g, seed, n, mode, min_len, max_len, &w, &o))
      return set_error(PUCK_ERR_INTERNAL, "sentence sampler exhausted its attempt budget");
    if ((int64_t)w.size() > words_cap) return set_error(PUCK_ERR_ARG, "words_out too small");
    std::memcpy(words_out, w.data(), w.size() * 4);
    std::memcpy(offsets_out, o.data(), o.size() * 4);
    return PUCK_OK;
  });
}

int32_t puck_parser_create(const puck_grammar_t* const* layers, int32_t n_layers, int32_t device,
                           int64_t mem_bytes, int32_t flags, puck_parser_t** out) {
  return guarded([&]() -> int32_t {
    if (!layers || n_layers < 1 || !out) return set_error(PUCK_ERR_ARG, "bad argument");
    std::vector<std::shared_ptr<puck::Grammar>> gs;
    for (int32_t i = 0; i < n_layers; ++i) {
      if (!layers[i]) return set_error(PUCK_ERR_ARG, "null grammar layer");
      gs.push_back(layers[i]->g);
    }
    const int32_t ipl = (flags & PUCK_FLAG_ONE_ITEM_PER_LANE) ? 1 : 2;
    if (flags & PUCK_FLAG_DEVICE_SCHEDULE) return set_error(PUCK_ERR_ARG, puck::kDeviceScheduleDisabled);
    auto* h = new puck_parser;
    try {
      h->p.reset(new puck::Parser(gs, device, mem_bytes > 0 ? mem_bytes : (int64_t)16 << 30, ipl));
      h->p->set_persistent((flags & PUCK_FLAG_PERSISTENT) != 0);
      h->p->use_graphs = (flags & PUCK_FLAG_GRAPHS) != 0;
      h->p->set_outside_final((flags & PUCK_FLAG_OUTSIDE_FINAL) != 0);
      h->p->set_device_schedule((flags & PUCK_FLAG_DEVICE_SCHEDULE) != 0);
      h->p->set_fold((flags & PUCK_FLAG_SLOT_PARTIALS) == 0);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return PUCK_OK;
  });
}

int32_t puck_parser_configure(puck_parser_t* p, const puck_launch_config_t* cfg) {
  return guarded([&]() -> int32_t {
    if (!p || !cfg) return set_error(PUCK_ERR_ARG, "bad argument");
    if (cfg->piece_target < 1 || cfg->queue_wgs_per_cu < 1 || cfg->queue_wgs_per_cu > 8)
      return set_error(PUCK_ERR_ARG, "piece_target must be >= 1 and queue_wgs_per_cu in [1, 8]");
    const int32_t rt = cfg->reduce_targets;
    if (rt != 1 && rt != 2 && rt != 4 && rt != 8 && rt != 16)
      return set_error(PUCK_ERR_ARG, "reduce_targets must be 1, 2, 4, 8 or 16");
    std::lock_guard<std::mutex> lk(p->mu);
    p->p->set_persistent(cfg->persistent != 0);
    p->p->set_launch_shape(cfg->piece_target, cfg->queue_wgs_per_cu);
    p->p->set_reduce_targets(rt);
    return PUCK_OK;
  });
}

void puck_parser_free(puck_parser_t* p) { delete p; }

int32_t puck_parse_viterbi(puck_parser_t* h, const int32_t* words, const int32_t* offsets, int32_t n,
                           puck_tree_node_t* nodes, int64_t cap) {
  return guarded([&]() -> int32_t {
    if (!h || !words || !offsets || !nodes || n < 0) return set_error(PUCK_ERR_ARG, "bad argument");
    if (nodes_needed(offsets, n) > cap) return set_error(PUCK_ERR_ARG, "node buffer too small");
    std::vector<std::vector<puck::TreeNode>> trees;
    std::lock_guard<std::mutex> lk(h->mu);
    h->p->parse(words, offsets, n, &trees, nullptr);
    int64_t o = 0;
    for (auto& t : trees) {
      std::memcpy(nodes + o, t.data(), t.size() * sizeof(puck_tree_node_t));
      o += (int64_t)t.size();
    }
    return PUCK_OK;
  });
}

int32_t puck_parse_viterbi_noalloc(puck_parser_t* h, const int32_t* words, const int32_t* offsets, int32_t n,
                                   puck_tree_node_t* nodes, int64_t cap) {
  return guarded([&]() -> int32_t {
    if (!h || !words || !offsets || !nodes || n < 0) return set_error(PUCK_ERR_ARG, "bad argument");
    if (nodes_needed(offsets, n) > cap) return set_error(PUCK_ERR_ARG, "node buffer too small");
    std::vector<std::vector<puck::TreeNode>> trees;
    std::lock_guard<std::mutex> lk(h->mu);
    h->p->parse_noalloc(words, offsets, n, &trees, nullptr);
    int64_t o = 0;
    for (auto& t : trees) {
      std::memcpy(nodes + o, t.data(), t.size() * sizeof(puck_tree_node_t));
      o += (int64_t)t.size();
    }
    return PUCK_OK;
  });
}

int32_t puck_parse_mbr(puck_parser_t* h, const int32_t* words, const int32_t* offsets, int32_t n,
                       puck_mbr_node_t* nodes, int64_t cap) {
  return guarded([&]() -> int32_t {
    if (!h || !words || !offsets || !nodes || n < 0) return set_error(PUCK_ERR_ARG, "bad argument");
    if (nodes_needed(offsets, n) > cap) return set_error(PUCK_ERR_ARG, "node buffer too small");
    std::vector<std::vector<puck::MbrNode>> trees;
    std::lock_guard<std::mutex> lk(h->mu);
    h->p->parse(words, offsets, n, nullptr, &trees);
    int64_t o = 0;
    for (auto& t : trees) {
      std::memcpy(nodes + o, t.data(), t.size() * sizeof(puck_mbr_node_t));
      o += (int64_t)t.size();
    }
    return PUCK_OK;
  });
}

int32_t puck_partitions(puck_parser_t* h, const int32_t* words, const int32_t* offsets, int32_t n, float* out) {
  return guarded([&]() -> int32_t {
    if (!h || !words || !offsets || !out || n < 0) return set_error(PUCK_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(h->mu);
    h->p->partitions(words, offsets, n, out);
    return PUCK_OK;
  });
}

int32_t puck_inside_outside(puck_parser_t* h, const int32_t* words, const int32_t* offsets, int32_t n) {
  return guarded([&]() -> int32_t {
    if (!h || !words || !offsets || n < 0) return set_error(PUCK_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(h->mu);
    h->p->inside_outside(words, offsets, n);
    return PUCK_OK;
  });
}

int32_t puck_masks(puck_parser_t* h, const int32_t* words, const int32_t* offsets, int32_t n, uint32_t* out,
                   int64_t cap) {
  return guarded([&]() -> int32_t {
    if (!h || !words || !offsets || !out || n < 0) return set_error(PUCK_ERR_ARG, "bad argument");
    std::vector<uint32_t> m;
    std::vector<int32_t> co;
    std::lock_guard<std::mutex> lk(h->mu);
    h->p->masks(words, offsets, n, &m, &co);
    if ((int64_t)m.size() > cap) return set_error(PUCK_ERR_ARG, "mask buffer too small");
    std::memcpy(out, m.data(), m.size() * 4);
    return PUCK_OK;
  });
}

int32_t puck_debug_charts(puck_parser_t* h, const int32_t* words, const int32_t* offsets, int32_t n,
                          int32_t layer, int32_t with_outside, float* inside, float* outside,
                          float* in_scale, float* out_scale, uint32_t* masks, double* partition_scales,
                          int64_t chart_cap) {
  return guarded([&]() -> int32_t {
    if (!h || !words || !offsets || n < 1) return set_error(PUCK_ERR_ARG, "bad argument");
    puck::Parser::Charts c;
    std::lock_guard<std::mutex> lk(h->mu);
    h->p->debug_charts(words, offsets, n, layer, with_outside != 0, &c);
    if ((int64_t)c.inside.size() > chart_cap) return set_error(PUCK_ERR_ARG, "chart buffer too small");
    if (inside) std::memcpy(inside, c.inside.data(), c.inside.size() * 4);
    if (outside) std::memcpy(outside, c.outside.data(), c.outside.size() * 4);
    if (in_scale) std::memcpy(in_scale, c.in_scale.data(), c.in_scale.size() * 4);
    if (out_scale) std::memcpy(out_scale, c.out_scale.data(), c.out_scale.size() * 4);
    if (masks && !c.masks.empty()) std::memcpy(masks, c.masks.data(), c.masks.size() * 4);
    if (partition_scales)
      for (int32_t i = 0; i < n; ++i)
        partition_scales[i] = i < (int32_t)c.partition_scales.size() ? c.partition_scales[i] : 0.0;
    return PUCK_OK;
  });
}

int32_t puck_parser_stats(const puck_parser_t* h, puck_stats_t* out) {
  if (!h || !out) return set_error(PUCK_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(h->mu);
  const puck::PassStats& s = h->p->stats;
  out->rule_kernel_ms = s.rule_kernel_ms;
  out->reduce_ms = s.reduce_ms;
  out->other_ms = s.other_ms;
  out->total_ms = s.total_ms;
  out->rule_apps = s.rule_apps;
  out->rule_launches = s.rule_launches;
  out->device_ms = s.device_ms;
  return PUCK_OK;
}

int64_t puck_parser_launches(const puck_parser_t* h, float* ms, double* apps, int32_t* family, int64_t cap) {
  if (!h) return set_error(PUCK_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(h->mu);
  const puck::PassStats& s = h->p->stats;
  const int64_t n = (int64_t)s.launch_ms.size();
  for (int64_t i = 0; i < n && i < cap; ++i) {
    if (ms) ms[i] = s.launch_ms[i];
    if (apps) apps[i] = s.launch_apps[i];
    if (family) family[i] = s.launch_family[i];
  }
  return n;
}

}  // extern "C"

extern "C" {

int32_t puck_server_create(const puck_grammar_t* const* layers, int32_t n_layers, int32_t device,
                           int64_t mem_bytes, int32_t max_inflight, int32_t max_batch, int32_t flags,
                           puck_server_t** out) {
  return guarded([&]() -> int32_t {
    if (!layers || n_layers < 1 || !out || max_inflight < 1) return set_error(PUCK_ERR_ARG, "bad argument");
    std::vector<std::shared_ptr<puck::Grammar>> gs;
    for (int32_t i = 0; i < n_layers; ++i) {
      if (!layers[i]) return set_error(PUCK_ERR_ARG, "null grammar layer");
      gs.push_back(layers[i]->g);
    }
    auto* h = new puck_server;
    h->mbr = gs.back()->semiring != puck::kViterbi;
    try {
      h->s.reset(new puck::Server(gs, device, mem_bytes > 0 ? mem_bytes : (int64_t)16 << 30, max_inflight,
                                  max_batch > 0 ? max_batch : 20000, (flags & PUCK_FLAG_ONE_ITEM_PER_LANE) ? 1 : 2,
                                  (flags & PUCK_FLAG_PERSISTENT) != 0));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return PUCK_OK;
  });
}

void puck_server_free(puck_server_t* s) { delete s; }

int32_t puck_server_submit(puck_server_t* h, const int32_t* words, int32_t len, int64_t* ticket) {
  return guarded([&]() -> int32_t {
    if (!h || !words || !ticket) return set_error(PUCK_ERR_ARG, "bad argument");
    *ticket = h->s->submit(words, len);
    return PUCK_OK;
  });
}

int32_t puck_server_result(puck_server_t* h, int64_t ticket, int32_t wait, void* nodes, int64_t cap) {
  return guarded([&]() -> int32_t {
    if (!h) return set_error(PUCK_ERR_ARG, "bad argument");
    std::vector<puck::TreeNode> t;
    std::vector<puck::MbrNode> m;
    int32_t status = 0;
    std::string err;
    if (!h->s->result(ticket, wait != 0, &t, &m, &status, &err)) return 0;
    if (status < 0) return set_error(status, err);
    if (h->mbr) {
      if ((int64_t)m.size() > cap) return set_error(PUCK_ERR_ARG, "node buffer too small");
      if (nodes) std::memcpy(nodes, m.data(), m.size() * sizeof(puck_mbr_node_t));
    } else {
      if ((int64_t)t.size() > cap) return set_error(PUCK_ERR_ARG, "node buffer too small");
      if (nodes) std::memcpy(nodes, t.data(), t.size() * sizeof(puck_tree_node_t));
    }
    return 1;
  });
}

int32_t puck_server_flush(puck_server_t* h) {
  return guarded([&]() -> int32_t {
    if (!h) return set_error(PUCK_ERR_ARG, "bad argument");
    h->s->flush();
    return PUCK_OK;
  });
}

int32_t puck_server_stats(puck_server_t* h, int32_t reset, puck_server_stats_t* out) {
  return guarded([&]() -> int32_t {
    if (!h || !out) return set_error(PUCK_ERR_ARG, "bad argument");
    puck::ServerStats s = h->s->stats();
    out->completed = s.completed;
    out->batches = s.batches;
    out->mean_batch = s.mean_batch;
    out->p50_ms = s.p50_ms;
    out->p90_ms = s.p90_ms;
    out->p99_ms = s.p99_ms;
    out->max_ms = s.max_ms;
    if (reset) h->s->reset_stats();
    return PUCK_OK;
  });
}

}  // extern "C"

int32_t puck_batch_create(const puck_grammar_t* const* layers, int32_t n_layers, int32_t device, int64_t mem_bytes,
                          const puck_batch_config_t* cfg, puck_batch_t** out) {
  return guarded([&]() -> int32_t {
    if (!layers || n_layers < 1 || !out || !cfg) return set_error(PUCK_ERR_ARG, "bad argument");
    std::vector<std::shared_ptr<puck::Grammar>> gs;
    for (int32_t i = 0; i < n_layers; ++i) {
      if (!layers[i]) return set_error(PUCK_ERR_ARG, "null grammar layer");
      gs.push_back(layers[i]->g);
    }
    if (cfg->items_per_lane != 1 && cfg->items_per_lane != 2) return set_error(PUCK_ERR_ARG, "items_per_lane must be 1 or 2");
    if (cfg->piece_target < 1) return set_error(PUCK_ERR_ARG, "piece_target must be >= 1");
    if (cfg->device_schedule) return set_error(PUCK_ERR_ARG, puck::kDeviceScheduleDisabled);
    const int32_t rt = cfg->reduce_targets;
    if (rt != 1 && rt != 2 && rt != 4 && rt != 8 && rt != 16)
      return set_error(PUCK_ERR_ARG, "reduce_targets must be 1, 2, 4, 8 or 16");
    puck::BatchConfig c;
    c.streams = cfg->streams;
    c.bucket_cells = cfg->bucket_cells;
    c.persistent = cfg->persistent != 0;
    c.graphs = cfg->graphs != 0;
    c.piece_target = cfg->piece_target;
    c.items_per_lane = cfg->items_per_lane;
    c.outside_final = cfg->outside_final != 0;
    c.device_schedule = cfg->device_schedule != 0;
    c.reduce_targets = rt;
    c.fold = cfg->slot_partials == 0;
    auto* h = new puck_batch;
    try {
      h->b.reset(new puck::BatchParser(gs, device, mem_bytes > 0 ? mem_bytes : (int64_t)16 << 30, c));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return PUCK_OK;
  });
}

void puck_batch_free(puck_batch_t* b) { delete b; }

int32_t puck_batch_parse_viterbi(puck_batch_t* h, const int32_t* words, const int32_t* offsets, int32_t n,
                                 puck_tree_node_t* nodes, int64_t nodes_cap) {
  return guarded([&]() -> int32_t {
    if (!h || !words || !offsets || n < 0 || !nodes) return set_error(PUCK_ERR_ARG, "bad argument");
    if (nodes_needed(offsets, n) > nodes_cap) return set_error(PUCK_ERR_ARG, "nodes buffer too small");
    std::vector<std::vector<puck::TreeNode>> trees;
    std::lock_guard<std::mutex> lk(h->mu);
    h->b->parse(words, offsets, n, &trees, nullptr);
    int64_t o = 0;
    for (auto& t : trees) {
      std::memcpy(nodes + o, t.data(), t.size() * sizeof(puck::TreeNode));
      o += (int64_t)t.size();
    }
    return PUCK_OK;
  });
}

int32_t puck_batch_parse_mbr(puck_batch_t* h, const int32_t* words, const int32_t* offsets, int32_t n,
                             puck_mbr_node_t* nodes, int64_t nodes_cap) {
  return guarded([&]() -> int32_t {
    if (!h || !words || !offsets || n < 0 || !nodes) return set_error(PUCK_ERR_ARG, "bad argument");
    if (nodes_needed(offsets, n) > nodes_cap) return set_error(PUCK_ERR_ARG, "nodes buffer too small");
    std::vector<std::vector<puck::MbrNode>> trees;
    std::lock_guard<std::mutex> lk(h->mu);
    h->b->parse(words, offsets, n, nullptr, &trees);
    int64_t o = 0;
    for (auto& t : trees) {
      std::memcpy(nodes + o, t.data(), t.size() * sizeof(puck::MbrNode));
      o += (int64_t)t.size();
    }
    return PUCK_OK;
  });
}

int32_t puck_batch_stats(puck_batch_t* h, puck_batch_stats_t* out) {
  if (!h || !out) return set_error(PUCK_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(h->mu);
  out->device_ms = h->b->stats.device_ms;
  out->buckets = h->b->stats.buckets;
  out->graph_captures = h->b->stats.graph_captures;
  out->graph_replays = h->b->stats.graph_replays;
  return PUCK_OK;
}
