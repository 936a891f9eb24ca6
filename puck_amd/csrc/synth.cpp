// Synthetic wsj_6-shape grammar and PCFG sentence sampling. See synth.h.
#include "synth.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

namespace puck {

double Rng::normal() {
  double u1 = uniform(), u2 = uniform();
  if (u1 < 1e-300) u1 = 1e-300;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

namespace {

[[noreturn]] void fail(const std::string& m) { throw std::runtime_error(m); }

std::string base_of(const std::string& s, int* sub) {
  size_t u = s.rfind('_');
  if (u == std::string::npos) fail("symbol without substate: " + s);
  *sub = std::atoi(s.c_str() + u + 1);
  return s.substr(0, u);
}

struct Coarse {
  std::vector<std::string> names;
  std::unordered_map<std::string, int> id;
  std::vector<int> n6, n1;
};

int sub_map(int j, int n6, int n1) { return (int)((int64_t)j * n1 / n6); }

}  // namespace

std::string synth_wsj6(const std::string& dir, const std::string& out, uint64_t seed) {
  Rng rng(seed);
  Coarse C;
  {
    std::ifstream in(dir + "/wsj_6.gr.numstates");
    if (!in) fail("missing " + dir + "/wsj_6.gr.numstates");
    std::string s;
    int n;
    while (in >> s >> n) {
      C.id[s] = (int)C.names.size();
      C.names.push_back(s);
      C.n6.push_back(n);
    }
    C.n1.assign(C.names.size(), 1);
    std::ifstream in1(dir + "/wsj_1.gr.numstates");
    if (!in1) fail("missing wsj_1.gr.numstates");
    while (in1 >> s >> n) C.n1.at(C.id.at(s)) = n;
  }
  const int NC = (int)C.names.size();

  // wsj_1 refined probabilities.
  std::unordered_map<std::string, double> p1b, p1u;
  {
    std::ifstream in(dir + "/wsj_1.gr.binary");
    std::string a, b, c;
    double p;
    while (in >> a >> b >> c >> p) p1b[a + " " + b + " " + c] = p;
    std::ifstream iu(dir + "/wsj_1.gr.unary");
    while (iu >> a >> b >> p) p1u[a + " " + b] = p;
  }
  auto rname = [&](int c, int s) { return C.names[c] + "_" + std::to_string(s); };

  // Coarse rules from wsj_0 in file order; refined counts from wsj_6.gr.num.binary.
  struct CB { int a, b, c; int64_t count; };
  std::vector<CB> cbin;
  {
    std::ifstream in(dir + "/wsj_0.gr.binary");
    std::string a, b, c;
    double p;
    int s;
    while (in >> a >> b >> c >> p) cbin.push_back({C.id.at(base_of(a, &s)), C.id.at(base_of(b, &s)), C.id.at(base_of(c, &s)), 0});
  }
  std::vector<int64_t> counts;
  {
    std::ifstream in(dir + "/wsj_6.gr.num.binary");
    int64_t v;
    while (in >> v) counts.push_back(v);
  }
  int64_t target_binary = 0;
  for (auto v : counts) target_binary += v;
  int64_t have = 0;
  for (size_t i = 0; i < cbin.size(); ++i) {
    const int64_t cap = (int64_t)C.n6[cbin[i].a] * C.n6[cbin[i].b] * C.n6[cbin[i].c];
    cbin[i].count = std::min<int64_t>(i < counts.size() ? counts[i] : 1, cap);
    have += cbin[i].count;
  }
  {
    // Redistribute what clipping (and the count without a wsj_0 rule) removed.
    std::vector<size_t> order(cbin.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    auto spare = [&](size_t i) {
      return (int64_t)C.n6[cbin[i].a] * C.n6[cbin[i].b] * C.n6[cbin[i].c] - cbin[i].count;
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return spare(x) > spare(y); });
    int64_t rem = target_binary - have;
    while (rem > 0) {
      bool moved = false;
      for (size_t k = 0; k < order.size() && rem > 0; ++k) {
        int64_t add = std::min<int64_t>({spare(order[k]), rem, std::max<int64_t>(1, rem / 64)});
        if (add > 0) {
          cbin[order[k]].count += add;
          rem -= add;
          moved = true;
        }
      }
      if (!moved) fail("wsj_6 binary counts exceed the coarse rule capacity");
    }
  }

  struct CU { int a, c; int64_t count; };
  std::vector<CU> cun;
  {
    std::ifstream in(dir + "/wsj_0.gr.unary");
    std::string a, c;
    double p;
    int s;
    while (in >> a >> c >> p)
      if (a != c) cun.push_back({C.id.at(base_of(a, &s)), C.id.at(base_of(c, &s)), 0});
  }
  int64_t target_unary = 0;
  {
    std::ifstream in(dir + "/wsj_6.gr.num.unary");
    int64_t v;
    while (in >> v) target_unary += v;
  }
  int64_t total_syms = 0;
  for (int c = 0; c < NC; ++c) total_syms += C.n6[c];
  {
    // Identity unaries (one per refined symbol) are part of the published unary count.
    const int64_t want = target_unary - total_syms;
    double tot_cap = 0;
    for (auto& u : cun) tot_cap += (double)C.n6[u.a] * C.n6[u.c];
    int64_t got = 0;
    for (auto& u : cun) {
      const int64_t cap = (int64_t)C.n6[u.a] * C.n6[u.c];
      u.count = std::min<int64_t>(cap, std::max<int64_t>(1, (int64_t)std::floor(want * (cap / tot_cap))));
      got += u.count;
    }
    int64_t rem = want - got;
    for (size_t k = 0; rem != 0; k = (k + 1) % cun.size()) {
      const int64_t cap = (int64_t)C.n6[cun[k].a] * C.n6[cun[k].c];
      if (rem > 0 && cun[k].count < cap) { cun[k].count++; rem--; }
      else if (rem < 0 && cun[k].count > 1) { cun[k].count--; rem++; }
    }
  }

  // Refined rules: distinct tuples per coarse rule; the first tuples walk the substates diagonally
  // so that every substate of every symbol of the coarse rule occurs.
  struct RB { int a, b, c; double w; };
  struct RU { int a, c; double w; };
  std::vector<std::vector<RB>> rb(cbin.size());
  std::vector<std::vector<RU>> ru(cun.size());
  for (size_t i = 0; i < cbin.size(); ++i) {
    const CB& r = cbin[i];
    const int na = C.n6[r.a], nb = C.n6[r.b], nc = C.n6[r.c];
    std::unordered_set<int64_t> seen;
    auto add = [&](int a, int b, int c) {
      const int64_t key = ((int64_t)a * nb + b) * nc + c;
      if (!seen.insert(key).second) return false;
      const int a1 = sub_map(a, na, C.n1[r.a]), b1 = sub_map(b, nb, C.n1[r.b]), c1 = sub_map(c, nc, C.n1[r.c]);
      auto it = p1b.find(rname(r.a, a1) + " " + rname(r.b, b1) + " " + rname(r.c, c1));
      double base = it == p1b.end() ? 1e-6 : it->second;
      rb[i].push_back({a, b, c, base * std::exp(0.5 * rng.normal())});
      return true;
    };
    const int diag = std::max({na, nb, nc});
    for (int k = 0; k < diag && (int64_t)rb[i].size() < r.count; ++k) add(k % na, k % nb, k % nc);
    while ((int64_t)rb[i].size() < r.count) add((int)rng.below(na), (int)rng.below(nb), (int)rng.below(nc));
  }
  for (size_t i = 0; i < cun.size(); ++i) {
    const CU& r = cun[i];
    const int na = C.n6[r.a], nc = C.n6[r.c];
    std::unordered_set<int64_t> seen;
    auto add = [&](int a, int c) {
      if (r.a == r.c && a == c) return false;
      if (!seen.insert((int64_t)a * nc + c).second) return false;
      const int a1 = sub_map(a, na, C.n1[r.a]), c1 = sub_map(c, nc, C.n1[r.c]);
      auto it = p1u.find(rname(r.a, a1) + " " + rname(r.c, c1));
      double base = it == p1u.end() ? 1e-6 : it->second;
      ru[i].push_back({a, c, base * std::exp(0.5 * rng.normal())});
      return true;
    };
    const int diag = std::max(na, nc);
    for (int k = 0; k < diag && (int64_t)ru[i].size() < r.count; ++k) add(k % na, (k + (r.a == r.c)) % nc);
    int guard = 0;
    while ((int64_t)ru[i].size() < r.count && guard++ < 100000000) add((int)rng.below(na), (int)rng.below(nc));
  }
  // Normalise per refined parent to a proper distribution, keep every rule above e^-11.5 so the
  // loader's -12 threshold keeps the published counts.
  std::vector<std::vector<double>> mass(NC);
  for (int c = 0; c < NC; ++c) mass[c].assign(C.n6[c], 0.0);
  for (size_t i = 0; i < cbin.size(); ++i)
    for (auto& x : rb[i]) mass[cbin[i].a][x.a] += x.w;
  for (size_t i = 0; i < cun.size(); ++i)
    for (auto& x : ru[i]) mass[cun[i].a][x.a] += x.w;
  const double floor_p = std::exp(-11.5);
  auto norm = [&](int c, int a, double w) { return std::max(floor_p, w / mass[c][a]); };

  std::string summary;
  {
    FILE* f = std::fopen((out + ".numstates").c_str(), "w");
    if (!f) fail("cannot write " + out + ".numstates");
    for (int c = 0; c < NC; ++c) std::fprintf(f, "%s\t%d\n", C.names[c].c_str(), C.n6[c]);
    std::fclose(f);
    f = std::fopen((out + ".binary").c_str(), "w");
    if (!f) fail("cannot write " + out + ".binary");
    std::unique_ptr<char[]> buf(new char[1 << 22]);
    std::setvbuf(f, buf.get(), _IOFBF, 1 << 22);
    int64_t nb = 0;
    for (size_t i = 0; i < cbin.size(); ++i)
      for (auto& x : rb[i]) {
        std::fprintf(f, "%s_%d %s_%d %s_%d %.17g\n", C.names[cbin[i].a].c_str(), x.a, C.names[cbin[i].b].c_str(), x.b,
                     C.names[cbin[i].c].c_str(), x.c, norm(cbin[i].a, x.a, x.w));
        ++nb;
      }
    std::fclose(f);
    f = std::fopen((out + ".unary").c_str(), "w");
    if (!f) fail("cannot write " + out + ".unary");
    int64_t nu = 0;
    for (int c = 0; c < NC; ++c)
      for (int s = 0; s < C.n6[c]; ++s) {
        std::fprintf(f, "%s_%d %s_%d 1.0\n", C.names[c].c_str(), s, C.names[c].c_str(), s);
        ++nu;
      }
    for (size_t i = 0; i < cun.size(); ++i)
      for (auto& x : ru[i]) {
        std::fprintf(f, "%s_%d %s_%d %.17g\n", C.names[cun[i].a].c_str(), x.a, C.names[cun[i].c].c_str(), x.c,
                     norm(cun[i].a, x.a, x.w));
        ++nu;
      }
    std::fclose(f);
    summary = "binary=" + std::to_string(nb) + " unary=" + std::to_string(nu);
  }
  // Lexicon: P6(w | T_j) from P1(w | T_h(j)) with seeded perturbation, normalised per refined tag.
  {
    std::ifstream in(dir + "/wsj_1.gr.lexicon");
    if (!in) fail("missing wsj_1.gr.lexicon");
    struct Line { int tag; std::string word; std::vector<double> p; };
    std::vector<Line> lines;
    std::string line;
    while (std::getline(in, line)) {
      if (line.empty()) continue;
      size_t sp = line.find(' '), lb = line.rfind('['), rb2 = line.rfind(']');
      Line L;
      L.tag = C.id.at(line.substr(0, sp));
      L.word = line.substr(sp + 1, lb - sp - 2);
      std::vector<double> p1;
      const char* p = line.c_str() + lb + 1;
      const char* e = line.c_str() + rb2;
      while (p < e) {
        while (p < e && (*p == ' ' || *p == ',')) ++p;
        if (p >= e) break;
        char* q;
        p1.push_back(std::strtod(p, &q));
        p = q;
      }
      const int n6 = C.n6[L.tag], n1 = (int)p1.size();
      L.p.resize(n6);
      for (int j = 0; j < n6; ++j) L.p[j] = p1[sub_map(j, n6, n1)] * std::exp(0.3 * rng.normal());
      lines.push_back(std::move(L));
    }
    std::vector<std::vector<double>> colsum(NC);
    for (int c = 0; c < NC; ++c) colsum[c].assign(C.n6[c], 0.0);
    for (auto& L : lines)
      for (size_t j = 0; j < L.p.size(); ++j) colsum[L.tag][j] += L.p[j];
    FILE* f = std::fopen((out + ".lexicon").c_str(), "w");
    if (!f) fail("cannot write " + out + ".lexicon");
    std::unique_ptr<char[]> buf(new char[1 << 22]);
    std::setvbuf(f, buf.get(), _IOFBF, 1 << 22);
    for (auto& L : lines) {
      std::fprintf(f, "%s %s [", C.names[L.tag].c_str(), L.word.c_str());
      for (size_t j = 0; j < L.p.size(); ++j)
        std::fprintf(f, "%s%.17g", j ? ", " : "", colsum[L.tag][j] > 0 ? L.p[j] / colsum[L.tag][j] : 0.0);
      std::fprintf(f, "]\n");
    }
    std::fclose(f);
  }
  {
    // The split hierarchy of wsj_6: the synthetic symbols have its substate counts, so its trees
    // describe them (used by the hierarchical reprojection of a coarse/fine pair).
    std::ifstream in(dir + "/wsj_6.gr.hierarchy", std::ios::binary);
    if (!in) fail("missing wsj_6.gr.hierarchy");
    std::ofstream o(out + ".hierarchy", std::ios::binary);
    o << in.rdbuf();
    if (!o) fail("cannot write " + out + ".hierarchy");
  }
  return summary;
}

// -------------------------------------------------------------------------------------------------
bool sample_sentences(const Grammar& g, uint64_t seed, int32_t n, int32_t mode, int32_t min_len,
                      int32_t max_len, std::vector<int32_t>* words, std::vector<int32_t>* offsets) {
  Rng rng(seed);
  const int nN = g.num_nonterms(), nT = g.num_terms();
  auto prob = [&](float w) { return g.semiring == kReal ? (double)w : std::exp((double)w); };
  // Expansions of each nonterminal: binary rules (any family) and non-identity unaries.
  struct Exp { int kind; int l, r; };  // kind 0 binary (l, r with terminal flags), 1 unary
  std::vector<std::vector<Exp>> ex(nN);
  std::vector<std::vector<double>> cdf(nN);
  for (int f = 0; f < 4; ++f) {
    const bool lt = (f == kTN || f == kTT), rt = (f == kNT || f == kTT);
    for (auto& r : g.binaries[f]) {
      ex[r.p].push_back({0, lt ? -1 - r.l : r.l, rt ? -1 - r.r : r.r});
      cdf[r.p].push_back(prob(r.w));
    }
  }
  for (auto& u : g.unary_n)
    if (u.p != u.c) {
      ex[u.p].push_back({1, u.c, 0});
      cdf[u.p].push_back(prob(u.w));
    }
  for (auto& u : g.unary_t) {
    ex[u.p].push_back({1, -1 - u.c, 0});
    cdf[u.p].push_back(prob(u.w));
  }
  for (int p = 0; p < nN; ++p)
    for (size_t i = 1; i < cdf[p].size(); ++i) cdf[p][i] += cdf[p][i - 1];
  // Emissions per tag.
  std::vector<std::vector<int32_t>> em_w(nT);
  std::vector<std::vector<double>> em_c(nT);
  for (int32_t w = 0; w + 1 < (int32_t)g.lex_offsets.size(); ++w)
    for (int e = g.lex_offsets[w]; e < g.lex_offsets[w + 1]; ++e) {
      em_w[g.lex_tags[e]].push_back(w);
      em_c[g.lex_tags[e]].push_back(std::exp(g.lex_logp[e]));
    }
  for (int t = 0; t < nT; ++t)
    for (size_t i = 1; i < em_c[t].size(); ++i) em_c[t][i] += em_c[t][i - 1];
  auto pick = [&](const std::vector<double>& c) -> int {
    if (c.empty()) return -1;
    const double u = rng.uniform() * c.back();
    return (int)(std::upper_bound(c.begin(), c.end(), u) - c.begin() >= (long)c.size()
                     ? c.size() - 1
                     : std::upper_bound(c.begin(), c.end(), u) - c.begin());
  };
  // Sample one sentence with at most `cap` words; returns false on overflow or dead ends.
  std::vector<int32_t> out;
  auto sample_one = [&](int cap) -> bool {
    out.clear();
    std::vector<int> stack{g.root};
    int steps = 0;
    while (!stack.empty()) {
      if (++steps > 20000) return false;
      const int s = stack.back();
      stack.pop_back();
      if (s < 0) {
        const int t = -1 - s;
        const int k = pick(em_c[t]);
        if (k < 0) return false;
        out.push_back(em_w[t][k]);
        if ((int)out.size() > cap) return false;
        continue;
      }
      const int k = pick(cdf[s]);
      if (k < 0) return false;
      const Exp& e = ex[s][k];
      if (e.kind == 0) {
        stack.push_back(e.r);
        stack.push_back(e.l);
      } else {
        stack.push_back(e.l);
      }
    }
    return true;
  };
  words->clear();
  offsets->assign(1, 0);
  for (int32_t i = 0; i < n; ++i) {
    int target = -1, lo = min_len, hi = max_len;
    if (mode == kLenLogNormal || mode == kLenMixed) {
      bool longer = (mode == kLenMixed) && rng.uniform() < 0.3;
      if (longer) {
        lo = 41;
        hi = 80;
        target = lo + (int)rng.below(hi - lo + 1);
      } else {
        if (mode == kLenMixed) { lo = 1; hi = 40; }
        do {
          target = (int)std::lround(std::exp(3.049 + 0.416 * rng.normal()));
        } while (target < lo || target > hi);
      }
    }
    bool ok = false;
    for (int attempt = 0; attempt < 2000000 && !ok; ++attempt) {
      const int cap = target > 0 ? target : hi;
      if (!sample_one(cap)) continue;
      const int len = (int)out.size();
      ok = target > 0 ? len == target : (len >= lo && len <= hi);
    }
    if (!ok) return false;
    words->insert(words->end(), out.begin(), out.end());
    offsets->push_back((int32_t)words->size());
  }
  return true;
}

}  // namespace puck
