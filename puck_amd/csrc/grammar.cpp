// Grammar compilation for the CKY engine. See grammar.h for the reference entities mirrored here.
#include "grammar.h"

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <unordered_set>

namespace puck {

namespace {

[[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(msg); }

std::ifstream open_or_fail(const std::string& path) {
  std::ifstream in(path);
  if (!in) fail("cannot open grammar file " + path);
  return in;
}

// "NP_12" -> ("NP", 12). Symbols are split at the last underscore (Berkeley refined names).
bool split_refined(const std::string& s, std::string* base, int* sub) {
  size_t u = s.rfind('_');
  if (u == std::string::npos || u + 1 >= s.size()) return false;
  char* end = nullptr;
  long v = std::strtol(s.c_str() + u + 1, &end, 10);
  if (*end != '\0') return false;
  *base = s.substr(0, u);
  *sub = (int)v;
  return true;
}

struct RawBinary { std::string p, l, r; double prob; };
struct RawUnary { std::string p, c; double prob; };

}  // namespace

// ---------------------------------------------------------------------------------------------
// Java Formatter "%f" (SimpleGenRuleMultiply.java:151 formats rule scores with String.format("%ff")).
// Java widens the float to double, takes the shortest decimal digits of that double
// (FloatingDecimal), rounds them HALF_UP at 6 fraction digits (FormattedFloatingDecimal
// .applyPrecision) and zero-pads. The OpenCL compiler then parses the literal as a float.
std::string java_format_f6_string(float v) {
  const double d = (double)v;
  if (std::isnan(d)) return "NaN";
  if (std::isinf(d)) return d < 0 ? "-Infinity" : "Infinity";
  const bool neg = std::signbit(d);
  std::string out = neg ? "-" : "";
  if (d == 0.0) return out + "0.000000";

  char buf[64];
  auto res = std::to_chars(buf, buf + sizeof(buf), std::fabs(d), std::chars_format::scientific);
  *res.ptr = '\0';
  // buf = "D.DDDDe[+-]XX"
  std::string sci(buf);
  size_t epos = sci.find('e');
  std::string mant = sci.substr(0, epos);
  int x = std::atoi(sci.c_str() + epos + 1);
  std::string digits;
  for (char c : mant)
    if (c != '.') digits.push_back(c);
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  int dec_exp = x + 1;  // value = 0.digits * 10^dec_exp
  const int nd = (int)digits.size();
  const int prec = dec_exp + 6;
  if (prec >= 0 && prec < nd) {
    if (prec == 0) {
      if (digits[0] >= '5') {
        digits = "1";
        dec_exp += 1;
      } else {
        digits = "0";
      }
    } else {
      char q = digits[prec];
      if (q >= '5') {
        int i = prec - 1;
        while (i >= 0 && digits[i] == '9') --i;
        if (i < 0) {
          digits = "1";
          dec_exp += 1;
        } else {
          digits[i] = (char)(digits[i] + 1);
          digits.resize(i + 1);
        }
      } else {
        digits.resize(prec);
      }
    }
  }
  // fillDecimal: integer part then exactly 6 fraction digits.
  std::string ip, fp;
  if (dec_exp > 0) {
    for (int i = 0; i < dec_exp; ++i) ip.push_back(i < (int)digits.size() ? digits[i] : '0');
    for (int i = dec_exp; i < dec_exp + 6; ++i) fp.push_back(i < (int)digits.size() ? digits[i] : '0');
  } else {
    ip = "0";
    for (int i = 0; i < 6; ++i) {
      int di = i + dec_exp;  // digit index for fraction position i
      fp.push_back(di >= 0 && di < (int)digits.size() ? digits[di] : '0');
    }
  }
  while (ip.size() > 1 && ip[0] == '0') ip.erase(0, 1);
  return out + ip + "." + fp;
}

float java_format_f6(float v) {
  std::string s = java_format_f6_string(v);
  return std::strtof(s.c_str(), nullptr);
}

// ---------------------------------------------------------------------------------------------
// CoarseParentSymbolGrammarClusterer.segmentBinaries (CoarseParentSymbolGrammarClusterer.java:38-121):
// group rules by coarse parent, sort groups by size descending; groups with >= 300 rules are a
// segment each, the first group below 300 absorbs every remaining group. Each segment becomes one
// kernel whose parent mask is its set of coarse parents (JavaFriendlyGenRuleMultiply.scala:248-256);
// kernels with identical masks form a block (CLParser.scala:784). Equal-size ties (HashMap order in
// the reference) are broken by ascending coarse id.
BlockPartition coarse_parent_partition(const std::vector<int32_t>& coarse_key_per_rule,
                                       int32_t num_coarse, int32_t mask_words) {
  std::vector<int64_t> counts(num_coarse, 0);
  for (int32_t c : coarse_key_per_rule) counts[c]++;
  std::vector<int32_t> present;
  for (int32_t c = 0; c < num_coarse; ++c)
    if (counts[c] > 0) present.push_back(c);
  std::stable_sort(present.begin(), present.end(),
                   [&](int32_t a, int32_t b) { return counts[a] > counts[b]; });
  BlockPartition bp;
  bp.block_of_coarse.assign(num_coarse, -1);
  size_t i = 0;
  for (; i < present.size() && counts[present[i]] >= 300; ++i) {
    bp.block_of_coarse[present[i]] = bp.num_blocks++;
  }
  if (i < present.size()) {
    int32_t lump = bp.num_blocks++;
    for (; i < present.size(); ++i) bp.block_of_coarse[present[i]] = lump;
  }
  bp.masks.assign((size_t)bp.num_blocks * mask_words, 0u);
  for (int32_t c = 0; c < num_coarse; ++c) {
    int32_t b = bp.block_of_coarse[c];
    if (b >= 0) bp.masks[(size_t)b * mask_words + c / 32] |= 1u << (c % 32);
  }
  return bp;
}

// ---------------------------------------------------------------------------------------------
// Tables derived from the rules and the symbol projections: CoarseParent partitions and the
// Viterbi decoder arrays. Rebuilt after a reprojection changes the projections or rule scores.
void build_tables(Grammar& g) {
  const int32_t num_coarse = (int32_t)g.coarse_names.size();
  g.mask_words = (num_coarse + 31) / 32;
  g.dec_offsets.clear();
  g.dec_scores.clear();
  g.dec_left.clear();
  g.dec_right.clear();
  // CoarseParent partitions for the 4 inside families and the 8 rotated outside families
  // (CLOutside.scala:61-71 rotates the updated child into the parent slot).
  for (int f = 0; f < 4; ++f) {
    const bool lt = (f == kTN || f == kTT), rt = (f == kNT || f == kTT);
    std::vector<int32_t> kp, kl, kr;
    for (auto& r : g.binaries[f]) {
      kp.push_back(g.proj_n[r.p]);
      kl.push_back(lt ? g.proj_t[r.l] : g.proj_n[r.l]);
      kr.push_back(rt ? g.proj_t[r.r] : g.proj_n[r.r]);
    }
    g.inside_blocks[f] = coarse_parent_partition(kp, num_coarse, g.mask_words);
    g.outside_l_blocks[f] = coarse_parent_partition(kl, num_coarse, g.mask_words);
    g.outside_r_blocks[f] = coarse_parent_partition(kr, num_coarse, g.mask_words);
  }

  // Viterbi decoder tables (CLViterbi.scala:287-309, 610-628).
  {
    const int32_t nn = g.num_nonterms();
    g.dec_offsets.assign(1, 0);
    auto add_group = [&](std::vector<int32_t> parents, std::vector<float> scores,
                         std::vector<int32_t> lefts, std::vector<int32_t> rights, bool binary) {
      std::vector<int32_t> order(parents.size());
      for (size_t i = 0; i < order.size(); ++i) order[i] = (int32_t)i;
      std::stable_sort(order.begin(), order.end(),
                       [&](int32_t a, int32_t b) { return parents[a] < parents[b]; });
      const int32_t start = (int32_t)g.dec_scores.size();
      std::vector<int32_t> sorted_parents;
      for (int32_t i : order) {
        sorted_parents.push_back(parents[i]);
        g.dec_scores.push_back(scores[i]);
        g.dec_left.push_back(lefts[i]);
        if (binary) g.dec_right.push_back(rights[i]);
      }
      for (int32_t sym = 0; sym < nn; ++sym) {
        auto it = std::upper_bound(sorted_parents.begin(), sorted_parents.end(), sym);
        g.dec_offsets.push_back(start + (int32_t)(it - sorted_parents.begin()));
      }
    };
    for (int f = 0; f < 4; ++f) {
      std::vector<int32_t> p, l, r;
      std::vector<float> s;
      for (auto& br : g.binaries[f]) {
        p.push_back(br.p);
        l.push_back(br.l);
        r.push_back(br.r);
        s.push_back(br.w);
      }
      add_group(p, s, l, r, true);
    }
    for (auto* us : {&g.unary_n, &g.unary_t}) {
      std::vector<int32_t> p, c;
      std::vector<float> s;
      for (auto& ur : *us) {
        p.push_back(ur.p);
        c.push_back(ur.c);
        s.push_back(ur.w);
      }
      add_group(p, s, c, {}, false);
    }
  }
}

Grammar load_text_grammar(const std::string& prefix, Semiring semiring, double log_threshold) {
  Grammar g;
  g.prefix = prefix;
  g.semiring = semiring;

  // .numstates: coarse symbols in line order.
  std::unordered_map<std::string, int32_t> coarse_id;
  {
    auto in = open_or_fail(prefix + ".numstates");
    std::string sym;
    int n;
    while (in >> sym >> n) {
      coarse_id[sym] = (int32_t)g.coarse_names.size();
      g.coarse_names.push_back(sym);
      g.coarse_substates.push_back(n);
    }
    if (g.coarse_names.empty()) fail("empty numstates: " + prefix);
  }
  const int32_t num_coarse = (int32_t)g.coarse_names.size();
  g.mask_words = (num_coarse + 31) / 32;

  std::vector<RawBinary> raw_b;
  std::vector<RawUnary> raw_u;
  {
    auto in = open_or_fail(prefix + ".binary");
    std::string line;
    while (std::getline(in, line)) {
      if (line.empty()) continue;
      std::istringstream ss(line);
      RawBinary r;
      if (!(ss >> r.p >> r.l >> r.r >> r.prob)) fail("bad binary rule line: " + line);
      raw_b.push_back(r);
    }
  }
  {
    auto in = open_or_fail(prefix + ".unary");
    std::string line;
    while (std::getline(in, line)) {
      if (line.empty()) continue;
      std::istringstream ss(line);
      RawUnary r;
      if (!(ss >> r.p >> r.c >> r.prob)) fail("bad unary rule line: " + line);
      raw_u.push_back(r);
    }
  }
  g.num_rules_read = (int32_t)(raw_b.size() + raw_u.size());

  // Threshold (parseBerkeleyText(prefix, -12, ...), CLParser.scala:1026).
  auto keep = [&](double p) { return p > 0 && std::log(p) >= log_threshold; };
  std::vector<RawBinary> kb;
  std::vector<RawUnary> ku;
  for (auto& r : raw_b)
    if (keep(r.prob)) kb.push_back(r);
  for (auto& r : raw_u)
    if (r.p == r.c || keep(r.prob)) ku.push_back(r);
  g.num_rules_dropped = g.num_rules_read - (int32_t)(kb.size() + ku.size());

  // RuleStructure.scala:150-160: nonterminals are parents of binaries and of non-identity unaries;
  // terminals are the remaining right-hand-side symbols.
  std::set<std::string> nonterm_set, rhs_set;
  for (auto& r : kb) {
    nonterm_set.insert(r.p);
    rhs_set.insert(r.l);
    rhs_set.insert(r.r);
  }
  for (auto& r : ku)
    if (r.p != r.c) {
      nonterm_set.insert(r.p);
      rhs_set.insert(r.c);
    }
  struct Key { int32_t coarse, sub; std::string name; };
  std::vector<Key> nk, tk;
  auto mk = [&](const std::string& s) {
    std::string base;
    int sub;
    if (!split_refined(s, &base, &sub)) fail("symbol without substate: " + s);
    auto it = coarse_id.find(base);
    if (it == coarse_id.end()) fail("symbol not in numstates: " + s);
    return Key{it->second, sub, s};
  };
  for (auto& s : nonterm_set) nk.push_back(mk(s));
  for (auto& s : rhs_set)
    if (!nonterm_set.count(s)) tk.push_back(mk(s));
  // GPU id order: (coarse id, substate). The reference orders by symbols_dump, which only moves
  // symbols around in the chart (RuleStructure.scala:162).
  auto by_cs = [](const Key& a, const Key& b) {
    return a.coarse != b.coarse ? a.coarse < b.coarse : a.sub < b.sub;
  };
  std::sort(nk.begin(), nk.end(), by_cs);
  std::sort(tk.begin(), tk.end(), by_cs);
  std::unordered_map<std::string, int32_t> nid, tid;
  for (auto& k : nk) {
    nid[k.name] = (int32_t)g.nonterms.size();
    g.nonterms.push_back(Symbol{k.name, k.coarse, k.sub, false, (int32_t)g.nonterms.size()});
    g.proj_n.push_back(k.coarse);
  }
  for (auto& k : tk) {
    tid[k.name] = (int32_t)g.terms.size();
    g.terms.push_back(Symbol{k.name, k.coarse, k.sub, true, (int32_t)g.terms.size()});
    g.proj_t.push_back(k.coarse);
  }
  auto root_it = nid.find("ROOT_0");
  if (root_it == nid.end()) fail("grammar has no ROOT_0 nonterminal: " + prefix);
  g.root = root_it->second;
  g.num_syms = std::max(g.num_nonterms(), g.num_terms());
  g.cell_size = (g.num_syms + 31) / 32 * 32;

  auto to_w = [&](double prob) {
    float s = (float)std::log(prob);
    if (semiring == kReal) return (float)std::exp((double)s);
    return s;
  };

  int32_t rule_id = 0;
  for (auto& r : kb) {
    bool lt = tid.count(r.l) > 0, rt = tid.count(r.r) > 0;
    BinaryRule br;
    br.p = nid.at(r.p);
    br.l = lt ? tid.at(r.l) : nid.at(r.l);
    br.r = rt ? tid.at(r.r) : nid.at(r.r);
    br.w = to_w(r.prob);
    br.q = java_format_f6(br.w);
    br.id = rule_id++;
    int fam = lt ? (rt ? kTT : kTN) : (rt ? kNT : kNN);
    g.binaries[fam].push_back(br);
  }
  for (auto& r : ku) {
    if (r.p == r.c) {
      // Identity unaries: kept for nonterminals (RuleStructure.scala:191), terminal identities are
      // unused (:194), identities of symbols that occur in no other rule cannot be indexed.
      if (!nid.count(r.p)) continue;
    }
    UnaryRule ur;
    ur.p = nid.at(r.p);
    ur.w = to_w(r.prob);
    ur.q = java_format_f6(ur.w);
    ur.id = rule_id++;
    if (nid.count(r.c)) {
      ur.c = nid.at(r.c);
      g.unary_n.push_back(ur);
    } else {
      ur.c = tid.at(r.c);
      g.unary_t.push_back(ur);
    }
  }

  build_tables(g);

  // Lexicon: "TAG word [p_0, ..., p_k]" = P(word | TAG_i).
  {
    auto in = open_or_fail(prefix + ".lexicon");
    std::string line;
    std::vector<std::vector<std::pair<int32_t, double>>> entries;
    while (std::getline(in, line)) {
      if (line.empty()) continue;
      size_t sp = line.find_first_of(" \t");
      size_t lb = line.rfind('[');
      size_t rb = line.rfind(']');
      if (sp == std::string::npos || lb == std::string::npos || rb == std::string::npos || lb < sp)
        fail("bad lexicon line: " + line);
      std::string tag = line.substr(0, sp);
      size_t ws = line.find_first_not_of(" \t", sp);
      size_t we = line.find_last_not_of(" \t", lb - 1);
      if (ws == std::string::npos || we == std::string::npos || we < ws) fail("bad lexicon line: " + line);
      std::string word = line.substr(ws, we - ws + 1);
      auto wit = g.word_ids.find(word);
      int32_t w;
      if (wit == g.word_ids.end()) {
        w = (int32_t)g.words.size();
        g.word_ids.emplace(word, w);
        g.words.push_back(word);
        entries.emplace_back();
      } else {
        w = wit->second;
      }
      const char* p = line.c_str() + lb + 1;
      const char* end = line.c_str() + rb;
      int sub = 0;
      while (p < end) {
        while (p < end && (*p == ' ' || *p == ',')) ++p;
        if (p >= end) break;
        char* q = nullptr;
        double prob = std::strtod(p, &q);
        if (q == p) fail("bad lexicon probability: " + line);
        p = q;
        if (prob > 0) {
          auto it = tid.find(tag + "_" + std::to_string(sub));
          if (it != tid.end()) entries[w].emplace_back(it->second, std::log(prob));
        }
        ++sub;
      }
    }
    g.lex_offsets.assign(1, 0);
    for (auto& e : entries) {
      std::sort(e.begin(), e.end());
      for (auto& pr : e) {
        g.lex_tags.push_back(pr.first);
        g.lex_logp.push_back(pr.second);
      }
      g.lex_offsets.push_back((int32_t)g.lex_tags.size());
    }
  }
  return g;
}

namespace {

// One bracketed node of a .hierarchy line, read the way readHierarchy reads it with
// PennTreeReader.parseHard (CLParser.scala:1187-1188): a node whose brackets hold only words is a
// leaf spanning those words; other nodes hold only subtrees.
struct HierNode {
  std::string label;
  std::vector<int32_t> kids;
  int32_t w0 = 0, w1 = 0;  // the node's words are [w0, w1) of the line's word list
  bool has_words = false;
};

struct HierTree {
  std::vector<HierNode> nodes;  // preorder, nodes[0] is the root
  std::vector<std::string> words;
};

HierTree parse_hier_tree(const std::string& text) {
  HierTree t;
  size_t i = 0;
  auto skip = [&]() {
    while (i < text.size() && std::isspace((unsigned char)text[i])) ++i;
  };
  auto atom = [&]() {
    skip();
    size_t j = i;
    while (j < text.size() && !std::isspace((unsigned char)text[j]) && text[j] != '(' && text[j] != ')') ++j;
    if (j == i) fail("bad hierarchy tree: " + text);
    std::string a = text.substr(i, j - i);
    i = j;
    return a;
  };
  std::vector<int32_t> open;
  skip();
  if (i >= text.size() || text[i] != '(') fail("hierarchy tree must start with '(': " + text);
  do {
    skip();
    if (i >= text.size()) fail("unbalanced hierarchy tree: " + text);
    if (text[i] == '(') {
      ++i;
      HierNode nd;
      nd.label = atom();
      nd.w0 = (int32_t)t.words.size();
      const int32_t id = (int32_t)t.nodes.size();
      if (!open.empty()) {
        if (t.nodes[open.back()].has_words) fail("hierarchy node mixes words and subtrees: " + text);
        t.nodes[open.back()].kids.push_back(id);
      }
      t.nodes.push_back(std::move(nd));
      open.push_back(id);
    } else if (text[i] == ')') {
      ++i;
      t.nodes[open.back()].w1 = (int32_t)t.words.size();
      open.pop_back();
    } else {
      HierNode& nd = t.nodes[open.back()];
      if (!nd.kids.empty()) fail("hierarchy node mixes words and subtrees: " + text);
      nd.has_words = true;
      t.words.push_back(atom());
    }
  } while (!open.empty());
  skip();
  if (i != text.size()) fail("trailing text after hierarchy tree: " + text);
  return t;
}

}  // namespace

std::map<std::string, std::vector<std::string>> read_hierarchy(const std::string& path, int32_t coarse_level,
                                                               int32_t fine_level) {
  std::map<std::string, std::vector<std::string>> out;
  auto in = open_or_fail(path);
  std::string line;
  while (std::getline(in, line)) {
    size_t a = 0;
    while (a < line.size() && std::isspace((unsigned char)line[a])) ++a;
    if (a == line.size()) continue;
    size_t b = a;
    while (b < line.size() && !std::isspace((unsigned char)line[b])) ++b;
    const std::string sym = line.substr(a, b - a);
    const HierTree t = parse_hier_tree(line.substr(b));
    const int32_t nn = (int32_t)t.nodes.size();
    // leftHeight: 0 for a leaf, else 1 + leftHeight of the first child. Children follow their
    // parent in preorder, so one backward sweep fills heights and subtree sizes.
    std::vector<int32_t> lh(nn, 0), size(nn, 1);
    for (int32_t v = nn - 1; v >= 0; --v) {
      const HierNode& nd = t.nodes[v];
      if (!nd.kids.empty()) lh[v] = 1 + lh[nd.kids[0]];
      for (int32_t k : nd.kids) size[v] += size[k];
    }
    const int32_t total = lh[0];
    for (int32_t v = 0; v < nn; ++v) {
      if (lh[v] != total - coarse_level) continue;
      std::vector<std::string> fines;
      if (fine_level - 1 == total) {
        for (int32_t w = t.nodes[v].w0; w < t.nodes[v].w1; ++w) fines.push_back(sym + "_" + t.words[w]);
      } else {
        for (int32_t d = v; d < v + size[v]; ++d)
          if (lh[d] == total - fine_level) fines.push_back(sym + "_" + t.nodes[d].label);
      }
      out[sym + "_" + t.nodes[v].label] = std::move(fines);  // a later node of one name replaces it
    }
  }
  return out;
}

int32_t grammar_level_from_name(const std::string& name) {
  // readHierarchy's "^.*(?:^|[^0-9])([0-9]+)([^0-9]*)$" -> the last run of digits.
  size_t end = name.size();
  while (end > 0 && !std::isdigit((unsigned char)name[end - 1])) --end;
  if (end == 0) fail("no split level in grammar name: " + name);
  size_t beg = end;
  while (beg > 0 && std::isdigit((unsigned char)name[beg - 1])) --beg;
  return std::stoi(name.substr(beg, end - beg));
}

void reproject_grammars(Grammar& coarse, Grammar& fine, const std::string& hierarchy_path, int32_t coarse_level,
                        int32_t fine_level) {
  if (coarse_level < 0 || fine_level <= coarse_level)
    fail("reprojection needs 0 <= coarse level < fine level (got " + std::to_string(coarse_level) + ", " +
         std::to_string(fine_level) + ")");
  // fine refined symbol -> coarse-level symbol (reverseSymMap). The root has no hierarchy line;
  // the reference maps its root label to itself (TOP_0 there, ROOT_0 in the text files).
  std::unordered_map<std::string, std::string> up;
  for (const auto& kv : read_hierarchy(hierarchy_path, coarse_level, fine_level))
    for (const auto& f : kv.second) {
      auto ins = up.emplace(f, kv.first);
      if (!ins.second && ins.first->second != kv.first)
        fail("hierarchy puts " + f + " under both " + ins.first->second + " and " + kv.first);
    }
  up["ROOT_0"] = "ROOT_0";

  // The coarse layer's refined symbols become the coarse symbols of both layers (the label index
  // of coarseGrammar.topology), in (base symbol, substate) order.
  std::vector<const Symbol*> cs;
  for (const auto& x : coarse.nonterms) cs.push_back(&x);
  for (const auto& x : coarse.terms) cs.push_back(&x);
  std::sort(cs.begin(), cs.end(), [](const Symbol* a, const Symbol* b) {
    return a->coarse != b->coarse ? a->coarse < b->coarse : a->sub < b->sub;
  });
  if ((int32_t)cs.size() > kMaxCoarseSymbols)
    fail("reprojection gives " + std::to_string(cs.size()) + " coarse symbols; at most " +
         std::to_string(kMaxCoarseSymbols) + " are supported");
  std::vector<std::string> names;
  std::unordered_map<std::string, int32_t> cid;
  for (const Symbol* x : cs) {
    cid[x->name] = (int32_t)names.size();
    names.push_back(x->name);
  }

  // Coarse layer: identity refinements (GrammarRefinements.identity), scores unchanged.
  for (size_t i = 0; i < coarse.nonterms.size(); ++i)
    coarse.proj_n[i] = coarse.nonterms[i].coarse = cid.at(coarse.nonterms[i].name);
  for (size_t i = 0; i < coarse.terms.size(); ++i)
    coarse.proj_t[i] = coarse.terms[i].coarse = cid.at(coarse.terms[i].name);
  coarse.coarse_names = names;
  coarse.coarse_substates.assign(names.size(), 1);

  // Rules of the coarse layer over the new symbol ids: a fine rule keeps its score only if its
  // projection is one of them (GrammarRefinements(..., skipMissingCoarseRules = true) leaves the
  // others unprojected, and CLParserData.make scores those fromLogSpace(-12), CLParser.scala:1266-1274).
  auto key3 = [](int64_t p, int64_t l, int64_t r) { return (p << 40) | (l << 20) | r; };
  std::unordered_set<int64_t> cbin, cun;
  for (int f = 0; f < 4; ++f) {
    const bool lt = (f == kTN || f == kTT), rt = (f == kNT || f == kTT);
    for (const auto& r : coarse.binaries[f])
      cbin.insert(key3(coarse.proj_n[r.p], lt ? coarse.proj_t[r.l] : coarse.proj_n[r.l],
                       rt ? coarse.proj_t[r.r] : coarse.proj_n[r.r]));
  }
  for (const auto& r : coarse.unary_n) cun.insert(key3(coarse.proj_n[r.p], coarse.proj_n[r.c], 0));
  for (const auto& r : coarse.unary_t) cun.insert(key3(coarse.proj_n[r.p], coarse.proj_t[r.c], 0));

  // Fine layer: each refined symbol projects to its coarse-level ancestor.
  auto to_coarse = [&](const std::string& name) {
    auto it = up.find(name);
    if (it == up.end()) fail("symbol " + name + " of " + fine.prefix + " is not in " + hierarchy_path);
    auto jt = cid.find(it->second);
    if (jt == cid.end()) fail("coarse-level symbol " + it->second + " is not a symbol of " + coarse.prefix);
    return jt->second;
  };
  std::vector<int32_t> per_coarse(names.size(), 0);
  for (size_t i = 0; i < fine.nonterms.size(); ++i) {
    fine.proj_n[i] = fine.nonterms[i].coarse = to_coarse(fine.nonterms[i].name);
    per_coarse[fine.proj_n[i]]++;
  }
  for (size_t i = 0; i < fine.terms.size(); ++i) {
    fine.proj_t[i] = fine.terms[i].coarse = to_coarse(fine.terms[i].name);
    per_coarse[fine.proj_t[i]]++;
  }
  fine.coarse_names = names;
  fine.coarse_substates = per_coarse;
  const float w12 = fine.semiring == kReal ? (float)std::exp(-12.0) : -12.0f;
  const float q12 = java_format_f6(w12);
  for (int f = 0; f < 4; ++f) {
    const bool lt = (f == kTN || f == kTT), rt = (f == kNT || f == kTT);
    for (auto& r : fine.binaries[f])
      if (!cbin.count(key3(fine.proj_n[r.p], lt ? fine.proj_t[r.l] : fine.proj_n[r.l],
                           rt ? fine.proj_t[r.r] : fine.proj_n[r.r]))) {
        r.w = w12;
        r.q = q12;
      }
  }
  for (auto& r : fine.unary_n)
    if (!cun.count(key3(fine.proj_n[r.p], fine.proj_n[r.c], 0))) {
      r.w = w12;
      r.q = q12;
    }
  for (auto& r : fine.unary_t)
    if (!cun.count(key3(fine.proj_n[r.p], fine.proj_t[r.c], 0))) {
      r.w = w12;
      r.q = q12;
    }
  build_tables(coarse);
  build_tables(fine);
}

}  // namespace puck
