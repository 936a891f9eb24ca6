// grammar.grz container: zip (stored entries) of binary grammar tables. See grz.h.
#include "grz.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <type_traits>

namespace puck {

namespace {

constexpr uint32_t kMagic = 0x5a524750;  // "PGRZ"
constexpr uint32_t kVersion = 2;

uint32_t crc32(const std::string& s) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  uint32_t c = 0xFFFFFFFFu;
  for (unsigned char ch : s) c = table[(c ^ ch) & 0xFF] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

// Little-endian field writer/reader (the hosts this runs on are little-endian x86-64).
struct Out {
  std::string b;
  template <class T> void pod(const T& v) {
    static_assert(std::is_trivially_copyable<T>::value, "pod");
    b.append(reinterpret_cast<const char*>(&v), sizeof(T));
  }
  template <class T> void vec(const std::vector<T>& v) {
    pod<uint64_t>(v.size());
    if (!v.empty()) b.append(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
  }
  void str(const std::string& s) {
    pod<uint64_t>(s.size());
    b.append(s);
  }
};

struct In {
  const std::string& b;
  size_t o = 0;
  explicit In(const std::string& s) : b(s) {}
  void need(size_t n) {
    if (o + n > b.size()) throw std::runtime_error("bad grz payload: truncated");
  }
  template <class T> T pod() {
    need(sizeof(T));
    T v;
    std::memcpy(&v, b.data() + o, sizeof(T));
    o += sizeof(T);
    return v;
  }
  template <class T> std::vector<T> vec() {
    const uint64_t n = pod<uint64_t>();
    if (n > (b.size() - o) / std::max<size_t>(1, sizeof(T))) throw std::runtime_error("bad grz payload: length");
    std::vector<T> v(n);
    if (n) std::memcpy(v.data(), b.data() + o, n * sizeof(T));
    o += n * sizeof(T);
    return v;
  }
  std::string str() {
    const uint64_t n = pod<uint64_t>();
    need(n);
    std::string s = b.substr(o, n);
    o += n;
    return s;
  }
};

void put_symbols(Out& w, const std::vector<Symbol>& syms) {
  w.pod<uint64_t>(syms.size());
  for (const Symbol& s : syms) {
    w.str(s.name);
    w.pod<int32_t>(s.coarse);
    w.pod<int32_t>(s.sub);
    w.pod<int32_t>(s.terminal ? 1 : 0);
    w.pod<int32_t>(s.gpu);
  }
}

std::vector<Symbol> get_symbols(In& r) {
  const uint64_t n = r.pod<uint64_t>();
  std::vector<Symbol> v;
  for (uint64_t i = 0; i < n; ++i) {
    Symbol s;
    s.name = r.str();
    s.coarse = r.pod<int32_t>();
    s.sub = r.pod<int32_t>();
    s.terminal = r.pod<int32_t>() != 0;
    s.gpu = r.pod<int32_t>();
    v.push_back(std::move(s));
  }
  return v;
}

void put_blocks(Out& w, const BlockPartition& p) {
  w.pod<int32_t>(p.num_blocks);
  w.vec(p.block_of_coarse);
  w.vec(p.masks);
}

BlockPartition get_blocks(In& r) {
  BlockPartition p;
  p.num_blocks = r.pod<int32_t>();
  p.block_of_coarse = r.vec<int32_t>();
  p.masks = r.vec<uint32_t>();
  return p;
}

// parser-i/grammar: header, coarse symbols, refined symbols and projections, rule families, lexicon.
std::string grammar_payload(const Grammar& g) {
  Out w;
  w.pod(kMagic);
  w.pod(kVersion);
  w.str(g.prefix);
  w.pod<uint64_t>(g.coarse_names.size());
  for (const auto& s : g.coarse_names) w.str(s);
  w.vec(g.coarse_substates);
  put_symbols(w, g.nonterms);
  put_symbols(w, g.terms);
  w.vec(g.proj_n);
  w.vec(g.proj_t);
  for (int32_t v : {g.root, g.num_syms, g.cell_size, g.mask_words, g.num_rules_read, g.num_rules_dropped}) w.pod(v);
  for (int f = 0; f < 4; ++f) w.vec(g.binaries[f]);
  w.vec(g.unary_n);
  w.vec(g.unary_t);
  w.pod<uint64_t>(g.words.size());
  for (const auto& s : g.words) w.str(s);
  w.vec(g.lex_offsets);
  w.vec(g.lex_tags);
  w.vec(g.lex_logp);
  return w.b;
}

// parser-i/structure: CoarseParent partitions (inside, outside-left, outside-right per family).
std::string structure_payload(const Grammar& g) {
  Out w;
  w.pod(kMagic);
  w.pod(kVersion);
  for (int f = 0; f < 4; ++f) put_blocks(w, g.inside_blocks[f]);
  for (int f = 0; f < 4; ++f) put_blocks(w, g.outside_l_blocks[f]);
  for (int f = 0; f < 4; ++f) put_blocks(w, g.outside_r_blocks[f]);
  return w.b;
}

// Raw little-endian array payload (the Viterbi table entries and the rule-block lists).
template <class T>
std::string raw(const std::vector<T>& v) {
  return std::string(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}

template <class T>
std::vector<T> unraw(const std::string& b, const std::string& entry) {
  if (b.size() % sizeof(T)) throw std::runtime_error("bad grz payload: " + entry + " size");
  std::vector<T> v(b.size() / sizeof(T));
  if (!v.empty()) std::memcpy(v.data(), b.data(), b.size());
  return v;
}

// What a kernel entry holds here: the device code lives in libpuck.so, so `<name>.name` is the
// kernel that replaces the reference's and `<name>.program` names the code object.
void kernel_entry(std::vector<std::pair<std::string, std::string>>& e, const std::string& name,
                  const std::string& kernel) {
  e.push_back({name + ".program", "gfx950 code object in libpuck.so"});
  e.push_back({name + ".name", kernel});
}

// A rule-kernel bundle (CLBinaryRuleUpdater.write / CLUnaryRuleUpdater.write,
// CLBinaryRuleUpdater.scala:60-67, CLUnaryRuleUpdater.scala:34-39): numKernels = the CoarseParent
// blocks of the family, then per block its coarse-symbol mask (bits), the indices of its rules in
// the family's rule table (rules), the workgroup shape and the kernel; binary bundles also carry
// the work-list kernel (enqueuer: the mask plan), extra and directWrite.
void bundle_entries(std::vector<std::pair<std::string, std::string>>& e, const std::string& name,
                    const BlockPartition* bp, const std::vector<int32_t>& rule_coarse, int32_t mask_words) {
  const int32_t nb = bp ? std::max<int32_t>(bp->num_blocks, 1) : 1;
  e.push_back({name + "/numKernels", std::to_string(nb)});
  for (int32_t j = 0; j < nb; ++j) {
    const std::string d = name + "/" + std::to_string(j);
    std::vector<int32_t> rules;
    for (size_t r = 0; r < rule_coarse.size(); ++r)
      if (!bp || bp->num_blocks == 0 || bp->block_of_coarse[rule_coarse[r]] == j) rules.push_back((int32_t)r);
    std::vector<uint32_t> bits(mask_words, 0u);
    if (bp && bp->num_blocks > 0)
      for (int32_t w = 0; w < mask_words; ++w) bits[w] = bp->masks[(size_t)j * mask_words + w];
    else
      for (int32_t w = 0; w < mask_words; ++w) bits[w] = ~0u;
    kernel_entry(e, d + "/0", "k_rules_q");
    e.push_back({d + "/entries", d + "/0"});
    e.push_back({d + "/wgSize", "1024"});
    e.push_back({d + "/bits", raw(bits)});
    e.push_back({d + "/rules", raw(rules)});
  }
  if (bp) {
    kernel_entry(e, name + "/enqueuer/enqueue", "k_plan");
    e.push_back({name + "/enqueuer/ints", raw(std::vector<int32_t>{mask_words})});
    e.push_back({name + "/extra", ""});
    e.push_back({name + "/directWrite", "false"});
  }
}

void check_header(In& r, const std::string& entry) {
  if (r.pod<uint32_t>() != kMagic) throw std::runtime_error("bad grz payload: " + entry + " magic");
  if (r.pod<uint32_t>() != kVersion) throw std::runtime_error("bad grz payload: " + entry + " version");
}

void read_grammar(const std::string& b, Grammar& g) {
  In r(b);
  check_header(r, "grammar");
  g.prefix = r.str();
  const uint64_t nc = r.pod<uint64_t>();
  for (uint64_t i = 0; i < nc; ++i) g.coarse_names.push_back(r.str());
  g.coarse_substates = r.vec<int32_t>();
  g.nonterms = get_symbols(r);
  g.terms = get_symbols(r);
  g.proj_n = r.vec<int32_t>();
  g.proj_t = r.vec<int32_t>();
  g.root = r.pod<int32_t>();
  g.num_syms = r.pod<int32_t>();
  g.cell_size = r.pod<int32_t>();
  g.mask_words = r.pod<int32_t>();
  g.num_rules_read = r.pod<int32_t>();
  g.num_rules_dropped = r.pod<int32_t>();
  for (int f = 0; f < 4; ++f) g.binaries[f] = r.vec<BinaryRule>();
  g.unary_n = r.vec<UnaryRule>();
  g.unary_t = r.vec<UnaryRule>();
  const uint64_t nw = r.pod<uint64_t>();
  for (uint64_t i = 0; i < nw; ++i) {
    g.words.push_back(r.str());
    g.word_ids[g.words.back()] = (int32_t)i;
  }
  g.lex_offsets = r.vec<int32_t>();
  g.lex_tags = r.vec<int32_t>();
  g.lex_logp = r.vec<double>();
  if (g.lex_offsets.size() != g.words.size() + 1) throw std::runtime_error("bad grz payload: lexicon offsets");
  if (g.proj_n.size() != g.nonterms.size() || g.proj_t.size() != g.terms.size())
    throw std::runtime_error("bad grz payload: projections");
}

void read_structure(const std::string& b, Grammar& g) {
  In r(b);
  check_header(r, "structure");
  for (int f = 0; f < 4; ++f) g.inside_blocks[f] = get_blocks(r);
  for (int f = 0; f < 4; ++f) g.outside_l_blocks[f] = get_blocks(r);
  for (int f = 0; f < 4; ++f) g.outside_r_blocks[f] = get_blocks(r);
}

// ---- minimal zip (stored entries) ----
void le16(std::string& s, uint16_t v) { s.append(reinterpret_cast<const char*>(&v), 2); }
void le32(std::string& s, uint32_t v) { s.append(reinterpret_cast<const char*>(&v), 4); }
uint16_t rd16(const std::string& s, size_t o) { uint16_t v; std::memcpy(&v, s.data() + o, 2); return v; }
uint32_t rd32(const std::string& s, size_t o) { uint32_t v; std::memcpy(&v, s.data() + o, 4); return v; }

void write_zip(const std::string& path, const std::vector<std::pair<std::string, std::string>>& entries) {
  std::string out, central;
  for (const auto& [name, data] : entries) {
    if (data.size() >= 0xFFFFFFFFull || out.size() >= 0xFFFFFFFFull)
      throw std::runtime_error("grz entry exceeds the 4 GiB zip limit");
    const uint32_t crc = crc32(data);
    const uint32_t off = (uint32_t)out.size();
    le32(out, 0x04034b50);
    le16(out, 20);  // version needed
    le16(out, 0);   // flags
    le16(out, 0);   // method: stored
    le16(out, 0);   // time
    le16(out, 0x21);  // date 1980-01-01
    le32(out, crc);
    le32(out, (uint32_t)data.size());
    le32(out, (uint32_t)data.size());
    le16(out, (uint16_t)name.size());
    le16(out, 0);
    out += name;
    out += data;
    le32(central, 0x02014b50);
    le16(central, 20);
    le16(central, 20);
    le16(central, 0);
    le16(central, 0);
    le16(central, 0);
    le16(central, 0x21);
    le32(central, crc);
    le32(central, (uint32_t)data.size());
    le32(central, (uint32_t)data.size());
    le16(central, (uint16_t)name.size());
    le16(central, 0);
    le16(central, 0);
    le16(central, 0);
    le16(central, 0);
    le32(central, 0);
    le32(central, off);
    central += name;
  }
  const uint32_t cd_off = (uint32_t)out.size();
  out += central;
  le32(out, 0x06054b50);
  le16(out, 0);
  le16(out, 0);
  le16(out, (uint16_t)entries.size());
  le16(out, (uint16_t)entries.size());
  le32(out, (uint32_t)central.size());
  le32(out, cd_off);
  le16(out, 0);
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path + " for writing");
  f.write(out.data(), (std::streamsize)out.size());
  if (!f) throw std::runtime_error("cannot write " + path);
}

std::map<std::string, std::string> read_zip(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  if (s.size() < 22) throw std::runtime_error("bad grz: not a zip file");
  size_t eocd = std::string::npos;
  for (size_t i = s.size() - 22 + 1; i-- > 0 && s.size() - i <= 22 + 65535;)
    if (rd32(s, i) == 0x06054b50) {
      eocd = i;
      break;
    }
  if (eocd == std::string::npos) throw std::runtime_error("bad grz: no zip directory");
  const uint16_t n = rd16(s, eocd + 10);
  size_t p = rd32(s, eocd + 16);
  std::map<std::string, std::string> out;
  for (uint16_t k = 0; k < n; ++k) {
    if (p + 46 > s.size() || rd32(s, p) != 0x02014b50) throw std::runtime_error("bad grz: directory entry");
    const uint16_t method = rd16(s, p + 10);
    const uint32_t crc = rd32(s, p + 16), csize = rd32(s, p + 20), usize = rd32(s, p + 24);
    const uint16_t nlen = rd16(s, p + 28), xlen = rd16(s, p + 30), clen = rd16(s, p + 32);
    const uint32_t loc = rd32(s, p + 42);
    if (p + 46 + nlen > s.size()) throw std::runtime_error("bad grz: directory entry name");
    const std::string name = s.substr(p + 46, nlen);
    if (method != 0 || csize != usize) throw std::runtime_error("bad grz: entry " + name + " is compressed");
    if (loc + 30 > s.size() || rd32(s, loc) != 0x04034b50) throw std::runtime_error("bad grz: local header");
    const size_t data = loc + 30 + rd16(s, loc + 26) + rd16(s, loc + 28);
    if (data + usize > s.size()) throw std::runtime_error("bad grz: entry " + name + " truncated");
    std::string body = s.substr(data, usize);
    if (crc32(body) != crc) throw std::runtime_error("bad grz: CRC mismatch in " + name);
    out[name] = std::move(body);
    p += 46 + nlen + xlen + clen;
  }
  return out;
}

}  // namespace

void save_grz(const std::string& path, const std::vector<const Grammar*>& layers) {
  if (layers.empty()) throw std::runtime_error("no grammar layers to save");
  std::vector<std::pair<std::string, std::string>> entries;
  std::string names;
  for (size_t i = 0; i < layers.size(); ++i) names += (i ? "\n" : "") + std::string("parser-") + std::to_string(i);
  entries.push_back({"parserNames", names});
  for (size_t i = 0; i < layers.size(); ++i) {
    const Grammar& g = *layers[i];
    const std::string d = "parser-" + std::to_string(i) + "/";
    entries.push_back({d + "grammar", grammar_payload(g)});
    entries.push_back({d + "structure", structure_payload(g)});
    entries.push_back({d + "semiring", g.semiring == kViterbi ? "Viterbi" : (g.semiring == kReal ? "Real" : "LogSum")});
    // Rule-kernel bundles (CLInside.scala:16-23, CLOutside.scala:27-38): inside blocks key on the
    // parent's coarse symbol, outside-left on the left child's, outside-right on the right child's.
    static const char* kFam[4] = {"NN", "NT", "TN", "TT"};
    auto coarse_of = [&](int f, int role) {  // role 0 parent, 1 left, 2 right
      std::vector<int32_t> c;
      const bool lt = f == kTN || f == kTT, rt = f == kNT || f == kTT;
      for (const BinaryRule& r : g.binaries[f]) {
        if (role == 0) c.push_back(g.proj_n[r.p]);
        else if (role == 1) c.push_back(lt ? g.proj_t[r.l] : g.proj_n[r.l]);
        else c.push_back(rt ? g.proj_t[r.r] : g.proj_n[r.r]);
      }
      return c;
    };
    for (int f : {kNT, kNN, kTN, kTT})
      bundle_entries(entries, d + "inside" + kFam[f], &g.inside_blocks[f], coarse_of(f, 0), g.mask_words);
    std::vector<int32_t> un(g.unary_n.size(), 0), ut(g.unary_t.size(), 0);
    bundle_entries(entries, d + "insideNU", nullptr, un, g.mask_words);
    bundle_entries(entries, d + "insideTU", nullptr, ut, g.mask_words);
    for (int f : {kNN, kNT, kTN, kTT}) {
      bundle_entries(entries, d + "outside_L_" + kFam[f], &g.outside_l_blocks[f], coarse_of(f, 1), g.mask_words);
      bundle_entries(entries, d + "outside_R_" + kFam[f], &g.outside_r_blocks[f], coarse_of(f, 2), g.mask_words);
    }
    bundle_entries(entries, d + "outsideNU", nullptr, un, g.mask_words);
    bundle_entries(entries, d + "outsideTU", nullptr, ut, g.mask_words);
    // CLParserUtils.write (CLParserUtils.scala:14-16): root scores.
    kernel_entry(entries, d + "setRootScoresKernel", "k_reduce");
    kernel_entry(entries, d + "getRootScoresKernel", "k_get_root");
    // CLMaskKernels.write (CLMaskKernels.scala:23-25): maskSize, blocks per sentence, block size.
    kernel_entry(entries, d + "computeMasksKernel", "k_masks");
    entries.push_back({d + "MasksInts", raw(std::vector<int32_t>{g.mask_words, 1, 64})});
    // CLMBRKernels.write (CLMBRKernels.scala:20-22).
    kernel_entry(entries, d + "computeMBRKernel", "k_mbr");
    entries.push_back({d + "MBRInts", raw(std::vector<int32_t>{g.mask_words})});
    // CLViterbi.write (CLViterbi.scala:73-80): decoder tables, the six group offsets NN, NT, TN,
    // TT, U, U_T into ViterbiParents.
    kernel_entry(entries, d + "computeViterbiKernel", "k_viterbi");
    entries.push_back({d + "ViterbiWGSize", raw(std::vector<int32_t>{1, 256})});
    entries.push_back({d + "ViterbiParents", raw(g.dec_offsets)});
    entries.push_back({d + "ViterbiLeft", raw(g.dec_left)});
    entries.push_back({d + "ViterbiRight", raw(g.dec_right)});
    entries.push_back({d + "ViterbiScores", raw(g.dec_scores)});
    const int32_t nN = (int32_t)g.nonterms.size();
    entries.push_back({d + "ViterbiOffsets", raw(std::vector<int32_t>{0, nN, 2 * nN, 3 * nN, 4 * nN, 5 * nN})});
    // CLScalingKernels.write (CLScalingKernels.scala:19-20).
    kernel_entry(entries, d + "getScalingConstantsKernel", "k_scaling");
  }
  write_zip(path, entries);
}

std::vector<std::shared_ptr<Grammar>> load_grz(const std::string& path) {
  auto z = read_zip(path);
  auto get = [&](const std::string& name) -> const std::string& {
    auto it = z.find(name);
    if (it == z.end()) throw std::runtime_error("bad grz: missing entry " + name);
    return it->second;
  };
  // CLParserData.readSequence (CLParser.scala:1302-1305): each trimmed line of parserNames names
  // a layer's entry directory.
  std::istringstream names(get("parserNames"));
  std::vector<std::shared_ptr<Grammar>> out;
  std::string line;
  while (std::getline(names, line)) {
    const size_t a = line.find_first_not_of(" \t\r"), b = line.find_last_not_of(" \t\r");
    if (a == std::string::npos) continue;
    const std::string d = line.substr(a, b - a + 1) + "/";
    auto g = std::make_shared<Grammar>();
    read_grammar(get(d + "grammar"), *g);
    read_structure(get(d + "structure"), *g);
    g->dec_offsets = unraw<int32_t>(get(d + "ViterbiParents"), "ViterbiParents");
    g->dec_left = unraw<int32_t>(get(d + "ViterbiLeft"), "ViterbiLeft");
    g->dec_right = unraw<int32_t>(get(d + "ViterbiRight"), "ViterbiRight");
    g->dec_scores = unraw<float>(get(d + "ViterbiScores"), "ViterbiScores");
    if (g->dec_offsets.size() != 6 * g->nonterms.size() + 1 || g->dec_scores.size() != g->dec_left.size())
      throw std::runtime_error("bad grz payload: Viterbi tables");
    const std::string& sr = get(d + "semiring");
    if (sr == "Viterbi") g->semiring = kViterbi;
    else if (sr == "Real") g->semiring = kReal;
    else if (sr == "LogSum") g->semiring = kLogSum;
    else throw std::runtime_error("bad grz: semiring " + sr);
    out.push_back(g);
  }
  if (out.empty()) throw std::runtime_error("bad grz: no parser layers");
  return out;
}

}  // namespace puck
