// Grammar compilation: Berkeley split-merge text grammars -> rule tables for the CKY engine.
//
// Mirrors the host half of CLParserData.make (CLParser.scala:1264-1285), RuleStructure
// (RuleStructure.scala:15-128), the score-literal formatting of the rule kernels
// (SimpleGenRuleMultiply.java:151,304), the CoarseParent rule partition
// (CoarseParentSymbolGrammarClusterer.java:38-121) and the rule arrays of the Viterbi decoder
// (CLViterbi.scala:274-342,610-628).
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

namespace puck {

enum Semiring : int32_t { kViterbi = 0, kReal = 1, kLogSum = 2 };

// Binary rule over GPU symbol ids. For inside families p is a nonterminal id; l and r are
// nonterminal or terminal ids according to the family (NN, NT, TN, TT).
struct BinaryRule {
  int32_t p, l, r;
  float w;   // semiring score (Viterbi/LogSum: ln p, Real: p) as fp32 -- the decoder's value
  float q;   // w after Java "%f" formatting and fp32 re-parse -- the rule kernels' value
  int32_t id;
};

struct UnaryRule {
  int32_t p, c;
  float w, q;
  int32_t id;
};

enum BinaryFamily : int32_t { kNN = 0, kNT = 1, kTN = 2, kTT = 3 };

// CoarseParent partition of one rule family: coarse parents with >= 300 refined rules each form
// a block of their own, all others share one block. block_of_coarse maps a coarse parent id to its
// block (-1 if it has no rule in the family).
struct BlockPartition {
  int32_t num_blocks = 0;
  std::vector<int32_t> block_of_coarse;
  std::vector<uint32_t> masks;  // num_blocks x mask_words, bit c set iff coarse c in block
};

struct Symbol {
  std::string name;  // refined name, e.g. "NP_3"
  int32_t coarse;    // index into coarse_names (.numstates line order)
  int32_t sub;       // substate
  bool terminal;
  int32_t gpu;       // index within the nonterminal or terminal space
};

struct Grammar {
  std::string prefix;
  Semiring semiring = kViterbi;

  std::vector<std::string> coarse_names;
  std::vector<int32_t> coarse_substates;

  std::vector<Symbol> nonterms;   // by gpu id
  std::vector<Symbol> terms;      // by gpu id
  std::vector<int32_t> proj_n;    // nonterminal gpu id -> coarse id
  std::vector<int32_t> proj_t;    // terminal gpu id -> coarse id
  int32_t root = -1;              // nonterminal gpu id of ROOT_0
  int32_t num_syms = 0;           // max(#N, #T)  (CLParserData.numSyms)
  int32_t cell_size = 0;          // roundUp(num_syms, 32) (CLParser.scala:185)
  int32_t mask_words = 0;         // roundUp(#coarse, 32) / 32 (CLMaskKernels.scala:77)

  std::vector<BinaryRule> binaries[4];   // NN, NT, TN, TT in rule-index order
  std::vector<UnaryRule> unary_n;        // N parent, N child (identity X->X included)
  std::vector<UnaryRule> unary_t;        // N parent, T child
  int32_t num_rules_read = 0;
  int32_t num_rules_dropped = 0;

  // CoarseParent partitions. inside[f] keys on the parent, outside_l[f] on the left child
  // (rotateLeftToParent), outside_r[f] on the right child (rotateRightToParent).
  BlockPartition inside_blocks[4], outside_l_blocks[4], outside_r_blocks[4];

  // Viterbi decoder tables: 6 groups (NN, NT, TN, TT, U, U_T), each stable-sorted by parent.
  std::vector<int32_t> dec_offsets;  // 6 * #N + 1
  std::vector<float> dec_scores;
  std::vector<int32_t> dec_left, dec_right;  // right only for the 4 binary groups

  // Lexicon: per word, the (terminal gpu id, ln P(w|t)) pairs with P > 0.
  std::vector<std::string> words;
  std::unordered_map<std::string, int32_t> word_ids;
  std::vector<int32_t> lex_offsets;  // #words + 1
  std::vector<int32_t> lex_tags;
  std::vector<double> lex_logp;

  int32_t num_nonterms() const { return (int32_t)nonterms.size(); }
  int32_t num_terms() const { return (int32_t)terms.size(); }
};

// Loads prefix.{numstates,binary,unary,lexicon}. Rules with ln p < log_threshold are dropped
// (the threshold argument of SimpleGrammar.parseBerkeleyText, -12 at every call site).
// Throws std::runtime_error on malformed input.
Grammar load_text_grammar(const std::string& prefix, Semiring semiring, double log_threshold = -12.0);

// Java Formatter "%f" of a float (widened to double, shortest digits, HALF_UP at 6 decimals),
// re-parsed as an fp32 literal.
float java_format_f6(float v);
std::string java_format_f6_string(float v);

BlockPartition coarse_parent_partition(const std::vector<int32_t>& coarse_key_per_rule,
                                       int32_t num_coarse, int32_t mask_words);

// Recomputes mask_words, the CoarseParent partitions and the decoder tables from the rules and the
// symbol projections.
void build_tables(Grammar& g);

// Coarse symbols per grammar supported by the mask and MBR kernels (8 mask words).
constexpr int32_t kMaxCoarseSymbols = 256;

// readHierarchy (CLParser.scala:1182-1208): per coarse-level symbol "SYM_k" of a .hierarchy file,
// the fine-level symbols below it. Levels count split rounds (0 = unsplit).
std::map<std::string, std::vector<std::string>> read_hierarchy(const std::string& path, int32_t coarse_level,
                                                               int32_t fine_level);
// The split level in a grammar's name: its last run of digits ("wsj_6.gr" -> 6), as readHierarchy
// derives both levels from the grammar names.
int32_t grammar_level_from_name(const std::string& name);
// reprojectGrammar (CLParser.scala:1153-1180), applied in place to a coarse/fine pair: the coarse
// layer's refined symbols become the coarse symbols of both layers (identity projection in the
// coarse layer, the hierarchy's ancestor at coarse_level in the fine layer), so masks keep one bit
// per coarse substate; fine rules whose projection is not a coarse rule are scored
// fromLogSpace(-12). Throws std::runtime_error on a symbol the hierarchy or the coarse grammar lacks.
void reproject_grammars(Grammar& coarse, Grammar& fine, const std::string& hierarchy_path, int32_t coarse_level,
                        int32_t fine_level);

}  // namespace puck
