// text.cpp -- host parsing of PhyloCSF parameter files.
//
//   Newick tokens ........ lib/CamlPaml/NewickLexer.mll:5-14
//   Newick grammar ....... lib/CamlPaml/NewickParser.mly:7-24
//   subtree / length ..... lib/CamlPaml/Newick.ml:34-61
//   T numbering .......... lib/CamlPaml/T.ml:57-112
//   ECM layout ........... src/ECM.ml:18-72
#include "text.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <sstream>

#include "../../../include/phylocsf_hip.h"

namespace pcsf {
namespace {

// OCaml float_of_string: strtod syntax, the whole token consumed.
bool to_double(const std::string& tok, double* out) {
  if (tok.empty()) return false;
  char* end = nullptr;
  double v = std::strtod(tok.c_str(), &end);
  if (end != tok.c_str() + tok.size()) return false;
  *out = v;
  return true;
}

enum Tok { LP, RP, COMMA, COLON, BRANCHLEN, LABEL, END };

struct Lexer {
  const std::string& s;
  size_t pos = 0;
  Tok tok = END;
  std::string text;
  double value = 0;

  explicit Lexer(const std::string& src) : s(src) {}

  static bool label_char(char c) {
    return std::isalnum(static_cast<unsigned char>(c)) || c == '_' || c == '.';
  }
  static bool num_char(char c) { return std::isdigit(static_cast<unsigned char>(c)) || c == '.'; }

  void next() {
    // ' ', '\t', '\r', '\n' and ';' are skipped by the first lexer rule
    while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t' || s[pos] == '\r' ||
                              s[pos] == '\n' || s[pos] == ';'))
      ++pos;
    if (pos >= s.size()) {
      tok = END;
      return;
    }
    char c = s[pos];
    switch (c) {
      case '(': tok = LP; ++pos; return;
      case ')': tok = RP; ++pos; return;
      case ',': tok = COMMA; ++pos; return;
      case ':': tok = COLON; ++pos; return;
      default: break;
    }
    if (!label_char(c)) throw ParseError{PCSF_EFAIL, "Newick: unexpected character"};
    size_t run = pos, numrun = pos;
    while (run < s.size() && label_char(s[run])) ++run;
    while (numrun < s.size() && num_char(s[numrun])) ++numrun;
    // longest match; on equal length the BRANCHLEN rule (listed first) wins
    if (numrun == run) {
      text = s.substr(pos, run - pos);
      if (!to_double(text, &value)) throw ParseError{PCSF_EFAIL, "float_of_string " + text};
      tok = BRANCHLEN;
    } else {
      text = s.substr(pos, run - pos);
      tok = LABEL;
    }
    pos = run;
  }
};

void parse_label(Lexer& L, NewickNode* x) {
  if (L.tok == LABEL) {
    x->label = L.text;
    L.next();
    if (L.tok == COLON) {
      L.next();
      if (L.tok != BRANCHLEN) throw ParseError{PCSF_EFAIL, "Newick: Parse_error"};
      x->has_branch = true;
      x->branch = L.value;
      L.next();
    }
    return;
  }
  if (L.tok == COLON) {
    L.next();
    if (L.tok != BRANCHLEN) throw ParseError{PCSF_EFAIL, "Newick: Parse_error"};
    x->has_branch = true;
    x->branch = L.value;
    L.next();
    return;
  }
  throw ParseError{PCSF_EFAIL, "Newick: Parse_error"};
}

std::unique_ptr<NewickNode> parse_node(Lexer& L) {
  auto x = std::make_unique<NewickNode>();
  if (L.tok == LP) {
    L.next();
    for (;;) {
      x->kids.push_back(parse_node(L));
      if (L.tok == COMMA) {
        L.next();
        continue;
      }
      break;
    }
    if (L.tok != RP) throw ParseError{PCSF_EFAIL, "Newick: Parse_error"};
    L.next();
    if (L.tok == LABEL || L.tok == COLON) parse_label(L, x.get());
    return x;
  }
  parse_label(L, x.get());
  return x;
}

}  // namespace

std::unique_ptr<NewickNode> parse_newick(const std::string& text) {
  Lexer L(text);
  L.next();
  auto t = parse_node(L);
  if (L.tok != END) throw ParseError{PCSF_EFAIL, "Newick: Parse_error"};
  return t;
}

std::unique_ptr<NewickNode> newick_subtree(std::unique_ptr<NewickNode> t,
                                           const std::vector<std::string>& keep) {
  auto kept = [&](const std::string& l) {
    if (l.empty()) return true;
    for (auto& k : keep)
      if (k == l) return true;
    return false;
  };
  if (!kept(t->label)) return nullptr;
  if (t->kids.empty()) return t;
  std::vector<std::unique_ptr<NewickNode>> out;
  for (auto& k : t->kids) {
    auto s = newick_subtree(std::move(k), keep);
    if (s) out.push_back(std::move(s));
  }
  if (out.empty()) return nullptr;
  if (out.size() == 1) {
    auto c = std::move(out[0]);
    // Node (sst, snm, maybe_add bl sbl)
    if (t->has_branch && c->has_branch) c->branch = t->branch + c->branch;
    else c->has_branch = false;
    return c;
  }
  t->kids = std::move(out);
  return t;
}

int newick_leaves(const NewickNode& t) {
  if (t.kids.empty()) return 1;
  int s = 0;
  for (auto& k : t.kids) s += newick_leaves(*k);
  return s;
}

double newick_total_length(const NewickNode& t) {
  std::function<double(const NewickNode&, bool)> go = [&](const NewickNode& x, bool root) {
    double bl;
    if (root) bl = 0.0;
    else if (x.has_branch) bl = x.branch;
    else throw ParseError{PCSF_EINVAL, "CamlPaml.Newick.total_length: unspecified branch length"};
    double s = 0.0;
    for (auto& k : x.kids) s += go(*k, false);
    return bl + s;
  };
  return go(t, true);
}

Tree tree_of_newick(const NewickNode& t) {
  const char* bad = "CamlPaml.T.of_newick: input is not a rooted, bifurcating tree";
  std::function<int(const NewickNode&)> size = [&](const NewickNode& x) {
    int s = 1;
    for (auto& k : x.kids) s += size(*k);
    return s;
  };
  Tree T;
  T.n = size(t);
  if (T.n < 3 || T.n % 2 == 0) throw ParseError{PCSF_EINVAL, bad};
  T.nl = (T.n + 1) / 2;
  T.parent.assign(T.n, -1);
  T.lc.assign(T.n, -1);
  T.rc.assign(T.n, -1);
  T.branch.assign(T.n, NAN);
  T.labels.assign(T.n, "");
  int next_leaf = 0, next_internal = T.nl;
  std::function<int(const NewickNode&)> fill = [&](const NewickNode& x) -> int {
    if (x.kids.empty()) {
      if (next_leaf >= T.nl) throw ParseError{PCSF_EINVAL, bad};
      int i = next_leaf++;
      T.labels[i] = x.label;
      T.branch[i] = x.has_branch ? x.branch : NAN;
      return i;
    }
    if (x.kids.size() != 2) throw ParseError{PCSF_EINVAL, bad};
    int l = fill(*x.kids[0]);
    int r = fill(*x.kids[1]);
    int i = next_internal++;
    T.labels[i] = x.label;
    T.branch[i] = x.has_branch ? x.branch : NAN;
    T.parent[l] = i;
    T.parent[r] = i;
    T.lc[i] = l;
    T.rc[i] = r;
    return i;
  };
  int root = fill(t);
  if (root != T.n - 1 || next_leaf != T.nl) throw ParseError{PCSF_EINVAL, bad};
  return T;
}

namespace {
std::vector<std::string> space_fields(const std::string& line) {
  // Str.split on single spaces, each piece String.trim'ed, empties dropped
  std::vector<std::string> out;
  size_t i = 0;
  while (i <= line.size()) {
    size_t j = line.find(' ', i);
    if (j == std::string::npos) j = line.size();
    size_t a = i, b = j;
    while (a < b && std::isspace(static_cast<unsigned char>(line[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(line[b - 1]))) --b;
    if (b > a) out.push_back(line.substr(a, b - a));
    i = j + 1;
  }
  return out;
}

int nuc_index(char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return -1;
  }
}
}  // namespace

void read_ecm(const std::string& path, double* S, double* pi) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw ParseError{PCSF_EFAIL, "cannot open " + path};
  std::vector<std::string> lines;
  std::string line;
  while (std::getline(f, line)) lines.push_back(line);
  if (lines.size() < 71) throw ParseError{PCSF_EINVAL, "ECM.import_parameters: short file"};
  // lines 1..63: strict lower triangle, line i holds i entries
  std::vector<std::vector<double>> raw(63);
  for (int r = 0; r < 63; ++r) {
    for (auto& tok : space_fields(lines[r])) {
      double v;
      if (!to_double(tok, &v)) throw ParseError{PCSF_EFAIL, "float_of_string " + tok};
      raw[r].push_back(v);
    }
    if (static_cast<int>(raw[r].size()) < r + 1)
      throw ParseError{PCSF_EINVAL, "ECM.import_parameters: index out of bounds"};
  }
  for (int i = 0; i < 64; ++i)
    for (int j = 0; j < 64; ++j)
      S[i * 64 + j] = (i > j) ? raw[i - 1][j] : (i < j) ? raw[j - 1][i] : 0.0;
  if (!space_fields(lines[63]).empty()) throw ParseError{PCSF_EFAIL, "ECM.import_parameters"};
  auto fr = space_fields(lines[64]);
  if (fr.size() != 64) throw ParseError{PCSF_EINVAL, "ECM.import_parameters: frequencies"};
  for (int j = 0; j < 64; ++j)
    if (!to_double(fr[j], &pi[j])) throw ParseError{PCSF_EFAIL, "float_of_string " + fr[j]};
  int idx = 0;
  for (int li = 67; li <= 70; ++li)
    for (auto& c : space_fields(lines[li])) {
      if (c.size() != 3) throw ParseError{PCSF_EFAIL, "ECM.import_parameters: codon"};
      int a = nuc_index(c[0]), b = nuc_index(c[1]), d = nuc_index(c[2]);
      if (a < 0 || b < 0 || d < 0) throw ParseError{PCSF_EINVAL, "unrecognized nucleotide"};
      if (16 * a + 4 * b + d != idx)
        throw ParseError{PCSF_EFAIL, "ECM.import_parameters: incorrect codon order"};
      ++idx;
    }
  if (idx != 64) throw ParseError{PCSF_EFAIL, "ECM.import_parameters: incorrect codon order"};
}

}  // namespace pcsf
