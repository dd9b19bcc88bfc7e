// text.hpp -- host parsing of PhyloCSF parameter files (Newick trees, ECMs).
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace pcsf {

// A parse failure (the reference raises Failure / Invalid_argument).
struct ParseError {
  int code;  // PCSF_EINVAL or PCSF_EFAIL
  std::string msg;
};

// Newick.t (lib/CamlPaml/Newick.mli): children, label, optional branch.
struct NewickNode {
  std::vector<std::unique_ptr<NewickNode>> kids;
  std::string label;
  bool has_branch = false;
  double branch = 0.0;
};

// NewickLexer.mll + NewickParser.mly.  Throws ParseError.
std::unique_ptr<NewickNode> parse_newick(const std::string& text);

// Newick.subtree (lib/CamlPaml/Newick.ml:34-41); returns nullptr for None.
std::unique_ptr<NewickNode> newick_subtree(std::unique_ptr<NewickNode> t,
                                           const std::vector<std::string>& keep);

// Newick.total_length ~count_root:false (lib/CamlPaml/Newick.ml:52-61).
double newick_total_length(const NewickNode& t);
int newick_leaves(const NewickNode& t);

// T.t (lib/CamlPaml/T.ml): array tree in the reference's numbering.
struct Tree {
  int n = 0, nl = 0;
  std::vector<int> parent;  // parent[root] = -1
  std::vector<int> lc, rc;  // indexed by node; -1 for leaves
  std::vector<double> branch;
  std::vector<std::string> labels;
  int root() const { return n - 1; }
};

// T.of_newick (lib/CamlPaml/T.ml:57-112).  Throws ParseError.
Tree tree_of_newick(const NewickNode& t);

// ECM.import_parameters (src/ECM.ml:18-72).  Throws ParseError.
void read_ecm(const std::string& path, double* S /*64*64*/, double* pi /*64*/);

}  // namespace pcsf
