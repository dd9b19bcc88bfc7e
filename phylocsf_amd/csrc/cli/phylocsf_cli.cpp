// phylocsf_cli.cpp -- the PhyloCSF command-line program over the MI355X path.
//
// Same command line, input handling and output as the reference executable
// (src/PhyloCSF.ml): options (:17-50), MFA reader (:83-117),
// --removeRefGaps (:121-131), ORF search and candidate regions (:133-217),
// leaf codes per codon column (:219-246), BLS (:250-262), translation
// (:268-279), per-alignment processing and report format (:280-389),
// parameter-set lookup and the input loop (:407-470).  What differs is how
// regions are scored: instead of one PhyloCSFModel.score per region (forked
// over -p processes), the candidate regions of many alignments are packed
// into one batch and scored by one pcsf_score call on the GPU; output is
// still written alignment by alignment, in input order.
//
// Extra options (no counterpart in the reference):
//   --precision=fp64m|fp64|fp32|fp16s   arithmetic of the pruning (default
//                                 fp64m: FP64 on matrix cores, the reference's
//                                 digits; fp64 the plain FP64 kernel; fp32 and
//                                 split-fp16 agree with them to 1e-3 dB per
//                                 region at fixed scales)
//   --device=N                    GPU index (default 0)
//   --batchBytes=N                leaf-code bytes packed per GPU call (default 1 GiB)
// The omega strategy (src/OmegaModel.ml) runs every candidate region's two
// MAP fits in lock step on the GPU (pcsf_score_omega), with omega_H1 and
// sigma_H1 from the parameter set's optional _omega file (:452-463).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "../../../include/phylocsf_hip.h"
#include "../host/text.hpp"

namespace {

// ---- OCaml-style exceptions (the text Printexc.to_string gives) ------------

std::string ocaml_quote(const std::string& s) {  // sprintf "%S"
  std::string o = "\"";
  for (unsigned char c : s) {
    if (c == '"' || c == '\\') {
      o += '\\';
      o += (char)c;
    } else if (c == '\n') {
      o += "\\n";
    } else if (c == '\t') {
      o += "\\t";
    } else if (c == '\r') {
      o += "\\r";
    } else if (c < 32 || c >= 127) {
      char b[8];
      std::snprintf(b, sizeof b, "\\%03d", c);
      o += b;
    } else {
      o += (char)c;
    }
  }
  return o + "\"";
}

struct OcamlError {
  std::string text;  // Printexc.to_string of the exception
};
[[noreturn]] void failwith(const std::string& m) { throw OcamlError{"Failure(" + ocaml_quote(m) + ")"}; }
[[noreturn]] void invalid_arg(const std::string& m) {
  throw OcamlError{"Invalid_argument(" + ocaml_quote(m) + ")"};
}
[[noreturn]] void sys_error(const std::string& m) { throw OcamlError{"Sys_error(" + ocaml_quote(m) + ")"}; }

// ---- options (src/PhyloCSF.ml:17-50) ---------------------------------------

enum class Strategy { MaxLik, FixedLik, Omega, Nop };
enum class OrfMode { AsIs, ATGStop, StopStop, StopStop3, ToFirstStop, FromLastStop, ToOrFromStop };

struct Options {
  Strategy strategy = Strategy::MaxLik;
  bool files = false, remove_ref_gaps = false, allow_ref_gaps = false;
  bool species_set = false;
  std::string species;
  int frames = 1;
  OrfMode orf = OrfMode::AsIs;
  int min_codons = 25;
  bool all_scores = false;
  int procs = 1;
  bool bls = false, anc_comp = false, dna = false, aa = false, debug = false;
  int precision = PCSF_FP64M;
  int device = 0;
  long long batch_bytes = 1LL << 30;
  std::vector<std::string> positional;
};

const char* kUsage =
    "usage: PhyloCSF parameter_set [file1 file2 ...]\n"
    "input will be read from stdin if no filenames are given.\n\n"
    "options:\n"
    "  --strategy=mle|fixed|omega   evaluation strategy (default mle)\n\n"
    "  input interpretation:\n"
    "    --files                    input list(s) of alignment filenames instead of individual "
    "alignment(s)\n"
    "    --removeRefGaps            automatically remove any alignment columns that are gapped "
    "in the reference sequence\n"
    "    --species=Species1,Species2,...\n"
    "                               hint that only this subset of species will be used\n\n"
    "  searching mulitple reading frames and ORFs:\n"
    "    -f 1|3|6, --frames=1|3|6   how many reading frames to search (default 1)\n"
    "    --orf=AsIs|ATGStop|StopStop|StopStop3|ToFirstStop|FromLastStop|ToOrFromStop\n"
    "                               search for ORFs (default AsIs)\n"
    "    --minCodons=MINCODONS      minimum ORF length for searching over ORFs (default 25 "
    "codons)\n"
    "    --allScores                report scores of all regions evaluated, not just the max\n"
    "    -p P                       accepted for compatibility (regions are batched on the GPU)\n\n"
    "  output control:\n"
    "    --bls                      include alignment branch length score (BLS)\n"
    "    --ancComp                  include ancestral sequence composition score in output\n"
    "    --dna                      include DNA sequence in output\n"
    "    --aa                       include amino acid translation in output\n"
    "    --debug                    print extra information about parameters and errors\n\n"
    "  GPU:\n"
    "    --precision=fp64m|fp64|fp32|fp16s  arithmetic of the pruning kernel (default fp64m)\n"
    "    --device=N                 GPU index (default 0)\n"
    "    --batchBytes=N             leaf-code bytes packed per GPU call (default 1 GiB)\n";

[[noreturn]] void usage_error(const std::string& msg) {
  std::fprintf(stderr, "%s\nPhyloCSF: error: %s\n", kUsage, msg.c_str());
  std::exit(2);
}

std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

int to_int(const std::string& v, const std::string& opt) {
  char* end = nullptr;
  const long x = std::strtol(v.c_str(), &end, 10);
  if (v.empty() || *end) usage_error("option " + opt + ": invalid integer value '" + v + "'");
  return (int)x;
}

Options parse_options(int argc, char** argv) {
  Options o;
  const std::set<std::string> flags = {"files", "removeRefGaps", "allowRefGaps", "allScores",
                                       "bls", "ancComp", "dna", "aa", "debug"};
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--") {
      for (++i; i < argc; ++i) o.positional.push_back(argv[i]);
      break;
    }
    if (a == "-h" || a == "--help") {
      std::fputs(kUsage, stdout);
      std::exit(0);
    }
    std::string name, val;
    bool has_val = false;
    if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
      const size_t eq = a.find('=');
      name = a.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
      if (eq != std::string::npos) {
        val = a.substr(eq + 1);
        has_val = true;
      }
    } else if (a.size() >= 2 && a[0] == '-' && (a[1] == 'f' || a[1] == 'p')) {
      name = (a[1] == 'f') ? "frames" : "p";
      if (a.size() > 2) {
        val = a.substr(2);
        has_val = true;
      }
    } else if (a.size() > 1 && a[0] == '-') {
      usage_error("no such option: " + a);
    } else {
      o.positional.push_back(a);
      continue;
    }
    if (flags.count(name)) {
      if (has_val) usage_error("option --" + name + " takes no value");
      if (name == "files") o.files = true;
      if (name == "removeRefGaps") o.remove_ref_gaps = true;
      if (name == "allowRefGaps") o.allow_ref_gaps = true;
      if (name == "allScores") o.all_scores = true;
      if (name == "bls") o.bls = true;
      if (name == "ancComp") o.anc_comp = true;
      if (name == "dna") o.dna = true;
      if (name == "aa") o.aa = true;
      if (name == "debug") o.debug = true;
      continue;
    }
    static const std::set<std::string> valued = {"strategy", "species", "frames", "orf",
                                                 "minCodons", "p", "precision", "device",
                                                 "batchBytes"};
    if (!valued.count(name)) usage_error("no such option: " + a);
    if (!has_val) {
      if (i + 1 >= argc) usage_error("option " + a + " needs a value");
      val = argv[++i];
    }
    if (name == "strategy") {
      const std::string v = lower(val);
      if (v == "mle") o.strategy = Strategy::MaxLik;
      else if (v == "fixed") o.strategy = Strategy::FixedLik;
      else if (v == "nop") o.strategy = Strategy::Nop;
      else if (v == "omega") o.strategy = Strategy::Omega;
      else usage_error("option --strategy: invalid strategy " + val);
    } else if (name == "species") {
      o.species = val;
      o.species_set = true;
    } else if (name == "frames") {
      if (val == "1" || val == "3" || val == "6") o.frames = val[0] - '0';
      else usage_error("option --frames: invalid reading frame " + val);
    } else if (name == "orf") {
      const std::string v = lower(val);
      if (v == "asis") o.orf = OrfMode::AsIs;
      else if (v == "atgstop") o.orf = OrfMode::ATGStop;
      else if (v == "stopstop") o.orf = OrfMode::StopStop;
      else if (v == "stopstop3") o.orf = OrfMode::StopStop3;
      else if (v == "tofirststop") o.orf = OrfMode::ToFirstStop;
      else if (v == "fromlaststop") o.orf = OrfMode::FromLastStop;
      else if (v == "toorfromstop") o.orf = OrfMode::ToOrFromStop;
      else usage_error("option --orf: invalid ORF search mode " + val);
    } else if (name == "minCodons") {
      o.min_codons = to_int(val, a);
    } else if (name == "p") {
      o.procs = to_int(val, a);
    } else if (name == "precision") {
      const std::string v = lower(val);
      if (v == "fp64") o.precision = PCSF_FP64;
      else if (v == "fp64m") o.precision = PCSF_FP64M;
      else if (v == "fp32") o.precision = PCSF_FP32;
      else if (v == "fp16s") o.precision = PCSF_FP16S;
      else usage_error("option --precision: fp64, fp64m, fp32 or fp16s");
    } else if (name == "device") {
      o.device = to_int(val, a);
    } else if (name == "batchBytes") {
      o.batch_bytes = std::max(1, to_int(val, a));
    }
  }
  return o;
}

// ---- sequences (src/PhyloCSF.ml:83-131,264-279; lib/CamlPaml/Code.ml) -----

bool is_nuc(char c) {  // Code.DNA.is
  switch (c) {
    case 'A': case 'a': case 'C': case 'c': case 'G': case 'g': case 'T': case 't': return true;
    default: return false;
  }
}
int nuc_index(char c) {  // Code.DNA.index
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    default: return 3;
  }
}
char comp(char c) {  // Code.DNA.comp
  switch (c) {
    case 'A': return 'T'; case 'G': return 'C'; case 'C': return 'G'; case 'T': return 'A';
    case 'a': return 't'; case 'g': return 'c'; case 'c': return 'g'; case 't': return 'a';
    case 'N': return 'N'; case 'n': return 'n'; case '-': return '-';
    default: invalid_arg(std::string("unrecognized nucleotide ") + c);
  }
}
std::string revcomp(const std::string& s) {
  std::string r(s.size(), ' ');
  for (size_t i = 0; i < s.size(); ++i) r[s.size() - 1 - i] = comp(s[i]);
  return r;
}
// standard genetic code in codon-index order (Code.Codon64.translation_table)
const char kCode[65] = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF";
std::string translate(const std::string& dna) {
  std::string pp(dna.size() / 3, '?');
  for (size_t i = 0; i < pp.size(); ++i) {
    const char a = dna[3 * i], b = dna[3 * i + 1], c = dna[3 * i + 2];
    if (is_nuc(a) && is_nuc(b) && is_nuc(c))
      pp[i] = kCode[16 * nuc_index(a) + 4 * nuc_index(b) + nuc_index(c)];
  }
  return pp;
}

std::string trim(const std::string& s) {  // String.trim: spaces, tabs, newlines
  const char* ws = " \t\n\r\x0c";
  const size_t a = s.find_first_not_of(ws);
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(ws);
  return s.substr(a, b - a + 1);
}

struct Alignment {
  std::vector<std::string> species, rows;
};

// input_mfa: headers '>name[|...]', sequence lines trimmed and joined.
Alignment input_mfa(const std::vector<std::string>& lines) {
  try {
    Alignment a;
    size_t i = 0;
    bool in_seq = false;
    for (;;) {
      if (!in_seq) {  // hdr
        while (i < lines.size() && lines[i].empty()) ++i;
        if (i >= lines.size()) invalid_arg("Option.get");
        const std::string& line = lines[i++];
        if (line[0] != '>') failwith("bad header");
        const std::string hdr = line.substr(1);
        const size_t bar = hdr.find('|');
        a.species.push_back(trim(bar == std::string::npos ? hdr : hdr.substr(0, bar)));
        a.rows.emplace_back();
        in_seq = true;
        continue;
      }
      if (i >= lines.size()) break;
      const std::string line = trim(lines[i]);
      if (line.empty()) {
        ++i;
        continue;
      }
      if (line[0] != '>') {
        a.rows.back() += line;
        ++i;
        continue;
      }
      in_seq = false;
    }
    const size_t len = a.rows[0].size();
    bool bad = len == 0;
    for (size_t k = 0; k < a.rows.size(); ++k)
      bad = bad || a.species[k].empty() || a.rows[k].size() != len;
    if (bad) failwith("empty species name or sequence, or sequence length mismatch");
    return a;
  } catch (const OcamlError& e) {
    const std::string pre = "Failure(\"";
    if (e.text.compare(0, pre.size(), pre) == 0) {
      // Failure msg -> failwith ("invalid MFA alignment: " ^ msg)
      std::string msg = e.text.substr(pre.size(), e.text.size() - pre.size() - 2);
      failwith("invalid MFA alignment: " + msg);
    }
    failwith("invalid MFA alignment: " + e.text);
  }
}

void remove_ref_gaps(Alignment& a) {
  const std::string ref = a.rows[0];
  for (auto& r : a.rows) {
    std::string o;
    for (size_t j = 0; j < ref.size(); ++j)
      if (ref[j] != '-') o += r[j];
    r = o;
  }
}

// ---- ORFs and candidate regions (src/PhyloCSF.ml:133-217) ----------------

struct Region {
  bool rc;
  int lo, hi;
};

std::vector<std::pair<int, int>> find_orfs(const std::string& dna, int ofs, OrfMode mode,
                                           int min_codons) {
  const bool atg = mode == OrfMode::ATGStop;
  auto up = [&](int p) { return (char)std::toupper((unsigned char)dna[p]); };
  auto is_start = [&](int p) { return up(p) == 'A' && up(p + 1) == 'T' && up(p + 2) == 'G'; };
  auto is_stop = [&](int p) {
    const char a = up(p), b = up(p + 1), c = up(p + 2);
    return a == 'T' && ((b == 'A' && (c == 'A' || c == 'G')) || (b == 'G' && c == 'A'));
  };
  const int len = (int)dna.size();
  std::vector<std::pair<int, int>> orfs;  // newest first, as the reference's list
  std::vector<int> starts;                // newest first
  for (int lo = ofs; lo <= len - 3; ++lo) {
    if ((lo - ofs) % 3 != 0) continue;
    const int hi = lo + 2;
    if ((!atg && starts.empty() && !is_stop(lo)) || (atg && is_start(lo)))
      starts.insert(starts.begin(), lo);
    if (hi + 3 < len && is_stop(hi + 1)) {
      for (int st : starts)
        if (hi > st + 2) orfs.insert(orfs.begin(), {st, hi});
      starts.clear();
    }
  }
  if (!atg)
    for (int st : starts) orfs.insert(orfs.begin(), {st, st + ((len - st) / 3) * 3 - 1});
  if (mode == OrfMode::StopStop3) {
    std::vector<std::pair<int, int>> all;
    for (auto [lo, hi] : orfs) {
      std::vector<std::pair<int, int>> sub = {{lo, hi}};
      const int codons = (hi - lo + 1) / 3;
      const int lo2 = lo + (codons / 3) * 3;
      if (lo2 > lo) sub.insert(sub.begin(), {lo2, hi});
      const int lo3 = lo + (2 * codons / 3) * 3;
      if (lo3 > lo2 && lo3 > lo) sub.insert(sub.begin(), {lo3, hi});
      all.insert(all.end(), sub.begin(), sub.end());
    }
    orfs = all;
  }
  if (mode == OrfMode::ToFirstStop && !orfs.empty()) {
    const auto first = orfs.back();
    orfs.clear();
    if (first.first == ofs) orfs.push_back(first);
  }
  if (mode == OrfMode::FromLastStop && !orfs.empty()) {
    const auto last = orfs.front();
    orfs.clear();
    if (len - last.second <= 3) orfs.push_back(last);
  }
  if (mode == OrfMode::ToOrFromStop && !orfs.empty()) {
    const auto first = orfs.back(), last = orfs.front();
    orfs.clear();
    if (first.first == ofs) orfs.push_back(first);
    if (len - last.second <= 3 && first != last) orfs.insert(orfs.begin(), last);
  }
  std::vector<std::pair<int, int>> out;
  for (auto it = orfs.rbegin(); it != orfs.rend(); ++it)
    if ((it->second - it->first + 1) / 3 >= min_codons) out.push_back(*it);
  return out;
}

std::vector<Region> candidate_regions(const std::string& dna, const Options& o) {
  std::vector<Region> r;
  if (o.orf == OrfMode::AsIs) {
    const int hi = (int)dna.size() - 1;
    r.push_back({false, 0, hi});
    if (o.frames != 1) {
      r.push_back({false, 1, hi});
      r.push_back({false, 2, hi});
    }
    if (o.frames == 6)
      for (int k = 0; k < 3; ++k) r.push_back({true, k, hi});
    return r;
  }
  auto add = [&](bool rc, const std::string& s, int ofs) {
    for (auto [lo, hi] : find_orfs(s, ofs, o.orf, o.min_codons)) r.push_back({rc, lo, hi});
  };
  add(false, dna, 0);
  if (o.frames != 1) {
    add(false, dna, 1);
    add(false, dna, 2);
  }
  if (o.frames == 6) {
    const std::string rcd = revcomp(dna);
    for (int k = 0; k < 3; ++k) add(true, rcd, k);
  }
  return r;
}

// ---- BLS (src/PhyloCSF.ml:250-262; Newick.subtree, Newick.ml:34-61) --------

struct Sub {
  bool some = false;
  double inner = 0.0;  // total length strictly below the (collapsed) node
  bool has_bl = false;
  double bl = 0.0;
};

Sub subtree_len(const pcsf::NewickNode& n, const std::function<bool(const std::string&)>& keep) {
  Sub r;
  if (n.kids.empty()) {
    if (n.label.empty() || keep(n.label)) {
      r.some = true;
      r.has_bl = n.has_branch;
      r.bl = n.branch;
    }
    return r;
  }
  if (!(n.label.empty() || keep(n.label))) return r;
  std::vector<Sub> kept;
  for (const auto& k : n.kids) {
    Sub s = subtree_len(*k, keep);
    if (s.some) kept.push_back(s);
  }
  if (kept.empty()) return r;
  r.some = true;
  if (kept.size() == 1) {  // collapse: branch lengths add (maybe_add)
    r.inner = kept[0].inner;
    r.has_bl = n.has_branch && kept[0].has_bl;
    r.bl = n.branch + kept[0].bl;
    return r;
  }
  for (const Sub& s : kept) {
    if (!s.has_bl) invalid_arg("CamlPaml.Newick.total_length: unspecified branch length");
    r.inner += s.inner + s.bl;
  }
  r.has_bl = n.has_branch;
  r.bl = n.branch;
  return r;
}

double bls(const pcsf::NewickNode& nt, const std::vector<std::string>& rows,
           const std::map<std::string, int>& which_row, int lo, int hi) {
  double total = 0.0;
  for (int i = lo; i <= hi; ++i) {
    auto keep = [&](const std::string& sp) {
      auto it = which_row.find(sp);
      if (it == which_row.end()) return false;
      const char c = rows[it->second][i];
      return !(c == '-' || c == '.' || c == 'N');
    };
    const Sub s = subtree_len(nt, keep);
    if (s.some) total += s.inner;
  }
  return total / (pcsf::newick_total_length(nt) * (double)(hi - lo + 1));
}

// ---- parameters, model and the scoring batch --------------------------------

std::string exe_dir() {
  char buf[4096];
  const ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
  if (n <= 0) return ".";
  buf[n] = 0;
  std::string p(buf);
  return p.substr(0, p.find_last_of('/'));
}

bool is_dir(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}
bool file_exists(const std::string& p) {
  std::ifstream f(p);
  return (bool)f;
}

// initialize_strategy's paramfile (src/PhyloCSF.ml:407-421).  A bare name
// resolves under $PHYLOCSF_BASE/PhyloCSF_Parameters or, when that is unset,
// under the parameter sets shipped with this package.
std::string paramfile(const std::string& paramset, const std::string& suffix, bool required) {
  std::string fn;
  if (paramset.find('/') == std::string::npos) {
    const char* base = std::getenv("PHYLOCSF_BASE");
    std::string dir;
    if (base && is_dir(base)) dir = std::string(base) + "/PhyloCSF_Parameters";
    else if (!base && is_dir(exe_dir() + "/../data/PhyloCSF_Parameters"))
      dir = exe_dir() + "/../data/PhyloCSF_Parameters";
    else
      failwith("PHYLOCSF_BASE environment variable must be set to the root directory of the "
               "source or executable distribution");
    fn = dir + "/" + paramset + suffix;
  } else {
    fn = paramset + suffix;
  }
  if (required && !file_exists(fn)) failwith("could not find required parameter file " + fn);
  return fn;
}

std::string read_file(const std::string& fn) {
  std::ifstream in(fn, std::ios::binary);
  if (!in) sys_error(fn + ": No such file or directory");
  std::stringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

std::vector<std::string> split_lines(const std::string& text) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < text.size()) {
    size_t j = text.find('\n', i);
    if (j == std::string::npos) j = text.size();
    std::string l = text.substr(i, j - i);
    if (!l.empty() && l.back() == '\r') l.pop_back();
    out.push_back(l);
    i = j + 1;
  }
  return out;
}

struct Scorer {
  const Options& o;
  int nl = 0;
  pcsf_ctx* ctx = nullptr;
  pcsf_model* model = nullptr;
  std::vector<int32_t> children;
  std::vector<double> branch, S, pi;
  double omega_h1 = 0.2, sigma_h1 = 0.01;  // OmegaModel.score's defaults

  explicit Scorer(const Options& opt) : o(opt) {}
  ~Scorer() {
    if (model) pcsf_model_destroy(model);
    if (ctx) pcsf_ctx_destroy(ctx);
  }
  [[noreturn]] void device_fail(const char* what, int rc) {
    std::fprintf(stderr, "PhyloCSF: %s failed (%d): %s\n", what, rc,
                 ctx ? pcsf_last_error(ctx) : "");
    std::exit(1);
  }
  void ensure_model() {
    if (model) return;
    int rc = pcsf_ctx_create(o.device, &ctx);
    if (rc) device_fail("GPU context", rc);
    const int n = (int)branch.size();
    if (o.strategy == Strategy::Omega)  // a tree only: the omega model is made per region
      rc = pcsf_model_create(ctx, n, children.data(), branch.data(), 0, nullptr, nullptr, &model);
    else
      rc = pcsf_model_create(ctx, n, children.data(), branch.data(), 2, S.data(), pi.data(),
                             &model);
    if (rc) device_fail("model", rc);
  }
  struct Result {
    // fixed / mle: rho and lpr of the coding and noncoding models; omega: of
    // H0 and H1, with kappa
    double score = 0, anc = 0, rho_c = 1, rho_n = 1, lpr_c = 0, lpr_n = 0;
    double kappa_0 = 0, kappa_1 = 0;
    int status = PCSF_OK;
  };
  // scores regions [codes rows per region given by offsets]
  std::vector<Result> score(const std::vector<uint8_t>& codes, const std::vector<int64_t>& off,
                            std::string* err) {
    const int32_t R = (int32_t)off.size() - 1;
    std::vector<Result> res(R);
    if (o.strategy == Strategy::Nop || R == 0) return res;
    ensure_model();
    pcsf_batch* b = nullptr;
    int rc = pcsf_batch_create(model, codes.data(), off.back(), off.data(), R, &b);
    if (rc) device_fail("batch", rc);
    std::vector<double> sc(R), an(R, NAN), rho(2 * R), lpr(2 * R), kap(2 * R, 0.0);
    std::vector<int32_t> st(R);
    if (o.strategy == Strategy::Omega)
      rc = pcsf_score_omega(b, omega_h1, sigma_h1, o.precision, sc.data(), lpr.data(), rho.data(),
                            kap.data(), st.data(), nullptr);
    else
      rc = pcsf_score(b, o.strategy == Strategy::MaxLik ? PCSF_STRATEGY_MLE : PCSF_STRATEGY_FIXED,
                      o.precision, sc.data(), an.data(), rho.data(), lpr.data(), nullptr,
                      st.data(), nullptr);
    if (rc) device_fail("scoring", rc);
    *err = pcsf_last_error(ctx);
    pcsf_batch_destroy(b);
    for (int32_t r = 0; r < R; ++r)
      res[r] = Result{sc[r],         an[r],       rho[2 * r],  rho[2 * r + 1], lpr[2 * r],
                      lpr[2 * r + 1], kap[2 * r], kap[2 * r + 1], st[r]};
    return res;
  }
};

// ---- per-alignment processing (src/PhyloCSF.ml:280-389) --------------------

struct Pending {
  std::string id;
  std::string failure;  // non-empty: print "<id>\tfailure\t<failure>"
  Alignment aln;
  std::vector<std::string> rc_rows;
  std::map<std::string, int> which_row;
  std::vector<Region> regions;
  int64_t first = 0;  // first region of this alignment in the batch
};

std::string sf(double x) {
  char b[64];
  std::snprintf(b, sizeof b, "%.2f", x);
  return b;
}
double db(double x) { return 10.0 * x / std::log(10.0); }

class Driver {
 public:
  Driver(const Options& o, const std::string& paramset) : o_(o), scorer_(o) {
    const std::string fn_tree = paramfile(paramset, ".nh", true);
    nt_ = pcsf::parse_newick(read_file(fn_tree));
    if (o.species_set) {
      std::vector<std::string> keep;
      std::stringstream ss(o.species);
      std::string x;
      while (std::getline(ss, x, ',')) keep.push_back(x);
      nt_ = pcsf::newick_subtree(std::move(nt_), keep);
      if (!nt_ || pcsf::newick_leaves(*nt_) <= 1) failwith("specify at least two available --species");
    }
    t_ = pcsf::tree_of_newick(*nt_);
    for (int i = 0; i < t_.nl; ++i) t_species_.insert(t_.labels[i]);
    if (o.strategy == Strategy::Omega) {
      // initialize_strategy's OmegaTest (src/PhyloCSF.ml:452-463): an optional
      // key<TAB>value file naming omega_H1 and sigma_H1
      const std::string fn = paramfile(paramset, "_omega", false);
      if (file_exists(fn)) {
        std::map<std::string, std::string> kv;
        for (std::string line : split_lines(read_file(fn))) {
          const size_t a = line.find_first_not_of(" \t"), b = line.find_last_not_of(" \t");
          line = (a == std::string::npos) ? "" : line.substr(a, b - a + 1);
          if (line.empty() || line[0] == '#') continue;
          const size_t tab = line.find('\t');
          if (tab == std::string::npos) kv[line] = "";
          else kv[line.substr(0, tab)] = line.substr(tab + 1);
        }
        auto num = [&](const char* key, double* out) {
          auto it = kv.find(key);
          if (it == kv.end()) return false;
          char* end = nullptr;
          *out = std::strtod(it->second.c_str(), &end);
          return end && end != it->second.c_str() && *end == '\0';
        };
        if (!num("omega_H1", &scorer_.omega_h1) || !num("sigma_H1", &scorer_.sigma_h1))
          failwith("configuration file " + fn +
                   " exists but does not specify valid omega_H1 and sigma_H1 values");
      }
    }
    if (o.strategy == Strategy::Omega) {
      scorer_.nl = t_.nl;
      for (int i = t_.nl; i < t_.n; ++i) {
        scorer_.children.push_back(t_.lc[i]);
        scorer_.children.push_back(t_.rc[i]);
      }
      scorer_.branch = t_.branch;
    } else if (o.strategy != Strategy::Nop) {
      const std::string fc = paramfile(paramset, "_coding.ECM", true);
      const std::string fnc = paramfile(paramset, "_noncoding.ECM", true);
      scorer_.S.resize(2 * 4096);
      scorer_.pi.resize(2 * 64);
      pcsf::read_ecm(fc, scorer_.S.data(), scorer_.pi.data());
      pcsf::read_ecm(fnc, scorer_.S.data() + 4096, scorer_.pi.data() + 64);
      scorer_.nl = t_.nl;
      for (int i = t_.nl; i < t_.n; ++i) {
        scorer_.children.push_back(t_.lc[i]);
        scorer_.children.push_back(t_.rc[i]);
      }
      scorer_.branch = t_.branch;
    }
  }

  // process_alignment; returns false after an abort (the caller exits 255)
  bool process(const std::string& fn) {
    const std::string id = fn.empty() ? "(STDIN)" : fn;
    Pending pa;
    pa.id = id;
    try {
      std::vector<std::string> lines;
      if (fn.empty()) {
        std::string l;
        while (std::getline(std::cin, l)) {
          if (!l.empty() && l.back() == '\r') l.pop_back();
          lines.push_back(l);
        }
      } else {
        lines = split_lines(read_file(fn));
      }
      pa.aln = input_mfa(lines);
      if (o_.remove_ref_gaps) remove_ref_gaps(pa.aln);
      for (auto& r : pa.aln.rows)
        for (char& c : r) c = (c == 'u') ? 't' : (c == 'U') ? 'T' : c;
      if (!o_.allow_ref_gaps && pa.aln.rows[0].find('-') != std::string::npos)
        failwith("the reference sequence (first alignment row) must be ungapped");
      for (auto& r : pa.aln.rows) pa.rc_rows.push_back(revcomp(r));
      std::set<std::string> wtf;
      for (auto& sp : pa.aln.species)
        if (!t_species_.count(sp)) wtf.insert(sp);
      if (!wtf.empty()) {
        std::string m;
        for (auto& s : wtf) m += (m.empty() ? "" : " ") + s;
        failwith("parameters not available for species: " + m);
      }
      for (size_t i = 0; i < pa.aln.species.size(); ++i) pa.which_row[pa.aln.species[i]] = (int)i;
      pa.regions = candidate_regions(pa.aln.rows[0], o_);
      if (o_.procs > 1)
        std::stable_sort(pa.regions.begin(), pa.regions.end(), [](const Region& a, const Region& b) {
          return (a.hi - a.lo) > (b.hi - b.lo);
        });
    } catch (const OcamlError& e) {
      flush();
      std::printf("%s\tabort\t%s\n", id.c_str(), e.text.c_str());
      std::fflush(stdout);
      return false;
    }
    if (pa.regions.empty()) {
      pa.failure = "Failure(\"no sufficiently long ORFs found\")";
    } else {
      std::vector<int> leaf_row(t_.nl, -1);
      for (int l = 0; l < t_.nl; ++l) {
        auto it = pa.which_row.find(t_.labels[l]);
        if (it != pa.which_row.end()) leaf_row[l] = it->second;
      }
      pa.first = (int64_t)off_.size() - 1;
      for (const Region& rg : pa.regions) {
        const auto& rows = rg.rc ? pa.rc_rows : pa.aln.rows;
        for (int pos = rg.lo; pos + 2 <= rg.hi; pos += 3)  // pleaves
          for (int l = 0; l < t_.nl; ++l) {
            uint8_t code = PCSF_MARGINALIZE;
            if (leaf_row[l] >= 0) {
              const std::string& s = rows[leaf_row[l]];
              const char a = s[pos], b = s[pos + 1], c = s[pos + 2];
              if (is_nuc(a) && is_nuc(b) && is_nuc(c))
                code = (uint8_t)(16 * nuc_index(a) + 4 * nuc_index(b) + nuc_index(c));
            }
            codes_.push_back(code);
          }
        off_.push_back((int64_t)(codes_.size() / t_.nl));
      }
    }
    pending_.push_back(std::move(pa));
    if ((long long)codes_.size() >= o_.batch_bytes) flush();
    return true;
  }

  // scores everything pending and writes it out in input order
  void flush() {
    if (pending_.empty()) return;
    std::string err;
    const auto res = scorer_.score(codes_, off_, &err);
    for (Pending& pa : pending_) {
      if (!pa.failure.empty()) {
        std::printf("%s\tfailure\t%s\n", pa.id.c_str(), pa.failure.c_str());
        std::fflush(stdout);
        continue;
      }
      try {
        report(pa, res, err);
      } catch (const OcamlError& e) {
        std::printf("%s\tfailure\t%s\n", pa.id.c_str(), e.text.c_str());
      }
      std::fflush(stdout);
    }
    pending_.clear();
    codes_.clear();
    off_.assign(1, 0);
  }

 private:
  using Diag = std::vector<std::pair<std::string, std::string>>;
  Diag diagnostics(const Scorer::Result& r) const {
    const auto dbs = [](double x) { return sf(db(x)); };
    if (o_.strategy == Strategy::MaxLik)
      return {{"rho_0", sf(1.0)}, {"rho_C", sf(r.rho_c)}, {"rho_N", sf(r.rho_n)},
              {"L(C)", dbs(r.lpr_c)}, {"L(NC)", dbs(r.lpr_n)}};
    if (o_.strategy == Strategy::FixedLik)
      return {{"rho", sf(1.0)}, {"L(C)", dbs(r.lpr_c)}, {"L(NC)", dbs(r.lpr_n)}};
    if (o_.strategy == Strategy::Omega)  // src/OmegaModel.ml:210-215
      return {{"L(H0)", dbs(r.lpr_c)},         {"rho_H0", sf(r.rho_c)},
              {"kappa_H0", sf(r.kappa_0)},     {"omega_H0", sf(1.0)},
              {"sigma_H0", sf(1.0)},           {"L(H1)", dbs(r.lpr_n)},
              {"rho_H1", sf(r.rho_n)},         {"kappa_H1", sf(r.kappa_1)},
              {"omega_H1", sf(scorer_.omega_h1)}, {"sigma_H1", sf(scorer_.sigma_h1)}};
    return {};
  }

  void report(Pending& pa, const std::vector<Scorer::Result>& res, const std::string& err) {
    struct Scored {
      Scorer::Result r;
      Diag diag;
      Region rg;
    };
    std::vector<Scored> ok;
    for (size_t k = 0; k < pa.regions.size(); ++k) {
      const Scorer::Result& r = res[pa.first + k];
      const Region& rg = pa.regions[k];
      if (r.status != PCSF_OK) {  // a region that raised: reported, then skipped
        std::printf("%s\texception\t%d\t%d%s\t%s\n", pa.id.c_str(), rg.lo, rg.hi,
                    o_.frames == 6 ? (rg.rc ? "\t-" : "\t+") : "",
                    ("Failure(" + ocaml_quote(err) + ")").c_str());
        continue;
      }
      ok.push_back({r, diagnostics(r), rg});
    }
    if (ok.empty()) failwith("no regions successfully evaluated");
    const bool several = o_.frames != 1 || o_.orf != OrfMode::AsIs;
    auto line = [&](const char* ty, const Scored& s) {
      std::printf("%s\t%s\t%.4f", pa.id.c_str(), ty, s.r.score);
      if (several) std::printf("\t%d\t%d", s.rg.lo, s.rg.hi);
      if (o_.frames == 6) std::printf("\t%c", s.rg.rc ? '-' : '+');
      const auto& rows = s.rg.rc ? pa.rc_rows : pa.aln.rows;
      if (o_.bls) std::printf("\t%.4f", bls(*nt_, rows, pa.which_row, s.rg.lo, s.rg.hi));
      if (o_.anc_comp) std::printf("\t%.4f", s.r.anc);
      const std::string refdna = rows[0].substr(s.rg.lo, s.rg.hi - s.rg.lo + 1);
      if (o_.dna) std::printf("\t%s", refdna.c_str());
      if (o_.aa) std::printf("\t%s", translate(refdna).c_str());
      if (o_.debug) {
        std::printf("\t#");
        for (const auto& kv : s.diag) std::printf(" %s=%s", kv.first.c_str(), kv.second.c_str());
      }
      std::printf("\n");
    };
    if (o_.all_scores)
      for (const Scored& s : ok) line("orf_score(decibans)", s);
    // List.reduce max: OCaml's structural order on (score record, rc, lo, hi)
    auto less = [](const Scored& a, const Scored& b) {
      if (a.r.score != b.r.score) return a.r.score < b.r.score;
      // OCaml's compare: nan equals nan (the omega test's anc_comp_score)
      if (!(std::isnan(a.r.anc) && std::isnan(b.r.anc)) && a.r.anc != b.r.anc)
        return a.r.anc < b.r.anc;
      if (a.diag != b.diag) return a.diag < b.diag;
      if (a.rg.rc != b.rg.rc) return !a.rg.rc;
      if (a.rg.lo != b.rg.lo) return a.rg.lo < b.rg.lo;
      return a.rg.hi < b.rg.hi;
    };
    size_t best = 0;
    for (size_t k = 1; k < ok.size(); ++k)
      if (less(ok[best], ok[k])) best = k;
    line(several ? "max_score(decibans)" : "score(decibans)", ok[best]);
  }

  const Options& o_;
  Scorer scorer_;
  std::unique_ptr<pcsf::NewickNode> nt_;
  pcsf::Tree t_;
  std::set<std::string> t_species_;
  std::vector<Pending> pending_;
  std::vector<uint8_t> codes_;
  std::vector<int64_t> off_ = {0};
};

}  // namespace

int main(int argc, char** argv) {
  const Options o = parse_options(argc, argv);
  if (o.positional.empty()) {
    std::fputs(kUsage, stderr);
    return 255;
  }
  if (o.orf != OrfMode::AsIs && o.allow_ref_gaps) {
    std::fprintf(stderr, "--allowRefGaps should not be used with --orf\n");
    std::fputs(kUsage, stderr);
    return 255;
  }
  if (o.orf != OrfMode::AsIs && o.frames == 1)
    std::fprintf(stderr,
                 "Warning: --orf with --frames=1; are you sure you don't want to search for ORFs "
                 "in three or six frames?\n");
  const std::string paramset = o.positional[0];
  const std::vector<std::string> inputs(o.positional.begin() + 1, o.positional.end());
  try {
    Driver drv(o, paramset);
    std::vector<std::string> fns;
    if (o.files) {
      if (inputs.empty()) {
        if (isatty(0)) {
          std::printf("Input alignment filename(s) then EOF (Ctrl+D): ");
          std::fflush(stdout);
        }
        std::string l;
        while (std::getline(std::cin, l)) fns.push_back(l);
      } else {
        for (const auto& lf : inputs)
          for (auto& l : split_lines(read_file(lf))) fns.push_back(l);
      }
    } else if (inputs.empty()) {
      if (isatty(0)) {
        std::printf("Input alignment then EOF (Ctrl+D):\n");
        std::fflush(stdout);
      }
      fns.push_back("");
    } else {
      fns = inputs;
    }
    for (const auto& fn : fns)
      if (!drv.process(fn)) return 255;
    drv.flush();
  } catch (const OcamlError& e) {
    std::fflush(stdout);
    std::fprintf(stderr, "Fatal error: exception %s\n", e.text.c_str());
    return 2;
  } catch (const pcsf::ParseError& e) {
    std::fflush(stdout);
    std::fprintf(stderr, "Fatal error: exception Failure(%s)\n", ocaml_quote(e.msg).c_str());
    return 2;
  }
  return 0;
}
