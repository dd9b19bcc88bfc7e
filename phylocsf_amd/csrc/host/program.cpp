// program.cpp -- the op program of the fused pruning kernel.
#include "program.hpp"

#include <algorithm>

namespace pcsf {

namespace {

// Compile the tree into the fused kernel's op program: post-order, the
// subtree needing more live tiles first (Sethi-Ullman order), so the number
// of live 64x32 tiles is minimal.
struct ProgramBuilder {
  const std::vector<int>& lc;
  const std::vector<int>& rc;
  int nl;
  std::vector<int> need;
  std::vector<int32_t> ops;
  int max_depth = 0;
  std::vector<int> gemm_pos;  // index into ops (pairs) of each GEMM

  // cherry node -> table index when messages above cherries come from tables
  std::vector<int> table_of;

  ProgramBuilder(const std::vector<int>& l, const std::vector<int>& r, int nleaves)
      : lc(l), rc(r), nl(nleaves), need(l.size(), 0) {}

  bool leaf(int i) const { return i < nl; }
  bool tabled(int i) const { return !table_of.empty() && table_of[i] >= 0; }

  // walk(y, d) followed by the GEMM of branch y, or the table read replacing both
  void message(int y, int d) {
    if (tabled(y)) {
      emit(OP_CHERRYT, d, lc[y], rc[y]);
      ops[ops.size() - 2] = table_of[y];
      return;
    }
    walk(y, d);
    gemm(d, y);
  }

  void compute_need() {
    for (int i = nl; i < static_cast<int>(lc.size()); ++i) {
      const int a = lc[i], b = rc[i];
      if (leaf(a) && leaf(b)) need[i] = 1;
      else if (leaf(a)) need[i] = std::max(1, need[b]);
      else if (leaf(b)) need[i] = std::max(1, need[a]);
      else {
        const int x = need[a], y = need[b];
        need[i] = std::max(std::max(x, y), std::min(x, y) + 1);
      }
    }
  }

  void emit(int op, int d, int a0, int a1) {
    ops.push_back(op | (d << 8));
    ops.push_back((a0 & 0xffff) | ((a1 & 0xffff) << 16));
    ops.push_back(kNoBranch);
    ops.push_back(0);
    max_depth = std::max(max_depth, d);
  }

  int count() const { return static_cast<int>(ops.size() / kOpWords); }
  int opcode(int k) const { return ops[kOpWords * k] & 0xff; }

  void gemm(int d, int br) {
    gemm_pos.push_back(count());
    emit(OP_GEMM, d, br, kNoBranch);
  }

  void walk(int i, int d) {
    const int a = lc[i], b = rc[i];
    if (leaf(a) && leaf(b)) {
      emit(OP_CHERRY, d, a, b);
    } else if (leaf(a) || leaf(b)) {
      const int x = leaf(a) ? a : b, y = leaf(a) ? b : a;
      message(y, d);
      emit(OP_LEAFMUL, d, x, 0);
    } else {
      const int first = (need[b] > need[a]) ? b : a, second = (first == a) ? b : a;
      message(first, d);
      message(second, d + 1);
      emit(OP_MULPOP, d, 0, 0);
    }
  }

  void build(int root) {
    compute_need();
    walk(root, 0);
    emit(OP_ROOT, 0, 0, 0);
    emit(OP_END, 0, 0, 0);
    // each GEMM prefetches the next GEMM's tile
    for (size_t k = 0; k + 1 < gemm_pos.size(); ++k) {
      int32_t& w1 = ops[kOpWords * gemm_pos[k] + 1];
      const int br = w1 & 0xffff;
      const int nbr = ops[kOpWords * gemm_pos[k + 1] + 1] & 0xffff;
      w1 = br | (nbr << 16);
    }
    // the last GEMM before a gather op (with no gather op in between) issues
    // that op's first leaf gather or its table row, unless the GEMM runs at
    // the deepest GEMM level (the tile issued ahead is live through the
    // GEMM, so the program's peak of live tiles stays where it was)
    int deepest = -1;
    for (int g : gemm_pos) deepest = std::max(deepest, (ops[kOpWords * g] >> 8) & 0xff);
    int last_gemm = -1;
    for (int k = 0; k < count(); ++k) {
      const int op = opcode(k);
      if (op == OP_GEMM) {
        last_gemm = ((ops[kOpWords * k] >> 8) & 0xff) < deepest ? k : -1;
      } else if (op == OP_CHERRYT) {
        const int a0 = ops[kOpWords * k + 1] & 0xffff, a1 = (ops[kOpWords * k + 1] >> 16) & 0xffff;
        const int tab = ops[kOpWords * k + 2];
        if (last_gemm >= 0 && a0 <= kPfLeafMax && a1 <= kPfLeafMax && tab <= kPfTableMax) {
          ops[kOpWords * last_gemm + 3] = static_cast<int32_t>(
              kPfTable | (static_cast<uint32_t>(tab) << 24) | (static_cast<uint32_t>(a1) << 12) |
              static_cast<uint32_t>(a0));
          ops[kOpWords * k] |= 1 << 16;
        }
        last_gemm = -1;
      } else if (op == OP_CHERRY || op == OP_LEAFMUL) {
        if (last_gemm >= 0) {
          ops[kOpWords * last_gemm + 2] = ops[kOpWords * k + 1] & 0xffff;
          ops[kOpWords * k] |= 1 << 16;
        }
        last_gemm = -1;
      }
    }
  }
};

}  // namespace

Program make_program(const std::vector<int>& lc, const std::vector<int>& rc, int nl,
                     bool with_tables) {
  const int n = static_cast<int>(lc.size());
  Program pg;
  ProgramBuilder pb(lc, rc, nl);
  if (with_tables) {
    // cherries whose message feeds a GEMM (every cherry but a root cherry)
    pb.table_of.assign(n, -1);
    for (int i = nl; i < n - 1; ++i)
      if (lc[i] < nl && rc[i] < nl) {
        pb.table_of[i] = static_cast<int>(pg.cherries.size() / 3);
        pg.cherries.push_back(i);
        pg.cherries.push_back(lc[i]);
        pg.cherries.push_back(rc[i]);
      }
  }
  pb.build(n - 1);
  pg.ops = pb.ops;
  pg.slots = pb.max_depth + 1;
  pg.n_gemm = static_cast<int>(pb.gemm_pos.size());
  pg.first_gemm_br = pb.gemm_pos.empty() ? kNoBranch : (pg.ops[kOpWords * pb.gemm_pos[0] + 1] & 0xffff);
  return pg;
}

}  // namespace pcsf
