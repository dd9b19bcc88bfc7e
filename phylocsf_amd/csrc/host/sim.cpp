// sim.cpp -- synthetic codon alignments for the bench and the parity tests.
//
// Semantics of PhyloModel.simulate (lib/CamlPaml/PhyloModel.ml:38-54): the
// root codon is drawn from the prior, then every node (root-1 down to 0)
// from the row of P(t) of its parent's codon, by inverse CDF on the
// cumulative row (lib/CamlPaml/Tools.ml:22-41).  The random numbers come from
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011) keyed by the seed and counted by
// (region, column, draw), so any thread count gives the same alignment.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../internal.hpp"

namespace {

struct Philox {
  uint32_t k0, k1;
  static void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    const uint64_t p = (uint64_t)a * b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
  }
  void block(const uint32_t ctr[4], uint32_t out[4]) const {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t key0 = k0, key1 = k1;
    for (int r = 0; r < 10; ++r) {
      uint32_t hi0, lo0, hi1, lo1;
      mulhilo(0xD2511F53u, c0, &hi0, &lo0);
      mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
      const uint32_t n0 = hi1 ^ c1 ^ key0, n2 = hi0 ^ c3 ^ key1;
      c0 = n0;
      c1 = lo1;
      c2 = n2;
      c3 = lo0;
      key0 += 0x9E3779B9u;
      key1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
  }
};

// A stream of uniforms in [0,1) for one (region, column).
struct Stream {
  const Philox* ph;
  uint32_t ctr[4];
  uint32_t buf[4];
  int used = 4;
  Stream(const Philox* p, uint32_t region, uint64_t col) : ph(p) {
    ctr[0] = 0;
    ctr[1] = (uint32_t)col;
    ctr[2] = (uint32_t)(col >> 32);
    ctr[3] = region;
  }
  uint32_t next32() {
    if (used == 4) {
      ph->block(ctr, buf);
      ++ctr[0];
      used = 0;
    }
    return buf[used++];
  }
  double uniform() {  // 53-bit
    const uint64_t a = next32() >> 5, b = next32() >> 6;
    return (double)(a * 67108864ull + b) * (1.0 / 9007199254740992.0);
  }
};

int draw(const double* cum, int k, double u) {
  // smallest i with u*total < cum[i]
  const double x = u * cum[k - 1];
  int lo = 0, hi = k - 1;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (x < cum[mid]) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

}  // namespace

// Writes codes (n_cols x nl) or, if nuc is given instead, nucleotide
// characters: leaf j's row at nuc + j * 3 * n_cols, codon column c at
// characters 3c..3c+2; a gap is "---" and an N codon has 'N' at one of its
// three positions (chosen from the same uniform that decided the N).
static int simulate(pcsf_model* m, int32_t nreg, const int64_t* off, const int32_t* region_model,
             const double* region_rho, int gap_mode, double gap_p, double gap_run, double n_p,
             uint64_t seed, int n_threads, uint8_t* codes, uint8_t* nuc) {
  if (!m || m->tree_only || nreg < 0 || !off || !region_model || !region_rho ||
      (!codes && !nuc))
    return PCSF_EINVAL;
  if (gap_p < 0 || gap_p >= 1 || n_p < 0 || n_p > 1 || (gap_mode == 1 && !(gap_run >= 1)))
    return PCSF_EINVAL;
  // runs of mean gap_run can cover at most gap_run / (gap_run + 1)
  if (gap_mode == 1 && gap_p / (1 - gap_p) > gap_run) return PCSF_EINVAL;
  for (int32_t r = 0; r < nreg; ++r)
    if (region_model[r] < 0 || region_model[r] >= m->nm || !(region_rho[r] > 0) ||
        off[r + 1] < off[r])
      return PCSF_EINVAL;
  const int n = m->n, nl = m->nl, nbr = m->nbr;
  // cumulative rows of P per (model, distinct rho) -- computed once per region here
  Philox ph{(uint32_t)seed, (uint32_t)(seed >> 32)};
  if (n_threads <= 0) n_threads = (int)std::max(1u, std::thread::hardware_concurrency());
  n_threads = std::min(n_threads, 64);
  int status = PCSF_OK;
  std::vector<std::thread> pool;
  std::vector<int> fails(n_threads, 0);
  for (int t = 0; t < n_threads; ++t) {
    pool.emplace_back([&, t]() {
      std::vector<double> cum((size_t)nbr * 4096), P(4096), root(64);
      std::vector<int> state(n);
      int cached_q = -1;
      double cached_rho = -1;
      for (int32_t r = t; r < nreg; r += n_threads) {
        const int q = region_model[r];
        const double rho = region_rho[r];
        if (q != cached_q || rho != cached_rho) {
          for (int br = 0; br < nbr; ++br) {
            std::string msg;
            if (!pcsf::transition_host(m->eig[q], rho * m->branch[br], 1e-6, P.data(), &msg)) {
              fails[t] = 1;
              return;
            }
            double* cb = cum.data() + (size_t)br * 4096;
            for (int a = 0; a < 64; ++a) {
              double s = 0;
              for (int b = 0; b < 64; ++b) {
                s += P[a * 64 + b];
                cb[a * 64 + b] = s;
              }
            }
          }
          double s = 0;
          for (int x = 0; x < 64; ++x) {
            s += m->pi[q * 64 + x];
            root[x] = s;
          }
          cached_q = q;
          cached_rho = rho;
        }
        // gap-run state per species for this region (gap_mode 1): a two-state
        // chain that leaves a run with probability 1 / gap_run per column
        // (geometric run lengths of mean gap_run) and enters one with the
        // probability that makes its stationary gap fraction gap_p:
        // p_start / (p_start + p_stop) = gap_p
        std::vector<int> in_gap(nl, 0);
        const double p_stop = (gap_mode == 1) ? 1.0 / gap_run : 0;
        const double p_start = (gap_mode == 1) ? gap_p * p_stop / (1 - gap_p) : 0;
        for (int64_t c = off[r]; c < off[r + 1]; ++c) {
          Stream st(&ph, (uint32_t)r, (uint64_t)c);
          state[n - 1] = draw(root.data(), 64, st.uniform());
          for (int i = n - 2; i >= 0; --i)
            state[i] = draw(cum.data() + (size_t)i * 4096 + state[m->parent[i]] * 64, 64,
                            st.uniform());
          static const char kNuc[4] = {'A', 'C', 'G', 'T'};
          const int64_t ntot = off[nreg];
          auto put = [&](int j, int code, bool gap, int npos) {
            if (codes) {
              codes[(size_t)c * nl + j] = (gap || npos >= 0) ? (uint8_t)PCSF_MARGINALIZE
                                                             : (uint8_t)code;
              return;
            }
            uint8_t* w = nuc + (size_t)j * 3 * ntot + 3 * (size_t)c;
            for (int k = 0; k < 3; ++k)
              w[k] = gap ? '-' : (k == npos) ? 'N' : kNuc[(code >> (4 - 2 * k)) & 3];
          };
          put(0, state[0], false, -1);
          for (int j = 1; j < nl; ++j) {
            bool gap;
            const double u = st.uniform();
            if (gap_mode == 1) {
              if (c == off[r]) in_gap[j] = u < gap_p;  // stationary start
              else in_gap[j] = in_gap[j] ? (u >= p_stop) : (u < p_start);
              gap = in_gap[j];
            } else {
              gap = u < gap_p;
            }
            const double un = st.uniform();
            const int npos = (un < n_p) ? std::min(2, (int)(3.0 * un / n_p)) : -1;
            put(j, state[j], gap, gap ? -1 : npos);
          }
        }
      }
    });
  }
  for (auto& th : pool) th.join();
  for (int f : fails)
    if (f) status = PCSF_EFAIL;
  return status;
}

extern "C" int pcsf_simulate(pcsf_model* m, int32_t nreg, const int64_t* off,
                             const int32_t* region_model, const double* region_rho,
                             int gap_mode, double gap_p, double gap_run, double n_p,
                             uint64_t seed, int n_threads, uint8_t* codes) {
  if (!codes) return PCSF_EINVAL;
  return simulate(m, nreg, off, region_model, region_rho, gap_mode, gap_p, gap_run, n_p, seed,
                  n_threads, codes, nullptr);
}

extern "C" int pcsf_simulate_nucleotides(pcsf_model* m, int32_t nreg, const int64_t* off,
                                         const int32_t* region_model, const double* region_rho,
                                         int gap_mode, double gap_p, double gap_run,
                                         double n_p, uint64_t seed, int n_threads,
                                         uint8_t* nuc) {
  if (!nuc) return PCSF_EINVAL;
  return simulate(m, nreg, off, region_model, region_rho, gap_mode, gap_p, gap_run, n_p, seed,
                  n_threads, nullptr, nuc);
}
