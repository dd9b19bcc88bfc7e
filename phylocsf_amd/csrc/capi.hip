// capi.hip -- the extern "C" boundary (include/phylocsf_hip.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <thread>

#include "host/maximize.hpp"
#include "host/omega.hpp"
#include "host/program.hpp"
#include "host/text.hpp"
#include "internal.hpp"

namespace {

int set_err(pcsf_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return set_err((ctx), e_ == hipErrorOutOfMemory ? PCSF_ENOMEM : PCSF_EDEVICE,     \
                     std::string(#expr) + ": " + hipGetErrorString(e_));                \
  } while (0)

// PCSF_DEVICE_ALLOC_LIMIT=<bytes>: any single device allocation larger than
// that fails as if the device were out of memory (read at every allocation;
// lets a test drive the library's recovery from a failed allocation).
bool alloc_over_limit(size_t bytes) {
  const char* lim = std::getenv("PCSF_DEVICE_ALLOC_LIMIT");
  if (!lim || !*lim) return false;
  return bytes > std::strtoull(lim, nullptr, 10);
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return hipSuccess;
  if (alloc_over_limit(count * sizeof(T))) return hipErrorOutOfMemory;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

bool rho_valid(double r) { return r > 0.0 && std::isfinite(r); }

bool known_precision(int p) { return p >= PCSF_FP64 && p <= PCSF_FP64M; }
// FP64 arithmetic throughout (plain or on matrix cores)
bool fp64_precision(int p) { return p == PCSF_FP64 || p == PCSF_FP64M; }

constexpr const char* kTreeOnly =
    "a tree-only model (made with n_models = 0) serves the omega test only";

void free_psets(PSets& ps);

hipError_t alloc_psets(PSets& ps, int cap, int nbr, bool fp64, int nctab = 0,
                       bool tangent = false, bool own_models = false) {
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) {
    if (e == hipSuccess) e = x;
  };
  const size_t n = (size_t)cap * nbr;
  chk(dalloc(&ps.d_PT32, n * kLeafRows * 64));
  chk(dalloc(&ps.d_Pfrag, n * kTileFloats));
  chk(dalloc(&ps.d_lnPcc, n * kLeafRows));
  chk(dalloc(&ps.d_rho, (size_t)cap));
  chk(dalloc(&ps.d_active, (size_t)cap));
  if (fp64) chk(dalloc(&ps.d_P64T, n * 4096));
  if (fp64) chk(dalloc(&ps.d_P64F, n * kTile64));
  if (fp64) chk(dalloc(&ps.d_P64FT, n * 4096));
  if (tangent) {
    chk(dalloc(&ps.d_DT32, n * kLeafRows * 64));
    chk(dalloc(&ps.d_Dfrag, n * 4096));
  }
  if (nctab > 0) {
    chk(dalloc(&ps.d_ctab, (size_t)cap * nctab * kTabCodes * kTabCodes * 64));
    chk(dalloc(&ps.d_clnt, (size_t)cap * nctab * kTabCodes * kTabCodes));
  }
  if (own_models) {
    chk(dalloc(&ps.d_eVT, (size_t)cap * 4096));
    chk(dalloc(&ps.d_eVinv, (size_t)cap * 4096));
    chk(dalloc(&ps.d_elam, (size_t)cap * 64));
    chk(dalloc(&ps.d_epi, (size_t)cap * 64));
    chk(dalloc(&ps.d_elnpi, (size_t)cap * 64));
  }
  if (e != hipSuccess) {
    free_psets(ps);  // nothing half-allocated survives: cap stays 0
    return e;
  }
  ps.cap = cap;
  ps.fp64 = fp64;
  ps.fmt = -1;
  return e;
}

void free_psets(PSets& ps) {
  dfree(ps.d_P64T);
  dfree(ps.d_P64F);
  dfree(ps.d_P64FT);
  dfree(ps.d_PT32);
  dfree(ps.d_Pfrag);
  dfree(ps.d_lnPcc);
  dfree(ps.d_rho);
  dfree(ps.d_active);
  dfree(ps.d_ctab);
  dfree(ps.d_clnt);
  dfree(ps.d_DT32);
  dfree(ps.d_Dfrag);
  dfree(ps.d_eVT);
  dfree(ps.d_eVinv);
  dfree(ps.d_elam);
  dfree(ps.d_epi);
  dfree(ps.d_elnpi);
  ps.cap = 0;
  ps.fp64 = false;
  ps.fmt = -1;
}

// The reference's message for a P(t) failure recorded by the P builder as
// key (kernels/pbuild.hip: ((set * nbr + br) * 64 + row) * 65 + col, then
// two bits of the check).  The failing matrix is rebuilt on the host in FP64
// (Q.Diag.real_to_Pt, lib/CamlPaml/Q.ml:211-249), which also gives the
// offending value; should the host copy pass where the device's failed (a
// rounding-level difference), the device's position is reported.
int pbuild_failure(pcsf_model* m, unsigned long long key, const double* rho_host,
                   bool own_models) {
  pcsf_ctx* c = m->ctx;
  const int check = (int)(key & 3);
  unsigned long long pos = key >> 2;
  const int col = (int)(pos % 65);
  pos /= 65;
  const int row = (int)(pos % 64);
  pos /= 64;
  const int br = (int)(pos % (unsigned long long)m->nbr);
  const int set = (int)(pos / (unsigned long long)m->nbr);
  const double t = rho_host[set] * m->branch[br];
  if (check == 0) return set_err(c, PCSF_EINVAL, "CamlPaml.Q.to_Pt");
  std::vector<double> P(4096);
  std::string msg;
  if (!own_models && !pcsf::transition_host(m->eig[set % m->nm], t, 1e-6, P.data(), &msg)) {
    if (msg.rfind("Assert_failure", 0) == 0) msg = "Assert_failure (Q.ml:245)";
    return set_err(c, PCSF_EFAIL, msg);
  }
  char buf[256];
  if (check == 1)
    std::snprintf(buf, sizeof buf, "CamlPaml.Q.substition_matrix: expm(%.2e*Q)[%d,%d] < 0", t,
                  row, col);
  else if (check == 2)
    std::snprintf(buf, sizeof buf, "CamlPaml.Q.substitution matrix: sum(expm(%.2e*Q)[%d,] > 1",
                  t, row);
  else
    std::snprintf(buf, sizeof buf, "Assert_failure (Q.ml:245)");
  return set_err(c, PCSF_EFAIL, buf);
}

// Runs the P builder over `nsets` sets whose rho/active are already on the
// device.  A failure comes back as the key of the first failing (set,
// branch, row) in the reference's order; its message is made on the host
// from the same P(t) built in FP64 by Q.Diag.real_to_Pt's rules.
int run_pbuild(pcsf_model* m, PSets& ps, int nsets, const double* rho_host, int fmt,
               float* ms_out,
               unsigned outputs = kBuildP64 | kBuildP64M | kBuildP64MT | kBuildMfma |
                                  kBuildTangent) {
  pcsf_ctx* c = m->ctx;
  HIP_TRY(c, hipMemsetAsync(m->d_perr, 0xff, sizeof(unsigned long long), c->stream));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing && ms_out) {
    HIP_TRY(c, hipEventCreate(&e0));
    HIP_TRY(c, hipEventCreate(&e1));
    HIP_TRY(c, hipEventRecord(e0, c->stream));
  }
  ps.fmt = -1;
  HIP_TRY(c, launch_pbuild(m, ps, nsets, fmt, outputs, c->stream));
  if (e1) HIP_TRY(c, hipEventRecord(e1, c->stream));
  unsigned long long key = ~0ull;
  HIP_TRY(c, hipMemcpyAsync(&key, m->d_perr, sizeof key, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (e0) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_out += ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (key != ~0ull) return pbuild_failure(m, key, rho_host, ps.d_eVT != nullptr);
  ps.fmt = (outputs & kBuildMfma) ? fmt : -1;
  return PCSF_OK;
}

// Transition sets and cherry tables of the fixed scales rho[0 .. nm) (one
// per model) with P tiles in format fmt; rebuilt only when either changes.
// Allocates nothing: the sets are the model's own and cached_rho's capacity
// is reserved when the model is made.
int ensure_transitions(pcsf_model* m, const double* rho, int fmt, float* ms_out) {
  pcsf_ctx* c = m->ctx;
  if (ms_out) *ms_out = 0.f;
  const bool same_rho = (int)m->cached_rho.size() == m->nm &&
                        std::equal(rho, rho + m->nm, m->cached_rho.begin());
  if (same_rho && m->fixed.fmt == fmt) return PCSF_OK;
  m->cached_rho.clear();
  HIP_TRY(c, hipMemcpyAsync(m->fixed.d_rho, rho, sizeof(double) * m->nm, hipMemcpyHostToDevice,
                            c->stream));
  HIP_TRY(c, hipMemsetAsync(m->fixed.d_active, 1, (size_t)m->nm, c->stream));
  int rc = run_pbuild(m, m->fixed, m->nm, rho, fmt, ms_out);
  if (rc) return rc;
  if (!same_rho) HIP_TRY(c, launch_cherry_tables(m, m->fixed, m->nm, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  m->cached_rho.assign(rho, rho + m->nm);
  return PCSF_OK;
}

// Builds the region-aligned copy of the batch's codes (first per-region call).
int ensure_aligned(pcsf_batch* b) {
  if (b->al.d_codes || b->nreg == 0) return PCSF_OK;
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  Aligned& al = b->al;
  al.astart.assign(b->nreg, 0);
  al.block0.assign(b->nreg, 0);
  int64_t pos = 0;
  for (int32_t r = 0; r < b->nreg; ++r) {
    al.astart[r] = pos;
    al.block0[r] = pos / kAlignCols;
    const int64_t len = b->offsets[r + 1] - b->offsets[r];
    pos += (len + kAlignCols - 1) / kAlignCols * kAlignCols;
  }
  al.ld = std::max<int64_t>(kGroupCols, (pos + kGroupCols - 1) / kGroupCols * kGroupCols);
  std::vector<int32_t> br(al.ld / kAlignCols, -1);
  for (int32_t r = 0; r < b->nreg; ++r) {
    const int64_t len = b->offsets[r + 1] - b->offsets[r];
    for (int64_t k = 0; k < (len + kAlignCols - 1) / kAlignCols; ++k) br[al.block0[r] + k] = r;
  }
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) {
    if (e == hipSuccess) e = x;
  };
  chk(dalloc(&al.d_codes, (size_t)m->nl * al.ld));
  chk(dalloc(&al.d_block_region, br.size()));
  chk(dalloc(&al.d_astart, al.astart.size()));
  chk(dalloc(&b->d_alogz, (size_t)m->nm * al.ld));
  chk(dalloc(&b->d_acolelpr, (size_t)m->nm * al.ld));
  if (e == hipSuccess)
    chk(hipMemcpy(al.d_block_region, br.data(), br.size() * 4, hipMemcpyHostToDevice));
  if (e == hipSuccess)
    chk(hipMemcpy(al.d_astart, al.astart.data(), al.astart.size() * 8, hipMemcpyHostToDevice));
  if (e == hipSuccess) chk(launch_align_codes(b, c->stream));
  if (e == hipSuccess) chk(hipStreamSynchronize(c->stream));
  if (e != hipSuccess) {
    // leave the batch as before the call: the next per-region call retries
    dfree(al.d_codes);
    dfree(al.d_block_region);
    dfree(al.d_astart);
    dfree(b->d_alogz);
    dfree(b->d_acolelpr);
    al.astart.clear();
    al.block0.clear();
    al.ld = 0;
  }
  HIP_TRY(c, e);
  return PCSF_OK;
}

// Device memory for the per-region transition sets of a batch: regions go
// through in chunks of as many sets as fit, 24 GiB (a fraction of the 288 GB
// of one MI355X, so that several large batches can live side by side) unless
// PHYLOCSF_SET_BUDGET_MB says otherwise (read on every pass; small values
// force many chunks, which the tests use).
size_t set_budget() {
  const char* e = std::getenv("PHYLOCSF_SET_BUDGET_MB");
  if (e && *e) {
    const long long mb = std::atoll(e);
    if (mb > 0) return (size_t)mb << 20;
  }
  return (size_t)24 << 30;
}

// Bytes of one transition set on the device.
size_t set_bytes(const pcsf_model* m, bool fp64, bool tangent = false, bool own_model = false) {
  return (size_t)m->nbr * (kLeafRows * 64 * 4 + kTileFloats * 4 + kLeafRows * 8 +
                           (fp64 ? (2 * 4096 + kTile64) * 8 : 0) +
                           (tangent ? (kLeafRows + 64) * 64 * 4 : 0)) +
         (own_model ? (2 * 4096 + 3 * 64) * 8 : 0);
}

// Rate models of their own for every (region, model) pair k = r * nm + q:
// the omega model's twelve settings per pair (host array, np * 12), from
// which each chunk's eigensystems and root priors are built on the GPU.
struct SetModels {
  const double* qs;
};

// Node tiles of the FP64 outside pass (adjoint_mfma.hip): one
// set per resident wave, allocated on first use.
int ensure_adj_scratch(pcsf_batch* b) {
  const pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  const size_t need = adjoint_mfma_scratch_doubles(m, adjoint_mfma_grid(m));
  if (b->adj_scratch >= need) return PCSF_OK;
  dfree(b->d_adj_scratch);
  b->adj_scratch = 0;
  HIP_TRY(c, dalloc(&b->d_adj_scratch, need));
  b->adj_scratch = need;
  return PCSF_OK;
}

// Forward with one tree scale per (region, model): rho[r*nm+q].  active
// (optional, r*nm+q) restricts the work to the listed pairs; results of the
// others are left as they were.  Regions go through in chunks whose
// transition sets fit the memory budget.
int forward_regions(pcsf_batch* b, const double* rho, const uint8_t* active, int precision,
                    const SetModels* sm = nullptr) {
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  const int nm = m->nm;
  int rc = ensure_aligned(b);
  if (rc) return rc;
  const bool fp64 = fp64_precision(precision);
  const int fmt = fp64 ? kFragF32 : frag_format(precision);
  const bool own = sm != nullptr;
  // a reallocated pool keeps what the old one served (FP64 and tangent
  // data), so that maximisers alternating passes do not rebuild it each time
  const bool keep = own == (b->rs.d_eVT != nullptr);
  const bool want_fp64 = fp64 || (keep && b->rs.fp64);
  const bool want_tangent = keep && b->rs.d_Dfrag != nullptr;
  int cap = (int)std::max<size_t>(
      nm, set_budget() / set_bytes(m, want_fp64, want_tangent, own) / nm * nm);
  cap = std::min<int>(cap, std::max(nm, b->nreg * nm));
  if (b->rs.cap < cap || (fp64 && !b->rs.fp64) || !keep) {
    free_psets(b->rs);
    b->eig_ok.clear();
    b->eig_ref_ok.clear();
    HIP_TRY(c, alloc_psets(b->rs, cap, m->nbr, want_fp64, 0, want_tangent, own));
  }
  const int chunk_regions = b->rs.cap / nm;
  // one chunk: set k is pair k on every pass, so a set's eigensystem can
  // seed (or stand for) the next one's
  const bool one_chunk = chunk_regions >= b->nreg;
  if (!one_chunk) {
    b->eig_ok.clear();
    b->eig_ref_ok.clear();
  }
  int variant = fp64 ? 0 : m->mfma_variant[precision];
  if (16 * kMfmaVariants[variant].wpg * kMfmaVariants[variant].g > kAlignCols) variant = 1;
  std::vector<uint8_t> act, eigmode;
  b->t_pbuild = b->t_prune = b->t_reduce = 0.f;
  for (int32_t r0 = 0; r0 < b->nreg; r0 += chunk_regions) {
    const int32_t nr = std::min<int32_t>(chunk_regions, b->nreg - r0);
    act.assign((size_t)nr * nm, 1);
    bool any = false;
    for (size_t k = 0; k < act.size(); ++k) {
      if (active) act[k] = active[(size_t)r0 * nm + k] ? 1 : 0;
      any = any || act[k];
    }
    if (!any) continue;
    RegionChunk ch;
    ch.r0 = r0;
    ch.nreg = nr;
    ch.block0 = b->al.block0[r0];
    const int64_t end_block =
        (r0 + nr < b->nreg) ? b->al.block0[r0 + nr] : (b->al.ld / kAlignCols);
    ch.nblocks = end_block - ch.block0;
    ch.sets = &b->rs;
    HIP_TRY(c, hipMemcpyAsync(b->rs.d_rho, rho + (size_t)r0 * nm, sizeof(double) * nr * nm,
                              hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(b->rs.d_active, act.data(), act.size(), hipMemcpyHostToDevice,
                              c->stream));
    if (own) {  // the chunk's rate models, built from their settings
      const size_t k0 = (size_t)r0 * nm, nk = (size_t)nr * nm;
      if (b->setqs_cap < (size_t)b->rs.cap * 12) {
        dfree(b->d_setqs);
        dfree(b->d_eigmode);
        dfree(b->d_eigref);
        b->setqs_cap = 0;
        b->eig_ref_ok.clear();
        HIP_TRY(c, dalloc(&b->d_setqs, (size_t)b->rs.cap * 12));
        HIP_TRY(c, dalloc(&b->d_eigmode, (size_t)b->rs.cap));
        HIP_TRY(c, dalloc(&b->d_eigref, (size_t)b->rs.cap * 4096));
        b->setqs_cap = (size_t)b->rs.cap * 12;
      }
      // per set: keep (inactive, or settings as last built), warm (the
      // reference basis has the same pi: only kappa / omega differ), or cold
      if (one_chunk && b->eig_ok.size() != nk) {
        b->eig_ok.assign(nk, 0);
        b->eig_qs.assign(nk * 12, 0.0);
      }
      if (one_chunk && b->eig_ref_ok.size() != nk) {
        b->eig_ref_ok.assign(nk, 0);
        b->eig_ref_qs.assign(nk * 12, 0.0);
      }
      eigmode.assign(nk, kEigCold);
      bool any = false;
      for (size_t k = 0; k < nk; ++k) {
        const double* q = sm->qs + (k0 + k) * 12;
        if (!act[k]) {
          eigmode[k] = kEigKeep;
          continue;
        }
        if (one_chunk && b->eig_ok[k] && std::memcmp(&b->eig_qs[k * 12], q, 12 * 8) == 0)
          eigmode[k] = kEigKeep;
        else if (one_chunk && b->eig_ref_ok[k] &&
                 std::memcmp(&b->eig_ref_qs[k * 12 + 2], q + 2, 10 * 8) == 0)
          eigmode[k] = kEigWarm;
        any = any || eigmode[k] != kEigKeep;
      }
      if (any) {
        if (one_chunk)  // invalid until the launch below is issued
          for (size_t k = 0; k < nk; ++k) {
            if (eigmode[k] != kEigKeep) b->eig_ok[k] = 0;
            if (eigmode[k] == kEigCold) b->eig_ref_ok[k] = 0;
          }
        HIP_TRY(c, hipMemcpyAsync(b->d_setqs, sm->qs + k0 * 12, nk * 12 * 8,
                                  hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(b->d_eigmode, eigmode.data(), nk, hipMemcpyHostToDevice,
                                  c->stream));
        HIP_TRY(c, launch_omega_eigen(b->d_setqs, b->d_eigmode, b->d_eigref, (int)nk, b->rs,
                                      c->stream));
        if (one_chunk)
          for (size_t k = 0; k < nk; ++k) {
            const double* q = sm->qs + (k0 + k) * 12;
            if (eigmode[k] != kEigKeep) {
              std::memcpy(&b->eig_qs[k * 12], q, 12 * 8);
              b->eig_ok[k] = 1;
            }
            if (eigmode[k] == kEigCold) {
              std::memcpy(&b->eig_ref_qs[k * 12], q, 12 * 8);
              b->eig_ref_ok[k] = 1;
            }
          }
      }
    }
    float tp = 0.f;
    rc = run_pbuild(m, b->rs, nr * nm, rho + (size_t)r0 * nm, fmt, &tp,
                    precision == PCSF_FP64M ? kBuildP64M : fp64 ? kBuildP64 : kBuildMfma);
    if (rc) {
      char buf[64];
      std::snprintf(buf, sizeof buf, " (region block from %d)", r0);
      c->err += buf;
      return rc;
    }
    b->t_pbuild += tp;
    if (c->timing) HIP_TRY(c, hipEventRecord(b->ev[0], c->stream));
    hipError_t e =
        (precision == PCSF_FP64M)
            ? launch_forward_fp64m(m, b, ch, c->stream)
        : fp64 ? launch_prune_fp64_regions(m, b, ch, c->stream)
               : launch_prune_mfma_regions(m, b, precision, variant, ch, c->stream);
    if (e == hipErrorInvalidValue && precision == PCSF_FP64)
      return set_err(c, PCSF_EINVAL, "FP64 kernel: tree workspace exceeds 64 KiB of LDS");
    HIP_TRY(c, e);
    if (c->timing) HIP_TRY(c, hipEventRecord(b->ev[1], c->stream));
    HIP_TRY(c, launch_region_reduce_aligned(b, ch, b->d_loglik, b->d_elpr, c->stream));
    if (c->timing) {
      HIP_TRY(c, hipEventRecord(b->ev[2], c->stream));
      HIP_TRY(c, hipEventSynchronize(b->ev[2]));
      float a = 0.f, r = 0.f;
      HIP_TRY(c, hipEventElapsedTime(&a, b->ev[0], b->ev[1]));
      HIP_TRY(c, hipEventElapsedTime(&r, b->ev[1], b->ev[2]));
      b->t_prune += a;
      b->t_reduce += r;
    }
  }
  return PCSF_OK;
}

// Gradient pass at tree scales rho (nreg * nm) over the pairs in `active`
// (optional): per-region lpr into d_gloglik, d lpr / d rho into d_grad and
// the sum over columns of (d ln z / d rho)^2 into d_gsq
// (the per-column buffers of the forward hold the column values).  tangent = false: the FP64
// outside-pass adjoint, which also accumulates, if d_gbr is given, the
// per-region sums of d ln z / d t_br (nreg * nm * nbr); tangent = true: the
// FP32 MFMA tangent pass (d lpr / d rho only).
int adjoint_regions(pcsf_batch* b, const double* rho, const uint8_t* active, double* d_gbr,
                    bool tangent = false) {
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  const int nm = m->nm;
  int rc = ensure_aligned(b);
  if (rc) return rc;
  // (sets with rate models of their own are not kept: the gradient passes
  // run on the models' eigensystems)
  const bool keep = b->rs.d_eVT == nullptr;
  const bool want_fp64 = !tangent || (keep && b->rs.fp64);
  const bool want_tangent = tangent || (keep && b->rs.d_Dfrag != nullptr);
  int cap = (int)std::max<size_t>(
      nm, set_budget() / set_bytes(m, want_fp64, want_tangent) / nm * nm);
  cap = std::min<int>(cap, std::max(nm, b->nreg * nm));
  if (b->rs.cap < cap || (!tangent && !b->rs.fp64) || (tangent && !b->rs.d_Dfrag) || !keep) {
    free_psets(b->rs);
    b->eig_ok.clear();
    b->eig_ref_ok.clear();
    HIP_TRY(c, alloc_psets(b->rs, cap, m->nbr, want_fp64, 0, want_tangent));
  }
  const int chunk_regions = b->rs.cap / nm;
  std::vector<uint8_t> act;
  for (int32_t r0 = 0; r0 < b->nreg; r0 += chunk_regions) {
    const int32_t nr = std::min<int32_t>(chunk_regions, b->nreg - r0);
    act.assign((size_t)nr * nm, 1);
    bool any = false;
    for (size_t k = 0; k < act.size(); ++k) {
      if (active) act[k] = active[(size_t)r0 * nm + k] ? 1 : 0;
      any = any || act[k];
    }
    if (!any) continue;
    RegionChunk ch;
    ch.r0 = r0;
    ch.nreg = nr;
    ch.block0 = b->al.block0[r0];
    const int64_t end_block =
        (r0 + nr < b->nreg) ? b->al.block0[r0 + nr] : (b->al.ld / kAlignCols);
    ch.nblocks = end_block - ch.block0;
    ch.sets = &b->rs;
    HIP_TRY(c, hipMemcpyAsync(b->rs.d_rho, rho + (size_t)r0 * nm, sizeof(double) * nr * nm,
                              hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(b->rs.d_active, act.data(), act.size(), hipMemcpyHostToDevice,
                              c->stream));
    float tp = 0.f;
    rc = run_pbuild(m, b->rs, nr * nm, rho + (size_t)r0 * nm, kFragF32, &tp,
                    tangent ? (kBuildMfma | kBuildTangent) : (kBuildP64M | kBuildP64MT));
    if (rc) return rc;
    hipError_t e;
    if (tangent) {
      e = launch_prune_tangent_regions(m, b, ch, c->stream);
    } else {
      rc = ensure_adj_scratch(b);
      if (rc) return rc;
      e = launch_adjoint_mfma(m, b, ch, d_gbr ? d_gbr + (size_t)r0 * nm * m->nbr : nullptr,
                              b->d_adj_scratch, adjoint_mfma_grid(m), c->stream);
    }
    HIP_TRY(c, e);
    HIP_TRY(c, launch_region_reduce_aligned(b, ch, b->d_gloglik, b->d_grad, c->stream,
                                            b->d_gsq));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PCSF_OK;
}

// ---- scratch of pcsf_forward_noalloc ---------------------------------------
// One fixed device array per GPU, held until the process ends: the
// leaf-major codes, the per-column ln z and elpr of a call, and a status
// word.  On the host the name `scratch` stands for a host-side shadow, not a
// device address, so host code never hands it to a kernel or a copy; each
// device's address comes from hipGetSymbolAddress, taken once per device
// with that device current (pcsf_ctx_create) and kept in g_scratch.
}  // namespace

static __device__ __align__(256) unsigned char scratch[PCSF_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};          // device address of `scratch`, per device
hipError_t g_scratch_err[kMaxDevices] = {};  // status of that lookup
bool g_scratch_done[kMaxDevices] = {};

// Looks the array up on `device` (which must be current) the first time;
// afterwards returns the kept address.
void* scratch_address(int device, hipError_t* err) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  if (!g_scratch_done[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch[device] = (e == hipSuccess) ? ptr : nullptr;
    g_scratch_err[device] = e;
    g_scratch_done[device] = true;
  }
  *err = g_scratch_err[device];
  return g_scratch[device];
}

// Region offsets on the device: off[0] = 0, nondecreasing, off[nreg] =
// ncols; a violation sets bit 1 of *status.
__global__ void noalloc_check_offsets_kernel(const int64_t* __restrict__ off, int nreg,
                                             int64_t ncols, int* __restrict__ status) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > nreg) return;
  bool bad = (r == 0) ? off[0] != 0 : off[r] < off[r - 1];
  if (r == nreg) bad = bad || off[nreg] != ncols;
  if (bad) atomicOr(status, 2);
}

}  // namespace

extern "C" {

int pcsf_ctx_create(int device, pcsf_ctx** out) {
  if (!out) return PCSF_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PCSF_EDEVICE;
  if (device < 0 || device >= ndev) return PCSF_EINVAL;
  auto* c = new (std::nothrow) pcsf_ctx();
  if (!c) return PCSF_ENOMEM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return PCSF_EDEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    c->num_cus = prop.multiProcessorCount;
  // the no-allocation forward's scratch on this device (looked up once per
  // device; a failure is reported by pcsf_forward_noalloc)
  hipError_t se = hipSuccess;
  (void)scratch_address(device, &se);
  *out = c;
  return PCSF_OK;
}

void pcsf_ctx_destroy(pcsf_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  EstepCache& es = c->es;
  if (es.copy) (void)hipStreamSynchronize(es.copy);
  dfree(es.U);
  dfree(es.V);
  dfree(es.scr);
  for (int i = 0; i < 2; ++i) {
    dfree(es.leaves[i]);
    if (es.staged[i]) (void)hipHostFree(es.staged[i]);
    if (es.uploaded[i]) (void)hipEventDestroy(es.uploaded[i]);
    if (es.consumed[i]) (void)hipEventDestroy(es.consumed[i]);
  }
  if (es.copy) (void)hipStreamDestroy(es.copy);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* pcsf_last_error(const pcsf_ctx* c) { return c ? c->err.c_str() : "no context"; }

int pcsf_ctx_set_timing(pcsf_ctx* c, int enable) {
  if (!c) return PCSF_EINVAL;
  c->timing = enable != 0;
  return PCSF_OK;
}

void* pcsf_ctx_stream(pcsf_ctx* c) { return c ? reinterpret_cast<void*>(c->stream) : nullptr; }

int pcsf_ctx_synchronize(pcsf_ctx* c) {
  if (!c) return PCSF_EINVAL;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PCSF_OK;
}

int pcsf_program(int n, const int32_t* children, int with_tables, int32_t* ops, int cap,
                 int* n_ops, int* slots) {
  if (n < 3 || n % 2 == 0 || !children || !n_ops) return PCSF_EINVAL;
  const int nl = (n + 1) / 2;
  std::vector<int> lc(n, -1), rc(n, -1), par(n, -1);
  for (int i = nl; i < n; ++i) {
    const int a = children[(i - nl) * 2], b = children[(i - nl) * 2 + 1];
    if (a < 0 || b < 0 || a >= i || b >= i || a == b || par[a] >= 0 || par[b] >= 0)
      return PCSF_EINVAL;
    lc[i] = a;
    rc[i] = b;
    par[a] = par[b] = i;
  }
  pcsf::Program pg = pcsf::make_program(lc, rc, nl, with_tables != 0);
  *n_ops = static_cast<int>(pg.ops.size() / kOpWords);
  if (slots) *slots = pg.slots;
  if (ops) {
    if (cap < (int)pg.ops.size()) return PCSF_EINVAL;
    std::copy(pg.ops.begin(), pg.ops.end(), ops);
  }
  return PCSF_OK;
}

int pcsf_ecm_read(const char* path, double* S, double* pi, char* err, int errlen) {
  try {
    pcsf::read_ecm(path, S, pi);
  } catch (const pcsf::ParseError& e) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", e.msg.c_str());
    return e.code;
  }
  return PCSF_OK;
}

int pcsf_newick_tree(const char* text, const char* keep_csv, int max_nodes, int* n_out,
                     int32_t* children, double* branch, char* labels_buf, int labels_len,
                     char* err, int errlen) {
  try {
    auto t = pcsf::parse_newick(text);
    if (keep_csv) {
      std::vector<std::string> keep;
      std::string s(keep_csv);
      size_t i = 0;
      while (i <= s.size()) {
        size_t j = s.find(',', i);
        if (j == std::string::npos) j = s.size();
        keep.push_back(s.substr(i, j - i));
        i = j + 1;
      }
      t = pcsf::newick_subtree(std::move(t), keep);
      if (!t || pcsf::newick_leaves(*t) <= 1)
        throw pcsf::ParseError{PCSF_EFAIL, "specify at least two available --species"};
    }
    pcsf::Tree T = pcsf::tree_of_newick(*t);
    if (T.n > max_nodes) throw pcsf::ParseError{PCSF_EINVAL, "tree larger than max_nodes"};
    *n_out = T.n;
    for (int i = T.nl; i < T.n; ++i) {
      children[(i - T.nl) * 2] = T.lc[i];
      children[(i - T.nl) * 2 + 1] = T.rc[i];
    }
    for (int i = 0; i < T.n; ++i) branch[i] = T.branch[i];
    if (labels_buf && labels_len > 0) {
      std::string all;
      for (int i = 0; i < T.n; ++i) {
        if (i) all += '\n';
        all += T.labels[i];
      }
      if (static_cast<int>(all.size()) >= labels_len)
        throw pcsf::ParseError{PCSF_EINVAL, "labels buffer too small"};
      std::memcpy(labels_buf, all.c_str(), all.size() + 1);
    }
  } catch (const pcsf::ParseError& e) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", e.msg.c_str());
    return e.code;
  }
  return PCSF_OK;
}

int pcsf_model_create(pcsf_ctx* c, int n, const int32_t* children, const double* branch,
                      int nm, const double* S, const double* pi, pcsf_model** out) {
  if (!out) return PCSF_EINVAL;
  *out = nullptr;
  if (n < 3 || n % 2 == 0 || !children || !branch || nm < 0 || nm > 8 ||
      (nm > 0 && (!S || !pi)))
    return set_err(c, PCSF_EINVAL, "CamlPaml.PhyloModel.make: bad tree or model count");
  // no rate models: a tree for models carried per set (the omega test's two
  // hypotheses per region)
  const bool tree_only = nm == 0;
  const int nrm = nm;  // rate models given
  if (tree_only) nm = 2;
  const int nl = (n + 1) / 2;
  if (n - 1 >= kNoBranch) return set_err(c, PCSF_EINVAL, "tree too large");
  auto m = new (std::nothrow) pcsf_model();
  if (!m) return PCSF_ENOMEM;
  std::unique_ptr<pcsf_model, void (*)(pcsf_model*)> guard(m, pcsf_model_destroy);
  m->ctx = c;
  m->n = n;
  m->nl = nl;
  m->nm = nm;
  m->nbr = n - 1;
  m->tree_only = tree_only;
  m->lc.assign(n, -1);
  m->rc.assign(n, -1);
  m->parent.assign(n, -1);
  for (int i = nl; i < n; ++i) {
    const int a = children[(i - nl) * 2], b = children[(i - nl) * 2 + 1];
    if (a < 0 || b < 0 || a >= i || b >= i || a == b || m->parent[a] >= 0 || m->parent[b] >= 0)
      return set_err(c, PCSF_EINVAL, "CamlPaml.T: children must precede parents (T order)");
    m->lc[i] = a;
    m->rc[i] = b;
    m->parent[a] = i;
    m->parent[b] = i;
  }
  for (int i = 0; i < n - 1; ++i)
    if (m->parent[i] < 0) return set_err(c, PCSF_EINVAL, "CamlPaml.T: orphan node");
  m->branch.assign(branch, branch + n - 1);
  for (double bl : m->branch)
    if (!(bl >= 0.0) || !std::isfinite(bl))
      return set_err(c, PCSF_EINVAL, "CamlPaml.PhyloModel.make: negative or missing branch");
  // The eigensystems are those of reversible models (ratematrix.cpp), which
  // an ECM's symmetric exchangeabilities give; a non-symmetric S would need
  // the reference's complex path (lib/CamlPaml/Q.ml:124-141,218-223), which
  // this build does not carry, so it is refused rather than symmetrised.
  for (int q = 0; q < nrm; ++q)
    for (int i = 0; i < 64; ++i)
      for (int j = 0; j < i; ++j) {
        const double a = S[q * 4096 + i * 64 + j], b = S[q * 4096 + j * 64 + i];
        if (!(std::fabs(a - b) <= 1e-12 * std::max(std::fabs(a), std::fabs(b)))) {
          char buf[192];
          std::snprintf(buf, sizeof buf,
                        "CamlPaml.PhyloModel.make: model %d is not reversible (S[%d,%d] = %g, "
                        "S[%d,%d] = %g); only symmetric exchangeabilities are supported",
                        q, i, j, a, j, i, b);
          return set_err(c, PCSF_EINVAL, buf);
        }
      }
  m->eig.resize(nrm);
  m->cached_rho.reserve(nm);
  if (nrm) m->pi.assign(pi, pi + 64 * nrm);
  std::vector<double> V(nrm * 4096), Vi(nrm * 4096), lam(nrm * 64), lnpi(nrm * 64),
      QT(nrm * 4096);
  for (int q = 0; q < nrm; ++q) {
    std::vector<double> Q(4096);
    std::string msg;
    if (!pcsf::build_rate_matrix(64, S + q * 4096, pi + q * 64, Q.data(), &msg))
      return set_err(c, PCSF_EFAIL, msg);
    if (!pcsf::eigen_reversible(64, Q.data(), pi + q * 64, &m->eig[q], &msg))
      return set_err(c, PCSF_EINVAL, msg);
    for (int a = 0; a < 64; ++a)
      for (int k = 0; k < 64; ++k) QT[q * 4096 + k * 64 + a] = Q[a * 64 + k];
    for (int i = 0; i < 64; ++i)
      for (int k = 0; k < 64; ++k) V[q * 4096 + k * 64 + i] = m->eig[q].V[i * 64 + k];
    std::copy(m->eig[q].Vinv.begin(), m->eig[q].Vinv.end(), Vi.begin() + q * 4096);
    std::copy(m->eig[q].lam.begin(), m->eig[q].lam.end(), lam.begin() + q * 64);
    for (int x = 0; x < 64; ++x) lnpi[q * 64 + x] = std::log(pi[q * 64 + x]);
  }
  const pcsf::Program pg = pcsf::make_program(m->lc, m->rc, nl, false);
  const pcsf::Program pt = pcsf::make_program(m->lc, m->rc, nl, true);
  m->prog = pg.ops;
  m->prog_t = pt.ops;
  m->cherries = pt.cherries;
  m->first_gemm_br_t = pt.first_gemm_br;
  m->first_gemm_br = pg.first_gemm_br;
  m->slots = std::max(pg.slots, pt.slots);
  m->n_gemm = pg.n_gemm;
  m->tree_p = find_tree_program(m->prog, false);
  m->tree_t = find_tree_program(m->prog_t, true);
  // default launch shape until pcsf_autotune: three waves per SIMD for the
  // straight-line programs (the fastest shape on every shipped tree measured),
  // two for the interpreter, whose switch needs more registers
  if (m->tree_p && m->tree_t)
    for (int& v : m->mfma_variant) v = 2;
  if (m->slots > kMaxSlots)
    return set_err(c, PCSF_EINVAL, "tree needs more than 6 live partial tiles");
  if (!c) {  // host-only model: simulation and model facts, no device work
    *out = guard.release();
    return PCSF_OK;
  }

  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, dalloc(&m->d_VT, V.size()));
  HIP_TRY(c, dalloc(&m->d_Vinv, Vi.size()));
  HIP_TRY(c, dalloc(&m->d_lam, lam.size()));
  HIP_TRY(c, dalloc(&m->d_pi, m->pi.size()));
  HIP_TRY(c, dalloc(&m->d_lnpi, lnpi.size()));
  HIP_TRY(c, dalloc(&m->d_QT, QT.size()));
  HIP_TRY(c, dalloc(&m->d_branch, m->branch.size()));
  HIP_TRY(c, dalloc(&m->d_lc, (size_t)n));
  HIP_TRY(c, dalloc(&m->d_rc, (size_t)n));
  HIP_TRY(c, dalloc(&m->d_prog, m->prog.size()));
  HIP_TRY(c, dalloc(&m->d_prog_t, m->prog_t.size()));
  HIP_TRY(c, dalloc(&m->d_cherries, std::max<size_t>(1, m->cherries.size())));
  // leaves each program gathers (CHERRY operands, LEAFMUL operands)
  auto gathered = [&](const std::vector<int32_t>& ops) {
    std::vector<int16_t> out;
    for (size_t k = 0; k + kOpWords <= ops.size(); k += kOpWords) {
      const int op = ops[k] & 0xff, a0 = ops[k + 1] & 0xffff, a1 = (ops[k + 1] >> 16) & 0xffff;
      if (op == OP_CHERRY) {
        out.push_back((int16_t)a0);
        out.push_back((int16_t)a1);
      } else if (op == OP_LEAFMUL) {
        out.push_back((int16_t)a0);
      }
    }
    return out;
  };
  const std::vector<int16_t> lsl = gathered(m->prog), lsl_t = gathered(m->prog_t);
  m->n_lsl = (int)lsl.size();
  m->n_lsl_t = (int)lsl_t.size();
  HIP_TRY(c, dalloc(&m->d_lsl, std::max<size_t>(1, lsl.size())));
  HIP_TRY(c, dalloc(&m->d_lsl_t, std::max<size_t>(1, lsl_t.size())));
  HIP_TRY(c, dalloc(&m->d_perr, (size_t)1));
  if (!tree_only)
    HIP_TRY(c, alloc_psets(m->fixed, nm, m->nbr, true, (int)m->cherries.size() / 3));
  std::vector<int32_t> lc32(m->lc.begin(), m->lc.end()), rc32(m->rc.begin(), m->rc.end());
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
  };
  HIP_TRY(c, up(m->d_VT, V.data(), V.size() * 8));
  HIP_TRY(c, up(m->d_Vinv, Vi.data(), Vi.size() * 8));
  HIP_TRY(c, up(m->d_lam, lam.data(), lam.size() * 8));
  HIP_TRY(c, up(m->d_pi, m->pi.data(), m->pi.size() * 8));
  HIP_TRY(c, up(m->d_lnpi, lnpi.data(), lnpi.size() * 8));
  HIP_TRY(c, up(m->d_QT, QT.data(), QT.size() * 8));
  HIP_TRY(c, up(m->d_branch, m->branch.data(), m->branch.size() * 8));
  HIP_TRY(c, up(m->d_lc, lc32.data(), n * 4));
  HIP_TRY(c, up(m->d_rc, rc32.data(), n * 4));
  HIP_TRY(c, up(m->d_prog, m->prog.data(), m->prog.size() * 4));
  HIP_TRY(c, up(m->d_prog_t, m->prog_t.data(), m->prog_t.size() * 4));
  if (!m->cherries.empty())
    HIP_TRY(c, up(m->d_cherries, m->cherries.data(), m->cherries.size() * 4));
  if (!lsl.empty()) HIP_TRY(c, up(m->d_lsl, lsl.data(), lsl.size() * 2));
  if (!lsl_t.empty()) HIP_TRY(c, up(m->d_lsl_t, lsl_t.data(), lsl_t.size() * 2));
  *out = guard.release();
  return PCSF_OK;
}

void pcsf_model_destroy(pcsf_model* m) {
  if (!m) return;
  dfree(m->d_VT);
  dfree(m->d_Vinv);
  dfree(m->d_lam);
  dfree(m->d_pi);
  dfree(m->d_lnpi);
  dfree(m->d_QT);
  dfree(m->d_branch);
  dfree(m->d_lc);
  dfree(m->d_rc);
  dfree(m->d_prog);
  dfree(m->d_prog_t);
  dfree(m->d_cherries);
  dfree(m->d_lsl);
  dfree(m->d_lsl_t);
  dfree(m->d_perr);
  free_psets(m->fixed);
  delete m;
}

int pcsf_model_info(const pcsf_model* m, int* n, int* nl, int* nm, int* slots) {
  if (!m) return PCSF_EINVAL;
  if (n) *n = m->n;
  if (nl) *nl = m->nl;
  if (nm) *nm = m->nm;
  if (slots) *slots = m->slots;
  return PCSF_OK;
}

int pcsf_model_transition(pcsf_model* m, int model, double rho, double* P_out) {
  if (!m || !P_out || !m->ctx) return PCSF_EINVAL;
  pcsf_ctx* c = m->ctx;
  if (m->tree_only) return set_err(c, PCSF_EINVAL, kTreeOnly);
  if (model < 0 || model >= m->nm) return set_err(c, PCSF_EINVAL, "model index out of range");
  if (!rho_valid(rho))
    return set_err(c, PCSF_EINVAL, "CamlPaml.P14n.instantiate_tree: domain violation on variable 0");
  std::vector<double> r(m->nm, rho);
  int rc = ensure_transitions(m, r.data(), m->fixed.fmt < 0 ? kFragF32 : m->fixed.fmt, nullptr);
  if (rc) return rc;
  std::vector<double> T((size_t)m->nbr * 4096);
  HIP_TRY(c, hipMemcpy(T.data(), m->fixed.d_P64T + (size_t)model * m->nbr * 4096, T.size() * 8,
                       hipMemcpyDeviceToHost));
  for (int br = 0; br < m->nbr; ++br)
    for (int a = 0; a < 64; ++a)
      for (int b = 0; b < 64; ++b)
        P_out[(size_t)br * 4096 + a * 64 + b] = T[(size_t)br * 4096 + b * 64 + a];
  return PCSF_OK;
}

int pcsf_batch_create(pcsf_model* m, const uint8_t* codes, int64_t ncols,
                      const int64_t* off, int32_t nreg, pcsf_batch** out) {
  if (!m || !out || !m->ctx) return PCSF_EINVAL;
  pcsf_ctx* c = m->ctx;
  *out = nullptr;
  if (ncols < 0 || nreg < 0 || (ncols > 0 && !codes) || !off)
    return set_err(c, PCSF_EINVAL, "pcsf_batch_create: bad arguments");
  if (off[0] != 0 || off[nreg] != ncols)
    return set_err(c, PCSF_EINVAL, "pcsf_batch_create: region offsets must span [0, n_cols]");
  for (int32_t r = 0; r < nreg; ++r)
    if (off[r + 1] < off[r])
      return set_err(c, PCSF_EINVAL, "pcsf_batch_create: region offsets must not decrease");
  auto b = new (std::nothrow) pcsf_batch();
  if (!b) return PCSF_ENOMEM;
  std::unique_ptr<pcsf_batch, void (*)(pcsf_batch*)> guard(b, pcsf_batch_destroy);
  b->m = m;
  b->ncols = ncols;
  b->nreg = nreg;
  b->offsets.assign(off, off + nreg + 1);
  b->ld = std::max<int64_t>(kGroupCols, (ncols + kGroupCols - 1) / kGroupCols * kGroupCols);
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, dalloc(&b->d_codes, (size_t)m->nl * b->ld));
  HIP_TRY(c, dalloc(&b->d_offsets, (size_t)nreg + 1));
  HIP_TRY(c, dalloc(&b->d_logz, (size_t)m->nm * b->ld));
  HIP_TRY(c, dalloc(&b->d_colelpr, (size_t)m->nm * b->ld));
  HIP_TRY(c, dalloc(&b->d_loglik, std::max<size_t>(1, (size_t)nreg * m->nm)));
  HIP_TRY(c, dalloc(&b->d_elpr, std::max<size_t>(1, (size_t)nreg * m->nm)));
  HIP_TRY(c, dalloc(&b->d_gloglik, std::max<size_t>(1, (size_t)nreg * m->nm)));
  HIP_TRY(c, dalloc(&b->d_grad, std::max<size_t>(1, (size_t)nreg * m->nm)));
  HIP_TRY(c, dalloc(&b->d_gsq, std::max<size_t>(1, (size_t)nreg * m->nm)));
  HIP_TRY(c, hipMemcpy(b->d_offsets, off, sizeof(int64_t) * (nreg + 1), hipMemcpyHostToDevice));
  // upload through a staging buffer (with a flag word for invalid codes at
  // its end), then make leaf-major
  uint8_t* stage = nullptr;
  const size_t bytes = (size_t)ncols * m->nl;
  const size_t flag_at = (bytes + 3) / 4 * 4;
  HIP_TRY(c, dalloc(&stage, flag_at + 4));
  int* d_bad = reinterpret_cast<int*>(stage + flag_at);
  int bad = 0;
  hipError_t e = hipMemsetAsync(d_bad, 0, 4, c->stream);
  if (e == hipSuccess && bytes) e = hipMemcpyAsync(stage, codes, bytes, hipMemcpyHostToDevice,
                                                    c->stream);
  if (e == hipSuccess)
    e = launch_transpose_codes(stage, ncols, m->nl, b->d_codes, b->ld, d_bad, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(stage);
  HIP_TRY(c, e);
  if (bad)
    return set_err(c, PCSF_EINVAL,
                   "pcsf_batch_create: a leaf code is neither a codon index (0..63) nor "
                   "PCSF_MARGINALIZE (255)");
  for (auto& ev : b->ev) HIP_TRY(c, hipEventCreate(&ev));
  *out = guard.release();
  return PCSF_OK;
}

int pcsf_batch_create_frames(pcsf_model* m, const uint8_t* nuc, int64_t n_nuc,
                             const int64_t* roff, int32_t nreg, int frames, pcsf_batch** out) {
  if (!m || !out || !m->ctx) return PCSF_EINVAL;
  pcsf_ctx* c = m->ctx;
  *out = nullptr;
  if (n_nuc < 0 || nreg < 0 || (n_nuc > 0 && !nuc) || !roff ||
      !(frames == 1 || frames == 3 || frames == 6))
    return set_err(c, PCSF_EINVAL, "pcsf_batch_create_frames: bad arguments");
  if (roff[0] != 0 || roff[nreg] != n_nuc)
    return set_err(c, PCSF_EINVAL,
                   "pcsf_batch_create_frames: region offsets must span [0, n_nuc]");
  std::vector<int64_t> foff((size_t)nreg * frames + 1, 0);
  for (int32_t r = 0; r < nreg; ++r) {
    const int64_t len = roff[r + 1] - roff[r];
    if (len < 0)
      return set_err(c, PCSF_EINVAL,
                     "pcsf_batch_create_frames: region offsets must not decrease");
    for (int f = 0; f < frames; ++f) {
      const int64_t k = (size_t)r * frames + f, rest = len - f % 3;
      foff[k + 1] = foff[k] + (rest >= 3 ? rest / 3 : 0);
    }
  }
  const int64_t ncols = foff.back();
  const int32_t nbr = (int32_t)(foff.size() - 1);
  auto b = new (std::nothrow) pcsf_batch();
  if (!b) return PCSF_ENOMEM;
  std::unique_ptr<pcsf_batch, void (*)(pcsf_batch*)> guard(b, pcsf_batch_destroy);
  b->m = m;
  b->ncols = ncols;
  b->nreg = nbr;
  b->offsets = foff;
  b->ld = std::max<int64_t>(kGroupCols, (ncols + kGroupCols - 1) / kGroupCols * kGroupCols);
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, dalloc(&b->d_codes, (size_t)m->nl * b->ld));
  HIP_TRY(c, dalloc(&b->d_offsets, (size_t)nbr + 1));
  HIP_TRY(c, dalloc(&b->d_logz, (size_t)m->nm * b->ld));
  HIP_TRY(c, dalloc(&b->d_colelpr, (size_t)m->nm * b->ld));
  HIP_TRY(c, dalloc(&b->d_loglik, std::max<size_t>(1, (size_t)nbr * m->nm)));
  HIP_TRY(c, dalloc(&b->d_elpr, std::max<size_t>(1, (size_t)nbr * m->nm)));
  HIP_TRY(c, dalloc(&b->d_gloglik, std::max<size_t>(1, (size_t)nbr * m->nm)));
  HIP_TRY(c, dalloc(&b->d_grad, std::max<size_t>(1, (size_t)nbr * m->nm)));
  HIP_TRY(c, dalloc(&b->d_gsq, std::max<size_t>(1, (size_t)nbr * m->nm)));
  HIP_TRY(c, hipMemcpy(b->d_offsets, foff.data(), sizeof(int64_t) * foff.size(),
                       hipMemcpyHostToDevice));
  uint8_t* d_nuc = nullptr;
  int64_t* d_roff = nullptr;
  const size_t bytes = (size_t)n_nuc * m->nl;
  hipError_t e = dalloc(&d_nuc, std::max<size_t>(bytes, 1));
  if (e == hipSuccess) e = dalloc(&d_roff, (size_t)nreg + 1);
  if (e == hipSuccess && bytes) e = hipMemcpy(d_nuc, nuc, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d_roff, roff, sizeof(int64_t) * (nreg + 1), hipMemcpyHostToDevice);
  // padding columns stay `Marginalize
  if (e == hipSuccess)
    e = hipMemsetAsync(b->d_codes, PCSF_MARGINALIZE, (size_t)m->nl * b->ld, c->stream);
  if (e == hipSuccess)
    e = launch_frames(d_nuc, n_nuc, d_roff, b->d_offsets, nreg, frames, m->nl, b->d_codes, b->ld,
                      c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dfree(d_nuc);
  dfree(d_roff);
  HIP_TRY(c, e);
  for (auto& ev : b->ev) HIP_TRY(c, hipEventCreate(&ev));
  *out = guard.release();
  return PCSF_OK;
}

void pcsf_batch_destroy(pcsf_batch* b) {
  if (!b) return;
  dfree(b->d_codes);
  dfree(b->d_offsets);
  dfree(b->d_logz);
  dfree(b->d_colelpr);
  dfree(b->d_loglik);
  dfree(b->d_elpr);
  dfree(b->al.d_codes);
  dfree(b->al.d_block_region);
  dfree(b->al.d_astart);
  dfree(b->d_alogz);
  dfree(b->d_acolelpr);
  dfree(b->d_gloglik);
  dfree(b->d_grad);
  dfree(b->d_gsq);
  dfree(b->d_adj_scratch);
  dfree(b->d_setqs);
  dfree(b->d_eigmode);
  dfree(b->d_eigref);
  free_psets(b->rs);
  for (auto& ev : b->ev)
    if (ev) (void)hipEventDestroy(ev);
  delete b;
}

int64_t pcsf_batch_columns(const pcsf_batch* b) { return b ? b->ncols : -1; }

int pcsf_batch_codes(pcsf_batch* b, uint8_t* codes, int64_t* region_offsets) {
  if (!b) return PCSF_EINVAL;
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  if (region_offsets) std::copy(b->offsets.begin(), b->offsets.end(), region_offsets);
  if (!codes || b->ncols == 0) return PCSF_OK;
  HIP_TRY(c, hipSetDevice(c->device));
  // leaf-major rows of ld bytes on the device -> n_cols x nl on the host
  std::vector<uint8_t> rows((size_t)m->nl * b->ncols);
  HIP_TRY(c, hipMemcpy2D(rows.data(), b->ncols, b->d_codes, b->ld, b->ncols, m->nl,
                         hipMemcpyDeviceToHost));
  for (int l = 0; l < m->nl; ++l)
    for (int64_t k = 0; k < b->ncols; ++k)
      codes[(size_t)k * m->nl + l] = rows[(size_t)l * b->ncols + k];
  return PCSF_OK;
}

int pcsf_forward(pcsf_batch* b, const double* rho, int rho_per_region, int precision,
                 double* loglik, double* elpr, int save_for_backward) {
  if (!b || !rho) return PCSF_EINVAL;
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  if (m->tree_only) return set_err(c, PCSF_EINVAL, kTreeOnly);
  (void)save_for_backward;
  if (!known_precision(precision))
    return set_err(c, PCSF_EINVAL, "pcsf_forward: unknown precision");
  if (rho_per_region) {
    const size_t nr = (size_t)b->nreg * m->nm;
    for (size_t k = 0; k < nr; ++k)
      if (!rho_valid(rho[k]))
        return set_err(c, PCSF_EINVAL,
                       "CamlPaml.P14n.instantiate_tree: domain violation on variable 0");
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = forward_regions(b, rho, nullptr, precision);
    if (rc) return rc;
    b->last_rho.assign(rho, rho + nr);
    b->last_precision = precision;
    if (loglik && nr)
      HIP_TRY(c, hipMemcpyAsync(loglik, b->d_loglik, nr * 8, hipMemcpyDeviceToHost, c->stream));
    if (elpr && nr)
      HIP_TRY(c, hipMemcpyAsync(elpr, b->d_elpr, nr * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return PCSF_OK;
  }
  std::vector<double> r(rho, rho + m->nm);
  for (double x : r)
    if (!rho_valid(x))
      return set_err(c, PCSF_EINVAL,
                     "CamlPaml.P14n.instantiate_tree: domain violation on variable 0");
  HIP_TRY(c, hipSetDevice(c->device));
  float tp = 0.f;
  const int fmt = fp64_precision(precision) ? (m->fixed.fmt < 0 ? kFragF32 : m->fixed.fmt)
                                            : frag_format(precision);
  int rc = ensure_transitions(m, r.data(), fmt, &tp);
  if (rc) return rc;
  if (precision == PCSF_FP64M && b->nreg > 0) {
    // region-aligned columns, the model's sets for every region
    if ((rc = ensure_aligned(b))) return rc;
  }
  b->t_pbuild = tp;
  b->last_rho.resize((size_t)b->nreg * m->nm);
  b->last_precision = precision;
  for (int32_t i = 0; i < b->nreg; ++i)
    for (int q = 0; q < m->nm; ++q) b->last_rho[(size_t)i * m->nm + q] = r[q];
  if (c->timing) HIP_TRY(c, hipEventRecord(b->ev[0], c->stream));
  hipError_t e;
  RegionChunk all;
  all.r0 = 0;
  all.nreg = b->nreg;
  all.block0 = 0;
  all.nblocks = b->al.ld / kAlignCols;
  all.sets = &m->fixed;
  if (precision == PCSF_FP64M) {
    e = b->nreg > 0 ? launch_forward_fp64m(m, b, all, c->stream) : hipSuccess;
  } else if (precision == PCSF_FP64) {
    e = launch_prune_fp64(m, b, c->stream);
    if (e == hipErrorInvalidValue)
      return set_err(c, PCSF_EINVAL, "FP64 kernel: tree workspace exceeds 64 KiB of LDS");
  } else if (frag_format(precision) >= 0) {
    e = launch_prune_mfma(m, b, precision, m->mfma_variant[precision], c->stream);
  } else {
    return set_err(c, PCSF_EINVAL, "pcsf_forward: unknown precision");
  }
  HIP_TRY(c, e);
  if (c->timing) HIP_TRY(c, hipEventRecord(b->ev[1], c->stream));
  if (precision == PCSF_FP64M) {
    if (b->nreg > 0) HIP_TRY(c, launch_region_reduce_aligned(b, all, b->d_loglik, b->d_elpr,
                                                              c->stream));
  } else {
    HIP_TRY(c, launch_region_reduce(b, c->stream));
  }
  if (c->timing) HIP_TRY(c, hipEventRecord(b->ev[2], c->stream));
  const size_t nres = (size_t)b->nreg * m->nm;
  if (loglik && nres)
    HIP_TRY(c, hipMemcpyAsync(loglik, b->d_loglik, nres * 8, hipMemcpyDeviceToHost, c->stream));
  if (elpr && nres)
    HIP_TRY(c, hipMemcpyAsync(elpr, b->d_elpr, nres * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (c->timing) {
    HIP_TRY(c, hipEventElapsedTime(&b->t_prune, b->ev[0], b->ev[1]));
    HIP_TRY(c, hipEventElapsedTime(&b->t_reduce, b->ev[1], b->ev[2]));
  }
  return PCSF_OK;
}

int pcsf_forward_noalloc(pcsf_model* m, const uint8_t* d_codes, int64_t n_cols,
                         const int64_t* d_region_offsets, int32_t n_regions, const double* rho,
                         int precision, double* d_loglik, double* d_elpr) {
  if (!m || !m->ctx || !rho) return PCSF_EINVAL;
  pcsf_ctx* c = m->ctx;
  if (m->tree_only) return set_err(c, PCSF_EINVAL, kTreeOnly);
  if (precision == PCSF_FP64M)
    return set_err(c, PCSF_EINVAL,
                   "pcsf_forward_noalloc: PCSF_FP64M reads P in a layout of its own and the "
                   "codes region-aligned, both in device memory it allocates; use PCSF_FP64 "
                   "here");
  if (!known_precision(precision))
    return set_err(c, PCSF_EINVAL, "pcsf_forward_noalloc: unknown precision");
  if (n_cols < 0 || n_regions < 0 || (n_cols > 0 && !d_codes) || !d_region_offsets ||
      (n_regions > 0 && (!d_loglik || !d_elpr)))
    return set_err(c, PCSF_EINVAL, "pcsf_forward_noalloc: bad arguments");
  for (int q = 0; q < m->nm; ++q)
    if (!rho_valid(rho[q]))
      return set_err(c, PCSF_EINVAL,
                     "CamlPaml.P14n.instantiate_tree: domain violation on variable 0");
  // scratch layout: leaf-major codes (nl x ld bytes), ln z and elpr per
  // column (nm x ld doubles each), status word
  constexpr size_t kCap = PCSF_NOALLOC_SCRATCH_BYTES;
  const size_t per_col = (size_t)m->nl + 16 * (size_t)m->nm;
  if ((size_t)n_cols > kCap / per_col) {
    char buf[200];
    std::snprintf(buf, sizeof buf,
                  "pcsf_forward_noalloc: %lld columns of a %d-leaf, %d-model tree exceed the "
                  "%zu-byte scratch (at most %zu columns)",
                  (long long)n_cols, m->nl, m->nm, kCap, kCap / per_col - kGroupCols);
    return set_err(c, PCSF_EINVAL, buf);
  }
  const int64_t ld = std::max<int64_t>(kGroupCols, (n_cols + kGroupCols - 1) / kGroupCols *
                                                       kGroupCols);
  const size_t codes_bytes = ((size_t)m->nl * ld + 255) / 256 * 256;
  const size_t col_bytes = (size_t)m->nm * ld * sizeof(double);
  if (codes_bytes + 2 * col_bytes + 256 > kCap) {
    char buf[200];
    std::snprintf(buf, sizeof buf,
                  "pcsf_forward_noalloc: %lld columns of a %d-leaf, %d-model tree exceed the "
                  "%zu-byte scratch (at most %zu columns)",
                  (long long)n_cols, m->nl, m->nm, kCap, kCap / per_col - kGroupCols);
    return set_err(c, PCSF_EINVAL, buf);
  }
  HIP_TRY(c, hipSetDevice(c->device));
  hipError_t se = hipSuccess;
  unsigned char* base = static_cast<unsigned char*>(scratch_address(c->device, &se));
  if (!base)
    return set_err(c, PCSF_EDEVICE,
                   std::string("pcsf_forward_noalloc: scratch lookup failed: ") +
                       hipGetErrorString(se));
  // transition sets at these scales: the model's own (allocated when it was made)
  const int fmt = precision == PCSF_FP64 ? (m->fixed.fmt < 0 ? kFragF32 : m->fixed.fmt)
                                         : frag_format(precision);
  int rc = ensure_transitions(m, rho, fmt, nullptr);
  if (rc) return rc;
  // a view of the scratch as a batch (owns nothing; its vectors stay empty)
  pcsf_batch v;
  v.m = m;
  v.ncols = n_cols;
  v.ld = ld;
  v.nreg = n_regions;
  v.d_codes = base;
  v.d_offsets = const_cast<int64_t*>(d_region_offsets);
  v.d_logz = reinterpret_cast<double*>(base + codes_bytes);
  v.d_colelpr = reinterpret_cast<double*>(base + codes_bytes + col_bytes);
  v.d_loglik = d_loglik;
  v.d_elpr = d_elpr;
  int* d_status = reinterpret_cast<int*>(base + codes_bytes + 2 * col_bytes);
  // leaf codes into the scratch, checked; region offsets checked
  HIP_TRY(c, hipMemsetAsync(d_status, 0, sizeof(int), c->stream));
  HIP_TRY(c, launch_transpose_codes(d_codes, n_cols, m->nl, v.d_codes, ld, d_status, c->stream));
  hipLaunchKernelGGL(noalloc_check_offsets_kernel, dim3((unsigned)((n_regions + 256) / 256)),
                     dim3(256), 0, c->stream, d_region_offsets, (int)n_regions, n_cols,
                     d_status);
  HIP_TRY(c, hipGetLastError());
  int status = 0;
  HIP_TRY(c, hipMemcpyAsync(&status, d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (status & 2)
    return set_err(c, PCSF_EINVAL,
                   "pcsf_forward_noalloc: region offsets must span [0, n_cols] and not decrease");
  if (status & 1)
    return set_err(c, PCSF_EINVAL,
                   "pcsf_forward_noalloc: a leaf code is neither a codon index (0..63) nor "
                   "PCSF_MARGINALIZE (255)");
  hipError_t e;
  if (precision == PCSF_FP64) {
    e = launch_prune_fp64(m, &v, c->stream);
    if (e == hipErrorInvalidValue)
      return set_err(c, PCSF_EINVAL, "FP64 kernel: tree workspace exceeds 64 KiB of LDS");
  } else {
    e = launch_prune_mfma(m, &v, precision, m->mfma_variant[precision], c->stream);
  }
  HIP_TRY(c, e);
  HIP_TRY(c, launch_region_reduce(&v, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PCSF_OK;
}

int pcsf_autotune(pcsf_batch* b, int precision, int reps, double* ms_out, int* chosen) {
  if (!b || reps < 1) return PCSF_EINVAL;
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  if (m->tree_only) return set_err(c, PCSF_EINVAL, kTreeOnly);
  if (precision == PCSF_FP64 || frag_format(precision) < 0)
    return set_err(c, PCSF_EINVAL, "pcsf_autotune: not an MFMA precision");
  HIP_TRY(c, hipSetDevice(c->device));
  {
    std::vector<double> one(m->nm, 1.0);
    int rc = ensure_transitions(m, one.data(), frag_format(precision), nullptr);
    if (rc) return rc;
  }
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  double best = 1e300;
  int best_v = m->mfma_variant[precision];
  hipError_t err = hipSuccess;
  for (int v = 0; v < kNumMfmaVariants && err == hipSuccess; ++v) {
    err = launch_prune_mfma(m, b, precision, v, c->stream);  // warm-up
    if (err == hipSuccess) err = hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps && err == hipSuccess; ++r)
      err = launch_prune_mfma(m, b, precision, v, c->stream);
    if (err == hipSuccess) err = hipEventRecord(e1, c->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    const double per = ms / reps;
    if (ms_out) ms_out[v] = per;
    if (err == hipSuccess && per < best) {
      best = per;
      best_v = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIP_TRY(c, err);
  m->mfma_variant[precision] = best_v;
  if (chosen) *chosen = best_v;
  return PCSF_OK;
}

int pcsf_model_specialized(pcsf_model* m, int enable, int* available) {
  if (!m) return PCSF_EINVAL;
  if (enable == 0 || enable == 1) m->specialized = enable == 1;
  if (available) *available = (m->tree_p ? 1 : 0) | (m->tree_t ? 2 : 0);
  return PCSF_OK;
}

int pcsf_set_variant(pcsf_model* m, int precision, int variant) {
  if (!m) return PCSF_EINVAL;
  if (precision == PCSF_FP64 || frag_format(precision) < 0)
    return set_err(m->ctx, PCSF_EINVAL, "pcsf_set_variant: not an MFMA precision");
  if (variant < 0 || variant >= kNumMfmaVariants)
    return set_err(m->ctx, PCSF_EINVAL, "pcsf_set_variant: no such launch variant");
  m->mfma_variant[precision] = variant;
  return PCSF_OK;
}

namespace {

struct PairState {
  enum Phase { RETRY, BRENT, FINAL, DONE } phase = RETRY;
  int retry = 0;         // retries done
  double flo, fhi;       // lpr at lo, hi
  double fx, ex, x;      // find_init's current point
  pcsf::BrentMin br;
  double pending = 0;    // abscissa to evaluate next
  double res_x = 0, res_lpr = 0, res_elpr = 0;
  int evals = 0;
  int status = PCSF_OK;
  // values already evaluated (x -> lpr, elpr), newest last
  std::vector<double> cx, cl, ce;
  void remember(double xx, double l, double e) {
    cx.push_back(xx);
    cl.push_back(l);
    ce.push_back(e);
  }
  bool recall(double xx, double* l, double* e) const {
    for (size_t k = cx.size(); k-- > 0;)
      if (cx[k] == xx) {
        *l = cl[k];
        *e = ce[k];
        return true;
      }
    return false;
  }
};

// Evaluates lpr (and elpr) of every region and model at one shared tree scale.
int eval_shared(pcsf_batch* b, double x, int precision, std::vector<double>& l,
                std::vector<double>& e) {
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  std::vector<double> r(m->nm, x);
  const size_t n = (size_t)b->nreg * m->nm;
  l.assign(n, 0.0);
  e.assign(n, 0.0);
  int rc = pcsf_forward(b, r.data(), 0, precision, l.data(), e.data(), 0);
  (void)c;
  return rc;
}

// The reference's maximize_lpr for every (region, model) pair at once
// (src/PhyloCSFModel.ml:84-99): Fit.find_init (init 1, [0.01, 10], log
// space, at most 250 tries), then GSL Brent on -lpr until the bracket is
// within 1% of the estimate, then lpr at the estimate.
int score_mle(pcsf_batch* b, int precision, std::vector<PairState>& st) {
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  const int nm = m->nm;
  const size_t np = (size_t)b->nreg * nm;
  const double lo = 1e-2, hi = 10.0, init = 1.0, accuracy = 0.01;
  std::vector<double> Llo, Elo, Lhi, Ehi, L1, E1, Lv, Ev;
  int rc;
  if ((rc = eval_shared(b, lo, precision, Llo, Elo))) return rc;
  if ((rc = eval_shared(b, hi, precision, Lhi, Ehi))) return rc;
  if ((rc = eval_shared(b, init, precision, L1, E1))) return rc;
  pcsf::BrentMin probe;
  probe.x_lo = lo;
  probe.x_hi = hi;
  const double v = probe.golden_point();
  if ((rc = eval_shared(b, v, precision, Lv, Ev))) return rc;
  const std::vector<double> retries = pcsf::find_init_retries(lo, hi, 250);
  st.assign(np, PairState());
  auto start_brent = [&](PairState& s, size_t k) {
    // gsl_min_fminimizer_set: lower, upper, minimum must be finite and the
    // endpoints must enclose the minimum of -lpr; then brent_init's f(v)
    const double fmin = -s.fx, flo = -s.flo, fhi = -s.fhi, fv = -Lv[k];
    if (!std::isfinite(fmin) || !std::isfinite(flo) || !std::isfinite(fhi) ||
        !std::isfinite(fv) || !(fmin < flo && fmin < fhi)) {
      s.status = PCSF_EFAIL;
      s.phase = PairState::DONE;
      return;
    }
    s.br.init(s.x, fmin, lo, flo, hi, fhi, fv);
    // With FP32 / FP16 likelihoods (relative rounding noise ~1e-8, i.e. ~1e-5
    // nats on a lpr of 1000) a step of sqrt(DBL_EPSILON)|x| changes lpr by
    // less than that noise, and GSL's comparison of f(x +- tol) with f(x)
    // would be decided by rounding; such steps are 1e-5 |x| instead, still
    // far inside the maximiser's 1% stopping interval.
    if (!fp64_precision(precision)) s.br.rel_tol = 1e-5;
    s.br.propose();
    s.pending = s.br.u;
    s.phase = PairState::BRENT;
  };
  auto finish_at = [&](PairState& s, double x) {
    double l, e;
    if (s.recall(x, &l, &e)) {
      s.res_x = x;
      s.res_lpr = l;
      s.res_elpr = e;
      s.phase = PairState::DONE;
    } else {
      s.pending = x;
      s.phase = PairState::FINAL;
    }
  };
  for (size_t k = 0; k < np; ++k) {
    PairState& s = st[k];
    s.flo = Llo[k];
    s.fhi = Lhi[k];
    s.x = init;
    s.fx = L1[k];
    s.ex = E1[k];
    s.evals = 4;
    s.remember(lo, Llo[k], Elo[k]);
    s.remember(hi, Lhi[k], Ehi[k]);
    s.remember(v, Lv[k], Ev[k]);
    s.remember(init, L1[k], E1[k]);
    if (!(s.fx <= s.flo || s.fx <= s.fhi)) {
      start_brent(s, k);
    } else {
      s.pending = retries[0];
      s.phase = PairState::RETRY;
    }
  }
  std::vector<double> rho(np, 1.0), L(np), E(np);
  std::vector<uint8_t> act(np);
  // PCSF_TRACE_MLE=<pair index r*nm+q>: every evaluation of that pair to stderr
  const char* tr = std::getenv("PCSF_TRACE_MLE");
  const long trace = tr ? std::strtol(tr, nullptr, 10) : -1;
  if (trace >= 0 && (size_t)trace < np)
    std::fprintf(stderr, "mle pair %ld: lo %.17g hi %.17g init %.17g v %.17g (%.17g) phase %d\n",
                 trace, Llo[trace], Lhi[trace], L1[trace], v, Lv[trace], (int)st[trace].phase);
  for (int round = 0; round < 100000; ++round) {
    bool any = false;
    for (size_t k = 0; k < np; ++k) {
      act[k] = st[k].phase != PairState::DONE;
      rho[k] = act[k] ? st[k].pending : 1.0;
      any = any || act[k];
    }
    if (!any) break;
    rc = forward_regions(b, rho.data(), act.data(), precision);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(L.data(), b->d_loglik, np * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(E.data(), b->d_elpr, np * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < np; ++k) {
      if (!act[k]) continue;
      PairState& s = st[k];
      const double x = s.pending, l = L[k], e = E[k];
      if ((long)k == trace)
        std::fprintf(stderr, "mle pair %ld round %d phase %d: x %.17g lpr %.17g\n", trace,
                     round, (int)s.phase, x, l);
      s.evals++;
      s.remember(x, l, e);
      switch (s.phase) {
        case PairState::RETRY: {
          s.x = x;
          s.fx = l;
          s.ex = e;
          s.retry++;
          // Fit.ml:36-47: after the 250th try the boundary is taken whatever it gave
          if (s.retry == 250) {
            finish_at(s, (s.flo > s.fhi) ? lo : hi);
          } else if (!(s.fx <= s.flo || s.fx <= s.fhi)) {
            start_brent(s, k);
          } else {
            s.pending = retries[s.retry];
          }
          break;
        }
        case PairState::BRENT: {
          const double fu = -l;
          if (!std::isfinite(fu)) {
            s.status = PCSF_EFAIL;
            s.phase = PairState::DONE;
            break;
          }
          s.br.accept(fu);
          if ((s.br.x_hi - s.br.x_lo) / s.br.x_min > accuracy) {
            s.br.propose();
            s.pending = s.br.u;
          } else {
            finish_at(s, s.br.x_min);
          }
          break;
        }
        case PairState::FINAL:
          s.res_x = x;
          s.res_lpr = l;
          s.res_elpr = e;
          s.phase = PairState::DONE;
          break;
        default:
          break;
      }
    }
  }
  if (!fp64_precision(precision)) {
    // the reference's final `f x` in FP64: the search ran on the MFMA
    // precision's likelihoods, the score is FP64's at its estimate
    for (size_t k = 0; k < np; ++k) {
      act[k] = st[k].status == PCSF_OK;
      rho[k] = act[k] ? st[k].res_x : 1.0;
    }
    rc = forward_regions(b, rho.data(), act.data(), PCSF_FP64M);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpy(L.data(), b->d_loglik, np * 8, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(E.data(), b->d_elpr, np * 8, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < np; ++k)
      if (act[k]) {
        st[k].res_lpr = L[k];
        st[k].res_elpr = E[k];
        st[k].evals++;
      }
  }
  return PCSF_OK;
}

// Gradient maximiser of lpr over the tree scale for every (region, model)
// pair: a safeguarded root search on h(x) = d lpr / d ln rho over
// [ln 0.01, ln 10] (the reference's bounds): from rho = 1 a first step of
// one unit of ln rho toward the side the gradient points to, then secant
// Newton steps, the boundary when the gradient keeps its sign there, and
// bisection when a step leaves the bracket or the bracket stops halving;
// done when a step is below tol.  Each step is one gradient launch giving
// lpr and d lpr / d rho; a final forward gives elpr.  st[k].evals counts the
// launches of pair k, the final forward included.
int score_mle_grad(pcsf_batch* b, int precision, std::vector<PairState>& st) {
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  const size_t np = (size_t)b->nreg * m->nm;
  const double xlo = std::log(1e-2), xhi = std::log(10.0);
  // ln rho tolerance: lpr is flat to second order at the maximum, so an
  // error e in ln rho costs ~e^2/2 |d2 lpr / d (ln rho)^2| -- 1e-12 and 1e-8
  // of it for the FP64 and the FP32-gradient tolerances.  FP32 gradients are
  // noisy at the 1e-3 level near the root (a few thousand nats of lpr), so
  // asking them for more than 1e-4 only buys bisection steps.
  const double tol = fp64_precision(precision) ? 1e-6 : 1e-4;
  // A pair also stops when the Newton step from its last point promises a
  // gain G = h^2 / (2 curv) that leaves nothing to measure, and takes that
  // step.  With the secant's curvature within a factor 1 +- d of lpr's the
  // step leaves ~d^2 G, d estimated by the change of the curvature between
  // the last two steps (the grid parabola's before the first secant; at
  // least 0.1): FP64 asks d^2 G <=
  // 1e-12 |lpr| (the oracle maximiser's digits), the MFMA precisions
  // d^2 G <= 2e-5 nats (1e-4 dB; their final lpr is FP64's at the estimate).
  const bool strict = fp64_precision(precision);
  // Per pair: first lpr on a grid of seven tree scales shared by every pair
  // (ln rho = -2 .. 2, through the fixed-scale path: one transition set per
  // model and grid point, as find_init's shared abscissae), the start at the
  // vertex of the parabola through the best grid point and its neighbours,
  // inside the bracket they span.  Then Newton steps on h = d lpr / d ln rho,
  // inside the bracket [lo, hi] that the signs of h leave for the maximum,
  // with h' from the last two evaluations (secant) where that is concave,
  // else from the parabola's curvature (first step) or the outer product of
  // the column scores, sum_col (d ln z / d ln rho)^2 (BHHH: the expected
  // information).  While the last two points are more than
  // 0.01 apart the step goes to the maximum of the cubic through their lpr
  // and h instead (Hermite interpolation: one order above the secant, and
  // the lpr differences still far above their rounding there).  Bisection
  // when a step leaves the bracket or two steps in a row fail to halve |h|.
  struct G {
    bool done = false, have_prev = false, seen_pos = false, seen_neg = false;
    double x = 0, px = 0, ph = 0, pL = 0;
    double pcurv = 0;  // the previous secant curvature (0: none yet)
    double icurv = 0;  // the grid parabola's curvature (0: none)
    double lo = 0, hi = 0;
    int slow = 0;  // consecutive steps that did not halve |h|
  };
  std::vector<G> g(np);
  for (G& s : g) {
    s.lo = xlo;
    s.hi = xhi;
  }
  st.assign(np, PairState());
  {
    constexpr int K = 7;
    const double grid[K] = {-2.0, -1.2, -0.6, 0.0, 0.6, 1.2, 2.0};
    std::vector<double> Lk[K], e;
    // FP64 evaluates the grid on the matrix cores (the plain kernel is the
    // slow one; both are FP64)
    const int gp = (precision == PCSF_FP64) ? PCSF_FP64M : precision;
    for (int k = 0; k < K; ++k) {
      const int rc = eval_shared(b, std::exp(grid[k]), gp, Lk[k], e);
      if (rc) return rc;
    }
    for (size_t q = 0; q < np; ++q) {
      st[q].evals += K;
      int best = 0;
      bool finite = true;
      for (int k = 0; k < K; ++k) {
        finite = finite && std::isfinite(Lk[k][q]);
        if (Lk[k][q] > Lk[best][q]) best = k;
      }
      if (!finite) continue;  // the search starts at rho = 1 over the whole domain
      G& s = g[q];
      if (best == 0) {
        s.x = grid[0];
        s.hi = grid[1];
      } else if (best == K - 1) {
        s.x = grid[K - 1];
        s.lo = grid[K - 2];
      } else {
        const double x0 = grid[best - 1], x1 = grid[best], x2 = grid[best + 1];
        const double d1 = (Lk[best][q] - Lk[best - 1][q]) / (x1 - x0);
        const double d2 = (Lk[best + 1][q] - Lk[best][q]) / (x2 - x1);
        const double a2 = (d2 - d1) / (x2 - x0);  // half the parabola's second derivative
        s.lo = x0;
        s.hi = x2;
        s.seen_pos = s.seen_neg = true;  // lpr rises into x1 and falls after it
        s.x = x1;
        if (a2 < 0.0) {
          double xv = 0.5 * (x0 + x1) - d1 / (2.0 * a2);
          double cv = -2.0 * a2;
          // with a fourth point (the better outer neighbour) the cubic's
          // maximum and curvature there: one order closer than the parabola's
          const int o = (best >= 2 && (best + 2 >= K || Lk[best - 2][q] >= Lk[best + 2][q]))
                            ? best - 2
                            : (best + 2 < K ? best + 2 : -1);
          if (o >= 0) {
            const double x3 = grid[o];
            const double d3 = (o < best) ? (Lk[best - 1][q] - Lk[o][q]) / (x0 - x3)
                                         : (Lk[o][q] - Lk[best + 1][q]) / (x3 - x2);
            // Newton form over x0 < x1 < x2 with x3 outside: p(x) = f0 + d1 (x - x0)
            // + a2 (x - x0)(x - x1) + a3 (x - x0)(x - x1)(x - x2)
            const double b2 = (o < best) ? (d1 - d3) / (x1 - x3) : (d3 - d2) / (x3 - x1);
            const double a3 = (o < best) ? (a2 - b2) / (x2 - x3) : (b2 - a2) / (x3 - x0);
            // p'(x) = d1 + a2 (2x - x0 - x1) + a3 q'(x), q = (x - x0)(x - x1)(x - x2)
            auto dp = [&](double x) {
              const double qd = (x - x1) * (x - x2) + (x - x0) * (x - x2) + (x - x0) * (x - x1);
              return d1 + a2 * (2.0 * x - x0 - x1) + a3 * qd;
            };
            auto d2p = [&](double x) { return 2.0 * a2 + a3 * (6.0 * x - 2.0 * (x0 + x1 + x2)); };
            // the root of p' in (x0, x2) by Newton from the parabola's vertex
            double xc = (xv > x0 && xv < x2) ? xv : x1;
            for (int it = 0; it < 8; ++it) {
              const double dd = d2p(xc);
              if (!(dd < 0.0)) break;
              xc = std::min(std::max(xc - dp(xc) / dd, x0), x2);
            }
            if (xc > x0 && xc < x2 && d2p(xc) < 0.0 && std::fabs(dp(xc)) < 1e-6 * std::fabs(cv)) {
              xv = xc;
              cv = -d2p(xc);
            }
          }
          if (xv > x0 && xv < x2) s.x = xv;
          // also the first reference for the secant curvature's change
          s.icurv = s.pcurv = cv;
        }
      }
    }
  }
  std::vector<double> rho(np, 1.0), L(np), D(np), S(np);
  std::vector<uint8_t> act(np, 1);
  auto eval = [&]() -> int {
    // FP64: the outside-pass adjoint; MFMA precisions: the FP32 tangent pass
    int rc = adjoint_regions(b, rho.data(), act.data(), nullptr, !fp64_precision(precision));
    if (rc) return rc;
    HIP_TRY(c, hipMemcpy(L.data(), b->d_gloglik, np * 8, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(D.data(), b->d_grad, np * 8, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(S.data(), b->d_gsq, np * 8, hipMemcpyDeviceToHost));
    return PCSF_OK;
  };
  for (int it = 0; it < 100; ++it) {
    bool any = false;
    for (size_t k = 0; k < np; ++k) {
      act[k] = !g[k].done;
      rho[k] = std::exp(g[k].x);
      any = any || act[k];
    }
    if (!any) break;
    int rc = eval();
    if (rc) return rc;
    for (size_t k = 0; k < np; ++k) {
      if (!act[k]) continue;
      G& s = g[k];
      st[k].evals++;
      const double h = std::exp(s.x) * D[k];  // d lpr / d ln rho
      if (!std::isfinite(h) || !std::isfinite(L[k])) {
        st[k].status = PCSF_EFAIL;
        s.done = true;
        continue;
      }
      // h = 0, or (FP32 gradients) below their rounding noise, ~1e-8 |lpr|:
      // the root is as close as these gradients can place it
      if (h == 0.0 || (!fp64_precision(precision) && std::fabs(h) <= 1e-7 * std::fabs(L[k]))) {
        s.done = true;
        continue;
      }
      if (h > 0) {
        s.seen_pos = true;
        s.lo = std::max(s.lo, s.x);
        if (s.x >= xhi) {  // increasing at the upper bound
          s.done = true;
          continue;
        }
      } else {
        s.seen_neg = true;
        s.hi = std::min(s.hi, s.x);
        if (s.x <= xlo) {
          s.done = true;
          continue;
        }
      }
      if (s.hi - s.lo <= tol) {
        s.x = 0.5 * (s.lo + s.hi);
        s.done = true;
        continue;
      }
      double curv = 0.0;  // -h' in ln rho
      bool secant = false;
      if (s.have_prev) {
        const double hp = (h - s.ph) / (s.x - s.px);
        if (hp < 0) {
          curv = -hp;
          secant = true;
        }
      }
      // the outer product of the column scores about their mean, n times
      // their variance: the mean's square, h^2 / n, would dominate the plain
      // sum away from the maximum
      const double n_cols = (double)(b->offsets[k / m->nm + 1] - b->offsets[k / m->nm]);
      const double info = std::exp(2.0 * s.x) * S[k] - h * h / std::max(1.0, n_cols);
      if (!secant && s.icurv > 0.0 && !s.have_prev) curv = s.icurv;
      else if (!secant && std::isfinite(info) && info > 0.0) curv = info;
      double xn;
      if (curv > 0.0) {
        xn = s.x + h / curv;
        if (secant) {
          const double d = std::max(0.1, s.pcurv > 0.0 ? std::fabs(curv - s.pcurv) / curv : 1.0);
          const double left = d * d * h * h / (2.0 * curv);  // what the step leaves
          s.pcurv = curv;
          if (left <= (strict ? 1e-12 * std::max(1.0, std::fabs(L[k])) : 2e-5)) {
            if (xn > s.lo && xn < s.hi) s.x = xn;
            s.done = true;
            continue;
          }
        }
        if (s.have_prev && std::fabs(s.x - s.px) > 1e-2) {
          // maximum of the cubic with lpr and h at px and x: Nocedal and
          // Wright's interpolation step (Numerical Optimization, eq. 3.59)
          // on phi = -lpr
          const double f0 = -s.pL, g0 = -s.ph, f1 = -L[k], g1 = -h;
          const double d1 = g0 + g1 - 3.0 * (f0 - f1) / (s.px - s.x);
          const double disc = d1 * d1 - g0 * g1;
          if (disc >= 0.0) {
            const double d2 = std::copysign(std::sqrt(disc), s.x - s.px);
            const double den = g1 - g0 + 2.0 * d2;
            const double xc = s.x - (s.x - s.px) * (g1 + d2 - d1) / den;
            if (den != 0.0 && std::isfinite(xc)) xn = xc;
          }
        }
        xn = std::min(std::max(xn, s.x - 2.0), s.x + 2.0);  // at most e^2 in rho
      } else if (s.have_prev) {
        xn = s.x + (h > 0 ? 2.0 : -2.0) * std::fabs(s.x - s.px);
      } else {
        xn = s.x + (h > 0 ? 1.0 : -1.0);  // a factor e in rho
      }
      const bool bracketed = s.seen_pos && s.seen_neg;
      // Newton steps that converge from one side leave the bracket wide, so
      // progress is judged by |h|: bisection after two steps in a row that
      // did not halve it
      s.slow = (s.have_prev && std::fabs(h) > 0.5 * std::fabs(s.ph)) ? s.slow + 1 : 0;
      if (bracketed && s.slow >= 2) {
        xn = 0.5 * (s.lo + s.hi);
        s.slow = 0;
      }
      if (!(xn > s.lo && xn < s.hi))
        xn = bracketed ? 0.5 * (s.lo + s.hi) : std::min(std::max(xn, s.lo), s.hi);
      if (std::fabs(xn - s.x) <= tol) {
        s.x = xn;
        s.done = true;
        continue;
      }
      s.px = s.x;
      s.ph = h;
      s.pL = L[k];
      s.have_prev = true;
      s.x = xn;
    }
  }
  // lpr and elpr at the estimates, in FP64 whatever the search ran in: the
  // MFMA precisions' rounding (up to ~6e-9 of lpr on 100vertebrates regions
  // of a few thousand codons, ~3e-3 dB) would otherwise reach the score
  for (size_t k = 0; k < np; ++k) rho[k] = std::exp(g[k].x);
  std::vector<uint8_t> all(np, 1);
  int rc = forward_regions(b, rho.data(), all.data(), PCSF_FP64M);
  if (rc) return rc;
  HIP_TRY(c, hipMemcpy(L.data(), b->d_loglik, np * 8, hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(D.data(), b->d_elpr, np * 8, hipMemcpyDeviceToHost));
  for (size_t k = 0; k < np; ++k) {
    st[k].res_x = rho[k];
    st[k].res_lpr = L[k];
    st[k].res_elpr = D[k];
    st[k].evals++;
  }
  return PCSF_OK;
}

// The omega test (src/OmegaModel.ml:160-219) for every region of a batch:
// per region, H0 (omega = sigma = 1) and then H1 (omega_H1, sigma_H1, from
// H0's estimates) are each fitted by kr_map, three rounds of coordinate
// ascent -- maximize_lpr over the tree scale in [0.001, 10], then over kappa
// in [1, 10] -- of the log prior plus lpr.  Every (region, hypothesis) pair
// is a state machine; all pairs advance in lock step, one batched
// per-region forward per step, with the rate model of each pair's current
// settings as the set's own (its eigensystem made on the host when kappa
// changes).
struct OmegaPair {
  double qs[pcsf::kOmegaSettings];
  double rho = 1.0, kappa = 2.5;
  int round = 0, phase = 0;  // phase 0: tree scale, 1: kappa, 2: done
  bool waiting = false;      // H1 before its H0 is done
  pcsf::MaximizeLpr mx;
  double lpr = 0;
  int status = PCSF_OK;
  int evals = 0;
};

int score_omega(pcsf_batch* b, int precision, double omega_h1, double sigma_h1,
                std::vector<OmegaPair>& st) {
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  const int R = b->nreg;
  const size_t np = (size_t)R * 2;
  const double rel_tol = fp64_precision(precision) ? 1.4901161193847656e-08 : 1e-5;
  // F3x4 settings per region, from the batch's codes
  std::vector<uint8_t> codes((size_t)b->ncols * m->nl);
  int rc = pcsf_batch_codes(b, codes.data(), nullptr);
  if (rc) return rc;
  st.assign(np, OmegaPair());
  for (int r = 0; r < R; ++r) {
    double f3[9];
    pcsf::omega_f3x4(codes.data() + (size_t)b->offsets[r] * m->nl,
                     b->offsets[r + 1] - b->offsets[r], m->nl, f3);
    for (int h = 0; h < 2; ++h) {
      OmegaPair& p = st[(size_t)r * 2 + h];
      p.qs[0] = 2.5;  // new_instance ~kappa:2.5 (OmegaModel.ml:197)
      p.qs[1] = 1.0;
      p.qs[2] = 1.0;
      std::copy(f3, f3 + 9, p.qs + 3);
      p.waiting = h == 1;
      if (h == 0) p.mx.start(p.rho, 0.001, 10.0, 0.01, rel_tol);
    }
  }
  std::vector<double> qs_all(np * pcsf::kOmegaSettings);
  std::vector<double> rho(np, 1.0), L(np);
  std::vector<uint8_t> act(np);
  SetModels sm{qs_all.data()};
  // the settings of pair k's rate model at kappa x (the GPU builds Q, its
  // eigensystem and the root prior from them); false where the reference's
  // rate-matrix construction would fail: a codon frequency or the mean rate
  // not positive
  auto settings = [&](size_t k, double x) -> bool {
    double* qs = qs_all.data() + k * pcsf::kOmegaSettings;
    std::copy(st[k].qs, st[k].qs + pcsf::kOmegaSettings, qs);
    qs[0] = x;
    double pi[64];
    pcsf::omega_pi(qs, pi);
    for (double v : pi)
      if (!(v > 0.0) || !std::isfinite(v)) return false;
    return x > 0.0 && qs[1] > 0.0;
  };
  for (int step = 0; step < 1000000; ++step) {
    bool any = false;
    for (size_t k = 0; k < np; ++k) {
      OmegaPair& p = st[k];
      act[k] = p.phase < 2 && !p.waiting && p.status == PCSF_OK;
      if (!act[k]) continue;
      const double x = p.mx.pending();
      rho[k] = p.phase == 0 ? x : p.rho;
      if (!settings(k, p.phase == 1 ? x : p.kappa)) {
        p.status = PCSF_EFAIL;
        act[k] = 0;
        continue;
      }
      any = true;
    }
    if (!any) break;
    rc = forward_regions(b, rho.data(), act.data(), precision, &sm);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpy(L.data(), b->d_loglik, np * 8, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < np; ++k) {
      if (!act[k]) continue;
      OmegaPair& p = st[k];
      const double x = p.mx.pending();
      // f_rho / f_kappa (OmegaModel.ml:161-170): log prior + lpr
      const double prior = p.phase == 0 ? pcsf::omega_lpr_rho(x) : pcsf::omega_lpr_kappa(x);
      p.evals++;
      p.mx.feed(prior + L[k]);
      if (!p.mx.done()) continue;
      if (p.mx.failed) {
        p.status = PCSF_EFAIL;
        continue;
      }
      if (p.phase == 0) {
        p.rho = p.mx.x;
        p.phase = 1;
        p.mx.start(p.kappa, 1.0, 10.0, 0.01, rel_tol);
        continue;
      }
      p.kappa = p.mx.x;
      p.qs[0] = p.kappa;
      p.lpr = p.mx.fx;
      if (++p.round < 3) {
        p.phase = 0;
        p.mx.start(p.rho, 0.001, 10.0, 0.01, rel_tol);
        continue;
      }
      p.phase = 2;
      if (k % 2 == 0) {  // H0 done: H1 starts from its estimates (OmegaModel.ml:202-206)
        OmegaPair& h1 = st[k + 1];
        std::copy(p.qs, p.qs + pcsf::kOmegaSettings, h1.qs);
        h1.qs[1] = omega_h1;
        h1.qs[2] = sigma_h1;
        h1.rho = p.rho;
        h1.kappa = p.kappa;
        h1.waiting = false;
        h1.mx.start(h1.rho, 0.001, 10.0, 0.01, rel_tol);
      }
    }
    for (size_t k = 0; k + 1 < np; k += 2)  // an H0 that failed takes its H1 along
      if (st[k].status != PCSF_OK && st[k + 1].waiting) {
        st[k + 1].status = st[k].status;
        st[k + 1].waiting = false;
      }
  }
  if (!fp64_precision(precision)) {
    // each hypothesis's lpr at its estimates in FP64 (the fits ran on the
    // MFMA precision's likelihoods): the last kappa step's f = lpr_kappa + lpr
    for (size_t k = 0; k < np; ++k) {
      act[k] = st[k].status == PCSF_OK;
      rho[k] = st[k].rho;
      if (act[k] && !settings(k, st[k].kappa)) {
        st[k].status = PCSF_EFAIL;
        act[k] = 0;
      }
    }
    rc = forward_regions(b, rho.data(), act.data(), PCSF_FP64M, &sm);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpy(L.data(), b->d_loglik, np * 8, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < np; ++k)
      if (act[k]) {
        st[k].lpr = pcsf::omega_lpr_kappa(st[k].kappa) + L[k];
        st[k].evals++;
      }
  }
  return PCSF_OK;
}

}  // namespace

int pcsf_score(pcsf_batch* b, int strategy, int precision, double* score, double* anc_comp,
               double* rho_out, double* lpr_out, double* elpr_out, int32_t* status,
               int32_t* n_evals) {
  if (!b) return PCSF_EINVAL;
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  if (m->tree_only) return set_err(c, PCSF_EINVAL, kTreeOnly);
  if (m->nm != 2)
    return set_err(c, PCSF_EINVAL, "pcsf_score: needs exactly two models (coding, noncoding)");
  const int R = b->nreg;
  const double db = 10.0 / std::log(10.0);
  std::vector<double> l(2 * (size_t)R), e(2 * (size_t)R), r(2 * (size_t)R, 1.0);
  std::vector<int32_t> stt(R, PCSF_OK), ev(R, 0);
  if (strategy == PCSF_STRATEGY_FIXED) {
    const double one[2] = {1.0, 1.0};
    int rc = pcsf_forward(b, one, 0, precision, l.data(), e.data(), 0);
    if (rc) return rc;
    for (int i = 0; i < R; ++i) ev[i] = 2;
  } else if (strategy == PCSF_STRATEGY_MLE || strategy == PCSF_STRATEGY_MLE_GRAD) {
    std::vector<PairState> st;
    int rc = (strategy == PCSF_STRATEGY_MLE) ? score_mle(b, precision, st)
                                             : score_mle_grad(b, precision, st);
    if (rc) return rc;
    for (int i = 0; i < R; ++i)
      for (int q = 0; q < 2; ++q) {
        const PairState& s = st[(size_t)i * 2 + q];
        l[i * 2 + q] = s.res_lpr;
        e[i * 2 + q] = s.res_elpr;
        r[i * 2 + q] = s.res_x;
        ev[i] += s.evals;
        if (s.status) stt[i] = s.status;
      }
  } else {
    return set_err(c, PCSF_EINVAL, "pcsf_score: unknown strategy");
  }
  for (int i = 0; i < R; ++i) {
    const bool ok = stt[i] == PCSF_OK;
    // score = db (lpr_c - lpr_n), anc_comp = db (elpr_c - elpr_n)
    // (src/PhyloCSFModel.ml:113,122-136)
    if (score) score[i] = ok ? db * (l[2 * i] - l[2 * i + 1]) : NAN;
    if (anc_comp) anc_comp[i] = ok ? db * (e[2 * i] - e[2 * i + 1]) : NAN;
    if (status) status[i] = stt[i];
    if (n_evals) n_evals[i] = ev[i];
  }
  if (rho_out) std::copy(r.begin(), r.end(), rho_out);
  if (lpr_out) std::copy(l.begin(), l.end(), lpr_out);
  if (elpr_out) std::copy(e.begin(), e.end(), elpr_out);
  return PCSF_OK;
}

int pcsf_score_omega(pcsf_batch* b, double omega_h1, double sigma_h1, int precision,
                     double* score, double* lpr, double* rho, double* kappa, int32_t* status,
                     int32_t* n_evals) {
  if (!b) return PCSF_EINVAL;
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  if (m->nm != 2)
    return set_err(c, PCSF_EINVAL, "pcsf_score_omega: needs a model with two hypothesis slots");
  if (!known_precision(precision))
    return set_err(c, PCSF_EINVAL, "pcsf_score_omega: unknown precision");
  if (!(omega_h1 >= 0.0) || !(sigma_h1 >= 0.0) || !std::isfinite(omega_h1) ||
      !std::isfinite(sigma_h1))
    return set_err(c, PCSF_EINVAL,
                   "CamlPaml.P14n.instantiate_q: domain violation on variable 1");
  HIP_TRY(c, hipSetDevice(c->device));
  std::vector<OmegaPair> st;
  int rc = score_omega(b, precision, omega_h1, sigma_h1, st);
  if (rc) return rc;
  const double db = 10.0 / std::log(10.0);
  for (int r = 0; r < b->nreg; ++r) {
    const OmegaPair &h0 = st[(size_t)r * 2], &h1 = st[(size_t)r * 2 + 1];
    const int s = h0.status ? h0.status : h1.status;
    if (score) score[r] = s ? NAN : db * (h1.lpr - h0.lpr);
    for (int h = 0; h < 2; ++h) {
      const OmegaPair& p = st[(size_t)r * 2 + h];
      if (lpr) lpr[r * 2 + h] = s ? NAN : p.lpr;
      if (rho) rho[r * 2 + h] = p.rho;
      if (kappa) kappa[r * 2 + h] = p.kappa;
    }
    if (status) status[r] = s;
    if (n_evals) n_evals[r] = h0.evals + h1.evals;
  }
  return PCSF_OK;
}

int pcsf_backward(pcsf_batch* b, const double* grad_loglik, double* grad_rho,
                  double* grad_branch) {
  if (!b || !grad_loglik) return PCSF_EINVAL;
  pcsf_model* m = b->m;
  pcsf_ctx* c = m->ctx;
  if (m->tree_only) return set_err(c, PCSF_EINVAL, kTreeOnly);
  const size_t np = (size_t)b->nreg * m->nm;
  if (b->last_rho.size() != np)
    return set_err(c, PCSF_EINVAL, "pcsf_backward: no forward on this batch to differentiate");
  HIP_TRY(c, hipSetDevice(c->device));
  double* d_gbr = nullptr;
  const size_t ngb = np * m->nbr;
  if (grad_branch && ngb) {
    HIP_TRY(c, dalloc(&d_gbr, ngb));
    HIP_TRY(c, hipMemsetAsync(d_gbr, 0, ngb * 8, c->stream));
  }
  // an MFMA-precision forward without per-branch gradients is differentiated
  // by the FP32 tangent pass; everything else by the FP64 outside pass
  const bool tangent = !fp64_precision(b->last_precision) && !d_gbr;
  int rc = adjoint_regions(b, b->last_rho.data(), nullptr, d_gbr, tangent);
  std::vector<double> dl(np), gb;
  if (!rc && np) {
    hipError_t e = hipMemcpy(dl.data(), b->d_grad, np * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && d_gbr) {
      gb.resize(ngb);
      e = hipMemcpy(gb.data(), d_gbr, ngb * 8, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) rc = set_err(c, PCSF_EDEVICE, hipGetErrorString(e));
  }
  dfree(d_gbr);
  if (rc) return rc;
  for (size_t k = 0; k < np; ++k) {
    if (grad_rho) grad_rho[k] = grad_loglik[k] * dl[k];
    if (grad_branch)
      for (int br = 0; br < m->nbr; ++br)
        grad_branch[k * m->nbr + br] = grad_loglik[k] * gb[k * m->nbr + br];
  }
  return PCSF_OK;
}

int pcsf_batch_device_results(pcsf_batch* b, const double** d_loglik, const double** d_elpr) {
  if (!b) return PCSF_EINVAL;
  if (d_loglik) *d_loglik = b->d_loglik;
  if (d_elpr) *d_elpr = b->d_elpr;
  return PCSF_OK;
}

int pcsf_batch_last_timing(const pcsf_batch* b, double* prune_ms, double* reduce_ms,
                           double* pbuild_ms) {
  if (!b) return PCSF_EINVAL;
  if (prune_ms) *prune_ms = b->t_prune;
  if (reduce_ms) *reduce_ms = b->t_reduce;
  if (pbuild_ms) *pbuild_ms = b->t_pbuild;
  return PCSF_OK;
}

}  // extern "C"

namespace {

// PhyloLik.prepare + likelihood + node_posterior + add_branch_posteriors
// (lib/CamlPaml/PhyloLik.ml:46-174) over n_cols columns: the host side of
// kernels/lik.hip.  Columns go in chunks whose device buffers stay within
// ~512 MB; post_mode 1 = the root's posterior, 2 = every node's.
// The E-step for 64 states without `Distribution leaves on the FP64 matrix
// cores (kernels/estep.hip): columns in chunks of up to 32,768, u and v of
// every (branch, column) of a chunk in device memory, the counts added per
// chunk.
// The context's E-step buffers for chunks of `chunk` columns (grown, never
// shrunk; a failed growth leaves them released).
hipError_t estep_cache(pcsf_ctx* c, int n, int64_t ldu, int64_t chunk, int grid) {
  EstepCache& es = c->es;
  const int nl = (n + 1) / 2, nbr = n - 1;
  hipError_t e = hipSuccess;
  if (!es.copy) e = hipStreamCreateWithFlags(&es.copy, hipStreamNonBlocking);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) {
    if (!es.uploaded[i]) e = hipEventCreateWithFlags(&es.uploaded[i], hipEventDisableTiming);
    if (e == hipSuccess && !es.consumed[i])
      e = hipEventCreateWithFlags(&es.consumed[i], hipEventDisableTiming);
  }
  if (e != hipSuccess) return e;
  const size_t uv = (size_t)nbr * ldu * 64;
  if (es.cap_uv < uv) {
    dfree(es.U);
    dfree(es.V);
    es.cap_uv = 0;
    e = dalloc(&es.U, uv);
    if (e == hipSuccess) e = dalloc(&es.V, uv);
    if (e != hipSuccess) {
      dfree(es.U);
      dfree(es.V);
      return e;
    }
    es.cap_uv = uv;
  }
  const size_t scr = estep_scratch_doubles(n, grid);
  if (es.cap_scr < scr) {
    dfree(es.scr);
    es.cap_scr = 0;
    if ((e = dalloc(&es.scr, scr)) != hipSuccess) return e;
    es.cap_scr = scr;
  }
  const size_t lv = (size_t)chunk * nl;
  if (es.cap_leaves < lv) {
    for (int i = 0; i < 2; ++i) {
      dfree(es.leaves[i]);
      if (es.staged[i]) (void)hipHostFree(es.staged[i]);
      es.staged[i] = nullptr;
    }
    es.cap_leaves = 0;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      e = dalloc(&es.leaves[i], lv);
      if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&es.staged[i]), lv * 4, hipHostMallocDefault);
    }
    if (e != hipSuccess) {
      for (int i = 0; i < 2; ++i) {
        dfree(es.leaves[i]);
        if (es.staged[i]) (void)hipHostFree(es.staged[i]);
        es.staged[i] = nullptr;
      }
      return e;
    }
    es.cap_leaves = lv;
  }
  return hipSuccess;
}

int estep_mfma(pcsf_ctx* c, int n, const std::vector<int32_t>& lc,
               const std::vector<int32_t>& rc, const double* P, const double* prior,
               const int32_t* leaves, int64_t ncols, const double* weights, double* z,
               double* ecounts) {
  const int nl = (n + 1) / 2, nbr = n - 1;
  const int64_t chunk = std::min<int64_t>(ncols, 32768);
  const int64_t ldu = (chunk + 63) / 64 * 64;
  const int grid = estep_grid(c);
  std::vector<double> rowsum((size_t)nbr * 64);
  for (int br = 0; br < nbr; ++br)
    for (int a = 0; a < 64; ++a) {
      double t = 0.0;
      for (int b = 0; b < 64; ++b) t += P[((size_t)br * 64 + a) * 64 + b];
      rowsum[(size_t)br * 64 + a] = t;
    }
  const std::vector<int32_t> seq = estep_sequence(n, lc, rc);
  int32_t *d_lc = nullptr, *d_rc = nullptr, *d_seq = nullptr;
  double *d_P = nullptr, *d_rs = nullptr, *d_prior = nullptr, *d_w = nullptr, *d_z = nullptr,
         *d_ec = nullptr, *d_Pf = nullptr, *d_PTf = nullptr;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) {
    if (e == hipSuccess) e = x;
  };
  chk(estep_cache(c, n, ldu, chunk, grid));
  EstepCache& es = c->es;
  chk(dalloc(&d_lc, (size_t)n));
  chk(dalloc(&d_rc, (size_t)n));
  chk(dalloc(&d_P, (size_t)nbr * 4096));
  chk(dalloc(&d_Pf, (size_t)nbr * 4096));
  chk(dalloc(&d_PTf, (size_t)nbr * 4096));
  chk(dalloc(&d_seq, std::max<size_t>(1, seq.size())));
  chk(dalloc(&d_rs, (size_t)nbr * 64));
  chk(dalloc(&d_prior, 64));
  if (weights) chk(dalloc(&d_w, (size_t)ncols));
  chk(dalloc(&d_z, (size_t)ncols));
  chk(dalloc(&d_ec, (size_t)nbr * 4096));
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  auto down = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  };
  up(d_lc, lc.data(), (size_t)n * 4);
  up(d_rc, rc.data(), (size_t)n * 4);
  up(d_P, P, (size_t)nbr * 4096 * 8);
  up(d_seq, seq.data(), seq.size() * 4);
  if (e == hipSuccess) e = launch_estep_frags(d_P, nbr, d_Pf, d_PTf, c->stream);
  up(d_rs, rowsum.data(), rowsum.size() * 8);
  up(d_prior, prior, 64 * 8);
  up(d_ec, ecounts, (size_t)nbr * 4096 * 8);  // add_branch_posteriors adds
  if (weights) up(d_w, weights, (size_t)ncols * 8);
  // Chunk k's leaf states go host -> pinned slot k % 2 -> device slot k % 2 on
  // the copy stream while chunk k - 1 computes; a slot is refilled once the
  // chunk two back has been copied out of it (host) and computed (device).
  const int64_t nchunks = (ncols + chunk - 1) / chunk;
  auto stage = [&](int64_t k) {
    const int sl = (int)(k & 1);
    const int64_t c0 = k * chunk, nc = std::min(chunk, ncols - c0);
    if (k >= 2) {
      chk(hipEventSynchronize(es.uploaded[sl]));
      chk(hipStreamWaitEvent(es.copy, es.consumed[sl], 0));
    }
    if (e != hipSuccess) return;
    std::memcpy(es.staged[sl], leaves + c0 * nl, (size_t)nc * nl * 4);
    chk(hipMemcpyAsync(es.leaves[sl], es.staged[sl], (size_t)nc * nl * 4,
                       hipMemcpyHostToDevice, es.copy));
    chk(hipEventRecord(es.uploaded[sl], es.copy));
  };
  if (nchunks > 0) stage(0);
  for (int64_t k = 0; k < nchunks && e == hipSuccess; ++k) {
    if (k + 1 < nchunks) stage(k + 1);
    const int sl = (int)(k & 1);
    const int64_t c0 = k * chunk, nc = std::min(chunk, ncols - c0);
    chk(hipStreamWaitEvent(c->stream, es.uploaded[sl], 0));
    EstepArgs a;
    a.n = n;
    a.nl = nl;
    a.lc = d_lc;
    a.rc = d_rc;
    a.P = d_P;
    a.Pf = d_Pf;
    a.PTf = d_PTf;
    a.seq = d_seq;
    a.nseq = (int)seq.size();
    a.rowsum = d_rs;
    a.prior = d_prior;
    a.leaves = es.leaves[sl];
    a.weights = d_w ? d_w + c0 : nullptr;
    a.ncols = nc;
    a.z = d_z + c0;
    a.U = es.U;
    a.V = es.V;
    a.ldu = ldu;
    a.scratch = es.scr;
    if (e == hipSuccess) e = launch_estep(a, d_P, d_ec, grid, c->stream);
    chk(hipEventRecord(es.consumed[sl], c->stream));
  }
  // both streams drain before any buffer of this call is released
  const hipError_t s1 = hipStreamSynchronize(c->stream);
  const hipError_t s2 = es.copy ? hipStreamSynchronize(es.copy) : hipSuccess;
  chk(s1);
  chk(s2);
  down(z, d_z, (size_t)ncols * 8);
  down(ecounts, d_ec, (size_t)nbr * 4096 * 8);
  for (auto* q : {d_lc, d_rc, d_seq}) {
    int32_t* x = q;
    dfree(x);
  }
  for (auto* q : {d_P, d_rs, d_prior, d_w, d_z, d_ec, d_Pf, d_PTf}) {
    double* x = q;
    dfree(x);
  }
  HIP_TRY(c, e);
  return PCSF_OK;
}

int lik_impl(pcsf_ctx* c, int k, int n, const int32_t* children, const double* P,
             const double* prior, const int32_t* leaves, const double* leaf_dist, int64_t ncols,
             const double* weights, double* z, int post_mode, double* post, double* ecounts) {
  if (!c) return PCSF_EINVAL;
  if (k < 1 || k > 64 || n < 3 || n % 2 == 0 || !children || !P || !prior || !z ||
      (ncols > 0 && !leaves) || ncols < 0)
    return set_err(c, PCSF_EINVAL, "CamlPaml.Infer.prepare: bad arguments");
  const int nl = (n + 1) / 2;
  std::vector<int32_t> lc(n, -1), rc(n, -1), par(n, -1), sib(n, -1);
  for (int i = nl; i < n; ++i) {
    const int a = children[(i - nl) * 2], b = children[(i - nl) * 2 + 1];
    if (a < 0 || b < 0 || a >= i || b >= i || par[a] >= 0 || par[b] >= 0 || a == b)
      return set_err(c, PCSF_EINVAL, "CamlPaml.T: children must precede parents (T order)");
    lc[i] = a;
    rc[i] = b;
    par[a] = par[b] = i;
    sib[a] = b;
    sib[b] = a;
  }
  // a state >= k is a leaf vector of zeros, as raw_leaf makes it
  // (PhyloLik.ml:11); below -2 is no leaf kind.  One branch-free pass (it
  // vectorises) over what can be a few hundred MB of states.
  int32_t vmin = 0, ndist = 0;
  for (int64_t e = 0; e < ncols * nl; ++e) {
    const int32_t v = leaves[e];
    vmin = std::min(vmin, v);
    ndist |= (int32_t)(v == -2);
  }
  if (vmin < -2) return set_err(c, PCSF_EINVAL, "CamlPaml.Infer.prepare: leaf state out of range");
  const bool dist = ndist != 0;
  if (dist && !leaf_dist)
    return set_err(c, PCSF_EINVAL, "CamlPaml.Infer.prepare: `Distribution leaves need leaf_dist");
  if (ncols == 0) return PCSF_OK;
  HIP_TRY(c, hipSetDevice(c->device));
  const bool estep = ecounts != nullptr;
  if (estep && k == 64 && !dist && post_mode == 0)
    return estep_mfma(c, n, lc, rc, P, prior, leaves, ncols, weights, z, ecounts);
  const size_t per_col = (size_t)nl * 4 + (dist ? (size_t)nl * k * 8 : 0) + 16 +
                         (post_mode == 2 ? (size_t)n * k * 8 : post_mode == 1 ? (size_t)k * 8 : 0);
  const int64_t chunk =
      std::max<int64_t>(1, std::min<int64_t>(estep ? 512 : 65536, ((size_t)512 << 20) / per_col));
  const int grid_max = (int)std::min<int64_t>(estep ? chunk : 2048, std::min(ncols, chunk));
  int32_t *d_lc = nullptr, *d_rc = nullptr, *d_par = nullptr, *d_sib = nullptr,
          *d_leaves = nullptr;
  double *d_P = nullptr, *d_prior = nullptr, *d_dist = nullptr, *d_w = nullptr, *d_z = nullptr,
         *d_post = nullptr, *d_ws = nullptr, *d_U = nullptr, *d_V = nullptr, *d_ec = nullptr;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) {
    if (e == hipSuccess) e = x;
  };
  const int64_t cmax = std::min(ncols, chunk);
  chk(dalloc(&d_lc, (size_t)n));
  chk(dalloc(&d_rc, (size_t)n));
  chk(dalloc(&d_par, (size_t)n));
  chk(dalloc(&d_sib, (size_t)n));
  chk(dalloc(&d_P, (size_t)(n - 1) * k * k));
  chk(dalloc(&d_prior, (size_t)k));
  chk(dalloc(&d_leaves, (size_t)cmax * nl));
  if (dist) chk(dalloc(&d_dist, (size_t)cmax * nl * k));
  if (weights) chk(dalloc(&d_w, (size_t)cmax));
  chk(dalloc(&d_z, (size_t)cmax));
  if (post_mode) chk(dalloc(&d_post, (size_t)cmax * (post_mode == 2 ? n : 1) * k));
  chk(dalloc(&d_ws, (size_t)grid_max * 2 * n * k));
  if (estep) {
    chk(dalloc(&d_U, (size_t)(n - 1) * grid_max * k));
    chk(dalloc(&d_V, (size_t)(n - 1) * grid_max * k));
    chk(dalloc(&d_ec, (size_t)(n - 1) * k * k));
  }
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  auto down = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  };
  up(d_lc, lc.data(), n * 4);
  up(d_rc, rc.data(), n * 4);
  up(d_par, par.data(), n * 4);
  up(d_sib, sib.data(), n * 4);
  up(d_P, P, (size_t)(n - 1) * k * k * 8);
  up(d_prior, prior, (size_t)k * 8);
  if (estep) up(d_ec, ecounts, (size_t)(n - 1) * k * k * 8);  // add_branch_posteriors adds
  for (int64_t c0 = 0; c0 < ncols && e == hipSuccess; c0 += chunk) {
    const int64_t nc = std::min(chunk, ncols - c0);
    up(d_leaves, leaves + c0 * nl, (size_t)nc * nl * 4);
    if (dist) up(d_dist, leaf_dist + (size_t)c0 * nl * k, (size_t)nc * nl * k * 8);
    if (weights) up(d_w, weights + c0, (size_t)nc * 8);
    LikArgs a;
    a.k = k;
    a.n = n;
    a.lc = d_lc;
    a.rc = d_rc;
    a.parent = d_par;
    a.sib = d_sib;
    a.P = d_P;
    a.prior = d_prior;
    a.leaves = d_leaves;
    a.leaf_dist = d_dist;
    a.ncols = nc;
    a.weights = d_w;
    a.ws = d_ws;
    a.z = d_z;
    a.post_mode = post_mode;
    a.post = d_post;
    a.U = d_U;
    a.V = d_V;
    a.ecounts = d_ec;
    // one column per workgroup when the E-step factors are kept per slot
    const int grid = estep ? (int)nc : (int)std::min<int64_t>(nc, grid_max);
    if (e == hipSuccess) e = launch_lik_posteriors(a, grid, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    down(z + c0, d_z, (size_t)nc * 8);
    if (post_mode)
      down(post + (size_t)c0 * (post_mode == 2 ? n : 1) * k, d_post,
           (size_t)nc * (post_mode == 2 ? n : 1) * k * 8);
  }
  if (estep) down(ecounts, d_ec, (size_t)(n - 1) * k * k * 8);
  for (auto* p : {d_lc, d_rc, d_par, d_sib, d_leaves}) {
    int32_t* q = p;
    dfree(q);
  }
  for (auto* p : {d_P, d_prior, d_dist, d_w, d_z, d_post, d_ws, d_U, d_V, d_ec}) {
    double* q = p;
    dfree(q);
  }
  HIP_TRY(c, e);
  return PCSF_OK;
}

}  // namespace

extern "C" {

int pcsf_lik_columns(pcsf_ctx* c, int k, int n, const int32_t* children, const double* P,
                     const double* prior, const int32_t* leaves, int64_t ncols, double* z,
                     double* root_post) {
  if (!c) return PCSF_EINVAL;
  if (leaves)  // this entry takes `Certain and `Marginalize leaves only
    for (int64_t e = 0; e < ncols * ((n + 1) / 2); ++e)
      if (leaves[e] < -1)
        return set_err(c, PCSF_EINVAL, "CamlPaml.Infer.prepare: leaf state out of range");
  return lik_impl(c, k, n, children, P, prior, leaves, nullptr, ncols, nullptr, z,
                  root_post ? 1 : 0, root_post, nullptr);
}

int pcsf_lik_posteriors(pcsf_ctx* c, int k, int n, const int32_t* children, const double* P,
                        const double* prior, const int32_t* leaves, const double* leaf_dist,
                        int64_t ncols, const double* weights, double* z, double* node_post,
                        double* ecounts) {
  return lik_impl(c, k, n, children, P, prior, leaves, leaf_dist, ncols, weights, z,
                  node_post ? 2 : 0, node_post, ecounts);
}

}  // extern "C"
