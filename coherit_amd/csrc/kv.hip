// Matrix-free operator family K_g V = Z_g (Z_g^T V) / eff_g on the resident
// panel (replaces kv_impl.py:58-218 per-call kernels and the streaming
// drivers kv_impl.py:268-439, 911-1339).
//
// Phase 1 ("xtv"):  T[slot, :] = factor(slot) * sum_i diff[i, slot] V[i, :]
// Phase 2 ("zu"):   out_g[i, :] (+)= sum_{slot in g} diff[i, slot] T[slot, :]
// with diff = (code == 3 ? 0 : code - mean) (missing imputed to the mean and
// centred, kv_impl.py:77-91) and factor = inv^2 * comp_scale[g] for K.V,
// inv * scale for X^T V.  Both phases split their contraction axis over
// blocks and reduce the partials in a fixed order (deterministic results).
//
// Precision COH_PREC_FP32 runs on the VALU: the centred value is formed in
// FP32 exactly as the reference does and multiplied in FP32; FMAs accumulate
// in FP32 over runs of kRun terms that are flushed into FP64 accumulators, and
// split partials are reduced in FP64.  A plain FP32 running sum over ~1e5
// samples loses ~1e-2 absolute on sums of centred codes whose exact value is
// ~1e-3 (the intercept column of X^T V), which puts a ~1e-4 floor under PCG
// residuals at n = 1e5; the blocked sum removes it at ~3% extra VALU work.
// The int8-limb MFMA precisions live in kv_mfma.hip.
#include "common.hpp"
#include "kv_plan.hpp"
#include "timing.hpp"

#include <algorithm>
#include <cstdio>

namespace coh {

// ------------------------------------------------------------ VALU kernels

constexpr int kRun = 32;  // FP32 accumulation run before an FP64 flush

template <int KC>
__device__ inline void flush_run(double tot[4][KC], float acc[4][KC]) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      tot[s][c] += static_cast<double>(acc[s][c]);
      acc[s][c] = 0.0f;
    }
}

template <int KC>
__device__ inline void p1_accum(uint32_t b, const float* __restrict__ vrow, int kvalid,
                                const float mu[4], float acc[4][KC]) {
  float v[KC];
#pragma unroll
  for (int c = 0; c < KC; ++c) v[c] = (c < kvalid) ? vrow[c] : 0.0f;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const uint32_t code = (b >> (2 * s)) & 3u;
    const float d = (code == 3u) ? 0.0f : (static_cast<float>(code) - mu[s]);
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[s][c] = fmaf(d, v[c], acc[s][c]);
  }
}

// One thread per panel byte column (4 slots); loops over a sample range.
template <int KC>
__global__ __launch_bounds__(256) void k_p1_valu(const uint8_t* __restrict__ smaj, int64_t sbytes,
                                                 int64_t n, int64_t byte0, int64_t nbytes,
                                                 const float* __restrict__ mean,
                                                 const float* __restrict__ V, int64_t ldv,
                                                 int kvalid, int64_t rows_per_split,
                                                 double* __restrict__ part) {
  const int64_t qb = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool act = qb < nbytes;
  const int64_t q = byte0 + (act ? qb : 0);
  const int64_t i0 = (int64_t)blockIdx.y * rows_per_split;
  const int64_t i1 = min(n, i0 + rows_per_split);
  float mu[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) mu[s] = mean[4 * q + s];
  float acc[4][KC];
  double tot[4][KC];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      acc[s][c] = 0.0f;
      tot[s][c] = 0.0;
    }
  const uint8_t* col = smaj + q;
  int64_t i = i0;
  while (i < i1) {
    const int64_t iend = min(i1, i + kRun);
    for (; i + 4 <= iend; i += 4) {
      const uint32_t b0 = col[(i + 0) * sbytes];
      const uint32_t b1 = col[(i + 1) * sbytes];
      const uint32_t b2 = col[(i + 2) * sbytes];
      const uint32_t b3 = col[(i + 3) * sbytes];
      p1_accum<KC>(b0, V + (i + 0) * ldv, kvalid, mu, acc);
      p1_accum<KC>(b1, V + (i + 1) * ldv, kvalid, mu, acc);
      p1_accum<KC>(b2, V + (i + 2) * ldv, kvalid, mu, acc);
      p1_accum<KC>(b3, V + (i + 3) * ldv, kvalid, mu, acc);
    }
    for (; i < iend; ++i) p1_accum<KC>(col[i * sbytes], V + i * ldv, kvalid, mu, acc);
    flush_run<KC>(tot, acc);
  }
  if (!act) return;
  double* out = part + ((int64_t)blockIdx.y * nbytes + qb) * 4 * KC;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < KC; ++c) out[s * KC + c] = tot[s][c];
}

template <int KC>
__device__ inline void p2_accum(uint32_t byte, float mu, const float* __restrict__ urow,
                                int kvalid, float acc[4][KC]) {
  float u[KC];
#pragma unroll
  for (int c = 0; c < KC; ++c) u[c] = (c < kvalid) ? urow[c] : 0.0f;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t code = (byte >> (2 * r)) & 3u;
    const float d = (code == 3u) ? 0.0f : (static_cast<float>(code) - mu);
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[r][c] = fmaf(d, u[c], acc[r][c]);
  }
}

// One thread per variant-major byte (4 samples); loops over a slot chunk.
template <int KC>
__global__ __launch_bounds__(256) void k_p2_valu(const uint8_t* __restrict__ vmaj, int64_t vbytes,
                                                 int64_t n, const float* __restrict__ mean,
                                                 const float* __restrict__ U, int64_t ldu,
                                                 int64_t slot_base, int kvalid, ChunkTable tab,
                                                 double* __restrict__ part) {
  const int64_t nb = (n + 3) / 4;
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool act = b < nb;
  int64_t j0, j1;
  tab.chunk_range(blockIdx.y, &j0, &j1);
  float acc[4][KC];
  double tot[4][KC];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      acc[r][c] = 0.0f;
      tot[r][c] = 0.0;
    }
  const uint8_t* col = vmaj + (act ? b : 0);
  int64_t j = j0;
  while (j < j1) {
    const int64_t jend = min(j1, j + kRun);
    for (; j + 4 <= jend; j += 4) {
      const uint32_t y0 = col[(j + 0) * vbytes];
      const uint32_t y1 = col[(j + 1) * vbytes];
      const uint32_t y2 = col[(j + 2) * vbytes];
      const uint32_t y3 = col[(j + 3) * vbytes];
      p2_accum<KC>(y0, mean[j + 0], U + (j + 0 - slot_base) * ldu, kvalid, acc);
      p2_accum<KC>(y1, mean[j + 1], U + (j + 1 - slot_base) * ldu, kvalid, acc);
      p2_accum<KC>(y2, mean[j + 2], U + (j + 2 - slot_base) * ldu, kvalid, acc);
      p2_accum<KC>(y3, mean[j + 3], U + (j + 3 - slot_base) * ldu, kvalid, acc);
    }
    for (; j < jend; ++j)
      p2_accum<KC>(col[j * vbytes], mean[j], U + (j - slot_base) * ldu, kvalid, acc);
    flush_run<KC>(tot, acc);
  }
  if (!act) return;
  double* out = part + (int64_t)blockIdx.y * nb * 4 * KC + b * 4 * KC;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < KC; ++c) out[r * KC + c] = tot[r][c];
}

// U[local, c0+c] = factor(slot) * sum_split part[split][local][c]
__global__ void k_reduce_p1(const double* __restrict__ part, int nsplit, int64_t nslots, int KC,
                            int kvalid, int c0, const float* __restrict__ inv, int64_t slot0,
                            CompScale cs, int inv_pow, float* __restrict__ U, int64_t ldu) {
  const int64_t total = nslots * KC;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t l = idx / KC;
    const int c = static_cast<int>(idx % KC);
    if (c >= kvalid) continue;
    double sum = 0.0;
    for (int s = 0; s < nsplit; ++s) sum += part[((int64_t)s * nslots + l) * KC + c];
    const int64_t slot = slot0 + l;
    const float iv = inv[slot];
    const float f = (inv_pow == 2 ? iv * iv : iv) * cs.scale_of(slot);
    U[l * ldu + c0 + c] = static_cast<float>(sum * static_cast<double>(f));
  }
}

// out_g[i, c0+c] = beta*out_g + sum over g's chunks of part[chunk][i][c]
__global__ void k_reduce_p2(const double* __restrict__ part, int64_t n, int KC, int kvalid, int c0,
                            ChunkTable tab, int nout, int stacked, float* __restrict__ out,
                            int64_t ldo, int64_t out_stride, float beta) {
  const int64_t nb4 = ((n + 3) / 4) * 4;
  const int64_t total = (int64_t)nout * n * KC;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int c = static_cast<int>(idx % KC);
    const int64_t i = (idx / KC) % n;
    const int g = static_cast<int>(idx / (KC * n));
    if (c >= kvalid) continue;
    int cb = 0, ce = tab.nchunks;
    if (stacked) {
      cb = tab.chunk_off[g];
      ce = tab.chunk_off[g + 1];
    }
    double sum = 0.0;
    for (int ch = cb; ch < ce; ++ch) sum += part[((int64_t)ch * nb4 + i) * KC + c];
    float* o = out + (int64_t)g * out_stride + i * ldo + c0 + c;
    const float v = static_cast<float>(sum);
    *o = (beta != 0.0f) ? beta * (*o) + v : v;
  }
}

__global__ void k_gather_rows(const float* __restrict__ U, int64_t ldu, int64_t slot0,
                              const int32_t* __restrict__ cache_to_slot, int64_t s0, int64_t count,
                              int k, float* __restrict__ out, int64_t ldo) {
  const int64_t total = count * k;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx / k;
    const int c = static_cast<int>(idx % k);
    out[r * ldo + c] = U[(cache_to_slot[s0 + r] - slot0) * ldu + c];
  }
}

__global__ void k_copy_rows(const float* __restrict__ U, int64_t ldu, int64_t rows, int k,
                            float* __restrict__ out, int64_t ldo) {
  const int64_t total = rows * k;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx / k;
    const int c = static_cast<int>(idx % k);
    out[r * ldo + c] = U[r * ldu + c];
  }
}

// U[l, c] = inv[slot0+l] * B[cache(slot0+l) - s0, c]; padding slots -> 0.
__global__ void k_scatter_effects(const float* __restrict__ B, int64_t ldb, int64_t s0,
                                  const int32_t* __restrict__ slot_to_cache,
                                  const float* __restrict__ inv, int64_t slot0, int64_t nslots,
                                  int k, float* __restrict__ U, int64_t ldu) {
  const int64_t total = nslots * k;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t l = idx / k;
    const int c = static_cast<int>(idx % k);
    const int64_t s = slot_to_cache[slot0 + l];
    U[l * ldu + c] = (s >= 0) ? inv[slot0 + l] * B[(s - s0) * ldb + c] : 0.0f;
  }
}

static unsigned grid_for(int64_t total) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, 256), 8192));
}

// ------------------------------------------------------------- planning

static int kc_for(int64_t k) {
  if (k <= 1) return 1;
  if (k <= 2) return 2;
  if (k <= 4) return 4;
  if (k <= 8) return 8;
  return 16;
}

int plan_valu(const coh_panel* p, int64_t k, int32_t cb, int32_t ce, KvPlan* pl) {
  pl->slot0 = p->comp_slot_begin[cb];
  pl->slot1 = p->comp_slot_end[ce - 1];
  const int64_t nslots = pl->slot1 - pl->slot0;
  pl->KC = kc_for(k);
  const int64_t nbytes1 = nslots / 4;
  const int64_t xb1 = std::max<int64_t>(1, ceil_div(nbytes1, 256));
  int64_t nsplit = std::max<int64_t>(1, ceil_div(kTargetBlocks, xb1));
  nsplit = std::min<int64_t>(nsplit, std::max<int64_t>(1, p->n / 64));
  nsplit = std::min<int64_t>(nsplit, 256);
  pl->nsplit1 = static_cast<int>(nsplit);
  pl->rows_per_split = ceil_div(p->n, nsplit);
  const int64_t nb = ceil_div(p->n, 4);
  const int64_t xb2 = std::max<int64_t>(1, ceil_div(nb, 256));
  const int64_t target_chunks = std::max<int64_t>(1, ceil_div(kTargetBlocks, xb2));
  pl->chunk_len = std::max<int64_t>(256, round_up(ceil_div(nslots, target_chunks), 256));
  build_chunks(p, cb, ce, pl->chunk_len, &pl->tab);
  pl->off_U = 0;
  pl->ldu = k;
  int64_t off = round_up(nslots * k * 4, 256);
  pl->off_p1 = off;
  off += round_up((int64_t)pl->nsplit1 * nslots * pl->KC * 8, 256);
  pl->off_p2 = off;
  off += round_up((int64_t)pl->tab.nchunks * nb * 4 * pl->KC * 8, 256);
  pl->total = off;
  return COH_OK;
}

int run_phase1_valu(const coh_panel* p, const KvPlan& pl, const float* V, int64_t k, int64_t ldv,
                    const CompScale& cs, int inv_pow, uint8_t* ws, hipStream_t st) {
  const int64_t nslots = pl.slot1 - pl.slot0;
  const int64_t nbytes1 = nslots / 4;
  float* U = reinterpret_cast<float*>(ws + pl.off_U);
  double* part = reinterpret_cast<double*>(ws + pl.off_p1);
  dim3 grid((unsigned)ceil_div(nbytes1, 256), (unsigned)pl.nsplit1);
  for (int64_t c0 = 0; c0 < k; c0 += pl.KC) {
    const int kvalid = static_cast<int>(std::min<int64_t>(pl.KC, k - c0));
    const float* Vc = V + c0;
#define COH_P1(KCV)                                                                          \
  hipLaunchKernelGGL(k_p1_valu<KCV>, grid, dim3(256), 0, st, p->smaj, p->sbytes, p->n,        \
                     pl.slot0 / 4, nbytes1, p->mean, Vc, ldv, kvalid, pl.rows_per_split, part)
    char name1[64];
    snprintf(name1, sizeof(name1), "k_p1_valu_KC%d", pl.KC);
    const double fl1 = 2.0 * (double)p->n * (double)pl.m_sel * (double)kvalid;
    KernelTimer tm1(st, name1, (double)p->n * (double)ceil_div(pl.m_sel, 4), fl1, fl1);
    switch (pl.KC) {
      case 1: COH_P1(1); break;
      case 2: COH_P1(2); break;
      case 4: COH_P1(4); break;
      case 8: COH_P1(8); break;
      default: COH_P1(16); break;
    }
    tm1.stop();
#undef COH_P1
    COH_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reduce_p1, dim3(grid_for(nslots * pl.KC)), dim3(256), 0, st, part,
                       pl.nsplit1, nslots, pl.KC, kvalid, (int)c0, p->inv, pl.slot0, cs, inv_pow,
                       U, pl.ldu);
    COH_CHECK_LAUNCH();
  }
  return COH_OK;
}

int run_phase2_valu(const coh_panel* p, const KvPlan& pl, int64_t k, int stacked, float* out,
                    int64_t ldo, int64_t out_stride, float beta, uint8_t* ws, hipStream_t st) {
  const float* U = reinterpret_cast<const float*>(ws + pl.off_U);
  double* part = reinterpret_cast<double*>(ws + pl.off_p2);
  const int64_t nb = ceil_div(p->n, 4);
  dim3 grid((unsigned)ceil_div(nb, 256), (unsigned)pl.tab.nchunks);
  const int nout = stacked ? pl.tab.ncomp : 1;
  for (int64_t c0 = 0; c0 < k; c0 += pl.KC) {
    const int kvalid = static_cast<int>(std::min<int64_t>(pl.KC, k - c0));
    const float* Uc = U + c0;
#define COH_P2(KCV)                                                                          \
  hipLaunchKernelGGL(k_p2_valu<KCV>, grid, dim3(256), 0, st, p->vmaj, p->vbytes, p->n, p->mean, \
                     Uc, pl.ldu, pl.slot0, kvalid, pl.tab, part)
    char name2[64];
    snprintf(name2, sizeof(name2), "k_p2_valu_KC%d", pl.KC);
    const double fl2 = 2.0 * (double)p->n * (double)pl.m_sel * (double)kvalid;
    KernelTimer tm2(st, name2, (double)p->n * (double)ceil_div(pl.m_sel, 4), fl2, fl2);
    switch (pl.KC) {
      case 1: COH_P2(1); break;
      case 2: COH_P2(2); break;
      case 4: COH_P2(4); break;
      case 8: COH_P2(8); break;
      default: COH_P2(16); break;
    }
    tm2.stop();
#undef COH_P2
    COH_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reduce_p2, dim3(grid_for((int64_t)nout * p->n * pl.KC)), dim3(256), 0, st,
                       part, p->n, pl.KC, kvalid, (int)c0, pl.tab, nout, stacked, out, ldo,
                       out_stride, beta);
    COH_CHECK_LAUNCH();
  }
  return COH_OK;
}

void build_chunks(const coh_panel* p, int32_t cb, int32_t ce, int64_t chunk_len, ChunkTable* t) {
  t->ncomp = ce - cb;
  t->chunk_len = chunk_len;
  int nch = 0;
  for (int g = 0; g < t->ncomp; ++g) {
    t->begin[g] = p->comp_slot_begin[cb + g];
    t->end[g] = p->comp_slot_end[cb + g];
    t->chunk_off[g] = nch;
    const int64_t len = t->end[g] - t->begin[g];
    nch += static_cast<int>(ceil_div(len, chunk_len));
  }
  t->chunk_off[t->ncomp] = nch;
  t->nchunks = nch;
}

int validate_common(const coh_panel* p, const float* V, int64_t k, int64_t ldv, int32_t cb,
                    int32_t ce) {
  COH_REQUIRE(p != nullptr, COH_ERR_VALUE, "NULL panel");
  COH_REQUIRE(p->finalized, COH_ERR_RUNTIME, "panel statistics are not available (not finalized)");
  COH_REQUIRE(V != nullptr, COH_ERR_VALUE, "NULL V");
  COH_REQUIRE(k >= 1, COH_ERR_VALUE, "V must have at least one column (k=%lld)", (long long)k);
  COH_REQUIRE(ldv >= k, COH_ERR_VALUE, "ldv=%lld < k=%lld", (long long)ldv, (long long)k);
  COH_REQUIRE(cb >= 0 && ce <= p->G && cb < ce, COH_ERR_INDEX,
              "component range [%d, %d) out of range for %d components", cb, ce, p->G);
  COH_REQUIRE(ce - cb <= kMaxComp, COH_ERR_VALUE, "at most %d components per call", kMaxComp);
  return COH_OK;
}

}  // namespace coh

using namespace coh;

extern "C" int coh_workspace_bytes(const coh_panel* p, int32_t op, int32_t precision, int64_t k,
                                   int32_t cb, int32_t ce, uint64_t* bytes) {
  COH_REQUIRE(p && bytes, COH_ERR_VALUE, "NULL argument");
  COH_REQUIRE(op == COH_OP_KV || op == COH_OP_XTV, COH_ERR_VALUE, "unknown op %d", op);
  COH_REQUIRE(k >= 1, COH_ERR_VALUE, "k must be positive");
  COH_REQUIRE(cb >= 0 && ce <= p->G && cb < ce, COH_ERR_INDEX, "component range out of range");
  KvPlan pl{};
  int st = make_plan(p, precision, k, cb, ce, &pl);
  if (st != COH_OK) return st;
  *bytes = static_cast<uint64_t>(pl.total);
  return COH_OK;
}

extern "C" int coh_kv(const coh_panel* p, int32_t precision, const float* V, int64_t k,
                      int64_t ldv, int32_t cb, int32_t ce, const float* comp_scale,
                      int32_t stacked, float* out, int64_t ldo, int64_t out_stride, float beta,
                      void* ws, uint64_t ws_bytes, void* stream) {
  int st = validate_common(p, V, k, ldv, cb, ce);
  if (st != COH_OK) return st;
  COH_REQUIRE(comp_scale != nullptr, COH_ERR_VALUE, "NULL comp_scale");
  COH_REQUIRE(out != nullptr, COH_ERR_VALUE, "NULL out");
  COH_REQUIRE(ldo >= k, COH_ERR_VALUE, "ldo < k");
  COH_REQUIRE(beta == 0.0f || beta == 1.0f, COH_ERR_VALUE, "beta must be 0 or 1");
  COH_REQUIRE(!stacked || out_stride >= p->n * ldo || ce - cb == 1, COH_ERR_VALUE,
              "stacked outputs overlap (out_stride too small)");
  KvPlan pl{};
  st = make_plan(p, precision, k, cb, ce, &pl);
  if (st != COH_OK) return st;
  COH_REQUIRE(ws != nullptr && ws_bytes >= (uint64_t)pl.total, COH_ERR_VALUE,
              "workspace too small: %llu < %lld bytes", (unsigned long long)ws_bytes,
              (long long)pl.total);
  CompScale cs{};
  cs.ncomp = ce - cb;
  for (int g = 0; g < cs.ncomp; ++g) {
    cs.end[g] = p->comp_slot_end[cb + g];
    cs.scale[g] = comp_scale[cb + g];
  }
  hipStream_t s = (hipStream_t)stream;
  return run_kv(p, pl, precision, V, k, ldv, cs, stacked, out, ldo, out_stride, beta,
                static_cast<uint8_t*>(ws), s);
}

extern "C" int coh_xtv(const coh_panel* p, int32_t precision, const float* V, int64_t k,
                       int64_t ldv, int32_t cb, int32_t ce, float scale, int32_t slot_space,
                       float* out, int64_t ldo, void* ws, uint64_t ws_bytes, void* stream) {
  int st = validate_common(p, V, k, ldv, cb, ce);
  if (st != COH_OK) return st;
  COH_REQUIRE(out != nullptr, COH_ERR_VALUE, "NULL out");
  COH_REQUIRE(ldo >= k, COH_ERR_VALUE, "ldo < k");
  KvPlan pl{};
  st = make_plan(p, precision, k, cb, ce, &pl);
  if (st != COH_OK) return st;
  COH_REQUIRE(ws != nullptr && ws_bytes >= (uint64_t)pl.total, COH_ERR_VALUE,
              "workspace too small: %llu < %lld bytes", (unsigned long long)ws_bytes,
              (long long)pl.total);
  CompScale cs{};
  cs.ncomp = ce - cb;
  for (int g = 0; g < cs.ncomp; ++g) {
    cs.end[g] = p->comp_slot_end[cb + g];
    cs.scale[g] = scale;
  }
  hipStream_t s = (hipStream_t)stream;
  uint8_t* w = static_cast<uint8_t*>(ws);
  st = run_xtv(p, pl, precision, V, k, ldv, cs, w, s);
  if (st != COH_OK) return st;
  const float* U = reinterpret_cast<const float*>(w + pl.off_U);
  const int64_t nslots = pl.slot1 - pl.slot0;
  if (slot_space) {
    hipLaunchKernelGGL(k_copy_rows, dim3(grid_for(nslots * k)), dim3(256), 0, s, U, pl.ldu, nslots,
                       (int)k, out, ldo);
  } else {
    // cache SNPs of components [cb, ce): from the first call of cb
    int64_t s0 = -1, s1 = 0;
    for (int32_t c = 0; c < p->n_calls; ++c)
      if (p->call_comp[c] >= cb && p->call_comp[c] < ce) {
        if (s0 < 0) s0 = p->call_start[c];
        s1 = p->call_start[c] + p->call_width[c];
      }
    if (s0 < 0) return COH_OK;
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_for((s1 - s0) * k)), dim3(256), 0, s, U, pl.ldu,
                       pl.slot0, p->cache_to_slot, s0, s1 - s0, (int)k, out, ldo);
  }
  COH_CHECK_LAUNCH();
  return COH_OK;
}

extern "C" int coh_zb(const coh_panel* p, int32_t precision, const float* B, int64_t k, int64_t ldb,
                      int32_t cb, int32_t ce, int32_t stacked, float* out, int64_t ldo,
                      int64_t out_stride, float beta, void* ws, uint64_t ws_bytes, void* stream) {
  int st = validate_common(p, B, k, ldb, cb, ce);
  if (st != COH_OK) return st;
  COH_REQUIRE(out != nullptr, COH_ERR_VALUE, "NULL out");
  COH_REQUIRE(ldo >= k, COH_ERR_VALUE, "ldo < k");
  COH_REQUIRE(beta == 0.0f || beta == 1.0f, COH_ERR_VALUE, "beta must be 0 or 1");
  KvPlan pl{};
  st = make_plan(p, precision, k, cb, ce, &pl);
  if (st != COH_OK) return st;
  COH_REQUIRE(ws != nullptr && ws_bytes >= (uint64_t)pl.total, COH_ERR_VALUE,
              "workspace too small: %llu < %lld bytes", (unsigned long long)ws_bytes,
              (long long)pl.total);
  int64_t s0 = -1;
  for (int32_t c = 0; c < p->n_calls; ++c)
    if (p->call_comp[c] >= cb && p->call_comp[c] < ce) {
      s0 = p->call_start[c];
      break;
    }
  hipStream_t s = (hipStream_t)stream;
  uint8_t* w = static_cast<uint8_t*>(ws);
  float* U = reinterpret_cast<float*>(w + pl.off_U);
  const int64_t nslots = pl.slot1 - pl.slot0;
  hipLaunchKernelGGL(k_scatter_effects, dim3(grid_for(nslots * k)), dim3(256), 0, s, B, ldb,
                     s0 < 0 ? 0 : s0, p->slot_to_cache, p->inv, pl.slot0, nslots, (int)k, U,
                     pl.ldu);
  COH_CHECK_LAUNCH();
  if (precision == COH_PREC_FP32)
    return run_phase2_valu(p, pl, k, stacked, out, ldo, out_stride, beta, w, s);
  return run_phase2_mfma(p, pl, k, stacked, out, ldo, out_stride, beta, w, s);
}
