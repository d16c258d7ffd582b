// Plain FP32 contraction on the matrix cores (v_mfma_f32_16x16x4f32).
//
// Same arithmetic as the reference (kv_impl.py:77-113): the centred value is
// formed in FP32 as fl(code - mean) (missing -> 0) and multiplied in FP32;
// the matrix cores accumulate in FP32 within each 64-element block, and the
// block sums are flushed into FP64 accumulators (a plain FP32 running sum
// over ~1e5 samples loses ~1e-2 on sums of centred codes whose exact value is
// ~1e-3, which breaks the operator's linearity on the intercept column).
// Split partials are FP64 and reduced in a fixed order by the reduction
// kernels of kv.hip.
//
// Operand mapping: the contraction axis is walked in 64-element blocks; one
// dword of packed codes per lane covers elements 16g .. 16g+15 of a block
// (g = lane / 16), and MFMA K-step t (0..15) takes element 16g + t from lane
// group g.  The right-hand block is staged in LDS as [t][g][col16][nb] so a
// lane fetches its NB column blocks with one ds_read.  C layout: lane holds
// column lane % 16 of rows 4 * (lane / 16) + v.
#include "kv_plan.hpp"
#include "timing.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>

namespace coh {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kXcdsF = 8;

struct WgMapF {
  int64_t bx, by;
  int bz;
  bool valid;
};

// Column groups of one (tile, split) unit on the same XCD, adjacent in its
// queue (ids 8 apart), so they share the genotype rows through its L2.
__device__ inline WgMapF xcd_map_f(int64_t ux, int64_t uy, int groups) {
  const int64_t w = blockIdx.x;
  const int64_t q = w / kXcdsF;
  const int64_t u = (q / groups) * kXcdsF + w % kXcdsF;
  WgMapF m;
  m.bz = static_cast<int>(q % groups);
  m.valid = u < ux * uy;
  m.bx = u % ux;
  m.by = u / ux;
  return m;
}

static unsigned xcd_grid_f(int64_t ux, int64_t uy, int64_t groups) {
  return (unsigned)(ceil_div(ux * uy, kXcdsF) * kXcdsF * groups);
}

template <int NB>
__device__ inline void load_cols(const float* __restrict__ s, float b[NB]);

template <>
__device__ inline void load_cols<1>(const float* __restrict__ s, float b[1]) {
  b[0] = s[0];
}
template <>
__device__ inline void load_cols<2>(const float* __restrict__ s, float b[2]) {
  const float2 v = *reinterpret_cast<const float2*>(s);
  b[0] = v.x;
  b[1] = v.y;
}
template <>
__device__ inline void load_cols<4>(const float* __restrict__ s, float b[4]) {
  const float4 v = *reinterpret_cast<const float4*>(s);
  b[0] = v.x;
  b[1] = v.y;
  b[2] = v.z;
  b[3] = v.w;
}

// Stage a 64 x (NB*16) block of a row-major FP32 matrix X (rows r0..r0+63,
// columns c0..c0+NB*16) into sB[t][g][j][nb] for element 16g + t, column
// nb*16 + j; rows >= rows_end and columns >= ncols are zero.  Optionally the
// 64 per-row means go to sM[t * 4 + g].
template <int NB>
__device__ inline void stage_block(const float* __restrict__ X, int64_t ldx, int64_t r0,
                                   int64_t rows_end, int64_t c0, int64_t ncols,
                                   float* __restrict__ sB, const float* __restrict__ mean,
                                   int64_t mean_off, float* __restrict__ sM) {
  constexpr int COLS = NB * 16;
  for (int e = threadIdx.x; e < 64 * COLS; e += 256) {
    const int rl = e / COLS, cl = e % COLS;
    const int64_t r = r0 + rl, c = c0 + cl;
    const float x = (r < rows_end && c < ncols) ? X[r * ldx + c] : 0.0f;
    const int g = rl / 16, t = rl % 16, nb = cl / 16, j = cl % 16;
    sB[((t * 4 + g) * 16 + j) * NB + nb] = x;
  }
  if (sM != nullptr && threadIdx.x < 64) {
    const int rl = threadIdx.x;
    const int64_t r = r0 + rl;
    sM[(rl % 16) * 4 + rl / 16] = (r < rows_end) ? mean[mean_off + r] : 0.0f;
  }
}

struct F1Args {
  const uint8_t* vmaj;
  int64_t vbytes;
  int64_t slot0, nslots;
  int64_t n;
  int64_t nblk;           // 64-sample blocks
  int64_t blk_per_split;
  const float* V;
  int64_t ldv, k;
  const float* mean;
  double* part;           // [split][nslots][k]
  int64_t ux, uy;
  int groups;
};

// part[split][slot][col] = sum over the split's samples of fl(code - mean) * V
template <int NB, int MB>
__global__ __launch_bounds__(256) void k_p1_f32(F1Args a) {
  __shared__ float sB[2][64 * NB * 16];
  const WgMapF wg = xcd_map_f(a.ux, a.uy, a.groups);
  if (!wg.valid) return;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int r = lane % 16, g = lane / 16;
  const int64_t tile0 = wg.bx * (64 * MB) + wave * (16 * MB);
  const int64_t col0 = (int64_t)wg.bz * NB * 16;
  const int64_t b0 = wg.by * a.blk_per_split;
  const int64_t b1 = min(a.nblk, b0 + a.blk_per_split);
  const uint8_t* rowp[MB];
  float mu[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) {
    const int64_t slot = a.slot0 + tile0 + mb * 16 + r;
    rowp[mb] = a.vmaj + slot * a.vbytes + 4 * g;
    mu[mb] = a.mean[slot];
  }
  v4f acc[MB][NB];
  double tot[MB][NB][4];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      acc[mb][nb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int v = 0; v < 4; ++v) tot[mb][nb][v] = 0.0;
    }
  uint32_t cur[MB], nxt[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) cur[mb] = nxt[mb] = 0u;
  if (b0 < b1) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) cur[mb] = *reinterpret_cast<const uint32_t*>(rowp[mb] + b0 * 16);
    stage_block<NB>(a.V, a.ldv, b0 * 64, a.n, col0, a.k, sB[0], nullptr, 0, nullptr);
  }
  __syncthreads();
  int buf = 0;
  for (int64_t b = b0; b < b1; ++b) {
    const int64_t bn = min(b + 1, b1 - 1);
    stage_block<NB>(a.V, a.ldv, bn * 64, a.n, col0, a.k, sB[buf ^ 1], nullptr, 0, nullptr);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) nxt[mb] = *reinterpret_cast<const uint32_t*>(rowp[mb] + bn * 16);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      float bv[NB];
      load_cols<NB>(&sB[buf][((t * 4 + g) * 16 + r) * NB], bv);
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const uint32_t code = (cur[mb] >> (2 * t)) & 3u;
        const float d = (code == 3u) ? 0.0f : (static_cast<float>(code) - mu[mb]);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(d, bv[nb], acc[mb][nb], 0, 0, 0);
      }
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
        for (int v = 0; v < 4; ++v) tot[mb][nb][v] += static_cast<double>(acc[mb][nb][v]);
        acc[mb][nb] = v4f{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) cur[mb] = nxt[mb];
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int64_t col = col0 + nb * 16 + r;
      if (col >= a.k) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t jl = tile0 + mb * 16 + 4 * g + v;
        a.part[(wg.by * a.nslots + jl) * a.k + col] = tot[mb][nb][v];
      }
    }
}

struct F2Args {
  const uint8_t* smaj;
  int64_t sbytes;
  int64_t n;
  int64_t slot0;
  const float* U;        // (nslots, ldu) rows of slot - slot0
  int64_t ldu, k;
  const float* mean;
  ChunkTable tab;
  double* part;          // [chunk][round_up(n, 4)][k]
  int64_t ux, uy;
  int groups;
};

// part[chunk][i][col] = sum over the chunk's slots of fl(code - mean) * U
template <int NB, int MB>
__global__ __launch_bounds__(256) void k_p2_f32(F2Args a) {
  __shared__ float sB[2][64 * NB * 16];
  __shared__ float sM[2][64];
  const WgMapF wg = xcd_map_f(a.ux, a.uy, a.groups);
  if (!wg.valid) return;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int r = lane % 16, g = lane / 16;
  const int64_t i0 = wg.bx * (64 * MB) + wave * (16 * MB);
  const int64_t col0 = (int64_t)wg.bz * NB * 16;
  int64_t j0, j1;
  a.tab.chunk_range(static_cast<int>(wg.by), &j0, &j1);
  const int64_t b0 = (j0 - a.slot0) / 64, b1 = (j1 - a.slot0 + 63) / 64;
  const int64_t rows_end = j1 - a.slot0;  // U rows beyond the chunk are not read
  const uint8_t* rowp[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) {
    const int64_t i = i0 + mb * 16 + r;
    rowp[mb] = a.smaj + (i < a.n ? i : 0) * a.sbytes + a.slot0 / 4 + 4 * g;
  }
  v4f acc[MB][NB];
  double tot[MB][NB][4];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      acc[mb][nb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int v = 0; v < 4; ++v) tot[mb][nb][v] = 0.0;
    }
  uint32_t cur[MB], nxt[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) cur[mb] = nxt[mb] = 0u;
  if (b0 < b1) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) cur[mb] = *reinterpret_cast<const uint32_t*>(rowp[mb] + b0 * 16);
    stage_block<NB>(a.U, a.ldu, b0 * 64, rows_end, col0, a.k, sB[0], a.mean, a.slot0, sM[0]);
  }
  __syncthreads();
  int buf = 0;
  for (int64_t b = b0; b < b1; ++b) {
    const int64_t bn = min(b + 1, b1 - 1);
    stage_block<NB>(a.U, a.ldu, bn * 64, rows_end, col0, a.k, sB[buf ^ 1], a.mean, a.slot0,
                    sM[buf ^ 1]);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) nxt[mb] = *reinterpret_cast<const uint32_t*>(rowp[mb] + bn * 16);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      float bv[NB];
      load_cols<NB>(&sB[buf][((t * 4 + g) * 16 + r) * NB], bv);
      const float muj = sM[buf][t * 4 + g];
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const uint32_t code = (cur[mb] >> (2 * t)) & 3u;
        const float d = (code == 3u) ? 0.0f : (static_cast<float>(code) - muj);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(d, bv[nb], acc[mb][nb], 0, 0, 0);
      }
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
        for (int v = 0; v < 4; ++v) tot[mb][nb][v] += static_cast<double>(acc[mb][nb][v]);
        acc[mb][nb] = v4f{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) cur[mb] = nxt[mb];
    __syncthreads();
    buf ^= 1;
  }
  const int64_t nb4 = ((a.n + 3) / 4) * 4;
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int64_t col = col0 + nb * 16 + r;
      if (col >= a.k) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = i0 + mb * 16 + 4 * g + v;
        if (i >= a.n) continue;
        a.part[(wg.by * nb4 + i) * a.k + col] = tot[mb][nb][v];
      }
    }
}

// ------------------------------------------------------------- planning

static int f32_nb(int64_t k) { return k <= 16 ? 1 : (k <= 32 ? 2 : 4); }

// Row tiles per wave of phase 1 / phase 2: defaults from a sweep at n = 1e5
// (4 for one column group, where the B operand is small and the genotype
// stream dominates; 2 above, where 4 spills), or COH_F32_MB="mb1,mb2" (1, 2
// or 4) for tuning sweeps.
static void f32_mb(int64_t k, int* mb1, int* mb2) {
  *mb1 = k <= 16 ? 4 : 2;
  *mb2 = k <= 16 ? 4 : 2;
  if (const char* e = std::getenv("COH_F32_MB")) {
    int a = 0, b = 0;
    if (std::sscanf(e, "%d,%d", &a, &b) == 2 && (a == 1 || a == 2 || a == 4) &&
        (b == 1 || b == 2 || b == 4)) {
      *mb1 = a;
      *mb2 = b;
    }
  }
}

int plan_f32(const coh_panel* p, int64_t k, int32_t cb, int32_t ce, KvPlan* pl) {
  pl->slot0 = p->comp_slot_begin[cb];
  pl->slot1 = p->comp_slot_end[ce - 1];
  const int64_t nslots = pl->slot1 - pl->slot0;
  const int NB = f32_nb(k);
  const int64_t groups = ceil_div(ceil_div(k, 16), NB);
  pl->KC = NB * 16;
  int mb1, mb2;
  f32_mb(k, &mb1, &mb2);
  pl->shape[0] = mb1;
  pl->shape[4] = mb2;
  const int64_t nblk = p->vbytes / 16;  // 64-sample blocks of the padded rows
  const int64_t tiles1 = std::max<int64_t>(1, nslots / (64 * mb1));
  int64_t nsplit = std::max<int64_t>(1, ceil_div(kTargetBlocks, tiles1 * groups));
  nsplit = std::min<int64_t>(nsplit, std::max<int64_t>(1, nblk / 4));
  nsplit = std::min<int64_t>(nsplit, 64);
  pl->rows_per_split = ceil_div(nblk, nsplit);  // in 64-sample blocks
  pl->nsplit1 = static_cast<int>(ceil_div(nblk, pl->rows_per_split));
  const int64_t xb2 = std::max<int64_t>(1, ceil_div(p->n, 64 * mb2)) * groups;
  const int64_t target_chunks = std::max<int64_t>(1, ceil_div(kTargetBlocks, xb2));
  pl->chunk_len = std::max<int64_t>(256, round_up(ceil_div(nslots, target_chunks), 256));
  build_chunks(p, cb, ce, pl->chunk_len, &pl->tab);
  pl->ldu = k;
  const int64_t nb4 = round_up(p->n, 4);
  int64_t off = 0;
  pl->off_U = off;
  off += round_up(nslots * k * 4, 256);
  pl->off_p1 = off;  // phase-1 and phase-2 partials share one buffer
  pl->off_p2 = off;
  const int64_t p1 = (int64_t)pl->nsplit1 * nslots * k * 8;
  const int64_t p2 = (int64_t)pl->tab.nchunks * nb4 * k * 8;
  off += round_up(std::max(p1, p2), 256);
  pl->total = off;
  return COH_OK;
}

int run_phase1_f32(const coh_panel* p, const KvPlan& pl, const float* V, int64_t k, int64_t ldv,
                   const CompScale& cs, int inv_pow, uint8_t* ws, hipStream_t st) {
  const int64_t nslots = pl.slot1 - pl.slot0;
  const int NB = f32_nb(k);
  const int64_t groups = ceil_div(ceil_div(k, 16), NB);
  F1Args a;
  a.vmaj = p->vmaj;
  a.vbytes = p->vbytes;
  a.slot0 = pl.slot0;
  a.nslots = nslots;
  a.n = p->n;
  a.nblk = p->vbytes / 16;
  a.blk_per_split = pl.rows_per_split;
  a.V = V;
  a.ldv = ldv;
  a.k = k;
  a.mean = p->mean;
  a.part = reinterpret_cast<double*>(ws + pl.off_p1);
  const int MB = pl.shape[0];
  a.ux = nslots / (64 * MB);
  a.uy = pl.nsplit1;
  a.groups = (int)groups;
  char name[64];
  snprintf(name, sizeof(name), "k_p1_f32_NB%d_MB%d", NB, MB);
  const double fl = 2.0 * (double)p->n * (double)pl.m_sel * (double)k;
  KernelTimer tm(st, name, (double)p->n * (double)ceil_div(pl.m_sel, 4), fl,
                 2.0 * (double)nslots * (double)(a.nblk * 64) * (double)(groups * NB * 16));
  dim3 grid(xcd_grid_f(a.ux, a.uy, groups));
#define COH_F1(NB_, MB_) \
  if (NB == NB_ && MB == MB_) hipLaunchKernelGGL((k_p1_f32<NB_, MB_>), grid, dim3(256), 0, st, a);
  COH_F1(1, 1) COH_F1(1, 2) COH_F1(1, 4) COH_F1(2, 1) COH_F1(2, 2) COH_F1(2, 4)
  COH_F1(4, 1) COH_F1(4, 2) COH_F1(4, 4)
#undef COH_F1
  tm.stop();
  COH_CHECK_LAUNCH();
  return reduce_p1_partials(pl, a.part, pl.nsplit1, nslots, k, p->inv, cs, inv_pow,
                            reinterpret_cast<float*>(ws + pl.off_U), st);
}

int run_phase2_f32(const coh_panel* p, const KvPlan& pl, int64_t k, int stacked, float* out,
                   int64_t ldo, int64_t out_stride, float beta, uint8_t* ws, hipStream_t st) {
  const int NB = f32_nb(k);
  const int64_t groups = ceil_div(ceil_div(k, 16), NB);
  F2Args a;
  a.smaj = p->smaj;
  a.sbytes = p->sbytes;
  a.n = p->n;
  a.slot0 = pl.slot0;
  a.U = reinterpret_cast<const float*>(ws + pl.off_U);
  a.ldu = pl.ldu;
  a.k = k;
  a.mean = p->mean;
  a.tab = pl.tab;
  a.part = reinterpret_cast<double*>(ws + pl.off_p2);
  const int MB = pl.shape[4];
  a.ux = ceil_div(p->n, 64 * MB);
  a.uy = pl.tab.nchunks;
  a.groups = (int)groups;
  char name[64];
  snprintf(name, sizeof(name), "k_p2_f32_NB%d_MB%d", NB, MB);
  const double fl = 2.0 * (double)p->n * (double)pl.m_sel * (double)k;
  KernelTimer tm(st, name, (double)p->n * (double)ceil_div(pl.m_sel, 4), fl,
                 2.0 * (double)(pl.slot1 - pl.slot0) * (double)round_up(p->n, 64 * MB) *
                     (double)(groups * NB * 16));
  dim3 grid(xcd_grid_f(a.ux, a.uy, groups));
#define COH_F2(NB_, MB_) \
  if (NB == NB_ && MB == MB_) hipLaunchKernelGGL((k_p2_f32<NB_, MB_>), grid, dim3(256), 0, st, a);
  COH_F2(1, 1) COH_F2(1, 2) COH_F2(1, 4) COH_F2(2, 1) COH_F2(2, 2) COH_F2(2, 4)
  COH_F2(4, 1) COH_F2(4, 2) COH_F2(4, 4)
#undef COH_F2
  tm.stop();
  COH_CHECK_LAUNCH();
  return reduce_p2_partials(pl, a.part, p->n, k, stacked, out, ldo, out_stride, beta, st);
}

}  // namespace coh
