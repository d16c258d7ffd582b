// Standardization statistics of sample subsets of a resident panel.
//
// The reference restricts a streamer to a sample subset with `sample_mask`
// (geno_stream.py:1161-1247): genotype blocks keep the masked columns
// (geno_stream.py:2597-2598) and the per-SNP stats are taken over the kept
// samples only (geno_stream.py:1780-1792).  Here the panel stays resident and
// every subset's stats come from one pass over the variant-major codes:
//   cnt = #kept observed, s1 = sum of kept values, s2 = sum of kept squares
// as popcounts of (code plane & subset mask).  The subset masks are stored in
// the panel's own 2-bit sample packing (sample i at bit 2*(i % 16) of dword
// i / 16), so a code dword and a mask dword line up bit for bit.
//
// Stats are computed from the packed codes (0/1/2, 3 = missing).  That is the
// reference's result whenever the raw genotype values are 0/1/2 or missing,
// which holds for BED and packed input.
#include "common.hpp"
#include "stats.hpp"

#include <algorithm>

namespace coh {

constexpr int kSubSlots = 64;   // slots per block
constexpr int kSubSets = 16;    // subsets per block (4 per wave)
constexpr int kSubDw = 32;      // code dwords per stage

// spread[set][d] bit 2s = keep[set][16d + s] (samples beyond n: 0)
__global__ __launch_bounds__(256) void k_spread_masks(const uint8_t* __restrict__ keep, int64_t n,
                                                      int64_t nsets, int64_t ndw,
                                                      uint32_t* __restrict__ spread) {
  const int64_t total = nsets * ndw;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t set = idx / ndw, d = idx % ndw;
    uint32_t w = 0;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int64_t i = 16 * d + s;
      if (i < n && keep[set * n + i] != 0) w |= 1u << (2 * s);
    }
    spread[idx] = w;
  }
}

// One block: kSubSlots slots x kSubSets subsets.  Thread (wave q, lane r):
// slot slot0 + r, subsets set0 + 4q .. +3.
__global__ __launch_bounds__(256) void k_subset_stats(const uint8_t* __restrict__ vmaj,
                                                      int64_t vbytes, int64_t m_slot,
                                                      const int32_t* __restrict__ slot_to_cache,
                                                      const uint32_t* __restrict__ spread,
                                                      int64_t nsets, float* __restrict__ mean,
                                                      float* __restrict__ inv,
                                                      int64_t* __restrict__ cnt_out,
                                                      int64_t stride) {
  __shared__ uint32_t sR[kSubDw][kSubSlots + 1];
  __shared__ uint32_t sM[kSubSets][kSubDw];
  const int tid = threadIdx.x;
  const int r = tid % 64, q = tid / 64;
  const int64_t slot0 = (int64_t)blockIdx.x * kSubSlots;
  const int64_t set0 = (int64_t)blockIdx.y * kSubSets;
  const int64_t ndw = vbytes / 4;
  uint32_t cn[4] = {0, 0, 0, 0}, c1[4] = {0, 0, 0, 0}, c2[4] = {0, 0, 0, 0};
  for (int64_t d0 = 0; d0 < ndw; d0 += kSubDw) {
    for (int e = tid; e < kSubSlots * kSubDw; e += 256) {
      const int rr = e / kSubDw, c = e % kSubDw;
      const int64_t slot = slot0 + rr, d = d0 + c;
      sR[c][rr] = (slot < m_slot && d < ndw)
                      ? reinterpret_cast<const uint32_t*>(vmaj + slot * vbytes)[d]
                      : 0u;
    }
    for (int e = tid; e < kSubSets * kSubDw; e += 256) {
      const int t = e / kSubDw, c = e % kSubDw;
      const int64_t set = set0 + t, d = d0 + c;
      sM[t][c] = (set < nsets && d < ndw) ? spread[set * ndw + d] : 0u;
    }
    __syncthreads();
#pragma unroll 4
    for (int c = 0; c < kSubDw; ++c) {
      const uint32_t w = sR[c][r];
      const uint32_t lo = w & 0x55555555u, hi = (w >> 1) & 0x55555555u;
      const uint32_t b1 = lo & ~hi, b2 = hi & ~lo;
      const uint32_t ob = ~(lo & hi) & 0x55555555u;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t mk = sM[4 * q + j][c];
        cn[j] += __popc(mk & ob);
        c1[j] += __popc(mk & b1);
        c2[j] += __popc(mk & b2);
      }
    }
    __syncthreads();
  }
  const int64_t slot = slot0 + r;
  if (slot >= m_slot) return;
  const bool real = slot_to_cache[slot] >= 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t set = set0 + 4 * q + j;
    if (set >= nsets) continue;
    const int64_t o = set * stride + slot;
    const int64_t c = real ? (int64_t)cn[j] : 0;
    float mu = 0.0f, iv = 0.0f;
    if (real) f32_stats(c, (int64_t)c1[j] + 2 * (int64_t)c2[j], (int64_t)c1[j] + 4 * (int64_t)c2[j], &mu, &iv);
    mean[o] = mu;
    inv[o] = iv;
    if (cnt_out) cnt_out[o] = c;
  }
}

}  // namespace coh

using namespace coh;

extern "C" int coh_subset_stats_workspace_bytes(const coh_panel* p, int32_t nsets,
                                                uint64_t* bytes) {
  COH_REQUIRE(p && bytes, COH_ERR_VALUE, "NULL argument");
  COH_REQUIRE(nsets >= 1, COH_ERR_VALUE, "nsets must be positive");
  *bytes = (uint64_t)round_up((int64_t)nsets * p->vbytes, 256);
  return COH_OK;
}

extern "C" int coh_panel_subset_stats(const coh_panel* p, int32_t nsets, const uint8_t* keep,
                                      float* mean, float* inv, int64_t* cnt, int64_t stride,
                                      void* ws, uint64_t ws_bytes, void* stream) {
  COH_REQUIRE(p, COH_ERR_VALUE, "NULL panel");
  COH_REQUIRE(p->finalized, COH_ERR_RUNTIME, "panel is not finalized");
  COH_REQUIRE(nsets >= 1, COH_ERR_VALUE, "nsets must be positive");
  COH_REQUIRE(keep && mean && inv && ws, COH_ERR_VALUE, "NULL argument");
  COH_REQUIRE(stride >= p->m_slot, COH_ERR_VALUE, "stride %lld < m_slot %lld", (long long)stride,
              (long long)p->m_slot);
  uint64_t need = 0;
  coh_subset_stats_workspace_bytes(p, nsets, &need);
  COH_REQUIRE(ws_bytes >= need, COH_ERR_VALUE, "workspace too small: %llu < %llu bytes",
              (unsigned long long)ws_bytes, (unsigned long long)need);
  hipStream_t s = (hipStream_t)stream;
  uint32_t* spread = static_cast<uint32_t*>(ws);
  const int64_t ndw = p->vbytes / 4;
  const int64_t total = (int64_t)nsets * ndw;
  hipLaunchKernelGGL(k_spread_masks,
                     dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, 256), 8192))),
                     dim3(256), 0, s, keep, p->n, (int64_t)nsets, ndw, spread);
  COH_CHECK_LAUNCH();
  dim3 grid((unsigned)ceil_div(p->m_slot, kSubSlots), (unsigned)ceil_div(nsets, kSubSets));
  hipLaunchKernelGGL(k_subset_stats, grid, dim3(256), 0, s, p->vmaj, p->vbytes, p->m_slot,
                     p->slot_to_cache, spread, (int64_t)nsets, mean, inv, cnt, stride);
  COH_CHECK_LAUNCH();
  return COH_OK;
}
