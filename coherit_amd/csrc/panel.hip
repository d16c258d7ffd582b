// Resident genotype panel: creation, ingest, generation, layout transposes
// and per-SNP statistics.  Replaces the packed tmpfile cache + pinned H2D
// ring of geno_stream.py (1167-1453, 148-327): the panel is written into HBM
// once, in two layouts:
//   smaj (sample-major): row i = sample, byte q = slots 4q..4q+3 at bits 2t
//        (the reference cache layout, geno_stream.py:885-904);
//   vmaj (variant-major): row j = slot, byte b = samples 4b..4b+3 at bits 2r
//        (the BED row layout with reference codes).
// Each operator phase reads the layout whose contraction axis is contiguous.
#include "common.hpp"
#include "stats.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstring>

namespace coh {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

int fail(int status, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return status;
}

// ---------------------------------------------------------------- kernels

// int8 variant-major rows -> vmaj codes + raw int64 stats
// (geno_stream.py:958-1007 stats loop, LUT geno_stream.py:1263-1270).
__global__ __launch_bounds__(256) void k_ingest_i8(const int8_t* __restrict__ block, int64_t ld,
                                                   int64_t n, int64_t count, int64_t cache_start,
                                                   const int32_t* __restrict__ cache_to_slot,
                                                   int32_t missing_raw, uint8_t* __restrict__ vmaj,
                                                   int64_t vbytes, int64_t* __restrict__ cnt,
                                                   int64_t* __restrict__ s1,
                                                   int64_t* __restrict__ s2,
                                                   uint8_t* __restrict__ has_stats) {
  const int64_t r = blockIdx.x;
  if (r >= count) return;
  const int64_t slot = cache_to_slot[cache_start + r];
  const int8_t* row = block + r * ld;
  const uint8_t mu8 = static_cast<uint8_t>(static_cast<int8_t>(missing_raw));
  int64_t c = 0, a1 = 0, a2 = 0;
  const int64_t nbytes = (n + 3) / 4;
  for (int64_t b = threadIdx.x; b < nbytes; b += blockDim.x) {
    uint32_t packed = 0;
    for (int t = 0; t < 4; ++t) {
      const int64_t i = 4 * b + t;
      if (i >= n) break;
      const int8_t v = row[i];
      const uint8_t u = static_cast<uint8_t>(v);
      uint32_t code = (u == mu8) ? 3u : (u == 1 ? 1u : (u == 2 ? 2u : 0u));
      packed |= code << (2 * t);
      if (v != static_cast<int8_t>(missing_raw)) {
        c += 1;
        a1 += v;
        a2 += static_cast<int64_t>(v) * v;
      }
    }
    vmaj[slot * vbytes + b] = static_cast<uint8_t>(packed);
  }
  __shared__ int64_t red[3][256];
  red[0][threadIdx.x] = c;
  red[1][threadIdx.x] = a1;
  red[2][threadIdx.x] = a2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
      red[2][threadIdx.x] += red[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cnt[slot] = red[0][0];
    s1[slot] = red[1][0];
    s2[slot] = red[2][0];
    has_stats[slot] = 1;
  }
}

// Raw BED rows -> vmaj codes for the kept samples
// (geno_stream.py:748-848; BED code -> value LUT [2,3,1,0]).
__global__ __launch_bounds__(256) void k_ingest_bed(const uint8_t* __restrict__ rows, int64_t bps,
                                                    int64_t n, int64_t count, int64_t cache_start,
                                                    const int32_t* __restrict__ cache_to_slot,
                                                    const int32_t* __restrict__ sample_idx,
                                                    uint8_t* __restrict__ vmaj, int64_t vbytes) {
  const int64_t nbytes = (n + 3) / 4;
  const int64_t total = count * nbytes;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx / nbytes, b = idx % nbytes;
    const uint8_t* row = rows + r * bps;
    uint32_t packed = 0;
    for (int t = 0; t < 4; ++t) {
      const int64_t i = 4 * b + t;
      if (i >= n) break;
      const int64_t fi = sample_idx[i];
      const uint32_t g = (row[fi >> 2] >> ((fi & 3) << 1)) & 3u;
      const uint32_t code = (0x1Eu >> (2 * g)) & 3u;  // [2,3,1,0] packed: 2 | 3<<2 | 1<<4 | 0<<6
      packed |= code << (2 * t);
    }
    vmaj[(int64_t)cache_to_slot[cache_start + r] * vbytes + b] = static_cast<uint8_t>(packed);
  }
}

__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Counter-based synthetic genotypes straight into vmaj.  Sample i of SNP j:
// genotype word (i&3) of Philox((i>>2, j, 0, 0)), missingness word (i&3) of
// Philox((i>>2, j, 1, 0)); key = (seed lo, seed hi).
__global__ __launch_bounds__(256) void k_generate(uint64_t seed, int64_t n, int64_t m,
                                                  const uint32_t* __restrict__ snp_ids,
                                                  const uint32_t* __restrict__ t0,
                                                  const uint32_t* __restrict__ t1, uint32_t tmiss,
                                                  const int32_t* __restrict__ cache_to_slot,
                                                  uint8_t* __restrict__ vmaj, int64_t vbytes) {
  const int64_t nbytes = (n + 3) / 4;
  const int64_t total = m * nbytes;
  const uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = idx / nbytes, q = idx % nbytes;
    const uint32_t j = snp_ids[s];
    uint32_t g[4] = {static_cast<uint32_t>(q), j, 0u, 0u};
    uint32_t mm[4] = {static_cast<uint32_t>(q), j, 1u, 0u};
    philox4x32_10(g, k0, k1);
    philox4x32_10(mm, k0, k1);
    const uint32_t a = t0[s], b = t1[s];
    uint32_t packed = 0;
    for (int w = 0; w < 4; ++w) {
      if (4 * q + w >= n) break;
      uint32_t code = g[w] < a ? 0u : (g[w] < b ? 1u : 2u);
      if (mm[w] < tmiss) code = 3u;
      packed |= code << (2 * w);
    }
    vmaj[(int64_t)cache_to_slot[s] * vbytes + q] = static_cast<uint8_t>(packed);
  }
}

// 256 slots x 256 samples tile transposes between the two layouts.
__global__ __launch_bounds__(256) void k_vmaj_to_smaj(const uint8_t* __restrict__ vmaj,
                                                      int64_t vbytes, uint8_t* __restrict__ smaj,
                                                      int64_t sbytes, int64_t n) {
  __shared__ uint8_t tile[256][64 + 4];
  const int64_t slot0 = (int64_t)blockIdx.x * 256, samp0 = (int64_t)blockIdx.y * 256;
  for (int idx = threadIdx.x; idx < 256 * 16; idx += 256) {
    const int r = idx / 16, d = idx % 16;
    const uint32_t w =
        *reinterpret_cast<const uint32_t*>(vmaj + (slot0 + r) * vbytes + samp0 / 4 + 4 * d);
    tile[r][4 * d + 0] = w & 0xff;
    tile[r][4 * d + 1] = (w >> 8) & 0xff;
    tile[r][4 * d + 2] = (w >> 16) & 0xff;
    tile[r][4 * d + 3] = (w >> 24) & 0xff;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 256 * 16; idx += 256) {
    const int i = idx / 16, d = idx % 16;  // sample i of tile, output dword d (bytes 4d..4d+3)
    if (samp0 + i >= n) continue;
    uint32_t out = 0;
    for (int bq = 0; bq < 4; ++bq) {
      const int q = 4 * d + bq;  // output byte: slots 4q..4q+3
      uint32_t byte = 0;
      for (int t = 0; t < 4; ++t) {
        const uint32_t code = (tile[4 * q + t][i >> 2] >> (2 * (i & 3))) & 3u;
        byte |= code << (2 * t);
      }
      out |= byte << (8 * bq);
    }
    *reinterpret_cast<uint32_t*>(smaj + (samp0 + i) * sbytes + slot0 / 4 + 4 * d) = out;
  }
}

__global__ __launch_bounds__(256) void k_smaj_to_vmaj(const uint8_t* __restrict__ smaj,
                                                      int64_t sbytes, uint8_t* __restrict__ vmaj,
                                                      int64_t vbytes, int64_t n) {
  __shared__ uint8_t tile[256][64 + 4];
  const int64_t slot0 = (int64_t)blockIdx.x * 256, samp0 = (int64_t)blockIdx.y * 256;
  for (int idx = threadIdx.x; idx < 256 * 16; idx += 256) {
    const int i = idx / 16, d = idx % 16;
    uint32_t w = 0;
    if (samp0 + i < n)
      w = *reinterpret_cast<const uint32_t*>(smaj + (samp0 + i) * sbytes + slot0 / 4 + 4 * d);
    tile[i][4 * d + 0] = w & 0xff;
    tile[i][4 * d + 1] = (w >> 8) & 0xff;
    tile[i][4 * d + 2] = (w >> 16) & 0xff;
    tile[i][4 * d + 3] = (w >> 24) & 0xff;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 256 * 16; idx += 256) {
    const int j = idx / 16, d = idx % 16;  // slot j of tile, output dword d
    uint32_t out = 0;
    for (int bb = 0; bb < 4; ++bb) {
      const int b = 4 * d + bb;  // output byte: samples 4b..4b+3
      uint32_t byte = 0;
      for (int r = 0; r < 4; ++r) {
        const uint32_t code = (tile[4 * b + r][j >> 2] >> (2 * (j & 3))) & 3u;
        byte |= code << (2 * r);
      }
      out |= byte << (8 * bb);
    }
    *reinterpret_cast<uint32_t*>(vmaj + (slot0 + j) * vbytes + samp0 / 4 + 4 * d) = out;
  }
}

// Per-slot integer stats from vmaj codes (popcounts) + float32 epilogue.
__global__ __launch_bounds__(256) void k_stats(const uint8_t* __restrict__ vmaj, int64_t vbytes,
                                               int64_t n, const int32_t* __restrict__ slot_to_cache,
                                               const uint8_t* __restrict__ has_stats,
                                               int64_t* __restrict__ cnt, int64_t* __restrict__ s1,
                                               int64_t* __restrict__ s2, float* __restrict__ mean,
                                               float* __restrict__ inv, int use_override) {
  const int64_t slot = blockIdx.x;
  const bool real = slot_to_cache[slot] >= 0;
  __shared__ int64_t red[3][256];
  int64_t c1 = 0, c2 = 0, c3 = 0;
  if (real && !has_stats[slot]) {
    const uint32_t* row = reinterpret_cast<const uint32_t*>(vmaj + slot * vbytes);
    for (int64_t d = threadIdx.x; d < vbytes / 4; d += blockDim.x) {
      const uint32_t w = row[d];
      const uint32_t lo = w & 0x55555555u, hi = (w >> 1) & 0x55555555u;
      c1 += __popc(lo & ~hi);
      c2 += __popc(hi & ~lo);
      c3 += __popc(lo & hi);
    }
  }
  red[0][threadIdx.x] = c1;
  red[1][threadIdx.x] = c2;
  red[2][threadIdx.x] = c3;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
      red[2][threadIdx.x] += red[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (!real) {
    cnt[slot] = s1[slot] = s2[slot] = 0;
    mean[slot] = 0.0f;
    inv[slot] = 0.0f;
    return;
  }
  if (!has_stats[slot]) {
    cnt[slot] = n - red[2][0];
    s1[slot] = red[0][0] + 2 * red[1][0];
    s2[slot] = red[0][0] + 4 * red[1][0];
  }
  if (!use_override) f32_stats(cnt[slot], s1[slot], s2[slot], &mean[slot], &inv[slot]);
}

__global__ void k_override(const float* __restrict__ mo, const float* __restrict__ io, int64_t m,
                           const int32_t* __restrict__ cache_to_slot, float* __restrict__ mean,
                           float* __restrict__ inv) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < m;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int64_t slot = cache_to_slot[s];
    mean[slot] = mo[s];
    inv[slot] = io[s];
  }
}

__global__ void k_gather_stats(int64_t m, const int32_t* __restrict__ cache_to_slot,
                               const float* __restrict__ mean, const float* __restrict__ inv,
                               const int64_t* __restrict__ cnt, const int64_t* __restrict__ s1,
                               const int64_t* __restrict__ s2, float* om, float* oi, int64_t* oc,
                               int64_t* o1, int64_t* o2) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < m;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int64_t slot = cache_to_slot[s];
    if (om) om[s] = mean[slot];
    if (oi) oi[s] = inv[slot];
    if (oc) oc[s] = cnt[slot];
    if (o1) o1[s] = s1[slot];
    if (o2) o2[s] = s2[slot];
  }
}

// Per-sample sum_j z_ij^2 (geno_stream.py:929-955), FP64 accumulation.
__global__ __launch_bounds__(256) void k_exact_diag(const uint8_t* __restrict__ smaj,
                                                    int64_t sbytes, const float* __restrict__ mean,
                                                    const float* __restrict__ inv, double inv_eff,
                                                    float* __restrict__ out) {
  const int64_t i = blockIdx.x;
  const uint8_t* row = smaj + i * sbytes;
  double acc = 0.0;
  for (int64_t q = threadIdx.x; q < sbytes; q += blockDim.x) {
    const uint32_t b = row[q];
    for (int t = 0; t < 4; ++t) {
      const uint32_t code = (b >> (2 * t)) & 3u;
      const int64_t j = 4 * q + t;
      if (code == 3u) continue;
      const double z = (static_cast<double>(code) - static_cast<double>(mean[j])) *
                       static_cast<double>(inv[j]);
      acc += z * z;
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = static_cast<float>(red[0] * inv_eff);
}

}  // namespace coh

using namespace coh;

extern "C" const char* coh_last_error(void) { return g_err.c_str(); }
extern "C" int coh_version(void) { return 1; }

extern "C" int coh_panel_create(int64_t n, int32_t n_calls, const int64_t* call_snp_starts,
                                const int32_t* call_true_widths,
                                const int32_t* call_component_ids, int32_t n_components,
                                coh_panel** out) {
  COH_REQUIRE(out != nullptr, COH_ERR_VALUE, "out handle pointer is NULL");
  *out = nullptr;
  COH_REQUIRE(n > 0, COH_ERR_VALUE, "panel requires at least one sample (n=%lld)", (long long)n);
  COH_REQUIRE(n <= (int64_t)1 << 30, COH_ERR_VALUE, "n=%lld exceeds the supported sample count",
              (long long)n);
  COH_REQUIRE(n_calls > 0, COH_ERR_VALUE, "panel requires at least one call");
  COH_REQUIRE(n_components > 0, COH_ERR_VALUE, "panel requires at least one component");
  COH_REQUIRE(call_snp_starts && call_true_widths && call_component_ids, COH_ERR_VALUE,
              "call geometry arrays must not be NULL");
  auto p = new coh_panel();
  p->n = n;
  p->n_calls = n_calls;
  p->G = n_components;
  p->call_start.assign(call_snp_starts, call_snp_starts + n_calls);
  p->call_width.assign(call_true_widths, call_true_widths + n_calls);
  p->call_comp.assign(call_component_ids, call_component_ids + n_calls);
  p->call_slot.resize(n_calls);
  p->comp_slot_begin.assign(n_components, 0);
  p->comp_slot_end.assign(n_components, 0);
  int64_t expect_start = 0;
  int32_t prev_comp = 0;
  for (int32_t c = 0; c < n_calls; ++c) {
    const int32_t g = p->call_comp[c];
    if (p->call_width[c] <= 0 || p->call_start[c] != expect_start || g < prev_comp ||
        g >= n_components) {
      delete p;
      return fail(COH_ERR_VALUE,
                  "invalid call plan at call %d (start=%lld width=%d component=%d)", c,
                  (long long)call_snp_starts[c], call_true_widths[c], g);
    }
    expect_start += p->call_width[c];
    prev_comp = g;
  }
  p->m = expect_start;
  int64_t slot = 0;
  int32_t c = 0;
  for (int32_t g = 0; g < n_components; ++g) {
    p->comp_slot_begin[g] = slot;
    while (c < n_calls && p->call_comp[c] == g) {
      p->call_slot[c] = slot;
      slot += 4 * ceil_div(p->call_width[c], 4);
      ++c;
    }
    slot = round_up(slot, kSlotAlign);
    p->comp_slot_end[g] = slot;
  }
  p->m_slot = slot;
  p->sbytes = p->m_slot / 4;
  p->vbytes = round_up(n, kSampleAlign) / 4;
  std::vector<int32_t> s2c(p->m_slot, -1);
  p->h_cache_to_slot.assign(p->m, 0);
  for (int32_t cc = 0; cc < n_calls; ++cc)
    for (int32_t t = 0; t < p->call_width[cc]; ++t) {
      const int64_t s = p->call_start[cc] + t, sl = p->call_slot[cc] + t;
      s2c[sl] = static_cast<int32_t>(s);
      p->h_cache_to_slot[s] = static_cast<int32_t>(sl);
    }
  hipGetDevice(&p->device);
  auto cleanup_fail = [&](hipError_t e, const char* what) {
    int st = fail(COH_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
    coh_panel_destroy(p);
    return st;
  };
  hipError_t e;
  if ((e = hipMalloc(&p->smaj, (size_t)p->n * p->sbytes)) != hipSuccess) return cleanup_fail(e, "hipMalloc(smaj)");
  if ((e = hipMalloc(&p->vmaj, (size_t)p->m_slot * p->vbytes)) != hipSuccess) return cleanup_fail(e, "hipMalloc(vmaj)");
  if ((e = hipMalloc(&p->mean, p->m_slot * sizeof(float))) != hipSuccess) return cleanup_fail(e, "hipMalloc(mean)");
  if ((e = hipMalloc(&p->inv, p->m_slot * sizeof(float))) != hipSuccess) return cleanup_fail(e, "hipMalloc(inv)");
  if ((e = hipMalloc(&p->cnt, p->m_slot * sizeof(int64_t))) != hipSuccess) return cleanup_fail(e, "hipMalloc(cnt)");
  if ((e = hipMalloc(&p->s1, p->m_slot * sizeof(int64_t))) != hipSuccess) return cleanup_fail(e, "hipMalloc(s1)");
  if ((e = hipMalloc(&p->s2, p->m_slot * sizeof(int64_t))) != hipSuccess) return cleanup_fail(e, "hipMalloc(s2)");
  if ((e = hipMalloc(&p->slot_to_cache, p->m_slot * sizeof(int32_t))) != hipSuccess) return cleanup_fail(e, "hipMalloc(slot_to_cache)");
  if ((e = hipMalloc(&p->cache_to_slot, p->m * sizeof(int32_t))) != hipSuccess) return cleanup_fail(e, "hipMalloc(cache_to_slot)");
  if ((e = hipMalloc(&p->row_has_stats, p->m_slot)) != hipSuccess) return cleanup_fail(e, "hipMalloc(row_has_stats)");
  if ((e = hipMemset(p->smaj, 0, (size_t)p->n * p->sbytes)) != hipSuccess) return cleanup_fail(e, "hipMemset(smaj)");
  if ((e = hipMemset(p->vmaj, 0, (size_t)p->m_slot * p->vbytes)) != hipSuccess) return cleanup_fail(e, "hipMemset(vmaj)");
  if ((e = hipMemset(p->mean, 0, p->m_slot * sizeof(float))) != hipSuccess) return cleanup_fail(e, "hipMemset(mean)");
  if ((e = hipMemset(p->inv, 0, p->m_slot * sizeof(float))) != hipSuccess) return cleanup_fail(e, "hipMemset(inv)");
  if ((e = hipMemset(p->row_has_stats, 0, p->m_slot)) != hipSuccess) return cleanup_fail(e, "hipMemset(row_has_stats)");
  if ((e = hipMemcpy(p->slot_to_cache, s2c.data(), p->m_slot * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess) return cleanup_fail(e, "hipMemcpy(slot_to_cache)");
  if ((e = hipMemcpy(p->cache_to_slot, p->h_cache_to_slot.data(), p->m * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess) return cleanup_fail(e, "hipMemcpy(cache_to_slot)");
  *out = p;
  return COH_OK;
}

extern "C" int coh_panel_destroy(coh_panel* p) {
  if (!p) return COH_OK;
  for (void* ptr : {(void*)p->smaj, (void*)p->vmaj, (void*)p->mean, (void*)p->inv, (void*)p->cnt,
                    (void*)p->s1, (void*)p->s2, (void*)p->slot_to_cache,
                    (void*)p->cache_to_slot, (void*)p->row_has_stats})
    if (ptr) (void)hipFree(ptr);
  delete p;
  return COH_OK;
}

extern "C" int coh_panel_info(const coh_panel* p, int64_t* info) {
  COH_REQUIRE(p && info, COH_ERR_VALUE, "NULL panel or info pointer");
  info[0] = p->n;
  info[1] = p->m;
  info[2] = p->n_calls;
  info[3] = p->G;
  info[4] = p->m_slot;
  info[5] = p->sbytes;
  info[6] = p->vbytes;
  info[7] = p->finalized ? 1 : 0;
  return COH_OK;
}

extern "C" int coh_panel_component_slots(const coh_panel* p, int64_t* begin, int64_t* end) {
  COH_REQUIRE(p && begin && end, COH_ERR_VALUE, "NULL argument");
  std::copy(p->comp_slot_begin.begin(), p->comp_slot_begin.end(), begin);
  std::copy(p->comp_slot_end.begin(), p->comp_slot_end.end(), end);
  return COH_OK;
}

extern "C" int coh_panel_call_slots(const coh_panel* p, int64_t* call_slot) {
  COH_REQUIRE(p && call_slot, COH_ERR_VALUE, "NULL argument");
  std::copy(p->call_slot.begin(), p->call_slot.end(), call_slot);
  return COH_OK;
}

extern "C" int coh_panel_ingest_packed_call(coh_panel* p, int32_t c, const uint8_t* block,
                                            int64_t ld, void* stream) {
  COH_REQUIRE(p && block, COH_ERR_VALUE, "NULL panel or block");
  COH_REQUIRE(!p->finalized, COH_ERR_RUNTIME, "panel is already finalized");
  COH_REQUIRE(c >= 0 && c < p->n_calls, COH_ERR_INDEX, "call index %d out of range", c);
  const int64_t pw = ceil_div(p->call_width[c], 4);
  COH_REQUIRE(ld >= pw, COH_ERR_VALUE, "row stride %lld < packed width %lld", (long long)ld,
              (long long)pw);
  COH_HIP(hipMemcpy2DAsync(p->smaj + p->call_slot[c] / 4, p->sbytes, block, ld, pw, p->n,
                           hipMemcpyDeviceToDevice, (hipStream_t)stream));
  p->smaj_written = true;
  return COH_OK;
}

extern "C" int coh_panel_ingest_i8_varmaj(coh_panel* p, int64_t start, int64_t count,
                                          const int8_t* block, int64_t ld, int32_t missing_raw,
                                          void* stream) {
  COH_REQUIRE(p && block, COH_ERR_VALUE, "NULL panel or block");
  COH_REQUIRE(!p->finalized, COH_ERR_RUNTIME, "panel is already finalized");
  COH_REQUIRE(start >= 0 && count >= 0 && start + count <= p->m, COH_ERR_INDEX,
              "SNP rows [%lld, %lld) out of range for m=%lld", (long long)start,
              (long long)(start + count), (long long)p->m);
  COH_REQUIRE(ld >= p->n, COH_ERR_VALUE, "row stride %lld < n", (long long)ld);
  if (count == 0) return COH_OK;
  hipLaunchKernelGGL(k_ingest_i8, dim3((unsigned)count), dim3(256), 0, (hipStream_t)stream, block,
                     ld, p->n, count, start, p->cache_to_slot, missing_raw, p->vmaj, p->vbytes,
                     p->cnt, p->s1, p->s2, p->row_has_stats);
  COH_CHECK_LAUNCH();
  p->vmaj_written = true;
  return COH_OK;
}

extern "C" int coh_panel_ingest_bed(coh_panel* p, int64_t start, int64_t count,
                                    const uint8_t* rows, int64_t bps,
                                    const int32_t* sample_idx, void* stream) {
  COH_REQUIRE(p && rows && sample_idx, COH_ERR_VALUE, "NULL argument");
  COH_REQUIRE(!p->finalized, COH_ERR_RUNTIME, "panel is already finalized");
  COH_REQUIRE(start >= 0 && count >= 0 && start + count <= p->m, COH_ERR_INDEX,
              "SNP rows out of range");
  COH_REQUIRE(bps > 0, COH_ERR_VALUE, "bytes_per_snp must be positive");
  if (count == 0) return COH_OK;
  const int64_t total = count * ceil_div(p->n, 4);
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(total, 256), 65536);
  hipLaunchKernelGGL(k_ingest_bed, dim3(grid), dim3(256), 0, (hipStream_t)stream, rows, bps, p->n,
                     count, start, p->cache_to_slot, sample_idx, p->vmaj, p->vbytes);
  COH_CHECK_LAUNCH();
  p->vmaj_written = true;
  return COH_OK;
}

extern "C" int coh_panel_generate(coh_panel* p, uint64_t seed, const uint32_t* snp_ids,
                                  const uint32_t* t0, const uint32_t* t1, uint32_t tmiss,
                                  void* stream) {
  COH_REQUIRE(p && snp_ids && t0 && t1, COH_ERR_VALUE, "NULL argument");
  COH_REQUIRE(!p->finalized, COH_ERR_RUNTIME, "panel is already finalized");
  const int64_t total = p->m * ceil_div(p->n, 4);
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(total, 256), 1 << 20);
  hipLaunchKernelGGL(k_generate, dim3(grid), dim3(256), 0, (hipStream_t)stream, seed, p->n, p->m,
                     snp_ids, t0, t1, tmiss, p->cache_to_slot, p->vmaj, p->vbytes);
  COH_CHECK_LAUNCH();
  p->vmaj_written = true;
  return COH_OK;
}

extern "C" int coh_panel_finalize(coh_panel* p, const float* mo, const float* io, void* stream) {
  COH_REQUIRE(p, COH_ERR_VALUE, "NULL panel");
  COH_REQUIRE(!p->finalized, COH_ERR_RUNTIME, "panel is already finalized");
  COH_REQUIRE(p->smaj_written != p->vmaj_written, COH_ERR_RUNTIME,
              p->smaj_written ? "panel was written through both packed calls and row ingest"
                              : "no genotypes were ingested");
  COH_REQUIRE((mo == nullptr) == (io == nullptr), COH_ERR_VALUE,
              "mean and inv_sd overrides must be given together");
  hipStream_t s = (hipStream_t)stream;
  dim3 tgrid((unsigned)(p->m_slot / 256), (unsigned)(p->vbytes * 4 / 256));
  if (p->vmaj_written)
    hipLaunchKernelGGL(k_vmaj_to_smaj, tgrid, dim3(256), 0, s, p->vmaj, p->vbytes, p->smaj,
                       p->sbytes, p->n);
  else
    hipLaunchKernelGGL(k_smaj_to_vmaj, tgrid, dim3(256), 0, s, p->smaj, p->sbytes, p->vmaj,
                       p->vbytes, p->n);
  COH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_stats, dim3((unsigned)p->m_slot), dim3(256), 0, s, p->vmaj, p->vbytes, p->n,
                     p->slot_to_cache, p->row_has_stats, p->cnt, p->s1, p->s2, p->mean, p->inv,
                     mo ? 1 : 0);
  COH_CHECK_LAUNCH();
  if (mo) {
    hipLaunchKernelGGL(k_override, dim3(256), dim3(256), 0, s, mo, io, p->m, p->cache_to_slot,
                       p->mean, p->inv);
    COH_CHECK_LAUNCH();
  }
  p->finalized = true;
  return COH_OK;
}

extern "C" int coh_panel_copy_stats(const coh_panel* p, float* om, float* oi, int64_t* oc,
                                    int64_t* o1, int64_t* o2, void* stream) {
  COH_REQUIRE(p, COH_ERR_VALUE, "NULL panel");
  COH_REQUIRE(p->finalized, COH_ERR_RUNTIME, "panel statistics are not available before finalize");
  hipLaunchKernelGGL(k_gather_stats, dim3(256), dim3(256), 0, (hipStream_t)stream, p->m,
                     p->cache_to_slot, p->mean, p->inv, p->cnt, p->s1, p->s2, om, oi, oc, o1, o2);
  COH_CHECK_LAUNCH();
  return COH_OK;
}

extern "C" int coh_panel_set_standardization(coh_panel* p, const float* mo, const float* io,
                                             void* stream) {
  COH_REQUIRE(p && mo && io, COH_ERR_VALUE, "NULL argument");
  COH_REQUIRE(p->finalized, COH_ERR_RUNTIME, "panel is not finalized");
  hipLaunchKernelGGL(k_override, dim3(256), dim3(256), 0, (hipStream_t)stream, mo, io, p->m,
                     p->cache_to_slot, p->mean, p->inv);
  COH_CHECK_LAUNCH();
  return COH_OK;
}

extern "C" int coh_panel_read_packed_call(const coh_panel* p, int32_t c, uint8_t* out, int64_t ld,
                                          void* stream) {
  COH_REQUIRE(p && out, COH_ERR_VALUE, "NULL argument");
  COH_REQUIRE(p->finalized, COH_ERR_RUNTIME, "panel is not finalized");
  COH_REQUIRE(c >= 0 && c < p->n_calls, COH_ERR_INDEX, "call index %d out of range", c);
  const int64_t pw = ceil_div(p->call_width[c], 4);
  COH_REQUIRE(ld >= pw, COH_ERR_VALUE, "row stride too small");
  COH_HIP(hipMemcpy2DAsync(out, ld, p->smaj + p->call_slot[c] / 4, p->sbytes, pw, p->n,
                           hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return COH_OK;
}

extern "C" int coh_exact_diag(const coh_panel* p, double eff, float* out, void* stream) {
  COH_REQUIRE(p && out, COH_ERR_VALUE, "NULL argument");
  COH_REQUIRE(p->finalized, COH_ERR_RUNTIME, "panel is not finalized");
  const double inv_eff = eff > 0.0 ? 1.0 / eff : 0.0;
  hipLaunchKernelGGL(k_exact_diag, dim3((unsigned)p->n), dim3(256), 0, (hipStream_t)stream,
                     p->smaj, p->sbytes, p->mean, p->inv, inv_eff, out);
  COH_CHECK_LAUNCH();
  return COH_OK;
}
