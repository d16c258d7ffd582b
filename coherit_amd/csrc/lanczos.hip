// Re-orthogonalisation primitives of the batched SLQ Lanczos (slq.py:49-74)
// and of its reverse sweep (slq.py:77-81, 124-140).
//
// Vectors are probe-major: basis[s][j][i] = Lanczos vector j of probe s
// (element stride 1, vector stride `vs`, probe stride `ps`), and a block of
// per-probe vectors X[s][i] has row stride ldx.  Each step of the two-pass
// full re-orthogonalisation is
//     c[s][j] = sum_i P[s][j][i] X[s][i]          (coh_probe_dots)
//     Y[s][i] = X[s][i] - sum_j P[s][j][i] c[s][j] (coh_probe_project_out)
// and the reverse sweep adds a rank-q update to the basis cotangent
//     B[s][j][i] += sum_q coeff[s][j][q] V[s][q][i] (coh_probe_rank_update).
// These are GEMV-shaped (r <= a few hundred, n ~ 1e5..1e6) and bound by
// streaming the basis from HBM: one block per (probe, 2048-element chunk),
// coalesced along i.  Dot products accumulate FP32 products in FP64 with a
// fixed two-stage order (deterministic).
//
// The rest of the step (coh_probe_step_prep / coh_probe_step_finish) and of
// its adjoint (coh_probe_step_adjoint_head / _tail) are fused element-wise
// passes with the same chunking and FP64 partial sums, so one step takes a
// dozen launches instead of the ~45 of its element-wise form.
#include "common.hpp"

#include <algorithm>

namespace coh {

constexpr int kEPT = 8;               // elements per thread
constexpr int kChunk = 256 * kEPT;    // elements per block
constexpr int kJT = 8;                // basis vectors per register tile
constexpr int kRedAll = 128;          // vectors whose partials fit LDS at once

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Body of k_probe_dots_part; FULL: the whole chunk lies inside [0, n).
template <bool FULL>
__device__ inline void probe_dots_body(int64_t n, int r, const float* __restrict__ P,
                                       int64_t ps, int64_t vs, const float* __restrict__ X,
                                       int64_t ldx, double* __restrict__ part, int nch) {
  const int s = blockIdx.y, ch = blockIdx.x, t = threadIdx.x;
  const int wave = t / 64, lane = t % 64;
  const int64_t i0 = (int64_t)ch * kChunk + t;
  float x[kEPT];
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const int64_t i = i0 + 256 * e;
    x[e] = (FULL || i < n) ? X[(int64_t)s * ldx + i] : 0.0f;
  }
  const float* Ps = P + (int64_t)s * ps;
  if (r <= kRedAll) {
    // every vector's wave partials go to LDS; one barrier at the end (the
    // same per-wave and four-wave sums as the tiled loop below)
    __shared__ double all[kRedAll][4];
    for (int j0 = 0; j0 < r; j0 += kJT) {
      double acc[kJT];
#pragma unroll
      for (int jj = 0; jj < kJT; ++jj) {
        acc[jj] = 0.0;
        const int j = j0 + jj;
        if (j < r) {
#pragma unroll
          for (int e = 0; e < kEPT; ++e) {
            const int64_t i = i0 + 256 * e;
            if (FULL || i < n) acc[jj] += static_cast<double>(Ps[(int64_t)j * vs + i]) * static_cast<double>(x[e]);
          }
        }
      }
#pragma unroll
      for (int jj = 0; jj < kJT; ++jj) {
        const double v = wave_sum(acc[jj]);
        if (lane == 0 && j0 + jj < r) all[j0 + jj][wave] = v;
      }
    }
    __syncthreads();
    if (t < r) part[((int64_t)s * r + t) * nch + ch] = all[t][0] + all[t][1] + all[t][2] + all[t][3];
    return;
  }
  __shared__ double red[kJT][4];
  for (int j0 = 0; j0 < r; j0 += kJT) {
    double acc[kJT];
#pragma unroll
    for (int jj = 0; jj < kJT; ++jj) {
      acc[jj] = 0.0;
      const int j = j0 + jj;
      if (j < r) {
#pragma unroll
        for (int e = 0; e < kEPT; ++e) {
          const int64_t i = i0 + 256 * e;
          if (FULL || i < n) acc[jj] += static_cast<double>(Ps[(int64_t)j * vs + i]) * static_cast<double>(x[e]);
        }
      }
    }
#pragma unroll
    for (int jj = 0; jj < kJT; ++jj) {
      const double v = wave_sum(acc[jj]);
      if (lane == 0) red[jj][wave] = v;
    }
    __syncthreads();
    if (t < kJT && j0 + t < r)
      part[((int64_t)s * r + j0 + t) * nch + ch] = red[t][0] + red[t][1] + red[t][2] + red[t][3];
    __syncthreads();
  }
}

// part[(s * r + j) * nch + ch] = sum over chunk ch of P[s][j][i] X[s][i];
// chunks wholly inside [0, n) take the body without per-element checks
__global__ __launch_bounds__(256) void k_probe_dots_part(int64_t n, int r, const float* __restrict__ P,
                                                         int64_t ps, int64_t vs,
                                                         const float* __restrict__ X, int64_t ldx,
                                                         double* __restrict__ part, int nch) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= n)
    probe_dots_body<true>(n, r, P, ps, vs, X, ldx, part, nch);
  else
    probe_dots_body<false>(n, r, P, ps, vs, X, ldx, part, nch);
}

__global__ void k_probe_dots_final(int64_t S, int r, int nch, const double* __restrict__ part,
                                   float* __restrict__ C, int64_t ldc) {
  const int64_t total = S * r;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = idx / r;
    const int j = static_cast<int>(idx % r);
    double sum = 0.0;
    for (int ch = 0; ch < nch; ++ch) sum += part[idx * nch + ch];
    C[s * ldc + j] = static_cast<float>(sum);
  }
}

// Y[s][i] = X[s][i] - sum_j P[s][j][i] C[s][j]; Y may alias X.
template <bool FULL>
__device__ inline void probe_project_body(int64_t n, int r, const float* __restrict__ P,
                                          int64_t ps, int64_t vs, const float* __restrict__ C,
                                          int64_t ldc, const float* X, int64_t ldx, float* Y,
                                          int64_t ldy) {
  const int s = blockIdx.y;
  const int64_t i0 = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  const float* Ps = P + (int64_t)s * ps;
  const float* Cs = C + (int64_t)s * ldc;
  float acc[kEPT];
#pragma unroll
  for (int e = 0; e < kEPT; ++e) acc[e] = 0.0f;
#pragma unroll 4
  for (int j = 0; j < r; ++j) {
    const float cj = Cs[j];
    const float* Pj = Ps + (int64_t)j * vs;
#pragma unroll
    for (int e = 0; e < kEPT; ++e) {
      const int64_t i = i0 + 256 * e;
      if (FULL || i < n) acc[e] = fmaf(Pj[i], cj, acc[e]);
    }
  }
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const int64_t i = i0 + 256 * e;
    if (FULL || i < n) Y[(int64_t)s * ldy + i] = X[(int64_t)s * ldx + i] - acc[e];
  }
}

__global__ __launch_bounds__(256) void k_probe_project_out(int64_t n, int r, const float* __restrict__ P,
                                                           int64_t ps, int64_t vs,
                                                           const float* __restrict__ C, int64_t ldc,
                                                           const float* X, int64_t ldx, float* Y,
                                                           int64_t ldy) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= n)
    probe_project_body<true>(n, r, P, ps, vs, C, ldc, X, ldx, Y, ldy);
  else
    probe_project_body<false>(n, r, P, ps, vs, C, ldc, X, ldx, Y, ldy);
}

// B[s][j][i] += sum_q coeff[s][j][q] V[s][q][i] for j < r, q < Q <= 4.
template <bool FULL>
__device__ inline void probe_rank_body(int64_t n, int r, int Q, const float* __restrict__ coeff,
                                       int64_t cps, int64_t cvs, const float* __restrict__ V,
                                       int64_t vps, int64_t vvs, float* __restrict__ B,
                                       int64_t bps, int64_t bvs) {
  const int s = blockIdx.y;
  const int64_t i0 = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  float v[4][kEPT];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int e = 0; e < kEPT; ++e) {
      const int64_t i = i0 + 256 * e;
      v[q][e] = (q < Q && (FULL || i < n)) ? V[(int64_t)s * vps + (int64_t)q * vvs + i] : 0.0f;
    }
  float* Bs = B + (int64_t)s * bps;
  const float* Cs = coeff + (int64_t)s * cps;
#pragma unroll 4
  for (int j = 0; j < r; ++j) {
    float c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) c[q] = (q < Q) ? Cs[(int64_t)j * cvs + q] : 0.0f;
    float* Bj = Bs + (int64_t)j * bvs;
#pragma unroll
    for (int e = 0; e < kEPT; ++e) {
      const int64_t i = i0 + 256 * e;
      if (FULL || i < n) {
        float add = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q) add = fmaf(c[q], v[q][e], add);
        Bj[i] += add;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_probe_rank_update(int64_t n, int r, int Q,
                                                           const float* __restrict__ coeff,
                                                           int64_t cps, int64_t cvs,
                                                           const float* __restrict__ V, int64_t vps,
                                                           int64_t vvs, float* __restrict__ B,
                                                           int64_t bps, int64_t bvs) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= n)
    probe_rank_body<true>(n, r, Q, coeff, cps, cvs, V, vps, vvs, B, bps, bvs);
  else
    probe_rank_body<false>(n, r, Q, coeff, cps, cvs, V, vps, vvs, B, bps, bvs);
}

static int check_dims(int64_t S, int64_t r, int64_t n) {
  if (S < 0 || r < 0 || n < 0) return fail(COH_ERR_VALUE, "negative dimension");
  if (S > 65535) return fail(COH_ERR_VALUE, "at most 65535 probes per call");
  if (r > (1 << 20)) return fail(COH_ERR_VALUE, "too many basis vectors");
  return COH_OK;
}

// ----------------------------------------------------- fused Lanczos step
//
// One re-orthogonalised Lanczos step (slq.py:49-74) around the two passes
// above, with the step's reductions in FP64 partials:
//   prep:   w0 = product - beta_prev * prev, alpha = q . w0, w1 = w0 - alpha q,
//           active = (q . q > 0), scale = |product|
//   (two re-orthogonalisation passes: w1 -> w2 -> w3)
//   finish: square = |w3|^2, norm = sqrt(square > 0 ? square : 1),
//           keep = active and square > (16 eps max(scale, |alpha|))^2,
//           beta = keep ? norm : 0, alpha_out = active ? alpha : 1,
//           next = keep ? w3 / norm : 0
// q = basis[s][index], prev = basis[s][index - 1] (zero at index 0).  The
// FP32 element-wise expressions are the reference's, without contraction.

constexpr float kEps32 = 1.1920928955078125e-07f;

// Sum of one double per thread over the block: result valid in thread 0.
__device__ inline void block_sums3(double a, double b, double c, double* out3) {
  __shared__ double red[3][4];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  if (lane == 0) {
    red[0][wave] = a;
    red[1][wave] = b;
    red[2][wave] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < 3; ++v) out3[v] = red[v][0] + red[v][1] + red[v][2] + red[v][3];
  }
}

template <bool FULL>
__device__ inline void lz_prep_part_body(int64_t n, int64_t index, const float* __restrict__ basis,
                                         int64_t ps, int64_t vs, const float* __restrict__ product,
                                         int64_t ldp, const float* __restrict__ prev_beta,
                                         double* __restrict__ part, int nch) {
#pragma clang fp contract(off)
  const int s = blockIdx.y, ch = blockIdx.x;
  const int64_t i0 = (int64_t)ch * kChunk + threadIdx.x;
  const float* q = basis + (int64_t)s * ps + index * vs;
  const float* pv = index > 0 ? q - vs : nullptr;
  const float* pr = product + (int64_t)s * ldp;
  const float b = prev_beta[s];
  double qq = 0.0, qw = 0.0, pp = 0.0;
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const int64_t i = i0 + 256 * e;
    if (FULL || i < n) {
      const float qi = q[i], pi = pr[i];
      const float w0 = pv ? pi - pv[i] * b : pi;
      qq += static_cast<double>(qi) * static_cast<double>(qi);
      qw += static_cast<double>(qi) * static_cast<double>(w0);
      pp += static_cast<double>(pi) * static_cast<double>(pi);
    }
  }
  double sums[3];
  block_sums3(qq, qw, pp, sums);
  if (threadIdx.x == 0) {
    double* dst = part + ((int64_t)s * nch + ch) * 3;
    dst[0] = sums[0];
    dst[1] = sums[1];
    dst[2] = sums[2];
  }
}

__global__ __launch_bounds__(256) void k_lz_prep_part(int64_t n, int64_t index,
                                                      const float* __restrict__ basis, int64_t ps,
                                                      int64_t vs, const float* __restrict__ product,
                                                      int64_t ldp, const float* __restrict__ prev_beta,
                                                      double* __restrict__ part, int nch) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= n)
    lz_prep_part_body<true>(n, index, basis, ps, vs, product, ldp, prev_beta, part, nch);
  else
    lz_prep_part_body<false>(n, index, basis, ps, vs, product, ldp, prev_beta, part, nch);
}

// One thread per probe: chunk partials summed in chunk order.
__global__ void k_lz_prep_final(int64_t S, int nch, const double* __restrict__ part,
                                float* __restrict__ alpha, uint8_t* __restrict__ active,
                                float* __restrict__ scale) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= S) return;
  double qq = 0.0, qw = 0.0, pp = 0.0;
  for (int ch = 0; ch < nch; ++ch) {
    const double* p = part + (s * nch + ch) * 3;
    qq += p[0];
    qw += p[1];
    pp += p[2];
  }
  alpha[s] = static_cast<float>(qw);
  active[s] = qq > 0.0 ? 1 : 0;
  scale[s] = sqrtf(static_cast<float>(pp));
}

template <bool FULL>
__device__ inline void lz_prep_w_body(int64_t n, int64_t index, const float* __restrict__ basis,
                                      int64_t ps, int64_t vs, const float* __restrict__ product,
                                      int64_t ldp, const float* __restrict__ prev_beta,
                                      const float* __restrict__ alpha, float* __restrict__ w0,
                                      float* __restrict__ w1, int64_t ldw) {
#pragma clang fp contract(off)
  const int s = blockIdx.y;
  const int64_t i0 = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  const float* q = basis + (int64_t)s * ps + index * vs;
  const float* pv = index > 0 ? q - vs : nullptr;
  const float* pr = product + (int64_t)s * ldp;
  const float b = prev_beta[s], a = alpha[s];
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const int64_t i = i0 + 256 * e;
    if (FULL || i < n) {
      const float v0 = pv ? pr[i] - pv[i] * b : pr[i];
      w0[(int64_t)s * ldw + i] = v0;
      w1[(int64_t)s * ldw + i] = v0 - q[i] * a;
    }
  }
}

__global__ __launch_bounds__(256) void k_lz_prep_w(int64_t n, int64_t index,
                                                   const float* __restrict__ basis, int64_t ps,
                                                   int64_t vs, const float* __restrict__ product,
                                                   int64_t ldp, const float* __restrict__ prev_beta,
                                                   const float* __restrict__ alpha,
                                                   float* __restrict__ w0, float* __restrict__ w1,
                                                   int64_t ldw) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= n)
    lz_prep_w_body<true>(n, index, basis, ps, vs, product, ldp, prev_beta, alpha, w0, w1, ldw);
  else
    lz_prep_w_body<false>(n, index, basis, ps, vs, product, ldp, prev_beta, alpha, w0, w1, ldw);
}

template <bool FULL>
__device__ inline void lz_sq_part_body(int64_t n, const float* __restrict__ w3, int64_t ldw,
                                       double* __restrict__ part, int nch) {
  const int s = blockIdx.y, ch = blockIdx.x;
  const int64_t i0 = (int64_t)ch * kChunk + threadIdx.x;
  double sq = 0.0;
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const int64_t i = i0 + 256 * e;
    if (FULL || i < n) {
      const double v = static_cast<double>(w3[(int64_t)s * ldw + i]);
      sq += v * v;
    }
  }
  double sums[3];
  block_sums3(sq, 0.0, 0.0, sums);
  if (threadIdx.x == 0) part[(int64_t)s * nch + ch] = sums[0];
}

__global__ __launch_bounds__(256) void k_lz_sq_part(int64_t n, const float* __restrict__ w3,
                                                    int64_t ldw, double* __restrict__ part, int nch) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= n)
    lz_sq_part_body<true>(n, w3, ldw, part, nch);
  else
    lz_sq_part_body<false>(n, w3, ldw, part, nch);
}

__global__ void k_lz_finish_final(int64_t S, int nch, const double* __restrict__ part,
                                  const float* __restrict__ alpha, const uint8_t* __restrict__ active,
                                  const float* __restrict__ scale, float* __restrict__ norm,
                                  uint8_t* __restrict__ keep, float* __restrict__ beta,
                                  float* __restrict__ alpha_out) {
#pragma clang fp contract(off)
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= S) return;
  double sq = 0.0;
  for (int ch = 0; ch < nch; ++ch) sq += part[s * nch + ch];
  const float square = static_cast<float>(sq);
  const float nm = sqrtf(square > 0.0f ? square : 1.0f);
  const float threshold = 16.0f * kEps32 * fmaxf(scale[s], fabsf(alpha[s]));
  const bool k = active[s] && square > threshold * threshold;
  norm[s] = nm;
  keep[s] = k ? 1 : 0;
  if (beta) beta[s] = k ? nm : 0.0f;
  if (alpha_out) alpha_out[s] = active[s] ? alpha[s] : 1.0f;
}

template <bool FULL>
__device__ inline void lz_next_body(int64_t n, const float* __restrict__ w3, int64_t ldw,
                                    const float* __restrict__ norm, const uint8_t* __restrict__ keep,
                                    float* __restrict__ next, int64_t ldn) {
  const int s = blockIdx.y;
  const int64_t i0 = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  const bool k = keep[s] != 0;
  const float nm = norm[s];
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const int64_t i = i0 + 256 * e;
    if (FULL || i < n) next[(int64_t)s * ldn + i] = k ? w3[(int64_t)s * ldw + i] / nm : 0.0f;
  }
}

__global__ __launch_bounds__(256) void k_lz_next(int64_t n, const float* __restrict__ w3,
                                                 int64_t ldw, const float* __restrict__ norm,
                                                 const uint8_t* __restrict__ keep,
                                                 float* __restrict__ next, int64_t ldn) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= n)
    lz_next_body<true>(n, w3, ldw, norm, keep, next, ldn);
  else
    lz_next_body<false>(n, w3, ldw, norm, keep, next, ldn);
}

// ---- Adjoint of the step (slq.py:49-74 reversed) ----------------------
// Head:  nxt = w3 / norm, proj = nxt . q_bar_next,
//        w3_bar = keep ? (q_bar_next - nxt * proj) / norm + beta_bar * nxt : 0.
// Tail:  a_tot = (active ? alpha_bar : 0) - q . w1_bar,
//        w0_bar = w1_bar + q * a_tot,
//        basis_bar[index]     += -alpha * w1_bar + a_tot * w0,
//        basis_bar[index - 1] += -prev_beta * w0_bar,
//        prev_beta_bar = -(prev . w0_bar).
// Every block sums the chunk partials of its probe itself, in one fixed
// order, so all blocks of a probe see the same total without a separate
// launch.

// Fixed-order sum of part[0..nch) by the whole block, returned to every thread.
__device__ inline double block_total(const double* __restrict__ part, int nch) {
  __shared__ double red[4];
  __shared__ double total;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  double v = 0.0;
  for (int c = threadIdx.x; c < nch; c += 256) v += part[c];
  v = wave_sum(v);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) total = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return total;
}

template <bool FULL>
__device__ inline void lz_adj_head_part_body(int64_t n, const float* __restrict__ w3, int64_t ldw,
                                             const float* __restrict__ norm,
                                             const float* __restrict__ qbn, int64_t ldq,
                                             double* __restrict__ part, int nch) {
#pragma clang fp contract(off)
  const int s = blockIdx.y, ch = blockIdx.x;
  const int64_t i0 = (int64_t)ch * kChunk + threadIdx.x;
  const float nm = norm[s];
  double d = 0.0;
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const int64_t i = i0 + 256 * e;
    if (FULL || i < n) {
      const float nx = w3[(int64_t)s * ldw + i] / nm;
      d += static_cast<double>(nx) * static_cast<double>(qbn[(int64_t)s * ldq + i]);
    }
  }
  double sums[3];
  block_sums3(d, 0.0, 0.0, sums);
  if (threadIdx.x == 0) part[(int64_t)s * nch + ch] = sums[0];
}

__global__ __launch_bounds__(256) void k_lz_adj_head_part(int64_t n, const float* __restrict__ w3,
                                                          int64_t ldw, const float* __restrict__ norm,
                                                          const float* __restrict__ qbn, int64_t ldq,
                                                          double* __restrict__ part, int nch) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= n)
    lz_adj_head_part_body<true>(n, w3, ldw, norm, qbn, ldq, part, nch);
  else
    lz_adj_head_part_body<false>(n, w3, ldw, norm, qbn, ldq, part, nch);
}

template <bool FULL>
__device__ inline void lz_adj_head_w_body(int64_t n, const float* __restrict__ w3, int64_t ldw,
                                          const float* __restrict__ norm,
                                          const uint8_t* __restrict__ keep,
                                          const float* __restrict__ qbn, int64_t ldq,
                                          const float* __restrict__ beta_bar,
                                          const double* __restrict__ part, int nch,
                                          float* __restrict__ w3b, int64_t ldb) {
#pragma clang fp contract(off)
  const int s = blockIdx.y;
  const int64_t i0 = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  const float proj = static_cast<float>(block_total(part + (int64_t)s * nch, nch));
  const float nm = norm[s], bb = beta_bar[s];
  const bool k = keep[s] != 0;
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const int64_t i = i0 + 256 * e;
    if (FULL || i < n) {
      float v = 0.0f;
      if (k) {
        const float nx = w3[(int64_t)s * ldw + i] / nm;
        const float t = (qbn[(int64_t)s * ldq + i] - nx * proj) / nm;
        v = t + bb * nx;
      }
      w3b[(int64_t)s * ldb + i] = v;
    }
  }
}

__global__ __launch_bounds__(256) void k_lz_adj_head_w(int64_t n, const float* __restrict__ w3,
                                                       int64_t ldw, const float* __restrict__ norm,
                                                       const uint8_t* __restrict__ keep,
                                                       const float* __restrict__ qbn, int64_t ldq,
                                                       const float* __restrict__ beta_bar,
                                                       const double* __restrict__ part, int nch,
                                                       float* __restrict__ w3b, int64_t ldb) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= n)
    lz_adj_head_w_body<true>(n, w3, ldw, norm, keep, qbn, ldq, beta_bar, part, nch, w3b, ldb);
  else
    lz_adj_head_w_body<false>(n, w3, ldw, norm, keep, qbn, ldq, beta_bar, part, nch, w3b, ldb);
}

template <bool FULL>
__device__ inline void lz_adj_tail_part_body(int64_t n, const float* __restrict__ q, int64_t ps,
                                             const float* __restrict__ w1b, int64_t ld1,
                                             double* __restrict__ part, int nch) {
  const int s = blockIdx.y, ch = blockIdx.x;
  const int64_t i0 = (int64_t)ch * kChunk + threadIdx.x;
  double d = 0.0;
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const int64_t i = i0 + 256 * e;
    if (FULL || i < n)
      d += static_cast<double>(q[(int64_t)s * ps + i]) *
           static_cast<double>(w1b[(int64_t)s * ld1 + i]);
  }
  double sums[3];
  block_sums3(d, 0.0, 0.0, sums);
  if (threadIdx.x == 0) part[(int64_t)s * nch + ch] = sums[0];
}

__global__ __launch_bounds__(256) void k_lz_adj_tail_part(int64_t n, const float* __restrict__ q,
                                                          int64_t ps, const float* __restrict__ w1b,
                                                          int64_t ld1, double* __restrict__ part,
                                                          int nch) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= n)
    lz_adj_tail_part_body<true>(n, q, ps, w1b, ld1, part, nch);
  else
    lz_adj_tail_part_body<false>(n, q, ps, w1b, ld1, part, nch);
}

struct AdjTailArgs {
  int64_t n, index, ps, vs, ld0, ld1, bps, bvs, ldo;
  const float* basis;
  const float* w0;
  const float* w1b;
  const float* alpha;
  const uint8_t* active;
  const float* alpha_bar;
  const float* prev_beta;
  float* basis_bar;
  float* w0b;
  const double* part_in;
  double* part_out;
  int nch;
};

template <bool FULL>
__device__ inline void lz_adj_tail_w_body(const AdjTailArgs& a) {
#pragma clang fp contract(off)
  const int s = blockIdx.y, ch = blockIdx.x;
  const int64_t i0 = (int64_t)ch * kChunk + threadIdx.x;
  const float dot = static_cast<float>(block_total(a.part_in + (int64_t)s * a.nch, a.nch));
  const float a_tot = (a.active[s] ? a.alpha_bar[s] : 0.0f) - dot;
  const float neg_alpha = -a.alpha[s];
  const float neg_pb = -a.prev_beta[s];
  const float* q = a.basis + (int64_t)s * a.ps + a.index * a.vs;
  const float* pv = a.index > 0 ? q - a.vs : nullptr;
  float* bq = a.basis_bar + (int64_t)s * a.bps + a.index * a.bvs;
  float* bp = a.index > 0 ? bq - a.bvs : nullptr;
  double d = 0.0;
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const int64_t i = i0 + 256 * e;
    if (FULL || i < a.n) {
      const float w1b = a.w1b[(int64_t)s * a.ld1 + i];
      const float w0b = w1b + q[i] * a_tot;
      a.w0b[(int64_t)s * a.ldo + i] = w0b;
      const float t1 = neg_alpha * w1b;
      const float t2 = a_tot * a.w0[(int64_t)s * a.ld0 + i];
      bq[i] = bq[i] + (t1 + t2);
      if (pv) {
        bp[i] = bp[i] + neg_pb * w0b;
        d += static_cast<double>(pv[i]) * static_cast<double>(w0b);
      }
    }
  }
  double sums[3];
  block_sums3(d, 0.0, 0.0, sums);
  if (threadIdx.x == 0) a.part_out[(int64_t)s * a.nch + ch] = sums[0];
}

__global__ __launch_bounds__(256) void k_lz_adj_tail_w(AdjTailArgs a) {
  if ((int64_t)(blockIdx.x + 1) * kChunk <= a.n)
    lz_adj_tail_w_body<true>(a);
  else
    lz_adj_tail_w_body<false>(a);
}

__global__ void k_lz_adj_tail_final(int64_t S, int nch, const double* __restrict__ part,
                                    float* __restrict__ prev_beta_bar) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= S) return;
  double d = 0.0;
  for (int ch = 0; ch < nch; ++ch) d += part[s * nch + ch];
  prev_beta_bar[s] = -static_cast<float>(d);
}

}  // namespace coh

using namespace coh;

extern "C" int coh_probe_dots_workspace_bytes(int64_t S, int64_t r, int64_t n, uint64_t* bytes) {
  COH_REQUIRE(bytes != nullptr, COH_ERR_VALUE, "bytes is NULL");
  int st = check_dims(S, r, n);
  if (st != COH_OK) return st;
  const int64_t nch = std::max<int64_t>(1, ceil_div(n, kChunk));
  *bytes = (uint64_t)round_up(std::max<int64_t>(1, S * r * nch) * 8, 256);
  return COH_OK;
}

extern "C" int coh_probe_dots(int64_t S, int64_t r, int64_t n, const float* P, int64_t ps,
                              int64_t vs, const float* X, int64_t ldx, float* C, int64_t ldc,
                              void* ws, uint64_t ws_bytes, void* stream) {
  int st = check_dims(S, r, n);
  if (st != COH_OK) return st;
  if (S == 0 || r == 0) return COH_OK;
  COH_REQUIRE(P && X && C && ws, COH_ERR_VALUE, "NULL pointer argument");
  COH_REQUIRE(ldx >= n && ldc >= r && vs >= n && ps >= r * vs, COH_ERR_VALUE,
              "inconsistent strides");
  uint64_t need = 0;
  coh_probe_dots_workspace_bytes(S, r, n, &need);
  COH_REQUIRE(ws_bytes >= need, COH_ERR_VALUE, "workspace too small");
  const int nch = (int)std::max<int64_t>(1, ceil_div(n, kChunk));
  hipStream_t hs = static_cast<hipStream_t>(stream);
  double* part = static_cast<double*>(ws);
  hipLaunchKernelGGL(k_probe_dots_part, dim3(nch, (unsigned)S), dim3(256), 0, hs, n, (int)r, P, ps,
                     vs, X, ldx, part, nch);
  COH_CHECK_LAUNCH();
  const int64_t total = S * r;
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, 256), 4096));
  hipLaunchKernelGGL(k_probe_dots_final, dim3(blocks), dim3(256), 0, hs, S, (int)r, nch, part, C,
                     ldc);
  COH_CHECK_LAUNCH();
  return COH_OK;
}

extern "C" int coh_probe_project_out(int64_t S, int64_t r, int64_t n, const float* P, int64_t ps,
                                     int64_t vs, const float* C, int64_t ldc, const float* X,
                                     int64_t ldx, float* Y, int64_t ldy, void* stream) {
  int st = check_dims(S, r, n);
  if (st != COH_OK) return st;
  if (S == 0 || n == 0) return COH_OK;
  COH_REQUIRE(X && Y && (r == 0 || (P && C)), COH_ERR_VALUE, "NULL pointer argument");
  COH_REQUIRE(ldx >= n && ldy >= n && (r == 0 || (ldc >= r && vs >= n && ps >= r * vs)),
              COH_ERR_VALUE, "inconsistent strides");
  const int nch = (int)std::max<int64_t>(1, ceil_div(n, kChunk));
  hipLaunchKernelGGL(k_probe_project_out, dim3(nch, (unsigned)S), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, (int)r, P, ps, vs, C, ldc, X, ldx, Y,
                     ldy);
  COH_CHECK_LAUNCH();
  return COH_OK;
}

extern "C" int coh_probe_rank_update(int64_t S, int64_t r, int64_t n, int32_t q,
                                     const float* coeff, int64_t cps, int64_t cvs, const float* V,
                                     int64_t vps, int64_t vvs, float* B, int64_t bps, int64_t bvs,
                                     void* stream) {
  int st = check_dims(S, r, n);
  if (st != COH_OK) return st;
  COH_REQUIRE(q >= 1 && q <= 4, COH_ERR_VALUE, "rank update takes 1 to 4 vectors");
  if (S == 0 || r == 0 || n == 0) return COH_OK;
  COH_REQUIRE(coeff && V && B, COH_ERR_VALUE, "NULL pointer argument");
  COH_REQUIRE(cvs >= q && cps >= r * cvs && vvs >= n && vps >= q * vvs && bvs >= n &&
                  bps >= r * bvs,
              COH_ERR_VALUE, "inconsistent strides");
  const int nch = (int)std::max<int64_t>(1, ceil_div(n, kChunk));
  hipLaunchKernelGGL(k_probe_rank_update, dim3(nch, (unsigned)S), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, (int)r, (int)q, coeff, cps, cvs, V, vps,
                     vvs, B, bps, bvs);
  COH_CHECK_LAUNCH();
  return COH_OK;
}

extern "C" int coh_probe_step_workspace_bytes(int64_t S, int64_t n, uint64_t* bytes) {
  COH_REQUIRE(bytes != nullptr, COH_ERR_VALUE, "bytes is NULL");
  int st = check_dims(S, 0, n);
  if (st != COH_OK) return st;
  const int64_t nch = std::max<int64_t>(1, ceil_div(n, kChunk));
  *bytes = (uint64_t)round_up(std::max<int64_t>(1, S * nch * 3) * 8, 256);
  return COH_OK;
}

extern "C" int coh_probe_step_prep(int64_t S, int64_t n, int64_t index, const float* basis,
                                int64_t ps, int64_t vs, const float* product, int64_t ldp,
                                const float* prev_beta, float* w0, float* w1, int64_t ldw,
                                float* alpha, uint8_t* active, float* scale, void* ws,
                                uint64_t ws_bytes, void* stream) {
  int st = check_dims(S, 0, n);
  if (st != COH_OK) return st;
  if (S == 0) return COH_OK;
  COH_REQUIRE(basis && product && prev_beta && w0 && w1 && alpha && active && scale && ws,
              COH_ERR_VALUE, "NULL pointer argument");
  COH_REQUIRE(index >= 0 && vs >= n && ps >= (index + 1) * vs && ldp >= n && ldw >= n,
              COH_ERR_VALUE, "inconsistent strides");
  uint64_t need = 0;
  coh_probe_step_workspace_bytes(S, n, &need);
  COH_REQUIRE(ws_bytes >= need, COH_ERR_VALUE, "workspace too small");
  const int nch = (int)std::max<int64_t>(1, ceil_div(n, kChunk));
  hipStream_t hs = static_cast<hipStream_t>(stream);
  double* part = static_cast<double*>(ws);
  hipLaunchKernelGGL(k_lz_prep_part, dim3(nch, (unsigned)S), dim3(256), 0, hs, n, index, basis,
                     ps, vs, product, ldp, prev_beta, part, nch);
  COH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_lz_prep_final, dim3((unsigned)ceil_div(S, 64)), dim3(64), 0, hs, S, nch,
                     part, alpha, active, scale);
  COH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_lz_prep_w, dim3(nch, (unsigned)S), dim3(256), 0, hs, n, index, basis, ps,
                     vs, product, ldp, prev_beta, alpha, w0, w1, ldw);
  COH_CHECK_LAUNCH();
  return COH_OK;
}

extern "C" int coh_probe_step_finish(int64_t S, int64_t n, const float* w3, int64_t ldw,
                                  const float* alpha, const uint8_t* active, const float* scale,
                                  float* norm, uint8_t* keep, float* beta, float* alpha_out,
                                  float* next, int64_t ldn, void* ws, uint64_t ws_bytes,
                                  void* stream) {
  int st = check_dims(S, 0, n);
  if (st != COH_OK) return st;
  if (S == 0) return COH_OK;
  COH_REQUIRE(w3 && alpha && active && scale && norm && keep && ws, COH_ERR_VALUE,
              "NULL pointer argument");
  COH_REQUIRE(ldw >= n && (next == nullptr || ldn >= n), COH_ERR_VALUE, "inconsistent strides");
  uint64_t need = 0;
  coh_probe_step_workspace_bytes(S, n, &need);
  COH_REQUIRE(ws_bytes >= need, COH_ERR_VALUE, "workspace too small");
  const int nch = (int)std::max<int64_t>(1, ceil_div(n, kChunk));
  hipStream_t hs = static_cast<hipStream_t>(stream);
  double* part = static_cast<double*>(ws);
  hipLaunchKernelGGL(k_lz_sq_part, dim3(nch, (unsigned)S), dim3(256), 0, hs, n, w3, ldw, part,
                     nch);
  COH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_lz_finish_final, dim3((unsigned)ceil_div(S, 64)), dim3(64), 0, hs, S, nch,
                     part, alpha, active, scale, norm, keep, beta, alpha_out);
  COH_CHECK_LAUNCH();
  if (next) {
    hipLaunchKernelGGL(k_lz_next, dim3(nch, (unsigned)S), dim3(256), 0, hs, n, w3, ldw, norm,
                       keep, next, ldn);
    COH_CHECK_LAUNCH();
  }
  return COH_OK;
}

extern "C" int coh_probe_step_adjoint_head(int64_t S, int64_t n, const float* w3, int64_t ldw,
                                          const float* norm, const uint8_t* keep,
                                          const float* q_bar_next, int64_t ldq,
                                          const float* beta_bar, float* w3_bar, int64_t ldb,
                                          void* ws, uint64_t ws_bytes, void* stream) {
  int st = check_dims(S, 0, n);
  if (st != COH_OK) return st;
  if (S == 0) return COH_OK;
  COH_REQUIRE(w3 && norm && keep && q_bar_next && beta_bar && w3_bar && ws, COH_ERR_VALUE,
              "NULL pointer argument");
  COH_REQUIRE(ldw >= n && ldq >= n && ldb >= n, COH_ERR_VALUE, "inconsistent strides");
  uint64_t need = 0;
  coh_probe_step_workspace_bytes(S, n, &need);
  COH_REQUIRE(ws_bytes >= need, COH_ERR_VALUE, "workspace too small");
  const int nch = (int)std::max<int64_t>(1, ceil_div(n, kChunk));
  hipStream_t hs = static_cast<hipStream_t>(stream);
  double* part = static_cast<double*>(ws);
  hipLaunchKernelGGL(k_lz_adj_head_part, dim3(nch, (unsigned)S), dim3(256), 0, hs, n, w3, ldw,
                     norm, q_bar_next, ldq, part, nch);
  COH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_lz_adj_head_w, dim3(nch, (unsigned)S), dim3(256), 0, hs, n, w3, ldw, norm,
                     keep, q_bar_next, ldq, beta_bar, part, nch, w3_bar, ldb);
  COH_CHECK_LAUNCH();
  return COH_OK;
}

extern "C" int coh_probe_step_adjoint_tail(int64_t S, int64_t n, int64_t index, const float* basis,
                                          int64_t ps, int64_t vs, const float* w0, int64_t ld0,
                                          const float* w1_bar, int64_t ld1, const float* alpha,
                                          const uint8_t* active, const float* alpha_bar,
                                          const float* prev_beta, float* basis_bar, int64_t bps,
                                          int64_t bvs, float* w0_bar, int64_t ldo,
                                          float* prev_beta_bar, void* ws, uint64_t ws_bytes,
                                          void* stream) {
  int st = check_dims(S, 0, n);
  if (st != COH_OK) return st;
  if (S == 0) return COH_OK;
  COH_REQUIRE(basis && w0 && w1_bar && alpha && active && alpha_bar && prev_beta && basis_bar &&
                  w0_bar && prev_beta_bar && ws,
              COH_ERR_VALUE, "NULL pointer argument");
  COH_REQUIRE(index >= 0 && vs >= n && ps >= (index + 1) * vs && bvs >= n &&
                  bps >= (index + 1) * bvs && ld0 >= n && ld1 >= n && ldo >= n,
              COH_ERR_VALUE, "inconsistent strides");
  uint64_t need = 0;
  coh_probe_step_workspace_bytes(S, n, &need);
  COH_REQUIRE(ws_bytes >= need, COH_ERR_VALUE, "workspace too small");
  const int nch = (int)std::max<int64_t>(1, ceil_div(n, kChunk));
  hipStream_t hs = static_cast<hipStream_t>(stream);
  double* part = static_cast<double*>(ws);
  double* part2 = part + (int64_t)S * nch;
  hipLaunchKernelGGL(k_lz_adj_tail_part, dim3(nch, (unsigned)S), dim3(256), 0, hs, n,
                     basis + index * vs, ps, w1_bar, ld1, part, nch);
  COH_CHECK_LAUNCH();
  AdjTailArgs a{n, index, ps, vs, ld0, ld1, bps, bvs, ldo, basis, w0, w1_bar, alpha, active,
                alpha_bar, prev_beta, basis_bar, w0_bar, part, part2, nch};
  hipLaunchKernelGGL(k_lz_adj_tail_w, dim3(nch, (unsigned)S), dim3(256), 0, hs, a);
  COH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_lz_adj_tail_final, dim3((unsigned)ceil_div(S, 64)), dim3(64), 0, hs, S,
                     nch, part2, prev_beta_bar);
  COH_CHECK_LAUNCH();
  return COH_OK;
}
