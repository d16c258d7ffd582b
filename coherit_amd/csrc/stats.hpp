// Per-SNP float32 standardization statistics from integer counts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace coh {

// float32 mean / inv_sd with the reference's operation order
// (geno_stream.py:1780-1792); contraction disabled so that mean*mean and the
// subtraction round separately, as in NumPy.
__device__ inline void f32_stats(int64_t c, int64_t a1, int64_t a2, float* mean_out,
                                 float* inv_out) {
#pragma clang fp contract(off)
  const float cnt_f = static_cast<float>(c);
  const float s1_f = static_cast<float>(a1);
  const float s2_f = static_cast<float>(a2);
  const float denom = fmaxf(cnt_f, 1.0f);
  const float mean = s1_f / denom;
  const float q = s2_f / denom;
  const float sq = mean * mean;
  const float var = fmaxf(q - sq, 0.0f);
  const bool valid = (cnt_f > 0.0f) && (var > 0.0f);
  float inv = 0.0f;
  if (valid) {
    // Correctly rounded FP32 sqrt and reciprocal: evaluate in FP64 and round
    // each result to FP32 (double rounding from 53 to 24 bits is innocuous
    // for sqrt and division since 53 >= 2*24 + 2), matching NumPy float32.
    const float t = static_cast<float>(__builtin_sqrt(static_cast<double>(fmaxf(var, 1e-6f))));
    inv = static_cast<float>(1.0 / static_cast<double>(t));
  }
  *mean_out = mean;
  *inv_out = inv;
}

}  // namespace coh
