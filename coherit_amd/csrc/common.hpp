// Shared definitions for the coherit_amd HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/coherit_amd.h"

namespace coh {

// Thread-local message behind coh_last_error().
void set_error(const char* fmt, ...);
int fail(int status, const char* fmt, ...);

#define COH_HIP(call)                                                              \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess)                                                          \
      return ::coh::fail(COH_ERR_HIP, "%s failed: %s (%s:%d)", #call,              \
                         hipGetErrorString(_e), __FILE__, __LINE__);               \
  } while (0)

#define COH_CHECK_LAUNCH()                                                         \
  do {                                                                             \
    hipError_t _e = hipGetLastError();                                             \
    if (_e != hipSuccess)                                                          \
      return ::coh::fail(COH_ERR_HIP, "kernel launch failed: %s (%s:%d)",          \
                         hipGetErrorString(_e), __FILE__, __LINE__);               \
  } while (0)

#define COH_REQUIRE(cond, status, ...)                                             \
  do {                                                                             \
    if (!(cond)) return ::coh::fail(status, __VA_ARGS__);                          \
  } while (0)

// Slot space: components are laid out one after the other, each component's
// calls concatenated with every call padded to a multiple of 4 SNPs (exactly
// the per-call packed width of the reference cache) and every component
// padded to a multiple of kSlotAlign slots, so that no kernel tile straddles
// two components.  Padding slots hold code 0 with mean = inv = 0.
constexpr int64_t kSlotAlign = 256;
// Variant-major rows (4 samples per byte) are padded to a multiple of
// kSampleAlign samples; padding samples hold code 0.
constexpr int64_t kSampleAlign = 256;

inline int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }
inline int64_t ceil_div(int64_t x, int64_t a) { return (x + a - 1) / a; }

}  // namespace coh

struct coh_panel {
  int device = 0;
  int64_t n = 0;          // samples
  int64_t m = 0;          // SNPs in cache order
  int32_t n_calls = 0;
  int32_t G = 0;          // components
  std::vector<int64_t> call_start;   // cache SNP start of each call
  std::vector<int32_t> call_width;   // true width
  std::vector<int32_t> call_comp;    // component of each call
  std::vector<int64_t> call_slot;    // first slot of each call
  std::vector<int64_t> comp_slot_begin, comp_slot_end;  // [G]
  int64_t m_slot = 0;     // total slots (multiple of kSlotAlign)
  int64_t sbytes = 0;     // sample-major row bytes = m_slot / 4
  int64_t vbytes = 0;     // variant-major row bytes
  uint8_t* smaj = nullptr;   // (n, sbytes)   4 SNP slots per byte
  uint8_t* vmaj = nullptr;   // (m_slot, vbytes) 4 samples per byte
  float* mean = nullptr;     // (m_slot)
  float* inv = nullptr;      // (m_slot)
  int64_t* cnt = nullptr;    // (m_slot) int64 stats
  int64_t* s1 = nullptr;
  int64_t* s2 = nullptr;
  int32_t* slot_to_cache = nullptr;  // (m_slot), -1 on padding
  int32_t* cache_to_slot = nullptr;  // (m)
  uint8_t* row_has_stats = nullptr;  // (m_slot) 1 when int8 ingest wrote raw stats
  bool smaj_written = false;
  bool vmaj_written = false;
  bool finalized = false;
  std::vector<int32_t> h_cache_to_slot;
};
