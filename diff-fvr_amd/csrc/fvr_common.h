// Shared host/device helpers of libfvr_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "../../include/fvr_hip.h"

namespace fvr {

// thread-local error text, set by the entry points before they return != FVR_OK
void set_error(const char* fmt, ...);

#define FVR_CHECK_ARG(cond, ...)               \
  do {                                         \
    if (!(cond)) {                             \
      ::fvr::set_error(__VA_ARGS__);           \
      return FVR_ERR_INVALID;                  \
    }                                          \
  } while (0)

#define FVR_CHECK_LAUNCH(what)                                                   \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) {                                                      \
      ::fvr::set_error("%s: %s", what, hipGetErrorString(e_));                   \
      return FVR_ERR_HIP;                                                        \
    }                                                                            \
  } while (0)

constexpr int kWave = 64;

// default rows of one fused-kernel tile (4 waves x 16 rows); tiles are 64 columns wide.
// A launch may use tile_h = 32 or 128 instead (item field tile_h).
constexpr int kSnccTileH = 64;

__host__ __device__ __forceinline__ bool sncc_tile_h_ok(int tile_h) {
  return tile_h == 32 || tile_h == 64 || tile_h == 128;
}

// Tile geometry: a tile is 64 columns x tile_h rows; a Hermitian-fold tile is 32 columns
// x 2 tile_h rows (the same 4096 samples at tile_h 64) and the fold covers h/2 rows.
// tile rows of one item; 0 for a bad tile_h
__host__ __device__ __forceinline__ int sncc_tiles_y(int h, uint32_t flags, int tile_h) {
  if (!sncc_tile_h_ok(tile_h)) return 0;
  const bool herm = flags & FVR_ITEM_HERM;
  const int rows = herm ? (h >> 1) : h, per = herm ? 2 * tile_h : tile_h;
  return (rows + per - 1) / per;
}

// fp16 / bf16 storage element -> float
__device__ __forceinline__ float h2f(uint16_t h) { return __half2float(__ushort_as_half(h)); }
__device__ __forceinline__ float b2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

__device__ __forceinline__ uint16_t f2h(float f) { return __half_as_ushort(__float2half_rn(f)); }
__device__ __forceinline__ uint16_t f2b(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return uint16_t(u >> 16);
}

}  // namespace fvr
