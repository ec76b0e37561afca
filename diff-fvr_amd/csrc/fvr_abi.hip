// ABI plumbing: version, thread-local error text, layout descriptors and the
// spectrum pack kernels (centred complex spectrum -> padded device layout).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "fvr_common.h"

namespace fvr {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------------------
// Pack: one thread per storage element; logical index = storage index - 1.
// transpose_xy swaps the mid and fast layout axes relative to (z, y, x).
// ---------------------------------------------------------------------------
struct PackArgs {
  fvr_layout lay;
  const float2* src;
  int nz, ny, nx;
  int transpose;
};

__device__ __forceinline__ float2 pack_src(const PackArgs& a, int s, int m, int f) {
  int z = s - 1, y = m - 1, x = f - 1;  // logical (slow, mid, fast)
  if (a.lay.flags & FVR_LAYOUT_WRAP) {  // logical index n is the periodic copy of index 0
    if (z == a.lay.n[0]) z = 0;
    if (y == a.lay.n[1]) y = 0;
    if (x == a.lay.n[2]) x = 0;
  }
  if (a.transpose) { int t = y; y = x; x = t; }
  if (z < 0 || z >= a.nz || y < 0 || y >= a.ny || x < 0 || x >= a.nx) return make_float2(0.f, 0.f);
  return a.src[(size_t(z) * a.ny + y) * a.nx + x];
}

// The spectrum's power-of-two scale sc (scaled dtypes): max|component| / sc in
// [2^14, 2^15), the maximum taken over the source (every stored element is a source
// element or 0).  Non-negative floats order as their bit patterns, so the workgroups
// combine their maxima with an integer atomicMax on the scale slot, which
// scale_clear_kernel zeroes first; scale_final_kernel turns the maximum into sc.
__global__ void scale_clear_kernel(uint32_t* slot) { *slot = 0u; }

__global__ __launch_bounds__(256) void absmax_kernel(const float2* __restrict__ src, size_t n,
                                                     uint32_t* __restrict__ slot) {
  float mx = 0.f;
  for (size_t k = size_t(blockIdx.x) * 256 + threadIdx.x; k < n; k += size_t(gridDim.x) * 256) {
    const float2 v = src[k];
    mx = fmaxf(mx, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(slot, __float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))));
}

__global__ void scale_final_kernel(float* slot) {
  const float m = __uint_as_float(*reinterpret_cast<const uint32_t*>(slot));
  *slot = (m > 0.f && isfinite(m)) ? ldexpf(1.f, ilogbf(m) - 14) : 1.f;
}

__global__ __launch_bounds__(256) void pack_kernel(PackArgs a) {
  const int f = blockIdx.x * 256 + threadIdx.x, m = blockIdx.y, s = blockIdx.z;  // storage index
  if (f >= a.lay.sdim[2]) return;
  const float2 v = pack_src(a, s, m, f);
  char* row = reinterpret_cast<char*>(a.lay.base) + size_t(s) * a.lay.plane_bytes + size_t(m) * a.lay.row_bytes;
  if (a.lay.dtype == FVR_C32) {
    reinterpret_cast<float2*>(row)[f] = v;
  } else {
    const float inv = 1.f / *reinterpret_cast<const float*>(a.lay.scales);  // a power of two: exact
    uint16_t re, im;
    if (a.lay.dtype == FVR_C16F) { re = f2h(v.x * inv); im = f2h(v.y * inv); }
    else { re = f2b(v.x * inv); im = f2b(v.y * inv); }
    reinterpret_cast<uint32_t*>(row)[f] = uint32_t(re) | (uint32_t(im) << 16);
  }
}

// Moments of two complex64 (h x w) slices a, b with uu = j - w/2, vv = i - h/2:
//   out = [Re sum a b*, sum |a|^2, sum |b|^2, sum uu Im(a b*), sum vv Im(a b*)]
// (float64, fixed order; one workgroup).  Serves compute_spectral_ncc
// (diff_fvr.py:184-195) and the translation closed form (gradients.py:27-74).
__global__ __launch_bounds__(1024) void c64_pair_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                                                        int h, int w, double* __restrict__ out) {
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t n = int64_t(h) * w;
  for (int64_t k = threadIdx.x; k < n; k += 1024) {
    const int i = int(k / w), j = int(k - int64_t(i) * w);
    const double ar = a[k].x, ai = a[k].y, br = b[k].x, bi = b[k].y;
    const double re = ar * br + ai * bi, im = ai * br - ar * bi;
    acc[0] += re;
    acc[1] += ar * ar + ai * ai;
    acc[2] += br * br + bi * bi;
    acc[3] += (double(j) - 0.5 * w) * im;
    acc[4] += (double(i) - 0.5 * h) * im;
  }
  __shared__ double red[1024];
  for (int q = 0; q < 5; ++q) {
    red[threadIdx.x] = acc[q];
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
      if (int(threadIdx.x) < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
    __syncthreads();
  }
}

static_assert(sizeof(fvr_layout) == 80, "fvr_layout ABI size");
static_assert(sizeof(fvr_sncc_item) == 200, "fvr_sncc_item ABI size");
static_assert(sizeof(fvr_adam_cfg) == 232, "fvr_adam_cfg ABI size");

}  // namespace fvr

using namespace fvr;

extern "C" {

int fvr_abi_version(void) { return FVR_ABI_VERSION; }

const char* fvr_last_error(void) { return g_err; }

int fvr_struct_sizes(int32_t* out) {
  if (!out) return FVR_ERR_INVALID;
  out[0] = int32_t(sizeof(fvr_layout));
  out[1] = int32_t(sizeof(fvr_sncc_item));
  out[2] = int32_t(sizeof(fvr_adam_cfg));
  out[3] = 0;
  return FVR_OK;
}

int fvr_c64_pair_moments(uint64_t a, uint64_t b, int32_t h, int32_t w, double* out, void* stream) {
  FVR_CHECK_ARG(a && b && out && h > 0 && w > 0, "fvr_c64_pair_moments: bad arguments");
  hipLaunchKernelGGL(c64_pair_kernel, dim3(1), dim3(1024), 0, reinterpret_cast<hipStream_t>(stream),
                     reinterpret_cast<const float2*>(a), reinterpret_cast<const float2*>(b), h, w, out);
  FVR_CHECK_LAUNCH("c64_pair_kernel");
  return FVR_OK;
}

int fvr_layout_init(fvr_layout* lay, const int32_t n[3], int32_t dtype, uint64_t* storage_bytes,
                    uint64_t* n_scales) {
  FVR_CHECK_ARG(lay && n, "fvr_layout_init: null argument");
  const uint32_t lflags = uint32_t(dtype) >> 8;
  dtype &= 0xFF;
  FVR_CHECK_ARG(dtype == FVR_C32 || dtype == FVR_C16F || dtype == FVR_C16B, "fvr_layout_init: bad dtype %d",
                dtype);
  FVR_CHECK_ARG((lflags & ~FVR_LAYOUT_WRAP) == 0, "fvr_layout_init: bad layout flags %u", lflags);
  for (int a = 0; a < 3; ++a) FVR_CHECK_ARG(n[a] >= 2 && n[a] < (1 << 14), "fvr_layout_init: bad dim %d", n[a]);
  fvr_layout L = {};
  for (int a = 0; a < 3; ++a) L.n[a] = n[a];
  L.dtype = dtype;
  L.flags = lflags;
  const int extra = (lflags & FVR_LAYOUT_WRAP) ? 1 : 0;  // the wrap index n, then the usual zero margins
  L.sdim[0] = n[0] + 3 + extra;
  L.sdim[1] = n[1] + 2 + extra;
  L.sdim[2] = n[2] + 2 + extra;
  L.elem_bytes = dtype == FVR_C32 ? 8u : 4u;
  const uint64_t row = uint64_t(L.sdim[2]) * L.elem_bytes;
  const uint64_t plane = row * uint64_t(L.sdim[1]);
  const uint64_t total = plane * uint64_t(L.sdim[0]);
  // kernels form 32-bit byte offsets with 24-bit multiplies
  FVR_CHECK_ARG(plane < (1ull << 24) && total < (1ull << 32),
                "fvr_layout_init: spectrum %dx%dx%d too large for 32-bit offsets", n[0], n[1], n[2]);
  L.row_bytes = uint32_t(row);
  L.plane_bytes = uint32_t(plane);
  L.total_bytes = total;
  *lay = L;
  if (storage_bytes) *storage_bytes = total;
  if (n_scales) *n_scales = dtype == FVR_C32 ? 0 : 1;
  return FVR_OK;
}

int fvr_spec_pack(const fvr_layout* lay, uint64_t src, int32_t nz, int32_t ny, int32_t nx, int32_t transpose_xy,
                  void* stream) {
  FVR_CHECK_ARG(lay && lay->base && src, "fvr_spec_pack: null pointer");
  const int ln1 = transpose_xy ? nx : ny, ln2 = transpose_xy ? ny : nx;
  FVR_CHECK_ARG(lay->n[0] == nz && lay->n[1] == ln1 && lay->n[2] == ln2,
                "fvr_spec_pack: layout dims (%d,%d,%d) do not match source (%d,%d,%d) transpose=%d", lay->n[0],
                lay->n[1], lay->n[2], nz, ny, nx, transpose_xy);
  FVR_CHECK_ARG(lay->dtype == FVR_C32 || lay->scales, "fvr_spec_pack: scaled dtype needs a scale buffer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  PackArgs a;
  a.lay = *lay;
  a.src = reinterpret_cast<const float2*>(src);
  a.nz = nz;
  a.ny = ny;
  a.nx = nx;
  a.transpose = transpose_xy;
  FVR_CHECK_ARG(lay->flags == (lay->flags & FVR_LAYOUT_WRAP) && lay->row_bytes == uint32_t(lay->sdim[2]) * lay->elem_bytes,
                "fvr_spec_pack: layout descriptor not from fvr_layout_init");
  if (lay->dtype != FVR_C32) {
    uint32_t* slot = reinterpret_cast<uint32_t*>(lay->scales);
    const size_t n = size_t(nz) * ny * nx;
    const int blocks = int(std::min<size_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(scale_clear_kernel, dim3(1), dim3(1), 0, st, slot);
    FVR_CHECK_LAUNCH("scale_clear_kernel");
    hipLaunchKernelGGL(absmax_kernel, dim3(blocks), dim3(256), 0, st, a.src, n, slot);
    FVR_CHECK_LAUNCH("absmax_kernel");
    hipLaunchKernelGGL(scale_final_kernel, dim3(1), dim3(1), 0, st, reinterpret_cast<float*>(slot));
    FVR_CHECK_LAUNCH("scale_final_kernel");
  }
  dim3 g((lay->sdim[2] + 255) / 256, lay->sdim[1], lay->sdim[0]);
  hipLaunchKernelGGL(pack_kernel, g, dim3(256), 0, st, a);
  FVR_CHECK_LAUNCH("pack_kernel");
  return FVR_OK;
}

}  // extern "C"
