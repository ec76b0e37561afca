// Generic batched 3D Fourier slice on the planar (B_vol, 2, D, H, W) layout the
// reference's operator takes (triton_slice.py:648-661 fourier_slice_3d):
//   p(n, q) = center + fx[q] u[n] + fy[q] v[n]      (x = W, y = H, z = D axes)
// trilinear with per-corner zero padding (grid_sample align_corners=True,
// padding 'zeros'; triton_slice.py:664-715).  The backward returns d/du, d/dv
// only (triton_slice.py:339), reduced without atomics: one partial per
// (pose, q-block), then a fixed-order sum.
#include "fvr_common.h"

namespace fvr {

template <int VT>
__device__ __forceinline__ float ldv(const void* p, size_t i) {
  if (VT == 0) return reinterpret_cast<const float*>(p)[i];
  const uint16_t h = reinterpret_cast<const uint16_t*>(p)[i];
  return VT == 1 ? h2f(h) : b2f(h);
}

template <int VT>
__device__ __forceinline__ void stv(void* p, size_t i, float v) {
  if (VT == 0) reinterpret_cast<float*>(p)[i] = v;
  else reinterpret_cast<uint16_t*>(p)[i] = VT == 1 ? f2h(v) : f2b(v);
}

struct SliceGeom {
  int b_vol, d, h, w, n_pose, k;
  float cx, cy, cz;
};

// the 8 corners (values masked to 0 outside the volume) of channel plane `base`
template <int VT>
__device__ __forceinline__ void corners(const void* vol, size_t base, int x0, int y0, int z0, const SliceGeom& g,
                                        float c[8]) {
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const int x = x0 + dx, y = y0 + dy, z = z0 + dz;
        const bool in = x >= 0 && x < g.w && y >= 0 && y < g.h && z >= 0 && z < g.d;
        c[dz * 4 + dy * 2 + dx] = in ? ldv<VT>(vol, base + (size_t(z) * g.h + y) * g.w + x) : 0.f;
      }
}

template <int VT>
__global__ __launch_bounds__(256) void slice_fwd_kernel(const void* __restrict__ vol, const float* __restrict__ u,
                                                        const float* __restrict__ v, const float* __restrict__ fx,
                                                        const float* __restrict__ fy, SliceGeom g,
                                                        void* __restrict__ out) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int bn = blockIdx.y;
  const int b = bn / g.n_pose, n = bn - b * g.n_pose;
  if (q >= g.k) return;
  const float a = fx[q], c = fy[q];
  const float px = g.cx + a * u[3 * n + 0] + c * v[3 * n + 0];
  const float py = g.cy + a * u[3 * n + 1] + c * v[3 * n + 1];
  const float pz = g.cz + a * u[3 * n + 2] + c * v[3 * n + 2];
  const float x0f = floorf(px), y0f = floorf(py), z0f = floorf(pz);
  const float tx = px - x0f, ty = py - y0f, tz = pz - z0f;
  const int x0 = int(x0f), y0 = int(y0f), z0 = int(z0f);
  const size_t dhw = size_t(g.d) * g.h * g.w;
  float w8[8];
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx)
        w8[dz * 4 + dy * 2 + dx] = (dx ? tx : 1.f - tx) * (dy ? ty : 1.f - ty) * (dz ? tz : 1.f - tz);
  const size_t obase = size_t(bn) * 2 * g.k;
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
    float cc[8];
    corners<VT>(vol, (size_t(b) * 2 + ch) * dhw, x0, y0, z0, g, cc);
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < 8; ++i) acc = fmaf(cc[i], w8[i], acc);
    stv<VT>(out, obase + size_t(ch) * g.k + q, acc);
  }
}

// one workgroup per (q-block, pose): loops over volumes, 6 partial sums
template <int VT>
__global__ __launch_bounds__(256) void slice_bwd_kernel(const void* __restrict__ vol, const void* __restrict__ go,
                                                        const float* __restrict__ u, const float* __restrict__ v,
                                                        const float* __restrict__ fx, const float* __restrict__ fy,
                                                        SliceGeom g, float* __restrict__ part) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int n = blockIdx.y;
  float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (q < g.k) {
    const float a = fx[q], c = fy[q];
    const float px = g.cx + a * u[3 * n + 0] + c * v[3 * n + 0];
    const float py = g.cy + a * u[3 * n + 1] + c * v[3 * n + 1];
    const float pz = g.cz + a * u[3 * n + 2] + c * v[3 * n + 2];
    const float x0f = floorf(px), y0f = floorf(py), z0f = floorf(pz);
    const float tx = px - x0f, ty = py - y0f, tz = pz - z0f;
    const float ox = 1.f - tx, oy = 1.f - ty, oz = 1.f - tz;
    const int x0 = int(x0f), y0 = int(y0f), z0 = int(z0f);
    const size_t dhw = size_t(g.d) * g.h * g.w;
    float gx = 0.f, gy = 0.f, gz = 0.f;
    for (int b = 0; b < g.b_vol; ++b) {
      const size_t ob = (size_t(b) * g.n_pose + n) * 2 * g.k;
#pragma unroll
      for (int ch = 0; ch < 2; ++ch) {
        const float gg = ldv<VT>(go, ob + size_t(ch) * g.k + q);
        float r[8];
        corners<VT>(vol, (size_t(b) * 2 + ch) * dhw, x0, y0, z0, g, r);
        // r index = dz*4 + dy*2 + dx
        const float ddx = ((r[1] - r[0]) * oy + (r[3] - r[2]) * ty) * oz + ((r[5] - r[4]) * oy + (r[7] - r[6]) * ty) * tz;
        const float ddy = ((r[2] - r[0]) * ox + (r[3] - r[1]) * tx) * oz + ((r[6] - r[4]) * ox + (r[7] - r[5]) * tx) * tz;
        const float ddz = ((r[4] - r[0]) * ox + (r[5] - r[1]) * tx) * oy + ((r[6] - r[2]) * ox + (r[7] - r[3]) * tx) * ty;
        gx = fmaf(gg, ddx, gx);
        gy = fmaf(gg, ddy, gy);
        gz = fmaf(gg, ddz, gz);
      }
    }
    acc[0] = gx * a;
    acc[1] = gy * a;
    acc[2] = gz * a;
    acc[3] = gx * c;
    acc[4] = gy * c;
    acc[5] = gz * c;
  }
  __shared__ float red[6][256];
#pragma unroll
  for (int i = 0; i < 6; ++i) red[i][threadIdx.x] = acc[i];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (int(threadIdx.x) < st)
#pragma unroll
      for (int i = 0; i < 6; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 6) part[(size_t(n) * gridDim.x + blockIdx.x) * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void slice_bwd_sum_kernel(const float* __restrict__ part, int nq, int n_pose, float* __restrict__ gu,
                                     float* __restrict__ gv) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_pose * 6) return;
  const int n = idx / 6, c = idx - n * 6;
  double s = 0.0;
  for (int b = 0; b < nq; ++b) s += double(part[(size_t(n) * nq + b) * 6 + c]);
  if (c < 3) gu[3 * n + c] = float(s);
  else gv[3 * n + c - 3] = float(s);
}

}  // namespace fvr

using namespace fvr;

extern "C" {

int fvr_slice_fwd(uint64_t vol, int32_t vol_dtype, int32_t b_vol, int32_t d, int32_t h, int32_t w,
                  const float* u_xyz, const float* v_xyz, int32_t n_pose, const float* fx, const float* fy,
                  int32_t k, float cx, float cy, float cz, uint64_t out, void* stream) {
  FVR_CHECK_ARG(vol && u_xyz && v_xyz && fx && fy && out, "fvr_slice_fwd: null pointer");
  FVR_CHECK_ARG(b_vol > 0 && d > 0 && h > 0 && w > 0 && n_pose > 0 && k > 0, "fvr_slice_fwd: empty shape");
  FVR_CHECK_ARG(int64_t(b_vol) * n_pose <= 65535, "fvr_slice_fwd: %lld (volume, pose) pairs exceed the grid's y limit",
                (long long)b_vol * n_pose);
  SliceGeom g{b_vol, d, h, w, n_pose, k, cx, cy, cz};
  dim3 grid((k + 255) / 256, b_vol * n_pose);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const void* vp = reinterpret_cast<const void*>(vol);
  void* op = reinterpret_cast<void*>(out);
  switch (vol_dtype) {
    case 0: hipLaunchKernelGGL(slice_fwd_kernel<0>, grid, dim3(256), 0, st, vp, u_xyz, v_xyz, fx, fy, g, op); break;
    case 1: hipLaunchKernelGGL(slice_fwd_kernel<1>, grid, dim3(256), 0, st, vp, u_xyz, v_xyz, fx, fy, g, op); break;
    case 2: hipLaunchKernelGGL(slice_fwd_kernel<2>, grid, dim3(256), 0, st, vp, u_xyz, v_xyz, fx, fy, g, op); break;
    default: set_error("fvr_slice_fwd: bad dtype %d", vol_dtype); return FVR_ERR_INVALID;
  }
  FVR_CHECK_LAUNCH("slice_fwd_kernel");
  return FVR_OK;
}

int64_t fvr_slice_bwd_workspace(int32_t n_pose, int32_t k) { return int64_t(n_pose) * ((k + 255) / 256) * 6; }

int fvr_slice_bwd(uint64_t vol, int32_t vol_dtype, int32_t b_vol, int32_t d, int32_t h, int32_t w, uint64_t grad_out,
                  const float* u_xyz, const float* v_xyz, int32_t n_pose, const float* fx, const float* fy,
                  int32_t k, float cx, float cy, float cz, float* part, float* gu, float* gv, void* stream) {
  FVR_CHECK_ARG(vol && grad_out && u_xyz && v_xyz && fx && fy && part && gu && gv, "fvr_slice_bwd: null pointer");
  FVR_CHECK_ARG(b_vol > 0 && d > 0 && h > 0 && w > 0 && n_pose > 0 && k > 0, "fvr_slice_bwd: empty shape");
  FVR_CHECK_ARG(n_pose <= 65535, "fvr_slice_bwd: %d poses exceed the grid's y limit", n_pose);
  SliceGeom g{b_vol, d, h, w, n_pose, k, cx, cy, cz};
  const int nq = (k + 255) / 256;
  dim3 grid(nq, n_pose);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const void* vp = reinterpret_cast<const void*>(vol);
  const void* gp = reinterpret_cast<const void*>(grad_out);
  switch (vol_dtype) {
    case 0: hipLaunchKernelGGL(slice_bwd_kernel<0>, grid, dim3(256), 0, st, vp, gp, u_xyz, v_xyz, fx, fy, g, part); break;
    case 1: hipLaunchKernelGGL(slice_bwd_kernel<1>, grid, dim3(256), 0, st, vp, gp, u_xyz, v_xyz, fx, fy, g, part); break;
    case 2: hipLaunchKernelGGL(slice_bwd_kernel<2>, grid, dim3(256), 0, st, vp, gp, u_xyz, v_xyz, fx, fy, g, part); break;
    default: set_error("fvr_slice_bwd: bad dtype %d", vol_dtype); return FVR_ERR_INVALID;
  }
  FVR_CHECK_LAUNCH("slice_bwd_kernel");
  hipLaunchKernelGGL(slice_bwd_sum_kernel, dim3((n_pose * 6 + 63) / 64), dim3(64), 0, st, part, nq, n_pose, gu, gv);
  FVR_CHECK_LAUNCH("slice_bwd_sum_kernel");
  return FVR_OK;
}

}  // extern "C"
