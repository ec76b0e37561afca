// Edge-dominant gradient NCC (cone_beam_fvr.py:71-112 of pareelamre/diff-fvr) on gfx950.
//
// Per image: Sobel x / y (3 x 3 cross-correlation, zero padding) -- and, for the
// multiscale term, Sobel x / y of the 3 x 3 box-blurred image (zero padding, divisor 9) --
// over the interior crop [crop, H - crop) x [crop, W - crop); NCC per channel with the
// crop mean removed and the denominator sqrt(max(Eu Ev, 1e-12)); the value is
// 0.5 (ncc_x + ncc_y), or (1 - w) of that plus w times the blurred pair.
//
// Every kernel works on 32 x 32 output tiles, one 256-thread workgroup per (tile, image),
// with the tile's neighbourhood of the image staged in LDS (halo 2 for the forward: the
// blur then the Sobel; halo 4 for the adjoint, which also recomputes the channel values it
// needs).  No atomics: moments are reduced per tile in a fixed order (float64), then per
// image in a fixed order, and the adjoint is a gather (each output pixel sums the stencil
// adjoints of its neighbourhood).
//
//   edge_channels_kernel  channel images of the crop (the target's: v_c)
//   edge_stats_kernel     per tile: sum u_c, sum u_c^2, sum u_c v_c (float64)
//   edge_combine_kernel   per image: NCCs, the value, and the adjoint coefficients
//   edge_bwd_kernel       d value / d image (scaled by the upstream gradient)
//
// Also here: the zero-mean NCC of whole images (ncc_*_kernel, fvr_ncc_fwd / fvr_ncc_bwd).

#include <cmath>

#include "fvr_common.h"

namespace fvr {

constexpr int kET = 32;        // output tile side
constexpr int kEThreads = 256;  // 4 pixels per thread
constexpr int kESums = 12;     // (sum u, sum u^2, sum u v) x 4 channels

struct EdgeGeom {
  int H, W, crop, Hc, Wc, nch;  // nch: 2 (Sobel x, y) or 4 (and of the blurred image)
  int ctx, cty;                 // crop tiles along x, y
};

__device__ __forceinline__ float ld_px(const float* __restrict__ img, int H, int W, int y, int x) {
  return (unsigned)y < (unsigned)H && (unsigned)x < (unsigned)W ? img[size_t(y) * W + x] : 0.f;
}

// Sobel x / y of an LDS array A (pitch PA) around (r, c)
template <int PA>
__device__ __forceinline__ void sobel_at(const float* A, int r, int c, float& gx, float& gy) {
  const float* m = A + (r - 1) * PA + c;
  const float* z = A + r * PA + c;
  const float* p = A + (r + 1) * PA + c;
  gx = (m[1] - m[-1]) + 2.f * (z[1] - z[-1]) + (p[1] - p[-1]);
  gy = (p[-1] - m[-1]) + 2.f * (p[0] - m[0]) + (p[1] - m[1]);
}

// stage the image around a crop tile: P (36 x 36, pitch 37) from image row y0 - 2, column
// x0 - 2; B (34 x 34, pitch 35) = the zero-padded box blur at image rows y0 - 1 .., zero
// where the blurred position lies outside the image (the Sobel's own zero padding)
__device__ __forceinline__ void edge_stage_fwd(const float* __restrict__ img, const EdgeGeom& g, int y0, int x0,
                                               bool blur, float* P, float* B) {
  for (int k = threadIdx.x; k < 36 * 36; k += kEThreads) {
    const int r = k / 36, c = k - r * 36;
    P[r * 37 + c] = ld_px(img, g.H, g.W, y0 - 2 + r, x0 - 2 + c);
  }
  __syncthreads();
  if (blur) {
    for (int k = threadIdx.x; k < 34 * 34; k += kEThreads) {
      const int r = k / 34, c = k - r * 34;
      const int y = y0 - 1 + r, x = x0 - 1 + c;
      float s = 0.f;
      if ((unsigned)y < (unsigned)g.H && (unsigned)x < (unsigned)g.W) {
        const float* a = P + r * 37 + c;  // P row r + 1 is image row y
        s = ((a[0] + a[1] + a[2]) + (a[37] + a[38] + a[39]) + (a[74] + a[75] + a[76])) * (1.f / 9.f);
      }
      B[r * 35 + c] = s;
    }
  }
  __syncthreads();
}

// the nch channel values at tile pixel (a, c) (image row y0 + a, column x0 + c)
__device__ __forceinline__ void edge_channels_at(const float* P, const float* B, int a, int c, int nch, float u[4]) {
  sobel_at<37>(P, a + 2, c + 2, u[0], u[1]);
  if (nch == 4) sobel_at<35>(B, a + 1, c + 1, u[2], u[3]);
}

__global__ __launch_bounds__(kEThreads) void edge_channels_kernel(const float* __restrict__ img, EdgeGeom g,
                                                                  float* __restrict__ ch) {
  __shared__ float P[36 * 37];
  __shared__ float B[34 * 35];
  const int tx = blockIdx.x % g.ctx, ty = blockIdx.x / g.ctx, n = blockIdx.y;
  const int y0 = g.crop + ty * kET, x0 = g.crop + tx * kET;
  edge_stage_fwd(img + size_t(n) * g.H * g.W, g, y0, x0, g.nch == 4, P, B);
  const size_t plane = size_t(g.Hc) * g.Wc;
  for (int k = threadIdx.x; k < kET * kET; k += kEThreads) {
    const int a = k / kET, c = k - a * kET;
    const int yc = ty * kET + a, xc = tx * kET + c;
    if (yc >= g.Hc || xc >= g.Wc) continue;
    float u[4];
    edge_channels_at(P, B, a, c, g.nch, u);
    for (int q = 0; q < g.nch; ++q) ch[(size_t(n) * g.nch + q) * plane + size_t(yc) * g.Wc + xc] = u[q];
  }
}

__global__ __launch_bounds__(kEThreads) void edge_stats_kernel(const float* __restrict__ img, EdgeGeom g,
                                                               const float* __restrict__ vch, int bt,
                                                               double* __restrict__ part) {
  __shared__ float P[36 * 37];
  __shared__ float B[34 * 35];
  __shared__ double wred[kEThreads / 64][kESums];
  const int tx = blockIdx.x % g.ctx, ty = blockIdx.x / g.ctx, n = blockIdx.y;
  const int y0 = g.crop + ty * kET, x0 = g.crop + tx * kET;
  edge_stage_fwd(img + size_t(n) * g.H * g.W, g, y0, x0, g.nch == 4, P, B);
  const size_t plane = size_t(g.Hc) * g.Wc;
  const float* v = vch + size_t(bt == 1 ? 0 : n) * g.nch * plane;
  double s[kESums];
#pragma unroll
  for (int q = 0; q < kESums; ++q) s[q] = 0.0;
  for (int k = threadIdx.x; k < kET * kET; k += kEThreads) {
    const int a = k / kET, c = k - a * kET;
    const int yc = ty * kET + a, xc = tx * kET + c;
    if (yc >= g.Hc || xc >= g.Wc) continue;
    float u[4] = {0.f, 0.f, 0.f, 0.f};
    edge_channels_at(P, B, a, c, g.nch, u);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < g.nch) {
        const double uu = u[q], vv = v[size_t(q) * plane + size_t(yc) * g.Wc + xc];
        s[3 * q] += uu;
        s[3 * q + 1] += uu * uu;
        s[3 * q + 2] += uu * vv;
      }
    }
  }
  // fixed-order reduction: wave butterflies, then the waves in order
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kESums; ++q) {
    double x = s[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (lane == 0) wred[wave][q] = x;
  }
  __syncthreads();
  if (threadIdx.x < kESums) {
    double x = wred[0][threadIdx.x];
    for (int w = 1; w < kEThreads / 64; ++w) x += wred[w][threadIdx.x];
    part[(size_t(n) * g.ctx * g.cty + blockIdx.x) * kESums + threadIdx.x] = x;
  }
}

// one wave per image: sums of its tiles (and of its target's), NCC per channel, the value
// and the adjoint coefficients coef[n][q] = (alpha, beta, mean u, mean v): d value / d u_q =
// alpha (v - mean v) - beta (u - mean u)
__global__ __launch_bounds__(64) void edge_combine_kernel(const double* __restrict__ part,
                                                          const double* __restrict__ tpart, int bt, EdgeGeom g,
                                                          double weight, double* __restrict__ val,
                                                          float* __restrict__ coef) {
  const int n = blockIdx.x, lane = threadIdx.x;
  const int tiles = g.ctx * g.cty;
  const double* pp = part + size_t(n) * tiles * kESums;
  const double* tp = tpart + size_t(bt == 1 ? 0 : n) * tiles * kESums;
  double s[kESums], t[kESums];
#pragma unroll
  for (int q = 0; q < kESums; ++q) s[q] = t[q] = 0.0;
  for (int k = lane; k < tiles; k += 64) {
#pragma unroll
    for (int q = 0; q < kESums; ++q) {
      s[q] += pp[size_t(k) * kESums + q];
      t[q] += tp[size_t(k) * kESums + q];
    }
  }
#pragma unroll
  for (int q = 0; q < kESums; ++q) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      s[q] += __shfl_xor(s[q], o);
      t[q] += __shfl_xor(t[q], o);
    }
  }
  if (lane != 0) return;
  const double npx = double(g.Hc) * double(g.Wc);
  double value = 0.0;
  for (int q = 0; q < 4; ++q) {
    float* cf = coef + (size_t(n) * 4 + q) * 4;
    if (q >= g.nch) {
      cf[0] = cf[1] = cf[2] = cf[3] = 0.f;
      continue;
    }
    const double wq = 0.5 * (g.nch == 4 ? (q < 2 ? 1.0 - weight : weight) : 1.0);
    const double su = s[3 * q], suu = s[3 * q + 1], suv = s[3 * q + 2];
    const double sv = t[3 * q], svv = t[3 * q + 1];  // the target's own sums (u = v there)
    const double A = suv - su * sv / npx;
    const double eu = suu - su * su / npx, ev = svv - sv * sv / npx;
    const double ee = eu * ev;
    const double D = sqrt(fmax(ee, 1e-12));
    value += wq * (A / D);
    cf[0] = float(wq / D);
    cf[1] = float(ee > 1e-12 && eu > 0.0 ? wq * A / (D * eu) : 0.0);
    cf[2] = float(su / npx);
    cf[3] = float(sv / npx);
  }
  val[n] = value;
}

// d value / d image for 32 x 32 output tiles over the whole image: the upstream gradient
// go[n] times the adjoints of the Sobel stencils (fine), of Sobel after the blur (coarse)
__global__ __launch_bounds__(kEThreads) void edge_bwd_kernel(const float* __restrict__ img, EdgeGeom g,
                                                             const float* __restrict__ vch, int bt,
                                                             const float* __restrict__ coef,
                                                             const double* __restrict__ go, int ftx,
                                                             float* __restrict__ grad) {
  __shared__ float P[40 * 41];        // image rows qy0 - 4 ..
  __shared__ float Bl[38 * 39];       // blurred image rows qy0 - 3 ..
  __shared__ float Gc[2][36 * 37];    // coarse channel gradients, rows qy0 - 2 ..
  __shared__ float DB[34 * 35];       // gradient of the blurred image, rows qy0 - 1 ..
  __shared__ float Gf[2][34 * 35];    // fine channel gradients, rows qy0 - 1 ..
  const int n = blockIdx.y;
  const int qx0 = (blockIdx.x % ftx) * kET, qy0 = (blockIdx.x / ftx) * kET;
  const float* im = img + size_t(n) * g.H * g.W;
  const size_t plane = size_t(g.Hc) * g.Wc;
  const float* v = vch + size_t(bt == 1 ? 0 : n) * g.nch * plane;
  const float scale = float(go[n]);
  float al[4], be[4], mu[4], mv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float* cf = coef + (size_t(n) * 4 + q) * 4;
    al[q] = cf[0] * scale;
    be[q] = cf[1] * scale;
    mu[q] = cf[2];
    mv[q] = cf[3];
  }
  const bool blur = g.nch == 4;
  for (int k = threadIdx.x; k < 40 * 40; k += kEThreads) {
    const int r = k / 40, c = k - r * 40;
    P[r * 41 + c] = ld_px(im, g.H, g.W, qy0 - 4 + r, qx0 - 4 + c);
  }
  __syncthreads();
  auto in_crop = [&](int y, int x) {
    return y >= g.crop && y < g.crop + g.Hc && x >= g.crop && x < g.crop + g.Wc;
  };
  if (blur) {
    for (int k = threadIdx.x; k < 38 * 38; k += kEThreads) {
      const int r = k / 38, c = k - r * 38;
      const int y = qy0 - 3 + r, x = qx0 - 3 + c;
      float s = 0.f;
      if ((unsigned)y < (unsigned)g.H && (unsigned)x < (unsigned)g.W) {
        const float* a = P + r * 41 + c;  // P row r + 1 is image row y
        s = ((a[0] + a[1] + a[2]) + (a[41] + a[42] + a[43]) + (a[82] + a[83] + a[84])) * (1.f / 9.f);
      }
      Bl[r * 39 + c] = s;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 36 * 36; k += kEThreads) {
      const int r = k / 36, c = k - r * 36;
      const int y = qy0 - 2 + r, x = qx0 - 2 + c;
      float gx = 0.f, gy = 0.f;
      if (in_crop(y, x)) {
        float ux, uy;
        sobel_at<39>(Bl, r + 1, c + 1, ux, uy);
        const size_t o = size_t(y - g.crop) * g.Wc + (x - g.crop);
        gx = al[2] * (v[2 * plane + o] - mv[2]) - be[2] * (ux - mu[2]);
        gy = al[3] * (v[3 * plane + o] - mv[3]) - be[3] * (uy - mu[3]);
      }
      Gc[0][r * 37 + c] = gx;
      Gc[1][r * 37 + c] = gy;
    }
    __syncthreads();
    // gradient of the blurred image: sum_k K(k) g(j - k) at j in the image
    for (int k = threadIdx.x; k < 34 * 34; k += kEThreads) {
      const int r = k / 34, c = k - r * 34;
      const int y = qy0 - 1 + r, x = qx0 - 1 + c;
      float s = 0.f;
      if ((unsigned)y < (unsigned)g.H && (unsigned)x < (unsigned)g.W) {
        const float* a = Gc[0] + (r + 1) * 37 + (c + 1);  // Gc index of j
        const float* b = Gc[1] + (r + 1) * 37 + (c + 1);
        // Sobel x adjoint: K(dy, dx) at g(j - (dy, dx))
        s = (a[37 - 1] - a[37 + 1]) + 2.f * (a[-1] - a[1]) + (a[-37 - 1] - a[-37 + 1]);
        s += (b[-37 + 1] - b[37 + 1]) + 2.f * (b[-37] - b[37]) + (b[-37 - 1] - b[37 - 1]);
      }
      DB[r * 35 + c] = s;
    }
  }
  for (int k = threadIdx.x; k < 34 * 34; k += kEThreads) {
    const int r = k / 34, c = k - r * 34;
    const int y = qy0 - 1 + r, x = qx0 - 1 + c;
    float gx = 0.f, gy = 0.f;
    if (in_crop(y, x)) {
      float ux, uy;
      sobel_at<41>(P, r + 3, c + 3, ux, uy);
      const size_t o = size_t(y - g.crop) * g.Wc + (x - g.crop);
      gx = al[0] * (v[o] - mv[0]) - be[0] * (ux - mu[0]);
      gy = al[1] * (v[plane + o] - mv[1]) - be[1] * (uy - mu[1]);
    }
    Gf[0][r * 35 + c] = gx;
    Gf[1][r * 35 + c] = gy;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kET * kET; k += kEThreads) {
    const int a = k / kET, c = k - a * kET;
    const int y = qy0 + a, x = qx0 + c;
    if (y >= g.H || x >= g.W) continue;
    const float* fx = Gf[0] + (a + 1) * 35 + (c + 1);
    const float* fy = Gf[1] + (a + 1) * 35 + (c + 1);
    float s = (fx[35 - 1] - fx[35 + 1]) + 2.f * (fx[-1] - fx[1]) + (fx[-35 - 1] - fx[-35 + 1]);
    s += (fy[-35 + 1] - fy[35 + 1]) + 2.f * (fy[-35] - fy[35]) + (fy[-35 - 1] - fy[35 - 1]);
    if (blur) {
      const float* d = DB + (a + 1) * 35 + (c + 1);
      s += ((d[-36] + d[-35] + d[-34]) + (d[-1] + d[0] + d[1]) + (d[34] + d[35] + d[36])) * (1.f / 9.f);
    }
    grad[size_t(n) * g.H * g.W + size_t(y) * g.W + x] = s;
  }
}

static int edge_geom(EdgeGeom& g, int32_t H, int32_t W, int32_t crop, int32_t multi, const char* who) {
  FVR_CHECK_ARG(H > 0 && W > 0 && crop >= 0, "%s: bad image shape %d x %d (crop %d)", who, H, W, crop);
  FVR_CHECK_ARG(H - 2 * crop > 0 && W - 2 * crop > 0, "%s: crop %d leaves nothing of a %d x %d image", who, crop, H,
                W);
  FVR_CHECK_ARG(int64_t(H) * W < (int64_t(1) << 28), "%s: image too large", who);
  g.H = H;
  g.W = W;
  g.crop = crop;
  g.Hc = H - 2 * crop;
  g.Wc = W - 2 * crop;
  g.nch = multi ? 4 : 2;
  g.ctx = (g.Wc + kET - 1) / kET;
  g.cty = (g.Hc + kET - 1) / kET;
  return FVR_OK;
}


// ---------------------------------------------------------------------------
// Zero-mean NCC of whole images, the NCC of the PoU path's registration loss
// (tests/test_cone_beam_fvr.py:24-27 of the reference: 1 - NCC):
//   val = sum(u0 v0) / (|u0| |v0| + 1e-8),  u0 = u - mean u, v0 = v - mean v.
//   ncc_stats_kernel    per chunk of kNccChunk pixels of an image: sum u, u^2, u v, v, v^2
//                       (float64, reduced in a fixed order)
//   ncc_combine_kernel  one wave per image: the chunk sums in a fixed order, the value and the
//                       adjoint coefficients
//   ncc_bwd_kernel      d val / d self = alpha (other - mean other) - beta (self - mean self),
//                       times the upstream gradient (elementwise)
// ---------------------------------------------------------------------------
constexpr int kNccChunk = 4096;
constexpr int kNccThreads = 256;
constexpr int kNccSums = 5;

__host__ __device__ inline long long ncc_chunks(long long K) { return (K + kNccChunk - 1) / kNccChunk; }

__global__ __launch_bounds__(kNccThreads) void ncc_stats_kernel(const float* __restrict__ u,
                                                               const float* __restrict__ v, int bt, long long K,
                                                               bool vec4, double* __restrict__ part) {
  const int n = blockIdx.y;
  const long long c0 = (long long)blockIdx.x * kNccChunk;
  const long long c1 = c0 + kNccChunk < K ? c0 + kNccChunk : K;
  const float* un = u + size_t(n) * size_t(K);
  const float* vn = v + size_t(bt == 1 ? 0 : n) * size_t(K);
  double s[kNccSums] = {0.0, 0.0, 0.0, 0.0, 0.0};
  auto add = [&](float a, float b) {
    const double x = a, y = b;
    s[0] += x;
    s[1] = fma(x, x, s[1]);
    s[2] = fma(x, y, s[2]);
    s[3] += y;
    s[4] = fma(y, y, s[4]);
  };
  if (vec4) {  // K a multiple of 4 and 16-byte aligned rows: float4 loads
    for (long long k = c0 + 4 * threadIdx.x; k < c1; k += 4 * kNccThreads) {
      const float4 a = *reinterpret_cast<const float4*>(un + k);
      const float4 b = *reinterpret_cast<const float4*>(vn + k);
      add(a.x, b.x);
      add(a.y, b.y);
      add(a.z, b.z);
      add(a.w, b.w);
    }
  } else {
    for (long long k = c0 + threadIdx.x; k < c1; k += kNccThreads) add(un[k], vn[k]);
  }
  __shared__ double wred[kNccThreads / 64][kNccSums];
#pragma unroll
  for (int q = 0; q < kNccSums; ++q) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s[q] += __shfl_xor(s[q], o);
  }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < kNccSums; ++q) wred[threadIdx.x >> 6][q] = s[q];
  __syncthreads();
  if (threadIdx.x < kNccSums) {
    double x = wred[0][threadIdx.x];
    for (int w = 1; w < kNccThreads / 64; ++w) x += wred[w][threadIdx.x];
    part[(size_t(n) * gridDim.x + blockIdx.x) * kNccSums + threadIdx.x] = x;
  }
}

// coef[n] = (alpha, beta_u, mean u, mean v, alpha, beta_v, mean v, mean u): the adjoint with
// respect to the image (first four) and to the target (last four)
__global__ __launch_bounds__(64) void ncc_combine_kernel(const double* __restrict__ part, int chunks, long long K,
                                                         double* __restrict__ val, float* __restrict__ coef) {
  const int n = blockIdx.x, lane = threadIdx.x;
  const double* pp = part + size_t(n) * chunks * kNccSums;
  double s[kNccSums] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int c = lane; c < chunks; c += 64)
#pragma unroll
    for (int q = 0; q < kNccSums; ++q) s[q] += pp[size_t(c) * kNccSums + q];
#pragma unroll
  for (int q = 0; q < kNccSums; ++q) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s[q] += __shfl_xor(s[q], o);
  }
  if (lane != 0) return;
  const double npx = double(K);
  const double su = s[0], suu = s[1], suv = s[2], sv = s[3], svv = s[4];
  const double A = suv - su * sv / npx;
  const double nu = sqrt(fmax(suu - su * su / npx, 0.0)), nv = sqrt(fmax(svv - sv * sv / npx, 0.0));
  const double D = nu * nv + 1e-8;
  val[n] = A / D;
  // d val / d u = (v0 - A nv u0 / (nu D)) / D (nothing from |u0| for a constant image)
  float* cf = coef + size_t(n) * 8;
  cf[0] = cf[4] = float(1.0 / D);
  cf[1] = float(nu > 0.0 ? A * nv / (nu * D * D) : 0.0);
  cf[5] = float(nv > 0.0 ? A * nu / (nv * D * D) : 0.0);
  cf[2] = cf[7] = float(su / npx);
  cf[3] = cf[6] = float(sv / npx);
}

// grad[n, k] = gval[n] (alpha (other_k - mean other) - beta (self_k - mean self)); self = the
// image n (wrt 0) or its target (wrt 1), other the other one
__global__ __launch_bounds__(kNccThreads) void ncc_bwd_kernel(const float* __restrict__ u,
                                                             const float* __restrict__ v, int bt, long long K,
                                                             const float* __restrict__ coef,
                                                             const double* __restrict__ gval, int wrt,
                                                             float* __restrict__ grad) {
  const int n = blockIdx.y;
  const float* un = u + size_t(n) * size_t(K);
  const float* vn = v + size_t(bt == 1 ? 0 : n) * size_t(K);
  const float* self = wrt ? vn : un;
  const float* other = wrt ? un : vn;
  const float* cf = coef + size_t(n) * 8 + 4 * wrt;
  const float g = float(gval[n]);
  const float al = g * cf[0], be = g * cf[1], ms = cf[2], mo = cf[3];
  float* gn = grad + size_t(n) * size_t(K);
  const long long c0 = (long long)blockIdx.x * kNccChunk;
  const long long c1 = c0 + kNccChunk < K ? c0 + kNccChunk : K;
  for (long long k = c0 + threadIdx.x; k < c1; k += kNccThreads)
    gn[k] = fmaf(al, other[k] - mo, -be * (self[k] - ms));
}

}  // namespace fvr

using namespace fvr;

extern "C" {

int32_t fvr_edge_tiles(int32_t H, int32_t W, int32_t crop) {
  EdgeGeom g;
  if (edge_geom(g, H, W, crop, 0, "fvr_edge_tiles") != FVR_OK) return -1;
  return g.ctx * g.cty;
}

int fvr_edge_channels(const float* img, int32_t n_img, int32_t H, int32_t W, int32_t crop, int32_t multi,
                      float* channels, void* stream) {
  FVR_CHECK_ARG(img && channels && n_img > 0 && n_img < 65536, "fvr_edge_channels: bad arguments");
  EdgeGeom g;
  if (edge_geom(g, H, W, crop, multi, "fvr_edge_channels") != FVR_OK) return FVR_ERR_INVALID;
  hipLaunchKernelGGL(edge_channels_kernel, dim3(g.ctx * g.cty, n_img), dim3(kEThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), img, g, channels);
  FVR_CHECK_LAUNCH("edge_channels_kernel");
  return FVR_OK;
}

int fvr_edge_stats(const float* img, int32_t n_img, int32_t H, int32_t W, int32_t crop, int32_t multi,
                   const float* channels, int32_t n_ch_img, double* partials, void* stream) {
  FVR_CHECK_ARG(img && channels && partials && n_img > 0 && n_img < 65536, "fvr_edge_stats: bad arguments");
  FVR_CHECK_ARG(n_ch_img == 1 || n_ch_img == n_img, "fvr_edge_stats: %d channel images for %d images", n_ch_img,
                n_img);
  EdgeGeom g;
  if (edge_geom(g, H, W, crop, multi, "fvr_edge_stats") != FVR_OK) return FVR_ERR_INVALID;
  hipLaunchKernelGGL(edge_stats_kernel, dim3(g.ctx * g.cty, n_img), dim3(kEThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), img, g, channels, n_ch_img, partials);
  FVR_CHECK_LAUNCH("edge_stats_kernel");
  return FVR_OK;
}

int fvr_edge_combine(const double* partials, int32_t n_img, const double* target_partials, int32_t n_target,
                     int32_t H, int32_t W, int32_t crop, int32_t multi, double weight, double* value, float* coef,
                     void* stream) {
  FVR_CHECK_ARG(partials && target_partials && value && coef && n_img > 0, "fvr_edge_combine: bad arguments");
  FVR_CHECK_ARG(n_target == 1 || n_target == n_img, "fvr_edge_combine: %d targets for %d images", n_target, n_img);
  EdgeGeom g;
  if (edge_geom(g, H, W, crop, multi, "fvr_edge_combine") != FVR_OK) return FVR_ERR_INVALID;
  hipLaunchKernelGGL(edge_combine_kernel, dim3(n_img), dim3(64), 0, reinterpret_cast<hipStream_t>(stream), partials,
                     target_partials, n_target, g, weight, value, coef);
  FVR_CHECK_LAUNCH("edge_combine_kernel");
  return FVR_OK;
}

int fvr_edge_bwd(const float* img, int32_t n_img, int32_t H, int32_t W, int32_t crop, int32_t multi,
                 const float* channels, int32_t n_ch_img, const float* coef, const double* grad_value, float* grad,
                 void* stream) {
  FVR_CHECK_ARG(img && channels && coef && grad_value && grad && n_img > 0 && n_img < 65536,
                "fvr_edge_bwd: bad arguments");
  FVR_CHECK_ARG(n_ch_img == 1 || n_ch_img == n_img, "fvr_edge_bwd: %d channel images for %d images", n_ch_img,
                n_img);
  EdgeGeom g;
  if (edge_geom(g, H, W, crop, multi, "fvr_edge_bwd") != FVR_OK) return FVR_ERR_INVALID;
  const int ftx = (W + kET - 1) / kET, fty = (H + kET - 1) / kET;
  hipLaunchKernelGGL(edge_bwd_kernel, dim3(ftx * fty, n_img), dim3(kEThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), img, g, channels, n_ch_img, coef, grad_value, ftx, grad);
  FVR_CHECK_LAUNCH("edge_bwd_kernel");
  return FVR_OK;
}

int64_t fvr_ncc_work_doubles(int32_t n_img, int64_t K) {
  if (n_img <= 0 || K <= 0) return -1;
  return int64_t(n_img) * ncc_chunks(K) * kNccSums;
}

int fvr_ncc_fwd(const float* img, int32_t n_img, const float* target, int32_t n_target, int64_t K, double* work,
                double* value, float* coef, void* stream) {
  FVR_CHECK_ARG(img && target && work && value && coef && n_img > 0 && n_img < 65536 && K > 0,
                "fvr_ncc_fwd: bad arguments");
  FVR_CHECK_ARG(n_target == 1 || n_target == n_img, "fvr_ncc_fwd: %d targets for %d images", n_target, n_img);
  const long long chunks = ncc_chunks(K);
  FVR_CHECK_ARG(chunks < (1ll << 31), "fvr_ncc_fwd: %lld pixels per image exceed the grid", (long long)K);
  const bool vec4 = (K % 4 == 0) && (reinterpret_cast<uintptr_t>(img) % 16 == 0) &&
                    (reinterpret_cast<uintptr_t>(target) % 16 == 0);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(ncc_stats_kernel, dim3(unsigned(chunks), n_img), dim3(kNccThreads), 0, st, img, target, n_target,
                     (long long)K, vec4, work);
  FVR_CHECK_LAUNCH("ncc_stats_kernel");
  hipLaunchKernelGGL(ncc_combine_kernel, dim3(n_img), dim3(64), 0, st, work, int(chunks), (long long)K, value, coef);
  FVR_CHECK_LAUNCH("ncc_combine_kernel");
  return FVR_OK;
}

int fvr_ncc_bwd(const float* img, int32_t n_img, const float* target, int32_t n_target, int64_t K, const float* coef,
                const double* grad_value, int32_t wrt_target, float* grad, void* stream) {
  FVR_CHECK_ARG(img && target && coef && grad_value && grad && n_img > 0 && n_img < 65536 && K > 0,
                "fvr_ncc_bwd: bad arguments");
  FVR_CHECK_ARG(n_target == 1 || n_target == n_img, "fvr_ncc_bwd: %d targets for %d images", n_target, n_img);
  const long long chunks = ncc_chunks(K);
  FVR_CHECK_ARG(chunks < (1ll << 31), "fvr_ncc_bwd: %lld pixels per image exceed the grid", (long long)K);
  hipLaunchKernelGGL(ncc_bwd_kernel, dim3(unsigned(chunks), n_img), dim3(kNccThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), img, target, n_target, (long long)K, coef, grad_value,
                     wrt_target ? 1 : 0, grad);
  FVR_CHECK_LAUNCH("ncc_bwd_kernel");
  return FVR_OK;
}

}  // extern "C"
