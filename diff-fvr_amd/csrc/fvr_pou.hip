// PoU-FVR perspective path (cone_beam_fvr.py, triton_slice.py:349-808).
//
//   fvr_pou_pack_blocks  planar (B, 2, P, P, P) block spectra -> the gather layout
//                        of the fused kernels (interleaved (re, im), zero margins).
//   fvr_pou_slice_ifft   one workgroup per (block, pose): the P x P central slice
//                        of the block spectrum (per-corner trilinear, packed
//                        storage) is gathered straight into LDS,
//                        inverse-transformed there (2D, radix-4 passes, unnormalised,
//                        centred: fftshift(ifft2(ifftshift(.)))) and its real part
//                        / P^2 written as the slab image (B, N, P, P) float.
//   fvr_pou_warp_fwd     detector pixel (n, q): sum_b bilinear(img_bn, p_bn(q)) / cos,
//                        p = (d_b r - c) / ds + P/2 (triton_slice.py:349-407).
//   fvr_pou_warp_bwd     one workgroup per (block, pose): d/d(d_b, cu, cv) and,
//                        optionally, d/d img (scatter into 64-bit fixed point:
//                        integer adds commute, so the result does not depend on
//                        the order in which waves arrive).
//   fvr_pou_render_bwd   the fused adjoint of slice -> iFFT -> warp for one
//                        (block, pose): warp adjoint scattered into an LDS
//                        fixed-point image, forward 2D FFT in LDS, then every
//                        slice sample is recomputed and contracted with its
//                        gradient: d/d(d_b, cu, cv, u, v) per (block, pose).
//   fvr_pou_ray_grads    d/d(rx, ry, inv_cos) per (pose, pixel) for per-pose rays.
//   fvr_pou_gscale       per-pose power-of-two fixed-point scale of |grad * inv_cos|.
//   fvr_pou_sum_blocks   fixed-order sum over blocks of per-(block, pose) partials.
//   fvr_pou_pose_geom(_bwd)  pose matrices -> detector axes and block offsets of the
//                        lattice path, and its adjoint from the fused adjoint's sums.
#include <atomic>
#include <cmath>
#include <type_traits>

#include "fvr_common.h"

namespace fvr {

constexpr int kPouThreads = 512;
constexpr int kPouSliceThreads = 1024;  // the fused adjoint (the full-size LDS image: one workgroup per CU)
constexpr int kPouHalfThreads = 512;    // the forward slice + iFFT (half-spectrum LDS: two workgroups per CU)
#ifndef FVR_POU_BATCH
#define FVR_POU_BATCH 4
#endif
constexpr int kPouBatch = FVR_POU_BATCH;  // slice samples in flight per thread during a gather
#ifndef FVR_POU_FWD_BATCH
#define FVR_POU_FWD_BATCH 4
#endif
// the same in the forward kernel (95 VGPRs: its occupancy is set by its LDS, not its registers)
constexpr int kFwdBatch = FVR_POU_FWD_BATCH;
constexpr int kPouMaxLogP = 7;          // LDS-resident transforms up to 128 x 128
constexpr int kPouMaxWaves = 16;        // waves of the largest workgroup (1024 threads)

typedef float pf4a8 __attribute__((ext_vector_type(4), aligned(8)));
typedef uint32_t pu2a4 __attribute__((ext_vector_type(2), aligned(4)));

// Gather layout of the fused kernels (fvr_pou_pack_blocks): per block (P+3) x (P+2) x (P+2)
// elements (z, y, x; storage index = logical index + 1 with zero margins, the slow axis one
// zero plane longer: the guard), 8 bytes each: the (re, im) pair as float2 for f32, or for
// f16 / bf16 the (re | im << 16) words of rows y and y + 1 (the row pair), so that a 16-byte
// load returns a cell's four (y, x) corners in one plane and a sample takes two loads instead
// of four pair loads (twice the footprint of one word per element: config 4's 55 blocks take
// 0.97 GB).  Every trilinear corner of a sample with
// floor(p) in [-1, P-1] is a valid address (out-of-range corners read the zero margins,
// i.e. grid_sample's per-corner zero padding); a sample with no corner inside reads the
// guard planes.
struct PouPk {
  uint32_t esz, row, plane;  // bytes
  unsigned long long block;  // bytes per block
};

__host__ __device__ inline PouPk pou_pk(int P, int vt) {
  PouPk k;
  k.esz = 8u;  // float2 (f32), or the 16-bit pairs of two mid-axis rows (f16 / bf16)
  k.row = uint32_t(P + 2) * k.esz;
  k.plane = k.row * uint32_t(P + 2);
  k.block = (unsigned long long)k.plane * (unsigned long long)(P + 3);
  return k;
}

template <int VT>
struct PkRaw;
template <>
struct PkRaw<0> {
  pf4a8 q[4];  // corner pairs (z, y) = (0,0), (0,1), (1,0), (1,1): (re x0, im x0, re x1, im x1)
};
typedef uint32_t pu4a8 __attribute__((ext_vector_type(4), aligned(8)));
template <int VT>
struct PkRaw {
  pu4a8 q[2];  // z0, z1: the 16-bit (re | im << 16) entries (y0, x0), (y1, x0), (y0, x1), (y1, x1)
};

template <int VT>
struct PkSmp {
  PkRaw<VT> r;
  float tx, ty, tz;
};

// whether storage index X = floor(p) + 1 of the low corner is 0, 1, P or P + 1, i.e. a corner of
// p's cell is storage index 0 or P (floor(p) in {-1, 0, P - 1, P}); P a power of two
__device__ __forceinline__ bool pou_edge_idx(int X, int P) {
  return (unsigned(X) <= unsigned(P + 1)) & ((X & (P - 1)) <= 1);
}

// issue the four pair loads of the sample at p (x, y, z block index space): a uniform block
// base plus 32-bit offsets (the block is below 4 GB); returns whether p's cell has a corner at
// storage index 0 or P on some axis (pou_edge_idx)
template <int VT>
__device__ __forceinline__ bool pk_issue(PkSmp<VT>& s, unsigned long long blk, const PouPk& k, int P, float px,
                                         float py, float pz) {
  const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
  s.tx = px - fx;
  s.ty = py - fy;
  s.tz = pz - fz;
  // storage index of the low corner (unsigned add: no overflow for a saturated conversion)
  const int X = int(unsigned(int(fx)) + 1u), Y = int(unsigned(int(fy)) + 1u), Z = int(unsigned(int(fz)) + 1u);
  const bool in = (unsigned(X) <= unsigned(P)) & (unsigned(Y) <= unsigned(P)) & (unsigned(Z) <= unsigned(P));
  const uint32_t o0 = in ? __umul24(uint32_t(Z), k.plane) + __umul24(uint32_t(Y), k.row) + uint32_t(X) * k.esz
                         : uint32_t(P + 1) * k.plane;
  const uint32_t o1 = o0 + k.row, o2 = o0 + k.plane, o3 = o2 + k.row;
  typedef const char __attribute__((address_space(1)))* gc;
  const gc b = reinterpret_cast<gc>(blk);
  if (VT == 0) {
    typedef const pf4a8 __attribute__((address_space(1)))* gp;
    PkRaw<0>& r = reinterpret_cast<PkRaw<0>&>(s.r);
    r.q[0] = *reinterpret_cast<gp>(b + o0);
    r.q[1] = *reinterpret_cast<gp>(b + o1);
    r.q[2] = *reinterpret_cast<gp>(b + o2);
    r.q[3] = *reinterpret_cast<gp>(b + o3);
  } else {  // row-pair layout: one 16-byte load per plane holds the cell's four (y, x) corners
    typedef const pu4a8 __attribute__((address_space(1)))* gp;
    PkRaw<1>& r = reinterpret_cast<PkRaw<1>&>(s.r);
    r.q[0] = *reinterpret_cast<gp>(b + o0);
    r.q[1] = *reinterpret_cast<gp>(b + o2);
  }
  return pou_edge_idx(X, P) | pou_edge_idx(Y, P) | pou_edge_idx(Z, P);
}

template <int VT>
__device__ __forceinline__ float pk_deq(uint32_t h) {
  return VT == 1 ? h2f(uint16_t(h)) : b2f(uint16_t(h));
}

// corner values (corner c = x + 2 y + 4 z) of an issued sample
template <int VT>
__device__ __forceinline__ void pk_corners(const PkSmp<VT>& s, float (&cr)[8], float (&ci)[8]) {
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    if (VT == 0) {
      const PkRaw<0>& r = reinterpret_cast<const PkRaw<0>&>(s.r);
      cr[2 * p] = r.q[p].x;
      ci[2 * p] = r.q[p].y;
      cr[2 * p + 1] = r.q[p].z;
      ci[2 * p + 1] = r.q[p].w;
    } else {
      // pair p = (z, y): plane z = p >> 1, row y = p & 1; entries (y, x0) at component y, (y, x1)
      // at component 2 + y of the plane's load
      const PkRaw<1>& r = reinterpret_cast<const PkRaw<1>&>(s.r);
      const pu4a8 v = r.q[p >> 1];
      const uint32_t lo = (p & 1) ? v.y : v.x, hi = (p & 1) ? v.w : v.z;
      cr[2 * p] = pk_deq<VT>(lo & 0xFFFFu);
      ci[2 * p] = pk_deq<VT>(lo >> 16);
      cr[2 * p + 1] = pk_deq<VT>(hi & 0xFFFFu);
      ci[2 * p + 1] = pk_deq<VT>(hi >> 16);
    }
  }
}

__device__ __forceinline__ int brev(int x, int logp) {
  return int(__builtin_bitreverse32(uint32_t(x)) >> (32 - logp));
}

__device__ __forceinline__ void make_twiddles(float2* tw, int P) {
  for (int k = threadIdx.x; k < P / 2; k += blockDim.x) {
    float s, c;
    sincospif(-2.f * float(k) / float(P), &s, &c);
    tw[k] = make_float2(c, s);
  }
}

// Batched 1D decimation-in-frequency FFTs of length L = 2^logL over nlines lines of an LDS
// complex array: element i of line l at a[l * lpitch + i * epitch]; natural-order input,
// bit-reversed output, unnormalised; SIGN = +1 inverse, -1 forward.  Twiddles of length L
// are tw[e * twstep] with tw[k] = exp(-2 pi i k / P) (twstep = P / L).  Pairs of radix-2 stages
// run as radix-4 passes: stages s and s - 1 of a block of 2^s elements act on the four elements
// k, k + q, k + 2q, k + 3q (q = 2^(s-2)) -- stage s pairs (k, k + 2q) and (k + q, k + 3q) with
// twiddles k ts and (k + q) ts (ts = L >> s), stage s - 1 pairs the results within each half with
// twiddle 2k ts -- so each element is read and written once per two stages (an odd logL ends
// with one radix-2 pass).  COLS: consecutive threads take
// consecutive lines (lines adjacent in memory); otherwise consecutive groups of one line.
template <int SIGN, bool COLS>
__device__ void fft_lines_dif4(float2* a, int nlines, int L, int logL, int lpitch, int epitch, const float2* tw,
                               int twstep) {
  auto twid = [&](int e) {
    float2 w = tw[e * twstep];
    if (SIGN > 0) w.y = -w.y;
    return w;
  };
  auto bfly = [&](float2& x0, float2& x1, float2 w) {  // x0 <- x0 + x1, x1 <- (x0 - x1) w
    const float dr = x0.x - x1.x, di = x0.y - x1.y;
    x0 = make_float2(x0.x + x1.x, x0.y + x1.y);
    x1 = make_float2(dr * w.x - di * w.y, dr * w.y + di * w.x);
  };
  // COLS with nlines - 1 a power of two (the P/2 + 1 columns of a half spectrum): no division
  const int nmain = nlines - 1;
  const bool pow2m1 = nmain > 0 && (nmain & (nmain - 1)) == 0;
  const int lmain = pow2m1 ? 31 - __builtin_clz(unsigned(nmain)) : 0;
  int s = logL;
#pragma unroll 1
  for (; s >= 2; s -= 2) {
    const int q = 1 << (s - 2), ts = L >> s, gpl = L >> 2;
    const int ng = nlines * gpl;
    for (int idx = threadIdx.x; idx < ng; idx += blockDim.x) {
      int line, g;
      if (COLS) {
        // lines first; with nlines - 1 a power of two the first nlines - 1 take shifts and the
        // last line the tail of the index range
        if (pow2m1) {
          if (idx < nmain * gpl) {
            line = idx & (nmain - 1);
            g = idx >> lmain;
          } else {
            line = nmain;
            g = idx - nmain * gpl;
          }
        } else {
          g = idx / nlines;
          line = idx - g * nlines;
        }
      } else {
        line = idx >> (logL - 2);
        g = idx & (gpl - 1);
      }
      const int k = g & (q - 1);
      const int i0 = ((g >> (s - 2)) << s) + k;
      float2* base = a + line * lpitch;
      float2& r0 = base[i0 * epitch];
      float2& r1 = base[(i0 + q) * epitch];
      float2& r2 = base[(i0 + 2 * q) * epitch];
      float2& r3 = base[(i0 + 3 * q) * epitch];
      float2 x0 = r0, x1 = r1, x2 = r2, x3 = r3;
      bfly(x0, x2, twid(k * ts));
      bfly(x1, x3, twid((k + q) * ts));
      const float2 w2 = twid(2 * k * ts);
      bfly(x0, x1, w2);
      bfly(x2, x3, w2);
      r0 = x0;
      r1 = x1;
      r2 = x2;
      r3 = x3;
    }
    __syncthreads();
  }
  if (s == 1) {  // the last radix-2 stage (twiddle 1)
    const int hl = L >> 1, ng = nlines * hl;
    for (int idx = threadIdx.x; idx < ng; idx += blockDim.x) {
      int line, k;
      if (COLS) {
        if (pow2m1) {
          if (idx < nmain * hl) {
            line = idx & (nmain - 1);
            k = idx >> lmain;
          } else {
            line = nmain;
            k = idx - nmain * hl;
          }
        } else {
          k = idx / nlines;
          line = idx - k * nlines;
        }
      } else {
        line = idx >> (logL - 1);
        k = idx & (hl - 1);
      }
      float2* base = a + line * lpitch;
      float2& r0 = base[(2 * k) * epitch];
      float2& r1 = base[(2 * k + 1) * epitch];
      float2 x0 = r0, x1 = r1;
      bfly(x0, x1, make_float2(1.f, 0.f));
      r0 = x0;
      r1 = x1;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// P = 128: the 2D transforms with register DFTs, two LDS passes per dimension
// (the LDS array: 128 rows x 65 columns, pitch 65).  SIGN = +1: exponent +2 pi i n k / n
// (inverse), -1: forward; unnormalised.
//   columns (length 128 = 16 x 8): pass A, task (column, n1 < 8), a 16-point DFT over
//   n2 of x[n1 + 8 n2] times W128^(n1 k2) back to row n1 + 8 k2; pass B, task (column,
//   k2 < 16), an 8-point DFT over the rows 8 k2 .. 8 k2 + 7: X[k2 + 16 k1] at row
//   8 k2 + k1 (pou_row128).
//   rows (length 64 = 8 x 8): the same with 8-point DFTs: X[k2 + 8 k1] at column 8 k2 + k1
//   (pou_col64).
// Columns 0 and 64 of a half spectrum are both real after the transform (inverse) or before
// it (forward), so they share one complex transform: column 0 carries col0 + i col64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 w) {
  return make_float2(fmaf(a.x, w.x, -a.y * w.y), fmaf(a.x, w.y, a.y * w.x));
}
// a * (SIGN i)
template <int SIGN>
__device__ __forceinline__ float2 cmul_i(float2 a) {
  return SIGN > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// a * (c + SIGN i s)
template <int SIGN>
__device__ __forceinline__ float2 cmul_cs(float2 a, float c, float s) {
  const float ss = SIGN > 0 ? s : -s;
  return make_float2(fmaf(a.x, c, -a.y * ss), fmaf(a.x, ss, a.y * c));
}

template <int SIGN>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3), d13 = cmul_i<SIGN>(csub(a1, a3));
  a0 = cadd(s02, s13);
  a2 = csub(s02, s13);
  a1 = cadd(d02, d13);
  a3 = csub(d02, d13);
}

constexpr float kR2 = 0.70710678118654752f;  // 1 / sqrt 2
constexpr float kC8 = 0.92387953251128676f;  // cos(pi / 8)
constexpr float kS8 = 0.38268343236508977f;  // sin(pi / 8)

// natural order in and out: even / odd 4-point DFTs, twiddles w8^k
template <int SIGN>
__device__ __forceinline__ void dft8(float2 (&v)[8]) {
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<SIGN>(e0, e1, e2, e3);
  dft4<SIGN>(o0, o1, o2, o3);
  o1 = cmul_cs<SIGN>(o1, kR2, kR2);
  o2 = cmul_i<SIGN>(o2);
  o3 = cmul_cs<SIGN>(o3, -kR2, kR2);
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

// natural order in and out: n = n1 + 4 n2, k = k2 + 4 k1; 4-point DFTs over n2, twiddles
// w16^(n1 k2), 4-point DFTs over n1
template <int SIGN>
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) dft4<SIGN>(v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]);
  // v[n1 + 4 k2] now holds A[n1][k2]
  v[1 + 4 * 1] = cmul_cs<SIGN>(v[5], kC8, kS8);     // w16^1
  v[1 + 4 * 2] = cmul_cs<SIGN>(v[9], kR2, kR2);     // w16^2
  v[1 + 4 * 3] = cmul_cs<SIGN>(v[13], kS8, kC8);    // w16^3
  v[2 + 4 * 1] = cmul_cs<SIGN>(v[6], kR2, kR2);     // w16^2
  v[2 + 4 * 2] = cmul_i<SIGN>(v[10]);               // w16^4
  v[2 + 4 * 3] = cmul_cs<SIGN>(v[14], -kR2, kR2);   // w16^6
  v[3 + 4 * 1] = cmul_cs<SIGN>(v[7], kS8, kC8);     // w16^3
  v[3 + 4 * 2] = cmul_cs<SIGN>(v[11], -kR2, kR2);   // w16^6
  v[3 + 4 * 3] = cmul_cs<SIGN>(v[15], -kC8, -kS8);  // w16^9
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) dft4<SIGN>(v[4 * k2], v[4 * k2 + 1], v[4 * k2 + 2], v[4 * k2 + 3]);
  // v[4 k2 + k1] holds X[k2 + 4 k1]: transpose to natural order
  float2 t[16];
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2)
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) t[k2 + 4 * k1] = v[4 * k2 + k1];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = t[k];
}

// W128^e with the sign of the transform, e in [0, 128), from tw[k] = exp(-2 pi i k / 128), k < 64
template <int SIGN>
__device__ __forceinline__ float2 tw128(const float2* tw, int e) {
  float2 w = tw[e & 63];
  if (e & 64) w = make_float2(-w.x, -w.y);
  if (SIGN > 0) w.y = -w.y;
  return w;
}

// storage row of frequency m after the column transform / column of n after the row transform
__device__ __forceinline__ int pou_row128(int m) { return 8 * (m & 15) + (m >> 4); }
__device__ __forceinline__ int pou_col64(int n) { return 8 * (n & 7) + (n >> 3); }

// the 64 column lines of a 128 x 65 half spectrum (column 0 packing col0 + i col64 when pack)
template <int SIGN, int NT>
__device__ void pou_cols128(float2* a, const float2* tw, bool pack) {
  constexpr int kP = 65;
  for (int task = threadIdx.x; task < 512; task += NT) {  // pass A: (v, n1)
    const int v = task & 63, n1 = task >> 6;
    float2 x[16];
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      x[n2] = a[(n1 + 8 * n2) * kP + v];
      if (pack && v == 0) {
        const float2 h = a[(n1 + 8 * n2) * kP + 64];
        x[n2] = make_float2(x[n2].x - h.y, x[n2].y + h.x);
      }
    }
    dft16<SIGN>(x);
#pragma unroll
    for (int k2 = 1; k2 < 16; ++k2) x[k2] = cmul(x[k2], tw128<SIGN>(tw, n1 * k2));
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) a[(n1 + 8 * k2) * kP + v] = x[k2];
  }
  __syncthreads();
  for (int task = threadIdx.x; task < 1024; task += NT) {  // pass B: (v, k2)
    const int v = task & 63, k2 = task >> 6;
    float2 y[8];
#pragma unroll
    for (int n1 = 0; n1 < 8; ++n1) y[n1] = a[(8 * k2 + n1) * kP + v];
    dft8<SIGN>(y);
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) a[(8 * k2 + k1) * kP + v] = y[k1];
  }
  __syncthreads();
}

// the 128 row lines (length 64) of a 128 x 65 array
template <int SIGN, int NT>
__device__ void pou_rows64(float2* a, const float2* tw) {
  constexpr int kP = 65;
  for (int task = threadIdx.x; task < 1024; task += NT) {  // pass A: (row, n1)
    const int r = task & 127, n1 = task >> 7;
    float2 x[8];
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2) x[n2] = a[r * kP + n1 + 8 * n2];
    dft8<SIGN>(x);
#pragma unroll
    for (int k2 = 1; k2 < 8; ++k2) x[k2] = cmul(x[k2], tw128<SIGN>(tw, 2 * n1 * k2));
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) a[r * kP + n1 + 8 * k2] = x[k2];
  }
  __syncthreads();
  for (int task = threadIdx.x; task < 1024; task += NT) {  // pass B: (row, k2)
    const int r = task & 127, k2 = task >> 7;
    float2 y[8];
#pragma unroll
    for (int n1 = 0; n1 < 8; ++n1) y[n1] = a[r * kP + 8 * k2 + n1];
    dft8<SIGN>(y);
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) a[r * kP + 8 * k2 + k1] = y[k1];
  }
  __syncthreads();
}

// trilinear value (and gradient) from the 8 corners (c = x + 2 y + 4 z) and the fractions:
// lerps along x (the difference d = high - low kept), then y, then z; the x-gradient is the
// bilinear (y, z) interpolation of the four x-differences, the y-gradient the z-lerp of the two
// y-differences, the z-gradient the last difference (the exact derivatives of the same trilinear
// interpolant, sharing its partial results)
template <bool GRAD>
__device__ __forceinline__ void interp8(const float (&cr)[8], const float (&ci)[8], float tx, float ty, float tz,
                                        float& re, float& im, float* gre, float* gim) {
  auto comp = [&](const float (&c)[8], float& val, float* g) {
    float d[4], e[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {  // k = y + 2 z
      d[k] = c[2 * k + 1] - c[2 * k];
      e[k] = fmaf(tx, d[k], c[2 * k]);
    }
    const float y0 = e[1] - e[0], y1 = e[3] - e[2];
    const float f0 = fmaf(ty, y0, e[0]), f1 = fmaf(ty, y1, e[2]);
    const float z = f1 - f0;
    val = fmaf(tz, z, f0);
    if (GRAD) {
      const float x0 = fmaf(ty, d[1] - d[0], d[0]), x1 = fmaf(ty, d[3] - d[2], d[2]);
      g[0] = fmaf(tz, x1 - x0, x0);
      g[1] = fmaf(tz, y1 - y0, y0);
      g[2] = z;
    }
  };
  comp(cr, re, gre);
  comp(ci, im, gim);
}

struct PouSlice {
  int B, N, P, logP;
  float c;   // P // 2, the centre voxel of every axis
  int mcap;  // entries of the workgroup's explicit-mirror queue (pou_mirror_queue)
};

// Capacity of the explicit-mirror queue of one workgroup: every stored item of the half plane
// (P x P/2 + P) up to 4096 (8 KB of 16-bit entries, so that two 512-thread workgroups of the
// forward kernel still fit in a CU's LDS at P = 128)
__host__ __device__ inline int pou_mirror_cap(int P) {
  const int n = P * (P / 2) + P;
  return n < 4096 ? n : 4096;
}
__host__ __device__ inline size_t pou_queue_bytes(int P) { return (size_t(pou_mirror_cap(P)) * 2 + 15) & ~size_t(15); }

// Deterministic compaction of the items whose DFT mirror is gathered explicitly: bit r of
// `bits` stands for the thread's item threadIdx.x + r * stride; the items go to
// queue[off ...] in bit order, off the exclusive prefix sum of the threads' counts in thread
// order (a shuffle scan per wave, the wave totals through `scan`, one int per wave).  Returns
// the number of entries, or -1 (nothing written) when they exceed cap: the caller then takes
// its per-thread pass over `bits`.  Contains two workgroup barriers.
__device__ __forceinline__ int pou_mirror_queue(uint32_t bits, int stride, int* scan, uint16_t* queue, int cap) {
  const int c = __popc(bits), lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int x = c;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) scan[w] = x;
  __syncthreads();
  int off = x - c, tot = 0;
  for (int k = 0; k < nw; ++k) {
    const int t = scan[k];
    off += k < w ? t : 0;
    tot += t;
  }
  if (tot <= cap)
    for (uint32_t r = bits; r != 0u; r &= r - 1u) queue[off++] = uint16_t(threadIdx.x + __builtin_ctz(r) * stride);
  __syncthreads();
  return tot <= cap ? tot : -1;
}

// (block, pose) of workgroup slot w: each XCD (w % 8) takes a contiguous block-major range,
// so the poses of one block run on one XCD and share its L2 copy of the block's slices
__device__ __forceinline__ int pou_bn(int w, int per_xcd) { return (w & 7) * per_xcd + (w >> 3); }

// Columns of a wave's slice patch in the PoU gathers: 2^kPatchLogC.  8 x 8 patches (the rows
// walk the block's mid axis, the columns its fast axis) touch fewer cache lines per load than
// the 2 x 32 patches of the parallel-beam kernel here: forward 0.51 -> 0.49 ms and the adjoint
// about 5 % faster at 55 blocks x 64 poses, P = 128 (4 x 16 and 16 x 4 measured the same).
#ifndef FVR_POU_PATCH_LOGC
#define FVR_POU_PATCH_LOGC 3
#endif
constexpr int kPatchLogC = FVR_POU_PATCH_LOGC;

// slice sample (row i, column jj) of the half plane, jj in [0, W) with W = 2^logW: every 64
// consecutive q form a (64 >> kPatchLogC)-row x 2^kPatchLogC-column patch (W >= 2^kPatchLogC),
// else row-major
__device__ __forceinline__ void pou_ij_w(int q, int logW, int& i, int& jj) {
  constexpr int C = 1 << kPatchLogC, R = 64 / C;
  if (logW >= kPatchLogC) {  // uniform
    const int w = q >> 6, lane = q & 63, sh = logW - kPatchLogC;
    i = R * (w >> sh) + (lane >> kPatchLogC);
    jj = C * (w & ((1 << sh) - 1)) + (lane & (C - 1));
  } else {
    i = q >> logW;
    jj = q & ((1 << logW) - 1);
  }
}

// position (block index space) of slice sample (i, j): c + (j - P/2) u + (i - P/2) v
__device__ __forceinline__ void pou_kpos(const PouSlice& g, float ux, float uy, float uz, float vx, float vy,
                                         float vz, int i, int j, float& px, float& py, float& pz) {
  const int hp = g.P / 2;
  const float kx = float(j - hp), ky = float(i - hp);
  px = fmaf(ky, vx, fmaf(kx, ux, g.c));
  py = fmaf(ky, vy, fmaf(kx, uy, g.c));
  pz = fmaf(ky, vz, fmaf(kx, uz, g.c));
}

// the four samples that are their own DFT mirror: i, j in {0, P/2} (bitwise operators: no
// branches)
__device__ __forceinline__ bool pou_self_mirror(int i, int j, int hp) {
  return ((i == 0) | (i == hp)) & ((j == 0) | (j == hp));
}

// Whether the DFT mirror of sample (i, j) at p must be gathered: the stored block spectrum
// of a real cube is Hermitian about the centre c = P/2 on the storage indices 1 .. P-1, so
// the trilinear sample at the reflection 2c - p is conj S(p) when no corner of p (hence of
// its reflection) is storage index 0 or P, i.e. floor(p) is not in [-1, 0] or [P-1, P] on
// any axis (corners further out read the zero padding on both sides): `edge`, from pk_issue.
// On row / column 0 (frequency -P/2) the DFT mirror is not the reflection, so it is always
// gathered.
__device__ __forceinline__ bool pou_mirror_explicit(int i, int j, bool edge) { return (i == 0) | (j == 0) | edge; }
// The same for the spatial gradient of the sample: at (or within float rounding of) a lattice
// plane the trilinear derivative is one-sided (the cell above), and the mirror's is the cell
// above the reflection, not the reflected cell, so those samples take the explicit mirror too
// (1e-4 is far above the rounding of the mirror's position, about 1e-5 at P = 128); f is the
// fraction p - floor(p) of an axis.
__device__ __forceinline__ bool pou_near_lattice(float f) { return !((f > 1e-4f) & (f < 1.f - 1e-4f)); }
template <int VT>
__device__ __forceinline__ bool pou_grad_mirror_explicit(int i, int j, bool edge, const PkSmp<VT>& s) {
  return pou_mirror_explicit(i, j, edge) | pou_near_lattice(s.tx) | pou_near_lattice(s.ty) | pou_near_lattice(s.tz);
}

// Real slab image of one (block, pose) through a Hermitian half spectrum.  The image is
// Re(fftshift(ifft2(ifftshift(S)))) = fftshift(ifft2(H)) with H(k) = (S(k) + conj S(-k)) / 2
// (indices mod P), a Hermitian array, so it is kept for the columns v in [0, P/2] only: each
// thread gathers a sample and its mirror and writes H into LDS (P x (P/2 + 1) complex); the
// columns take inverse FFTs of length P; each row of the result, Y(v), is Hermitian in v and
// its real transform comes from one complex FFT of length P/2 of
//   Z(k) = (Y(k) + conj Y(P/2 - k)) + i (Y(k) - conj Y(P/2 - k)) e^{2 pi i k / P},
// whose outputs z(n) = x(2n) + i x(2n + 1) are the even and odd image samples.  Half the
// LDS and half the transform work of the full complex 2D transform.
template <int VT>
__global__ __launch_bounds__(kPouHalfThreads, 4) void pou_slice_ifft_kernel(const char* __restrict__ Fp, PouSlice g,
                                                                         PouPk k, const float* __restrict__ u,
                                                                         const float* __restrict__ v, int per_xcd,
                                                                         float* __restrict__ imgs) {
  extern __shared__ float2 smem[];
  const int P = g.P, logP = g.logP, hp = P / 2, pitch = hp + 1;
  float2* a = smem;               // H, then Y, Z, z: P rows x (P/2 + 1) columns
  float2* tw = smem + P * pitch;  // exp(-2 pi i k / P), k < P/2
  int* scan = reinterpret_cast<int*>(tw + hp);                   // kPouMaxWaves
  uint16_t* queue = reinterpret_cast<uint16_t*>(scan + kPouMaxWaves);  // g.mcap entries
  const int bn = pou_bn(blockIdx.x, per_xcd);
  if (bn >= g.B * g.N) return;  // whole workgroup: padding slot of the XCD split
  const int b = bn / g.N, n = bn - b * g.N;
  make_twiddles(tw, P);
  const float ux = u[3 * n], uy = u[3 * n + 1], uz = u[3 * n + 2];
  const float vx = v[3 * n], vy = v[3 * n + 1], vz = v[3 * n + 2];
  const unsigned long long blk = reinterpret_cast<unsigned long long>(Fp) + (unsigned long long)b * k.block;
  const int stride = blockDim.x;
  // Sample (i, j) sits at c + (j - P/2) u + (i - P/2) v; its DFT mirror is ((P - i) mod P,
  // (P - j) mod P).  Pass 1 stores H = S(k) (Re S for the four self-mirrored samples); pass 2
  // completes H = (S(k) + conj S(mirror)) / 2 where the mirror is not conj S(k)
  // (pou_mirror_explicit): over the compacted queue of those items (a few hundred of the
  // P x P/2 + P at P = 128, mostly in bands along the box faces that every patch of a slice
  // column range crosses), or, when they overflow the queue, per thread from its bitmask.
  const int nq = P * hp, nitems = nq + P, logH = logP - 1;
  auto item = [&](int q, int& i, int& j, int& vcol) {  // stored entry q: columns v = jj < P/2, then v = P/2
    int ip, jj;
    pou_ij_w(q, logH, ip, jj);
    const bool main = q < nq;
    i = main ? ip : q - nq;
    j = main ? hp + jj : 0;
    vcol = main ? jj : hp;
  };
  auto sample = [&](const PkSmp<VT>& sm, float& re, float& im) {
    float cr[8], ci[8];
    pk_corners<VT>(sm, cr, ci);
    interp8<false>(cr, ci, sm.tx, sm.ty, sm.tz, re, im, nullptr, nullptr);
  };
  // (bit kFwdBatch * iteration + t of mirbits: item t of the thread's iteration needs its mirror;
  // at most 5 iterations of 4 items at P = 128: more samples in flight per thread measured slower)
  uint32_t mirbits = 0u;
  int it = 0;
  for (int q0 = threadIdx.x; q0 < nitems; q0 += kFwdBatch * stride, ++it) {
    PkSmp<VT> sm[kFwdBatch];
    int ii[kFwdBatch], jj[kFwdBatch], vc[kFwdBatch];
#pragma unroll
    for (int t = 0; t < kFwdBatch; ++t) {
      item(min(q0 + t * stride, nitems - 1), ii[t], jj[t], vc[t]);
      float px, py, pz;
      pou_kpos(g, ux, uy, uz, vx, vy, vz, ii[t], jj[t], px, py, pz);
      const bool edge = pk_issue<VT>(sm[t], blk, k, P, px, py, pz);
      const bool need = (q0 + t * stride < nitems) & !pou_self_mirror(ii[t], jj[t], hp) &
                        pou_mirror_explicit(ii[t], jj[t], edge);
      mirbits |= uint32_t(need) << (kFwdBatch * it + t);
    }
#pragma unroll
    for (int t = 0; t < kFwdBatch; ++t)
      if (q0 + t * stride < nitems) {
        float re, im;
        sample(sm[t], re, im);
        const bool self = pou_self_mirror(ii[t], jj[t], hp);
        a[((ii[t] + hp) & (P - 1)) * pitch + vc[t]] = make_float2(re, self ? 0.f : im);
      }
  }
  auto mirror_issue = [&](PkSmp<VT>& sm, int i, int j) {
    float px, py, pz;
    pou_kpos(g, ux, uy, uz, vx, vy, vz, (P - i) & (P - 1), (P - j) & (P - 1), px, py, pz);
    pk_issue<VT>(sm, blk, k, P, px, py, pz);
  };
  auto mirror_add = [&](const PkSmp<VT>& sm, int i, int vcol) {
    float re, im;
    sample(sm, re, im);
    float2& h = a[((i + hp) & (P - 1)) * pitch + vcol];
    h = make_float2(0.5f * (h.x + re), 0.5f * (h.y - im));
  };
  const int nmq = pou_mirror_queue(mirbits, stride, scan, queue, g.mcap);
  if (nmq >= 0) {
    for (int e0 = threadIdx.x; e0 < nmq; e0 += 2 * stride) {
      PkSmp<VT> sm[2];
      int ii[2], jj[2], vc[2];
#pragma unroll
      for (int t = 0; t < 2; ++t)
        if (e0 + t * stride < nmq) {
          item(queue[e0 + t * stride], ii[t], jj[t], vc[t]);
          mirror_issue(sm[t], ii[t], jj[t]);
        }
#pragma unroll
      for (int t = 0; t < 2; ++t)
        if (e0 + t * stride < nmq) mirror_add(sm[t], ii[t], vc[t]);
    }
  } else {
    it = 0;
    for (int q0 = threadIdx.x; q0 < nitems; q0 += kFwdBatch * stride, ++it) {
      const uint32_t bits = (mirbits >> (kFwdBatch * it)) & ((1u << kFwdBatch) - 1u);
      if (!__any(bits != 0u)) continue;  // no lane of the wave needs a mirror here
      PkSmp<VT> sm[kFwdBatch];
      int ii[kFwdBatch], jj[kFwdBatch], vc[kFwdBatch];
#pragma unroll
      for (int t = 0; t < kFwdBatch; ++t) {
        item(min(q0 + t * stride, nitems - 1), ii[t], jj[t], vc[t]);
        if ((bits >> t) & 1u) mirror_issue(sm[t], ii[t], jj[t]);
      }
#pragma unroll
      for (int t = 0; t < kFwdBatch; ++t)
        if ((bits >> t) & 1u) mirror_add(sm[t], ii[t], vc[t]);
    }
  }
  __syncthreads();
  // inverse FFTs of the P/2 + 1 columns (length P): Y at rows brev(m); at P = 128 the register
  // transforms, Y at rows pou_row128(m), column 0 holding Y(0) + i Y(P/2) (both real)
  const bool r128 = P == 128;
  if (r128)
    pou_cols128<+1, kPouHalfThreads>(a, tw, true);
  else
    fft_lines_dif4<+1, true>(a, pitch, P, logP, 1, pitch, tw, 1);
  // even / odd pre-pass per row: pairs (kk, P/2 - kk), kk in [0, P/4]
  const int qp = P >> 2, npair = P * (qp + 1);
  for (int idx = threadIdx.x; idx < npair; idx += stride) {
    const int r = idx / (qp + 1), kk = idx - r * (qp + 1);
    float2* row = a + r * pitch;
    float2 y0 = row[kk], y1 = row[hp - kk];
    if (r128 && kk == 0) {  // the packed pair
      y1 = make_float2(y0.y, 0.f);
      y0 = make_float2(y0.x, 0.f);
    }
    auto zk = [&](float2 ya, float2 yb, int e) {  // Z(e) from Y(e), Y(P/2 - e)
      const float er = ya.x + yb.x, ei = ya.y - yb.y;  // E = Y(e) + conj Y(P/2 - e)
      const float dr = ya.x - yb.x, di = ya.y + yb.y;  // Y(e) - conj Y(P/2 - e)
      const float2 w = tw[e];                          // exp(-2 pi i e / P); its conjugate is used
      const float orr = dr * w.x + di * w.y, oi = di * w.x - dr * w.y;
      return make_float2(er - oi, ei + orr);  // E + i O
    };
    row[kk] = zk(y0, y1, kk);
    if (kk != 0 && kk != hp - kk) row[hp - kk] = zk(y1, y0, hp - kk);
  }
  __syncthreads();
  // inverse FFTs of the rows (length P/2): z(n) = x(2n) + i x(2n + 1) at column brev(n)
  // (pou_col64(n) at P = 128)
  if (r128)
    pou_rows64<+1, kPouHalfThreads>(a, tw);
  else
    fft_lines_dif4<+1, false>(a, P, hp, logH, pitch, 1, tw, 2);
  const float inv = 1.f / float(P * P);
  float* out = imgs + size_t(bn) * P * P;
  for (int q = threadIdx.x; q < P * P; q += stride) {
    const int i = q >> logP, j = q & (P - 1);
    const int m = (i + hp) & (P - 1), nn = (j + hp) & (P - 1);  // fftshift
    const float2 z = r128 ? a[pou_row128(m) * pitch + pou_col64(nn >> 1)]
                          : a[brev(m, logP) * pitch + brev(nn >> 1, logH)];
    out[q] = ((nn & 1) ? z.y : z.x) * inv;
  }
}

struct PouWarp {
  int B, N, P, K;
  long long ray_stride;  // 0: rays shared by all poses, K: per pose
  float inv_ds, c_idx;
};

// image position of a detector ray: (d_b r - c) inv_ds + c_idx, with the roundings fixed
// (two FMAs) so that every kernel of the warp and of its adjoints forms the same coordinate
__device__ __forceinline__ float pou_img_pos(float db, float r, float c, float inv_ds, float c_idx) {
  return fmaf(fmaf(db, r, -c), inv_ds, c_idx);
}

// global-memory view of a uniform base pointer: loads at base + a 32-bit byte offset use the
// scalar base register and one offset register
typedef const char __attribute__((address_space(1)))* pou_gcptr;
__device__ __forceinline__ pou_gcptr pou_global(const void* p) {
  return reinterpret_cast<pou_gcptr>(reinterpret_cast<uintptr_t>(p));
}
__device__ __forceinline__ float pou_ld(pou_gcptr base, uint32_t byte_off) {
  return *reinterpret_cast<const float __attribute__((address_space(1)))*>(base + byte_off);
}
typedef float pou_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 pou_ld4(pou_gcptr base, uint32_t byte_off) {
  const pou_f4 v = *reinterpret_cast<const pou_f4 __attribute__((address_space(1)))*>(base + byte_off);
  return make_float4(v.x, v.y, v.z, v.w);
}

// the bilinear lerps of the warp with their roundings fixed (the x-lerps, then the y-lerp), so that
// the per-pixel and the tiled forward kernels give the same value bit for bit
__device__ __forceinline__ float pou_lerp(float lo, float hi, float t) { return fmaf(hi, t, lo * (1.f - t)); }
__device__ __forceinline__ float pou_bilerp(float v00, float v01, float v10, float v11, float tx, float ty) {
  return pou_lerp(pou_lerp(v00, v01, tx), pou_lerp(v10, v11, tx), ty);
}

__device__ __forceinline__ float img_at(const float* img, int P, int x, int y) {
  return ((unsigned)x < (unsigned)P && (unsigned)y < (unsigned)P) ? img[y * P + x] : 0.f;
}

// XCD-aware flat grid: each XCD takes a contiguous range of (pose, pixel-tile)s
__global__ __launch_bounds__(256) void pou_warp_fwd_kernel(const float* __restrict__ imgs,
                                                           const float* __restrict__ d_b,
                                                           const float* __restrict__ cu,
                                                           const float* __restrict__ cv,
                                                           const float* __restrict__ rx,
                                                           const float* __restrict__ ry,
                                                           const float* __restrict__ ic, PouWarp g, int qtiles,
                                                           int per_xcd, float* __restrict__ out) {
  const int t = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (t >= g.N * qtiles) return;
  const int n = t / qtiles, q = (t - n * qtiles) * 256 + threadIdx.x;
  if (q >= g.K) return;
  const size_t ro = size_t(n) * g.ray_stride + q;
  const float r_x = rx[ro], r_y = ry[ro];
  const size_t pp = size_t(g.P) * g.P;
  const unsigned P = unsigned(g.P), mask = P - 1u;
  const bool pow2 = (P & mask) == 0u;
  float acc = 0.f;
  for (int b = 0; b < g.B; ++b) {
    const int bn = b * g.N + n;
    const float px = pou_img_pos(d_b[bn], r_x, cu[bn], g.inv_ds, g.c_idx);
    const float py = pou_img_pos(d_b[bn], r_y, cv[bn], g.inv_ds, g.c_idx);
    const float x0f = floorf(px), y0f = floorf(py);
    const float tx = px - x0f, ty = py - y0f;
    const int x0 = int(x0f), y0 = int(y0f);
    float v00, v01, v10, v11;
    if (pow2) {
      // every corner address wrapped into the image (loads without a branch), the corners
      // outside it then counting 0: the same values as the guarded loads
      const unsigned ux = unsigned(x0), uy = unsigned(y0);
      const pou_gcptr img = pou_global(imgs + size_t(bn) * pp);
      const uint32_t ra = (uy & mask) * P, rb = ((uy + 1u) & mask) * P;
      const uint32_t xa = ux & mask, xb = (ux + 1u) & mask;
      const float f00 = pou_ld(img, (ra + xa) * 4u), f01 = pou_ld(img, (ra + xb) * 4u);
      const float f10 = pou_ld(img, (rb + xa) * 4u), f11 = pou_ld(img, (rb + xb) * 4u);
      const bool ix0 = ux < P, ix1 = ux + 1u < P, iy0 = uy < P, iy1 = uy + 1u < P;
      v00 = ix0 && iy0 ? f00 : 0.f;
      v01 = ix1 && iy0 ? f01 : 0.f;
      v10 = ix0 && iy1 ? f10 : 0.f;
      v11 = ix1 && iy1 ? f11 : 0.f;
    } else {
      const float* img = imgs + size_t(bn) * pp;
      v00 = img_at(img, g.P, x0, y0);
      v01 = img_at(img, g.P, x0 + 1, y0);
      v10 = img_at(img, g.P, x0, y0 + 1);
      v11 = img_at(img, g.P, x0 + 1, y0 + 1);
    }
    acc += pou_bilerp(v00, v01, v10, v11, tx, ty);
  }
  out[size_t(n) * g.K + q] = acc * ic[ro];
}

// Separable forward warp (rays shared by every pose, r_x per detector column and r_y per
// detector row, both monotone): one workgroup per (pose, 32 x 32 detector tile).  Per block the
// tile's pixels fall in a patch of the slab image spanning its first to its last column's / row's
// texels (at most 32 s + 2 texels a side for s texels per pixel); the patch is staged in LDS
// (zeros outside the image), the x-lerps are formed once per (patch row, tile column), then each
// pixel takes its y-lerp.  The image positions of the tile's rows and the patch of every block
// are formed once per workgroup (kWarpBlocks blocks at a time, one thread per (block, row) and
// per block) into LDS, not by each of the 32 threads of a row for every block.  The
// arithmetic per pixel is that of pou_warp_fwd_kernel (the same positions, the same lerps in the
// same order, out-of-image corners 0, blocks summed in order), so the projections are the same
// bit for bit; a (tile, block) whose patch exceeds kWarpPatch texels a side, or whose geometry is
// not finite, takes the per-pixel gather instead.
constexpr int kWarpTile = 32;
constexpr int kWarpPatch = 16;
#ifndef FVR_WARP_BLOCKS
#define FVR_WARP_BLOCKS 16
#endif
constexpr int kWarpBlocks = FVR_WARP_BLOCKS;  // blocks whose row positions are in LDS at a time

__global__ __launch_bounds__(256) void pou_warp_fwd_sep_kernel(const float* __restrict__ imgs,
                                                               const float* __restrict__ d_b,
                                                               const float* __restrict__ cu,
                                                               const float* __restrict__ cv,
                                                               const float* __restrict__ rx_col,
                                                               const float* __restrict__ ry_row,
                                                               const float* __restrict__ ic, PouWarp g, int H, int W,
                                                               int tiles_x, int ntile, int per_xcd,
                                                               float* __restrict__ out) {
  __shared__ float patch[kWarpPatch][kWarpPatch + 1];
  __shared__ float arow[kWarpPatch][kWarpTile + 1];
  __shared__ float2 rpos[kWarpBlocks][kWarpTile];  // (floor, fraction) of py per (block, tile row)
  __shared__ int3 tinfo[kWarpBlocks];               // the patch's first texel column, row; 0 skip, 1 patch, 2 per pixel
  const int t = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (t >= g.N * ntile) return;  // whole workgroup
  const int n = t / ntile, tile = t - n * ntile;
  const int tyi = tile / tiles_x, txi = tile - tyi * tiles_x;
  const int xs = txi * kWarpTile, ys = tyi * kWarpTile;
  const int xe = min(xs + kWarpTile, W) - 1, ye = min(ys + kWarpTile, H) - 1;  // the tile's last pixel
  const int tid = threadIdx.x, cx = tid & (kWarpTile - 1), rg = tid >> 5;     // column, row group
  const int x = xs + cx;
  const bool xok = x < W;
  const float rxv = rx_col[xok ? x : xe];
  bool yok[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) yok[k] = ys + rg + 8 * k < H;
  const float rxa = rx_col[xs], rxb = rx_col[xe], rya = ry_row[ys], ryb = ry_row[ye];
  const unsigned P = unsigned(g.P);
  const size_t pp = size_t(g.P) * g.P;
  const float lim = float(kWarpPatch - 1);
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int b0 = 0; b0 < g.B; b0 += kWarpBlocks) {
    const int nb = min(kWarpBlocks, g.B - b0);
    __syncthreads();  // the previous group's positions are no longer read
    for (int i = tid; i < nb * kWarpTile; i += 256) {
      const int bb = i >> 5, c = i & (kWarpTile - 1);
      const int bn = (b0 + bb) * g.N + n;
      const float py = pou_img_pos(d_b[bn], ry_row[min(ys + c, ye)], cv[bn], g.inv_ds, g.c_idx);
      const float fy = floorf(py);
      rpos[bb][c] = make_float2(fy, py - fy);
    }
    for (int bb = tid; bb < nb; bb += 256) {
      // the tile's texel columns [flo, fhi] and rows [glo, ghi] (monotone maps: the end pixels)
      const int bn = (b0 + bb) * g.N + n;
      const float db = d_b[bn], c_u = cu[bn], c_v = cv[bn];
      const float pa = pou_img_pos(db, rxa, c_u, g.inv_ds, g.c_idx), pb = pou_img_pos(db, rxb, c_u, g.inv_ds, g.c_idx);
      const float qa = pou_img_pos(db, rya, c_v, g.inv_ds, g.c_idx), qb = pou_img_pos(db, ryb, c_v, g.inv_ds, g.c_idx);
      const float flo = floorf(fminf(pa, pb)), fhi = floorf(fmaxf(pa, pb)) + 1.f;
      const float glo = floorf(fminf(qa, qb)), ghi = floorf(fmaxf(qa, qb)) + 1.f;
      const bool fit = fhi - flo <= lim && ghi - glo <= lim;  // false for NaN or huge positions
      // a patch outside the image adds exact zeros: skipped
      const bool out_img = fhi < 0.f || flo > float(P - 1) || ghi < 0.f || glo > float(P - 1);
      tinfo[bb] = make_int3(fit ? int(flo) : 0, fit ? int(glo) : 0, fit ? (out_img ? 0 : 1) : 2);
    }
    __syncthreads();
    for (int bb = 0; bb < nb; ++bb) {
      const int3 ti = tinfo[bb];  // the same for every thread
      if (ti.z == 0) continue;
      const int bn = (b0 + bb) * g.N + n;
      const float* img = imgs + size_t(bn) * pp;
      const float px = pou_img_pos(d_b[bn], rxv, cu[bn], g.inv_ds, g.c_idx);
      const float fxf = floorf(px), wx = px - fxf;
      if (ti.z == 1) {
        const int tx0 = ti.x, ty0 = ti.y;
        {
          const int pr = tid >> 4, pc = tid & (kWarpPatch - 1);  // 256 threads: one texel each
          const int ix = tx0 + pc, iy = ty0 + pr;
          patch[pr][pc] = (unsigned(ix) < P && unsigned(iy) < P) ? img[size_t(iy) * P + ix] : 0.f;
        }
        __syncthreads();
        const int c0 = int(fxf) - tx0;  // in [0, fhi - flo - 1]
#pragma unroll
        for (int r = rg; r < kWarpPatch; r += 8) arow[r][cx] = pou_lerp(patch[r][c0], patch[r][c0 + 1], wx);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float2 rp = rpos[bb][rg + 8 * k];
          const int r0 = int(rp.x) - ty0;  // in [0, ghi - glo - 1]
          acc[k] += pou_lerp(arow[r0][cx], arow[r0 + 1][cx], rp.y);
        }
      } else {
        const int x0 = int(fxf);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float2 rp = rpos[bb][rg + 8 * k];
          const int y0 = int(rp.x);
          const float v00 = img_at(img, g.P, x0, y0), v01 = img_at(img, g.P, x0 + 1, y0);
          const float v10 = img_at(img, g.P, x0, y0 + 1), v11 = img_at(img, g.P, x0 + 1, y0 + 1);
          acc[k] += pou_bilerp(v00, v01, v10, v11, wx, rp.y);
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (xok && yok[k]) {
      const size_t q = size_t(ys + rg + 8 * k) * W + x;
      out[size_t(n) * g.K + q] = acc[k] * ic[q];
    }
}

// fixed-order workgroup sum of NV per-thread values: an XOR butterfly inside every wave
// (lane 0's result is used), then the per-wave sums added in wave order by every thread;
// red holds NV x (waves per workgroup) floats
template <int NV>
__device__ __forceinline__ void block_sum(float (&v)[NV], float* red) {
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < NV; ++k) red[k * nw + wave] = v[k];
  __syncthreads();
  for (int k = 0; k < NV; ++k) {
    float t = 0.f;
    for (int w = 0; w < nw; ++w) t += red[k * nw + w];
    v[k] = t;
  }
  __syncthreads();
}

// warp adjoint of one (block, pose): geometry sums and (optionally) the image
// gradient as 64-bit fixed point, at index map(y, x) of `acc` (LDS or global)
#ifndef FVR_POU_SCAN_BATCH
#define FVR_POU_SCAN_BATCH 4
#endif
constexpr int kScanBatch = FVR_POU_SCAN_BATCH;  // detector pixels in flight per thread

// llrint(x) for |x| < 2^51 (round to nearest even): 1.5 * 2^52 + x, rounded in double, holds
// the integer two's complement in the bits above the constant's own bit pattern
__device__ __forceinline__ unsigned long long fix64(float x) {
  return (unsigned long long)__double_as_longlong(double(x) + 6755399441055744.0) - 0x4338000000000000ull;
}

// POW2: P is a power of two, so a corner's column / row is wrapped into the image with a mask
// (an out-of-range corner then reads some texel of the image and counts 0) instead of clamped
template <bool IMG, bool POW2>
__device__ __forceinline__ void warp_adjoint_scan(const float* img, const float* go, const float* rx,
                                                  const float* ry, const float* ic, const PouWarp& g, int n, float db,
                                                  float c_u, float c_v, float scale, unsigned long long* acc,
                                                  int acc_pitch, float (&s)[3]) {
  const int P = g.P, K = g.K, stride = blockDim.x;
  // a NaN scale marks a pose whose output gradient holds NaN or inf: no fixed-point scatter
  // (the conversion back multiplies by 1 / NaN, so that pose's image gradient is NaN)
  const bool scatter = IMG && scale == scale;
  const pou_gcptr imgb = pou_global(img);
  const pou_gcptr rxn = pou_global(rx + size_t(n) * g.ray_stride);
  const pou_gcptr ryn = pou_global(ry + size_t(n) * g.ray_stride);
  const pou_gcptr icn = pou_global(ic + size_t(n) * g.ray_stride);
  const pou_gcptr gon = pou_global(go + size_t(n) * K);
  const int mask = P - 1;
  for (int q0 = threadIdx.x; q0 < K; q0 += kScanBatch * stride) {
    // 1) the pixels' rays and output gradients
    float r_x[kScanBatch], r_y[kScanBatch], gg[kScanBatch];
#pragma unroll
    for (int t = 0; t < kScanBatch; ++t) {
      const uint32_t qb = uint32_t(min(q0 + t * stride, K - 1)) * 4u;
      r_x[t] = pou_ld(rxn, qb);
      r_y[t] = pou_ld(ryn, qb);
      gg[t] = q0 + t * stride < K ? pou_ld(gon, qb) * pou_ld(icn, qb) : 0.f;
    }
    // 2) their four image corners: addresses inside the image, out-of-range corners count 0
    float tx[kScanBatch], ty[kScanBatch], vv[kScanBatch][4];
    int xa[kScanBatch], xb[kScanBatch], ya[kScanBatch], yb[kScanBatch];
    bool in[kScanBatch][4];
#pragma unroll
    for (int t = 0; t < kScanBatch; ++t) {
      const float px = pou_img_pos(db, r_x[t], c_u, g.inv_ds, g.c_idx);
      const float py = pou_img_pos(db, r_y[t], c_v, g.inv_ds, g.c_idx);
      const float x0f = floorf(px), y0f = floorf(py);
      tx[t] = px - x0f;
      ty[t] = py - y0f;
      const int x0 = int(x0f), y0 = int(y0f);
      const bool ix0 = unsigned(x0) < unsigned(P), ix1 = unsigned(x0 + 1) < unsigned(P);
      const bool iy0 = unsigned(y0) < unsigned(P), iy1 = unsigned(y0 + 1) < unsigned(P);
      if (POW2) {
        xa[t] = x0 & mask;
        xb[t] = (x0 + 1) & mask;
        ya[t] = y0 & mask;
        yb[t] = (y0 + 1) & mask;
      } else {
        xa[t] = min(max(x0, 0), mask);
        xb[t] = min(max(x0 + 1, 0), mask);
        ya[t] = min(max(y0, 0), mask);
        yb[t] = min(max(y0 + 1, 0), mask);
      }
      in[t][0] = ix0 && iy0;
      in[t][1] = ix1 && iy0;
      in[t][2] = ix0 && iy1;
      in[t][3] = ix1 && iy1;
      const uint32_t ra = __umul24(ya[t], P), rb = __umul24(yb[t], P);
      const float f00 = pou_ld(imgb, (ra + xa[t]) * 4u), f01 = pou_ld(imgb, (ra + xb[t]) * 4u);
      const float f10 = pou_ld(imgb, (rb + xa[t]) * 4u), f11 = pou_ld(imgb, (rb + xb[t]) * 4u);
      vv[t][0] = in[t][0] ? f00 : 0.f;
      vv[t][1] = in[t][1] ? f01 : 0.f;
      vv[t][2] = in[t][2] ? f10 : 0.f;
      vv[t][3] = in[t][3] ? f11 : 0.f;
    }
    // 3) geometry sums and the image-gradient scatter (a zero added where a corner is out of range)
#pragma unroll
    for (int t = 0; t < kScanBatch; ++t) {
      const float ox = 1.f - tx[t], oy = 1.f - ty[t];
      const float v00 = vv[t][0], v01 = vv[t][1], v10 = vv[t][2], v11 = vv[t][3];
      const float dsx = ((v01 - v00) * oy + (v11 - v10) * ty[t]) * g.inv_ds;
      const float dsy = ((v10 - v00) * ox + (v11 - v01) * tx[t]) * g.inv_ds;
      s[0] = fmaf(gg[t], dsx * r_x[t] + dsy * r_y[t], s[0]);
      s[1] = fmaf(-gg[t], dsx, s[1]);
      s[2] = fmaf(-gg[t], dsy, s[2]);
      if (scatter && gg[t] != 0.f && (in[t][0] || in[t][1] || in[t][2] || in[t][3])) {
        const float gs = gg[t] * scale;  // a power of two: exact
        const float gox = gs * ox, gtx = gs * tx[t];
        const unsigned ra = __umul24(ya[t], acc_pitch), rb = __umul24(yb[t], acc_pitch);
        atomicAdd(acc + ra + xa[t], fix64(in[t][0] ? gox * oy : 0.f));
        atomicAdd(acc + ra + xb[t], fix64(in[t][1] ? gtx * oy : 0.f));
        atomicAdd(acc + rb + xa[t], fix64(in[t][2] ? gox * ty[t] : 0.f));
        atomicAdd(acc + rb + xb[t], fix64(in[t][3] ? gtx * ty[t] : 0.f));
      }
    }
  }
}

__global__ __launch_bounds__(kPouThreads) void pou_warp_bwd_kernel(
    const float* __restrict__ imgs, const float* __restrict__ go, const float* __restrict__ d_b,
    const float* __restrict__ cu, const float* __restrict__ cv, const float* __restrict__ rx,
    const float* __restrict__ ry, const float* __restrict__ ic, PouWarp g, const float* __restrict__ gscale,
    unsigned long long* __restrict__ gacc, float* __restrict__ part) {
  __shared__ float red[3 * kPouMaxWaves];
  const int bn = blockIdx.x;
  const int n = bn % g.N;
  const size_t pp = size_t(g.P) * g.P;
  float s[3] = {0.f, 0.f, 0.f};
  if (gacc)
    warp_adjoint_scan<true, false>(imgs + bn * pp, go, rx, ry, ic, g, n, d_b[bn], cu[bn], cv[bn], gscale[n],
                                   gacc + bn * pp, g.P, s);
  else
    warp_adjoint_scan<false, false>(imgs + bn * pp, go, rx, ry, ic, g, n, d_b[bn], cu[bn], cv[bn], 1.f, nullptr, g.P,
                                    s);
  block_sum<3>(s, red);
  if (threadIdx.x < 3) part[size_t(bn) * 3 + threadIdx.x] = s[threadIdx.x];
}

// int64 fixed point -> float, and reset of the accumulator to zero for the next call
__global__ void pou_fix_to_float_kernel(unsigned long long* __restrict__ acc, const float* __restrict__ gscale,
                                        int N, long long pp, long long total, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int n = int((i / pp) % N);
  out[i] = float(double((long long)acc[i]) / double(gscale[n]));
}

// Steps of the fused adjoint after the warp adjoint, shared by both warp adjoints: the image
// gradient g, ifftshifted and packed for a real-input transform in `a` (row y holds the P / 2
// complex values z(n) = g'(y, 2n) + i g'(y, 2n + 1) at pitch P / 2 + 1, g' = ifftshift(g)), is
// taken through the adjoint of fftshift(ifft2(ifftshift(.))).real, i.e.
// fftshift(fft2(ifftshift(g))) / P^2: forward FFTs of length P / 2 of the rows' z, then per row
//   X(k) = (Z(k) + conj Z(P/2 - k)) / 2 - (i / 2) e^{-2 pi i k / P} (Z(k) - conj Z(P/2 - k)),
// k in [0, P/2] (Z periodic in P / 2), the row transform of g for the half spectrum; then
// forward FFTs of length P of the P / 2 + 1 columns (the other half is conj X(-k)).  Every
// slice sample is then recomputed and its spatial gradient contracted with the slice
// gradient; the nine sums (the warp's three, d/du and d/dv) go to part[bn] in fixed order,
// each plus fl (0, or NaN to mark the pose's sums as undefined).  NT: threads of the workgroup.
template <int VT, int NT>
__device__ __forceinline__ void pou_adjoint_tail(float2* a, const float2* tw, float* red, uint16_t* queue,
                                                 const char* Fp,
                                                 const PouSlice& sg, const PouPk& k, int b, int n, int bn,
                                                 const float* __restrict__ u, const float* __restrict__ v,
                                                 const float (&s3)[3], float fl, float* __restrict__ part) {
  const int P = sg.P, logP = sg.logP, hp = P / 2, hpitch = hp + 1, logH = logP - 1;
  const bool r128 = P == 128;  // the register transforms: outputs at pou_col64 / pou_row128
  if (r128)
    pou_rows64<-1, NT>(a, tw);
  else
    fft_lines_dif4<-1, false>(a, P, hp, logH, hpitch, 1, tw, 2);
  {
    const int qp = P >> 2, npair = P * (qp + 1);
    constexpr int kPost = (128 * (128 / 4 + 1) + NT - 1) / NT;
    float2 xa[kPost], xb[kPost];
#pragma unroll
    for (int t = 0; t < kPost; ++t) {
      const int idx = threadIdx.x + t * NT;
      if (idx < npair) {
        const int r = idx / (qp + 1), kk = idx - r * (qp + 1);
        const float2* row = a + r * hpitch;
        const int c0 = r128 ? pou_col64(kk) : brev(kk, logH);
        const int c1 = r128 ? pou_col64((hp - kk) & (hp - 1)) : brev((hp - kk) & (hp - 1), logH);
        const float2 z0 = row[c0], z1 = row[c1];
        auto xk = [&](float2 za, float2 zb, float2 w) {  // X from Z(k), Z(P/2 - k), w = e^{-2 pi i k / P}
          const float er = 0.5f * (za.x + zb.x), ei = 0.5f * (za.y - zb.y);  // (Z + conj Z') / 2
          const float dr = 0.5f * (za.x - zb.x), di = 0.5f * (za.y + zb.y);  // (Z - conj Z') / 2
          const float pr = dr * w.x - di * w.y, pi = dr * w.y + di * w.x;    // w (Z - conj Z') / 2
          return make_float2(er + pi, ei - pr);                             // E - i p
        };
        xa[t] = xk(z0, z1, tw[kk]);
        xb[t] = kk == 0 ? xk(z0, z1, make_float2(-1.f, 0.f)) : xk(z1, z0, tw[hp - kk]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kPost; ++t) {
      const int idx = threadIdx.x + t * NT;
      if (idx < npair) {
        const int r = idx / (qp + 1), kk = idx - r * (qp + 1);
        float2* row = a + r * hpitch;
        row[kk] = xa[t];
        row[hp - kk] = xb[t];  // kk = 0: X(P/2); kk = P/4: the same entry as X(kk)
      }
    }
    __syncthreads();
  }
  if (r128) {
    // columns 0 and P/2 are real (the row transform's X(0), X(P/2) of real rows): one complex
    // transform of col0 + i col64, F = G0 + i G64, then G0 = (F(u) + conj F(-u)) / 2 and
    // G64 = -i (F(u) - conj F(-u)) / 2
    pou_cols128<-1, NT>(a, tw, true);
    const int u = threadIdx.x;
    float2 g0 = make_float2(0.f, 0.f), g64 = g0;
    if (u < 128) {
      const float2 f = a[pou_row128(u) * hpitch], fm = a[pou_row128((128 - u) & 127) * hpitch];
      g0 = make_float2(0.5f * (f.x + fm.x), 0.5f * (f.y - fm.y));
      g64 = make_float2(0.5f * (f.y + fm.y), -0.5f * (f.x - fm.x));
    }
    __syncthreads();
    if (u < 128) {
      a[pou_row128(u) * hpitch] = g0;
      a[pou_row128(u) * hpitch + 64] = g64;
    }
    __syncthreads();
  } else {
    fft_lines_dif4<-1, true>(a, hpitch, P, logP, 1, hpitch, tw, 1);
  }
  // every slice sample recomputed and contracted with its gradient.  The sum runs over one
  // representative of each DFT-mirror pair (the columns j >= P/2, then column 0; on the
  // self-mirrored columns j = P/2 and 0 the rows i > P/2 with i = 0, P/2): where the mirror is
  // the reflection of an interior sample its term equals the representative's (the slice
  // gradient G is Hermitian, the mirror sample conj S with the gradient -conj dS and the
  // frequencies negated), so the representative counts twice; otherwise
  // (pou_grad_mirror_explicit) the mirror is gathered in a second pass, over the compacted queue
  // of those representatives (sg.mcap entries at `queue`), or per thread from its bitmask when
  // they overflow it.
  const float ux = u[3 * n], uy = u[3 * n + 1], uz = u[3 * n + 2];
  const float vx = v[3 * n], vy = v[3 * n + 1], vz = v[3 * n + 2];
  const unsigned long long blk = reinterpret_cast<unsigned long long>(Fp) + (unsigned long long)b * k.block;
  const float inv = 1.f / float(P * P);
  float gu[3] = {0.f, 0.f, 0.f}, gv[3] = {0.f, 0.f, 0.f};
  const int nh = P * hp, nitems = nh + P;
  auto item = [&](int q, int& i, int& j) {  // (a branch: the select form makes the gather form spill)
    if (q < nh) {
      int jj;
      pou_ij_w(q, logH, i, jj);
      j = hp + jj;
    } else {
      i = q - nh;
      j = 0;
    }
  };
  // weight of a representative's own term: 0 (counted through its mirror), 1 (self-mirrored,
  // or its mirror gathered explicitly), 2 (the mirror's term equals its own)
  auto weight = [&](int i, int j, bool edge, const PkSmp<VT>& s) {
    const bool colself = (j == 0) | (j == hp);
    const bool one = pou_self_mirror(i, j, hp) | pou_grad_mirror_explicit<VT>(i, j, edge, s);
    return (colself & (i > 0) & (i < hp)) ? 0.f : (one ? 1.f : 2.f);
  };
  auto contract = [&](const PkSmp<VT>& sm, int i, int j, float w) {
    const float kx = float(j - hp), ky = float(i - hp);
    // G at frequency (uu, vv) (natural order): stored for vv <= P/2, else conj G(-uu, -vv)
    const int uu = (i + hp) & (P - 1), vv = (j + hp) & (P - 1);
    const bool lo = vv <= hp;
    const int gu_ = lo ? uu : (P - uu) & (P - 1);
    const float2 gs = a[(r128 ? pou_row128(gu_) : brev(gu_, logP)) * hpitch + (lo ? vv : P - vv)];
    const float gr = gs.x * inv * w, gi = (lo ? gs.y : -gs.y) * inv * w;
    float cr[8], ci[8], re, im, dre[3], dim[3];
    pk_corners<VT>(sm, cr, ci);
    interp8<true>(cr, ci, sm.tx, sm.ty, sm.tz, re, im, dre, dim);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float d = gr * dre[c] + gi * dim[c];
      gu[c] = fmaf(kx, d, gu[c]);
      gv[c] = fmaf(ky, d, gv[c]);
    }
  };
  // (bit kPouBatch * iteration + t of mirbits: the representative's mirror is gathered in the
  // second pass; at most 5 iterations of 4 items at P = 128 with 512 threads)
  uint32_t mirbits = 0u;
  int it = 0;
  for (int q0 = threadIdx.x; q0 < nitems; q0 += kPouBatch * NT, ++it) {
    PkSmp<VT> smp[kPouBatch];
    int ii[kPouBatch], jj[kPouBatch];
    float ww[kPouBatch];
#pragma unroll
    for (int t = 0; t < kPouBatch; ++t) {
      item(min(q0 + t * NT, nitems - 1), ii[t], jj[t]);
      float px, py, pz;
      pou_kpos(sg, ux, uy, uz, vx, vy, vz, ii[t], jj[t], px, py, pz);
      const bool edge = pk_issue<VT>(smp[t], blk, k, P, px, py, pz);
      ww[t] = q0 + t * NT < nitems ? weight(ii[t], jj[t], edge, smp[t]) : 0.f;
      mirbits |= uint32_t((ww[t] == 1.f) & !pou_self_mirror(ii[t], jj[t], hp)) << (kPouBatch * it + t);
    }
#pragma unroll
    for (int t = 0; t < kPouBatch; ++t)
      if (ww[t] != 0.f) contract(smp[t], ii[t], jj[t], ww[t]);
  }
  auto mirror_issue = [&](PkSmp<VT>& sm, int& i, int& j) {  // (i, j) <- the mirror's indices
    i = (P - i) & (P - 1);
    j = (P - j) & (P - 1);
    float px, py, pz;
    pou_kpos(sg, ux, uy, uz, vx, vy, vz, i, j, px, py, pz);
    pk_issue<VT>(sm, blk, k, P, px, py, pz);
  };
  const int nmq = pou_mirror_queue(mirbits, NT, reinterpret_cast<int*>(red), queue, sg.mcap);
  if (nmq >= 0) {
    for (int e0 = threadIdx.x; e0 < nmq; e0 += 2 * NT) {
      PkSmp<VT> smp[2];
      int ii[2], jj[2];
#pragma unroll
      for (int t = 0; t < 2; ++t)
        if (e0 + t * NT < nmq) {
          item(queue[e0 + t * NT], ii[t], jj[t]);
          mirror_issue(smp[t], ii[t], jj[t]);
        }
#pragma unroll
      for (int t = 0; t < 2; ++t)
        if (e0 + t * NT < nmq) contract(smp[t], ii[t], jj[t], 1.f);
    }
  } else {
    it = 0;
    for (int q0 = threadIdx.x; q0 < nitems; q0 += kPouBatch * NT, ++it) {
      const uint32_t bits = (mirbits >> (kPouBatch * it)) & ((1u << kPouBatch) - 1u);
      if (!__any(bits != 0u)) continue;  // no lane of the wave needs a mirror here
      PkSmp<VT> smp[kPouBatch];
      int ii[kPouBatch], jj[kPouBatch];
#pragma unroll
      for (int t = 0; t < kPouBatch; ++t) {
        item(min(q0 + t * NT, nitems - 1), ii[t], jj[t]);
        if ((bits >> t) & 1u) mirror_issue(smp[t], ii[t], jj[t]);
      }
#pragma unroll
      for (int t = 0; t < kPouBatch; ++t)
        if ((bits >> t) & 1u) contract(smp[t], ii[t], jj[t], 1.f);
    }
  }
  float all[9] = {s3[0], s3[1], s3[2], gu[0], gu[1], gu[2], gv[0], gv[1], gv[2]};
  __syncthreads();
  block_sum<9>(all, red);
  if (threadIdx.x < 9) part[size_t(bn) * 9 + threadIdx.x] = all[threadIdx.x] + fl;
}

// The fused adjoint for any ray arrays: the warp adjoint scattered into an LDS image of 64-bit
// fixed point (the whole P x (P + 1) image, one 1024-thread workgroup per CU), converted to
// float and packed for the shared steps above.
template <int VT>
__global__ __launch_bounds__(kPouSliceThreads) void pou_render_bwd_kernel(
    const char* __restrict__ Fp, PouSlice sg, PouPk k, const float* __restrict__ u, const float* __restrict__ v,
    const float* __restrict__ imgs, const float* __restrict__ go, const float* __restrict__ d_b,
    const float* __restrict__ cu, const float* __restrict__ cv, const float* __restrict__ rx,
    const float* __restrict__ ry, const float* __restrict__ ic, PouWarp g, const float* __restrict__ gscale,
    int per_xcd, float* __restrict__ part) {
  extern __shared__ float2 smem[];
  const int P = sg.P, logP = sg.logP, pitch = P + 1, hp = P / 2;
  float2* a = smem;  // P x pitch complex, first used as 64-bit fixed-point accumulators
  float2* tw = smem + P * pitch;
  float* red = reinterpret_cast<float*>(tw + P / 2);  // 9 * kPouMaxWaves floats
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(a);
  const int bn = pou_bn(blockIdx.x, per_xcd);
  if (bn >= sg.B * sg.N) return;  // whole workgroup: padding slot of the XCD split
  const int b = bn / sg.N, n = bn - b * sg.N;
  for (int i = threadIdx.x; i < P * pitch; i += blockDim.x) acc[i] = 0ull;
  make_twiddles(tw, P);
  __syncthreads();
  // 1) warp adjoint: geometry sums + image gradient in LDS fixed point
  float s3[3] = {0.f, 0.f, 0.f};
  const float scale = gscale[n];
  warp_adjoint_scan<true, true>(imgs + size_t(bn) * P * P, go, rx, ry, ic, g, n, d_b[bn], cu[bn], cv[bn], scale,
                                acc, pitch, s3);  // P: a power of two here
  __syncthreads();
  // 2) fixed point -> float, ifftshifted and repacked in place for the real-input transform
  // (every value is read into registers before any is written: the packed rows overlap the
  // 64-bit ones)
  const float inv_scale = 1.f / scale;
  const int hpitch = hp + 1, logH = logP - 1;
  constexpr int kConv = (128 * 128) / (2 * kPouSliceThreads);  // pairs per thread at P = 128
  float2 zv[kConv];
#pragma unroll
  for (int t = 0; t < kConv; ++t) {
    const int idx = threadIdx.x + t * blockDim.x;  // pair (y, n)
    if (idx < P * hp) {
      const int y = idx >> logH, n2 = idx & (hp - 1);
      const int src = ((y + hp) & (P - 1)) * pitch + ((2 * n2 + hp) & (P - 1));  // ifftshift
      zv[t] = make_float2(float(double((long long)acc[src])) * inv_scale,
                          float(double((long long)acc[src + 1])) * inv_scale);
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kConv; ++t) {
    const int idx = threadIdx.x + t * blockDim.x;
    if (idx < P * hp) a[(idx >> logH) * hpitch + (idx & (hp - 1))] = zv[t];
  }
  __syncthreads();
  // the queue: the upper half of the full-size image, unused by the half-spectrum steps
  pou_adjoint_tail<VT, kPouSliceThreads>(a, tw, red, reinterpret_cast<uint16_t*>(a + P * hpitch), Fp, sg, k, b, n,
                                         bn, u, v, s3, 0.f, part);
}

// ---------------------------------------------------------------------------
// The fused adjoint for separable rays (a zero-skew pinhole: r_x depends on the detector
// column only, r_y on the row only, rays shared by every pose).  The image position of pixel
// (y, x) is then (px(x), py(y)) and the bilinear warp weight factorises, so the warp adjoint
// is a gather: texel (ty, tx) of the image gradient is
//   G(ty, tx) = sum_y hat(py(y) - ty) A(y, tx),   A(y, tx) = sum_x hat(px(x) - tx) X(y, x),
// X = go * ic, over the pixels whose cell touches the texel (a rectangle of the detector for
// a monotone map), and the geometry sums of the scan factorise the same way:
//   dsx-sums from Bd(y, tx) = sum_x hat'(px - tx) X and Br(y, tx) = sum_x hat'(px - tx) r_x X,
//   dsy-sums from A with hat'(py - ty) = -1 (ty = floor py) / +1 (ty = floor py + 1).
// Every texel row ty is owned by one segment of the detector rows (the segment holding its
// first detector row; segments split the rows that reach the image evenly), and thread
// (tx, segment) walks its segment's rows, adding to the texels (fy, tx), (fy + 1, tx) it owns:
// every texel is written by one thread, in detector-row order -- no atomics, no fixed point,
// deterministic -- straight into the packed half-spectrum layout of the transform that
// follows (P x (P/2 + 1) complex, 66 KB at P = 128).
// Staged form (P >= 64, W a multiple of 4, W <= 512): each wave takes 64 consecutive texel
// columns from the first one the detector reaches (one or two column groups, 8 or 4 row
// segments), and per detector row loads the window of X its lanes read as one 16-byte load
// per lane into a per-wave LDS buffer, one row ahead of its use (the slab-image values the
// row needs are fetched with it), so that a row costs LDS latency, not a global round trip.
// Otherwise thread (tx, segment) reads its pixels from global memory directly.
// ---------------------------------------------------------------------------
constexpr int kPouSepThreads = 512;
constexpr int kPouSepWaves = kPouSepThreads / 64;
constexpr int kSepNX = 16;     // pixels per texel column held in registers (wider footprints: per-x path)
constexpr int kSepChunks = 2;  // 256-float chunks of a wave's row window in the staged form (W <= 512)
constexpr int kSepOrderPose = 1;  // launch option: workgroups in pose-major order

__host__ __device__ inline bool sep_staged(int P, int W) { return P >= 64 && W % 4 == 0 && W <= 256 * kSepChunks; }
// floats of one wave's staging buffer: its row window (256-float chunks) and kSepNX of zeros past it
__host__ __device__ inline int sep_stage_floats(int W) { return 256 * ((W + 255) / 256) + kSepNX; }

// entries of the segment table: thread (tid mod P, tid / P) segments, or up to one per wave
__host__ __device__ inline int sep_max_segments(int P) {
  return kPouSepThreads / P > kPouSepWaves ? kPouSepThreads / P : kPouSepWaves;
}

// bytes of the region after the packed image: the staged row windows during the gather, the
// explicit-mirror queue of the contraction afterwards
__host__ __device__ inline size_t sep_stage_bytes(int P, int W) {
  const size_t stage = sep_staged(P, W) ? size_t(kPouSepWaves) * sep_stage_floats(W) * sizeof(float) : 0;
  return stage > pou_queue_bytes(P) ? stage : pou_queue_bytes(P);
}

__host__ __device__ inline size_t sep_lds_bytes(int P, int W) {
  const int nseg = sep_max_segments(P);
  return size_t(P) * (P / 2 + 1) * sizeof(float2) + sep_stage_bytes(P, W) + (P / 2) * sizeof(float2) +
         9 * kPouMaxWaves * sizeof(float) + (5 * size_t(P) + 2 * nseg + 4) * sizeof(int);
}

template <int VT>
__global__ __launch_bounds__(kPouSepThreads, 4) void pou_render_bwd_sep_kernel(
    const char* __restrict__ Fp, PouSlice sg, PouPk k, const float* __restrict__ u, const float* __restrict__ v,
    const float* __restrict__ imgs, const float* __restrict__ X, const float* __restrict__ flag,
    const float* __restrict__ d_b, const float* __restrict__ cu, const float* __restrict__ cv,
    const float* __restrict__ rx_col, const float* __restrict__ ry_row, int H, int W, float inv_ds, float c_idx,
    int per_xcd, int opts, float* __restrict__ part) {
  constexpr int NT = kPouSepThreads;
  extern __shared__ float2 smem[];
  const int P = sg.P, logP = sg.logP, hp = P / 2, hpitch = hp + 1;
  const bool staged = sep_staged(P, W);
  float2* a = smem;                                     // the packed gradient image: P x hpitch
  float* xstage = reinterpret_cast<float*>(a + P * hpitch);  // staged: per-wave row windows; then the queue
  float2* tw = reinterpret_cast<float2*>(reinterpret_cast<char*>(xstage) + sep_stage_bytes(P, W));  // P/2
  float* red = reinterpret_cast<float*>(tw + hp);       // 9 * kPouMaxWaves
  int* xlo = reinterpret_cast<int*>(red + 9 * kPouMaxWaves);  // per texel column / row: first and
  int* xhi = xlo + P;                                   //   last pixel column / row touching it
  int* ylo = xhi + P;
  int* yhi = ylo + P;
  int* sown = yhi + P;                                  // segment owning texel row ty (-1: none)
  int* segy = sown + P;                                 // (first, last) detector row of each segment
  int* rng = segy + 2 * sep_max_segments(P);            // (first, last) row / texel column reached
  const int m = pou_bn(blockIdx.x, per_xcd);            // position in the launch order
  if (m >= sg.B * sg.N) return;  // whole workgroup: padding slot of the XCD split
  int b, n;
  if (opts & kSepOrderPose) {
    n = m / sg.B;
    b = m - n * sg.B;
  } else {
    b = m / sg.N;
    n = m - b * sg.N;
  }
  const int bn = b * sg.N + n;
  const int tid = threadIdx.x;
  const float db = d_b[bn], c_u = cu[bn], c_v = cv[bn];
  const float hi_lim = float(P) + 1.f;
  // floor and fraction of a row's image position (fy = -2: the row reaches no texel)
  auto rowpos = [&](int y, int& fy, float& wy, float& r) {
    r = ry_row[y];
    const float py = pou_img_pos(db, r, c_v, inv_ds, c_idx);
    fy = -2;
    wy = 0.f;
    if (py > -2.f && py < hi_lim) {  // false for NaN
      const float f = floorf(py);
      fy = int(f);
      wy = py - f;
    }
  };

  // 0) tables: the texel columns' / rows' pixel ranges, the segments, the owners
  for (int i = tid; i < P * hpitch; i += NT) a[i] = make_float2(0.f, 0.f);
  for (int t = tid; t < P; t += NT) {
    xlo[t] = ylo[t] = 0x7fffffff;
    xhi[t] = yhi[t] = -1;
  }
  for (int s = tid; s < sep_max_segments(P); s += NT) {
    segy[2 * s] = 0x7fffffff;
    segy[2 * s + 1] = -1;
  }
  if (tid < 2) {
    rng[2 * tid] = 0x7fffffff;
    rng[2 * tid + 1] = -1;
  }
  make_twiddles(tw, P);
  __syncthreads();
  for (int x = tid; x < W; x += NT) {
    const float px = pou_img_pos(db, rx_col[x], c_u, inv_ds, c_idx);
    if (px > -2.f && px < hi_lim) {
      const int fx = int(floorf(px));
      if (unsigned(fx) < unsigned(P)) {
        atomicMin(&xlo[fx], x);
        atomicMax(&xhi[fx], x);
      }
      if (unsigned(fx + 1) < unsigned(P)) {
        atomicMin(&xlo[fx + 1], x);
        atomicMax(&xhi[fx + 1], x);
      }
    }
  }
  for (int y = tid; y < H; y += NT) {
    int fy;
    float wy, r;
    rowpos(y, fy, wy, r);
    if (unsigned(fy) < unsigned(P)) {
      atomicMin(&ylo[fy], y);
      atomicMax(&yhi[fy], y);
    }
    if (unsigned(fy + 1) < unsigned(P)) {
      atomicMin(&ylo[fy + 1], y);
      atomicMax(&yhi[fy + 1], y);
    }
  }
  __syncthreads();
  for (int t = tid; t < P; t += NT) {
    if (ylo[t] <= yhi[t]) {
      atomicMin(&rng[0], ylo[t]);
      atomicMax(&rng[1], yhi[t]);
    }
    if (xlo[t] <= xhi[t]) {
      atomicMin(&rng[2], t);
      atomicMax(&rng[3], t);
    }
  }
  __syncthreads();
  const int tx0 = rng[2], tx1 = rng[3];
  // staged: column groups of 64 texel columns from tx0 (one or two at P <= 128), and the row
  // segments they leave to the waves; otherwise thread (tid mod P, tid / P)
  const int ncg = staged ? (tx1 >= tx0 ? (tx1 - tx0) / 64 + 1 : 0) : 1;
  const int nseg = staged ? kPouSepWaves / (ncg > 0 ? ncg : 1) : (NT >> logP);
  for (int t = tid; t < P; t += NT) {
    int s = -1;
    if (ylo[t] <= yhi[t]) {
      const long long span = (long long)rng[1] - rng[0] + 1;
      s = int(((long long)(ylo[t] - rng[0]) * nseg) / span);
      s = s < nseg - 1 ? s : nseg - 1;
      atomicMin(&segy[2 * s], ylo[t]);
      atomicMax(&segy[2 * s + 1], yhi[t]);
    }
    sown[t] = s;
  }
  __syncthreads();

  // 1) the gather
  const int wave = tid >> 6, lane = tid & 63;
  int tx, seg;
  bool active;
  if (staged) {
    const int cg = ncg > 0 ? wave % ncg : 0;
    seg = ncg > 0 ? wave / ncg : 0;
    tx = ncg > 0 ? tx0 + 64 * cg + lane : lane;  // (tx0 is a sentinel when no column is reached)
    active = ncg > 0 && tx <= tx1;               // tx1 < P
  } else {
    tx = tid & (P - 1);
    seg = tid >> logP;
    active = true;
  }
  const int xl = active ? xlo[tx] : 1, xh = active ? xhi[tx] : 0;
  const bool colok = xl <= xh;
  const int xb = colok ? xl : 0;
  const int nx = colok ? xh - xl + 1 : 0;
  int nxw = nx, wlo = colok ? xb : 0x7fffffff, whi = colok ? xb + nx - 1 : -1;  // all lanes active here
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    nxw = max(nxw, __shfl_xor(nxw, o));
    wlo = min(wlo, __shfl_xor(wlo, o));
    whi = max(whi, __shfl_xor(whi, o));
  }
  // the column weights of the lane's pixels (0 where a pixel's cell misses the texel)
  auto coef = [&](int x, float& c, float& d, float& e) {
    const float r = rx_col[x];
    const float px = pou_img_pos(db, r, c_u, inv_ds, c_idx);
    c = d = 0.f;
    if (px > -2.f && px < hi_lim) {
      const float f = floorf(px);
      const int fx = int(f);
      const float wx = px - f;
      if (fx == tx) {
        c = 1.f - wx;
        d = -1.f;
      } else if (fx == tx - 1) {
        c = wx;
        d = 1.f;
      }
    }
    e = d * r;
  };
  // the weights of the lane's first NXT pixels (0 past its footprint)
  auto coefs = [&](auto nxt, float* cc, float* dd, float* ee) {
    constexpr int NXT = decltype(nxt)::value;
#pragma unroll
    for (int q = 0; q < NXT; ++q) {
      cc[q] = dd[q] = ee[q] = 0.f;
      if (q < nx) coef(xb + q, cc[q], dd[q], ee[q]);
    }
  };
  const size_t K = size_t(H) * W;
  const float* Xn = X + size_t(n) * K;
  const float* img = imgs + size_t(bn) * P * P;
  float* gimg = reinterpret_cast<float*>(a);
  const int gcol = (tx + hp) & (P - 1);  // ifftshifted column: float index in the packed row
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
  // the texels (fy, tx), (fy + 1, tx) of detector row y that this thread owns, with the row's
  // column sums A, Bd, Br and the slab-image values I0, I1 at those texels
  auto accumulate = [&](int fy, float wy, float ry, float A, float Bd, float Br, float I0, float I1) {
    if (unsigned(fy) < unsigned(P) && sown[fy] == seg) {  // texel row fy: weight 1 - wy, d/dpy -1
      const float w0 = 1.f - wy;
      if (colok) gimg[2 * hpitch * ((fy + hp) & (P - 1)) + gcol] += w0 * A;
      s0 = fmaf(I0, fmaf(w0, Br, -ry * A), s0);
      s1 = fmaf(-I0 * w0, Bd, s1);
      s2 = fmaf(I0, A, s2);
    }
    if (unsigned(fy + 1) < unsigned(P) && sown[fy + 1] == seg) {  // texel row fy + 1: weight wy, d/dpy +1
      const float w1 = wy;
      if (colok) gimg[2 * hpitch * ((fy + 1 + hp) & (P - 1)) + gcol] += w1 * A;
      s0 = fmaf(I1, fmaf(w1, Br, ry * A), s0);
      s1 = fmaf(-I1 * w1, Bd, s1);
      s2 = fmaf(-I1, A, s2);
    }
  };
  const pou_gcptr imgg = pou_global(img);  // a uniform base and 32-bit offsets (no 64-bit lane pointers)
  auto img_at_row = [&](int ty) {
    return colok && unsigned(ty) < unsigned(P) ? pou_ld(imgg, uint32_t(ty * P + tx) * 4u) : 0.f;
  };
  const int x0a = wlo & ~3;
  const int nch = whi >= 0 ? (whi - x0a) / 256 + 1 : 0;
  if (staged && nch <= kSepChunks) {
    if (ncg > 0 && whi >= 0) {  // wave-uniform: every lane stages its part of the window
      const int yb = __builtin_amdgcn_readfirstlane(segy[2 * seg]);
      const int ye = __builtin_amdgcn_readfirstlane(segy[2 * seg + 1]);
      float* xw = xstage + wave * sep_stage_floats(W);
      // the buffer past this wave's window reads 0 (lanes whose footprint is narrower than the
      // unrolled count read there with weight 0)
      for (int i = 256 * nch + lane; i < sep_stage_floats(W); i += 64) xw[i] = 0.f;
      const int ib = colok ? xb - x0a : 0;  // the lane's first pixel in the window
      // a fetched row: its window, position and slab-image values; two of them alternate so that
      // a row's loads are consumed one row after they are issued, with no register copies
      struct Row {
        float4 win[kSepChunks];
        int fy;
        float wy, ry, I0, I1;
      };
      // row y's window and position, and its image values (those of `prev` when fy repeats)
      auto fetch = [&](Row& r, int y, const Row& prev) {
        const pou_gcptr rowp = pou_global(Xn + size_t(y) * W);  // uniform
#pragma unroll
        for (int c = 0; c < kSepChunks; ++c)
          if (c < nch) {
            const int x = x0a + 256 * c + 4 * lane;  // W % 4 == 0: a group is all in or all out
            r.win[c] = x < W ? pou_ld4(rowp, uint32_t(x) * 4u) : make_float4(0.f, 0.f, 0.f, 0.f);
          }
        rowpos(y, r.fy, r.wy, r.ry);
        if (r.fy == prev.fy) {
          r.I0 = prev.I0;
          r.I1 = prev.I1;
        } else {
          r.I0 = img_at_row(r.fy);
          r.I1 = img_at_row(r.fy + 1);
        }
      };
      // NXT: pixels per lane, unrolled (the lane's weights past its footprint are 0); 0: the
      // per-pixel path for footprints wider than kSepNX
      auto run = [&](auto nxt) {
        constexpr int NXT = decltype(nxt)::value;
        float cc[NXT > 0 ? NXT : 1], dd[NXT > 0 ? NXT : 1], ee[NXT > 0 ? NXT : 1];
        if constexpr (NXT > 0) coefs(nxt, cc, dd, ee);
        // the staged window of a fetched row, then its column sums and texel updates
        // the current run of rows with one fy (wave-uniform): its texel rows' sums acc0, acc1,
        // whether the wave's segment owns them, their slab-image values, and the geometry sums
        // t (d/dpx, d/dpy terms), d (d/dsx), sa (d/dsy) that the run's texel rows multiply
        int rf = -0x7fffffff;
        bool own0 = false, own1 = false;
        float rI0 = 0.f, rI1 = 0.f, acc0 = 0.f, acc1 = 0.f, t0 = 0.f, t1 = 0.f, d0 = 0.f, d1 = 0.f, sa = 0.f;
        auto flush = [&]() {
          if (own0) {  // texel row rf: weight 1 - wy, d/dpy -1
            if (colok) gimg[2 * hpitch * ((rf + hp) & (P - 1)) + gcol] += acc0;
            s0 = fmaf(rI0, t0, s0);
            s1 = fmaf(-rI0, d0, s1);
            s2 = fmaf(rI0, sa, s2);
          }
          if (own1) {  // texel row rf + 1: weight wy, d/dpy +1
            if (colok) gimg[2 * hpitch * ((rf + 1 + hp) & (P - 1)) + gcol] += acc1;
            s0 = fmaf(rI1, t1, s0);
            s1 = fmaf(-rI1, d1, s1);
            s2 = fmaf(-rI1, sa, s2);
          }
          acc0 = acc1 = t0 = t1 = d0 = d1 = sa = 0.f;
        };
        auto process = [&](const Row& r) {
#pragma unroll
          for (int c = 0; c < kSepChunks; ++c)
            if (c < nch) *reinterpret_cast<float4*>(xw + 256 * c + 4 * lane) = r.win[c];
          // (the wave's LDS operations complete in order: the reads below see this window)
          float A = 0.f, Bd = 0.f, Br = 0.f;
          if constexpr (NXT > 0) {
#pragma unroll
            for (int q = 0; q < NXT; ++q) {
              const float xv = xw[ib + q];
              A = fmaf(cc[q], xv, A);
              Bd = fmaf(dd[q], xv, Bd);
              Br = fmaf(ee[q], xv, Br);
            }
          } else {
            for (int q = 0; q < nx; ++q) {
              float c, d, e;
              coef(xb + q, c, d, e);
              const float xv = xw[ib + q];
              A = fmaf(c, xv, A);
              Bd = fmaf(d, xv, Bd);
              Br = fmaf(e, xv, Br);
            }
          }
          // the row's texel rows fy, fy + 1 are the wave's (one detector row per wave step):
          // consecutive rows of one fy form a run whose texel and geometry terms are summed in
          // registers and added once, at the run's end
          const int fy = __builtin_amdgcn_readfirstlane(r.fy);
          if (fy != rf) {
            flush();
            rf = fy;
            own0 = unsigned(fy) < unsigned(P) && __builtin_amdgcn_readfirstlane(sown[fy]) == seg;
            own1 = unsigned(fy + 1) < unsigned(P) && __builtin_amdgcn_readfirstlane(sown[fy + 1]) == seg;
            rI0 = r.I0;
            rI1 = r.I1;
          }
          const float w0 = 1.f - r.wy, w1 = r.wy;
          acc0 = fmaf(w0, A, acc0);
          acc1 = fmaf(w1, A, acc1);
          t0 = fmaf(w0, Br, fmaf(-r.ry, A, t0));
          t1 = fmaf(w1, Br, fmaf(r.ry, A, t1));
          d0 = fmaf(w0, Bd, d0);
          d1 = fmaf(w1, Bd, d1);
          sa += A;
        };
        Row ra, rb;
        rb.fy = -0x7fffffff;
        rb.I0 = rb.I1 = 0.f;
        if (yb <= ye) fetch(ra, yb, rb);
        for (int y = yb; y <= ye; y += 2) {
          if (y + 1 <= ye) fetch(rb, y + 1, ra);
          process(ra);
          if (y + 1 > ye) break;
          if (y + 2 <= ye) fetch(ra, y + 2, rb);
          process(rb);
        }
        flush();
      };
      if (nxw <= 8) run(std::integral_constant<int, 8>{});
      else if (nxw <= kSepNX) run(std::integral_constant<int, kSepNX>{});
      else run(std::integral_constant<int, 0>{});
    }
  } else if (active) {
    const int yb = segy[2 * seg], ye = segy[2 * seg + 1];
    int lastf = -0x7fffffff;
    float I0 = 0.f, I1 = 0.f;
    float cc[kSepNX], dd[kSepNX], ee[kSepNX];
    coefs(std::integral_constant<int, kSepNX>{}, cc, dd, ee);
    for (int y = yb; y <= ye; ++y) {
      int fy;
      float wy, ry;
      rowpos(y, fy, wy, ry);
      const float* xr = Xn + size_t(y) * W;
      float A = 0.f, Bd = 0.f, Br = 0.f;
      if (nxw <= kSepNX) {
#pragma unroll
        for (int q = 0; q < kSepNX; ++q)
          if (q < nxw) {
            const float xv = xr[min(xb + q, W - 1)];  // a pixel past the lane's range has weight 0
            A = fmaf(cc[q], xv, A);
            Bd = fmaf(dd[q], xv, Bd);
            Br = fmaf(ee[q], xv, Br);
          }
      } else {
        for (int q = 0; q < nx; ++q) {
          float c, d, e;
          coef(xb + q, c, d, e);
          const float xv = xr[xb + q];
          A = fmaf(c, xv, A);
          Bd = fmaf(d, xv, Bd);
          Br = fmaf(e, xv, Br);
        }
      }
      if (fy != lastf) {
        lastf = fy;
        I0 = img_at_row(fy);
        I1 = img_at_row(fy + 1);
      }
      accumulate(fy, wy, ry, A, Bd, Br, I0, I1);
    }
  }
  const float s3[3] = {s0 * inv_ds, s1 * inv_ds, s2 * inv_ds};
  __syncthreads();
  // a pose whose X holds a non-finite value (flag NaN) gets NaN sums
  pou_adjoint_tail<VT, NT>(a, tw, red, reinterpret_cast<uint16_t*>(xstage), Fp, sg, k, b, n, bn, u, v, s3, flag[n],
                           part);
}

// X[n][q] = go[n][q] ic[q] (ic per pose when per_pose_rays) and flag[n] = 0, or NaN when any
// X of pose n is not finite
__global__ __launch_bounds__(1024) void pou_gprep_kernel(const float* __restrict__ go, const float* __restrict__ ic,
                                                         long long ray_stride, int K, float* __restrict__ X,
                                                         float* __restrict__ flag) {
  __shared__ int bad;
  const int n = blockIdx.x;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const float* g = go + size_t(n) * K;
  const float* c = ic + size_t(n) * ray_stride;
  float* xo = X + size_t(n) * K;
  bool nf = false;
  auto one = [&](float x) {
    nf |= !(fabsf(x) <= 3.40282347e38f);
    return x;
  };
  // one workgroup streams a pose: 16-byte loads, four in flight per thread, when the rows allow
  const bool vec = (K % 4 == 0) && ((reinterpret_cast<uintptr_t>(g) | reinterpret_cast<uintptr_t>(c) |
                                     reinterpret_cast<uintptr_t>(xo)) % 16 == 0);
  int q0 = 0;
  if (vec) {
    const int K4 = K / 4;
    const float4* g4 = reinterpret_cast<const float4*>(g);
    const float4* c4 = reinterpret_cast<const float4*>(c);
    float4* x4 = reinterpret_cast<float4*>(xo);
    int q = threadIdx.x;
    for (; q + 3 * 1024 < K4; q += 4 * 1024) {
      float4 a[4], b[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        a[t] = g4[q + t * 1024];
        b[t] = c4[q + t * 1024];
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
        x4[q + t * 1024] = make_float4(one(a[t].x * b[t].x), one(a[t].y * b[t].y), one(a[t].z * b[t].z),
                                       one(a[t].w * b[t].w));
    }
    for (; q < K4; q += 1024) {
      const float4 a = g4[q], b = c4[q];
      x4[q] = make_float4(one(a.x * b.x), one(a.y * b.y), one(a.z * b.z), one(a.w * b.w));
    }
    q0 = K;
  }
  for (int q = q0 + threadIdx.x; q < K; q += 1024) xo[q] = one(g[q] * c[q]);
  if (__any(nf) && (threadIdx.x & 63) == 0) bad = 1;
  __syncthreads();
  if (threadIdx.x == 0) flag[n] = bad ? __builtin_nanf("") : 0.f;
}

// d out / d(rx, ry, inv_cos) for per-pose ray arrays (explicit-ray projections)
__global__ __launch_bounds__(256) void pou_ray_grads_kernel(const float* __restrict__ imgs,
                                                            const float* __restrict__ go,
                                                            const float* __restrict__ d_b,
                                                            const float* __restrict__ cu,
                                                            const float* __restrict__ cv,
                                                            const float* __restrict__ rx,
                                                            const float* __restrict__ ry,
                                                            const float* __restrict__ ic, PouWarp g,
                                                            float* __restrict__ grx, float* __restrict__ gry,
                                                            float* __restrict__ gic) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int n = blockIdx.y;
  if (q >= g.K) return;
  const size_t ro = size_t(n) * g.ray_stride + q;
  const float r_x = rx[ro], r_y = ry[ro], o = go[size_t(n) * g.K + q], icq = ic[ro];
  const size_t pp = size_t(g.P) * g.P;
  float acc = 0.f, sx = 0.f, sy = 0.f;
  for (int b = 0; b < g.B; ++b) {
    const int bn = b * g.N + n;
    const float db = d_b[bn];
    const float px = pou_img_pos(db, r_x, cu[bn], g.inv_ds, g.c_idx);
    const float py = pou_img_pos(db, r_y, cv[bn], g.inv_ds, g.c_idx);
    const float x0f = floorf(px), y0f = floorf(py);
    const float tx = px - x0f, ty = py - y0f, ox = 1.f - tx, oy = 1.f - ty;
    const int x0 = int(x0f), y0 = int(y0f);
    const float* img = imgs + size_t(bn) * pp;
    const float v00 = img_at(img, g.P, x0, y0), v01 = img_at(img, g.P, x0 + 1, y0);
    const float v10 = img_at(img, g.P, x0, y0 + 1), v11 = img_at(img, g.P, x0 + 1, y0 + 1);
    acc += pou_bilerp(v00, v01, v10, v11, tx, ty);
    sx = fmaf(((v01 - v00) * oy + (v11 - v10) * ty) * g.inv_ds, db, sx);
    sy = fmaf(((v10 - v00) * ox + (v11 - v01) * tx) * g.inv_ds, db, sy);
  }
  const size_t oq = size_t(n) * g.K + q;
  grx[oq] = o * icq * sx;
  gry[oq] = o * icq * sy;
  gic[oq] = o * acc;
}

// the fixed-point scale of a pose from its maximum |go * ic| (host and device: one function)
__host__ __device__ inline float pou_scale_of_max(float mx) {
  if (!(mx <= 3.40282347e38f)) return __builtin_nanf("");  // NaN or +inf in the output gradient
  if (!(mx > 0.f)) return 1.f;
  int e = 0;
  frexpf(mx, &e);
  const int k = 40 - e;
  return ldexpf(1.f, k < 126 ? k : 126);
}

// max of |x| that propagates NaN (fmaxf drops it): once NaN, the result stays NaN
__device__ __forceinline__ float nan_max(float m, float x) { return (x > m || x != x) ? x : m; }

// per pose: scale = 2^min(40 - e, 126) with max |go * ic| = m * 2^e (m in [0.5, 1)), 1 when all
// zero, NaN when any go * ic is NaN or infinite (a maximum is exact in any order: 1024 threads,
// four independent loads in flight each).  The cap keeps the scale a finite power of two for
// tiny maxima (below 2^-86 the fixed-point values simply use fewer than 40 bits).
__global__ __launch_bounds__(1024) void pou_gscale_kernel(const float* __restrict__ go, const float* __restrict__ ic,
                                                          long long ray_stride, int K, float* __restrict__ scale) {
  __shared__ float red[16];
  const int n = blockIdx.x;
  const float* g = go + size_t(n) * K;
  const float* c = ic + size_t(n) * ray_stride;
  float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;
  int q = threadIdx.x;
  for (; q + 3 * 1024 < K; q += 4 * 1024) {
    m0 = nan_max(m0, fabsf(g[q] * c[q]));
    m1 = nan_max(m1, fabsf(g[q + 1024] * c[q + 1024]));
    m2 = nan_max(m2, fabsf(g[q + 2048] * c[q + 2048]));
    m3 = nan_max(m3, fabsf(g[q + 3072] * c[q + 3072]));
  }
  for (; q < K; q += 1024) m0 = nan_max(m0, fabsf(g[q] * c[q]));
  float m = nan_max(nan_max(m0, m1), nan_max(m2, m3));
  for (int o = 32; o > 0; o >>= 1) m = nan_max(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float mx = red[0];
    for (int w = 1; w < 16; ++w) mx = nan_max(mx, red[w]);
    scale[n] = pou_scale_of_max(mx);
  }
}

__global__ void pou_sum_blocks_kernel(const float* __restrict__ part, int B, int N, int C, int c0, int nc,
                                      float* __restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * nc) return;
  const int n = idx / nc, c = idx - n * nc;
  double s = 0.0;
  for (int b = 0; b < B; ++b) s += double(part[(size_t(b) * N + n) * C + c0 + c]);
  out[idx] = float(s);
}

// Pose geometry of the lattice path (cone_beam_fvr.py:670-688 _forward_from_lattice,
// :507-521 _compute_block_offsets): pose n is M[n] = [R t; 0 1] (row-major 4 x 4), the
// camera-frame lattice lat = (src0, u0, v0, n0) (12 floats, world (z, y, x) components);
// src = R src0 + t, u = R u0, v = R v0, w = R n0 and, per block centre c_b,
//   d_b = (c_b - src) . w,  cu = (c_b - src) . u,  cv = (c_b - src) . v;
// the detector axes go to the slice kernels in (x, y, z) voxel order: u_xyz = (u2, u1, u0).
struct PoseGeom {
  float src[3], u[3], v[3], w[3];
};

__device__ __forceinline__ PoseGeom pose_geom(const float* __restrict__ M, const float* __restrict__ lat, int n) {
  const float* m = M + size_t(n) * 16;
  PoseGeom g;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float r0 = m[4 * i], r1 = m[4 * i + 1], r2 = m[4 * i + 2];
    g.src[i] = fmaf(r0, lat[0], fmaf(r1, lat[1], fmaf(r2, lat[2], m[4 * i + 3])));
    g.u[i] = fmaf(r0, lat[3], fmaf(r1, lat[4], r2 * lat[5]));
    g.v[i] = fmaf(r0, lat[6], fmaf(r1, lat[7], r2 * lat[8]));
    g.w[i] = fmaf(r0, lat[9], fmaf(r1, lat[10], r2 * lat[11]));
  }
  return g;
}

__global__ __launch_bounds__(256) void pou_pose_geom_kernel(const float* __restrict__ M, int N,
                                                            const float* __restrict__ lat,
                                                            const float* __restrict__ centers, int B,
                                                            float* __restrict__ u_xyz, float* __restrict__ v_xyz,
                                                            float* __restrict__ d_b, float* __restrict__ cu,
                                                            float* __restrict__ cv) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= B * N) return;
  const int b = idx / N, n = idx - b * N;
  const PoseGeom g = pose_geom(M, lat, n);
  const float d0 = centers[3 * b] - g.src[0], d1 = centers[3 * b + 1] - g.src[1], d2 = centers[3 * b + 2] - g.src[2];
  d_b[idx] = fmaf(d0, g.w[0], fmaf(d1, g.w[1], d2 * g.w[2]));
  cu[idx] = fmaf(d0, g.u[0], fmaf(d1, g.u[1], d2 * g.u[2]));
  cv[idx] = fmaf(d0, g.v[0], fmaf(d1, g.v[1], d2 * g.v[2]));
  if (b == 0) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      u_xyz[3 * n + i] = g.u[2 - i];
      v_xyz[3 * n + i] = g.v[2 - i];
    }
  }
}

// d(sum go * render)/dM from the fused adjoint's per-(block, pose) sums part[b][n][9]
// (d/d(d_b, cu, cv), d/du_xyz, d/dv_xyz), summed over the blocks in float64:
//   g_src = -sum_b (g_db w + g_cu u + g_cv v),  g_w = sum_b g_db (c_b - src),
//   g_u = flip(sum_b d/du_xyz) + sum_b g_cu (c_b - src),  g_v likewise,
//   dM[i][j] = g_src[i] src0[j] + g_u[i] u0[j] + g_v[i] v0[j] + g_w[i] n0[j],  dM[i][3] = g_src[i],
// and row 3 zero.  One wave per pose: lane l takes the blocks l, l + 64, ... in order, then a
// fixed XOR butterfly sums the lanes (the same order every call: deterministic).
__global__ __launch_bounds__(64) void pou_pose_geom_bwd_kernel(const float* __restrict__ M, int N,
                                                               const float* __restrict__ lat,
                                                               const float* __restrict__ centers, int B,
                                                               const float* __restrict__ part,
                                                               float* __restrict__ gM) {
  const int n = blockIdx.x, lane = threadIdx.x;
  const PoseGeom g = pose_geom(M, lat, n);
  double acc[12];  // g_src, g_w, g_u, g_v (3 each)
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = 0.0;
  for (int b = lane; b < B; b += 64) {
    const float* p = part + (size_t(b) * N + n) * 9;
    const double gdb = p[0], gcu = p[1], gcv = p[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double d = double(centers[3 * b + i]) - double(g.src[i]);
      acc[i] -= gdb * g.w[i] + gcu * g.u[i] + gcv * g.v[i];
      acc[3 + i] += gdb * d;
      acc[6 + i] += gcu * d + double(p[3 + 2 - i]);
      acc[9 + i] += gcv * d + double(p[6 + 2 - i]);
    }
  }
#pragma unroll
  for (int k = 0; k < 12; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o);
  float val = 0.f;  // entry `lane` of dM (row 3: 0)
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (lane == 4 * i + j)
        val = j == 3 ? float(acc[i])
                     : float(acc[i] * lat[j] + acc[6 + i] * lat[3 + j] + acc[9 + i] * lat[6 + j] +
                             acc[3 + i] * lat[9 + j]);
  if (lane < 16) gM[size_t(n) * 16 + lane] = val;
}

// planar (B, 2, P, P, P) -> gather layout; one thread per stored element (margins written as 0)
template <int VT>
__global__ __launch_bounds__(256) void pou_pack_kernel(const void* __restrict__ F, int P, PouPk k,
                                                       char* __restrict__ out) {
  const int gx = blockIdx.x * 256 + threadIdx.x, gy = blockIdx.y;
  if (gx >= P + 2) return;
  const int b = blockIdx.z / (P + 3), gz = blockIdx.z - b * (P + 3);
  const int x = gx - 1, y = gy - 1, z = gz - 1;
  const bool in = unsigned(x) < unsigned(P) && unsigned(y) < unsigned(P) && unsigned(z) < unsigned(P);
  const size_t plane = size_t(P) * P * P;
  const size_t o = size_t(b) * 2 * plane + (size_t(in ? z : 0) * P + (in ? y : 0)) * P + (in ? x : 0);
  char* dst = out + (unsigned long long)b * k.block + size_t(gz) * k.plane + size_t(gy) * k.row + size_t(gx) * k.esz;
  if (VT == 0) {
    const float* f = reinterpret_cast<const float*>(F);
    *reinterpret_cast<float2*>(dst) = in ? make_float2(f[o], f[o + plane]) : make_float2(0.f, 0.f);
  } else {  // element (z, y, x): the entries of rows y and y + 1 (row P + 2 is zero)
    const uint16_t* h = reinterpret_cast<const uint16_t*>(F);
    const bool in1 = unsigned(x) < unsigned(P) && unsigned(y + 1) < unsigned(P) && unsigned(z) < unsigned(P);
    const size_t o1 = size_t(b) * 2 * plane + (size_t(in1 ? z : 0) * P + (in1 ? y + 1 : 0)) * P + (in1 ? x : 0);
    *reinterpret_cast<uint2*>(dst) = make_uint2(in ? uint32_t(h[o]) | (uint32_t(h[o + plane]) << 16) : 0u,
                                                in1 ? uint32_t(h[o1]) | (uint32_t(h[o1 + plane]) << 16) : 0u);
  }
}

static size_t slice_lds_bytes(int P) { return (size_t(P) * (P + 1) + P / 2) * sizeof(float2); }
// the forward kernel: half spectrum, twiddles, the scan's wave totals, the explicit-mirror queue
static size_t half_lds_bytes(int P) {
  return (size_t(P) * (P / 2 + 1) + P / 2) * sizeof(float2) + kPouMaxWaves * sizeof(int) + pou_queue_bytes(P);
}

// Whether the gathers compact their explicit mirrors into a queue (1, the default) or take the
// per-thread pass for all of them (0); fvr_pou_set_mirror_queue
static std::atomic<int> g_pou_mirror_queue{1};

static PouSlice pou_slice(int B, int N, int P) {
  return PouSlice{B, N, P, 31 - __builtin_clz(uint32_t(P)), float(P / 2),
                  g_pou_mirror_queue.load(std::memory_order_relaxed) ? pou_mirror_cap(P) : 0};
}

__global__ void pou_zero_u64_kernel(unsigned long long* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// dynamic LDS above the default limit has to be allowed per kernel
template <typename K>
static bool allow_lds(K kernel, size_t bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             int(bytes)) == hipSuccess;
}

}  // namespace fvr

using namespace fvr;

extern "C" {

int fvr_pou_set_mirror_queue(int32_t enabled) { return g_pou_mirror_queue.exchange(enabled ? 1 : 0); }

int fvr_pou_lds_ok(int32_t P) {
  if (P < 4 || P > (1 << kPouMaxLogP) || (P & (P - 1))) return 0;
  return slice_lds_bytes(P) + 9 * kPouMaxWaves * sizeof(float) <= 160 * 1024 ? 1 : 0;
}

int64_t fvr_pou_packed_bytes(int32_t dtype, int32_t B, int32_t P) {
  if (B <= 0 || P < 2 || dtype < 0 || dtype > 2) return -1;
  return int64_t(B) * int64_t(pou_pk(P, dtype).block);
}

int fvr_pou_pack_blocks(uint64_t F, int32_t dtype, int32_t B, int32_t P, uint64_t packed, void* stream) {
  FVR_CHECK_ARG(F && packed && B > 0 && P >= 2 && P <= 1024, "fvr_pou_pack_blocks: bad arguments");
  FVR_CHECK_ARG(int64_t(B) * (P + 3) <= 65535, "fvr_pou_pack_blocks: %d blocks of P = %d exceed the grid", B, P);
  FVR_CHECK_ARG(pou_pk(P, dtype).plane < (1u << 24), "fvr_pou_pack_blocks: P = %d too large for 24-bit offsets", P);
  const PouPk k = pou_pk(P, dtype);
  dim3 grid((P + 2 + 255) / 256, P + 2, B * (P + 3));
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const void* fp = reinterpret_cast<const void*>(F);
  char* out = reinterpret_cast<char*>(packed);
  switch (dtype) {
    case 0: hipLaunchKernelGGL(pou_pack_kernel<0>, grid, dim3(256), 0, st, fp, P, k, out); break;
    case 1: hipLaunchKernelGGL(pou_pack_kernel<1>, grid, dim3(256), 0, st, fp, P, k, out); break;
    case 2: hipLaunchKernelGGL(pou_pack_kernel<2>, grid, dim3(256), 0, st, fp, P, k, out); break;
    default: set_error("fvr_pou_pack_blocks: bad dtype %d", dtype); return FVR_ERR_INVALID;
  }
  FVR_CHECK_LAUNCH("pou_pack_kernel");
  return FVR_OK;
}

int fvr_pou_slice_ifft(uint64_t F, int32_t dtype, int32_t B, int32_t P, const float* u_xyz, const float* v_xyz,
                       int32_t N, float* imgs, void* stream) {
  FVR_CHECK_ARG(F && u_xyz && v_xyz && imgs && B > 0 && N > 0, "fvr_pou_slice_ifft: bad arguments");
  FVR_CHECK_ARG(fvr_pou_lds_ok(P), "fvr_pou_slice_ifft: P = %d must be a power of two in [4, 128]", P);
  FVR_CHECK_ARG(int64_t(B) * N < (int64_t(1) << 30), "fvr_pou_slice_ifft: too many (block, pose) pairs");
  const PouSlice g = pou_slice(B, N, P);
  const PouPk k = pou_pk(P, dtype);
  const size_t lds = half_lds_bytes(P);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const char* fp = reinterpret_cast<const char*>(F);
  const int per_xcd = int((int64_t(B) * N + 7) / 8);
#define FVR_SI(VT)                                                                                            \
  FVR_CHECK_ARG(allow_lds(pou_slice_ifft_kernel<VT>, lds), "fvr_pou_slice_ifft: %zu B of LDS refused", lds); \
  hipLaunchKernelGGL(pou_slice_ifft_kernel<VT>, dim3(8 * per_xcd), dim3(kPouHalfThreads), lds, st, fp, g, k, \
                     u_xyz, v_xyz, per_xcd, imgs)
  switch (dtype) {
    case 0: FVR_SI(0); break;
    case 1: FVR_SI(1); break;
    case 2: FVR_SI(2); break;
    default: set_error("fvr_pou_slice_ifft: bad dtype %d", dtype); return FVR_ERR_INVALID;
  }
#undef FVR_SI
  FVR_CHECK_LAUNCH("pou_slice_ifft_kernel");
  return FVR_OK;
}

static int warp_geom(PouWarp& g, int32_t B, int32_t N, int32_t P, int32_t K, int32_t per_pose_rays, float inv_ds,
                     float c_idx) {
  FVR_CHECK_ARG(B > 0 && N > 0 && P > 1 && K > 0, "pou warp: empty shape");
  FVR_CHECK_ARG(int64_t(B) * N * P * P < (int64_t(1) << 40), "pou warp: image stack too large");
  g = PouWarp{B, N, P, K, per_pose_rays ? (long long)K : 0ll, inv_ds, c_idx};
  return FVR_OK;
}

int fvr_pou_warp_fwd(const float* imgs, int32_t B, int32_t N, int32_t P, const float* d_b, const float* cu,
                     const float* cv, const float* rx, const float* ry, const float* ic, int32_t per_pose_rays,
                     int32_t K, float inv_ds, float c_idx, float* out, void* stream) {
  FVR_CHECK_ARG(imgs && d_b && cu && cv && rx && ry && ic && out, "fvr_pou_warp_fwd: null pointer");
  PouWarp g;
  if (warp_geom(g, B, N, P, K, per_pose_rays, inv_ds, c_idx) != FVR_OK) return FVR_ERR_INVALID;
  const int qtiles = (K + 255) / 256;
  const int64_t tiles = int64_t(N) * qtiles;
  FVR_CHECK_ARG(tiles < (int64_t(1) << 30), "fvr_pou_warp_fwd: too many pixel tiles");
  const int per_xcd = int((tiles + 7) / 8);
  hipLaunchKernelGGL(pou_warp_fwd_kernel, dim3(per_xcd * 8), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     imgs, d_b, cu, cv, rx, ry, ic, g, qtiles, per_xcd, out);
  FVR_CHECK_LAUNCH("pou_warp_fwd_kernel");
  return FVR_OK;
}

int fvr_pou_warp_fwd_sep(const float* imgs, int32_t B, int32_t N, int32_t P, const float* d_b, const float* cu,
                         const float* cv, const float* rx_col, const float* ry_row, const float* ic, int32_t H,
                         int32_t W, float inv_ds, float c_idx, float* out, void* stream) {
  FVR_CHECK_ARG(imgs && d_b && cu && cv && rx_col && ry_row && ic && out, "fvr_pou_warp_fwd_sep: null pointer");
  FVR_CHECK_ARG(H > 0 && W > 0 && int64_t(H) * W < (int64_t(1) << 31), "fvr_pou_warp_fwd_sep: bad detector");
  PouWarp g;
  if (warp_geom(g, B, N, P, H * W, 0, inv_ds, c_idx) != FVR_OK) return FVR_ERR_INVALID;
  const int tiles_x = (W + kWarpTile - 1) / kWarpTile, ntile = tiles_x * ((H + kWarpTile - 1) / kWarpTile);
  const int64_t total = int64_t(N) * ntile;
  FVR_CHECK_ARG(total < (int64_t(1) << 30), "fvr_pou_warp_fwd_sep: too many tiles");
  const int per_xcd = int((total + 7) / 8);
  hipLaunchKernelGGL(pou_warp_fwd_sep_kernel, dim3(per_xcd * 8), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     imgs, d_b, cu, cv, rx_col, ry_row, ic, g, H, W, tiles_x, ntile, per_xcd, out);
  FVR_CHECK_LAUNCH("pou_warp_fwd_sep_kernel");
  return FVR_OK;
}

int fvr_pou_gscale(const float* go, const float* ic, int32_t N, int32_t K, int32_t per_pose_rays, float* scale,
                   void* stream) {
  FVR_CHECK_ARG(go && ic && scale && N > 0 && K > 0, "fvr_pou_gscale: bad arguments");
  hipLaunchKernelGGL(pou_gscale_kernel, dim3(N), dim3(1024), 0, reinterpret_cast<hipStream_t>(stream), go, ic,
                     per_pose_rays ? (long long)K : 0ll, K, scale);
  FVR_CHECK_LAUNCH("pou_gscale_kernel");
  return FVR_OK;
}

float fvr_pou_scale_of_max(float max_abs) { return pou_scale_of_max(max_abs); }

int64_t fvr_pou_warp_bwd_workspace(int32_t B, int32_t N, int32_t P) { return int64_t(B) * N * P * P * 8; }

int fvr_pou_warp_bwd(const float* imgs, int32_t B, int32_t N, int32_t P, const float* go, const float* d_b,
                     const float* cu, const float* cv, const float* rx, const float* ry, const float* ic,
                     int32_t per_pose_rays, int32_t K, float inv_ds, float c_idx, const float* gscale,
                     void* workspace, float* grad_imgs, float* part, void* stream) {
  FVR_CHECK_ARG(imgs && go && d_b && cu && cv && rx && ry && ic && part, "fvr_pou_warp_bwd: null pointer");
  FVR_CHECK_ARG(!grad_imgs || (workspace && gscale), "fvr_pou_warp_bwd: image gradient needs workspace + scale");
  PouWarp g;
  if (warp_geom(g, B, N, P, K, per_pose_rays, inv_ds, c_idx) != FVR_OK) return FVR_ERR_INVALID;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned long long* acc = grad_imgs ? reinterpret_cast<unsigned long long*>(workspace) : nullptr;
  const int64_t total = int64_t(B) * N * P * P;
  if (acc) {  // a fill kernel, not a memset node, so that the call can be graph-captured
    hipLaunchKernelGGL(pou_zero_u64_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, acc,
                       (long long)total);
    FVR_CHECK_LAUNCH("pou_zero_u64_kernel");
  }
  hipLaunchKernelGGL(pou_warp_bwd_kernel, dim3(B * N), dim3(kPouThreads), 0, st, imgs, go, d_b, cu, cv, rx, ry, ic,
                     g, gscale, acc, part);
  FVR_CHECK_LAUNCH("pou_warp_bwd_kernel");
  if (acc) {
    hipLaunchKernelGGL(pou_fix_to_float_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, acc, gscale,
                       N, (long long)P * P, (long long)total, grad_imgs);
    FVR_CHECK_LAUNCH("pou_fix_to_float_kernel");
  }
  return FVR_OK;
}

int fvr_pou_render_bwd(uint64_t F, int32_t dtype, int32_t B, int32_t P, const float* u_xyz, const float* v_xyz,
                       int32_t N, const float* imgs, const float* go, const float* d_b, const float* cu,
                       const float* cv, const float* rx, const float* ry, const float* ic, int32_t per_pose_rays,
                       int32_t K, float inv_ds, float c_idx, const float* gscale, float* part, void* stream) {
  FVR_CHECK_ARG(F && u_xyz && v_xyz && imgs && go && d_b && cu && cv && rx && ry && ic && gscale && part,
                "fvr_pou_render_bwd: null pointer");
  FVR_CHECK_ARG(fvr_pou_lds_ok(P), "fvr_pou_render_bwd: P = %d must be a power of two in [4, 128]", P);
  PouWarp g;
  if (warp_geom(g, B, N, P, K, per_pose_rays, inv_ds, c_idx) != FVR_OK) return FVR_ERR_INVALID;
  FVR_CHECK_ARG(int64_t(B) * N < (int64_t(1) << 30), "fvr_pou_render_bwd: too many (block, pose) pairs");
  const PouSlice sg = pou_slice(B, N, P);
  const PouPk k = pou_pk(P, dtype);
  const size_t lds = slice_lds_bytes(P) + 9 * kPouMaxWaves * sizeof(float);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const char* fp = reinterpret_cast<const char*>(F);
  const int per_xcd = int((int64_t(B) * N + 7) / 8);
#define FVR_RB(VT)                                                                                            \
  FVR_CHECK_ARG(allow_lds(pou_render_bwd_kernel<VT>, lds), "fvr_pou_render_bwd: %zu B of LDS refused", lds);  \
  hipLaunchKernelGGL(pou_render_bwd_kernel<VT>, dim3(8 * per_xcd), dim3(kPouSliceThreads), lds, st, fp, sg, k, u_xyz, \
                     v_xyz, imgs, go, d_b, cu, cv, rx, ry, ic, g, gscale, per_xcd, part)
  switch (dtype) {
    case 0: FVR_RB(0); break;
    case 1: FVR_RB(1); break;
    case 2: FVR_RB(2); break;
    default: set_error("fvr_pou_render_bwd: bad dtype %d", dtype); return FVR_ERR_INVALID;
  }
#undef FVR_RB
  FVR_CHECK_LAUNCH("pou_render_bwd_kernel");
  return FVR_OK;
}

int fvr_pou_sep_ok(int32_t P, int32_t H, int32_t W) {
  if (!fvr_pou_lds_ok(P) || H < 1 || W < 1 || int64_t(H) * W >= (int64_t(1) << 31)) return 0;
  return sep_lds_bytes(P, W) <= 160 * 1024 ? 1 : 0;
}

int fvr_pou_gprep(const float* go, const float* ic, int32_t N, int32_t K, int32_t per_pose_rays, float* X,
                  float* flag, void* stream) {
  FVR_CHECK_ARG(go && ic && X && flag && N > 0 && K > 0, "fvr_pou_gprep: bad arguments");
  hipLaunchKernelGGL(pou_gprep_kernel, dim3(N), dim3(1024), 0, reinterpret_cast<hipStream_t>(stream), go, ic,
                     per_pose_rays ? (long long)K : 0ll, K, X, flag);
  FVR_CHECK_LAUNCH("pou_gprep_kernel");
  return FVR_OK;
}

int fvr_pou_render_bwd_sep(uint64_t F, int32_t dtype, int32_t B, int32_t P, const float* u_xyz, const float* v_xyz,
                           int32_t N, const float* imgs, const float* X, const float* flag, const float* d_b,
                           const float* cu, const float* cv, const float* rx_col, const float* ry_row, int32_t H,
                           int32_t W, float inv_ds, float c_idx, int32_t opts, float* part, void* stream) {
  FVR_CHECK_ARG(F && u_xyz && v_xyz && imgs && X && flag && d_b && cu && cv && rx_col && ry_row && part,
                "fvr_pou_render_bwd_sep: null pointer");
  FVR_CHECK_ARG((opts & ~kSepOrderPose) == 0, "fvr_pou_render_bwd_sep: bad options %d", opts);
  FVR_CHECK_ARG(B > 0 && N > 0 && int64_t(B) * N < (int64_t(1) << 30), "fvr_pou_render_bwd_sep: bad batch");
  FVR_CHECK_ARG(fvr_pou_sep_ok(P, H, W), "fvr_pou_render_bwd_sep: P = %d with a %d x %d detector does not fit", P, H,
                W);
  const PouSlice sg = pou_slice(B, N, P);
  const PouPk k = pou_pk(P, dtype);
  const size_t lds = sep_lds_bytes(P, W);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const char* fp = reinterpret_cast<const char*>(F);
  const int per_xcd = int((int64_t(B) * N + 7) / 8);
#define FVR_RS(VT)                                                                                                 \
  FVR_CHECK_ARG(allow_lds(pou_render_bwd_sep_kernel<VT>, lds), "fvr_pou_render_bwd_sep: %zu B of LDS refused", lds); \
  hipLaunchKernelGGL(pou_render_bwd_sep_kernel<VT>, dim3(8 * per_xcd), dim3(kPouSepThreads), lds, st, fp, sg, k,     \
                     u_xyz, v_xyz, imgs, X, flag, d_b, cu, cv, rx_col, ry_row, H, W, inv_ds, c_idx, per_xcd, opts, part)
  switch (dtype) {
    case 0: FVR_RS(0); break;
    case 1: FVR_RS(1); break;
    case 2: FVR_RS(2); break;
    default: set_error("fvr_pou_render_bwd_sep: bad dtype %d", dtype); return FVR_ERR_INVALID;
  }
#undef FVR_RS
  FVR_CHECK_LAUNCH("pou_render_bwd_sep_kernel");
  return FVR_OK;
}

int fvr_pou_ray_grads(const float* imgs, int32_t B, int32_t N, int32_t P, const float* go, const float* d_b,
                      const float* cu, const float* cv, const float* rx, const float* ry, const float* ic,
                      int32_t per_pose_rays, int32_t K, float inv_ds, float c_idx, float* grx, float* gry,
                      float* gic, void* stream) {
  FVR_CHECK_ARG(imgs && go && d_b && cu && cv && rx && ry && ic && grx && gry && gic, "fvr_pou_ray_grads: null");
  FVR_CHECK_ARG(N <= 65535, "fvr_pou_ray_grads: %d poses exceed the grid's y limit", N);
  PouWarp g;
  if (warp_geom(g, B, N, P, K, per_pose_rays, inv_ds, c_idx) != FVR_OK) return FVR_ERR_INVALID;
  hipLaunchKernelGGL(pou_ray_grads_kernel, dim3((K + 255) / 256, N), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), imgs, go, d_b, cu, cv, rx, ry, ic, g, grx, gry, gic);
  FVR_CHECK_LAUNCH("pou_ray_grads_kernel");
  return FVR_OK;
}

int fvr_pou_pose_geom(const float* M, int32_t N, const float* lat, const float* centers, int32_t B, float* u_xyz,
                      float* v_xyz, float* d_b, float* cu, float* cv, void* stream) {
  FVR_CHECK_ARG(M && lat && centers && u_xyz && v_xyz && d_b && cu && cv, "fvr_pou_pose_geom: null pointer");
  FVR_CHECK_ARG(N > 0 && B > 0 && int64_t(B) * N < (int64_t(1) << 30), "fvr_pou_pose_geom: bad shape");
  hipLaunchKernelGGL(pou_pose_geom_kernel, dim3(unsigned((int64_t(B) * N + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), M, N, lat, centers, B, u_xyz, v_xyz, d_b, cu, cv);
  FVR_CHECK_LAUNCH("pou_pose_geom_kernel");
  return FVR_OK;
}

int fvr_pou_pose_geom_bwd(const float* M, int32_t N, const float* lat, const float* centers, int32_t B,
                          const float* part, float* gM, void* stream) {
  FVR_CHECK_ARG(M && lat && centers && part && gM && N > 0 && B > 0, "fvr_pou_pose_geom_bwd: bad arguments");
  hipLaunchKernelGGL(pou_pose_geom_bwd_kernel, dim3(unsigned(N)), dim3(64), 0,
                     reinterpret_cast<hipStream_t>(stream), M, N, lat, centers, B, part, gM);
  FVR_CHECK_LAUNCH("pou_pose_geom_bwd_kernel");
  return FVR_OK;
}

int fvr_pou_sum_blocks(const float* part, int32_t B, int32_t N, int32_t C, int32_t c0, int32_t nc, float* out,
                       void* stream) {
  FVR_CHECK_ARG(part && out && B > 0 && N > 0 && nc > 0 && c0 >= 0 && c0 + nc <= C, "fvr_pou_sum_blocks: bad args");
  hipLaunchKernelGGL(pou_sum_blocks_kernel, dim3((N * nc + 255) / 256), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), part, B, N, C, c0, nc, out);
  FVR_CHECK_LAUNCH("pou_sum_blocks_kernel");
  return FVR_OK;
}

}  // extern "C"
