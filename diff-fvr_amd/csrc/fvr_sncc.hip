// Fused Fourier central-slice sampling + phase ramp + Spectral-NCC partial sums
// + pose-adjoint moments, over a batch of (pose, view) work items.
//
// One workgroup (4 waves) evaluates a tile of 64 slice columns x 4*ROWS rows of
// one item; each lane owns one column and walks ROWS rows.  Per sample:
//   p      = c + j*ustep + i*vstep            (32.32 fixed point, exact steps)
//   I, dI  = trilinear value and spatial gradient of the packed spectrum
//   W      = exp(-2 pi i turns(i,j)) * conj(T(i,j)) [* high-pass weight]
// and the tile accumulates (deterministic fixed-order LDS reduction):
//   A  = sum Re(I W)            E  = sum |I|^2 (hp^2)
//   Mc_im = sum cw(j) Im(I W)   Mr_im = sum rw(i) Im(I W)
//   Mc_g  = sum cw(j) Re(dI W)  Mr_g  = sum rw(i) Re(dI W)      (3 axes)
//   Mc_h  = sum cw(j) Re(dI I*) Mr_h  = sum rw(i) Re(dI I*)     (3 axes)
// from which the host forms NCC and its exact derivative
// w.r.t. the slice axes and the in-plane translations:
//   dNCC = Re<dS,T>/(|S||T|) - NCC Re<dS,S>/|S|^2.
// Reference behaviour reproduced: diff_fvr.py:106-150 (slice + ramp),
// diff_fvr.py:184-195 (NCC), gradients.py:27-74 (translation closed form);
// torch_engine.py:176-290 with FVR_ITEM_HP (half-bin grid, HP filter).
#include <cmath>

#include "fvr_common.h"

namespace fvr {

constexpr int kSnccWaves = 4;
constexpr int kSnccRows = 16;  // rows per wave
static_assert(kSnccWaves * kSnccRows == kSnccTileH, "tile height");
constexpr float kTwoM32 = 2.3283064365386963e-10f;  // 2^-32

typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
typedef uint32_t u2a4 __attribute__((ext_vector_type(2), aligned(4)));
// global-address-space views (lets the compiler use SGPR-base + 32-bit offset loads)
typedef const char __attribute__((address_space(1)))* gcptr;
typedef const float __attribute__((address_space(1)))* gfptr;
typedef float f2v __attribute__((ext_vector_type(2)));
typedef const f2v __attribute__((address_space(1)))* gf2ptr;

struct cpx {
  float r, i;
};

// trilinear value + gradient from 4 fast-axis pairs (already dequantised)
struct Pair {
  float r0, i0, r1, i1;
};

template <int DT>
struct Gather;

template <>
struct Gather<FVR_C32> {
  static constexpr uint32_t kElem = 8;
  __device__ __forceinline__ static void load(gcptr b0, gcptr b1, uint32_t off, uint32_t row, gfptr, int, int, int,
                                              int, int, int, Pair q[4]) {
    typedef const f4a8 __attribute__((address_space(1)))* gp;
    const uint32_t offr = off + row;
    f4a8 a = *reinterpret_cast<gp>(b0 + off);
    f4a8 b = *reinterpret_cast<gp>(b0 + offr);
    f4a8 c = *reinterpret_cast<gp>(b1 + off);
    f4a8 d = *reinterpret_cast<gp>(b1 + offr);
    q[0] = {a.x, a.y, a.z, a.w};
    q[1] = {b.x, b.y, b.z, b.w};
    q[2] = {c.x, c.y, c.z, c.w};
    q[3] = {d.x, d.y, d.z, d.w};
  }
};

template <int DT>
__device__ __forceinline__ float deq(uint32_t half) {
  return DT == FVR_C16F ? h2f(uint16_t(half)) : b2f(uint16_t(half));
}

template <int DT>
struct GatherQ {
  // S, M, F: storage indices of the low corner; Bm, Bf: brick counts
  static constexpr uint32_t kElem = 4;
  __device__ __forceinline__ static void load(gcptr b0, gcptr b1, uint32_t off, uint32_t row, gfptr scales, int S,
                                              int M, int F, int Bm, int Bf, int, Pair q[4]) {
    typedef const u2a4 __attribute__((address_space(1)))* gp;
    const uint32_t offr = off + row;
    u2a4 a = *reinterpret_cast<gp>(b0 + off);
    u2a4 b = *reinterpret_cast<gp>(b0 + offr);
    u2a4 c = *reinterpret_cast<gp>(b1 + off);
    u2a4 d = *reinterpret_cast<gp>(b1 + offr);
    const uint32_t bs0 = uint32_t(S) >> 3, bs1 = uint32_t(S + 1) >> 3;
    const uint32_t bm0 = uint32_t(M) >> 3, bm1 = uint32_t(M + 1) >> 3;
    const uint32_t bf0 = uint32_t(F) >> 3, bf1 = uint32_t(F + 1) >> 3;
    const uint32_t r00 = __umul24(__umul24(bs0, Bm) + bm0, Bf);
    const uint32_t r01 = __umul24(__umul24(bs0, Bm) + bm1, Bf);
    const uint32_t r10 = __umul24(__umul24(bs1, Bm) + bm0, Bf);
    const uint32_t r11 = __umul24(__umul24(bs1, Bm) + bm1, Bf);
    const float s000 = scales[r00 + bf0], s001 = scales[r00 + bf1];
    const float s010 = scales[r01 + bf0], s011 = scales[r01 + bf1];
    const float s100 = scales[r10 + bf0], s101 = scales[r10 + bf1];
    const float s110 = scales[r11 + bf0], s111 = scales[r11 + bf1];
    q[0] = {deq<DT>(a.x & 0xFFFFu) * s000, deq<DT>(a.x >> 16) * s000, deq<DT>(a.y & 0xFFFFu) * s001,
            deq<DT>(a.y >> 16) * s001};
    q[1] = {deq<DT>(b.x & 0xFFFFu) * s010, deq<DT>(b.x >> 16) * s010, deq<DT>(b.y & 0xFFFFu) * s011,
            deq<DT>(b.y >> 16) * s011};
    q[2] = {deq<DT>(c.x & 0xFFFFu) * s100, deq<DT>(c.x >> 16) * s100, deq<DT>(c.y & 0xFFFFu) * s101,
            deq<DT>(c.y >> 16) * s101};
    q[3] = {deq<DT>(d.x & 0xFFFFu) * s110, deq<DT>(d.x >> 16) * s110, deq<DT>(d.y & 0xFFFFu) * s111,
            deq<DT>(d.y >> 16) * s111};
  }
};

template <>
struct Gather<FVR_C16F> : GatherQ<FVR_C16F> {};
template <>
struct Gather<FVR_C16B> : GatherQ<FVR_C16B> {};

__device__ __forceinline__ float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// exp(-2 pi i t / 2^32) for a 32-bit turn count t
__device__ __forceinline__ cpx phasor(uint32_t t) {
  float x = float(int32_t(t)) * kTwoM32;  // turns in [-0.5, 0.5)
  float s, c;
  sincospif(2.f * x, &s, &c);
  return {c, -s};
}

template <int DT, bool ADJ, bool HP, bool WS>
__global__ __launch_bounds__(256) void sncc_kernel(const fvr_layout* __restrict__ layouts,
                                                   int n_layouts, const fvr_sncc_item* __restrict__ items,
                                                   int n_items, int n_tiles, int per_xcd,
                                                   float* __restrict__ partials) {
  // XCD-aware remap: workgroup b runs on XCD b % 8, so each XCD takes a
  // contiguous range of tiles (whole items) and keeps their lines in its L2.
  const int b = blockIdx.x;
  const int tile = (b & 7) * per_xcd + (b >> 3);
  if (tile >= n_tiles) return;

  int lo = 0, hi = n_items - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].tile_base <= tile) lo = mid;
    else hi = mid - 1;
  }
  const fvr_sncc_item& it = items[lo];
  if (uint32_t(it.layout) >= uint32_t(n_layouts)) {  // malformed descriptor: contribute nothing
    if (threadIdx.x < FVR_SNCC_NMOM) partials[size_t(tile) * FVR_SNCC_NMOM + threadIdx.x] = 0.f;
    return;
  }
  const fvr_layout& L = layouts[it.layout];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // item fields -> registers once (uniform)
  const int ih = it.h, iw = it.w, lo_int = it.lo_int;
  const float rw0 = it.rw0, rws = it.rws;
  const int lt = tile - it.tile_base;
  const int tyi = lt / it.tiles_x, txi = lt - tyi * it.tiles_x;
  const int j = txi * 64 + lane;
  const int i0 = tyi * kSnccTileH + wave * kSnccRows;
  const bool colok = uint32_t(j) < uint32_t(iw);  // unsigned: a malformed tile index reads nothing

  constexpr uint32_t esz = Gather<DT>::kElem;
  const uint32_t rowb = L.row_bytes, planeb = L.plane_bytes;
  const uint32_t lo_off = uint32_t(lo_int) * (planeb + rowb + esz);
  const gcptr b0 = reinterpret_cast<gcptr>(L.base) + lo_off;
  const gcptr b1 = b0 + planeb;
  const uint32_t guard = it.guard_off - lo_off;
  const gfptr scales = reinterpret_cast<gfptr>(L.scales);
  const int Bm = L.bricks[1], Bf = L.bricks[2];
  const gcptr Tb = reinterpret_cast<gcptr>(it.target);
  float2* __restrict__ Sout = reinterpret_cast<float2*>(it.slice_out);

  int64_t p0 = it.c[0] + int64_t(j) * it.ustep[0] + int64_t(i0) * it.vstep[0];
  int64_t p1 = it.c[1] + int64_t(j) * it.ustep[1] + int64_t(i0) * it.vstep[1];
  int64_t p2 = it.c[2] + int64_t(j) * it.ustep[2] + int64_t(i0) * it.vstep[2];
  const int64_t v0 = it.vstep[0], v1 = it.vstep[1], v2 = it.vstep[2];
  const uint64_t h0 = it.box_hi[0], h1 = it.box_hi[1], h2 = it.box_hi[2];

  const cpx colph = phasor(it.ph_col0 + uint32_t(j) * it.ph_col);
  const cpx rowph = phasor(it.ph_row0 + uint32_t(i0 + lane) * it.ph_row);  // lanes < kSnccRows used
  float hx = 0.f, hy = 0.f;
  if (HP) {
    const float x = it.hp_x0 + float(j) * it.hp_dx;
    const float y = it.hp_y0 + float(i0 + lane) * it.hp_dy;
    hx = __expf(-x * x * it.hp_inv2s2);
    hy = __expf(-y * y * it.hp_inv2s2);
  }
  const float cw = it.cw0 + float(j) * it.cws;
  int tidx = i0 * iw + j;  // row-major index of (i, j) in the target / slice
  const uint32_t tstep = uint32_t(iw) * 8u;
  uint32_t toff = uint32_t(tidx) * 8u;  // its byte offset (targets < 4 GiB)

  float A = 0.f, E = 0.f, sim = 0.f, mrim = 0.f;
  float gs[3] = {0.f, 0.f, 0.f}, gr[3] = {0.f, 0.f, 0.f};
  float hs[3] = {0.f, 0.f, 0.f}, hr[3] = {0.f, 0.f, 0.f};

#pragma unroll 4
  for (int k = 0; k < kSnccRows; ++k) {
    const int i = i0 + k;
    const bool ok = colok & (uint32_t(i) < uint32_t(ih));
    const bool inb = ok & (uint64_t(p0) <= h0) & (uint64_t(p1) <= h1) & (uint64_t(p2) <= h2);
    const uint32_t fs = inb ? uint32_t(uint64_t(p0) >> 32) : 0u;
    const uint32_t fm = inb ? uint32_t(uint64_t(p1) >> 32) : 0u;
    const uint32_t ff = inb ? uint32_t(uint64_t(p2) >> 32) : 0u;
    const float ts = float(uint32_t(p0)) * kTwoM32;
    const float tm = float(uint32_t(p1)) * kTwoM32;
    const float tf = float(uint32_t(p2)) * kTwoM32;
    const uint32_t offi = __umul24(fs, planeb) + __umul24(fm, rowb) + ff * esz;
    const uint32_t off = inb ? offi : guard;

    Pair q[4];
    Gather<DT>::load(b0, b1, off, rowb, scales, int(fs) + lo_int, int(fm) + lo_int, int(ff) + lo_int, Bm, Bf, 0,
                     q);

    // fast-axis lerps (d = V1 - V0, e = V0 + tf d) for the four (slow, mid) pairs
    const float d00r = q[0].r1 - q[0].r0, d00i = q[0].i1 - q[0].i0;
    const float d01r = q[1].r1 - q[1].r0, d01i = q[1].i1 - q[1].i0;
    const float d10r = q[2].r1 - q[2].r0, d10i = q[2].i1 - q[2].i0;
    const float d11r = q[3].r1 - q[3].r0, d11i = q[3].i1 - q[3].i0;
    const float e00r = fmaf(tf, d00r, q[0].r0), e00i = fmaf(tf, d00i, q[0].i0);
    const float e01r = fmaf(tf, d01r, q[1].r0), e01i = fmaf(tf, d01i, q[1].i0);
    const float e10r = fmaf(tf, d10r, q[2].r0), e10i = fmaf(tf, d10i, q[2].i0);
    const float e11r = fmaf(tf, d11r, q[3].r0), e11i = fmaf(tf, d11i, q[3].i0);
    // mid-axis
    const float y0r = e01r - e00r, y0i = e01i - e00i;
    const float y1r = e11r - e10r, y1i = e11i - e10i;
    const float f0r = fmaf(tm, y0r, e00r), f0i = fmaf(tm, y0i, e00i);
    const float f1r = fmaf(tm, y1r, e10r), f1i = fmaf(tm, y1i, e10i);
    // slow-axis
    const float zr = f1r - f0r, zi = f1i - f0i;
    const float Ir = fmaf(ts, zr, f0r), Ii = fmaf(ts, zi, f0i);

    // target and phase
    const f2v t = *reinterpret_cast<gf2ptr>(Tb + (ok ? toff : 0u));
    const float prr = readlane_f(rowph.r, k), pri = readlane_f(rowph.i, k);
    const float phr = colph.r * prr - colph.i * pri;
    const float phi = fmaf(colph.r, pri, colph.i * prr);
    if (WS && ok) Sout[tidx] = make_float2(Ir * phr - Ii * phi, fmaf(Ir, phi, Ii * phr));
    // W = phase * conj(T)
    float Wr = fmaf(phr, t.x, phi * t.y);
    float Wi = fmaf(phi, t.x, -phr * t.y);
    float Jr = Ir, Ji = Ii;  // the factor carried by E and the H moments
    if (HP) {
      const float hw = fmaf(-hx, readlane_f(hy, k), 1.f);
      Wr *= hw;
      Wi *= hw;
      const float h2 = hw * hw;
      Jr *= h2;
      Ji *= h2;
    }
    const float rw = rw0 + float(i) * rws;
    const float IWr = fmaf(Ir, Wr, -Ii * Wi);
    const float IWi = fmaf(Ir, Wi, Ii * Wr);
    A += IWr;
    sim += IWi;
    mrim = fmaf(rw, IWi, mrim);
    E = fmaf(Ir, Jr, fmaf(Ii, Ji, E));

    if (ADJ) {
      // gradient: slow = z, mid = y0 + ts (y1 - y0), fast = bilerp of d
      const float gmr = fmaf(ts, y1r - y0r, y0r), gmi = fmaf(ts, y1i - y0i, y0i);
      const float x0r = fmaf(tm, d01r - d00r, d00r), x0i = fmaf(tm, d01i - d00i, d00i);
      const float x1r = fmaf(tm, d11r - d10r, d10r), x1i = fmaf(tm, d11i - d10i, d10i);
      const float gfr = fmaf(ts, x1r - x0r, x0r), gfi = fmaf(ts, x1i - x0i, x0i);
      const float g0 = fmaf(zr, Wr, -zi * Wi), g1 = fmaf(gmr, Wr, -gmi * Wi), g2 = fmaf(gfr, Wr, -gfi * Wi);
      const float k0 = fmaf(zr, Jr, zi * Ji), k1 = fmaf(gmr, Jr, gmi * Ji), k2 = fmaf(gfr, Jr, gfi * Ji);
      gs[0] += g0;
      gs[1] += g1;
      gs[2] += g2;
      gr[0] = fmaf(rw, g0, gr[0]);
      gr[1] = fmaf(rw, g1, gr[1]);
      gr[2] = fmaf(rw, g2, gr[2]);
      hs[0] += k0;
      hs[1] += k1;
      hs[2] += k2;
      hr[0] = fmaf(rw, k0, hr[0]);
      hr[1] = fmaf(rw, k1, hr[1]);
      hr[2] = fmaf(rw, k2, hr[2]);
    }
    p0 += v0;
    p1 += v1;
    p2 += v2;
    tidx += iw;
    toff += tstep;
  }

  // ---- deterministic workgroup reduction of the 16 moments ----
  __shared__ float red[FVR_SNCC_NMOM][257];
  const int tid = threadIdx.x;
  red[0][tid] = A;
  red[1][tid] = E;
  red[2][tid] = cw * sim;
  red[3][tid] = mrim;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    red[4 + a][tid] = cw * gs[a];
    red[7 + a][tid] = gr[a];
    red[10 + a][tid] = cw * hs[a];
    red[13 + a][tid] = hr[a];
  }
  __syncthreads();
  const int qn = tid >> 4, s = tid & 15;
  float part = 0.f;
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) part += red[qn][s * 16 + jj];
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) part += __shfl_xor(part, o, 16);
  if (s == 0) partials[size_t(tile) * FVR_SNCC_NMOM + qn] = part;
}

// ---------------------------------------------------------------------------
// Per-item fixed-order float64 reduction of the tile partials.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void sncc_reduce_kernel(const fvr_sncc_item* __restrict__ items,
                                                          const float* __restrict__ partials,
                                                          double* __restrict__ out) {
  const fvr_sncc_item& it = items[blockIdx.x];
  const int ntile = it.tiles_x * ((it.h + kSnccTileH - 1) / kSnccTileH);
  double acc[FVR_SNCC_NMOM];
#pragma unroll
  for (int q = 0; q < FVR_SNCC_NMOM; ++q) acc[q] = 0.0;
  for (int t = threadIdx.x; t < ntile; t += 256) {
    const float* pp = partials + size_t(it.tile_base + t) * FVR_SNCC_NMOM;
#pragma unroll
    for (int q = 0; q < FVR_SNCC_NMOM; ++q) acc[q] += double(pp[q]);
  }
  __shared__ double red[256];
  for (int q = 0; q < FVR_SNCC_NMOM; ++q) {
    red[threadIdx.x] = acc[q];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (int(threadIdx.x) < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[size_t(blockIdx.x) * FVR_SNCC_NMOM + q] = red[0];
    __syncthreads();
  }
}

template <int DT, bool ADJ, bool HP, bool WS>
static void launch_sncc(dim3 g, hipStream_t st, const fvr_layout* lay, int n_lay, const fvr_sncc_item* items,
                        int n_items, int n_tiles, int per_xcd, float* partials) {
  hipLaunchKernelGGL((sncc_kernel<DT, ADJ, HP, WS>), g, dim3(256), 0, st, lay, n_lay, items, n_items, n_tiles,
                     per_xcd, partials);
}

template <int DT>
static bool dispatch_flags(uint32_t flags, dim3 g, hipStream_t st, const fvr_layout* lay, int n_lay,
                           const fvr_sncc_item* items, int n_items, int n_tiles, int per_xcd, float* partials) {
  const bool adj = flags & FVR_ITEM_ADJOINT, hp = flags & FVR_ITEM_HP, ws = flags & FVR_ITEM_WRITE_SLICE;
#define FVR_L(a, h, w)                                                                     \
  if (adj == a && hp == h && ws == w) {                                                    \
    launch_sncc<DT, a, h, w>(g, st, lay, n_lay, items, n_items, n_tiles, per_xcd, partials); \
    return true;                                                                           \
  }
  FVR_L(true, false, false)
  FVR_L(false, false, false)
  FVR_L(false, false, true)
  FVR_L(true, true, false)
  FVR_L(false, true, false)
  FVR_L(true, false, true)
  FVR_L(false, true, true)
  FVR_L(true, true, true)
#undef FVR_L
  return false;
}

}  // namespace fvr

using namespace fvr;

extern "C" {

int fvr_sncc_tile_rows(void) { return kSnccTileH; }

int fvr_sncc_eval_v(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* items_dev, int32_t n_items,
                    int32_t n_tiles, float* partials, uint32_t variant, void* stream) {
  FVR_CHECK_ARG(layouts_dev && items_dev && partials, "fvr_sncc_eval: null pointer");
  FVR_CHECK_ARG(n_items > 0 && n_tiles > 0 && n_layouts > 0, "fvr_sncc_eval: empty batch");
  FVR_CHECK_ARG((variant >> 4) < 8u, "fvr_sncc_eval: bad flags %u", variant >> 4);
  const int dt = int(variant & 0xF);
  FVR_CHECK_ARG(dt == FVR_C32 || dt == FVR_C16F || dt == FVR_C16B, "fvr_sncc_eval: bad dtype %d", dt);
  const int per_xcd = (n_tiles + 7) / 8;
  dim3 g(per_xcd * 8);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const uint32_t flags = variant >> 4;
  bool launched;
  if (dt == FVR_C32)
    launched = dispatch_flags<FVR_C32>(flags, g, st, layouts_dev, n_layouts, items_dev, n_items, n_tiles, per_xcd,
                                       partials);
  else if (dt == FVR_C16F)
    launched = dispatch_flags<FVR_C16F>(flags, g, st, layouts_dev, n_layouts, items_dev, n_items, n_tiles, per_xcd,
                                        partials);
  else
    launched = dispatch_flags<FVR_C16B>(flags, g, st, layouts_dev, n_layouts, items_dev, n_items, n_tiles, per_xcd,
                                        partials);
  FVR_CHECK_ARG(launched, "fvr_sncc_eval: no kernel for variant %u", variant);
  FVR_CHECK_LAUNCH("sncc_kernel");
  return FVR_OK;
}

int fvr_sncc_reduce(const fvr_sncc_item* items_dev, int32_t n_items, const float* partials, double* out,
                    void* stream) {
  FVR_CHECK_ARG(items_dev && partials && out && n_items > 0, "fvr_sncc_reduce: bad arguments");
  hipLaunchKernelGGL(sncc_reduce_kernel, dim3(n_items), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     items_dev, partials, out);
  FVR_CHECK_LAUNCH("sncc_reduce_kernel");
  return FVR_OK;
}

}  // extern "C"
