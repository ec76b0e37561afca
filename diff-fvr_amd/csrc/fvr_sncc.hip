// Fused Fourier central-slice sampling + phase ramp + Spectral-NCC partial sums
// + pose-adjoint moments, over a batch of (pose, view) work items.
//
// One workgroup (4 waves) evaluates a tile of 64 slice columns x 4*R rows of
// one item (R = 8, 16 or 32 rows per wave, chosen per launch); each lane owns
// one column and walks R rows.  Per sample:
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
#include <type_traits>

#include "fvr_common.h"

namespace fvr {

constexpr float kTwoM32 = 2.3283064365386963e-10f;  // 2^-32

typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
typedef uint32_t u2a4 __attribute__((ext_vector_type(2), aligned(4)));
// global-address-space views (lets the compiler use SGPR-base + 32-bit offset loads)
typedef const char __attribute__((address_space(1)))* gcptr;
typedef const float __attribute__((address_space(1)))* gfptr;
typedef float f2v __attribute__((ext_vector_type(2)));
typedef const f2v __attribute__((address_space(1)))* gf2ptr;
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const f4v __attribute__((address_space(1)))* gf4ptr;

struct cpx {
  float r, i;
};

// Raw loads of one sample: the four fast-axis corner pairs (slow, mid) =
// (0,0), (0,1), (1,0), (1,1).  They are issued one sample ahead of their use
// (see the row loop).  Scaled storage holds value / sc with one power-of-two
// scale sc per spectrum: samples are interpolated in units of sc.
template <int DT>
struct Raw;

template <>
struct Raw<FVR_C32> {
  static constexpr uint32_t kElem = 8;
  f4a8 q[4];
  // byte offset of the storage corner (S, M, F) from the layout base
  static __device__ __forceinline__ uint32_t offset(uint32_t S, uint32_t M, uint32_t F, uint32_t planeb,
                                                    uint32_t rowb) {
    return __umul24(S, planeb) + __umul24(M, rowb) + F * 8u;
  }
  __device__ __forceinline__ void load(gcptr b0, gcptr b1, uint32_t off, uint32_t row) {
    typedef const f4a8 __attribute__((address_space(1)))* gp;
    const uint32_t offr = off + row;
    q[0] = *reinterpret_cast<gp>(b0 + off);
    q[1] = *reinterpret_cast<gp>(b0 + offr);
    q[2] = *reinterpret_cast<gp>(b1 + off);
    q[3] = *reinterpret_cast<gp>(b1 + offr);
  }
  // corner pair k along the fast axis: the fast-axis difference d = high - low and the
  // value e = low + t d at fraction t
  __device__ __forceinline__ void lerp(int k, float t, float& er, float& ei, float& dr, float& di) const {
    dr = q[k].z - q[k].x;
    di = q[k].w - q[k].y;
    er = fmaf(t, dr, q[k].x);
    ei = fmaf(t, di, q[k].y);
  }
};

template <int DT>
__device__ __forceinline__ float deq(uint32_t half) {
  return DT == FVR_C16F ? h2f(uint16_t(half)) : b2f(uint16_t(half));
}

template <int DT>
struct RawQ {
  static constexpr uint32_t kElem = 4;
  u2a4 q[4];
  static __device__ __forceinline__ uint32_t offset(uint32_t S, uint32_t M, uint32_t F, uint32_t planeb,
                                                    uint32_t rowb) {
    return __umul24(S, planeb) + __umul24(M, rowb) + F * 4u;
  }
  __device__ __forceinline__ void load(gcptr b0, gcptr b1, uint32_t off, uint32_t row) {
    typedef const u2a4 __attribute__((address_space(1)))* gp;
    const uint32_t offr = off + row;
    q[0] = *reinterpret_cast<gp>(b0 + off);
    q[1] = *reinterpret_cast<gp>(b0 + offr);
    q[2] = *reinterpret_cast<gp>(b1 + off);
    q[3] = *reinterpret_cast<gp>(b1 + offr);
  }
  // corner pair k along the fast axis, in units of the scale: the difference
  // d = high - low and the value e = low + t d at fraction t.  FP16: four
  // v_fma_mix_f32 that read the f16 halves directly (the conversion inside the FMA is
  // exact, so d and e round as the float32 forms high - low and fma(t, d, low) do).
  __device__ __forceinline__ void lerp(int k, float t, float& er, float& ei, float& dr, float& di) const {
    if (DT == FVR_C16F) {
      asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,1]" : "=v"(dr) : "v"(q[k].y), "v"(q[k].x));
      asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel:[1,0,1] op_sel_hi:[1,0,1]" : "=v"(di) : "v"(q[k].y), "v"(q[k].x));
      asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,0,1]" : "=v"(er) : "v"(t), "v"(dr), "v"(q[k].x));
      asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "=v"(ei) : "v"(t), "v"(di), "v"(q[k].x));
    } else {
      const float ar = deq<DT>(q[k].x & 0xFFFFu), ai = deq<DT>(q[k].x >> 16);
      dr = deq<DT>(q[k].y & 0xFFFFu) - ar;
      di = deq<DT>(q[k].y >> 16) - ai;
      er = fmaf(t, dr, ar);
      ei = fmaf(t, di, ai);
    }
  }
};

template <>
struct Raw<FVR_C16F> : RawQ<FVR_C16F> {};
template <>
struct Raw<FVR_C16B> : RawQ<FVR_C16B> {};

// one sample in flight: its raw loads, target value and fractional coordinates
template <int DT>
struct Smp {
  Raw<DT> raw;
  f2v t;
  float ts, tm, tf;
  bool ok;
};

// the 16 moment accumulators of one lane
struct Acc {
  float A, E, sim, mrim;
  float gs[3], gr[3], hs[3], hr[3];
};

// the spectrum's scale sc (1 for complex64 storage); a scaled layout without a scale
// buffer is malformed: 0, so that its items contribute nothing
template <int DT>
__device__ __forceinline__ float layout_scale(const fvr_layout& L) {
  if constexpr (DT == FVR_C32) return 1.f;
  else return L.scales ? *reinterpret_cast<gfptr>(reinterpret_cast<gcptr>(L.scales)) : 0.f;
}

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

// deterministic workgroup reduction of the 16 moments into partials[tile][16];
// the column-weighted sums are scaled by cw (1 when accumulated per sample) and the
// tile sums, accumulated in units of the spectrum's scale sc, are brought to spectrum
// units: sc for the sums linear in the slice, sc^2 for E and the H moments.  sc is a
// power of two (1 for complex64 storage), so this equals scaling every sample.
__device__ __forceinline__ void tile_reduce(const Acc& ac, float cw, float sc, int tile,
                                            float* __restrict__ partials) {
  __shared__ float red[FVR_SNCC_NMOM][257];
  const int tid = threadIdx.x;
  red[0][tid] = ac.A;
  red[1][tid] = ac.E;
  red[2][tid] = cw * ac.sim;
  red[3][tid] = ac.mrim;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    red[4 + a][tid] = cw * ac.gs[a];
    red[7 + a][tid] = ac.gr[a];
    red[10 + a][tid] = cw * ac.hs[a];
    red[13 + a][tid] = ac.hr[a];
  }
  __syncthreads();
  const int qn = tid >> 4, s = tid & 15;
  float part = 0.f;
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) part += red[qn][s * 16 + jj];
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) part += __shfl_xor(part, o, 16);
  part *= (qn == 1 || qn >= 10) ? sc * sc : sc;
  if (s == 0) partials[size_t(tile) * FVR_SNCC_NMOM + qn] = part;
  __syncthreads();  // red is reused by the workgroup's next tile
}

// Launch options (both kernels): kOrderBiplanar -- no map; the items are (AP, LAT) pairs
// (item 2 n + v, every item of view v with the tile count of items[v]) and the slots walk
// them in target-major order (view, tile, pose), split contiguously over the XCDs, the
// order a host-built 'target' map gives such a batch.
constexpr int kOrderBiplanar = 2;

// Workgroup slot w -> (item, tile within the item, global tile index).  With a
// map the host chose the order (entries with an out-of-range item are idle
// slots); with kOrderBiplanar the order is computed as above; otherwise each XCD
// (slot % 8) takes a contiguous range of tiles and the item is found by binary
// search over the items' tile_base.
template <int R>
__device__ __forceinline__ bool resolve_tile(int w, const int2* __restrict__ map,
                                             const fvr_sncc_item* __restrict__ items, int n_items, int n_tiles,
                                             int n_slots, int opts, int& item, int& lt, int& tile) {
  if (map) {
    const int2 e = map[w];
    if (uint32_t(e.x) >= uint32_t(n_items)) return false;
    const fvr_sncc_item& it = items[e.x];
    const int nt = it.tiles_x * sncc_tiles_y(it.h, it.flags, 4 * R);
    tile = it.tile_base + e.y;
    if (uint32_t(e.y) >= uint32_t(nt) || uint32_t(tile) >= uint32_t(n_tiles)) return false;
    item = e.x;
    lt = e.y;
    return true;
  }
  const int g = (w & 7) * (n_slots >> 3) + (w >> 3);  // position in the order
  if (g >= n_tiles) return false;
  if (opts & kOrderBiplanar) {
    const int np = n_items >> 1;
    if (np < 1) return false;
    const int t0 = items[0].tiles_x * sncc_tiles_y(items[0].h, items[0].flags, 4 * R);
    int v = 0, r = g;
    if (r >= np * t0) {
      v = 1;
      r -= np * t0;
    }
    const int t = r / np, pose = r - t * np;
    const fvr_sncc_item& it = items[2 * pose + v];
    const int nt = it.tiles_x * sncc_tiles_y(it.h, it.flags, 4 * R);
    tile = it.tile_base + t;
    if (uint32_t(t) >= uint32_t(nt) || uint32_t(tile) >= uint32_t(n_tiles)) return false;
    item = 2 * pose + v;
    lt = t;
    return true;
  }
  tile = g;
  int lo = 0, hi = n_items - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].tile_base <= tile) lo = mid;
    else hi = mid - 1;
  }
  item = lo;
  lt = tile - items[lo].tile_base;  // a malformed tile_base gives rows / columns outside the item: read nothing
  return true;
}

template <int DT, bool ADJ, bool HP, bool WS, int R>
__device__ __forceinline__ void sncc_tile(const fvr_layout* __restrict__ layouts, int n_layouts,
                                          const fvr_sncc_item& it, int lt, int tile, float* __restrict__ partials) {
  static_assert(R % 2 == 0 && R >= 4 && R <= 64, "rows per wave: even, <= 64 (row phasors live in lanes)");
  if (uint32_t(it.layout) >= uint32_t(n_layouts) || it.tile_h != 4 * R) {  // malformed: contribute nothing
    if (threadIdx.x < FVR_SNCC_NMOM) partials[size_t(tile) * FVR_SNCC_NMOM + threadIdx.x] = 0.f;
    return;
  }
  const fvr_layout& L = layouts[it.layout];

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // item fields -> registers once (uniform)
  const int ih = it.h, iw = it.w, lo_int = it.lo_int;
  const float rw0 = it.rw0, rws = it.rws;
  const int tyi = lt / it.tiles_x, txi = lt - tyi * it.tiles_x;
  const int j = txi * 64 + lane;
  const int i0 = tyi * (4 * R) + wave * R;
  const bool colok = uint32_t(j) < uint32_t(iw);  // unsigned: a malformed tile index reads nothing

  const uint32_t rowb = L.row_bytes, planeb = L.plane_bytes;
  const gcptr b0 = reinterpret_cast<gcptr>(L.base);
  const gcptr b1 = b0 + planeb;
  const uint32_t guard = it.guard_off;
  const float gsc = layout_scale<DT>(L);
  const gcptr Tb = reinterpret_cast<gcptr>(it.target);
  float2* __restrict__ Sout = reinterpret_cast<float2*>(it.slice_out);

  // position of the next sample to issue (row i0 first), 32.32 fixed point
  int64_t p0 = it.c[0] + int64_t(j) * it.ustep[0] + int64_t(i0) * it.vstep[0];
  int64_t p1 = it.c[1] + int64_t(j) * it.ustep[1] + int64_t(i0) * it.vstep[1];
  int64_t p2 = it.c[2] + int64_t(j) * it.ustep[2] + int64_t(i0) * it.vstep[2];
  const int64_t v0 = it.vstep[0], v1 = it.vstep[1], v2 = it.vstep[2];
  const uint64_t h0 = it.box_hi[0], h1 = it.box_hi[1], h2 = it.box_hi[2];

  const cpx colph = phasor(it.ph_col0 + uint32_t(j) * it.ph_col);
  const cpx rowph = phasor(it.ph_row0 + uint32_t(i0 + lane) * it.ph_row);  // lanes < R used
  float hx = 0.f, hy = 0.f;
  if (HP) {
    const float x = it.hp_x0 + float(j) * it.hp_dx;
    const float y = it.hp_y0 + float(i0 + lane) * it.hp_dy;
    hx = __expf(-x * x * it.hp_inv2s2);
    hy = __expf(-y * y * it.hp_inv2s2);
  }
  const float cw = it.cw0 + float(j) * it.cws;
  const uint32_t tstep = uint32_t(iw) * 8u;
  uint32_t toff = uint32_t(i0 * iw + j) * 8u;  // byte offset of the next issued sample's target (< 4 GiB)

  // stage 1: box test, corner addresses and the sample's loads (target included)
  auto issue = [&](Smp<DT>& sm, int i) {
    const bool ok = colok & (uint32_t(i) < uint32_t(ih));
    const bool inb = ok & (uint64_t(p0) <= h0) & (uint64_t(p1) <= h1) & (uint64_t(p2) <= h2);
    const uint32_t fs = inb ? uint32_t(uint64_t(p0) >> 32) : 0u;
    const uint32_t fm = inb ? uint32_t(uint64_t(p1) >> 32) : 0u;
    const uint32_t ff = inb ? uint32_t(uint64_t(p2) >> 32) : 0u;
    sm.ts = float(uint32_t(p0)) * kTwoM32;
    sm.tm = float(uint32_t(p1)) * kTwoM32;
    sm.tf = float(uint32_t(p2)) * kTwoM32;
    const uint32_t S = fs + lo_int, M = fm + lo_int, F = ff + lo_int;
    sm.raw.load(b0, b1, inb ? Raw<DT>::offset(S, M, F, planeb, rowb) : guard, rowb);
    sm.t = *reinterpret_cast<gf2ptr>(Tb + (ok ? toff : 0u));
    sm.ok = ok;
    p0 += v0;
    p1 += v1;
    p2 += v2;
    toff += tstep;
  };

  Acc ac = {};
  // stage 2: interpolation, phase, target product and the moment updates of row i (k = i - i0)
  auto consume = [&](const Smp<DT>& sm, int k) {
    const int i = i0 + k;
    const float ts = sm.ts, tm = sm.tm, tf = sm.tf;
    const Raw<DT>& q = sm.raw;
    // fast-axis lerps (d = V1 - V0, e = V0 + tf d) for the four (slow, mid) pairs
    float e00r, e00i, e01r, e01i, e10r, e10i, e11r, e11i, d00r, d00i, d01r, d01i, d10r, d10i, d11r, d11i;
    q.lerp(0, tf, e00r, e00i, d00r, d00i);
    q.lerp(1, tf, e01r, e01i, d01r, d01i);
    q.lerp(2, tf, e10r, e10i, d10r, d10i);
    q.lerp(3, tf, e11r, e11i, d11r, d11i);
    // mid-axis
    const float y0r = e01r - e00r, y0i = e01i - e00i;
    const float y1r = e11r - e10r, y1i = e11i - e10i;
    const float f0r = fmaf(tm, y0r, e00r), f0i = fmaf(tm, y0i, e00i);
    const float f1r = fmaf(tm, y1r, e10r), f1i = fmaf(tm, y1i, e10i);
    // slow-axis
    const float zr = f1r - f0r, zi = f1i - f0i;
    const float Ir = fmaf(ts, zr, f0r), Ii = fmaf(ts, zi, f0i);

    // phase and target
    const f2v t = sm.t;
    const float prr = readlane_f(rowph.r, k), pri = readlane_f(rowph.i, k);
    const float phr = colph.r * prr - colph.i * pri;
    const float phi = fmaf(colph.r, pri, colph.i * prr);
    // I, dI above are in units of the spectrum's scale sc, applied to the sums at the end
    if (WS && sm.ok) Sout[i * iw + j] = make_float2(gsc * (Ir * phr - Ii * phi), gsc * fmaf(Ir, phi, Ii * phr));
    // W = phase * conj(T) (times the high-pass weight hw); J = I (times hw^2)
    float Wr = fmaf(phr, t.x, phi * t.y);
    float Wi = fmaf(phi, t.x, -phr * t.y);
    float Jr = Ir, Ji = Ii;
    if (HP) {
      const float hw = fmaf(-hx, readlane_f(hy, k), 1.f), hw2 = hw * hw;
      Wr *= hw;
      Wi *= hw;
      Jr *= hw2;
      Ji *= hw2;
    }
    const float rw = rw0 + float(i) * rws;
    const float IWr = fmaf(Ir, Wr, -Ii * Wi);
    const float IWi = fmaf(Ir, Wi, Ii * Wr);
    ac.A += IWr;
    ac.sim += IWi;
    ac.mrim = fmaf(rw, IWi, ac.mrim);
    ac.E = fmaf(Ir, Jr, fmaf(Ii, Ji, ac.E));

    if (ADJ) {
      // gradient: slow = z, mid = y0 + ts (y1 - y0), fast = bilerp of d
      const float gmr = fmaf(ts, y1r - y0r, y0r), gmi = fmaf(ts, y1i - y0i, y0i);
      const float x0r = fmaf(tm, d01r - d00r, d00r), x0i = fmaf(tm, d01i - d00i, d00i);
      const float x1r = fmaf(tm, d11r - d10r, d10r), x1i = fmaf(tm, d11i - d10i, d10i);
      const float gfr = fmaf(ts, x1r - x0r, x0r), gfi = fmaf(ts, x1i - x0i, x0i);
      const float g0 = fmaf(zr, Wr, -zi * Wi), g1 = fmaf(gmr, Wr, -gmi * Wi), g2 = fmaf(gfr, Wr, -gfi * Wi);
      const float k0 = fmaf(zr, Jr, zi * Ji), k1 = fmaf(gmr, Jr, gmi * Ji), k2 = fmaf(gfr, Jr, gfi * Ji);
      ac.gs[0] += g0;
      ac.gs[1] += g1;
      ac.gs[2] += g2;
      ac.gr[0] = fmaf(rw, g0, ac.gr[0]);
      ac.gr[1] = fmaf(rw, g1, ac.gr[1]);
      ac.gr[2] = fmaf(rw, g2, ac.gr[2]);
      ac.hs[0] += k0;
      ac.hs[1] += k1;
      ac.hs[2] += k2;
      ac.hr[0] = fmaf(rw, k0, ac.hr[0]);
      ac.hr[1] = fmaf(rw, k1, ac.hr[1]);
      ac.hr[2] = fmaf(rw, k2, ac.hr[2]);
    }
  };

  // two samples in flight per lane: the loads of row k+1 are issued before row k is consumed
  Smp<DT> sa, sb;
  issue(sa, i0);
  // (the scheduling barriers keep each issue after the consume that frees its
  // registers, so the loop carries no register copies)
#pragma unroll 1
  for (int k = 0; k < R - 2; k += 2) {
    issue(sb, i0 + k + 1);
    consume(sa, k);
    __builtin_amdgcn_sched_barrier(0);
    issue(sa, i0 + k + 2);
    consume(sb, k + 1);
    __builtin_amdgcn_sched_barrier(0);
  }
  issue(sb, i0 + R - 1);
  consume(sa, R - 2);
  consume(sb, R - 1);

  tile_reduce(ac, cw, gsc, tile, partials);
}

// Slot loop: one tile per slot, slots walked with stride gridDim.x (a grid of
// n_slots workgroups takes one slot each).
template <int DT, bool ADJ, bool HP, bool WS, int R>
__global__ __launch_bounds__(256) void sncc_kernel(const fvr_layout* __restrict__ layouts, int n_layouts,
                                                   const fvr_sncc_item* __restrict__ items, int n_items,
                                                   int n_tiles, const int2* __restrict__ map, int n_slots,
                                                   int opts, float* __restrict__ partials) {
  for (int w = blockIdx.x; w < n_slots; w += gridDim.x) {
    int item, lt, tile;
    if (!resolve_tile<R>(w, map, items, n_items, n_tiles, n_slots, opts, item, lt, tile)) continue;
    sncc_tile<DT, ADJ, HP, WS, R>(layouts, n_layouts, items[item], lt, tile, partials);
  }
}

// ---------------------------------------------------------------------------
// Hermitian-fold variant (convention N, wrapped layout).
//
// For the spectrum of a real volume the stored (wrapped) lattice satisfies
// F(n - m) = conj F(m), so the trilinear sample at the mirror position n - p
// of p is conj I(p) and the slice's phase ramp is conjugated too: the mirror
// sample (h - i, w - j) of (i, j) is conj S(i, j).  Every moment is linear in
// S conj(T) or in |S|^2 with weights cw(j), rw(i) that are odd under the
// mirror exactly where the mirrored derivative is, so a pair contributes
//   Re(S (bp conj T(i,j) + bq T(h-i, w-j)))  and  (bp + bq) |S|^2
// (and the same for every moment), with bp / bq the scipy box tests of p and
// of its mirror -- the box [0, n-1] is not symmetric about n/2, so the two
// are tested separately and the pair is exact at the edges of the volume.
// The item covers rows h/2 .. h-1 (row h/2 only right of the centre); the
// unpaired row 0 is an extra row of the first wave, and the unpaired column 0
// of rows 1 .. h/2-1 is evaluated by the otherwise idle lanes 1 .. w/2-1 of
// row h/2.
// ---------------------------------------------------------------------------
// One folded sample in flight: its raw loads, folded target, fractional
// coordinates and the two box tests (lane masks).  Phase and moment weights are
// formed when the sample is consumed, so that little state rides the pipeline.
template <int DT>
struct SmpH {
  Raw<DT> raw;
  f4v t;  // (T(i,j), conj T(mirror))
  float ts, tm, tf;
  bool bp, bq;  // box tests of the sample and of its mirror
};

// opts bit kHermExactBox: every wave takes the per-sample box tests (no interior /
// outside classes); the sums are the same bit for bit
constexpr int kHermExactBox = 1;

template <int DT, bool ADJ, int R, int DEPTH>
__device__ __forceinline__ void sncc_herm_tile(const fvr_layout* __restrict__ layouts, int n_layouts,
                                               const fvr_sncc_item& it, int lt, int tile, int opts,
                                               float* __restrict__ partials) {
  static_assert(R >= 1 && 2 * R <= 64, "rows per wave: <= 64 (row phasors live in lanes)");
  static_assert(DEPTH >= 2 && DEPTH <= 4, "samples in flight per lane");
  if (uint32_t(it.layout) >= uint32_t(n_layouts) || it.tile_h != 4 * R) {
    if (threadIdx.x < FVR_SNCC_NMOM) partials[size_t(tile) * FVR_SNCC_NMOM + threadIdx.x] = 0.f;
    return;
  }
  const fvr_layout& L = layouts[it.layout];

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int H = it.h, W = it.w, hh = H >> 1, hw = W >> 1, lo_int = it.lo_int;
  const int tyi = lt / it.tiles_x, txi = lt - tyi * it.tiles_x;
  // One step of a wave covers 2 rows x 32 columns (lane = 32 rl + column): a tile is 32
  // columns x 8R rows and lane (rl, c) of wave w walks the rows h/2 + r0 + rl + 2t,
  // t < R.  (A 2 x 32 patch of a tilted slice crosses about half as many (slow, mid)
  // storage rows per load as a 1 x 64 row does.)
  const int rl = lane >> 5;
  const int j = txi * 32 + (lane & 31);
  const bool colok = uint32_t(j) < uint32_t(W);
  const int r0 = tyi * (8 * R) + wave * (2 * R);  // first row of the wave (offset from h/2)
  int nrow = hh - r0;                               // rows of this wave
  nrow = nrow < 0 ? 0 : (nrow > 2 * R ? 2 * R : nrow);
  // the first wave's step 0 on lane row 0 is row h/2: the sample stands for its mirror only
  // right of the centre and lanes 1 .. w/2 - 1 are the mirrors of the right half (box
  // weights below; the folded target's row h/2 is stored with the same masks)
  const bool first = r0 == 0;
  const int n_r = (nrow + 1) >> 1;  // steps
  const bool full = nrow == 2 * R;  // every step of every lane is a row of the fold
  const int vlim = colok ? nrow - rl : 0;  // step t of the lane is a row of the fold iff 2t < vlim

  const uint32_t rowb = L.row_bytes, planeb = L.plane_bytes;
  const gcptr b0 = reinterpret_cast<gcptr>(L.base);
  const gcptr b1 = b0 + planeb;
  const uint32_t guard = it.guard_off;
  const float gsc = layout_scale<DT>(L);
  const gcptr TT = reinterpret_cast<gcptr>(it.target);

  const uint64_t h0 = it.box_hi[0], h1 = it.box_hi[1], h2 = it.box_hi[2];
  // mirror of a coordinate: K - p with K = n 2^32 + 2 bias = box_hi + 2^32 + bias
  // (inside iff K - h <= p <= K: two compares instead of a 64-bit subtract and one)
  const uint64_t K0 = h0 + (1ull << 32) + (h0 & 0xFFFFFFFFull);
  const uint64_t K1 = h1 + (1ull << 32) + (h1 & 0xFFFFFFFFull);
  const uint64_t K2 = h2 + (1ull << 32) + (h2 & 0xFFFFFFFFull);
  const uint64_t M0 = K0 - h0, M1 = K1 - h1, M2 = K2 - h2;
  // per-lane column part of the position, phase and weight; lanes past the last column
  // take the positions and target of column W - 1 (their box weights are 0), so that
  // their loads stay on the rows and columns the wave class below checks
  const int jc = colok ? j : W - 1;
  const int64_t bj0 = it.c[0] + int64_t(jc) * it.ustep[0];
  const int64_t bj1 = it.c[1] + int64_t(jc) * it.ustep[1];
  const int64_t bj2 = it.c[2] + int64_t(jc) * it.ustep[2];
  // Phasors of the tile (exact 32-bit turn counts), computed once per workgroup and shared
  // through LDS: the lanes' column phasors, the same in every wave (each carries its lane's
  // row offset rl), by wave 0; the phasors of the tile's even rows h/2 + tyi 8R + 2m, m < 4R,
  // by waves 1 .. 4R/64.  The phase of step t of wave w is colph x rowph(m = w R + t).
  __shared__ cpx ph_tile[64 + 4 * R];
  if (wave == 0) {
    ph_tile[lane] = phasor(it.ph_col0 + uint32_t(j) * it.ph_col + uint32_t(rl) * it.ph_row);
  } else {
    const int m = (wave - 1) * 64 + lane;
    if (m < 4 * R) ph_tile[64 + m] = phasor(it.ph_row0 + uint32_t(hh + tyi * (8 * R) + 2 * m) * it.ph_row);
  }
  cpx colph, rowph;  // read after the barrier below (rowph: lane t <-> step t)
  const float cwj = it.cw0 + float(j) * it.cws;
  // a lane's rows are every other row: position, target offset and row weight are stepped
  // by two rows from the lane's first row (stepped positions carry the storage offset
  // lo_int in their high word: the high word is then the low corner's storage index)
  const int irow = hh + r0 + rl;
  const int64_t lob = int64_t(lo_int) << 32;
  int64_t q0 = bj0 + int64_t(irow) * it.vstep[0] + lob;
  int64_t q1 = bj1 + int64_t(irow) * it.vstep[1] + lob;
  int64_t q2 = bj2 + int64_t(irow) * it.vstep[2] + lob;
  const int64_t v0 = 2 * it.vstep[0], v1 = 2 * it.vstep[1], v2 = 2 * it.vstep[2];
  uint32_t toffr = uint32_t(irow * W + jc) * 16u;
  const uint32_t tstep = uint32_t(W) * 32u;
  float rwr = it.rw0 + float(irow) * it.rws;
  const float rws = 2.f * it.rws;

  const float wEl = j >= 1 ? 2.f : 1.f;  // interior waves: the sample and, but in column 0, its mirror
  const bool hvalid = j == 0 || j >= hw, hmirror = j > hw;  // row h/2
  const float wEh = first && rl == 0 ? (hvalid ? (hmirror ? 2.f : 1.f) : 0.f) : wEl;  // step 0 of the wave

  // box tests, loads and fractional coordinates of a sample at (p0, p1, p2); valid /
  // mirror: whether the sample / its mirror take part; toff: its folded-target offset
  auto load_tested = [&](SmpH<DT>& sm, int64_t p0, int64_t p1, int64_t p2, bool valid, bool mirror, uint32_t toff) {
    const bool bp = valid & (uint64_t(p0) <= h0) & (uint64_t(p1) <= h1) & (uint64_t(p2) <= h2);
    const bool bq = valid & mirror & (uint64_t(p0) >= M0) & (uint64_t(p0) <= K0) & (uint64_t(p1) >= M1) &
                    (uint64_t(p1) <= K1) & (uint64_t(p2) >= M2) & (uint64_t(p2) <= K2);
    const bool inb = bp | bq;
    const uint32_t fs = inb ? uint32_t(uint64_t(p0) >> 32) : 0u;
    const uint32_t fm = inb ? uint32_t(uint64_t(p1) >> 32) : 0u;
    const uint32_t ff = inb ? uint32_t(uint64_t(p2) >> 32) : 0u;
    sm.ts = float(uint32_t(p0)) * kTwoM32;
    sm.tm = float(uint32_t(p1)) * kTwoM32;
    sm.tf = float(uint32_t(p2)) * kTwoM32;
    const uint32_t S = fs + lo_int, M = fm + lo_int, F = ff + lo_int;
    sm.raw.load(b0, b1, inb ? Raw<DT>::offset(S, M, F, planeb, rowb) : guard, rowb);
    sm.t = *reinterpret_cast<gf4ptr>(TT + toff);
    sm.bp = bp;
    sm.bq = bq;
  };

  // stage 1 of step t (steps in order; position and target offset stepped): interior
  // waves skip the box tests (every sample and its mirror are inside); other waves test
  // every sample, drop the rows past h - 1 and apply row h/2's masks
  auto issue = [&](SmpH<DT>& sm, int t, auto interior) {
    if constexpr (decltype(interior)::value) {
      const uint32_t S = uint32_t(uint64_t(q0) >> 32);
      const uint32_t M = uint32_t(uint64_t(q1) >> 32);
      const uint32_t F = uint32_t(uint64_t(q2) >> 32);
      sm.ts = float(uint32_t(q0)) * kTwoM32;
      sm.tm = float(uint32_t(q1)) * kTwoM32;
      sm.tf = float(uint32_t(q2)) * kTwoM32;
      sm.raw.load(b0, b1, Raw<DT>::offset(S, M, F, planeb, rowb), rowb);
      sm.t = *reinterpret_cast<gf4ptr>(TT + toffr);
    } else {
      // (tested waves step their positions without the storage offset: it is taken out once
      // before their rows)
      const bool h2 = first && rl == 0 && t == 0;
      const bool valid = (2 * t < vlim) && (!h2 || hvalid);
      load_tested(sm, q0, q1, q2, valid, h2 ? hmirror : j >= 1, valid ? toffr : 0u);
    }
    q0 += v0;
    q1 += v1;
    q2 += v2;
    toffr += tstep;
  };

  Acc ac = {};
  // stage 2: interpolation and the moment updates of a sample with phase (phr, phi),
  // moment weights (cw, rw) and box weights (wp, wq) of the sample and its mirror
  // (interior: every lane valid, wp = wq = 1; column 0's mirror part of the folded target is 0)
  auto accumulate = [&](const SmpH<DT>& sm, float phr, float phi, float cw, float rw, float wp, float wq,
                        float wEi, auto interior, auto unit_we) {
    const float ts = sm.ts, tm = sm.tm, tf = sm.tf;
    const Raw<DT>& q = sm.raw;
    float e00r, e00i, e01r, e01i, e10r, e10i, e11r, e11i, d00r, d00i, d01r, d01i, d10r, d10i, d11r, d11i;
    q.lerp(0, tf, e00r, e00i, d00r, d00i);
    q.lerp(1, tf, e01r, e01i, d01r, d01i);
    q.lerp(2, tf, e10r, e10i, d10r, d10i);
    q.lerp(3, tf, e11r, e11i, d11r, d11i);
    const float y0r = e01r - e00r, y0i = e01i - e00i;
    const float y1r = e11r - e10r, y1i = e11i - e10i;
    const float f0r = fmaf(tm, y0r, e00r), f0i = fmaf(tm, y0i, e00i);
    const float f1r = fmaf(tm, y1r, e10r), f1i = fmaf(tm, y1i, e10i);
    const float zr = f1r - f0r, zi = f1i - f0i;
    const float Ir = fmaf(ts, zr, f0r), Ii = fmaf(ts, zi, f0i);

    // folded target T'' = bp T + bq conj T(mirror); W = phase conj(T''), with I, dI in
    // units of the spectrum's scale (applied to the sums at the end)
    float Tr, Ti, wE;
    if constexpr (decltype(interior)::value) {
      Tr = sm.t.x + sm.t.z;
      Ti = sm.t.y + sm.t.w;
      wE = wEi;
    } else {
      Tr = fmaf(wq, sm.t.z, wp * sm.t.x);
      Ti = fmaf(wq, sm.t.w, wp * sm.t.y);
      wE = wp + wq;
    }
    const float Wr = fmaf(phr, Tr, phi * Ti);
    const float Wi = fmaf(phi, Tr, -phr * Ti);
    // unit_we: the E weight is a per-lane power of two applied to E and the H moments after
    // the row loop (bit-identical: scaling by 1 or 2 commutes with every rounding here)
    const float Jr = decltype(unit_we)::value ? Ir : Ir * wE;
    const float Ji = decltype(unit_we)::value ? Ii : Ii * wE;
    const float IWr = fmaf(Ir, Wr, -Ii * Wi);
    const float IWi = fmaf(Ir, Wi, Ii * Wr);
    ac.A += IWr;
    ac.sim = fmaf(cw, IWi, ac.sim);
    ac.mrim = fmaf(rw, IWi, ac.mrim);
    ac.E = fmaf(Ir, Jr, fmaf(Ii, Ji, ac.E));
    if (ADJ) {
      const float gmr = fmaf(ts, y1r - y0r, y0r), gmi = fmaf(ts, y1i - y0i, y0i);
      const float x0r = fmaf(tm, d01r - d00r, d00r), x0i = fmaf(tm, d01i - d00i, d00i);
      const float x1r = fmaf(tm, d11r - d10r, d10r), x1i = fmaf(tm, d11i - d10i, d10i);
      const float gfr = fmaf(ts, x1r - x0r, x0r), gfi = fmaf(ts, x1i - x0i, x0i);
      const float g0 = fmaf(zr, Wr, -zi * Wi), g1 = fmaf(gmr, Wr, -gmi * Wi), g2 = fmaf(gfr, Wr, -gfi * Wi);
      const float k0 = fmaf(zr, Jr, zi * Ji), k1 = fmaf(gmr, Jr, gmi * Ji), k2 = fmaf(gfr, Jr, gfi * Ji);
      ac.gs[0] = fmaf(cw, g0, ac.gs[0]);
      ac.gs[1] = fmaf(cw, g1, ac.gs[1]);
      ac.gs[2] = fmaf(cw, g2, ac.gs[2]);
      ac.gr[0] = fmaf(rw, g0, ac.gr[0]);
      ac.gr[1] = fmaf(rw, g1, ac.gr[1]);
      ac.gr[2] = fmaf(rw, g2, ac.gr[2]);
      ac.hs[0] = fmaf(cw, k0, ac.hs[0]);
      ac.hs[1] = fmaf(cw, k1, ac.hs[1]);
      ac.hs[2] = fmaf(cw, k2, ac.hs[2]);
      ac.hr[0] = fmaf(rw, k0, ac.hr[0]);
      ac.hr[1] = fmaf(rw, k1, ac.hr[1]);
      ac.hr[2] = fmaf(rw, k2, ac.hr[2]);
    }
  };

  // stage 2 of step t (steps in order: rw is stepped)
  auto consume = [&](const SmpH<DT>& sm, int t, auto interior, auto unit_we) {
    const float prr = readlane_f(rowph.r, t), pri = readlane_f(rowph.i, t);
    const float phr = colph.r * prr - colph.i * pri;
    const float phi = fmaf(colph.r, pri, colph.i * prr);
    const float rw = rwr;
    rwr += rws;
    accumulate(sm, phr, phi, cwj, rw, sm.bp ? 1.f : 0.f, sm.bq ? 1.f : 0.f,
               decltype(unit_we)::value ? 1.f : (t == 0 ? wEh : wEl), interior, unit_we);
  };

  // The unpaired samples outside the rows h/2 .. h-1, one extra sample per lane of the first
  // wave: lane row 0 takes column 0 of row j (1 <= j <= h/2 - 1, whose mirror column w is not
  // on the grid), lane row 1 row 0 at column j (its mirror row h is not on the grid).  It
  // runs after the row loop (kept out of the loop, its registers would cost a wave per SIMD).
  const int xi = rl ? 0 : j, xj = rl ? j : 0;
  const bool xvalid = colok && (rl || (j >= 1 && j <= hh - 1));
  SmpH<DT> xs;
  auto extra_issue = [&]() {
    load_tested(xs, it.c[0] + int64_t(xj) * it.ustep[0] + int64_t(xi) * it.vstep[0],
                it.c[1] + int64_t(xj) * it.ustep[1] + int64_t(xi) * it.vstep[1],
                it.c[2] + int64_t(xj) * it.ustep[2] + int64_t(xi) * it.vstep[2], xvalid, false,
                xvalid ? uint32_t(xi * W + xj) * 16u : 0u);
  };
  auto extra_consume = [&]() {
    const cpx ph = phasor(it.ph_col0 + uint32_t(xj) * it.ph_col + it.ph_row0 + uint32_t(xi) * it.ph_row);
    accumulate(xs, ph.r, ph.i, it.cw0 + float(xj) * it.cws, it.rw0 + float(xi) * it.rws, xs.bp ? 1.f : 0.f,
               0.f, 0.f, std::false_type{}, std::false_type{});
  };

  // Class of a full wave (every step of every lane a regular row).  A position is an exact
  // affine integer function of (row, column), so over a lane's rows its extremes are at its
  // first and last row: when every lane is a column of the slice and both lie in [M, h] on
  // every axis for every lane, every sample and its mirror are inside the box (interior);
  // when on some axis both lie below 0 or both above K for every lane, none is (outside: the
  // rows contribute exactly 0 and are skipped).  Other waves test every sample.
  __syncthreads();
  colph = ph_tile[lane];
  rowph = ph_tile[64 + wave * R + (lane < R ? lane : 0)];

  int cls = 0;  // 0: per-sample tests, 1: interior, 2: outside
  if (full && !(opts & kHermExactBox)) {
    const int64_t f0 = q0 - lob, f1 = q1 - lob, f2 = q2 - lob;
    const int64_t l0 = f0 + int64_t(n_r - 1) * v0, l1 = f1 + int64_t(n_r - 1) * v1, l2 = f2 + int64_t(n_r - 1) * v2;
    const bool in = colok & (uint64_t(f0) >= M0) & (uint64_t(f0) <= h0) & (uint64_t(l0) >= M0) &
                    (uint64_t(l0) <= h0) & (uint64_t(f1) >= M1) & (uint64_t(f1) <= h1) & (uint64_t(l1) >= M1) & (uint64_t(l1) <= h1) &
                    (uint64_t(f2) >= M2) & (uint64_t(f2) <= h2) & (uint64_t(l2) >= M2) & (uint64_t(l2) <= h2);
    const bool out = !colok || (f0 < 0 && l0 < 0) || (f0 > int64_t(K0) && l0 > int64_t(K0)) ||
                     (f1 < 0 && l1 < 0) || (f1 > int64_t(K1) && l1 > int64_t(K1)) || (f2 < 0 && l2 < 0) ||
                     (f2 > int64_t(K2) && l2 > int64_t(K2));
    cls = __all(in) ? 1 : (__all(out) ? 2 : 0);
  }

  // DEPTH samples in flight per lane: the loads of row t + DEPTH - 1 are issued before
  // row t is consumed (ring of DEPTH buffers, indices fixed by the unrolled inner loop;
  // every bound below is wave-uniform)
  auto run = [&](auto interior, auto unit_we) {
    SmpH<DT> ring[DEPTH];
#pragma unroll
    for (int d = 0; d < DEPTH - 1; ++d)
      if (d < n_r) issue(ring[d], d, interior);
#pragma unroll 1
    for (int t0 = 0; t0 < n_r; t0 += DEPTH) {
#pragma unroll
      for (int d = 0; d < DEPTH; ++d) {
        const int t = t0 + d;
        if (t < n_r) {
          if (t + DEPTH - 1 < n_r) issue(ring[(d + DEPTH - 1) % DEPTH], t + DEPTH - 1, interior);
          consume(ring[d], t, interior, unit_we);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
  };

  // per-lane order of the sums: the steps, then the extra sample (first wave).  An interior
  // wave other than the first has one E weight per lane (wEl: 2, or 1 in column 0), applied
  // to E and the H moments once after its rows.
  if (cls == 1 && !first) {
    run(std::true_type{}, std::true_type{});
    ac.E *= wEl;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      ac.hs[a] *= wEl;
      ac.hr[a] *= wEl;
    }
  } else if (cls == 1) {
    run(std::true_type{}, std::false_type{});
  } else if (cls == 0) {
    q0 -= lob;
    q1 -= lob;
    q2 -= lob;
    run(std::false_type{}, std::false_type{});
  }
  if (first) {
    extra_issue();
    extra_consume();
  }
  tile_reduce(ac, 1.f, gsc, tile, partials);
}

template <int DT, bool ADJ, int R, int DEPTH>
__global__ __launch_bounds__(256) void sncc_herm_kernel(const fvr_layout* __restrict__ layouts, int n_layouts,
                                                        const fvr_sncc_item* __restrict__ items, int n_items,
                                                        int n_tiles, const int2* __restrict__ map, int n_slots,
                                                        int opts, float* __restrict__ partials) {
  for (int w = blockIdx.x; w < n_slots; w += gridDim.x) {
    int item, lt, tile;
    if (!resolve_tile<R>(w, map, items, n_items, n_tiles, n_slots, opts, item, lt, tile)) continue;
    sncc_herm_tile<DT, ADJ, R, DEPTH>(layouts, n_layouts, items[item], lt, tile, opts, partials);
  }
}

// ---------------------------------------------------------------------------
// Per-item fixed-order float64 reduction of the tile partials.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void sncc_reduce_kernel(const fvr_sncc_item* __restrict__ items,
                                                          const float* __restrict__ partials,
                                                          double* __restrict__ out) {
  const fvr_sncc_item& it = items[blockIdx.x];
  const int ntile = it.tiles_x * sncc_tiles_y(it.h, it.flags, it.tile_h);
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

struct Launch {
  dim3 grid;
  hipStream_t st;
  const fvr_layout* lay;
  int n_lay;
  const fvr_sncc_item* items;
  int n_items, n_tiles;
  const int2* map;
  int n_slots;
  float* partials;
  int depth;  // samples in flight per lane (Hermitian-fold kernel: 2 or 3)
  int opts;   // kHermExactBox (fold kernel), kOrderBiplanar
};

template <int DT, bool ADJ, bool HP, bool WS, int R>
static void launch_sncc(const Launch& a) {
  hipLaunchKernelGGL((sncc_kernel<DT, ADJ, HP, WS, R>), a.grid, dim3(256), 0, a.st, a.lay, a.n_lay, a.items,
                     a.n_items, a.n_tiles, a.map, a.n_slots, a.opts, a.partials);
}

template <int DT, bool ADJ, int R>
static void launch_herm(const Launch& a) {
  if (a.depth == 3)
    hipLaunchKernelGGL((sncc_herm_kernel<DT, ADJ, R, 3>), a.grid, dim3(256), 0, a.st, a.lay, a.n_lay, a.items,
                       a.n_items, a.n_tiles, a.map, a.n_slots, a.opts, a.partials);
  else
    hipLaunchKernelGGL((sncc_herm_kernel<DT, ADJ, R, 2>), a.grid, dim3(256), 0, a.st, a.lay, a.n_lay, a.items,
                       a.n_items, a.n_tiles, a.map, a.n_slots, a.opts, a.partials);
}

// Variants: every flag combination at tile_h 64; tile_h 32 / 128 for the batched
// adjoint forms (convention N plain or folded, convention T with high-pass); the
// folded kernel also with 3 samples in flight per lane (depth).
static bool has_variant(uint32_t flags, int tile_h, int depth = 2) {
  if (flags >= 16u || !sncc_tile_h_ok(tile_h)) return false;
  if (depth != 2 && !(depth == 3 && (flags & FVR_ITEM_HERM))) return false;
  const bool adj = flags & FVR_ITEM_ADJOINT, hp = flags & FVR_ITEM_HP, ws = flags & FVR_ITEM_WRITE_SLICE;
  if (flags & FVR_ITEM_HERM) return !hp && !ws;  // the fold is convention N without slice output
  return tile_h == 64 || (adj && !ws);
}

template <int DT, int R>
static void dispatch_rows(uint32_t flags, const Launch& a) {
  const bool adj = flags & FVR_ITEM_ADJOINT, hp = flags & FVR_ITEM_HP, ws = flags & FVR_ITEM_WRITE_SLICE;
  if (flags & FVR_ITEM_HERM) {
    if (adj) launch_herm<DT, true, R>(a);
    else launch_herm<DT, false, R>(a);
    return;
  }
  if (R != 16) {  // has_variant: adjoint without slice output
    if (hp) launch_sncc<DT, true, true, false, R>(a);
    else launch_sncc<DT, true, false, false, R>(a);
    return;
  }
#define FVR_L(x, h, w)                                   \
  if (adj == x && hp == h && ws == w) {                  \
    launch_sncc<DT, x, h, w, 16>(a);                     \
    return;                                              \
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
}

template <int DT>
static void dispatch_dt(uint32_t flags, int tile_h, const Launch& a) {
  if (tile_h == 32) dispatch_rows<DT, 8>(flags, a);
  else if (tile_h == 128) dispatch_rows<DT, 32>(flags, a);
  else dispatch_rows<DT, 16>(flags, a);
}

}  // namespace fvr

using namespace fvr;

extern "C" {

int fvr_sncc_tile_rows(void) { return kSnccTileH; }

int fvr_sncc_has_tile_h(uint32_t flags, int32_t tile_h) { return has_variant(flags, tile_h) ? 1 : 0; }

int fvr_sncc_has_depth(uint32_t flags, int32_t tile_h, int32_t depth) {
  return has_variant(flags, tile_h, depth) ? 1 : 0;
}

int fvr_sncc_eval_plan(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* items_dev,
                       int32_t n_items, int32_t n_tiles, float* partials, uint32_t variant,
                       const int32_t* tile_map, int32_t n_slots, int32_t grid, void* stream) {
  FVR_CHECK_ARG(layouts_dev && items_dev && partials, "fvr_sncc_eval: null pointer");
  FVR_CHECK_ARG(n_items > 0 && n_tiles > 0 && n_layouts > 0, "fvr_sncc_eval: empty batch");
  const int dt = int(variant & 0xF);
  const uint32_t flags = (variant >> 4) & 0xF;
  int tile_h = int((variant >> 8) & 0xFF);
  if (tile_h == 0) tile_h = kSnccTileH;
  int depth = int((variant >> 16) & 0xF);
  if (depth == 0) depth = 2;
  FVR_CHECK_ARG(dt == FVR_C32 || dt == FVR_C16F || dt == FVR_C16B, "fvr_sncc_eval: bad dtype %d", dt);
  const bool exact_box = (variant >> 20) & 1u;
  const bool biplanar = (variant >> 21) & 1u;
  FVR_CHECK_ARG((variant >> 22) == 0u, "fvr_sncc_eval: bad variant %u", variant);
  FVR_CHECK_ARG(!biplanar || (!tile_map && n_items % 2 == 0),
                "fvr_sncc_eval: the biplanar order takes (AP, LAT) item pairs and no map");
  FVR_CHECK_ARG(!exact_box || (flags & FVR_ITEM_HERM), "fvr_sncc_eval: exact-box bit without the fold flag");
  FVR_CHECK_ARG(has_variant(flags, tile_h, depth), "fvr_sncc_eval: no kernel for flags %u with tile_h %d, depth %d",
                flags, tile_h, depth);
  if (tile_map) {
    FVR_CHECK_ARG(n_slots > 0, "fvr_sncc_eval: empty tile map");
  } else {
    FVR_CHECK_ARG(n_slots == 8 * ((n_tiles + 7) / 8), "fvr_sncc_eval: %d slots for %d tiles without a map",
                  n_slots, n_tiles);
  }
  FVR_CHECK_ARG(grid >= 0 && grid <= n_slots, "fvr_sncc_eval: grid %d for %d slots", grid, n_slots);
  Launch a;
  a.grid = dim3(grid > 0 ? grid : n_slots);
  a.st = reinterpret_cast<hipStream_t>(stream);
  a.lay = layouts_dev;
  a.n_lay = n_layouts;
  a.items = items_dev;
  a.n_items = n_items;
  a.n_tiles = n_tiles;
  a.map = reinterpret_cast<const int2*>(tile_map);
  a.n_slots = n_slots;
  a.partials = partials;
  a.depth = depth;
  a.opts = (exact_box ? kHermExactBox : 0) | (biplanar ? kOrderBiplanar : 0);
  if (dt == FVR_C32) dispatch_dt<FVR_C32>(flags, tile_h, a);
  else if (dt == FVR_C16F) dispatch_dt<FVR_C16F>(flags, tile_h, a);
  else dispatch_dt<FVR_C16B>(flags, tile_h, a);
  FVR_CHECK_LAUNCH("sncc_kernel");
  return FVR_OK;
}

int fvr_sncc_eval_v(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* items_dev, int32_t n_items,
                    int32_t n_tiles, float* partials, uint32_t variant, void* stream) {
  FVR_CHECK_ARG(layouts_dev && items_dev && partials, "fvr_sncc_eval: null pointer");
  FVR_CHECK_ARG(n_tiles > 0 && n_tiles < (1 << 30), "fvr_sncc_eval: bad tile count %d", n_tiles);
  return fvr_sncc_eval_plan(layouts_dev, n_layouts, items_dev, n_items, n_tiles, partials, variant, nullptr,
                            8 * ((n_tiles + 7) / 8), 0, stream);
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
