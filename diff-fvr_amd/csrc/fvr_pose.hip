// Convention-N pose set-up and biplanar combine on the device, so that a batched
// loss + gradient evaluation is three stream-ordered launches with no host sync:
//   fvr_pn_build_items      poses -> (AP, LAT) work items + du/dtheta, dv/dtheta
//   fvr_sncc_eval_v         fused slice / Spectral-NCC / adjoint (fvr_sncc.hip)
//   fvr_pn_reduce_combine   per-pose fixed-order reduction of the tile partials
//                           and L = 1/2(1-NCC_AP) + 1/2(1-NCC_LAT) with its
//                           exact 6-DoF gradient (gradients.py:76-108).
// Geometry follows diff_fvr.py:56-150: rot = Rz(yaw) Ry(roll) Rx(pitch) in
// (Z, Y, X) index space, AP u = rot x_hat, LAT u = rot y_hat, v = rot z_hat,
// sample (i, j) at c + (j - W/2) u + (i - H/2) v with c = (pz, py, px)/2.
// Written as straight-line code: no dynamically indexed local arrays.
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

#include "fvr_common.h"

namespace fvr {

struct V3 {
  double a0, a1, a2;  // (z, y, x)
};

__device__ __forceinline__ V3 v3(double a, double b, double c) { return {a, b, c}; }

__device__ __forceinline__ int64_t to_fx(double x) { return llrint(x * 4294967296.0); }

// +2^-22 voxel on every sample coordinate: a coordinate that float64 puts exactly
// on a lattice plane stays on it after rounded fixed-point stepping
constexpr int64_t kSnapFx = 1 << 10;

__device__ __forceinline__ uint32_t turns_fx(double x) {
  x -= floor(x);
  return uint32_t(uint64_t(llrint(x * 4294967296.0)) & 0xFFFFFFFFull);
}

// layout axes: (z, y, x) or, for the transposed lateral copy, (z, x, y)
__device__ __forceinline__ void put_fx(int64_t dst[3], const V3& p, bool swap) {
  dst[0] = to_fx(p.a0);
  dst[1] = to_fx(swap ? p.a2 : p.a1);
  dst[2] = to_fx(swap ? p.a1 : p.a2);
}

__device__ __forceinline__ void put_jac(double* J, const V3& d0, const V3& d1, const V3& d2, bool swap) {
  // J[a*3 + j]: layout axis a, angle j
  J[0] = d0.a0;
  J[1] = d1.a0;
  J[2] = d2.a0;
  const double m0 = swap ? d0.a2 : d0.a1, m1 = swap ? d1.a2 : d1.a1, m2 = swap ? d2.a2 : d2.a1;
  const double f0 = swap ? d0.a1 : d0.a2, f1 = swap ? d1.a1 : d1.a2, f2 = swap ? d2.a1 : d2.a2;
  J[3] = m0;
  J[4] = m1;
  J[5] = m2;
  J[6] = f0;
  J[7] = f1;
  J[8] = f2;
}

// item idx = 2 pose + lat (AP, LAT) of one pose and its Jacobians
__device__ __forceinline__ void build_item(const double* P, int pose, bool lat, const fvr_sncc_item& tap, const fvr_sncc_item& tlat,
                           double pz, double py, double px, int lat_swap, int tile0, fvr_sncc_item* items,
                           double* jac) {
  const int idx = 2 * pose + (lat ? 1 : 0);
  const double deg = 3.14159265358979323846 / 180.0;
  const double y = P[0] * deg, p = P[1] * deg, r = P[2] * deg;
  const double cy = cos(y), sy = sin(y), cp = cos(p), sp = sin(p), cr = cos(r), sr = sin(r);
  // columns of R = Rz(y) Ry(r) Rx(p): R e_x (col 2), R e_y (col 1), R e_z (col 0)
  const V3 col0 = v3(cr * cp, cy * sp + sy * sr * cp, sy * sp - cy * sr * cp);
  const V3 col1 = v3(-cr * sp, cy * cp - sy * sr * sp, sy * cp + cy * sr * sp);
  const V3 col2 = v3(sr, -sy * cr, cy * cr);
  // d/dyaw:  dRz Ry Rx   d/dpitch: Rz Ry dRx   d/droll: Rz dRy Rx  (columns)
  const V3 dy0 = v3(0.0, -sy * sp + cy * sr * cp, cy * sp + sy * sr * cp);
  const V3 dy1 = v3(0.0, -sy * cp - cy * sr * sp, cy * cp - sy * sr * sp);
  const V3 dy2 = v3(0.0, -cy * cr, -sy * cr);
  const V3 dp0 = v3(-cr * sp, cy * cp - sy * sr * sp, sy * cp + cy * sr * sp);
  const V3 dp1 = v3(-cr * cp, -cy * sp - sy * sr * cp, -sy * sp + cy * sr * cp);
  const V3 dp2 = v3(0.0, 0.0, 0.0);
  const V3 dr0 = v3(-sr * cp, sy * cr * cp, -cy * cr * cp);
  const V3 dr1 = v3(sr * sp, -sy * cr * sp, cy * cr * sp);
  const V3 dr2 = v3(cr, sy * sr, -cy * sr);

  const V3 u = lat ? col1 : col2;
  const V3 v = col0;
  const double H = pz, W = lat ? py : px;
  const double uu0 = -W / 2.0, vv0 = -H / 2.0;
  const V3 cg = v3(pz / 2.0 + uu0 * u.a0 + vv0 * v.a0, py / 2.0 + uu0 * u.a1 + vv0 * v.a1,
                   px / 2.0 + uu0 * u.a2 + vv0 * v.a2);
  const bool swap = lat && lat_swap;
  // the view's template, copied word by word with a per-lane select (no private copy)
  static_assert(sizeof(fvr_sncc_item) % 4 == 0, "item words");
  fvr_sncc_item it;
  {
    const uint32_t* wa = reinterpret_cast<const uint32_t*>(&tap);
    const uint32_t* wb = reinterpret_cast<const uint32_t*>(&tlat);
    uint32_t* wd = reinterpret_cast<uint32_t*>(&it);
#pragma unroll
    for (int k = 0; k < int(sizeof(fvr_sncc_item) / 4); ++k) wd[k] = lat ? wb[k] : wa[k];
  }
  put_fx(it.c, cg, swap);
  for (int a = 0; a < 3; ++a) it.c[a] += kSnapFx;  // lattice-plane bias, as the host builder
  // a pose with a NaN or infinite entry: the item reads nothing (a layout index out of range
  // gives zero partials) and its Jacobians are NaN, which the combine turns into a NaN loss
  // and gradient for the pose
  bool finite = true;
  for (int k = 0; k < 6; ++k) finite = finite && isfinite(P[k]);
  if (!finite) it.layout = -1;
  put_fx(it.ustep, u, swap);
  put_fx(it.vstep, v, swap);
  const double a = lat ? P[4] : P[3];  // AP: tx, LAT: ty (gradients.py:90-94)
  const double b = P[5];               // tz
  it.ph_col = turns_fx(a / W);
  it.ph_col0 = turns_fx(uu0 * a / W);
  it.ph_row = turns_fx(b / H);
  it.ph_row0 = turns_fx(vv0 * b / H);
  const int t_ap = tap.tiles_x * sncc_tiles_y(tap.h, tap.flags, tap.tile_h);
  const int t_lat = tlat.tiles_x * sncc_tiles_y(tlat.h, tlat.flags, tlat.tile_h);
  it.tile_base = tile0 + pose * (t_ap + t_lat) + (lat ? t_ap : 0);
  items[idx] = it;
  double* J = jac + size_t(idx) * 18;
  const double k = finite ? deg : __builtin_nan("");
  if (lat) {
    put_jac(J, v3(dy1.a0 * k, dy1.a1 * k, dy1.a2 * k), v3(dp1.a0 * k, dp1.a1 * k, dp1.a2 * k),
            v3(dr1.a0 * k, dr1.a1 * k, dr1.a2 * k), swap);
  } else {
    put_jac(J, v3(dy2.a0 * k, dy2.a1 * k, dy2.a2 * k), v3(dp2.a0 * k, dp2.a1 * k, dp2.a2 * k),
            v3(dr2.a0 * k, dr2.a1 * k, dr2.a2 * k), swap);
  }
  put_jac(J + 9, v3(dy0.a0 * k, dy0.a1 * k, dy0.a2 * k), v3(dp0.a0 * k, dp0.a1 * k, dp0.a2 * k),
          v3(dr0.a0 * k, dr0.a1 * k, dr0.a2 * k), swap);
}

__global__ void pn_build_kernel(const double* __restrict__ poses, int n_poses, fvr_sncc_item tap, fvr_sncc_item tlat,
                                double pz, double py, double px, int lat_swap, int tile0,
                                fvr_sncc_item* __restrict__ items, double* __restrict__ jac) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 2 * n_poses) return;
  build_item(poses + 6 * (idx >> 1), idx >> 1, idx & 1, tap, tlat, pz, py, px, lat_swap, tile0, items, jac);
}

// What the reduction of one pose reads from its two items and Jacobians, staged in LDS
// by the workgroup (so that the next items can be rebuilt while the sums are formed).
struct PoseStage {
  double J[2][18];
  int tile_base[2], ntile[2], w[2], h[2];
};
__device__ __forceinline__ void stage_pose(const fvr_sncc_item* items, const double* jac, int pose, PoseStage& st) {
  const int tid = threadIdx.x;
  if (tid < 36) st.J[tid / 18][tid % 18] = jac[size_t(2 * pose) * 18 + tid];
  if (tid < 2) {
    const fvr_sncc_item& it = items[2 * pose + tid];
    st.tile_base[tid] = it.tile_base;
    st.ntile[tid] = it.tiles_x * sncc_tiles_y(it.h, it.flags, it.tile_h);
    st.w[tid] = it.w;
    st.h[tid] = it.h;
  }
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b & 0xFFFFFFFFll), l), hi = __builtin_amdgcn_readlane(int(b >> 32), l);
  return __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
}

// One wave per pose: fixed-order float64 reduction of each view's tile
// partials (half a wave per view: lane-strided partial sums, then an XOR butterfly
// inside the half), then the biplanar loss and exact gradient.
__device__ void reduce_combine_pose(const PoseStage& st, int pose, const float* __restrict__ partials, double t2_ap,
                                    double t2_lat, double* moments, double& L_out, double (&g_out)[6]) {
  const int lane = threadIdx.x & 63;
  const double two_pi = 6.283185307179586476925286766559;
  double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double L = 0.0;
  // lanes 0-31 sum the AP view's tiles and lanes 32-63 the LAT view's, each half in a fixed
  // order: lane-strided (stride 32) over its view's tiles, then an XOR butterfly inside the
  // half (every lane of a half ends with the same bits); then both sums are read from lanes
  // 0 and 32
  const int half = lane >> 5;
  double mh[FVR_SNCC_NMOM];
#pragma unroll
  for (int q = 0; q < FVR_SNCC_NMOM; ++q) mh[q] = 0.0;
  {
    const int ntile = st.ntile[half], tile_base = st.tile_base[half];
    for (int t = lane & 31; t < ntile; t += 32) {
      const float* pp = partials + size_t(tile_base + t) * FVR_SNCC_NMOM;
#pragma unroll
      for (int q = 0; q < FVR_SNCC_NMOM; ++q) mh[q] += double(pp[q]);
    }
  }
#pragma unroll
  for (int o = 16; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < FVR_SNCC_NMOM; ++q) mh[q] += __shfl_xor(mh[q], o);
  double mm[2][FVR_SNCC_NMOM];
#pragma unroll
  for (int view = 0; view < 2; ++view)
#pragma unroll
    for (int q = 0; q < FVR_SNCC_NMOM; ++q) mm[view][q] = readlane_d(mh[q], 32 * view);
#pragma unroll
  for (int view = 0; view < 2; ++view) {
    const double* m = mm[view];
    if (moments && lane < FVR_SNCC_NMOM) {
      double mv = 0.0;
#pragma unroll
      for (int q = 0; q < FVR_SNCC_NMOM; ++q) mv = q == lane ? m[q] : mv;
      moments[size_t(2 * pose + view) * FVR_SNCC_NMOM + lane] = mv;
    }
    const double* J = st.J[view];
    if (J[0] != J[0]) {  // a pose with a non-finite entry (set-up marks its Jacobians NaN)
      L = J[0];
      for (int a = 0; a < 6; ++a) g[a] = J[0];
      continue;
    }
    const double nt = sqrt(view ? t2_lat : t2_ap), ns = sqrt(fmax(m[1], 0.0));
    double ncc = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0, da = 0.0, db = 0.0;
    if (ns * nt != 0.0) {
      const double inv = 1.0 / (ns * nt);
      ncc = m[0] * inv;
      double G0 = 0.0, G1 = 0.0, G2 = 0.0, H0 = 0.0, H1 = 0.0, H2 = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        G0 += m[4 + a] * J[a * 3 + 0] + m[7 + a] * J[9 + a * 3 + 0];
        G1 += m[4 + a] * J[a * 3 + 1] + m[7 + a] * J[9 + a * 3 + 1];
        G2 += m[4 + a] * J[a * 3 + 2] + m[7 + a] * J[9 + a * 3 + 2];
        H0 += m[10 + a] * J[a * 3 + 0] + m[13 + a] * J[9 + a * 3 + 0];
        H1 += m[10 + a] * J[a * 3 + 1] + m[13 + a] * J[9 + a * 3 + 1];
        H2 += m[10 + a] * J[a * 3 + 2] + m[13 + a] * J[9 + a * 3 + 2];
      }
      d0 = G0 * inv - ncc * H0 / m[1];
      d1 = G1 * inv - ncc * H1 / m[1];
      d2 = G2 * inv - ncc * H2 / m[1];
      da = two_pi / double(st.w[view]) * m[2] * inv;  // d/da: 2 pi / W sum uu Im(S T*)
      db = two_pi / double(st.h[view]) * m[3] * inv;  // d/db: 2 pi / H sum vv Im(S T*)
    }
    L += 0.5 * (1.0 - ncc);
    g[0] -= 0.5 * d0;
    g[1] -= 0.5 * d1;
    g[2] -= 0.5 * d2;
    if (view == 0) g[3] -= 0.5 * da;
    else g[4] -= 0.5 * da;
    g[5] -= 0.5 * db;
  }
  L_out = L;
#pragma unroll
  for (int a = 0; a < 6; ++a) g_out[a] = g[a];
}

__device__ __forceinline__ void store_loss_grad(int pose, double L, const double (&g)[6], double* loss, double* grad) {
  const int lane = threadIdx.x & 63;
  if (lane == 0) loss[pose] = L;
  if (lane < 6) {
    double gv = 0.0;
#pragma unroll
    for (int a = 0; a < 6; ++a) gv = a == lane ? g[a] : gv;
    grad[6 * pose + lane] = gv;
  }
}

__global__ __launch_bounds__(64) void pn_reduce_combine_kernel(const fvr_sncc_item* __restrict__ items, int n_poses,
                                                               const float* __restrict__ partials,
                                                               const double* __restrict__ jac, double t2_ap,
                                                               double t2_lat, double* __restrict__ moments,
                                                               double* __restrict__ loss, double* __restrict__ grad) {
  const int pose = blockIdx.x;
  __shared__ PoseStage st;
  stage_pose(items, jac, pose, st);
  __syncthreads();
  double L, g[6];
  reduce_combine_pose(st, pose, partials, t2_ap, t2_lat, moments, L, g);
  store_loss_grad(pose, L, g, loss, grad);
}

// Batched Adam, one thread per pose (gradient_optimizer.py:45-83 per pose, with
// the gradient clamp and cosine learning-rate factor of the multistart driver
// of experiments/exp30_gpu_multistart_registration.py:140-183).
struct AdamState {
  double *m, *v;
  int32_t *t, *done;
  double *best_loss, *best_pose;
};

// One Adam step of pose k: P (its 6 values) is updated in place; every calling lane
// reads the same state and computes the same update, and only a lane with `store`
// writes the state back (all reads precede all writes).
__device__ __forceinline__ void adam_pose(int k, double (&P)[6], double L, const double* gk, const AdamState& S,
                                          const fvr_adam_cfg& c, bool store) {
  // no implicit FMA contraction: every kernel that inlines this computes the same bits
#pragma clang fp contract(off)
  if (S.done[k]) return;
  const int step = S.t[k] + 1;
  const bool better = L < S.best_loss[k];
  double mo[6], vo[6], P0[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    mo[a] = S.m[6 * k + a];
    vo[a] = S.v[6 * k + a];
    P0[a] = P[a];
  }
  double f = 1.0;
  if (c.cosine_t > 0) {
    const double s = double(step - 1 < c.cosine_t ? step - 1 : c.cosine_t);
    f = 0.1 + 0.45 * (1.0 + cos(3.14159265358979323846 * s / double(c.cosine_t)));
  }
  const double bc1 = 1.0 - pow(c.beta1, double(step)), bc2 = 1.0 - pow(c.beta2, double(step));
  double mx = 0.0;
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    double g = gk[a];
    if (c.gclip[a] > 0.0) g = fmin(fmax(g, -c.gclip[a]), c.gclip[a]);
    mo[a] = fma(c.beta1, mo[a], (1.0 - c.beta1) * g);
    vo[a] = fma(c.beta2, vo[a], ((1.0 - c.beta2) * g) * g);
    const double upd = f * c.lr[a] * (mo[a] / bc1) / (sqrt(vo[a] / bc2) + c.eps);
    mx = fmax(mx, fabs(upd));
    P[a] = fmin(fmax(P[a] - upd, c.lo[a]), c.hi[a]);
  }
  if (store) {
    S.t[k] = step;
    if (better) {
      S.best_loss[k] = L;
#pragma unroll
      for (int a = 0; a < 6; ++a) S.best_pose[6 * k + a] = P0[a];
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      S.m[6 * k + a] = mo[a];
      S.v[6 * k + a] = vo[a];
    }
    if (mx < c.tol) S.done[k] = 1;
  }
}

__global__ void adam_step_kernel(double* __restrict__ poses, const double* __restrict__ loss,
                                 const double* __restrict__ grad, AdamState S, int n, fvr_adam_cfg c) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double P[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) P[a] = poses[6 * k + a];
  adam_pose(k, P, loss[k], grad + 6 * k, S, c, true);
#pragma unroll
  for (int a = 0; a < 6; ++a) poses[6 * k + a] = P[a];
}

// ---------------------------------------------------------------------------
// Convention T (torch_engine.py:138-234) on the device: one thread per pose
// builds the item of one view from a host-validated template, and one wave per
// pose reduces its tiles and forms NCC = A / (sqrt(E E_t) + 1e-8) and d NCC /
// d pose (torch_engine.py:269-290).
// ---------------------------------------------------------------------------
struct M3 {
  double m[3][3];
};

__device__ __forceinline__ M3 mul3(const M3& a, const M3& b) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
  return r;
}

struct TGeom {
  double D, H, W;    // padded dims (z, y, x)
  double spacing0;   // spacing of the phase span R * spacing + 1e-6
  int R, lat;        // output resolution, lateral view (R Rz(90))
  int ord[3];        // layout axis -> xyz component
  unsigned long long target0, target_stride;  // target of pose p: target0 + p * stride
};

// v[k] for a runtime k in {0, 1, 2} without indexing private memory
__device__ __forceinline__ double pick3(const double (&v)[3], int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }

__global__ void pt_build_kernel(const double* __restrict__ poses, int n, fvr_sncc_item tmpl, TGeom G,
                                fvr_sncc_item* __restrict__ items, double* __restrict__ jac) {
  const int pose = blockIdx.x * blockDim.x + threadIdx.x;
  if (pose >= n) return;
  const double deg = 3.14159265358979323846 / 180.0;
  const double* P = poses + 6 * pose;
  const double cz = cos(P[0] * deg), sz = sin(P[0] * deg);  // yaw about z
  const double cy = cos(P[1] * deg), sy = sin(P[1] * deg);  // pitch about y
  const double cx = cos(P[2] * deg), sx = sin(P[2] * deg);  // roll about x
  const M3 Rz = {{{cz, -sz, 0.0}, {sz, cz, 0.0}, {0.0, 0.0, 1.0}}};
  const M3 Ry = {{{cy, 0.0, sy}, {0.0, 1.0, 0.0}, {-sy, 0.0, cy}}};
  const M3 Rx = {{{1.0, 0.0, 0.0}, {0.0, cx, -sx}, {0.0, sx, cx}}};
  const M3 dRz = {{{-sz, -cz, 0.0}, {cz, -sz, 0.0}, {0.0, 0.0, 0.0}}};
  const M3 dRy = {{{-sy, 0.0, cy}, {0.0, 0.0, 0.0}, {-cy, 0.0, -sy}}};
  const M3 dRx = {{{0.0, 0.0, 0.0}, {0.0, -sx, -cx}, {0.0, cx, -sx}}};
  const M3 Rz90 = {{{0.0, -1.0, 0.0}, {1.0, 0.0, 0.0}, {0.0, 0.0, 1.0}}};
  M3 R = mul3(mul3(Rz, Ry), Rx);
  M3 dR[3] = {mul3(mul3(dRz, Ry), Rx), mul3(mul3(Rz, dRy), Rx), mul3(mul3(Rz, Ry), dRx)};
  if (G.lat) {
    R = mul3(R, Rz90);
#pragma unroll
    for (int k = 0; k < 3; ++k) dR[k] = mul3(dR[k], Rz90);
  }
  // half extents (= the centre) of the padded volume in xyz
  const double s[3] = {0.5 * fmax(G.W - 1.0, 1.0), 0.5 * fmax(G.H - 1.0, 1.0), 0.5 * fmax(G.D - 1.0, 1.0)};
  const double f0 = -1.0, df = G.R > 1 ? 2.0 / double(G.R - 1) : 0.0;
  fvr_sncc_item it = tmpl;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int k = G.ord[a];
    const double sk = pick3(s, k);
    const double u = pick3(R.m[0], k) * sk, v = pick3(R.m[1], k) * sk;
    it.c[a] = to_fx(sk + f0 * u + f0 * v + 1.0) + kSnapFx;  // +1: per-corner coordinates (lo_int = 0)
    it.ustep[a] = to_fx(df * v);  // along a row: next column (v)
    it.vstep[a] = to_fx(df * u);  // down a column: next row (u)
  }
  const double span = double(G.R) * G.spacing0 + 1e-6;
  it.ph_row0 = turns_fx(f0 * P[3] / span);
  it.ph_row = turns_fx(df * P[3] / span);
  it.ph_col0 = turns_fx(f0 * P[4] / span);
  it.ph_col = turns_fx(df * P[4] / span);
  it.tile_base = pose * (tmpl.tiles_x * sncc_tiles_y(tmpl.h, tmpl.flags, tmpl.tile_h));
  it.target = G.target0 + (unsigned long long)pose * G.target_stride;
  // a non-finite angle or in-plane translation (tz does not enter convention T): the item reads
  // nothing and its Jacobians are NaN, which the combine turns into a NaN NCC and gradient
  bool finite = true;
  for (int k = 0; k < 5; ++k) finite = finite && isfinite(P[k]);
  if (!finite) it.layout = -1;
  items[pose] = it;
  // jac[p][0][a][theta] = d v / d theta, jac[p][1][a][theta] = d u / d theta (layout axes, per degree)
  double* J = jac + size_t(pose) * 18;
  const double jdeg = finite ? deg : __builtin_nan("");
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const int k = G.ord[a];
      J[a * 3 + t] = pick3(dR[t].m[1], k) * pick3(s, k) * jdeg;
      J[9 + a * 3 + t] = pick3(dR[t].m[0], k) * pick3(s, k) * jdeg;
    }
}

// one wave per pose: ncc[p] and dgrad[p][6] = -(1/n) d NCC_p / d pose (the gradient of
// 1 - mean NCC); e_t: target energy of pose p at e_t[p * et_stride]
__global__ __launch_bounds__(64) void pt_combine_kernel(const fvr_sncc_item* __restrict__ items, int n,
                                                        const float* __restrict__ partials,
                                                        const double* __restrict__ jac, const double* __restrict__ e_t,
                                                        int et_stride, double span, double* __restrict__ ncc_out,
                                                        double* __restrict__ grad) {
  const int pose = blockIdx.x, lane = threadIdx.x;
  const fvr_sncc_item& it = items[pose];
  const int ntile = it.tiles_x * sncc_tiles_y(it.h, it.flags, it.tile_h);
  double m[FVR_SNCC_NMOM];
#pragma unroll
  for (int q = 0; q < FVR_SNCC_NMOM; ++q) m[q] = 0.0;
  for (int t = lane; t < ntile; t += 64) {
    const float* pp = partials + size_t(it.tile_base + t) * FVR_SNCC_NMOM;
#pragma unroll
    for (int q = 0; q < FVR_SNCC_NMOM; ++q) m[q] += double(pp[q]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < FVR_SNCC_NMOM; ++q) m[q] += __shfl_xor(m[q], o);
  const double* J = jac + size_t(pose) * 18;
  const double E = fmax(m[1], 0.0), et = e_t[size_t(pose) * et_stride];
  const double Dn = sqrt(E * et) + 1e-8;
  const double ncc = m[0] / Dn;
  double d[6];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    double Gt = 0.0, Ht = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      Gt += m[4 + a] * J[a * 3 + t] + m[7 + a] * J[9 + a * 3 + t];
      Ht += m[10 + a] * J[a * 3 + t] + m[13 + a] * J[9 + a * 3 + t];
    }
    const double dD = E > 0.0 ? sqrt(et) * Ht / sqrt(E) : 0.0;
    d[t] = Gt / Dn - (m[0] / (Dn * Dn)) * dD;
  }
  const double two_pi = 6.283185307179586476925286766559;
  d[3] = two_pi / span * m[3] / Dn;  // d/dtx: row moments (rows follow u)
  d[4] = two_pi / span * m[2] / Dn;  // d/dty: column moments
  d[5] = 0.0;                        // tz does not enter convention T
  double ncc_p = ncc;
  if (J[0] != J[0]) {  // a pose with a non-finite entry (set-up marks its Jacobians NaN)
    ncc_p = J[0];
#pragma unroll
    for (int a = 0; a < 6; ++a) d[a] = J[0];
  }
  if (lane == 0) ncc_out[pose] = ncc_p;
  if (lane < 6) {
    double gv = 0.0;
#pragma unroll
    for (int a = 0; a < 6; ++a) gv = a == lane ? d[a] : gv;
    grad[6 * pose + lane] = -gv / double(n);
  }
}

// One workgroup of two waves per pose after the fused kernel: reduce + combine (loss,
// gradient at the pose the items were built from), then optionally one Adam step of the
// pose, then the pose's two items and Jacobians for the next fused launch.  The item
// fields and Jacobians the reduction reads are staged in LDS first; without Adam the
// second wave rebuilds the items while the first forms the sums (with Adam the rebuild
// needs the stepped pose and follows on the first wave).  A registration step is then two
// launches: the fused kernel and this one.
__global__ __launch_bounds__(128) void pn_step_kernel(fvr_sncc_item* items, int n_poses, const float* __restrict__ partials,
                                                     double* jac, double t2_ap, double t2_lat,
                                                     double* __restrict__ moments, double* __restrict__ loss,
                                                     double* __restrict__ grad, double* __restrict__ poses,
                                                     int use_adam, fvr_adam_cfg c, AdamState S, fvr_sncc_item tap,
                                                     fvr_sncc_item tlat, double pz, double py, double px,
                                                     int lat_swap, int tile0) {
  const int pose = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __shared__ PoseStage st;
  stage_pose(items, jac, pose, st);
  double P[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) P[a] = poses[6 * pose + a];
  __syncthreads();  // this pose's items and Jacobians are staged before they are rebuilt
  if (wave == 0) {
    double L, g[6];
    reduce_combine_pose(st, pose, partials, t2_ap, t2_lat, moments, L, g);  // every lane: same L, g
    store_loss_grad(pose, L, g, loss, grad);
    if (use_adam) {
      adam_pose(pose, P, L, g, S, c, lane == 0);  // lanes agree; lane 0 stores the state
      if (lane == 0)
#pragma unroll
        for (int a = 0; a < 6; ++a) poses[6 * pose + a] = P[a];
      if (lane < 2) build_item(P, pose, lane == 1, tap, tlat, pz, py, px, lat_swap, tile0, items, jac);
    }
  } else if (!use_adam && lane < 2) {
    build_item(P, pose, lane == 1, tap, tlat, pz, py, px, lat_swap, tile0, items, jac);
  }
}

// loss = 1 - mean(ncc[0 .. n)) as float32: one wave, lane-strided float64 sums in a fixed
// order, then an XOR butterfly
__global__ __launch_bounds__(64) void mean_loss_kernel(const double* __restrict__ ncc, int n, float* __restrict__ loss) {
  const int lane = threadIdx.x;
  double s = 0.0;
  for (int k = lane; k < n; k += 64) s += ncc[k];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) *loss = float(1.0 - s / double(n));
}

}  // namespace fvr

using namespace fvr;

extern "C" {

int fvr_pn_build_items(const double* poses, int32_t n_poses, const fvr_sncc_item* tmpl_host, double pz, double py,
                       double px, int32_t lat_swap, int32_t tile0, fvr_sncc_item* items_dev, double* jac_dev,
                       void* stream) {
  FVR_CHECK_ARG(poses && tmpl_host && items_dev && jac_dev && n_poses > 0 && tile0 >= 0,
                "fvr_pn_build_items: bad arguments");
  FVR_CHECK_ARG(tmpl_host[0].w == int(px) && tmpl_host[0].h == int(pz) && tmpl_host[1].h == int(pz) &&
                    tmpl_host[1].w == int(py),
                "fvr_pn_build_items: templates do not match the padded dims");
  FVR_CHECK_ARG(sncc_tile_h_ok(tmpl_host[0].tile_h) && tmpl_host[1].tile_h == tmpl_host[0].tile_h,
                "fvr_pn_build_items: bad template tile_h %d / %d", tmpl_host[0].tile_h, tmpl_host[1].tile_h);
  const int64_t t_pose =
      int64_t(tmpl_host[0].tiles_x) * sncc_tiles_y(tmpl_host[0].h, tmpl_host[0].flags, tmpl_host[0].tile_h) +
      int64_t(tmpl_host[1].tiles_x) * sncc_tiles_y(tmpl_host[1].h, tmpl_host[1].flags, tmpl_host[1].tile_h);
  FVR_CHECK_ARG(t_pose > 0 && int64_t(tile0) + int64_t(n_poses) * t_pose < (int64_t(1) << 31),
                "fvr_pn_build_items: %d poses x %lld tiles overflow the 32-bit tile index", n_poses,
                (long long)t_pose);
  const int n = 2 * n_poses;
  hipLaunchKernelGGL(pn_build_kernel, dim3((n + 63) / 64), dim3(64), 0, reinterpret_cast<hipStream_t>(stream), poses,
                     n_poses, tmpl_host[0], tmpl_host[1], pz, py, px, lat_swap, int(tile0), items_dev, jac_dev);
  FVR_CHECK_LAUNCH("pn_build_kernel");
  return FVR_OK;
}

int fvr_pn_reduce_combine(const fvr_sncc_item* items_dev, int32_t n_poses, const float* partials, const double* jac,
                          double t2_ap, double t2_lat, double* moments, double* loss, double* grad, void* stream) {
  FVR_CHECK_ARG(items_dev && partials && jac && loss && grad && n_poses > 0, "fvr_pn_reduce_combine: bad arguments");
  hipLaunchKernelGGL(pn_reduce_combine_kernel, dim3(n_poses), dim3(64), 0, reinterpret_cast<hipStream_t>(stream),
                     items_dev, n_poses, partials, jac, t2_ap, t2_lat, moments, loss, grad);
  FVR_CHECK_LAUNCH("pn_reduce_combine_kernel");
  return FVR_OK;
}

}  // extern "C"

namespace fvr {

static inline int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

// tiles of one pose (AP + LAT) for two validated templates
static inline int64_t pn_tiles_per_pose(const fvr_sncc_item* t) {
  return int64_t(t[0].tiles_x) * sncc_tiles_y(t[0].h, t[0].flags, t[0].tile_h) +
         int64_t(t[1].tiles_x) * sncc_tiles_y(t[1].h, t[1].flags, t[1].tile_h);
}

// workspace of fvr_pn_evaluate: items | Jacobians | tile partials, each 256-byte aligned
struct PnWs {
  int64_t items, jac, partials, total;
};
static inline PnWs pn_ws_layout(const fvr_sncc_item* t, int32_t n) {
  PnWs w;
  w.items = 0;
  w.jac = align256(int64_t(2) * n * int64_t(sizeof(fvr_sncc_item)));
  w.partials = w.jac + align256(int64_t(2) * n * 18 * int64_t(sizeof(double)));
  w.total = w.partials + align256(int64_t(n) * pn_tiles_per_pose(t) * FVR_SNCC_NMOM * int64_t(sizeof(float)));
  return w;
}

// argument checks shared by fvr_pn_evaluate and its allocation-free variant
static int pn_evaluate_check(const char* who, const fvr_layout* layouts_dev, int32_t n_layouts,
                             const fvr_sncc_item* tmpl_host, double pz, double py, double px, uint32_t variant,
                             const double* poses, int32_t n_poses, const double* loss, const double* grad) {
  FVR_CHECK_ARG(layouts_dev && n_layouts > 0 && tmpl_host && poses && loss && grad && n_poses > 0,
                "%s: bad arguments", who);
  FVR_CHECK_ARG(tmpl_host[0].w == int(px) && tmpl_host[0].h == int(pz) && tmpl_host[1].h == int(pz) &&
                    tmpl_host[1].w == int(py),
                "%s: templates do not match the padded dims", who);
  int tile_h = int((variant >> 8) & 0xFF);
  if (tile_h == 0) tile_h = 64;
  const uint32_t flags = (variant >> 4) & 0xF;
  FVR_CHECK_ARG(tmpl_host[0].tile_h == tile_h && tmpl_host[1].tile_h == tile_h,
                "%s: template tile_h %d / %d differs from the variant's %d", who, tmpl_host[0].tile_h,
                tmpl_host[1].tile_h, tile_h);
  FVR_CHECK_ARG(tmpl_host[0].flags == flags && tmpl_host[1].flags == flags,
                "%s: template flags %u / %u differ from the variant's %u", who, tmpl_host[0].flags,
                tmpl_host[1].flags, flags);
  FVR_CHECK_ARG(flags & FVR_ITEM_ADJOINT, "%s: the biplanar gradient needs the adjoint flag", who);
  const int64_t tp = pn_tiles_per_pose(tmpl_host);
  FVR_CHECK_ARG(tp > 0 && int64_t(n_poses) * tp < (int64_t(1) << 30), "%s: %d poses x %lld tiles too many", who,
                n_poses, (long long)tp);
  return FVR_OK;
}

// build -> fused kernel (XCD-contiguous tile order) -> reduce/combine over a workspace at ws
static int pn_evaluate_run(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* tmpl_host,
                           double pz, double py, double px, int32_t lat_swap, uint32_t variant, const double* poses,
                           int32_t n_poses, double t2_ap, double t2_lat, char* ws, double* moments, double* loss,
                           double* grad, void* stream) {
  const PnWs w = pn_ws_layout(tmpl_host, n_poses);
  fvr_sncc_item* items = reinterpret_cast<fvr_sncc_item*>(ws + w.items);
  double* jac = reinterpret_cast<double*>(ws + w.jac);
  float* partials = reinterpret_cast<float*>(ws + w.partials);
  const int32_t n_tiles = int32_t(int64_t(n_poses) * pn_tiles_per_pose(tmpl_host));
  int rc = fvr_pn_build_items(poses, n_poses, tmpl_host, pz, py, px, lat_swap, 0, items, jac, stream);
  if (rc != FVR_OK) return rc;
  rc = fvr_sncc_eval_plan(layouts_dev, n_layouts, items, 2 * n_poses, n_tiles, partials, variant, nullptr,
                          8 * ((n_tiles + 7) / 8), 0, stream);
  if (rc != FVR_OK) return rc;
  return fvr_pn_reduce_combine(items, n_poses, partials, jac, t2_ap, t2_lat, moments, loss, grad, stream);
}

// The allocation-free variant's workspace: one static array per GPU (each device that
// loads the library holds its own copy until the process ends).  Host code never uses the
// array's name as an address: the device address comes from hipGetSymbolAddress, once per
// device, and is kept in scratch_dev[].
static __device__ __align__(256) char scratch[FVR_PN_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static void* scratch_dev[kMaxDevices] = {};

// device address of the current device's scratch array (looked up on its first use)
static int scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("fvr_pn_evaluate_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    return FVR_ERR_HIP;
  }
  FVR_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "fvr_pn_evaluate_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("fvr_pn_evaluate_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                hipGetErrorString(e));
      return FVR_ERR_HIP;
    }
    scratch_dev[dev] = p;
  }
  *out = scratch_dev[dev];
  return FVR_OK;
}

}  // namespace fvr

extern "C" {

int64_t fvr_pn_evaluate_workspace(const fvr_sncc_item* tmpl_host, int32_t n_poses) {
  if (!tmpl_host || n_poses <= 0 || !sncc_tile_h_ok(tmpl_host[0].tile_h) ||
      tmpl_host[1].tile_h != tmpl_host[0].tile_h || tmpl_host[0].tiles_x <= 0 || tmpl_host[1].tiles_x <= 0)
    return -1;
  return pn_ws_layout(tmpl_host, n_poses).total;
}

int fvr_pn_evaluate(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* tmpl_host, double pz,
                    double py, double px, int32_t lat_swap, uint32_t variant, const double* poses, int32_t n_poses,
                    double t2_ap, double t2_lat, void* workspace, int64_t ws_bytes, double* moments, double* loss,
                    double* grad, void* stream) {
  if (pn_evaluate_check("fvr_pn_evaluate", layouts_dev, n_layouts, tmpl_host, pz, py, px, variant, poses, n_poses,
                        loss, grad) != FVR_OK)
    return FVR_ERR_INVALID;
  const int64_t need = pn_ws_layout(tmpl_host, n_poses).total;
  FVR_CHECK_ARG(workspace && ws_bytes >= need, "fvr_pn_evaluate: workspace of %lld bytes, %lld needed",
                (long long)ws_bytes, (long long)need);
  FVR_CHECK_ARG((reinterpret_cast<uintptr_t>(workspace) & 255) == 0, "fvr_pn_evaluate: workspace not 256-byte aligned");
  return pn_evaluate_run(layouts_dev, n_layouts, tmpl_host, pz, py, px, lat_swap, variant, poses, n_poses, t2_ap,
                         t2_lat, static_cast<char*>(workspace), moments, loss, grad, stream);
}

int fvr_pn_evaluate_noalloc(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* tmpl_host,
                            double pz, double py, double px, int32_t lat_swap, uint32_t variant, const double* poses,
                            int32_t n_poses, double t2_ap, double t2_lat, double* moments, double* loss,
                            double* grad, void* stream) {
  if (pn_evaluate_check("fvr_pn_evaluate_noalloc", layouts_dev, n_layouts, tmpl_host, pz, py, px, variant, poses,
                        n_poses, loss, grad) != FVR_OK)
    return FVR_ERR_INVALID;
  const int64_t need = pn_ws_layout(tmpl_host, n_poses).total;
  FVR_CHECK_ARG(need <= int64_t(FVR_PN_SCRATCH_BYTES),
                "fvr_pn_evaluate_noalloc: %d poses need %lld bytes of scratch, the static array holds %lld "
                "(use fvr_pn_evaluate with a workspace, or fewer poses per call)",
                n_poses, (long long)need, (long long)FVR_PN_SCRATCH_BYTES);
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc != FVR_OK) return rc;
  return pn_evaluate_run(layouts_dev, n_layouts, tmpl_host, pz, py, px, lat_swap, variant, poses, n_poses, t2_ap,
                         t2_lat, static_cast<char*>(ws), moments, loss, grad, stream);
}

// Launch-plan autotuner of fvr_pn_evaluate: the (tile height, samples in flight, tile order)
// variants the library has for the templates' flags, timed with HIP events on the stream;
// results cached per (device, dtype / flags / exact_box bits, padded dims, slice sizes,
// lateral axis order, pose count) under a mutex.
static std::mutex tune_mu;
static std::map<std::tuple<int, uint32_t, int, int, int, int, int, int, int, int>, std::pair<uint32_t, float>>
    tune_cache;

int fvr_autotune_pn_evaluate(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* tmpl_host,
                             double pz, double py, double px, int32_t lat_swap, uint32_t variant, const double* poses,
                             int32_t n_poses, double t2_ap, double t2_lat, void* workspace, int64_t ws_bytes,
                             double* loss, double* grad, int32_t reps, int32_t use_cache, uint32_t* best_variant,
                             float* best_ms, void* stream) {
  if (pn_evaluate_check("fvr_autotune_pn_evaluate", layouts_dev, n_layouts, tmpl_host, pz, py, px, variant, poses,
                        n_poses, loss, grad) != FVR_OK)
    return FVR_ERR_INVALID;
  FVR_CHECK_ARG(best_variant && best_ms && reps > 0 && reps <= 1000, "fvr_autotune_pn_evaluate: bad arguments");
  FVR_CHECK_ARG(workspace && (reinterpret_cast<uintptr_t>(workspace) & 255) == 0,
                "fvr_autotune_pn_evaluate: workspace missing or not 256-byte aligned");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("fvr_autotune_pn_evaluate: hipGetDevice failed");
    return FVR_ERR_HIP;
  }
  const uint32_t base = variant & (0xFu | (0xFu << 4) | (1u << 20));  // dtype, flags, exact_box
  const auto key = std::make_tuple(dev, base, n_poses, int(pz), int(py), int(px), int(lat_swap), tmpl_host[0].w,
                                   tmpl_host[1].w, tmpl_host[0].h);
  if (use_cache) {
    std::lock_guard<std::mutex> g(tune_mu);
    const auto hit = tune_cache.find(key);
    if (hit != tune_cache.end()) {
      *best_variant = hit->second.first;
      *best_ms = hit->second.second;
      return FVR_OK;
    }
  }
  const uint32_t flags = (variant >> 4) & 0xF;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess) {
    set_error("fvr_autotune_pn_evaluate: hipEventCreate failed");
    return FVR_ERR_HIP;
  }
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error("fvr_autotune_pn_evaluate: hipEventCreate failed");
    return FVR_ERR_HIP;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  uint32_t best = 0;
  float best_t = 0.f;
  int rc = FVR_OK;
  const int tile_hs[3] = {32, 64, 128};
  for (int a = 0; a < 3 && rc == FVR_OK; ++a) {
    for (int depth = 2; depth <= 3 && rc == FVR_OK; ++depth) {
      if (!fvr_sncc_has_depth(flags, tile_hs[a], depth)) continue;
      fvr_sncc_item tm[2] = {tmpl_host[0], tmpl_host[1]};
      tm[0].tile_h = tm[1].tile_h = tile_hs[a];
      if (pn_ws_layout(tm, n_poses).total > ws_bytes) continue;  // this tile height needs more workspace
      for (int order = 0; order < 2 && rc == FVR_OK; ++order) {
        const uint32_t v = base | (uint32_t(tile_hs[a]) << 8) | (uint32_t(depth) << 16) | (uint32_t(order) << 21);
        char* ws = static_cast<char*>(workspace);
        rc = pn_evaluate_run(layouts_dev, n_layouts, tm, pz, py, px, lat_swap, v, poses, n_poses, t2_ap, t2_lat, ws,
                             nullptr, loss, grad, stream);  // warm-up
        if (rc == FVR_OK && hipEventRecord(e0, st) != hipSuccess) rc = FVR_ERR_HIP;
        for (int k = 0; k < reps && rc == FVR_OK; ++k)
          rc = pn_evaluate_run(layouts_dev, n_layouts, tm, pz, py, px, lat_swap, v, poses, n_poses, t2_ap, t2_lat,
                               ws, nullptr, loss, grad, stream);
        float ms = 0.f;
        if (rc == FVR_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) {
          set_error("fvr_autotune_pn_evaluate: event timing failed");
          rc = FVR_ERR_HIP;
        }
        if (rc == FVR_OK && (best == 0 || ms / reps < best_t)) {
          best = v;
          best_t = ms / reps;
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != FVR_OK) return rc;
  FVR_CHECK_ARG(best != 0, "fvr_autotune_pn_evaluate: no variant fits a workspace of %lld bytes",
                (long long)ws_bytes);
  {
    std::lock_guard<std::mutex> g(tune_mu);
    tune_cache[key] = std::make_pair(best, best_t);
  }
  *best_variant = best;
  *best_ms = best_t;
  return FVR_OK;
}

int fvr_autotune_clear(void) {
  std::lock_guard<std::mutex> g(tune_mu);
  tune_cache.clear();
  return FVR_OK;
}

int fvr_pn_step(fvr_sncc_item* items_dev, int32_t n_poses, const float* partials, double* jac_dev, double t2_ap,
                double t2_lat, double* moments, double* loss, double* grad, double* poses,
                const fvr_adam_cfg* adam_cfg_host, double* m, double* v, int32_t* t, int32_t* done, double* best_loss,
                double* best_pose, const fvr_sncc_item* tmpl_host, double pz, double py, double px, int32_t lat_swap,
                int32_t tile0, void* stream) {
  FVR_CHECK_ARG(items_dev && partials && jac_dev && loss && grad && poses && tmpl_host && n_poses > 0 && tile0 >= 0,
                "fvr_pn_step: bad arguments");
  FVR_CHECK_ARG(!adam_cfg_host || (m && v && t && done && best_loss && best_pose), "fvr_pn_step: Adam state missing");
  FVR_CHECK_ARG(tmpl_host[0].w == int(px) && tmpl_host[0].h == int(pz) && tmpl_host[1].h == int(pz) &&
                    tmpl_host[1].w == int(py),
                "fvr_pn_step: templates do not match the padded dims");
  FVR_CHECK_ARG(sncc_tile_h_ok(tmpl_host[0].tile_h) && tmpl_host[1].tile_h == tmpl_host[0].tile_h,
                "fvr_pn_step: bad template tile_h");
  const int64_t t_pose =
      int64_t(tmpl_host[0].tiles_x) * sncc_tiles_y(tmpl_host[0].h, tmpl_host[0].flags, tmpl_host[0].tile_h) +
      int64_t(tmpl_host[1].tiles_x) * sncc_tiles_y(tmpl_host[1].h, tmpl_host[1].flags, tmpl_host[1].tile_h);
  FVR_CHECK_ARG(t_pose > 0 && int64_t(tile0) + int64_t(n_poses) * t_pose < (int64_t(1) << 31),
                "fvr_pn_step: tile index overflow");
  fvr_adam_cfg c = {};
  if (adam_cfg_host) {
    c = *adam_cfg_host;
    FVR_CHECK_ARG(c.beta1 >= 0.0 && c.beta1 < 1.0 && c.beta2 >= 0.0 && c.beta2 < 1.0 && c.eps >= 0.0 &&
                      c.cosine_t >= 0,
                  "fvr_pn_step: bad Adam hyper-parameters");
  }
  const AdamState S{m, v, t, done, best_loss, best_pose};
  hipLaunchKernelGGL(pn_step_kernel, dim3(n_poses), dim3(128), 0, reinterpret_cast<hipStream_t>(stream), items_dev,
                     n_poses, partials, jac_dev, t2_ap, t2_lat, moments, loss, grad, poses, adam_cfg_host ? 1 : 0, c,
                     S, tmpl_host[0], tmpl_host[1], pz, py, px, int(lat_swap), int(tile0));
  FVR_CHECK_LAUNCH("pn_step_kernel");
  return FVR_OK;
}

int fvr_pt_build_items(const double* poses, int32_t n_poses, const fvr_sncc_item* tmpl_host, double D, double H,
                       double W, double spacing0, int32_t out_res, int32_t lateral, const int32_t* layout_order,
                       uint64_t target0, uint64_t target_stride, fvr_sncc_item* items_dev, double* jac_dev,
                       void* stream) {
  FVR_CHECK_ARG(poses && tmpl_host && layout_order && items_dev && jac_dev && n_poses > 0 && target0,
                "fvr_pt_build_items: bad arguments");
  FVR_CHECK_ARG(tmpl_host->h == out_res && tmpl_host->w == out_res && out_res > 0, "fvr_pt_build_items: bad out_res");
  FVR_CHECK_ARG(sncc_tile_h_ok(tmpl_host->tile_h), "fvr_pt_build_items: bad template tile_h");
  int seen = 0;
  for (int a = 0; a < 3; ++a) {
    FVR_CHECK_ARG(layout_order[a] >= 0 && layout_order[a] < 3, "fvr_pt_build_items: bad layout order");
    seen |= 1 << layout_order[a];
  }
  FVR_CHECK_ARG(seen == 7, "fvr_pt_build_items: layout order is not a permutation");
  const int64_t t_pose = int64_t(tmpl_host->tiles_x) * sncc_tiles_y(tmpl_host->h, tmpl_host->flags, tmpl_host->tile_h);
  FVR_CHECK_ARG(t_pose > 0 && int64_t(n_poses) * t_pose < (int64_t(1) << 31), "fvr_pt_build_items: tile overflow");
  TGeom G;
  G.D = D;
  G.H = H;
  G.W = W;
  G.spacing0 = spacing0;
  G.R = out_res;
  G.lat = lateral ? 1 : 0;
  for (int a = 0; a < 3; ++a) G.ord[a] = layout_order[a];
  G.target0 = target0;
  G.target_stride = target_stride;
  hipLaunchKernelGGL(pt_build_kernel, dim3((n_poses + 63) / 64), dim3(64), 0, reinterpret_cast<hipStream_t>(stream),
                     poses, int(n_poses), *tmpl_host, G, items_dev, jac_dev);
  FVR_CHECK_LAUNCH("pt_build_kernel");
  return FVR_OK;
}

int fvr_pt_combine(const fvr_sncc_item* items_dev, int32_t n_poses, const float* partials, const double* jac,
                   const double* e_t, int32_t et_stride, double span, double* ncc, double* grad, void* stream) {
  FVR_CHECK_ARG(items_dev && partials && jac && e_t && ncc && grad && n_poses > 0 && et_stride >= 0,
                "fvr_pt_combine: bad arguments");
  hipLaunchKernelGGL(pt_combine_kernel, dim3(n_poses), dim3(64), 0, reinterpret_cast<hipStream_t>(stream), items_dev,
                     int(n_poses), partials, jac, e_t, int(et_stride), span, ncc, grad);
  FVR_CHECK_LAUNCH("pt_combine_kernel");
  return FVR_OK;
}

int fvr_mean_loss(const double* ncc, int32_t n, float* loss, void* stream) {
  FVR_CHECK_ARG(ncc && loss && n > 0, "fvr_mean_loss: bad arguments");
  hipLaunchKernelGGL(mean_loss_kernel, dim3(1), dim3(64), 0, reinterpret_cast<hipStream_t>(stream), ncc, int(n), loss);
  FVR_CHECK_LAUNCH("mean_loss_kernel");
  return FVR_OK;
}

int fvr_adam_step(double* poses, const double* loss, const double* grad, double* m, double* v, int32_t* t,
                  int32_t* done, double* best_loss, double* best_pose, int32_t n, const fvr_adam_cfg* cfg_host,
                  void* stream) {
  FVR_CHECK_ARG(poses && loss && grad && m && v && t && done && best_loss && best_pose && cfg_host && n > 0,
                "fvr_adam_step: bad arguments");
  const fvr_adam_cfg c = *cfg_host;
  FVR_CHECK_ARG(c.beta1 >= 0.0 && c.beta1 < 1.0 && c.beta2 >= 0.0 && c.beta2 < 1.0 && c.eps >= 0.0 &&
                    c.cosine_t >= 0,
                "fvr_adam_step: bad hyper-parameters");
  const AdamState S{m, v, t, done, best_loss, best_pose};
  hipLaunchKernelGGL(adam_step_kernel, dim3((n + 63) / 64), dim3(64), 0, reinterpret_cast<hipStream_t>(stream), poses,
                     loss, grad, S, int(n), c);
  FVR_CHECK_LAUNCH("adam_step_kernel");
  return FVR_OK;
}

}  // extern "C"
