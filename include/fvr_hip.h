/*
 * fvr_hip.h -- C ABI of libfvr_hip.so, the MI355X (gfx950) kernels behind the
 * Diff-FVR rendering hot path.
 *
 * Conventions for every entry point:
 *   - device buffers are caller-owned (passed as plain device addresses);
 *   - calls are stream-ordered and asynchronous: no host synchronisation and
 *     no allocation inside a launch entry point (graph-capturable);
 *   - the return value is an fvr_status; on failure fvr_last_error() returns a
 *     thread-local message.  No C++ exception or abort crosses the ABI.
 *
 * Reference interfaces replaced (pareelamre/diff-fvr, src/fvr/):
 *   fvr_sncc_eval       diff_fvr.py:106-150 extract_frequency_slice +
 *                       diff_fvr.py:184-195 compute_spectral_ncc +
 *                       gradients.py:27-135 (loss and 6-DoF gradient), and
 *                       torch_engine.py:176-307 (_extract_slice_spectrum,
 *                       _spectral_ncc_core) with the pose adjoint that
 *                       triton_slice.py:146-339 provides to autograd.
 *   fvr_spec_pack       the storage step of diff_fvr.py:22-54 /
 *                       torch_engine.py:76-88 (centred spectrum -> device
 *                       layout, optional FP16/BF16 with one power-of-two
 *                       scale per spectrum).
 *   fvr_slice_fwd/bwd   triton_slice.py:648-661 fourier_slice_3d (forward and
 *                       the u/v gradient of triton_slice.py:306-339).
 *   fvr_pou_warp_fwd/bwd  triton_slice.py:787-808 pou_pinhole_warp_sum (forward
 *                       K3 :349-407 and the imgs/d_b/cu/cv gradients of K4
 *                       :410-496, :548-591).
 *   fvr_pou_warp_fwd_sep  the same forward (K3) for the separable monotone rays of
 *                       a zero-skew pinhole lattice, tiled through LDS.
 *   fvr_pou_slice_ifft  cone_beam_fvr.py:480-505 _slice_and_ifft (slice +
 *                       centred inverse 2D FFT, fused in LDS).
 *   fvr_pou_render_bwd  the adjoint of cone_beam_fvr.py:547-573
 *                       _project_core_invariant (slice -> iFFT -> warp) w.r.t.
 *                       u_det, v_det, d_b, cu, cv, fused per (block, pose).
 *   fvr_pou_render_bwd_sep  the same adjoint for the separable rays of a zero-skew
 *                       pinhole lattice (cone_beam_fvr.py:706-728), the warp
 *                       adjoint as a gather; fvr_pou_gprep forms its input.
 *   fvr_pou_ray_grads   the ray-array gradients grid_sample gives autograd in
 *                       cone_beam_fvr.py:423-478 _project_core.
 *   fvr_pou_pose_geom(_bwd)  cone_beam_fvr.py:670-688 _forward_from_lattice and
 *                       :507-521 _compute_block_offsets (pose -> detector axes and
 *                       block offsets) and their adjoint to the pose matrices.
 *   fvr_edge_*          cone_beam_fvr.py:71-112 edge_dominant_ncc (forward moments,
 *                       combine, and the gradient w.r.t. the images).
 *   fvr_ncc_*           tests/test_cone_beam_fvr.py:24-27 _ncc_loss (the zero-mean
 *                       NCC of whole images) and its gradient.
 */
#ifndef FVR_HIP_H
#define FVR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVR_ABI_VERSION 4

typedef enum fvr_status {
  FVR_OK = 0,
  FVR_ERR_INVALID = 1,     /* bad argument (shape, size, null pointer) */
  FVR_ERR_HIP = 2,         /* HIP runtime error (launch / copy) */
  FVR_ERR_UNSUPPORTED = 3  /* valid request this build does not implement */
} fvr_status;

/* element type of a packed spectrum */
typedef enum fvr_dtype {
  FVR_C32 = 0,  /* complex64 interleaved (re, im) */
  FVR_C16F = 1, /* fp16 (re, im) pairs, one fp32 power-of-two scale per spectrum */
  FVR_C16B = 2  /* bf16 (re, im) pairs, one fp32 power-of-two scale per spectrum */
} fvr_dtype;

/* A packed spectrum: logical dims (slow, mid, fast) stored with a one-voxel
 * zero margin below and zero margins above every axis, so that out-of-range
 * trilinear corners read 0.  Storage index = logical index + 1.
 * Storage dims: slow n0+3, mid n1+2, fast n2+2 (zero rows/planes beyond).
 * With FVR_LAYOUT_WRAP the logical index n of every axis holds the periodic
 * copy of index 0 (the DFT's own wrap, so that the stored lattice is
 * Hermitian-symmetric about n/2 for the spectrum of a real volume) and the
 * storage dims are slow n0+4, mid n1+3, fast n2+3.
 *
 * Scaled dtypes (FVR_C16F / FVR_C16B) store every element as value / sc with one
 * power-of-two scale sc per spectrum: max|component| / sc in [2^14, 2^15).  A
 * power-of-two scale does not change a float's rounding, so an element rounds as
 * it would under any other power-of-two scale that keeps it a normal fp16 number;
 * the kernels work in units of sc and apply it once per tile.  Physical dims are
 * the storage dims; row_bytes and plane_bytes are the strides. */
#define FVR_LAYOUT_WRAP 1u
typedef struct fvr_layout {
  uint64_t base;        /* device address of storage element (0,0,0) */
  uint64_t scales;      /* device address of the float scale sc (one element; 0 for FVR_C32;
                           a scaled layout with 0 here contributes zero sums) */
  int32_t n[3];         /* logical dims (slow, mid, fast) */
  int32_t dtype;        /* fvr_dtype */
  int32_t sdim[3];      /* storage dims */
  int32_t reserved[3];  /* 0 */
  uint32_t elem_bytes;  /* 8 (C32) or 4 (C16F, C16B) */
  uint32_t row_bytes;   /* physical fast-axis length * elem_bytes */
  uint32_t plane_bytes; /* physical mid-axis length * row_bytes */
  uint32_t flags;       /* FVR_LAYOUT_* */
  uint64_t total_bytes; /* physical slow-axis length * plane_bytes */
} fvr_layout;

/* Fill a layout descriptor for dims n (slow, mid, fast) and dtype (| FVR_LAYOUT_WRAP << 8
 * for the wrapped layout); returns the storage bytes and the scale-array element
 * count (1 for scaled dtypes, 0 for FVR_C32). */
int fvr_layout_init(fvr_layout* lay, const int32_t n[3], int32_t dtype,
                    uint64_t* storage_bytes, uint64_t* n_scales);

/* Pack a dense complex64 spectrum src[(z*ny + y)*nx + x] (interleaved re, im)
 * into lay's storage.  transpose_xy = 0: layout axes (z, y, x); 1: (z, x, y).
 * lay->base (and lay->scales) must point at caller buffers of the sizes
 * fvr_layout_init returned.  Writes every storage element (margins = 0). */
int fvr_spec_pack(const fvr_layout* lay, uint64_t src, int32_t nz, int32_t ny, int32_t nx,
                  int32_t transpose_xy, void* stream);

/* ------------------------------------------------------------------------ */
/* Fused Spectral-NCC forward + pose adjoint over a batch of work items.      */
/* ------------------------------------------------------------------------ */

/* item flags */
#define FVR_ITEM_ADJOINT 1u     /* accumulate the pose-adjoint moments */
#define FVR_ITEM_HP 2u          /* apply the separable high-pass weight (convention T) */
#define FVR_ITEM_WRITE_SLICE 4u /* also store the phase-modulated slice S (complex64) */
/* Hermitian fold (convention N, wrapped layout, even h <= w): the item covers the
 * rows h/2 .. h-1, each sample also standing for its mirror (h - i, w - j), plus
 * the unpaired row 0 and column 0; `target` then points at float4 (T(i,j),
 * conj T(h-i, w-j)) per sample, with a zero mirror part in row 0 and column 0 and,
 * in row h/2, zeros in columns 1 .. w/2 - 1 and a zero mirror part in column w/2.
 * Sums are those of the full slice.  Tiles are 32 columns x 2 tile_h rows over the
 * h/2 rows (tiles_x = ceil(w / 32)). */
#define FVR_ITEM_HERM 8u

/* One (pose, view) evaluation.  Sample (row i, col j) sits at
 *   p = c + j * ustep + i * vstep        (32.32 fixed point, layout axes,
 *                                         already shifted by -box_lo)
 * and is inside iff 0 <= p[a] <= box_hi[a] for every axis a; inside samples
 * read storage corners floor(p)+lo_int .. floor(p)+lo_int+1.  The phase is
 *   exp(-2 pi i * (ph_col0 + j*ph_col + ph_row0 + i*ph_row) / 2^32).
 * Moment weights are cw(j) = cw0 + j*cws and rw(i) = rw0 + i*rws. */
typedef struct fvr_sncc_item {
  int64_t c[3];
  int64_t ustep[3];
  int64_t vstep[3];
  uint64_t box_hi[3];
  uint32_t ph_col0, ph_col, ph_row0, ph_row;
  float cw0, cws, rw0, rws;
  float hp_inv2s2;      /* 1/(2 sigma^2) of the high-pass weight */
  float hp_x0, hp_dx;   /* column coordinate x(j) = hp_x0 + j*hp_dx */
  float hp_y0, hp_dy;   /* row coordinate y(i) = hp_y0 + i*hp_dy */
  int32_t h, w;         /* slice rows, cols */
  int32_t layout;       /* index into the layouts array */
  int32_t lo_int;       /* storage offset of floor(p) (1 = scipy box, 0 = per-corner) */
  uint32_t guard_off;   /* byte offset (from layout base) of a zero region */
  int32_t tile_base;    /* global index of this item's first tile */
  int32_t tiles_x;      /* tiles per slice row (ceil(w / 64); ceil(w / 32) with FVR_ITEM_HERM) */
  uint32_t flags;
  int32_t tile_h;       /* rows per tile: 32, 64 or 128 (4 waves x 8, 16 or 32 rows); one value per launch */
  uint64_t target;      /* device address of the complex64 target (h x w, row-major) */
  uint64_t slice_out;   /* device address for S (FVR_ITEM_WRITE_SLICE), else 0 */
} fvr_sncc_item;

#define FVR_SNCC_NMOM 16 /* partial sums per tile */

/* Default tile geometry of fvr_sncc_eval: 64 columns x fvr_sncc_tile_rows() rows. */
int fvr_sncc_tile_rows(void);
/* 1 when the fused kernel has a variant for (flags, tile_h): every flag set with
 * tile_h = 64; 32 and 128 for the batched forms (adjoint, without slice output). */
int fvr_sncc_has_tile_h(uint32_t flags, int32_t tile_h);

/* 1 when the Hermitian-fold kernel has a variant with `depth` samples in flight per
 * lane (2 or 3; every other kernel has depth 2 only). */
int fvr_sncc_has_depth(uint32_t flags, int32_t tile_h, int32_t depth);

/* Launch the fused kernel over n_tiles tiles (the sum of items' tile counts).
 * variant = dtype | (flags << 4) | (tile_h << 8) | (depth << 16) | (exact_box << 20)
 *           | (biplanar_order << 21)
 * (tile_h 0 = 64, depth 0 = 2: samples in flight per lane, fvr_sncc_has_depth;
 * exact_box, fold kernel only: every wave takes the per-sample box tests instead of
 * the interior / outside wave classes -- the same sums bit for bit, slower; for
 * tests; biplanar_order, no tile map: the items are (AP, LAT) pairs, item 2n + v,
 * every item of view v with the tile count of item v (as fvr_pn_build_items makes
 * them), and the workgroup slots take the tiles in target-major order -- view, tile
 * within the item, pose -- split contiguously over the XCDs, so that an XCD's L2 holds
 * one band of one view's target while it walks every pose; the sums are the same bit
 * for bit as in the default order); every item of one launch shares flags and tile_h.  Items whose layout index is not below
 * n_layouts, or whose tile_h differs, contribute zero partials.
 * partials: float[n_tiles][FVR_SNCC_NMOM] (caller buffer).
 * Tile order: workgroup slot w (XCD w % 8) takes tile (w % 8) * ceil(n_tiles/8) + w / 8,
 * i.e. each XCD a contiguous range of tiles.  Replaces the per-slice extraction
 * and NCC of diff_fvr.py:106-195 / gradients.py:76-135 for a batch of items. */
int fvr_sncc_eval_v(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* items_dev,
                    int32_t n_items, int32_t n_tiles, float* partials, uint32_t variant, void* stream);

/* fvr_sncc_eval_v with a launch plan (no counterpart in the reference: the
 * tuned batched launch of the ragged multistart).  tile_map (device, may be
 * NULL): int32 pairs (item, tile within the item) per workgroup slot, n_slots
 * entries; entries whose item is out of range are idle slots.  Without a map,
 * n_slots must be 8 * ceil(n_tiles / 8) (the XCD-contiguous order above).
 * grid: workgroups launched (0 = n_slots); a smaller grid walks the slots with
 * stride grid (persistent workgroups; keep grid a multiple of 8 so that a
 * workgroup stays on its XCD's slots). */
int fvr_sncc_eval_plan(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* items_dev,
                       int32_t n_items, int32_t n_tiles, float* partials, uint32_t variant,
                       const int32_t* tile_map, int32_t n_slots, int32_t grid, void* stream);

/* Per-item reduction of the tile partials in fixed order (float64):
 * out[item][FVR_SNCC_NMOM] as double, ordered
 *   A, E, Mc_im, Mr_im, Mc_g[3], Mr_g[3], Mc_h[3], Mr_h[3]. */
int fvr_sncc_reduce(const fvr_sncc_item* items_dev, int32_t n_items, const float* partials,
                    double* out, void* stream);

/* Convention-N pose set-up on the device (float64): 2 items per pose (AP then
 * LAT) from poses[n][6] (yaw, pitch, roll in degrees; tx, ty, tz in voxels),
 * copying the static fields of tmpl_host[2] (host structs: AP, LAT, already
 * validated) and writing jac[2n][2][3][3] (du/dtheta, dv/dtheta per degree in
 * layout axes).  lat_swap = 1 when the LAT template uses the (z, x, y) copy;
 * tile0 = tile_base of the first item (several volumes share one launch).
 * Replaces the per-pose set-up of diff_fvr.py:56-150 / gradients.py:89-95. */
int fvr_pn_build_items(const double* poses, int32_t n_poses, const fvr_sncc_item* tmpl_host, double pz,
                       double py, double px, int32_t lat_swap, int32_t tile0, fvr_sncc_item* items_dev,
                       double* jac_dev, void* stream);

/* Per-pose fixed-order float64 reduction of the fused kernel's tile partials
 * and the biplanar loss L = 1/2(1-NCC_AP) + 1/2(1-NCC_LAT) with its exact
 * 6-DoF gradient (gradients.py:76-108): loss[n], grad[n][6]; optionally the
 * item moments[2n][FVR_SNCC_NMOM] (may be NULL). */
int fvr_pn_reduce_combine(const fvr_sncc_item* items_dev, int32_t n_poses, const float* partials,
                          const double* jac, double t2_ap, double t2_lat, double* moments, double* loss,
                          double* grad, void* stream);

/* Bytes of the workspace fvr_pn_evaluate needs for n_poses poses with these two
 * templates (items, Jacobians and tile partials, each 256-byte aligned); -1 for bad
 * templates. */
int64_t fvr_pn_evaluate_workspace(const fvr_sncc_item* tmpl_host, int32_t n_poses);

/* The batched biplanar loss and exact 6-DoF gradient of n_poses poses in one call
 * (replaces the per-pose SpectralGradientEngine.compute_biplanar_loss_and_gradients
 * of gradients.py:76-135 for a batch): fvr_pn_build_items (tile0 = 0), the fused
 * kernel (variant as in fvr_sncc_eval_plan, XCD-contiguous tile order, no map) and
 * fvr_pn_reduce_combine, stream-ordered, with the items, Jacobians and tile partials
 * in the caller's workspace (device, 256-byte aligned, ws_bytes >=
 * fvr_pn_evaluate_workspace).  The templates' flags and tile_h must be the variant's
 * (with FVR_ITEM_ADJOINT).  No allocation, no host sync. */
int fvr_pn_evaluate(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* tmpl_host,
                    double pz, double py, double px, int32_t lat_swap, uint32_t variant,
                    const double* poses, int32_t n_poses, double t2_ap, double t2_lat, void* workspace,
                    int64_t ws_bytes, double* moments, double* loss, double* grad, void* stream);

/* Launch-plan autotuner of fvr_pn_evaluate.  No counterpart in the reference, whose
 * Triton launches use a fixed BLOCK = 256 (triton_slice.py:281, 312, 524, 553); it is
 * the C form of the plan search fvr.multistart.autotune makes in Python: runs fvr_pn_evaluate
 * on these poses with every (tile height 32 / 64 / 128, samples in flight 2 / 3,
 * default or biplanar tile order) variant the library has for the variant's dtype,
 * flags and exact_box bits -- each once to warm up, then `reps` times between two HIP
 * events on the stream -- and returns the fastest variant and its mean time (ms).
 * The templates may carry any valid tile height: each candidate runs on copies whose
 * tile_h is the candidate's (a caller then builds its templates with
 * (*best_variant >> 8) & 0xFF).  Candidates whose fvr_pn_evaluate_workspace exceeds
 * ws_bytes are skipped; loss / grad receive the evaluations (n_poses and 6 n_poses
 * doubles).  Results are cached per (device, dtype / flags / exact_box, padded dims,
 * slice sizes, lateral axis order, pose count) behind a mutex; with use_cache a cached
 * workload returns at once without a launch.  A set-up call: it synchronises the
 * stream and is not for graph capture. */
int fvr_autotune_pn_evaluate(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* tmpl_host,
                             double pz, double py, double px, int32_t lat_swap, uint32_t variant, const double* poses,
                             int32_t n_poses, double t2_ap, double t2_lat, void* workspace, int64_t ws_bytes,
                             double* loss, double* grad, int32_t reps, int32_t use_cache, uint32_t* best_variant,
                             float* best_ms, void* stream);

/* Empty the autotuner's cache (every device). */
int fvr_autotune_clear(void);

/* fvr_pn_evaluate without a workspace: no counterpart in the reference; it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The items, Jacobians and tile partials live in a
 * static device array of FVR_PN_SCRATCH_BYTES per GPU (held by every GPU the library
 * runs on until the process ends); a call whose fvr_pn_evaluate_workspace exceeds it is
 * rejected (FVR_ERR_INVALID): about 2,400 poses with 512^2-sample views and the
 * Hermitian fold.  The stream must belong to the current device.  Every call on one GPU
 * shares that GPU's array (it keeps the last call's contents; each call overwrites all
 * it reads): calls that overlap on one GPU -- on two streams, or from two threads --
 * race.  No allocation, no host sync. */
#define FVR_PN_SCRATCH_BYTES (16 << 20)
int fvr_pn_evaluate_noalloc(const fvr_layout* layouts_dev, int32_t n_layouts, const fvr_sncc_item* tmpl_host,
                            double pz, double py, double px, int32_t lat_swap, uint32_t variant,
                            const double* poses, int32_t n_poses, double t2_ap, double t2_lat,
                            double* moments, double* loss, double* grad, void* stream);

/* Batched Adam over 6-DoF poses on the device, one thread per pose, so that a
 * registration loop is a fixed sequence of launches (graph-capturable).  Per
 * pose, with loss / grad evaluated at the current pose (gradient_optimizer.py:45-83;
 * experiments/exp30_gpu_multistart_registration.py:140-183 for the gradient
 * clamp and the cosine schedule):
 *   if done: nothing;  t += 1;  if loss < best_loss: best = (loss, pose)
 *   g = clamp(grad, -gclip, gclip) (gclip <= 0: no clamp)
 *   m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2
 *   f = cosine_T > 0 ? 0.1 + 0.45 (1 + cos(pi min(t-1, T) / T)) : 1
 *   step = f lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)
 *   pose = clip(pose - step, lo, hi);  done = max|step| < tol. */
typedef struct fvr_adam_cfg {
  double lr[6];
  double gclip[6];
  double lo[6], hi[6];
  double beta1, beta2, eps, tol;
  int32_t cosine_t;
  int32_t pad0;
} fvr_adam_cfg;
int fvr_adam_step(double* poses, const double* loss, const double* grad, double* m, double* v,
                  int32_t* t, int32_t* done, double* best_loss, double* best_pose, int32_t n,
                  const fvr_adam_cfg* cfg_host, void* stream);

/* One launch after the fused kernel of a registration step (no counterpart in the
 * reference: it fuses fvr_pn_reduce_combine, fvr_adam_step and fvr_pn_build_items).
 * Per pose: the fixed-order reduction and combine into loss[n], grad[n][6] (and
 * moments, may be NULL) at the pose the items were built from; with adam_cfg_host,
 * one Adam step of poses[n] (state m, v, t, done, best_loss, best_pose as in
 * fvr_adam_step; NULL: poses unchanged); then the pose's two items and Jacobians for
 * the next fused launch from tmpl_host (as fvr_pn_build_items).  A step is then two
 * launches: fvr_sncc_eval_plan and this. */
int fvr_pn_step(fvr_sncc_item* items_dev, int32_t n_poses, const float* partials, double* jac_dev, double t2_ap,
                double t2_lat, double* moments, double* loss, double* grad, double* poses,
                const fvr_adam_cfg* adam_cfg_host, double* m, double* v, int32_t* t, int32_t* done,
                double* best_loss, double* best_pose, const fvr_sncc_item* tmpl_host, double pz, double py,
                double px, int32_t lat_swap, int32_t tile0, void* stream);

/* Convention-T set-up on the device (float64): one item per pose of one view from
 * poses[n][6] (yaw, pitch, roll in degrees about z, y, x; tx, ty), copying the
 * static fields of the host-validated template tmpl_host; D, H, W: padded dims;
 * layout_order[a]: xyz component of layout axis a; target of pose p at
 * target0 + p * target_stride bytes; jac[n][2][3][3] = dv/dtheta, du/dtheta (layout
 * axes, per degree).  Replaces _extract_slice_spectrum's geometry and the
 * autograd through the rotation (torch_engine.py:138-234). */
int fvr_pt_build_items(const double* poses, int32_t n_poses, const fvr_sncc_item* tmpl_host, double D, double H,
                       double W, double spacing0, int32_t out_res, int32_t lateral, const int32_t* layout_order,
                       uint64_t target0, uint64_t target_stride, fvr_sncc_item* items_dev, double* jac_dev,
                       void* stream);
/* Per pose: fixed-order reduction of the fused kernel's tiles, NCC = A / (sqrt(E E_t)
 * + 1e-8) into ncc[n] and grad[n][6] = -(1/n) d NCC / d pose, i.e. the gradient of
 * 1 - mean NCC (torch_engine.py:269-290); e_t[p * et_stride]: target energies;
 * span = out_res * spacing + 1e-6. */
int fvr_pt_combine(const fvr_sncc_item* items_dev, int32_t n_poses, const float* partials, const double* jac,
                   const double* e_t, int32_t et_stride, double span, double* ncc, double* grad, void* stream);
/* loss[0] = 1 - mean(ncc[0 .. n)) as float32, summed in float64 in a fixed order (one
 * wave): the batch loss of torch_engine.py:290 (`1.0 - torch.mean(ncc)`) in one launch. */
int fvr_mean_loss(const double* ncc, int32_t n, float* loss, void* stream);

/* ------------------------------------------------------------------------ */
/* Generic fourier_slice_3d (grid_sample semantics, planar (B,2,D,H,W)).      */
/* ------------------------------------------------------------------------ */
/* vol dtype: 0 float32, 1 float16, 2 bfloat16; out has the vol dtype. */
int fvr_slice_fwd(uint64_t vol, int32_t vol_dtype, int32_t b_vol, int32_t d, int32_t h, int32_t w,
                  const float* u_xyz, const float* v_xyz, int32_t n_pose, const float* fx,
                  const float* fy, int32_t k, float cx, float cy, float cz, uint64_t out,
                  void* stream);
/* grad_out has the vol dtype; gu/gv: float[n_pose][3]; part: float workspace of
 * fvr_slice_bwd_workspace(n_pose, k) floats. */
int64_t fvr_slice_bwd_workspace(int32_t n_pose, int32_t k);
int fvr_slice_bwd(uint64_t vol, int32_t vol_dtype, int32_t b_vol, int32_t d, int32_t h, int32_t w,
                  uint64_t grad_out, const float* u_xyz, const float* v_xyz, int32_t n_pose,
                  const float* fx, const float* fy, int32_t k, float cx, float cy, float cz,
                  float* part, float* gu, float* gv, void* stream);

/* ------------------------------------------------------------------------ */
/* PoU-FVR perspective path (cone_beam_fvr.py, triton_slice.py:349-808).      */
/* Slab images imgs: float (B, N, P, P); d_b, cu, cv: float (B, N); rays rx,   */
/* ry, ic: float (K) shared by every pose (per_pose_rays = 0) or (N, K);       */
/* detector output / its gradient: float (N, K); K = H * W.                    */
/* ------------------------------------------------------------------------ */
/* 1 when P is a power of two in [4, 128] (the LDS-resident transforms). */
int fvr_pou_lds_ok(int32_t P);
/* Whether the PoU gathers (fvr_pou_slice_ifft, fvr_pou_render_bwd, fvr_pou_render_bwd_sep)
 * compact the samples whose DFT mirror they gather explicitly into a per-workgroup queue
 * (enabled != 0, the default: one dense pass over a few hundred samples) or take a pass of
 * every thread over its own (0; also what a workgroup does when its samples overflow the
 * queue).  Process-wide; returns the previous setting.  Forward results are the same bit for
 * bit; the adjoint's sums differ only in float rounding (another summation order).  No
 * counterpart in the reference (a switch for tests and timing). */
int fvr_pou_set_mirror_queue(int32_t enabled);
/* Gather layout of the fused PoU kernels: per block (P+3) x (P+2) x (P+2) elements
 * (z, y, x; storage index = logical + 1, zero margins, one extra zero plane on the
 * slow axis), 8 bytes each: the (re, im) pair as float2 (dtype 0), or (dtype 1 f16 /
 * 2 bf16) the (re, im) words of rows y and y + 1 (16-bit re in the low half of a 32-bit
 * word, im in the high half; row P + 2 reads zero), so that one 16-byte load returns the
 * four (y, x) corners of a cell in one plane.  No
 * counterpart in the reference (its kernel reads the planar (B, 2, P, P, P) array
 * of cone_beam_fvr.py:291-294 directly); bytes for B blocks, -1 on bad arguments: */
int64_t fvr_pou_packed_bytes(int32_t dtype, int32_t B, int32_t P);
/* planar F (B, 2, P, P, P) of dtype -> packed (caller buffer of fvr_pou_packed_bytes) */
int fvr_pou_pack_blocks(uint64_t F, int32_t dtype, int32_t B, int32_t P, uint64_t packed, void* stream);
/* F: packed block spectra (fvr_pou_pack_blocks) of dtype 0 f32 / 1 f16 / 2 bf16; u, v:
 * (N, 3) unit detector axes in (x, y, z) voxel order; sample (i, j) of slice
 * (b, n) sits at P/2 + (j - P/2) u + (i - P/2) v (per-corner zero padding, as
 * grid_sample); imgs = real(fftshift(ifft2(ifftshift(slice)))).  Replaces
 * _slice_and_ifft (cone_beam_fvr.py:480-505). */
int fvr_pou_slice_ifft(uint64_t F, int32_t dtype, int32_t B, int32_t P, const float* u_xyz,
                       const float* v_xyz, int32_t N, float* imgs, void* stream);
/* out[n][q] = ic[q] * sum_b bilinear(imgs[b][n], (d_b r_x - cu) inv_ds + c_idx,
 * (d_b r_y - cv) inv_ds + c_idx), zero outside the image. */
int fvr_pou_warp_fwd(const float* imgs, int32_t B, int32_t N, int32_t P, const float* d_b,
                     const float* cu, const float* cv, const float* rx, const float* ry,
                     const float* ic, int32_t per_pose_rays, int32_t K, float inv_ds, float c_idx,
                     float* out, void* stream);
/* fvr_pou_warp_fwd for separable monotone rays shared by every pose (r_x of pixel (y, x) is
 * rx_col[x], r_y is ry_row[y], each monotone; K = H * W): per (pose, 32 x 32 tile) the slab-image
 * patch each block contributes is staged in LDS; the same projections bit for bit. */
int fvr_pou_warp_fwd_sep(const float* imgs, int32_t B, int32_t N, int32_t P, const float* d_b,
                         const float* cu, const float* cv, const float* rx_col, const float* ry_row,
                         const float* ic, int32_t H, int32_t W, float inv_ds, float c_idx, float* out,
                         void* stream);
/* scale[n] = fvr_pou_scale_of_max(max_q |go[n][q] ic[q]|): the fixed-point scale of the
 * deterministic image-gradient scatter of fvr_pou_warp_bwd / fvr_pou_render_bwd. */
int fvr_pou_gscale(const float* go, const float* ic, int32_t N, int32_t K, int32_t per_pose_rays,
                   float* scale, void* stream);
/* Host-callable: the power of two 2^min(40 - e, 126) for max_abs = m 2^e (m in [0.5, 1)),
 * 1 for 0, NaN for NaN or +inf (a pose whose output gradient is not finite gets a NaN image
 * gradient).  The same function the gscale kernel evaluates on the device. */
float fvr_pou_scale_of_max(float max_abs);
/* Bytes of the 64-bit accumulator workspace of fvr_pou_warp_bwd. */
int64_t fvr_pou_warp_bwd_workspace(int32_t B, int32_t N, int32_t P);
/* part[b][n][3] = d/d(d_b, cu, cv); grad_imgs (B, N, P, P) when non-NULL (needs
 * gscale and the workspace; accumulated in 64-bit fixed point: deterministic). */
int fvr_pou_warp_bwd(const float* imgs, int32_t B, int32_t N, int32_t P, const float* go,
                     const float* d_b, const float* cu, const float* cv, const float* rx,
                     const float* ry, const float* ic, int32_t per_pose_rays, int32_t K, float inv_ds,
                     float c_idx, const float* gscale, void* workspace, float* grad_imgs, float* part,
                     void* stream);
/* part[b][n][9] = d/d(d_b, cu, cv), d/du (3), d/dv (3) of sum go * render, for
 * imgs produced by fvr_pou_slice_ifft with the same packed F, u, v (P as
 * fvr_pou_lds_ok).  Replaces the backward of _PouRender's reference
 * composition (K4 + FFT adjoint + K2, triton_slice.py:146-271, 410-496). */
int fvr_pou_render_bwd(uint64_t F, int32_t dtype, int32_t B, int32_t P, const float* u_xyz,
                       const float* v_xyz, int32_t N, const float* imgs, const float* go,
                       const float* d_b, const float* cu, const float* cv, const float* rx,
                       const float* ry, const float* ic, int32_t per_pose_rays, int32_t K,
                       float inv_ds, float c_idx, const float* gscale, float* part, void* stream);
/* 1 when fvr_pou_render_bwd_sep takes P with an H-row detector (P as fvr_pou_lds_ok, its
 * LDS tables within one CU's 160 KiB, H * W < 2^31). */
int fvr_pou_sep_ok(int32_t P, int32_t H, int32_t W);
/* X[n][q] = go[n][q] * ic[q] (ic[n][q] when per_pose_rays) and flag[n] = 0, or NaN when
 * any X of pose n is not finite: the input of fvr_pou_render_bwd_sep. */
int fvr_pou_gprep(const float* go, const float* ic, int32_t N, int32_t K, int32_t per_pose_rays, float* X,
                  float* flag, void* stream);
/* fvr_pou_render_bwd for separable rays shared by every pose: r_x of pixel (y, x) is
 * rx_col[x], r_y is ry_row[y] (a zero-skew pinhole lattice), X and flag from
 * fvr_pou_gprep, K = H * W.  The warp adjoint is a gather (each texel of the image
 * gradient sums its detector footprint, no atomics); the same part[b][n][9].  opts: 0
 * workgroups in block-major order, 1 pose-major (the results are the same). */
int fvr_pou_render_bwd_sep(uint64_t F, int32_t dtype, int32_t B, int32_t P, const float* u_xyz,
                           const float* v_xyz, int32_t N, const float* imgs, const float* X,
                           const float* flag, const float* d_b, const float* cu, const float* cv,
                           const float* rx_col, const float* ry_row, int32_t H, int32_t W,
                           float inv_ds, float c_idx, int32_t opts, float* part, void* stream);
/* grx/gry/gic (N, K): d/d(rx, ry, ic) of sum go * out for per-pose rays. */
int fvr_pou_ray_grads(const float* imgs, int32_t B, int32_t N, int32_t P, const float* go,
                      const float* d_b, const float* cu, const float* cv, const float* rx,
                      const float* ry, const float* ic, int32_t per_pose_rays, int32_t K,
                      float inv_ds, float c_idx, float* grx, float* gry, float* gic, void* stream);
/* Pose geometry of the lattice path (cone_beam_fvr.py:670-688, 507-521): M (N, 4, 4)
 * row-major pose matrices [R t; 0 1], lat = (src0, u0, v0, n0) (12 floats), centers (B, 3);
 * src = R src0 + t, u = R u0, v = R v0, w = R n0, d_b[b][n] = (c_b - src) . w, cu, cv the
 * same with u, v; u_xyz / v_xyz (N, 3) = u, v reversed (the slice kernels' (x, y, z)). */
int fvr_pou_pose_geom(const float* M, int32_t N, const float* lat, const float* centers, int32_t B,
                      float* u_xyz, float* v_xyz, float* d_b, float* cu, float* cv, void* stream);
/* gM (N, 4, 4): the adjoint of fvr_pou_pose_geom applied to the fused adjoint's sums part
 * (B, N, 9) of fvr_pou_render_bwd(_sep), the blocks summed in fixed order in float64. */
int fvr_pou_pose_geom_bwd(const float* M, int32_t N, const float* lat, const float* centers, int32_t B,
                          const float* part, float* gM, void* stream);
/* out[n][c] = sum_b part[b][n][c0 + c] in float64, fixed order (c < nc). */
int fvr_pou_sum_blocks(const float* part, int32_t B, int32_t N, int32_t C, int32_t c0, int32_t nc,
                       float* out, void* stream);

/* Edge-dominant gradient NCC (cone_beam_fvr.py:71-112 edge_dominant_ncc) of n_img float32
 * images (H x W, row-major, one channel): Sobel x / y (zero padded) over the crop
 * [crop, H - crop) x [crop, W - crop), NCC per channel with the crop mean removed and
 * denominator sqrt(max(Eu Ev, 1e-12)); with multi = 1 also Sobel x / y of the 3 x 3
 * box-blurred image (zero padding, divisor 9), weighted (1 - w, w) against the plain pair.
 * Tiles of the crop: 32 x 32 (fvr_edge_tiles; -1 for a shape the crop leaves empty).
 *   fvr_edge_channels: channels[n][2 or 4][H - 2 crop][W - 2 crop] (the target's)
 *   fvr_edge_stats:    partials[n][tiles][12] float64 = (sum u, sum u^2, sum u v) per
 *                      channel against the channel images of image n (n_ch_img = n_img)
 *                      or of one shared image (n_ch_img = 1)
 *   fvr_edge_combine:  value[n] and coef[n][4][4] = (alpha, beta, mean u, mean v) per
 *                      channel: d value / d u = alpha (v - mean v) - beta (u - mean u);
 *                      target_partials: fvr_edge_stats of the targets against themselves
 *   fvr_edge_bwd:      grad[n][H][W] = grad_value[n] x d value[n] / d image n
 * Deterministic (fixed-order sums, no atomics). */
int32_t fvr_edge_tiles(int32_t H, int32_t W, int32_t crop);
int fvr_edge_channels(const float* img, int32_t n_img, int32_t H, int32_t W, int32_t crop, int32_t multi,
                      float* channels, void* stream);
int fvr_edge_stats(const float* img, int32_t n_img, int32_t H, int32_t W, int32_t crop, int32_t multi,
                   const float* channels, int32_t n_ch_img, double* partials, void* stream);
int fvr_edge_combine(const double* partials, int32_t n_img, const double* target_partials, int32_t n_target,
                     int32_t H, int32_t W, int32_t crop, int32_t multi, double weight, double* value,
                     float* coef, void* stream);
int fvr_edge_bwd(const float* img, int32_t n_img, int32_t H, int32_t W, int32_t crop, int32_t multi,
                 const float* channels, int32_t n_ch_img, const float* coef, const double* grad_value,
                 float* grad, void* stream);

/* Zero-mean NCC of whole images, the NCC of the PoU path's registration loss (the
 * reference's tests/test_cone_beam_fvr.py:24-27 _ncc_loss = 1 - NCC, written there in torch):
 * n_img float32 images of K pixels against n_target (1 or n_img) targets,
 * value[n] = sum(u0 v0) / (|u0| |v0| + 1e-8), u0 / v0 less their means; float64 sums in a
 * fixed order (deterministic).
 *   fvr_ncc_work_doubles: size of the work array of fvr_ncc_fwd (-1 for bad sizes)
 *   fvr_ncc_fwd:  value[n] (float64) and coef[n][8] = (alpha, beta_u, mean u, mean v,
 *                 alpha, beta_v, mean v, mean u)
 *   fvr_ncc_bwd:  grad[n][K] = grad_value[n] x d value[n] / d image n (wrt_target = 0) or
 *                 d value[n] / d its target (wrt_target = 1; summed over n by the caller
 *                 for a shared target): alpha (other - mean) - beta (self - mean) */
int64_t fvr_ncc_work_doubles(int32_t n_img, int64_t K);
int fvr_ncc_fwd(const float* img, int32_t n_img, const float* target, int32_t n_target, int64_t K, double* work,
                double* value, float* coef, void* stream);
int fvr_ncc_bwd(const float* img, int32_t n_img, const float* target, int32_t n_target, int64_t K,
                const float* coef, const double* grad_value, int32_t wrt_target, float* grad, void* stream);

/* ------------------------------------------------------------------------ */
/* misc                                                                      */
/* ------------------------------------------------------------------------ */
int fvr_abi_version(void);
const char* fvr_last_error(void);
/* out[0] = sizeof(fvr_layout), out[1] = sizeof(fvr_sncc_item), out[2] = sizeof(fvr_adam_cfg), out[3] = 0 */
int fvr_struct_sizes(int32_t out[4]);
/* out[5] = [Re sum a conj(b), sum |a|^2, sum |b|^2, sum uu Im(a conj b), sum vv Im(a conj b)]
 * over two complex64 (h x w) row-major device slices, uu = j - w/2, vv = i - h/2
 * (float64, fixed order).  Replaces diff_fvr.py:184-195 compute_spectral_ncc and
 * the sums of gradients.py:27-74 compute_spectral_ncc_and_grad_translation. */
int fvr_c64_pair_moments(uint64_t a, uint64_t b, int32_t h, int32_t w, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FVR_HIP_H */
