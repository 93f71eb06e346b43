/*
 * fdkhip.h -- C ABI of the MI355X (gfx950) FDK cone-beam backprojector.
 *
 * Drop-in for the backprojection that MIRT's fbp/cbct_back.m hands to its
 * `back_call` (default fdk_mex_call, fbp/cbct_back.m:274-283):
 *
 *   img = fdk_mex('fdk,ts,back', int32([nx ny nz]), [dx dy dz], ...
 *                 [offset_x offset_y offset_z], uint8(mask), ...
 *                 dso, dsd, dfs, [ds dt], [offset_s offset_t], ...
 *                 proj, double(betas), int32(nthread));
 *
 * Conventions are those of the reference mex gateway
 * (mex/src/fdk/fdk-orig/fdk,mex.c:22-171) and kernel
 * (mex/src/fdk-helix/mex/fdk,ts.c:55-251):
 *   - proj  is [nt ns na], t (detector row) fastest, FP32;
 *   - image is [nz nx ny], z fastest ("zxy trick"), FP32, overwritten;
 *   - mask  is [nx ny], x fastest, uint8, 0 = column not reconstructed;
 *   - beta  is [na] source angles in radians (double);
 *   - dfs = 0 selects the arc (3rd generation) detector, dfs = inf a flat
 *     one, dso or dsd = inf the parallel-beam case; any other dfs is treated
 *     as flat with a warning, like the reference ("dfs not done - junk!");
 *   - dz/dt < 0 is an error, like the reference ("todo: dz/dt < 0 not done").
 * The vertical interpolation uses the debugged weights of
 * mex/src/fdk-helix/mex/fdk,ts.c:203 (p0 = wd*p1 + wu*p2), which the MATLAB
 * path also uses (fbp/cbct_back.m:251).
 *
 * Every function returns FDKHIP_OK (0) or a negative FDKHIP_E* code; the
 * message of the last failure on the calling thread is fdkhip_last_error().
 * Calls are synchronous with respect to the host unless they take a stream,
 * in which case they are asynchronous on that stream (a hipStream_t passed as
 * void*; NULL = the null stream).
 */
#ifndef FDKHIP_H
#define FDKHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FDKHIP_OK            0
#define FDKHIP_EARG         -1  /* bad argument (reference: Fail in fdk_mex_back, fdk,mex.c:39-90) */
#define FDKHIP_EGEOM        -2  /* "todo: dz/dt < 0 not done" (fdk,ts.c:153-154) */
#define FDKHIP_ETHREADS     -3  /* "too many threads" (cbct,mask2.c:738) */
#define FDKHIP_EHIP         -4  /* HIP runtime error */
#define FDKHIP_ENOMEM       -5  /* device or host allocation failed */
#define FDKHIP_EUNSUPPORTED -6  /* valid request this build does not serve */

/* storage format of filtered projections on the device */
#define FDKHIP_F32  0
#define FDKHIP_F16  1  /* IEEE binary16 with a per-view FP32 scale */
#define FDKHIP_BF16 2  /* bfloat16 with a per-view FP32 scale */

/* arithmetic mode of the FP32 path */
#define FDKHIP_MATH_EXACT 0  /* reference operation order, no FMA: bit-for-bit */
#define FDKHIP_MATH_FAST  1  /* FMA and folded weights: ~1e-7 relative L2 */

/* Scan geometry: the operands of 'fdk,ts,back' other than proj/beta/mask.
 * Field meaning as cbct_ig / cbct_cg (mex/src/fdk/fdk-orig/cbct,def.h:17-43),
 * kept in double on the host; the device narrows to float exactly where the
 * reference gateway does (fdk,mex.c:119-137). */
typedef struct fdkhip_geom {
	int32_t nx, ny, nz;              /* image size */
	double dx, dy, dz;               /* voxel size; dy may be < 0 */
	double offset_x, offset_y, offset_z; /* image center offsets [pixels] */
	double dso, dsd, dfs;            /* distances; dfs 0 = arc, inf = flat */
	int32_t ns, nt;                  /* detector columns, rows */
	double ds, dt;                   /* detector spacing */
	double offset_s, offset_t;       /* detector offsets [pixels] */
} fdkhip_geom;

/* Message of the last failure on this thread ("" if none). */
const char *fdkhip_last_error(void);

/* Library version string. */
const char *fdkhip_version(void);

/*
 * 'fdk,ts,back' on host memory: the compute of fdk_mex_back
 * (mex/src/fdk/fdk-orig/fdk,mex.c:22-171) and fdk_ts_back_t
 * (mex/src/fdk-helix/mex/fdk,ts,t.c:107-167), with the same 12 operands.
 * proj [nt ns na] FP32 and image [nz nx ny] FP32 are host pointers; mask may
 * be NULL (all ones).  nthread is accepted for signature compatibility and
 * checked as the reference does (1..250); the GPU ignores it.
 * Copies proj to the device, backprojects on the current HIP device, copies
 * the image back.  Synchronous.  Keeps, per device, the plan of its last call
 * and its device buffers (views, image; grown, never shrunk) for the next
 * call: a call with the same geometry, beta and mask skips the plan set-up,
 * and calls allocate no device memory once the buffers are large enough.
 * Calls on one device are serialised.  At clinical_arc size that holds
 * about 0.73 GB of device memory until fdkhip_release_host_cache.
 */
int fdkhip_fdk_ts_back(
	const int32_t nxyz[3], const double dxyz[3], const double cxyz[3],
	const uint8_t *mask,
	double dso, double dsd, double dfs,
	const double dst[2], const double ost[2],
	const float *proj, int32_t ns, int32_t nt, int32_t na,
	const double *beta, int32_t nthread,
	float *image);

/* Frees the plans, device buffers and pinned host chunks that
 * fdkhip_fdk_ts_back, fdkhip_cbct_back and fdkhip_fdk_ts_back_helix keep
 * between calls, on every device; the next call sets them up again.  No counterpart
 * in the reference (which allocates per call); for callers that need the
 * device memory back. */
int fdkhip_release_host_cache(void);

/*
 * The whole use_mex = 1 path of cbct_back_mex (fbp/cbct_back.m:74-137) on host
 * buffers in MATLAB order: proj_sta is `single(proj)`, [ns nt na] (s fastest,
 * g->nt detector rows); views 1:ia_skip:na with angles beta[0], beta[ia_skip],
 * ...; the edge rows are replicated extrapolate_t times at both ends
 * (cbct_back.m:86-94, which refuses offset_t != 0 then); the result is
 * image_xyz [nx ny nz] (x fastest) times `scale` (narrowed to single, as
 * MATLAB multiplies a single array by a double scalar; the reference's scale
 * is 0.5 * deg2rad(|orbit|) / (na / ia_skip), or 1 without scale_dang).  The
 * permutes and the row replication run on the device; the backprojection is
 * the EXACT kernel of fdkhip_fdk_ts_back.  mask may be NULL.  Synchronous.
 * Keeps its last plan between calls and shares the device buffers and pinned
 * transfer chunks of fdkhip_fdk_ts_back (released by
 * fdkhip_release_host_cache); calls on one device are serialised.
 */
int fdkhip_cbct_back(const fdkhip_geom *g, const float *proj_sta, int32_t na,
	const double *beta, int32_t ia_skip, const uint8_t *mask, int32_t extrapolate_t,
	double scale, float *image_xyz);

/*
 * 'fdk,st,back' on host memory (use_mex = 3 in fbp/cbct_back.m:121-127):
 * fdk_st_back1 (mex/src/fdk/fdk-orig/fdk,st.c:38-141) over all views, proj
 * [ns nt na] (s fastest), image [nx ny nz] (x fastest), no zero padding,
 * edge samples skipped and vertical weights as in that file.  Synchronous.
 */
int fdkhip_fdk_st_back(
	const int32_t nxyz[3], const double dxyz[3], const double cxyz[3],
	const uint8_t *mask,
	double dso, double dsd, double dfs,
	const double dst[2], const double ost[2],
	const float *proj, int32_t ns, int32_t nt, int32_t na,
	const double *beta, int32_t nthread,
	float *image);

/*
 * The 19-operand 'fdk,ts,back' of the helical mex gateway
 * (mex/src/fdk-helix/mex/fdk,mex.c:22-176, 227-236) on host memory: the
 * operands of fdkhip_fdk_ts_back plus source_zs [na] (source heights),
 * na1_half (views per half turn: each view's slice window reaches the heights
 * of views ia -/+ na1_half when pitch != 0, fdk,ts,t.c:78-85), cone_par (1:
 * views rebinned to cone-parallel geometry), w3d (the exponent of the w3d
 * view weights; cone_par only), source_dz_per_rad (narrowed to float, as the
 * gateway does) and pitch (single).  Computes fdk_ts_back_t of
 * mex/src/fdk-helix/mex/fdk,ts,t.c:107-167 with every branch of fdk_ts_back1
 * (fdk,ts.c:55-251), bit for bit.  w3d table entries past the reference's
 * allocation read 0.  Synchronous.  cone_par = 0 runs on a helical plan
 * (fdkhip_plan_create_helix), kept between calls with the same operands;
 * cone_par = 1 on the dedicated helical kernel.  Shares the cached device
 * buffers and pinned transfer chunks of fdkhip_fdk_ts_back.
 */
int fdkhip_fdk_ts_back_helix(
	const int32_t nxyz[3], const double dxyz[3], const double cxyz[3],
	const uint8_t *mask,
	double dso, double dsd, double dfs,
	const double dst[2], const double ost[2],
	const float *proj, int32_t ns, int32_t nt, int32_t na,
	const double *beta, const double *source_zs, int32_t na1_half,
	int32_t cone_par, int32_t w3d, double source_dz_per_rad, float pitch,
	int32_t nthread, float *image);

/*
 * fdkhip_fdk_ts_back_helix with proj_dev ([nt ns na] FP32) and image_dev
 * ([nz nx ny] FP32) in device memory and the geometry as fdkhip_geom; beta,
 * source_zs and mask stay host pointers.  Runs on `stream` and returns when
 * the kernel has finished; kernel_ms (may be NULL) receives its time from HIP
 * events on `stream`.  No counterpart in the reference (device-resident
 * form for callers that keep views on the GPU).
 */
int fdkhip_fdk_ts_back_helix_direct(const fdkhip_geom *g, const float *proj_dev,
	int32_t na, const double *beta, const double *source_zs, int32_t na1_half,
	int32_t cone_par, int32_t w3d, double source_dz_per_rad, float pitch,
	const uint8_t *mask, float *image_dev, void *stream, float *kernel_ms);

/*
 * cbct_back_mat (fbp/cbct_back.m:142-270), the use_mex = 0 route of cbct_back:
 * the reference's MATLAB backprojector, in double precision, with its own
 * edge rules -- a sample position outside detector columns [1, ns] adds 0,
 * detector rows are clamped to [1, nt] (replication, always on) -- and its
 * weights (mag^2 for flat and parallel, dsd^2 / (r_loop^2 + y_betas^2) for
 * the arc, r_loop = x_beta - offset_source).  proj_sta is [ns nt na] (s
 * fastest), float (proj_fp64 = 0) or double (1), as MATLAB's proj(:,:,ia)
 * enters double arithmetic; views 1:ia_skip:na with angles beta[ia] and
 * source heights source_zs[ia] (NULL = 0; the helix of cbct_back.m:152-154
 * and 219).  image_xyz [nx ny nz] double = scale * embed(sum, mask) (scale is
 * 0.5 * deg2rad(|orbit|) / (na / ia_skip), or 1 without scale_dang,
 * cbct_back.m:264-266).  mask may be NULL.  Unlike MATLAB, a NaN sample
 * position adds 0 instead of failing on the index.  Synchronous.
 */
int fdkhip_cbct_back_mat(const fdkhip_geom *g, const void *proj_sta, int32_t proj_fp64,
	int32_t na, const double *beta, const double *source_zs, int32_t ia_skip,
	double offset_source, const uint8_t *mask, double scale, double *image_xyz);

/*
 * The helical cbct_back_mat (mex/src/fdk-helix/cbct_back_helix.m:185-350),
 * the use_mex = 0 route of cbct_back_helix: as fdkhip_cbct_back_mat (double
 * precision, the MATLAB edge rules and weights, views at beta[ia] and
 * source heights source_zs[ia], all na views of proj_sta [ns nt na]) except
 * that slice iz sums views slice_views[2 iz] : ia_skip : slice_views[2 iz + 1]
 * (0-based, inclusive; first > last: none), the range the file picks around
 * the view whose source height is nearest the slice (:224-236), and is
 * multiplied by slice_scale[iz] (its "der angle" scale, :341-345; NULL: 1).
 * cone_par = 1: views rebinned to cone-parallel geometry (:245-265): mag =
 * dsd / (sqrt(dso^2 - x_beta^2) - (dso - y_betas)), sprime = x_beta, source
 * height source_zs[ia] + asin(-x_beta / dso) * source_dz_per_rad, no
 * distance weight.  (The file's w3d weighting is not offered: it calls Min
 * and Max, which the reference does not define.)  image_xyz [nx ny nz]
 * double, zero outside mask (may be NULL).  Synchronous.
 */
int fdkhip_cbct_back_helix_mat(const fdkhip_geom *g, const void *proj_sta,
	int32_t proj_fp64, int32_t na, const double *beta, const double *source_zs,
	int32_t ia_skip, double offset_source, const int32_t *slice_views,
	const double *slice_scale, int32_t cone_par, double source_dz_per_rad,
	const uint8_t *mask, double *image_xyz);

/*
 * fdkhip_cbct_back_mat with proj_dev (all na views, [ns nt na]) and image_dev
 * ([nx ny nz] double) in device memory; beta, source_zs and mask stay host
 * pointers.  Runs on `stream` and returns when the kernel has finished;
 * kernel_ms (may be NULL) receives the kernel's time from HIP events recorded
 * on `stream`.  No counterpart in the reference (device-resident form of
 * cbct_back_mat for callers that keep views on the GPU).
 */
int fdkhip_cbct_back_mat_direct(const fdkhip_geom *g, const void *proj_dev,
	int32_t proj_fp64, int32_t na, const double *beta, const double *source_zs,
	int32_t ia_skip, double offset_source, const uint8_t *mask, double scale,
	double *image_dev, void *stream, float *kernel_ms);

/*
 * Device-resident 'fdk,ts,back': a plan holds the per-view tables, the
 * device mask and the launch configuration for one geometry and one view
 * list, so repeated backprojections of new projections cost one launch
 * sequence (replayed from a HIP graph after the first call).  No counterpart
 * object exists in the reference; the arithmetic is fdk_ts_back_t's.
 *
 * fmt:  FDKHIP_F32 / F16 / BF16 storage of proj in fdkhip_plan_execute.
 * math: FDKHIP_MATH_EXACT or FDKHIP_MATH_FAST (F16/BF16 always use FAST).
 * mask_host may be NULL (all ones).  beta is copied.
 */
typedef struct fdkhip_plan fdkhip_plan;

int fdkhip_plan_create(fdkhip_plan **plan, const fdkhip_geom *g,
	int32_t na, const double *beta, const uint8_t *mask_host,
	int32_t fmt, int32_t math);

/*
 * A plan for a helical scan in native helical geometry (the cone_par = 0
 * branch of fdk_ts_back1, mex/src/fdk-helix/mex/fdk,ts.c:55-251): view ia
 * has source height source_zs[ia] (wzbeta = wz + source_zs/dz, :149-151) and,
 * when pitch != 0, the slice window between the heights of views
 * ia -+ na1_half (iz_min_beta, iz_max_beta of fdk,ts,t.c:78-85).  Same
 * execution, tuning and arithmetic modes as fdkhip_plan_create (EXACT is the
 * reference bit for bit); each z-group of the launch visits only the views
 * whose window meets it.  fdkhip_fdk_ts_back_helix and its _direct form run
 * cone_par = 0 calls on such a plan.
 */
int fdkhip_plan_create_helix(fdkhip_plan **plan, const fdkhip_geom *g, int32_t na,
	const double *beta, const double *source_zs, int32_t na1_half, float pitch,
	const uint8_t *mask_host, int32_t fmt, int32_t math);

/* Backproject device projections proj_dev ([nt ns na], plan's fmt) into the
 * device image image_dev ([nz nx ny] FP32, fully overwritten).  scale_dev is
 * the [na] FP32 per-view scale for F16/BF16 (ignored for F32).  Asynchronous
 * on `stream`.  Calls that share a plan must not overlap in time. */
int fdkhip_plan_execute(fdkhip_plan *plan, const void *proj_dev,
	const float *scale_dev, float *image_dev, void *stream);

/* Same as fdkhip_plan_execute but always launches directly (no graph). */
int fdkhip_plan_execute_direct(fdkhip_plan *plan, const void *proj_dev,
	const float *scale_dev, float *image_dev, void *stream);

/* Launch configuration chosen for the plan (for reports and the tuner):
 * cfg[0] = waves per workgroup, cfg[1] = slices per lane, cfg[2] = views per
 * LDS stage, cfg[3] = z origin, cfg[4] = z groups, cfg[5] = workgroups. */
int fdkhip_plan_config(const fdkhip_plan *plan, int32_t cfg[6]);

/* Override the launch configuration: waves per workgroup and slices per lane
 * must be a pair the build supports; vstage (views per LDS stage) and
 * slot_cap (largest staged footprint, floats) are upper bounds, 0 = chosen
 * automatically.  Views whose footprint exceeds the slot gather from HBM.
 * fdkhip_plan_config shows what was taken. */
int fdkhip_plan_set_config(fdkhip_plan *plan, int32_t waves, int32_t zper,
	int32_t vstage, int32_t slot_cap);

/* Minimum waves per SIMD of the kernel build the plan launches: 4 (the
 * default; up to 128 VGPRs a lane) or 8 (up to 64), built for one-chunk
 * FP16/BF16 launch shapes only, where small LDS stages then let four 8-wave
 * workgroups share a CU.  Shapes without an 8-wave build launch the 4-wave
 * one.  fdkhip_plan_autotune chooses it with the shape; results are the
 * same bits either way.  No reference counterpart (a launch parameter). */
int fdkhip_plan_set_occupancy(fdkhip_plan *plan, int32_t waves_per_simd);

/* *waves_per_simd = the minimum waves per SIMD of the build the plan's
 * current shape launches. */
int fdkhip_plan_occupancy(const fdkhip_plan *plan, int32_t *waves_per_simd);

/* Waits for the plan's device to finish its queued work (a launch of the
 * plan may still be running on any stream), then frees the plan. */
int fdkhip_plan_destroy(fdkhip_plan *plan);

/* Number of (voxel, view) updates the reference loop performs for this plan
 * (sum over in-mask, in-FOV columns of iz_max - iz_min, fdk,ts.c:159-188),
 * computed on the device.  Synchronous. */
int fdkhip_plan_count_updates(fdkhip_plan *plan, int64_t *n_updates);

/* Time every launch shape the build supports with each views-per-stage depth
 * that fits LDS and each kernel build (`reps` launches of the backprojection
 * kernel each, over the first 128 views, on `stream`), time the four fastest
 * again over all views (plans of more than 128 views), keep the fastest for
 * this plan and remember it for later plans of the same geometry, view
 * count, format and math mode in this process.  image_dev is overwritten
 * (its contents are undefined afterwards).  *best_ms (may be NULL) receives
 * the chosen launch's average time over the views of the last round.
 * Synchronous.  No counterpart in the reference, whose only knob is the
 * caller's nthread. */
int fdkhip_plan_autotune(fdkhip_plan *plan, const void *proj_dev,
	const float *scale_dev, float *image_dev, void *stream, int32_t reps,
	float *best_ms);

/* Average device time in milliseconds of the plan's three launch steps over
 * `reps` direct executions, from HIP events recorded on `stream` between
 * them: ms[0] zero fill of slices no view reaches, ms[1] zero-padded copy of
 * the views, ms[2] the backprojection kernel.  Synchronous.  No counterpart
 * in the reference (its timing is jf,time.c wall clock around the mex). */
int fdkhip_plan_profile(fdkhip_plan *plan, const void *proj_dev,
	const float *scale_dev, float *image_dev, void *stream, int32_t reps,
	float ms[3]);

/*
 * Batches of scans: many 'fdk,ts,back' calls -- different geometries, view
 * counts and masks, one storage format -- backprojected by one persistent
 * launch whose workgroups take (scan, column tile, z group) items from a
 * device work queue, so scans of uneven size keep every CU busy, or by one
 * launch per scan in turn (fdkhip_batch_set_per_scan), each in its own launch
 * shape; either way one graph replay per execute.  Each scan
 * i is what fdkhip_plan_execute would compute for geoms[i], na[i], beta[i]
 * and mask[i] (mask may be NULL, or mask[i] NULL for all ones), with the same
 * result bits for the same math mode.  The reference backprojects one scan
 * per mex call (fbp/cbct_back.m:99-103) and has no batch object.
 */
typedef struct fdkhip_batch fdkhip_batch;

int fdkhip_batch_create(fdkhip_batch **batch, int32_t nscan,
	const fdkhip_geom *geoms, const int32_t *na, const double *const *beta,
	const uint8_t *const *mask, int32_t fmt, int32_t math);

/* proj[i], scale[i] (F16/BF16 only; scale may be NULL for F32) and image[i]
 * are scan i's device buffers as in fdkhip_plan_execute.  Asynchronous on
 * `stream`; replayed from a HIP graph while the pointers stay the same. */
int fdkhip_batch_execute(fdkhip_batch *batch, const void *const *proj,
	const float *const *scale, float *const *image, void *stream);

/* Same as fdkhip_batch_execute but always launches directly (no graph). */
int fdkhip_batch_execute_direct(fdkhip_batch *batch, const void *const *proj,
	const float *const *scale, float *const *image, void *stream);

/* cfg[0] = waves per workgroup, cfg[1] = slices per lane, cfg[2] = views per
 * LDS stage, cfg[3] = floats per view slot, cfg[4] = persistent workgroups,
 * cfg[5] = work items.  One launch per scan: cfg[0..3] = 0 and cfg[4] =
 * cfg[5] = the workgroups of all the scans' launches. */
int fdkhip_batch_config(const fdkhip_batch *batch, int32_t cfg[6]);

/* One launch per scan instead of the persistent kernel, each scan in its
 * plan's own launch shape (the default of fdkhip_plan_create, or what
 * fdkhip_plan_autotune chose earlier in this process for that geometry), the
 * launches split over the caller's stream and a second stream of the batch
 * (forked from and joined back into the caller's), so that one scan's
 * workgroups fill what another launch's last round leaves idle.
 * Every workgroup of a launch starts together and walks the views in step
 * with its neighbouring tiles, which then share their detector footprints
 * in L2; the persistent kernel's workgroups drift apart from item to item.
 * fdkhip_batch_set_config and fdkhip_batch_set_zper return to the
 * persistent kernel.  Results: the same bits as the persistent kernel for
 * EXACT.  No reference counterpart (a launch parameter of the batch). */
int fdkhip_batch_set_per_scan(fdkhip_batch *batch);

/* Launch shape for the whole batch (a pair the build supports) and an upper
 * bound on views per LDS stage (0 = largest that fits).  Each scan's slices
 * per lane are then chosen anew: zper, or zper/2 (zper >= 16) for a scan
 * whose footprint over waves*zper slices would not fit the LDS slot. */
int fdkhip_batch_set_config(fdkhip_batch *batch, int32_t waves, int32_t zper,
	int32_t vstage);

/* Per-scan slices per lane under the current launch shape: zper[i] is the
 * shape's zper or, where that is >= 16, half of it; NULL restores the
 * automatic choice of fdkhip_batch_set_config.  Results do not depend on it
 * (EXACT: same bits); only the z-group height of scan i's work items does.
 * No reference counterpart (a launch parameter of the batch object). */
int fdkhip_batch_set_zper(fdkhip_batch *batch, const int32_t *zper);

/* zper[0..nscan) = each scan's slices per lane in the current configuration. */
int fdkhip_batch_scan_zper(const fdkhip_batch *batch, int32_t *zper);

/* As fdkhip_plan_autotune, over the batch's launch shapes and stage depths
 * (the persistent kernel over the first 128 views of every scan), each shape
 * with the automatic per-scan slices per lane and, where those differ from
 * the shape's, with the shape's for every scan; then one launch per scan,
 * each scan's plan tuned alone by fdkhip_plan_autotune.  Keeps the fastest. */
int fdkhip_batch_autotune(fdkhip_batch *batch, const void *const *proj,
	const float *const *scale, float *const *image, void *stream, int32_t reps,
	float *best_ms);

/* Average device milliseconds over `reps` direct executions, from HIP events
 * on `stream`: ms[0] the per-scan zero fills and padded copies, ms[1] the
 * persistent backprojection kernel (with its queue reset).  Synchronous. */
int fdkhip_batch_profile(fdkhip_batch *batch, const void *const *proj,
	const float *const *scale, float *const *image, void *stream, int32_t reps,
	float ms[2]);

/* Sum of fdkhip_plan_count_updates over the scans. */
int fdkhip_batch_count_updates(fdkhip_batch *batch, int64_t *n_updates);

/* Waits for the device as fdkhip_plan_destroy does, then frees the batch. */
int fdkhip_batch_destroy(fdkhip_batch *batch);

/*
 * Per-view quantisation of FP32 projections on the device, the storage step
 * for the F16/BF16 paths: scale[ia] = 2^e with e the smallest integer such
 * that max|proj[:,:,ia]| / 2^e <= 2^15 (so the scaling is exact), then
 * out = round_to_nearest_even(proj / scale) in the chosen format.
 * All pointers are device pointers.  Asynchronous on `stream`.
 */
int fdkhip_quantize_views(const float *proj_dev, int32_t ns, int32_t nt,
	int32_t na, int32_t fmt, void *out_dev, float *scale_dev, void *stream);

/*
 * The FDK steps before cbct_back on device-resident views: the cosine
 * pre-weight of feldkamp.m:103-132 (feldkamp_weight1) and the zero-padded
 * ramp filter along s of fdk_filter.m:1-63, in FP32, fused into one kernel.
 * proj is MATLAB's [ns nt na] (s fastest); ww1 is the [ns nt] weight table;
 * response holds the filter's npad-point frequency response (real and even:
 * fdk_filter's H times ds, with H[k] = H[npad - k]), npad the padded row
 * length, a power of two from max(ns, 2) to 4096 (fdk_filter's is
 * 2^ceil(log2(2 ns - 1))); tsa receives the filtered views as [nt ns na]
 * (t fastest), the layout fdkhip_plan_execute takes.  Each row is weighted,
 * transformed by an FFT of npad points in LDS, multiplied by the response,
 * transformed back and cropped to ns samples.  All pointers are device
 * pointers; tsa must not alias proj.  Asynchronous on `stream`.  The
 * library keeps one table of npad twiddles per device and length for the
 * life of the process (built on the first call with that length).
 */
int fdkhip_fdk_filter_views(const float *proj_dev, int32_t ns, int32_t nt,
	int32_t na, const float *ww1_dev, const float *response_dev, int32_t npad,
	float *tsa_dev, void *stream);

/*
 * ellipsoid_proj (fbp/ellipsoid_proj.m:35-125): analytic line integrals of
 * ellipsoids for a cone-beam (or parallel-beam) scan, in FP64, written to the
 * device buffer proj_dev [ns nt na] (s fastest, MATLAB order).  s [ns] and
 * t [nt] are the detector sample positions (ct_geom's cg.s, cg.t), beta [na]
 * the source angles in radians (cg.ar), source_zs [na] the source heights
 * (NULL: 0); dso = inf is parallel beam, else dfs = inf a flat and dfs = 0 an
 * arc detector (ir_coord_cb_{flat,arc}_to_par.m; other dfs: "not done").
 * params [nell][9] are MIRT's rows [cx cy cz rx ry rz xy_angle_deg
 * z_angle_deg value] (z rotation "not done"), at most 512.  oversample > 1
 * averages No x No sub-samples per sample as ellipsoid_proj.m:44-58 (uniform
 * spacing required).  Synchronous on `stream` when it returns.
 */
int fdkhip_ellipsoid_proj(const double *s, int32_t ns, const double *t, int32_t nt,
	const double *beta, const double *source_zs, int32_t na,
	double dso, double dod, double dfs, const double *params, int32_t nell,
	int32_t oversample, double *proj_dev, void *stream);

/*
 * Allocation-free backprojection.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same result as
 * fdkhip_plan_create(g, na, beta, mask, fmt, FDKHIP_MATH_FAST) followed by
 * fdkhip_plan_execute_direct(plan, proj_dev, scale_dev, image_dev, stream)
 * (bit for bit: the same launch shape, layout and kernels), for F16/BF16
 * views only (FP32 views need a plan's padded working copy).  The per-view
 * table, the mask and the set-up counters live in a static device array of
 * FDKHIP_NOALLOC_MAX_VIEWS * 8 + FDKHIP_NOALLOC_MAX_COLUMNS + 256 bytes
 * (about 4.3 MiB) that each GPU the library runs on holds until the process
 * ends; larger scans are refused (FDKHIP_EUNSUPPORTED).  The call copies its
 * tables and runs the set-up on `stream` and waits for them; the
 * backprojection itself is asynchronous on `stream`.  Every call on a GPU
 * shares that GPU's array: calls that overlap in time on one GPU (other
 * streams, other host threads) race; calls in sequence on one stream do not.
 */
#define FDKHIP_NOALLOC_MAX_VIEWS   16384
#define FDKHIP_NOALLOC_MAX_COLUMNS (2048 * 2048)
int fdkhip_back_noalloc(const fdkhip_geom *g, int32_t na, const double *beta,
	const uint8_t *mask, int32_t fmt, const void *proj_dev, const float *scale_dev,
	float *image_dev, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FDKHIP_H */
