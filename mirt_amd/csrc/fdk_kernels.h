// fdk_kernels.h -- device-side argument blocks and launchers of the FDK
// backprojector kernels (fdk_kernels.hip).  Internal to the library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fdk {

enum { FMT_F32 = 0, FMT_F16 = 1, FMT_BF16 = 2 };
enum { KIND_FLAT = 0, KIND_ARC = 1, KIND_PAR = 2 };

// Scan geometry as the reference kernel sees it: every scalar narrowed to
// float as in fdk,mex.c:119-137, w* as in fdk,ts.c:90-94.
struct GeomF {
	int nx, ny, nz, ns, nt, na;
	float dx, dy, dz, dso, dsd, ds, dt;
	float wx, wy, wz, ws, wt;
	int kind;
	// FP16/BF16 paths only (not in the reference): 1/ds and dz/dt
	float inv_ds, dz_dt;
};

struct BackArgs {
	GeomF g;
	const void *proj;        // [na][ns][nt] (t fastest), storage format fmt
	const void *padded;      // FP32: [na][ns+2][nt+2] zero-padded copy of proj
	const float *scale;      // [na] per-view scale (F16/BF16), else unused
	const float2 *trig;      // [na] {cosf(beta), sinf(beta)} of the float beta
	const uint8_t *mask;     // [ny][nx] or null
	float *image;            // [ny][nx][nz]
	int tiles_x, tiles_y, ntiles;
	int z0, ngroups;         // z-groups cover [z0, z0 + ngroups*NW*ZL)
	int vstage, slot;        // views per LDS stage, floats per view slot
	int zper;                // slices per consumer wave (ZL) of this scan's z-groups
	unsigned long long *count;   // survey: reference loop iterations
	int *maxarea;            // survey: largest staged footprint (floats)
	int *fov_any;            // survey: some column meets some view
	// helical scans (fdk-helix, cone_par = 0); NULL for an axial scan
	const float *vzs;        // [na] source_zs narrowed to float
	const int2 *izb;         // [na] iz_min_beta, iz_max_beta (fdk,ts,t.c:78-85)
	const int2 *gviews;      // [ngroups] views [first, last) whose window meets the z-group
};

// a batch of scans for the persistent kernel: scans[i] holds scan i's
// arguments; its work items are [item_begin[i], item_begin[i+1])
struct BatchArgs {
	const BackArgs *scans;         // [nscan] in device memory
	const long long *item_begin;   // [nscan + 1]
	int nscan;
	long long nitems;
	unsigned long long *counter;   // work-queue head, reset before each launch
};

struct StArgs {
	GeomF g;
	const float *proj;       // [na][nt][ns] (s fastest)
	const float2 *trig;
	const uint8_t *mask;
	float *image;            // [nz][ny][nx]
};

// helical / cone-parallel 'fdk,ts,back' of the fdk-helix kernel
// (mex/src/fdk-helix/mex/fdk,ts.c:55-251): per-view tables as
// fdk_ts_back_init (fdk,ts,t.c:39-100) computes them
struct HelixView {
	float cosb, sinb;        // cosf/sinf of the float-narrowed beta
	float zs;                // source_zs narrowed to float
	int izb0, izb1;          // iz_min_beta, iz_max_beta (fdk,ts,t.c:78-85)
	int pad[3];
};
#ifndef HX_ZC_SLICES
#define HX_ZC_SLICES 24
#endif
constexpr int HX_ZC = HX_ZC_SLICES;   // slices per lane
struct HelixArgs {
	GeomF g;
	const float *proj;       // [na][ns][nt] (t fastest)
	const HelixView *views;  // [na]
	const int2 *chunk_views; // [nzc] views whose window reaches the slice chunk
	const float *tab;        // w3d table (fdk,w3d,pow,tab.c) and a zero tail
	int tab_len;             // readable entries of tab
	int nt2, cone_par, w3d, nzc;
	float pitch, zs_min, zs_max;
	double dz_per_rad;       // float-narrowed, as the gateway narrows it
	const uint8_t *mask;     // [ny][nx] or null
	float *image;            // [ny][nx][nz]
};

// cbct_back_mat (fbp/cbct_back.m:142-270): the MATLAB backprojector, all in
// double as MATLAB holds it.  ws and wt are 1-based (cbct_back.m:185-186).
constexpr int MAT_ZC = 16;   // slices per lane
struct MatArgs {
	int nx, ny, nz, ns, nt, nv, nzc, kind;   // nv views used, nzc = ceil(nz / MAT_ZC)
	double dx, dy, dz, wx, wy, wz, ws, wt;
	double dso, dsd, ds, dt, offset_source, scale;
	const double4 *views;    // [nv] {cos beta, sin beta, source_zs, 0}
	const void *proj;        // views of [nt][ns] (s fastest), float or double
	long view_stride;        // elements from one used view to the next
	const uint8_t *mask;     // [ny][nx] or NULL
	double *image;           // [nz][ny][nx]
	// helical MATLAB backprojector (cbct_back_helix.m): NULL / 0 otherwise
	const int2 *zrange;      // [nz] first and last view of each slice (inclusive)
	int vstep;               // a slice uses views first, first + vstep, ... <= last
	int cone_par;            // views rebinned to cone-parallel geometry
	double dz_per_rad;       // source height step per radian (cone_par)
	const double *zscale;    // [nz] per-slice scale, or NULL for `scale`
};

size_t back_lds_bytes(int nw, int vstage, int slot, int fmt);
// nw consumer waves, zl slices per chunk, zr chunks per wave (interleaved);
// the survey takes the z-group's chunk count
bool back_config_supported(int nw, int zl, int zr, int fmt);
// minw: the kernel build's minimum waves per SIMD, 4, or 8 for one-chunk
// (zr = 1) FP16/BF16 kernels (at most 64 VGPRs)
bool back_minw_supported(int minw, int zr, int fmt);
hipError_t launch_back(const BackArgs &a, int nw, int zl, int zr, int fmt, bool fast,
	int minw, hipStream_t s);
hipError_t launch_survey(const BackArgs &a, int nchunk, int zl, bool direct, hipStream_t s);
hipError_t launch_batch(const BatchArgs &b, int nwg, int nw, int zl, int vstage, int slot,
	int fmt, bool fast, hipStream_t s);
// resident persistent workgroups per CU for this shape and LDS layout
int batch_blocks_per_cu(int nw, int zl, int vstage, int slot, int fmt, bool fast);
hipError_t launch_fill(float *image, long ncol, int nz, int zlo, int zhi, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long *p, int n, hipStream_t s);
// out[b*out_b + c*out_c + r] = scale * in[b*in_b + clamp(r - rshift, 0, R-1)*in_r + c]
hipError_t launch_transpose(const float *in, float *out, int B, int R, int C, long in_b,
	long in_r, long out_b, long out_c, int Rout, int rshift, float scale, hipStream_t s);
// zero-padded copy of FP32 views: [na][ns][nt] -> [na][ns+2][nt+2]
hipError_t launch_pad(const float *src, float *dst, int ns, int nt, int na, hipStream_t s);
hipError_t launch_zero_i32(int *p, int n, hipStream_t s);
hipError_t launch_st(const StArgs &a, hipStream_t s);
hipError_t launch_quantize(const float *proj, int ns, int nt, int na, int fmt,
	void *out, float *scale, int *maxbits, hipStream_t s);
hipError_t launch_mat(const MatArgs &a, bool proj_fp64, hipStream_t s);
hipError_t launch_helix(const HelixArgs &a, hipStream_t s);

} // namespace fdk
