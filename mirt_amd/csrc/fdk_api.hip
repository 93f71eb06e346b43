// fdk_api.hip -- host side of the C ABI declared in include/fdkhip.h.
//
// Argument checks and scalar narrowing mirror the reference mex gateway
// fdk_mex_back (mex/src/fdk/fdk-orig/fdk,mex.c:22-171); the per-view sines
// and cosines are those of fdk_ts_back1 (mex/src/fdk-helix/mex/fdk,ts.c:97-98:
// sinf/cosf of the beta narrowed to float), computed here on the host with
// the same libm so that the device sees the reference's exact values.

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "fdkhip.h"
#include "fdk_kernels.h"

using namespace fdk;

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_last_error;

static int fail(int code, const char *fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	g_last_error = buf;
	return code;
}

#define HIP_TRY(expr)                                                          \
	do {                                                                       \
		hipError_t e_ = (expr);                                                \
		if (e_ != hipSuccess)                                                  \
			return fail(FDKHIP_EHIP, "%s: %s", #expr, hipGetErrorString(e_));  \
	} while (0)

// the error-recording entry for the library's other source files
int fdkhip_fail_msg(int code, const char *msg)
{
	return fail(code, "%s", msg);
}

extern "C" const char *fdkhip_last_error(void)
{
	return g_last_error.c_str();
}

extern "C" const char *fdkhip_version(void)
{
	return "fdkhip 0.1 (gfx950)";
}

// ---------------------------------------------------------------------------
// geometry

static int check_geom(const fdkhip_geom *g)
{
	if (!g)
		return fail(FDKHIP_EARG, "geometry pointer is NULL");
	if (g->nx < 0 || g->ny < 0 || g->nz < 0)
		return fail(FDKHIP_EARG, "image size must be >= 0 (nx=%d ny=%d nz=%d)",
			g->nx, g->ny, g->nz);
	if (g->ns < 0 || g->nt < 0)
		return fail(FDKHIP_EARG, "detector size must be >= 0 (ns=%d nt=%d)", g->ns, g->nt);
	if ((long long) g->ns * g->nt > (1ll << 31) || (long long) g->nx * g->ny > (1ll << 31))
		return fail(FDKHIP_EARG, "detector or image plane too large");
	return FDKHIP_OK;
}

// fdk,mex.c:119-137 (double -> float fields) and fdk,ts.c:90-94 (w*).
static GeomF narrow_geom(const fdkhip_geom *g, int na)
{
	GeomF f;
	f.nx = g->nx; f.ny = g->ny; f.nz = g->nz;
	f.ns = g->ns; f.nt = g->nt; f.na = na;
	f.dx = (float) g->dx; f.dy = (float) g->dy; f.dz = (float) g->dz;
	f.dso = (float) g->dso; f.dsd = (float) g->dsd;
	f.ds = (float) g->ds; f.dt = (float) g->dt;
	const float ox = (float) g->offset_x, oy = (float) g->offset_y, oz = (float) g->offset_z;
	const float os = (float) g->offset_s, ot = (float) g->offset_t;
	f.wx = (float) ((g->nx - 1) / 2. + ox);
	f.wy = (float) ((g->ny - 1) / 2. + oy);
	f.wz = (float) ((g->nz - 1) / 2. + oz);
	f.ws = (float) ((g->ns - 1) / 2. + os);
	f.wt = (float) ((g->nt - 1) / 2. + ot);
	f.inv_ds = (float) (1.0 / g->ds);
	f.dz_dt = (float) (g->dz / g->dt);
	const float dfs = (float) g->dfs;
	if (isinf(f.dsd) || isinf(f.dso))
		f.kind = KIND_PAR;
	else if (dfs == 0)
		f.kind = KIND_ARC;
	else {
		if (!isinf(dfs))
			fprintf(stderr, "WARN fdkhip: dfs not done - junk!\n");
		f.kind = KIND_FLAT;
	}
	return f;
}

// [na] {cosf, sinf} of the float-narrowed beta (fdk,ts,t.c:94 -> fdk,ts.c:97-98)
static std::vector<float2> trig_table(const double *beta, int na)
{
	std::vector<float2> t((size_t) na);
	for (int ia = 0; ia < na; ++ia) {
		const float b = (float) beta[ia];
		t[ia].x = cosf(b);
		t[ia].y = sinf(b);
	}
	return t;
}

// Conservative z range [lo, hi) outside of which no column of the mask meets
// any view's detector rows: t_inc = mag*dz/dt with mag in
// [dsd/(dso + r), dsd/(dso - r)], r = largest distance of a masked column
// centre from the rotation axis.  Computed in double with a 2-slice margin.
// zsh_lo, zsh_hi: range of the source heights in slices (a helix; 0 axial)
static void support_z(const fdkhip_geom *g, const uint8_t *mask, int &lo, int &hi,
	double zsh_lo = 0, double zsh_hi = 0)
{
	lo = 0;
	hi = g->nz;
	const double wx = (g->nx - 1) / 2. + g->offset_x, wy = (g->ny - 1) / 2. + g->offset_y;
	const double wz = (g->nz - 1) / 2. + g->offset_z, wt = (g->nt - 1) / 2. + g->offset_t;
	double r2 = 0;
	for (int iy = 0; iy < g->ny; ++iy)
		for (int ix = 0; ix < g->nx; ++ix) {
			if (mask && !mask[(long) iy * g->nx + ix])
				continue;
			const double x = g->dx * (ix - wx), y = g->dy * (iy - wy);
			r2 = std::max(r2, x * x + y * y);
		}
	const double r = sqrt(r2) + fabs(g->dx) + fabs(g->dy);
	double mag_lo, mag_hi;
	if (isinf(g->dsd) || isinf(g->dso)) {
		mag_lo = mag_hi = 1;
	} else {
		if (!(g->dso - r > 0))
			return;   // volume reaches the source: no useful bound
		mag_lo = g->dsd / (g->dso + r);
		mag_hi = g->dsd / (g->dso - r);
	}
	const double rate = g->dz / g->dt;
	if (!(rate > 0) || !isfinite(rate))
		return;
	double zmin = 1e300, zmax = -1e300;
	for (double mag : {mag_lo, mag_hi}) {
		const double tinc = mag * rate;
		for (double tb : {-1.0 - 1.0, (double) g->nt + 1.0}) {
			const double z = wz + (tb - wt) / tinc;
			zmin = std::min(zmin, z);
			zmax = std::max(zmax, z);
		}
	}
	if (!isfinite(zmin) || !isfinite(zmax))
		return;
	zmin += std::min(zsh_lo, zsh_hi);
	zmax += std::max(zsh_lo, zsh_hi);
	lo = (int) std::max(0.0, floor(zmin) - 2);
	hi = (int) std::min((double) g->nz, ceil(zmax) + 3);
	if (lo >= hi) {
		lo = 0;
		hi = 0;
	}
}

// ---------------------------------------------------------------------------
// plan

struct fdkhip_plan {
	fdkhip_geom geom;
	GeomF gf;
	int na = 0, fmt = FMT_F32, math = FDKHIP_MATH_EXACT;
	int device = 0;
	float2 *d_trig = nullptr;
	uint8_t *d_mask = nullptr;
	unsigned long long *d_count = nullptr;
	int *d_maxarea = nullptr;
	int *d_any = nullptr;
	void *d_padded = nullptr;     // FP32: zero-padded copy of the views (see launch_pad)
	// survey results per launch shape (nw, zl): maxarea, fov_any, n_updates
	struct Survey {
		int nw, zl, zr, maxarea, fov_any;
		long long n_updates;
	};
	std::vector<Survey> surveys;
	size_t padded_bytes = 0;
	// launch configuration
	int nw = 8, zl = 16, vstage = 8, slot = 0;
	int zr = 1;                 // chunks of zl slices per consumer wave
	int minw = 4;               // kernel build: minimum waves per SIMD (4, or 8)
	int z0 = 0, ngroups = 0, tiles_x = 0, tiles_y = 0;
	int zs_lo = 0, zs_hi = 0;
	long long n_updates = -1;
	int maxarea = 0, fov_any = 0;
	size_t lds_limit = 65536;   // per workgroup, for the current configuration
	size_t lds_max = 65536;     // device limit per workgroup
	hipStream_t survey_stream = nullptr;   // stream of the survey pass (set-up)
	// helical scan (fdk-helix, cone_par = 0): per-view source heights and
	// slice windows, and per z-group the views whose window meets it
	bool helix = false;
	int na1_half = 0;
	float pitch = 0.f;
	std::vector<int2> h_izb;
	float *d_vzs = nullptr;
	int2 *d_izb = nullptr;
	int2 *d_gviews = nullptr;
	int gviews_cap = 0;
	// graph cache
	hipGraph_t graph = nullptr;
	hipGraphExec_t gexec = nullptr;
	const void *g_proj = nullptr;
	const float *g_scale = nullptr;
	float *g_image = nullptr;
	hipStream_t cap_stream = nullptr;
};

static BackArgs make_args(const fdkhip_plan *p, const void *proj, const float *scale,
	float *image)
{
	BackArgs a;
	memset(&a, 0, sizeof a);
	a.g = p->gf;
	a.proj = proj;
	a.padded = p->d_padded;
	a.scale = scale;
	a.trig = p->d_trig;
	a.mask = p->d_mask;
	a.image = image;
	a.tiles_x = p->tiles_x;
	a.tiles_y = p->tiles_y;
	a.ntiles = p->tiles_x * p->tiles_y;
	a.z0 = p->z0;
	a.ngroups = p->ngroups;
	a.vstage = p->vstage;
	a.slot = p->slot;
	a.zper = p->zl;
	a.count = p->d_count;
	a.maxarea = p->d_maxarea;
	a.fov_any = p->d_any;
	a.vzs = p->d_vzs;
	a.izb = p->d_izb;
	a.gviews = p->helix ? p->d_gviews : nullptr;
	return a;
}

// Waits for all work on `device`: a graph exec or a buffer about to be
// destroyed may still be in use by a launch queued on any stream.
static void quiesce(int device)
{
	int cur = 0;
	if (hipGetDevice(&cur) != hipSuccess)
		return;
	if (cur != device)
		(void) hipSetDevice(device);
	(void) hipDeviceSynchronize();
	if (cur != device)
		(void) hipSetDevice(cur);
}

static void drop_graph(fdkhip_plan *p)
{
	if (p->gexec || p->graph)
		quiesce(p->device);
	if (p->gexec) (void) hipGraphExecDestroy(p->gexec);
	if (p->graph) (void) hipGraphDestroy(p->graph);
	p->gexec = nullptr;
	p->graph = nullptr;
	p->g_proj = nullptr;
	p->g_scale = nullptr;
	p->g_image = nullptr;
}

// Tiling for (nw, zl), a survey pass over it, then the LDS slot size.
// vstage_req > 0 fixes the views per stage; slot_cap > 0 caps the slot (floats)
// -- views whose footprint does not fit gather from HBM instead of LDS.
// Stage depth (views per LDS stage) and slot size (floats per staged view)
// within `lds_limit` bytes.  A requested depth is kept and the slot shrunk
// to fit; otherwise the deepest stage up to 16 views whose slot holds the
// largest footprint, but never fewer than 4 views per stage while a smaller
// slot would allow 4: below that the producer waves, not the consumers, set
// the pace.  Views whose footprint exceeds the slot gather from HBM.
static bool choose_layout(int nw, int fmt, int maxarea, int vstage_req, int slot_cap,
	size_t lds_limit, int &vstage_out, int &slot_out)
{
	int slot = (std::max(maxarea, 64) + 63) & ~63;
	if (slot_cap > 0)
		slot = std::max(64, std::min(slot, slot_cap & ~63));
	const int floor_vs = vstage_req > 0 ? vstage_req : 4;
	int vstage = vstage_req > 0 ? vstage_req : 16;
	while (vstage > floor_vs && back_lds_bytes(nw, vstage, slot, fmt) > lds_limit)
		--vstage;
	while (slot > 64 && back_lds_bytes(nw, vstage, slot, fmt) > lds_limit)
		slot -= 64;
	while (vstage > 1 && back_lds_bytes(nw, vstage, slot, fmt) > lds_limit)
		--vstage;
	if (back_lds_bytes(nw, vstage, slot, fmt) > lds_limit)
		return false;
	vstage_out = vstage;
	slot_out = slot;
	return true;
}

// zper slices per consumer wave: one chunk of zper slices where that shape is
// built, else chunks of 16 interleaved over the waves
static bool split_zper(int nw, int zper, int fmt, int &zl, int &zr)
{
	if (back_config_supported(nw, zper, 1, fmt)) {
		zl = zper;
		zr = 1;
		return true;
	}
	if (zper % 16 == 0 && back_config_supported(nw, 16, zper / 16, fmt)) {
		zl = 16;
		zr = zper / 16;
		return true;
	}
	return false;
}

// the kernel build a plan launches: its chosen minimum waves per SIMD where
// the current shape has that build, else 4
static int plan_minw(const fdkhip_plan *p)
{
	return back_minw_supported(p->minw, p->zr, p->fmt) ? p->minw : 4;
}

static int configure(fdkhip_plan *p, int nw, int zper, int vstage_req, int slot_cap)
{
	int zl = 0, zr = 0;
	if (!split_zper(nw, zper, p->fmt, zl, zr))
		return fail(FDKHIP_EUNSUPPORTED, "configuration waves=%d zper=%d not built", nw, zper);
	p->nw = nw;
	p->zl = zl;
	p->zr = zr;
	p->tiles_x = (p->geom.nx + 7) / 8;
	p->tiles_y = (p->geom.ny + 7) / 8;
	const int gz = nw * zl * zr;
	p->z0 = p->zs_lo;
	p->ngroups = (p->zs_hi > p->zs_lo) ? (p->zs_hi - p->zs_lo + gz - 1) / gz : 0;
	drop_graph(p);
	if (p->helix && p->ngroups > 0) {
		// views [first, last) whose window [izb0, izb1) meets each z-group
		if (p->ngroups > p->gviews_cap) {
			(void) hipFree(p->d_gviews);
			p->d_gviews = nullptr;
			p->gviews_cap = 0;
			HIP_TRY(hipMalloc(&p->d_gviews, sizeof(int2) * p->ngroups));
			p->gviews_cap = p->ngroups;
		}
		std::vector<int2> gv((size_t) p->ngroups, make_int2(0, 0));
		for (int k = 0; k < p->ngroups; ++k) {
			const int z0 = p->z0 + k * gz, z1 = std::min(z0 + gz, p->geom.nz);
			int va = -1, vb = -1;
			for (int ia = 0; ia < p->na; ++ia)
				if (p->h_izb[ia].x < z1 && p->h_izb[ia].y > z0) {
					if (va < 0)
						va = ia;
					vb = ia + 1;
				}
			if (va >= 0)
				gv[k] = make_int2(va, vb);
		}
		HIP_TRY(hipMemcpy(p->d_gviews, gv.data(), sizeof(int2) * p->ngroups,
			hipMemcpyHostToDevice));
	}
	p->maxarea = 0;
	p->fov_any = 0;
	p->n_updates = 0;
	bool cached = false;
	for (const auto &sv : p->surveys)
		if (sv.nw == nw && sv.zl == zl && sv.zr == zr) {
			p->maxarea = sv.maxarea;
			p->fov_any = sv.fov_any;
			p->n_updates = sv.n_updates;
			cached = true;
		}
	if (!cached && p->ngroups > 0 && p->tiles_x > 0 && p->tiles_y > 0 && p->na > 0) {
		hipStream_t ss = p->survey_stream;
		HIP_TRY(launch_zero_u64(p->d_count, 1, ss));
		HIP_TRY(launch_zero_i32(p->d_maxarea, 1, ss));
		HIP_TRY(launch_zero_i32(p->d_any, 1, ss));
		BackArgs a = make_args(p, nullptr, nullptr, nullptr);
		HIP_TRY(launch_survey(a, nw * zr, zl, p->fmt != FDKHIP_F32, ss));
		unsigned long long cnt = 0;
		HIP_TRY(hipMemcpyAsync(&cnt, p->d_count, sizeof cnt, hipMemcpyDeviceToHost, ss));
		HIP_TRY(hipMemcpyAsync(&p->maxarea, p->d_maxarea, sizeof(int), hipMemcpyDeviceToHost,
			ss));
		HIP_TRY(hipMemcpyAsync(&p->fov_any, p->d_any, sizeof(int), hipMemcpyDeviceToHost, ss));
		HIP_TRY(hipStreamSynchronize(ss));
		p->n_updates = (long long) cnt;
		p->surveys.push_back({nw, zl, zr, p->maxarea, p->fov_any, p->n_updates});
	}
	// views per stage and slot size within the LDS budget: 8 consumer waves
	// (+8 producers) fill a CU with one workgroup, 4 (+4) fit two
	p->lds_limit = (nw >= 8) ? p->lds_max : p->lds_max / 2;
	int vstage = 0, slot = 0;
	if (!choose_layout(nw * zr, p->fmt, p->maxarea, vstage_req, slot_cap, p->lds_limit, vstage,
			slot))
		return fail(FDKHIP_EUNSUPPORTED, "LDS budget too small for waves=%d", nw);
	p->vstage = vstage;
	p->slot = slot;
	return FDKHIP_OK;
}

static int plan_free(fdkhip_plan *p)
{
	if (!p)
		return FDKHIP_OK;
	quiesce(p->device);
	drop_graph(p);
	if (p->cap_stream) (void) hipStreamDestroy(p->cap_stream);
	(void) hipFree(p->d_trig);
	(void) hipFree(p->d_mask);
	(void) hipFree(p->d_count);
	(void) hipFree(p->d_maxarea);
	(void) hipFree(p->d_any);
	(void) hipFree(p->d_padded);
	(void) hipFree(p->d_vzs);
	(void) hipFree(p->d_izb);
	(void) hipFree(p->d_gviews);
	delete p;
	return FDKHIP_OK;
}

static int apply_tuned(fdkhip_plan *p);
static int plan_default_config(fdkhip_plan *p);

constexpr size_t kPadTail = 256;   // zero bytes after a plan's padded copy

// helical scan of a plan: source heights and the window half-width
struct HelixPlanInfo {
	const double *source_zs;
	int na1_half;
	float pitch;
};

static int plan_create_impl(fdkhip_plan **out, const fdkhip_geom *g, int32_t na,
	const double *beta, const uint8_t *mask_host, int32_t fmt, int32_t math,
	const HelixPlanInfo *hx)
{
	if (!out)
		return fail(FDKHIP_EARG, "plan output pointer is NULL");
	*out = nullptr;
	int rc = check_geom(g);
	if (rc)
		return rc;
	if (na < 0)
		return fail(FDKHIP_EARG, "na must be >= 0 (na=%d)", na);
	if (na > 0 && !beta)
		return fail(FDKHIP_EARG, "beta is NULL");
	if (fmt != FDKHIP_F32 && fmt != FDKHIP_F16 && fmt != FDKHIP_BF16)
		return fail(FDKHIP_EARG, "unknown projection format %d", fmt);
	if (math != FDKHIP_MATH_EXACT && math != FDKHIP_MATH_FAST)
		return fail(FDKHIP_EARG, "unknown math mode %d", math);
	if (g->nz > 65535)
		return fail(FDKHIP_EUNSUPPORTED, "nz=%d: at most 65535 slices per volume", g->nz);

	if (hx && na > 0 && !hx->source_zs)
		return fail(FDKHIP_EARG, "source_zs is NULL");
	if (hx && hx->na1_half < 0)
		return fail(FDKHIP_EARG, "na1_half must be >= 0");

	fdkhip_plan *p = new fdkhip_plan;
	p->geom = *g;
	p->na = na;
	p->fmt = fmt;
	p->math = (fmt == FDKHIP_F32) ? math : FDKHIP_MATH_FAST;
	p->gf = narrow_geom(g, na);
	double zs_lo = 0, zs_hi = 0;   // source height range in slices
	std::vector<float> vzs;
	if (hx) {
		// fdk,ts,t.c:70-85: heights narrowed to float; the window of view ia
		// spans the slices between the heights of views ia -+ na1_half
		p->helix = true;
		p->na1_half = hx->na1_half;
		p->pitch = hx->pitch;
		vzs.resize((size_t) std::max(na, 1), 0.f);
		p->h_izb.assign((size_t) std::max(na, 1), make_int2(0, g->nz));
		const float wz = p->gf.wz, dz = p->gf.dz;
		for (int ia = 0; ia < na; ++ia) {
			vzs[ia] = (float) hx->source_zs[ia];
			if (hx->pitch != 0.f) {
				const int a0 = std::max(ia - hx->na1_half, 0);
				const int a1 = std::min(ia + hx->na1_half, na - 1);
				const float c0 = ceilf((float) (wz + hx->source_zs[a0] / dz));
				const float c1 = ceilf((float) (wz + hx->source_zs[a1] / dz));
				p->h_izb[ia] = make_int2((int) (c0 > 0 ? c0 : 0),
					(int) (c1 < g->nz ? c1 : g->nz));
			}
			const double zsl = hx->source_zs[ia] / g->dz;
			zs_lo = ia ? std::min(zs_lo, zsl) : zsl;
			zs_hi = ia ? std::max(zs_hi, zsl) : zsl;
		}
	}
	if (hipGetDevice(&p->device) != hipSuccess) {
		delete p;
		return fail(FDKHIP_EHIP, "no current HIP device");
	}
	int lds = 0;
	if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device)
		== hipSuccess && lds > 0)
		p->lds_max = (size_t) lds;

#define PLAN_TRY(expr)                                                          \
	do {                                                                        \
		hipError_t e_ = (expr);                                                 \
		if (e_ != hipSuccess) {                                                 \
			int c_ = fail(FDKHIP_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
			plan_free(p);                                                       \
			return c_;                                                          \
		}                                                                       \
	} while (0)

	if (na > 0) {
		std::vector<float2> trig = trig_table(beta, na);
		PLAN_TRY(hipMalloc(&p->d_trig, sizeof(float2) * na));
		PLAN_TRY(hipMemcpy(p->d_trig, trig.data(), sizeof(float2) * na, hipMemcpyHostToDevice));
		if (hx) {
			PLAN_TRY(hipMalloc(&p->d_vzs, sizeof(float) * na));
			PLAN_TRY(hipMemcpy(p->d_vzs, vzs.data(), sizeof(float) * na, hipMemcpyHostToDevice));
			PLAN_TRY(hipMalloc(&p->d_izb, sizeof(int2) * na));
			PLAN_TRY(hipMemcpy(p->d_izb, p->h_izb.data(), sizeof(int2) * na,
				hipMemcpyHostToDevice));
		}
	}
	const long nxy = (long) g->nx * g->ny;
	if (mask_host && nxy > 0) {
		PLAN_TRY(hipMalloc(&p->d_mask, nxy));
		PLAN_TRY(hipMemcpy(p->d_mask, mask_host, nxy, hipMemcpyHostToDevice));
	}
	PLAN_TRY(hipMalloc(&p->d_count, sizeof(unsigned long long)));
	PLAN_TRY(hipMalloc(&p->d_maxarea, sizeof(int)));
	PLAN_TRY(hipMalloc(&p->d_any, sizeof(int)));
	{
		// FP32: [na][ns+2][nt+2] floats; FP16/BF16 views are staged straight
		// from the caller's buffer and need no copy
		// plus a zero tail: row nt + 2 of the last padded column (read, as
		// a zero row, by slices at t_bin >= nt) lies just past the copy
		p->padded_bytes = (fmt == FDKHIP_F32)
			? (size_t) na * (g->ns + 2) * (g->nt + 2) * sizeof(float) : 0;
		if (p->padded_bytes > 0) {
			PLAN_TRY(hipMalloc(&p->d_padded, p->padded_bytes + kPadTail));
			PLAN_TRY(hipMemset((char *) p->d_padded + p->padded_bytes, 0, kPadTail));
		}
	}
#undef PLAN_TRY

	support_z(g, mask_host, p->zs_lo, p->zs_hi, zs_lo, zs_hi);
	rc = plan_default_config(p);
	if (rc) {
		plan_free(p);
		return rc;
	}
	// fdk,ts.c:153-154: the reference fails on the first in-FOV column
	if (p->gf.dz / p->gf.dt < 0 && p->fov_any) {
		plan_free(p);
		return fail(FDKHIP_EGEOM, "todo: dz/dt < 0 not done");
	}
	*out = p;
	return FDKHIP_OK;
}

extern "C" int fdkhip_plan_create(fdkhip_plan **out, const fdkhip_geom *g,
	int32_t na, const double *beta, const uint8_t *mask_host, int32_t fmt, int32_t math)
{
	return plan_create_impl(out, g, na, beta, mask_host, fmt, math, nullptr);
}

extern "C" int fdkhip_plan_create_helix(fdkhip_plan **out, const fdkhip_geom *g,
	int32_t na, const double *beta, const double *source_zs, int32_t na1_half, float pitch,
	const uint8_t *mask_host, int32_t fmt, int32_t math)
{
	const HelixPlanInfo hx{source_zs, na1_half, pitch};
	return plan_create_impl(out, g, na, beta, mask_host, fmt, math, &hx);
}

extern "C" int fdkhip_plan_destroy(fdkhip_plan *p)
{
	return plan_free(p);
}

extern "C" int fdkhip_plan_config(const fdkhip_plan *p, int32_t cfg[6])
{
	if (!p || !cfg)
		return fail(FDKHIP_EARG, "NULL argument");
	cfg[0] = p->nw;
	cfg[1] = p->zl * p->zr;
	cfg[2] = p->vstage;
	cfg[3] = p->z0;
	cfg[4] = p->ngroups;
	cfg[5] = p->tiles_x * p->tiles_y * p->ngroups;
	return FDKHIP_OK;
}

extern "C" int fdkhip_plan_set_config(fdkhip_plan *p, int32_t waves, int32_t zper,
	int32_t vstage, int32_t slot_cap)
{
	if (!p)
		return fail(FDKHIP_EARG, "plan is NULL");
	if (vstage < 0 || vstage > 64)
		return fail(FDKHIP_EARG, "views per stage out of range (%d)", vstage);
	if (slot_cap < 0)
		return fail(FDKHIP_EARG, "slot cap must be >= 0 (%d)", slot_cap);
	return configure(p, waves, zper, vstage, slot_cap);
}

extern "C" int fdkhip_plan_set_occupancy(fdkhip_plan *p, int32_t waves_per_simd)
{
	if (!p)
		return fail(FDKHIP_EARG, "plan is NULL");
	if (waves_per_simd != 4 && waves_per_simd != 8)
		return fail(FDKHIP_EARG, "waves per SIMD must be 4 or 8 (%d)", waves_per_simd);
	if (p->minw != waves_per_simd)
		drop_graph(p);
	p->minw = waves_per_simd;
	return FDKHIP_OK;
}

extern "C" int fdkhip_plan_occupancy(const fdkhip_plan *p, int32_t *waves_per_simd)
{
	if (!p || !waves_per_simd)
		return fail(FDKHIP_EARG, "NULL argument");
	*waves_per_simd = plan_minw(p);
	return FDKHIP_OK;
}

extern "C" int fdkhip_plan_count_updates(fdkhip_plan *p, int64_t *n)
{
	if (!p || !n)
		return fail(FDKHIP_EARG, "NULL argument");
	*n = p->n_updates;
	return FDKHIP_OK;
}

// The plan's launch sequence: zero fill of the slices no view reaches, the
// zero-padded copy of the views, the tiled backprojection.  With `ev`, the
// events ev[0..3] are recorded before the fill and after each step.
static int plan_launch(fdkhip_plan *p, const void *proj, const float *scale, float *image,
	hipStream_t s, hipEvent_t *ev = nullptr)
{
	const long ncol = (long) p->geom.nx * p->geom.ny;
	const int zend = std::min(p->z0 + p->ngroups * p->nw * p->zl * p->zr, p->geom.nz);
	const int zbeg = p->ngroups > 0 ? p->z0 : 0;
	const bool work = p->ngroups > 0 && ncol > 0 && p->na > 0;
	if (ev)
		HIP_TRY(hipEventRecord(ev[0], s));
	if (ncol > 0 && p->geom.nz > 0)   // no views: the z-groups are zeroed too
		HIP_TRY(launch_fill(image, ncol, p->geom.nz, work ? zbeg : 0, work ? zend : 0, s));
	if (ev)
		HIP_TRY(hipEventRecord(ev[1], s));
	if (work && p->fmt == FDKHIP_F32)
		HIP_TRY(launch_pad(static_cast<const float *>(proj), static_cast<float *>(p->d_padded),
			p->geom.ns, p->geom.nt, p->na, s));
	if (ev)
		HIP_TRY(hipEventRecord(ev[2], s));
	if (work) {
		BackArgs a = make_args(p, proj, scale, image);
		HIP_TRY(launch_back(a, p->nw, p->zl, p->zr, p->fmt, p->math == FDKHIP_MATH_FAST,
			plan_minw(p), s));
	}
	if (ev)
		HIP_TRY(hipEventRecord(ev[3], s));
	return FDKHIP_OK;
}

static int check_exec(fdkhip_plan *p, const void *proj, const float *scale, float *image)
{
	if (!p)
		return fail(FDKHIP_EARG, "plan is NULL");
	const long nvox = (long) p->geom.nx * p->geom.ny * p->geom.nz;
	if (nvox > 0 && !image)
		return fail(FDKHIP_EARG, "image is NULL");
	if (p->na > 0 && (long) p->geom.ns * p->geom.nt > 0 && !proj)
		return fail(FDKHIP_EARG, "proj is NULL");
	if (p->fmt != FDKHIP_F32 && p->na > 0 && !scale)
		return fail(FDKHIP_EARG, "per-view scale is required for F16/BF16 projections");
	return FDKHIP_OK;
}

extern "C" int fdkhip_plan_execute_direct(fdkhip_plan *p, const void *proj,
	const float *scale, float *image, void *stream)
{
	int rc = check_exec(p, proj, scale, image);
	if (rc)
		return rc;
	return plan_launch(p, proj, scale, image, (hipStream_t) stream);
}

extern "C" int fdkhip_plan_execute(fdkhip_plan *p, const void *proj, const float *scale,
	float *image, void *stream)
{
	int rc = check_exec(p, proj, scale, image);
	if (rc)
		return rc;
	if (!p->gexec || p->g_proj != proj || p->g_scale != scale || p->g_image != image) {
		drop_graph(p);
		if (!p->cap_stream)
			HIP_TRY(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
		HIP_TRY(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
		rc = plan_launch(p, proj, scale, image, p->cap_stream);
		hipGraph_t graph = nullptr;
		hipError_t e = hipStreamEndCapture(p->cap_stream, &graph);
		if (rc)
			return rc;
		if (e != hipSuccess)
			return fail(FDKHIP_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
		p->graph = graph;
		HIP_TRY(hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0));
		p->g_proj = proj;
		p->g_scale = scale;
		p->g_image = image;
	}
	HIP_TRY(hipGraphLaunch(p->gexec, (hipStream_t) stream));
	return FDKHIP_OK;
}

// ---------------------------------------------------------------------------
// backprojection of FP16/BF16 views without any device allocation

// Device scratch of fdkhip_back_noalloc: the per-view {cos, sin} table, the
// mask and the survey's three counters.  A file-scope device array: every
// GPU the library runs on holds its own copy until the process ends.  Its
// name is a host-side stand-in in host code; kernels and copies use the
// device address hipGetSymbolAddress returns (once per device, below).
constexpr size_t NOALLOC_TRIG = (size_t) FDKHIP_NOALLOC_MAX_VIEWS * sizeof(float2);
constexpr size_t NOALLOC_MASK = (size_t) FDKHIP_NOALLOC_MAX_COLUMNS;
constexpr size_t NOALLOC_COUNTERS = 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_TRIG + NOALLOC_MASK
	+ NOALLOC_COUNTERS];

static std::mutex g_scratch_mu;
static unsigned char *g_scratch[64];   // device address of `scratch`, per device id

// the device address of `scratch` on the current device `dev`, looked up on
// the first call on that device
static int scratch_address(int dev, unsigned char **out)
{
	if (dev < 0 || dev >= 64)
		return fail(FDKHIP_EUNSUPPORTED, "device id %d out of range", dev);
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch[dev]) {
		void *ptr = nullptr;
		const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ptr)
			return fail(FDKHIP_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
		g_scratch[dev] = static_cast<unsigned char *>(ptr);
	}
	*out = g_scratch[dev];
	return FDKHIP_OK;
}

extern "C" int fdkhip_back_noalloc(const fdkhip_geom *g, int32_t na, const double *beta,
	const uint8_t *mask_host, int32_t fmt, const void *proj_dev, const float *scale_dev,
	float *image_dev, void *stream)
{
	int rc = check_geom(g);
	if (rc)
		return rc;
	if (fmt != FDKHIP_F16 && fmt != FDKHIP_BF16)
		return fail(FDKHIP_EUNSUPPORTED, "fdkhip_back_noalloc takes F16 or BF16 views "
			"(FP32 views need the padded working copy of a plan)");
	if (na < 0)
		return fail(FDKHIP_EARG, "na must be >= 0 (na=%d)", na);
	if (na > FDKHIP_NOALLOC_MAX_VIEWS)
		return fail(FDKHIP_EUNSUPPORTED, "na=%d: at most %d views without allocation", na,
			FDKHIP_NOALLOC_MAX_VIEWS);
	if ((long long) g->nx * g->ny > FDKHIP_NOALLOC_MAX_COLUMNS)
		return fail(FDKHIP_EUNSUPPORTED, "nx*ny=%lld: at most %d columns without allocation",
			(long long) g->nx * g->ny, FDKHIP_NOALLOC_MAX_COLUMNS);
	if (g->nz > 65535)
		return fail(FDKHIP_EUNSUPPORTED, "nz=%d: at most 65535 slices per volume", g->nz);
	if (na > 0 && !beta)
		return fail(FDKHIP_EARG, "beta is NULL");
	const long nvox = (long) g->nx * g->ny * g->nz;
	if (nvox > 0 && !image_dev)
		return fail(FDKHIP_EARG, "image is NULL");
	if (na > 0 && (long) g->ns * g->nt > 0 && (!proj_dev || !scale_dev))
		return fail(FDKHIP_EARG, "proj and the per-view scale are required");
	if (nvox == 0)
		return FDKHIP_OK;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess)
		return fail(FDKHIP_EHIP, "no current HIP device");
	unsigned char *sp = nullptr;
	rc = scratch_address(dev, &sp);
	if (rc)
		return rc;
	hipStream_t s = (hipStream_t) stream;

	// a plan on the stack whose device tables live in the scratch
	fdkhip_plan p;
	p.geom = *g;
	p.na = na;
	p.fmt = fmt;
	p.math = FDKHIP_MATH_FAST;
	p.gf = narrow_geom(g, na);
	p.device = dev;
	int lds = 0;
	if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)
		== hipSuccess && lds > 0)
		p.lds_max = (size_t) lds;
	p.d_trig = reinterpret_cast<float2 *>(sp);
	p.d_mask = mask_host ? sp + NOALLOC_TRIG : nullptr;
	p.d_count = reinterpret_cast<unsigned long long *>(sp + NOALLOC_TRIG + NOALLOC_MASK);
	p.d_maxarea = reinterpret_cast<int *>(sp + NOALLOC_TRIG + NOALLOC_MASK + 64);
	p.d_any = reinterpret_cast<int *>(sp + NOALLOC_TRIG + NOALLOC_MASK + 128);
	p.survey_stream = s;
	// the tables, ordered after any earlier call's kernels on this stream
	// that still read them; waited for before the host copies go away
	const std::vector<float2> trig = na > 0 ? trig_table(beta, na) : std::vector<float2>();
	if (na > 0)
		HIP_TRY(hipMemcpyAsync(p.d_trig, trig.data(), sizeof(float2) * na,
			hipMemcpyHostToDevice, s));
	if (mask_host)
		HIP_TRY(hipMemcpyAsync(p.d_mask, mask_host, (size_t) g->nx * g->ny,
			hipMemcpyHostToDevice, s));
	HIP_TRY(hipStreamSynchronize(s));
	// the launch shape and layout a plan of this scan would use
	support_z(g, mask_host, p.zs_lo, p.zs_hi);
	rc = configure(&p, p.zs_hi - p.zs_lo <= 64 ? 4 : 8, 16, 0, 0);
	if (rc == FDKHIP_OK)
		rc = apply_tuned(&p);
	if (rc)
		return rc;
	if (p.gf.dz / p.gf.dt < 0 && p.fov_any)
		return fail(FDKHIP_EGEOM, "todo: dz/dt < 0 not done");
	return plan_launch(&p, proj_dev, scale_dev, image_dev, s);
}

// ---------------------------------------------------------------------------
// quantisation

extern "C" int fdkhip_quantize_views(const float *proj, int32_t ns, int32_t nt, int32_t na,
	int32_t fmt, void *out, float *scale, void *stream)
{
	if (fmt != FDKHIP_F16 && fmt != FDKHIP_BF16)
		return fail(FDKHIP_EARG, "quantisation format must be F16 or BF16 (got %d)", fmt);
	if (ns < 0 || nt < 0 || na < 0)
		return fail(FDKHIP_EARG, "negative size");
	if (na == 0 || (long) ns * nt == 0)
		return FDKHIP_OK;
	if (!proj || !out || !scale)
		return fail(FDKHIP_EARG, "NULL device pointer");
	hipStream_t s = (hipStream_t) stream;
	int *maxbits = nullptr;
	HIP_TRY(hipMallocAsync((void **) &maxbits, sizeof(int) * na, s));
	hipError_t e = launch_quantize(proj, ns, nt, na, fmt, out, scale, maxbits, s);
	(void) hipFreeAsync(maxbits, s);
	if (e != hipSuccess)
		return fail(FDKHIP_EHIP, "quantize: %s", hipGetErrorString(e));
	return FDKHIP_OK;
}

// ---------------------------------------------------------------------------
// host-memory drop-ins for 'fdk,ts,back' and 'fdk,st,back'

// fdk,mex.c:39-91 + cbct,mask2.c:738 checks shared by both commands
static int check_mex_args(const int32_t *nxyz, const double *dxyz, const double *cxyz,
	const double *dst, const double *ost, int32_t ns, int32_t nt, int32_t na,
	const double *beta, int32_t nthread, const float *proj, const float *image)
{
	if (!nxyz) return fail(FDKHIP_EARG, "mx_nxyz must have 3 elements");
	if (!dxyz) return fail(FDKHIP_EARG, "mx_dxyz must have 3 elements");
	if (!cxyz) return fail(FDKHIP_EARG, "mx_cxyz must have 3 elements");
	if (!dst) return fail(FDKHIP_EARG, "mx_dst must have 2 elements");
	if (!ost) return fail(FDKHIP_EARG, "mx_ost must have 2 elements");
	if (nthread > 250) return fail(FDKHIP_ETHREADS, "too many threads");
	if (nthread < 1) return fail(FDKHIP_EARG, "nthread must be >= 1 (got %d)", nthread);
	if (na < 0 || ns < 0 || nt < 0) return fail(FDKHIP_EARG, "negative projection size");
	if (na > 0 && !beta) return fail(FDKHIP_EARG, "beta length=0 vs na=size(proj,3)=%d", na);
	if (na > 0 && (long) ns * nt > 0 && !proj) return fail(FDKHIP_EARG, "proj is NULL");
	if ((long) nxyz[0] * nxyz[1] * nxyz[2] > 0 && !image)
		return fail(FDKHIP_EARG, "image is NULL");
	return FDKHIP_OK;
}

static fdkhip_geom mex_geom(const int32_t *nxyz, const double *dxyz, const double *cxyz,
	double dso, double dsd, double dfs, const double *dst, const double *ost,
	int32_t ns, int32_t nt)
{
	fdkhip_geom g;
	g.nx = nxyz[0]; g.ny = nxyz[1]; g.nz = nxyz[2];
	g.dx = dxyz[0]; g.dy = dxyz[1]; g.dz = dxyz[2];
	g.offset_x = cxyz[0]; g.offset_y = cxyz[1]; g.offset_z = cxyz[2];
	g.dso = dso; g.dsd = dsd; g.dfs = dfs;
	g.ns = ns; g.nt = nt;
	g.ds = dst[0]; g.dt = dst[1];
	g.offset_s = ost[0]; g.offset_t = ost[1];
	return g;
}

// RAII device buffer for the synchronous entry points
struct DevBuf {
	void *p = nullptr;
	~DevBuf() { if (p) (void) hipFree(p); }
};

// The host-array entries (fdkhip_fdk_ts_back, fdkhip_cbct_back,
// fdkhip_fdk_ts_back_helix) keep, per
// device, the plan of each one's last call and their device buffers (grown,
// never shrunk) until fdkhip_release_host_cache or the end of the process:
// a call with the same geometry, view angles and mask as that entry's call
// before skips the plan set-up (tables, surveys, the padded copy's
// allocation), and no call allocates device memory once the buffers are
// large enough.  Calls on one device are serialised by its mutex.  Views and
// image move through two pinned host chunks: host threads copy one chunk
// while the DMA engine moves the other, so that the caller's pageable
// buffers are neither registered nor copied by the runtime's own staging.
namespace {

constexpr size_t kStageChunk = (size_t) 32 << 20;   // bytes per pinned chunk
enum { kEntryTsBack = 0, kEntryCbctBack = 1, kEntryHelix = 2, kEntries = 3, kCacheBufs = 4 };

struct CachedPlan {
	fdkhip_plan *plan = nullptr;
	std::vector<unsigned char> key;    // geometry, na, beta and mask of `plan`
};

struct HostCache {
	std::mutex mu;
	CachedPlan plans[kEntries];
	void *buf[kCacheBufs] = {};        // device buffers
	size_t cap[kCacheBufs] = {};       // their sizes in bytes
	char *pinned = nullptr;            // 2 x kStageChunk, hipHostMalloc
	hipStream_t stream = nullptr;
	hipEvent_t ev[2] = {nullptr, nullptr};
};

constexpr int kCacheDevices = 64;
HostCache g_host_cache[kCacheDevices];

void cache_key(std::vector<unsigned char> &k, const fdkhip_geom &g, int32_t na,
	const double *beta, const uint8_t *mask, const std::vector<unsigned char> &extra = {})
{
	const size_t nxy = (size_t) g.nx * g.ny;
	k.resize(sizeof g + sizeof na + sizeof(double) * (size_t) na + 1 + (mask ? nxy : 0)
		+ extra.size());
	unsigned char *q = k.data();
	memcpy(q, &g, sizeof g);
	q += sizeof g;
	memcpy(q, &na, sizeof na);
	q += sizeof na;
	if (na > 0)
		memcpy(q, beta, sizeof(double) * (size_t) na);
	q += sizeof(double) * (size_t) na;
	*q++ = mask ? 1 : 0;
	if (mask)
		memcpy(q, mask, nxy);
	q += mask ? nxy : 0;
	if (!extra.empty())
		memcpy(q, extra.data(), extra.size());
}

// the cached plan of an entry for this geometry, view angles and mask
// (made anew when any of them differs from the entry's last call)
// (source_zs: a helical plan with na1_half and pitch, which join the key)
int cached_plan(CachedPlan &cp, const fdkhip_geom &g, int32_t na, const double *beta,
	const uint8_t *mask, fdkhip_plan *&out, const double *source_zs = nullptr,
	int32_t na1_half = 0, float pitch = 0.f)
{
	std::vector<unsigned char> key, extra;
	if (source_zs) {
		extra.resize(sizeof(double) * (size_t) na + sizeof na1_half + sizeof pitch);
		if (na > 0)
			memcpy(extra.data(), source_zs, sizeof(double) * (size_t) na);
		memcpy(extra.data() + sizeof(double) * (size_t) na, &na1_half, sizeof na1_half);
		memcpy(extra.data() + sizeof(double) * (size_t) na + sizeof na1_half, &pitch,
			sizeof pitch);
	}
	cache_key(key, g, na, beta, mask, extra);
	if (!cp.plan || key != cp.key) {
		if (cp.plan)
			plan_free(cp.plan);
		cp.plan = nullptr;
		cp.key.clear();
		const int rc = source_zs
			? fdkhip_plan_create_helix(&cp.plan, &g, na, beta, source_zs, na1_half, pitch,
				mask, FDKHIP_F32, FDKHIP_MATH_EXACT)
			: fdkhip_plan_create(&cp.plan, &g, na, beta, mask, FDKHIP_F32, FDKHIP_MATH_EXACT);
		if (rc) {
			cp.plan = nullptr;
			return rc;
		}
		cp.key.swap(key);
	}
	out = cp.plan;
	return FDKHIP_OK;
}

// grow a cached device buffer to at least `bytes`
int cache_grow(void *&buf, size_t &cap, size_t bytes)
{
	if (bytes <= cap)
		return FDKHIP_OK;
	if (buf)
		(void) hipFree(buf);
	buf = nullptr;
	cap = 0;
	if (hipMalloc(&buf, bytes) != hipSuccess)
		return fail(FDKHIP_ENOMEM, "device allocation of %zu bytes failed", bytes);
	cap = bytes;
	return FDKHIP_OK;
}

void cache_clear(HostCache &c)
{
	for (auto &cp : c.plans) {
		if (cp.plan)
			plan_free(cp.plan);   // waits for the device
		cp.plan = nullptr;
		cp.key.clear();
	}
	if (c.stream)
		(void) hipStreamSynchronize(c.stream);
	for (int k = 0; k < kCacheBufs; ++k) {
		if (c.buf[k]) (void) hipFree(c.buf[k]);
		c.buf[k] = nullptr;
		c.cap[k] = 0;
	}
	if (c.pinned) (void) hipHostFree(c.pinned);
	c.pinned = nullptr;
	for (auto &e : c.ev)
		if (e) (void) hipEventDestroy(e);
	c.ev[0] = c.ev[1] = nullptr;
	if (c.stream) (void) hipStreamDestroy(c.stream);
	c.stream = nullptr;
}

// the pinned chunks, the copy stream and its events, made once per device
int cache_staging(HostCache &c)
{
	if (!c.pinned && hipHostMalloc((void **) &c.pinned, 2 * kStageChunk) != hipSuccess) {
		c.pinned = nullptr;
		return fail(FDKHIP_ENOMEM, "pinned host allocation of %zu bytes failed",
			2 * kStageChunk);
	}
	if (!c.stream)
		HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
	for (auto &e : c.ev)
		if (!e)
			HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
	return FDKHIP_OK;
}

// A host byte stream: logical byte `off` is at src + (off / seg) * stride +
// off % seg (seg == stride: contiguous).
struct HostStream {
	const char *src;
	size_t seg, stride;
	const char *at(size_t off, size_t &avail) const
	{
		const size_t k = off / seg, r = off - k * seg;
		avail = seg - r;
		return src + k * stride + r;
	}
};

// bytes [off, off + n) of `in` into dst, on up to 8 host threads (pieces of
// at least 1 MiB)
void par_copy_in(char *dst, const HostStream &in, size_t off, size_t n)
{
	auto run = [&in, dst, off](size_t a, size_t b) {
		while (a < b) {
			size_t avail = 0;
			const char *p = in.at(off + a, avail);
			const size_t k = std::min(avail, b - a);
			memcpy(dst + a, p, k);
			a += k;
		}
	};
	const size_t piece = (size_t) 1 << 20;
	const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
	const size_t nt = std::min<size_t>({(size_t) 8, (size_t) hw, (n + piece - 1) / piece});
	if (nt <= 1) {
		run(0, n);
		return;
	}
	const size_t part = (n + nt - 1) / nt;
	std::vector<std::thread> th;
	th.reserve(nt);
	for (size_t t = 1; t < nt; ++t) {
		const size_t a = t * part, b = std::min(n, a + part);
		if (a >= b)
			continue;
		try {
			th.emplace_back(run, a, b);
		} catch (...) {
			run(a, b);   // no thread to be had: this piece on the calling thread
		}
	}
	run(0, std::min(n, part));
	for (auto &x : th)
		x.join();
}

void par_memcpy(void *dst, const void *src, size_t n)
{
	par_copy_in((char *) dst, HostStream{(const char *) src, n ? n : 1, n ? n : 1}, 0, n);
}

// host -> device through the pinned chunks, on c.stream: chunk i is copied
// into pinned half i % 2 once the DMA of chunk i - 2 has read it
int staged_h2d(HostCache &c, void *dst, const HostStream &in, size_t n)
{
	size_t i = 0;
	for (size_t off = 0; off < n; off += kStageChunk, ++i) {
		const size_t m = std::min(kStageChunk, n - off);
		char *buf = c.pinned + (i & 1) * kStageChunk;
		if (i >= 2)
			HIP_TRY(hipEventSynchronize(c.ev[i & 1]));
		par_copy_in(buf, in, off, m);
		HIP_TRY(hipMemcpyAsync((char *) dst + off, buf, m, hipMemcpyHostToDevice, c.stream));
		HIP_TRY(hipEventRecord(c.ev[i & 1], c.stream));
	}
	return FDKHIP_OK;
}

// device -> host through the pinned chunks, after the work queued on
// c.stream: the DMA of chunk i + 1 runs while host threads copy chunk i out
int staged_d2h(HostCache &c, void *dst, const void *src, size_t n)
{
	const size_t nchunk = (n + kStageChunk - 1) / kStageChunk;
	auto issue = [&](size_t i) -> int {
		const size_t off = i * kStageChunk, m = std::min(kStageChunk, n - off);
		HIP_TRY(hipMemcpyAsync(c.pinned + (i & 1) * kStageChunk, (const char *) src + off, m,
			hipMemcpyDeviceToHost, c.stream));
		HIP_TRY(hipEventRecord(c.ev[i & 1], c.stream));
		return FDKHIP_OK;
	};
	for (size_t i = 0; i < std::min<size_t>(2, nchunk); ++i) {
		const int rc = issue(i);
		if (rc)
			return rc;
	}
	for (size_t i = 0; i < nchunk; ++i) {
		const size_t off = i * kStageChunk, m = std::min(kStageChunk, n - off);
		HIP_TRY(hipEventSynchronize(c.ev[i & 1]));
		par_memcpy((char *) dst + off, c.pinned + (i & 1) * kStageChunk, m);
		if (i + 2 < nchunk) {
			const int rc = issue(i + 2);
			if (rc)
				return rc;
		}
	}
	return FDKHIP_OK;
}

}  // namespace

extern "C" int fdkhip_release_host_cache(void)
{
	int cur = 0;
	if (hipGetDevice(&cur) != hipSuccess)
		return FDKHIP_OK;   // no device: nothing was cached
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess)
		ndev = 0;
	for (int d = 0; d < std::min(ndev, kCacheDevices); ++d) {
		HostCache &c = g_host_cache[d];
		std::lock_guard<std::mutex> lk(c.mu);
		bool any = c.pinned != nullptr;
		for (const auto &cp : c.plans)
			any = any || cp.plan != nullptr;
		for (const void *b : c.buf)
			any = any || b != nullptr;
		if (!any)
			continue;
		(void) hipSetDevice(d);
		cache_clear(c);
	}
	(void) hipSetDevice(cur);
	return FDKHIP_OK;
}

extern "C" int fdkhip_fdk_ts_back(
	const int32_t nxyz[3], const double dxyz[3], const double cxyz[3],
	const uint8_t *mask, double dso, double dsd, double dfs,
	const double dst[2], const double ost[2],
	const float *proj, int32_t ns, int32_t nt, int32_t na,
	const double *beta, int32_t nthread, float *image)
{
	int rc = check_mex_args(nxyz, dxyz, cxyz, dst, ost, ns, nt, na, beta, nthread, proj, image);
	if (rc)
		return rc;
	const fdkhip_geom g = mex_geom(nxyz, dxyz, cxyz, dso, dsd, dfs, dst, ost, ns, nt);
	rc = check_geom(&g);
	if (rc)
		return rc;
	const size_t nvox = (size_t) g.nx * g.ny * g.nz;
	if (nvox == 0)
		return FDKHIP_OK;
	int dev = 0;
	HIP_TRY(hipGetDevice(&dev));
	if (dev < 0 || dev >= kCacheDevices)
		return fail(FDKHIP_EUNSUPPORTED, "device %d: at most %d devices", dev, kCacheDevices);
	HostCache &c = g_host_cache[dev];
	std::lock_guard<std::mutex> lk(c.mu);
	fdkhip_plan *p = nullptr;
	rc = cached_plan(c.plans[kEntryTsBack], g, na, beta, mask, p);
	if (rc)
		return rc;
	const size_t nproj = (size_t) ns * nt * na;
	rc = cache_grow(c.buf[0], c.cap[0], std::max(nproj, (size_t) 1) * sizeof(float));
	if (!rc)
		rc = cache_grow(c.buf[1], c.cap[1], nvox * sizeof(float));
	if (!rc)
		rc = cache_staging(c);
	const size_t pb = nproj * sizeof(float);
	if (!rc)
		rc = staged_h2d(c, c.buf[0], HostStream{(const char *) proj, pb ? pb : 1, pb ? pb : 1},
			pb);
	if (!rc)
		rc = fdkhip_plan_execute_direct(p, c.buf[0], nullptr, (float *) c.buf[1], c.stream);
	if (!rc)
		rc = staged_d2h(c, image, c.buf[1], nvox * sizeof(float));
	if (rc) {
		(void) hipStreamSynchronize(c.stream);   // no copy may still read a pinned chunk
		return rc;
	}
	HIP_TRY(hipStreamSynchronize(c.stream));
	return FDKHIP_OK;
}

// cbct_back_mex's whole use_mex = 1 path (fbp/cbct_back.m:74-137) on host
// buffers in MATLAB order, with the permutes, the edge-row replication and
// the angle scale done on the device.
extern "C" int fdkhip_cbct_back(const fdkhip_geom *g, const float *proj_sta, int32_t na,
	const double *beta, int32_t ia_skip, const uint8_t *mask, int32_t extrapolate_t,
	double scale, float *image_xyz)
{
	if (!g)
		return fail(FDKHIP_EARG, "geometry is NULL");
	int rc = check_geom(g);
	if (rc)
		return rc;
	if (na < 0)
		return fail(FDKHIP_EARG, "na must be >= 0 (na=%d)", na);
	if (ia_skip < 1)
		return fail(FDKHIP_EARG, "ia_skip must be >= 1 (ia_skip=%d)", ia_skip);
	if (extrapolate_t < 0)
		return fail(FDKHIP_EARG, "extrapolate_t must be >= 0");
	if (extrapolate_t > 0 && g->offset_t != 0)   // cbct_back.m:91-93
		return fail(FDKHIP_EARG, "offset_t not yet done for extrapolation");
	const size_t nvox = (size_t) g->nx * g->ny * g->nz;
	if (nvox > 0 && !image_xyz)
		return fail(FDKHIP_EARG, "image is NULL");
	if (na > 0 && (!beta || (!proj_sta && (size_t) g->ns * g->nt > 0)))
		return fail(FDKHIP_EARG, "proj or beta is NULL");
	if (g->nz > 65535)
		return fail(FDKHIP_EUNSUPPORTED, "nz=%d: at most 65535 slices per volume", g->nz);
	if (nvox == 0)
		return FDKHIP_OK;
	const int e = extrapolate_t;
	const int ns = g->ns, nt = g->nt, nte = nt + 2 * e;
	const int nsub = (na + ia_skip - 1) / ia_skip;   // views 1:ia_skip:na
	fdkhip_geom gk = *g;
	gk.nt = nte;
	std::vector<double> bsub((size_t) nsub);
	for (int i = 0; i < nsub; ++i)
		bsub[i] = beta[(size_t) i * ia_skip];
	int dev = 0;
	HIP_TRY(hipGetDevice(&dev));
	if (dev < 0 || dev >= kCacheDevices)
		return fail(FDKHIP_EUNSUPPORTED, "device %d: at most %d devices", dev, kCacheDevices);
	HostCache &c = g_host_cache[dev];
	std::lock_guard<std::mutex> lk(c.mu);
	fdkhip_plan *p = nullptr;
	rc = cached_plan(c.plans[kEntryCbctBack], gk, nsub, bsub.data(), mask, p);
	if (rc)
		return rc;
	const size_t view = (size_t) ns * nt, viewe = (size_t) ns * nte;
	const size_t nsta = view * nsub, ntsa = viewe * nsub;
	// buffers: [ns nt nsub] views as given, [nt+2e ns nsub] views, the
	// [nz nx ny] image, the [nx ny nz] image
	rc = cache_grow(c.buf[0], c.cap[0], std::max(nsta, (size_t) 1) * sizeof(float));
	if (!rc)
		rc = cache_grow(c.buf[2], c.cap[2], std::max(ntsa, (size_t) 1) * sizeof(float));
	if (!rc)
		rc = cache_grow(c.buf[1], c.cap[1], nvox * sizeof(float));
	if (!rc)
		rc = cache_grow(c.buf[3], c.cap[3], nvox * sizeof(float));
	if (!rc)
		rc = cache_staging(c);
	if (!rc && nsta > 0) {
		// views 0, ia_skip, 2 ia_skip, ... of [ns nt na]
		rc = staged_h2d(c, c.buf[0], HostStream{(const char *) proj_sta, view * sizeof(float),
			view * ia_skip * sizeof(float)}, nsta * sizeof(float));
		// [ns nt] -> [nt+2e ns] per view, edge rows replicated (cbct_back.m:86-98)
		if (!rc && launch_transpose((const float *) c.buf[0], (float *) c.buf[2], nsub, nt, ns,
				(long) view, ns, (long) viewe, nte, nte, e, 1.f, c.stream) != hipSuccess)
			rc = fail(FDKHIP_EHIP, "launch_transpose (views)");
	}
	if (!rc)
		rc = plan_launch(p, c.buf[2], nullptr, (float *) c.buf[1], c.stream);
	// [nz nx ny] -> [nx ny nz] with the angle scale (cbct_back.m:104,133-135);
	// the double scale multiplies a single image in single, as MATLAB does
	if (!rc && launch_transpose((const float *) c.buf[1], (float *) c.buf[3], g->ny, g->nx,
			g->nz, (long) g->nx * g->nz, g->nz, g->nx, (long) g->nx * g->ny, g->nx, 0,
			(float) scale, c.stream) != hipSuccess)
		rc = fail(FDKHIP_EHIP, "launch_transpose (image)");
	if (!rc)
		rc = staged_d2h(c, image_xyz, c.buf[3], nvox * sizeof(float));
	if (rc) {
		(void) hipStreamSynchronize(c.stream);   // no copy may still read a pinned chunk
		return rc;
	}
	HIP_TRY(hipStreamSynchronize(c.stream));
	return FDKHIP_OK;
}

extern "C" int fdkhip_fdk_st_back(
	const int32_t nxyz[3], const double dxyz[3], const double cxyz[3],
	const uint8_t *mask, double dso, double dsd, double dfs,
	const double dst[2], const double ost[2],
	const float *proj, int32_t ns, int32_t nt, int32_t na,
	const double *beta, int32_t nthread, float *image)
{
	int rc = check_mex_args(nxyz, dxyz, cxyz, dst, ost, ns, nt, na, beta, nthread, proj, image);
	if (rc)
		return rc;
	const fdkhip_geom g = mex_geom(nxyz, dxyz, cxyz, dso, dsd, dfs, dst, ost, ns, nt);
	rc = check_geom(&g);
	if (rc)
		return rc;
	if (nthread != 1)
		fprintf(stderr, "WARN fdkhip: only nthread=1 supported for fdk_st_, ignoring\n");
	const size_t nvox = (size_t) g.nx * g.ny * g.nz;
	if (nvox == 0)
		return FDKHIP_OK;
	const size_t nxy = (size_t) g.nx * g.ny;
	const size_t nproj = (size_t) ns * nt * na;
	std::vector<float2> trig = trig_table(beta, na);
	DevBuf dproj, dimg, dtrig, dmask;
	hipError_t e = hipMalloc(&dimg.p, nvox * sizeof(float));
	if (e == hipSuccess && nproj > 0)
		e = hipMalloc(&dproj.p, nproj * sizeof(float));
	if (e == hipSuccess && na > 0)
		e = hipMalloc(&dtrig.p, na * sizeof(float2));
	if (e == hipSuccess && mask)
		e = hipMalloc(&dmask.p, nxy);
	if (e != hipSuccess)
		return fail(FDKHIP_ENOMEM, "device allocation failed: %s", hipGetErrorString(e));
	if (nproj > 0)
		HIP_TRY(hipMemcpy(dproj.p, proj, nproj * sizeof(float), hipMemcpyHostToDevice));
	if (na > 0)
		HIP_TRY(hipMemcpy(dtrig.p, trig.data(), na * sizeof(float2), hipMemcpyHostToDevice));
	if (mask)
		HIP_TRY(hipMemcpy(dmask.p, mask, nxy, hipMemcpyHostToDevice));
	StArgs a;
	memset(&a, 0, sizeof a);
	a.g = narrow_geom(&g, na);
	a.proj = (const float *) dproj.p;
	a.trig = (const float2 *) dtrig.p;
	a.mask = (const uint8_t *) dmask.p;
	a.image = (float *) dimg.p;
	if (na == 0 || (long) ns * nt == 0) {
		HIP_TRY(launch_fill(a.image, (long) nxy, g.nz, 0, 0, nullptr));
	} else {
		HIP_TRY(launch_st(a, nullptr));
	}
	HIP_TRY(hipMemcpy(image, dimg.p, nvox * sizeof(float), hipMemcpyDeviceToHost));
	return FDKHIP_OK;
}

// ---------------------------------------------------------------------------
// helical / cone-parallel 'fdk,ts,back' (mex/src/fdk-helix/mex)

// Host tables of fdk_ts_back_t / fdk_ts_back_init (fdk,ts,t.c:39-141): per
// view trig, height and window; the w3d table with a zero tail; per slice
// chunk the range of views whose window reaches it.
struct HelixHost {
	std::vector<HelixView> views;
	std::vector<int2> chunks;
	std::vector<float> tab;
};

static int helix_setup(const fdkhip_geom *g, int na, const double *beta,
	const double *source_zs, int na1_half, int cone_par, int w3d, double dz_per_rad,
	float pitch, HelixArgs &a, HelixHost &h)
{
	int rc = check_geom(g);
	if (rc)
		return rc;
	if (na < 0)
		return fail(FDKHIP_EARG, "na must be >= 0 (na=%d)", na);
	if (na > 0 && (!beta || !source_zs))
		return fail(FDKHIP_EARG, "beta or source_zs is NULL");
	if (na1_half < 0 || w3d < 0)
		return fail(FDKHIP_EARG, "na1_half and w3d must be >= 0");
	memset(&a, 0, sizeof a);
	a.g = narrow_geom(g, na);
	const GeomF &f = a.g;
	if (na > 0 && f.dz / f.dt < 0)   // fdk,ts.c:153-154
		return fail(FDKHIP_EGEOM, "todo: dz/dt < 0 not done");
	a.cone_par = cone_par != 0;
	a.w3d = w3d;
	a.pitch = pitch;
	a.dz_per_rad = (double) (float) dz_per_rad;   // fdk,mex.c:128: a float
	a.nt2 = g->nt / 2;
	a.nzc = (g->nz + HX_ZC - 1) / HX_ZC;
	h.views.resize((size_t) std::max(na, 1));
	const float wz = f.wz, dz = f.dz;
	for (int ia = 0; ia < na; ++ia) {
		HelixView &v = h.views[ia];
		memset(&v, 0, sizeof v);
		const float b = (float) beta[ia];
		v.cosb = cosf(b);
		v.sinb = sinf(b);
		v.zs = (float) source_zs[ia];
		v.izb0 = 0;
		v.izb1 = g->nz;
		if (pitch != 0.f) {   // fdk,ts,t.c:80-85
			const int a0 = std::max(ia - na1_half, 0);
			const int a1 = std::min(ia + na1_half, na - 1);
			const float c0 = ceilf((float) (wz + source_zs[a0] / dz));
			const float c1 = ceilf((float) (wz + source_zs[a1] / dz));
			v.izb0 = (int) (c0 > 0 ? c0 : 0);
			v.izb1 = (int) (c1 < g->nz ? c1 : g->nz);
		}
	}
	if (na > 0) {
		a.zs_min = (float) source_zs[0];       // fdk,ts,t.c:70-71
		a.zs_max = (float) source_zs[na - 1];
	}
	// w3d table (fdk,ts,t.c:134-141 -> fdk,w3d,pow,tab.c:9-31, whose `dt`
	// argument receives wz) and a zero tail up to the largest index the
	// lookup can form: nt/2 + (nt/2) * trunc(max |iz - wz|)
	const int nt2 = a.nt2;
	const int rows = (int) ceilf((float) (g->nz / 2));
	const int ntab = pitch != 0.f ? nt2 : nt2 * rows;
	const float span = std::max(fabsf(wz), fabsf((float) (g->nz - 1) - wz));
	const long tail = (long) nt2 * ((long) span + 2) + 2;
	if (a.cone_par && w3d && ntab + tail > (1l << 28))
		return fail(FDKHIP_EUNSUPPORTED, "w3d table too large");
	h.tab.assign(a.cone_par && w3d ? (size_t) (ntab + tail) : 1, 0.f);
	if (a.cone_par && w3d) {
		const float dsd = f.dsd;
		if (pitch != 0.f) {
			for (int it = 0; it < nt2; ++it)
				h.tab[it] = (float) pow(it * wz / dsd, w3d * pitch);
		} else {
			size_t k = 0;
			for (int iz = 0; iz < rows; ++iz)
				for (int it = 0; it < nt2; ++it)
					h.tab[k++] = (float) pow(it * wz / dsd, w3d * iz * dz);
		}
	}
	a.tab_len = (int) h.tab.size();
	// views whose window [izb0, izb1) meets each chunk of HX_ZC slices
	h.chunks.assign((size_t) std::max(a.nzc, 1), make_int2(0, 0));
	for (int c = 0; c < a.nzc; ++c) {
		const int z0 = c * HX_ZC, z1 = std::min(z0 + HX_ZC, g->nz);
		int va = -1, vb = -1;
		for (int ia = 0; ia < na; ++ia)
			if (h.views[ia].izb0 < z1 && h.views[ia].izb1 > z0) {
				if (va < 0)
					va = ia;
				vb = ia + 1;
			}
		if (va >= 0)
			h.chunks[c] = make_int2(va, vb);
	}
	return FDKHIP_OK;
}

// Upload the tables, launch on `s`, wait; kernel_ms (optional) from events.
static int helix_run(HelixArgs &a, const HelixHost &h, const uint8_t *mask_host,
	hipStream_t s, float *kernel_ms)
{
	const size_t nxy = (size_t) a.g.nx * a.g.ny;
	DevBuf dviews, dchunks, dtab, dmask;
	hipError_t e = hipMalloc(&dviews.p, h.views.size() * sizeof(HelixView));
	if (e == hipSuccess)
		e = hipMalloc(&dchunks.p, h.chunks.size() * sizeof(int2));
	if (e == hipSuccess)
		e = hipMalloc(&dtab.p, h.tab.size() * sizeof(float));
	if (e == hipSuccess && mask_host && nxy > 0)
		e = hipMalloc(&dmask.p, nxy);
	if (e != hipSuccess)
		return fail(FDKHIP_ENOMEM, "device allocation failed: %s", hipGetErrorString(e));
	HIP_TRY(hipMemcpyAsync(dviews.p, h.views.data(), h.views.size() * sizeof(HelixView),
		hipMemcpyHostToDevice, s));
	HIP_TRY(hipMemcpyAsync(dchunks.p, h.chunks.data(), h.chunks.size() * sizeof(int2),
		hipMemcpyHostToDevice, s));
	HIP_TRY(hipMemcpyAsync(dtab.p, h.tab.data(), h.tab.size() * sizeof(float),
		hipMemcpyHostToDevice, s));
	if (dmask.p)
		HIP_TRY(hipMemcpyAsync(dmask.p, mask_host, nxy, hipMemcpyHostToDevice, s));
	a.views = (const HelixView *) dviews.p;
	a.chunk_views = (const int2 *) dchunks.p;
	a.tab = (const float *) dtab.p;
	a.mask = (const uint8_t *) dmask.p;
	hipEvent_t ev[2] = {nullptr, nullptr};
	if (kernel_ms) {
		HIP_TRY(hipEventCreate(&ev[0]));
		e = hipEventCreate(&ev[1]);
		if (e != hipSuccess) {
			(void) hipEventDestroy(ev[0]);
			return fail(FDKHIP_EHIP, "hipEventCreate: %s", hipGetErrorString(e));
		}
		e = hipEventRecord(ev[0], s);
	}
	if (e == hipSuccess)
		e = launch_helix(a, s);
	if (e == hipSuccess && kernel_ms)
		e = hipEventRecord(ev[1], s);
	if (e == hipSuccess)
		e = hipStreamSynchronize(s);
	if (e == hipSuccess && kernel_ms)
		e = hipEventElapsedTime(kernel_ms, ev[0], ev[1]);
	if (kernel_ms) {
		(void) hipEventDestroy(ev[0]);
		(void) hipEventDestroy(ev[1]);
	}
	if (e != hipSuccess)
		return fail(FDKHIP_EHIP, "helical backprojection launch: %s", hipGetErrorString(e));
	return FDKHIP_OK;
}

// Native helical geometry (cone_par = 0) runs on the axial pipeline: a
// helical plan (per-view heights and windows) for this one call.  kernel_ms:
// the plan's device time (fill, padded copy, backprojection).
static int helix_plan_run(const fdkhip_geom *g, const float *proj_dev, int na,
	const double *beta, const double *source_zs, int na1_half, float pitch,
	const uint8_t *mask, float *image_dev, hipStream_t s, float *kernel_ms)
{
	fdkhip_plan *p = nullptr;
	int rc = fdkhip_plan_create_helix(&p, g, na, beta, source_zs, na1_half, pitch, mask,
		FDKHIP_F32, FDKHIP_MATH_EXACT);
	if (rc)
		return rc;
	hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
	if (kernel_ms)
		for (auto &e : ev)
			if (!rc && hipEventCreate(&e) != hipSuccess)
				rc = fail(FDKHIP_EHIP, "hipEventCreate");
	if (!rc)
		rc = plan_launch(p, proj_dev, nullptr, image_dev, s, kernel_ms ? ev : nullptr);
	if (!rc && hipStreamSynchronize(s) != hipSuccess)
		rc = fail(FDKHIP_EHIP, "helical backprojection: %s", hipGetErrorString(hipGetLastError()));
	if (!rc && kernel_ms && hipEventElapsedTime(kernel_ms, ev[0], ev[3]) != hipSuccess)
		rc = fail(FDKHIP_EHIP, "hipEventElapsedTime");
	for (auto &e : ev)
		if (e)
			(void) hipEventDestroy(e);
	plan_free(p);
	return rc;
}

extern "C" int fdkhip_fdk_ts_back_helix(
	const int32_t nxyz[3], const double dxyz[3], const double cxyz[3],
	const uint8_t *mask, double dso, double dsd, double dfs,
	const double dst[2], const double ost[2],
	const float *proj, int32_t ns, int32_t nt, int32_t na,
	const double *beta, const double *source_zs, int32_t na1_half,
	int32_t cone_par, int32_t w3d, double source_dz_per_rad, float pitch,
	int32_t nthread, float *image)
{
	int rc = check_mex_args(nxyz, dxyz, cxyz, dst, ost, ns, nt, na, beta, nthread, proj, image);
	if (rc)
		return rc;
	if (na > 0 && !source_zs)
		return fail(FDKHIP_EARG, "source_zs length=0 vs na=%d", na);
	const fdkhip_geom g = mex_geom(nxyz, dxyz, cxyz, dso, dsd, dfs, dst, ost, ns, nt);
	HelixArgs a;
	HelixHost h;
	if ((rc = helix_setup(&g, na, beta, source_zs, na1_half, cone_par, w3d,
			source_dz_per_rad, pitch, a, h)))
		return rc;
	const size_t nvox = (size_t) g.nx * g.ny * g.nz;
	if (nvox == 0)
		return FDKHIP_OK;
	const size_t nproj = (size_t) ns * nt * na;
	// the cached buffers, pinned transfers and (native helical geometry) plan
	// of the host-array entries
	int dev = 0;
	HIP_TRY(hipGetDevice(&dev));
	if (dev < 0 || dev >= kCacheDevices)
		return fail(FDKHIP_EUNSUPPORTED, "device %d: at most %d devices", dev, kCacheDevices);
	HostCache &c = g_host_cache[dev];
	std::lock_guard<std::mutex> lk(c.mu);
	rc = cache_grow(c.buf[0], c.cap[0], std::max(nproj, (size_t) 1) * sizeof(float));
	if (!rc)
		rc = cache_grow(c.buf[1], c.cap[1], nvox * sizeof(float));
	if (!rc)
		rc = cache_staging(c);
	if (rc)
		return rc;
	a.proj = (const float *) c.buf[0];
	a.image = (float *) c.buf[1];
	const size_t pb = nproj * sizeof(float);
	if (nproj == 0) {
		if (launch_fill(a.image, (long) g.nx * g.ny, g.nz, 0, 0, c.stream) != hipSuccess)
			rc = fail(FDKHIP_EHIP, "launch_fill");
	} else {
		rc = staged_h2d(c, c.buf[0], HostStream{(const char *) proj, pb, pb}, pb);
		if (!rc && !cone_par) {
			fdkhip_plan *p = nullptr;
			rc = cached_plan(c.plans[kEntryHelix], g, na, beta, mask, p, source_zs, na1_half,
				pitch);
			if (!rc)
				rc = plan_launch(p, a.proj, nullptr, a.image, c.stream);
		} else if (!rc) {
			rc = helix_run(a, h, mask, c.stream, nullptr);
		}
	}
	if (!rc)
		rc = staged_d2h(c, image, a.image, nvox * sizeof(float));
	if (rc) {
		(void) hipStreamSynchronize(c.stream);   // no copy may still read a pinned chunk
		return rc;
	}
	HIP_TRY(hipStreamSynchronize(c.stream));
	return FDKHIP_OK;
}

extern "C" int fdkhip_fdk_ts_back_helix_direct(const fdkhip_geom *g, const float *proj_dev,
	int32_t na, const double *beta, const double *source_zs, int32_t na1_half,
	int32_t cone_par, int32_t w3d, double source_dz_per_rad, float pitch,
	const uint8_t *mask, float *image_dev, void *stream, float *kernel_ms)
{
	if (!g)
		return fail(FDKHIP_EARG, "geometry is NULL");
	HelixArgs a;
	HelixHost h;
	int rc = helix_setup(g, na, beta, source_zs, na1_half, cone_par, w3d, source_dz_per_rad,
		pitch, a, h);
	if (rc)
		return rc;
	if (kernel_ms)
		*kernel_ms = 0.f;
	const size_t nvox = (size_t) g->nx * g->ny * g->nz;
	if (nvox == 0)
		return FDKHIP_OK;
	if (!image_dev)
		return fail(FDKHIP_EARG, "image is NULL");
	if (na > 0 && (size_t) g->ns * g->nt > 0 && !proj_dev)
		return fail(FDKHIP_EARG, "proj is NULL");
	hipStream_t s = (hipStream_t) stream;
	a.proj = proj_dev;
	a.image = image_dev;
	if (na == 0 || (size_t) g->ns * g->nt == 0) {
		HIP_TRY(launch_fill(image_dev, (long) g->nx * g->ny, g->nz, 0, 0, s));
		HIP_TRY(hipStreamSynchronize(s));
		return FDKHIP_OK;
	}
	if (!cone_par)
		return helix_plan_run(g, proj_dev, na, beta, source_zs, na1_half, pitch, mask,
			image_dev, s, kernel_ms);
	return helix_run(a, h, mask, s, kernel_ms);
}

// ---------------------------------------------------------------------------
// cbct_back_mat (fbp/cbct_back.m:142-270): the MATLAB backprojector

// Argument checks of both mat entry points; fills the kernel arguments other
// than the device pointers.  MATLAB itself fails (an undefined sprime, or a
// row index 0) where this returns FDKHIP_EARG for dfs and nt.
static int mat_args(const fdkhip_geom *g, int na, const double *beta, int ia_skip,
	double offset_source, double scale, MatArgs &a)
{
	int rc = check_geom(g);
	if (rc)
		return rc;
	if (na < 0)
		return fail(FDKHIP_EARG, "na must be >= 0 (na=%d)", na);
	if (ia_skip < 1)
		return fail(FDKHIP_EARG, "ia_skip must be >= 1 (ia_skip=%d)", ia_skip);
	if (na > 0 && !beta)
		return fail(FDKHIP_EARG, "beta is NULL");
	const bool par = isinf(g->dsd) || isinf(g->dso);
	if (!par && !isinf(g->dfs) && g->dfs != 0)
		return fail(FDKHIP_EARG, "cbct_back_mat: dfs must be 0 (arc) or inf (flat), got %g",
			g->dfs);
	const long long nvox = (long long) g->nx * g->ny * g->nz;
	if (nvox > 0 && na > 0 && (g->ns < 1 || g->nt < 2))
		return fail(FDKHIP_EARG, "cbct_back_mat needs ns >= 1 and nt >= 2 (ns=%d nt=%d)",
			g->ns, g->nt);
	memset(&a, 0, sizeof a);
	a.nx = g->nx; a.ny = g->ny; a.nz = g->nz;
	a.ns = g->ns; a.nt = g->nt;
	a.nv = (na + ia_skip - 1) / ia_skip;           // views 1:ia_skip:na
	a.nzc = (g->nz + MAT_ZC - 1) / MAT_ZC;
	a.kind = par ? KIND_PAR : (g->dfs == 0 ? KIND_ARC : KIND_FLAT);
	a.dx = g->dx; a.dy = g->dy; a.dz = g->dz;
	// cbct_back.m:160-162, 185-186
	a.wx = (g->nx - 1) / 2. + g->offset_x;
	a.wy = (g->ny - 1) / 2. + g->offset_y;
	a.wz = (g->nz - 1) / 2. + g->offset_z;
	a.ws = (g->ns + 1) / 2. + g->offset_s;
	a.wt = (g->nt + 1) / 2. + g->offset_t;
	a.dso = g->dso; a.dsd = g->dsd; a.ds = g->ds; a.dt = g->dt;
	a.offset_source = offset_source;
	a.scale = scale;
	return FDKHIP_OK;
}

// [nv] {cos, sin, source_zs, 0} of views 0, ia_skip, 2 ia_skip, ...
static std::vector<double4> mat_views(const double *beta, const double *source_zs, int nv,
	int ia_skip)
{
	std::vector<double4> t((size_t) nv);
	for (int i = 0; i < nv; ++i) {
		const size_t ia = (size_t) i * ia_skip;
		t[i].x = cos(beta[ia]);
		t[i].y = sin(beta[ia]);
		t[i].z = source_zs ? source_zs[ia] : 0.0;
		t[i].w = 0.0;
	}
	return t;
}

// Upload the view table and the mask, launch, wait.  proj and image are
// device pointers; kernel_ms (optional) receives the kernel's time from HIP
// events recorded on `s` around the launch.
static int mat_run(MatArgs &a, const double *beta, const double *source_zs, int ia_skip,
	const uint8_t *mask_host, bool fp64, hipStream_t s, float *kernel_ms)
{
	const size_t nxy = (size_t) a.nx * a.ny;
	std::vector<double4> views = mat_views(beta, source_zs, a.nv, ia_skip);
	DevBuf dviews, dmask;
	hipError_t e = hipSuccess;
	if (a.nv > 0)
		e = hipMalloc(&dviews.p, views.size() * sizeof(double4));
	if (e == hipSuccess && mask_host && nxy > 0)
		e = hipMalloc(&dmask.p, nxy);
	if (e != hipSuccess)
		return fail(FDKHIP_ENOMEM, "device allocation failed: %s", hipGetErrorString(e));
	if (a.nv > 0)
		HIP_TRY(hipMemcpyAsync(dviews.p, views.data(), views.size() * sizeof(double4),
			hipMemcpyHostToDevice, s));
	if (dmask.p)
		HIP_TRY(hipMemcpyAsync(dmask.p, mask_host, nxy, hipMemcpyHostToDevice, s));
	a.views = (const double4 *) dviews.p;
	a.mask = (const uint8_t *) dmask.p;
	hipEvent_t ev[2] = {nullptr, nullptr};
	if (kernel_ms) {
		HIP_TRY(hipEventCreate(&ev[0]));
		e = hipEventCreate(&ev[1]);
		if (e != hipSuccess) {
			(void) hipEventDestroy(ev[0]);
			return fail(FDKHIP_EHIP, "hipEventCreate: %s", hipGetErrorString(e));
		}
		e = hipEventRecord(ev[0], s);
	}
	if (e == hipSuccess)
		e = launch_mat(a, fp64, s);
	if (e == hipSuccess && kernel_ms)
		e = hipEventRecord(ev[1], s);
	if (e == hipSuccess)
		e = hipStreamSynchronize(s);
	if (e == hipSuccess && kernel_ms)
		e = hipEventElapsedTime(kernel_ms, ev[0], ev[1]);
	if (kernel_ms) {
		(void) hipEventDestroy(ev[0]);
		(void) hipEventDestroy(ev[1]);
	}
	if (e != hipSuccess)
		return fail(FDKHIP_EHIP, "cbct_back_mat launch: %s", hipGetErrorString(e));
	return FDKHIP_OK;
}

extern "C" int fdkhip_cbct_back_mat(const fdkhip_geom *g, const void *proj_sta,
	int32_t proj_fp64, int32_t na, const double *beta, const double *source_zs,
	int32_t ia_skip, double offset_source, const uint8_t *mask, double scale,
	double *image_xyz)
{
	MatArgs a;
	int rc = mat_args(g, na, beta, ia_skip, offset_source, scale, a);
	if (rc)
		return rc;
	const size_t nvox = (size_t) g->nx * g->ny * g->nz;
	if (nvox == 0)
		return FDKHIP_OK;
	if (!image_xyz)
		return fail(FDKHIP_EARG, "image is NULL");
	const size_t view = (size_t) g->ns * g->nt;
	const size_t esz = proj_fp64 ? sizeof(double) : sizeof(float);
	if (a.nv > 0 && view > 0 && !proj_sta)
		return fail(FDKHIP_EARG, "proj is NULL");
	DevBuf dproj, dimg;
	hipError_t e = hipMalloc(&dimg.p, nvox * sizeof(double));
	if (e == hipSuccess && a.nv > 0 && view > 0)
		e = hipMalloc(&dproj.p, view * a.nv * esz);
	if (e != hipSuccess)
		return fail(FDKHIP_ENOMEM, "device allocation failed: %s", hipGetErrorString(e));
	if (dproj.p)   // views 0, ia_skip, 2 ia_skip, ... in one strided copy
		HIP_TRY(hipMemcpy2D(dproj.p, view * esz, proj_sta, view * ia_skip * esz, view * esz,
			a.nv, hipMemcpyHostToDevice));
	a.proj = dproj.p;
	a.view_stride = (long) view;
	a.image = (double *) dimg.p;
	if (a.nv == 0)   // img = zeros: no view adds anything (cbct_back.m:189)
		HIP_TRY(hipMemset(dimg.p, 0, nvox * sizeof(double)));
	else if ((rc = mat_run(a, beta, source_zs, ia_skip, mask, proj_fp64 != 0, nullptr,
			nullptr)))
		return rc;
	HIP_TRY(hipMemcpy(image_xyz, dimg.p, nvox * sizeof(double), hipMemcpyDeviceToHost));
	return FDKHIP_OK;
}

extern "C" int fdkhip_cbct_back_mat_direct(const fdkhip_geom *g, const void *proj_dev,
	int32_t proj_fp64, int32_t na, const double *beta, const double *source_zs,
	int32_t ia_skip, double offset_source, const uint8_t *mask, double scale,
	double *image_dev, void *stream, float *kernel_ms)
{
	MatArgs a;
	int rc = mat_args(g, na, beta, ia_skip, offset_source, scale, a);
	if (rc)
		return rc;
	if (kernel_ms)
		*kernel_ms = 0.f;
	const size_t nvox = (size_t) g->nx * g->ny * g->nz;
	if (nvox == 0)
		return FDKHIP_OK;
	if (!image_dev)
		return fail(FDKHIP_EARG, "image is NULL");
	const size_t view = (size_t) g->ns * g->nt;
	if (a.nv > 0 && view > 0 && !proj_dev)
		return fail(FDKHIP_EARG, "proj is NULL");
	hipStream_t s = (hipStream_t) stream;
	a.proj = proj_dev;
	a.view_stride = (long) (view * ia_skip);
	a.image = image_dev;
	if (a.nv == 0) {
		HIP_TRY(hipMemsetAsync(image_dev, 0, nvox * sizeof(double), s));
		HIP_TRY(hipStreamSynchronize(s));
		return FDKHIP_OK;
	}
	return mat_run(a, beta, source_zs, ia_skip, mask, proj_fp64 != 0, s, kernel_ms);
}

extern "C" int fdkhip_cbct_back_helix_mat(const fdkhip_geom *g, const void *proj_sta,
	int32_t proj_fp64, int32_t na, const double *beta, const double *source_zs,
	int32_t ia_skip, double offset_source, const int32_t *slice_views,
	const double *slice_scale, int32_t cone_par, double source_dz_per_rad,
	const uint8_t *mask, double *image_xyz)
{
	MatArgs a;
	int rc = mat_args(g, na, beta, 1, offset_source, 1.0, a);   // every view on the device
	if (rc)
		return rc;
	if (ia_skip < 1)
		return fail(FDKHIP_EARG, "ia_skip must be >= 1 (ia_skip=%d)", ia_skip);
	if (cone_par != 0 && cone_par != 1)
		return fail(FDKHIP_EARG, "cone_par must be 0 or 1 (cone_par=%d)", cone_par);
	const size_t nvox = (size_t) g->nx * g->ny * g->nz;
	if (nvox == 0)
		return FDKHIP_OK;
	if (!image_xyz)
		return fail(FDKHIP_EARG, "image is NULL");
	if (!slice_views)
		return fail(FDKHIP_EARG, "slice_views is NULL");
	if (na > 0 && !source_zs)
		return fail(FDKHIP_EARG, "source_zs is NULL");
	std::vector<int2> zr((size_t) g->nz);
	for (int z = 0; z < g->nz; ++z) {
		const int lo = slice_views[2 * z], hi = slice_views[2 * z + 1];
		if (lo <= hi && (lo < 0 || hi >= na))
			return fail(FDKHIP_EARG, "slice %d views [%d, %d] outside [0, %d)", z, lo, hi, na);
		zr[z] = make_int2(lo, hi);
	}
	const size_t view = (size_t) g->ns * g->nt;
	const size_t esz = proj_fp64 ? sizeof(double) : sizeof(float);
	if (na > 0 && view > 0 && !proj_sta)
		return fail(FDKHIP_EARG, "proj is NULL");
	DevBuf dproj, dimg, dzr, dzs;
	hipError_t e = hipMalloc(&dimg.p, nvox * sizeof(double));
	if (e == hipSuccess && na > 0 && view > 0)
		e = hipMalloc(&dproj.p, view * na * esz);
	if (e == hipSuccess)
		e = hipMalloc(&dzr.p, zr.size() * sizeof(int2));
	if (e == hipSuccess && slice_scale)
		e = hipMalloc(&dzs.p, (size_t) g->nz * sizeof(double));
	if (e != hipSuccess)
		return fail(FDKHIP_ENOMEM, "device allocation failed: %s", hipGetErrorString(e));
	if (dproj.p)
		HIP_TRY(hipMemcpy(dproj.p, proj_sta, view * na * esz, hipMemcpyHostToDevice));
	HIP_TRY(hipMemcpy(dzr.p, zr.data(), zr.size() * sizeof(int2), hipMemcpyHostToDevice));
	if (dzs.p)
		HIP_TRY(hipMemcpy(dzs.p, slice_scale, (size_t) g->nz * sizeof(double),
			hipMemcpyHostToDevice));
	a.proj = dproj.p;
	a.view_stride = (long) view;
	a.image = (double *) dimg.p;
	a.zrange = (const int2 *) dzr.p;
	a.vstep = ia_skip;
	a.cone_par = cone_par;
	a.dz_per_rad = source_dz_per_rad;
	a.zscale = (const double *) dzs.p;
	if (a.nv == 0 || view == 0) {   // no view adds anything: img2 = 0
		HIP_TRY(hipMemset(dimg.p, 0, nvox * sizeof(double)));
	} else if ((rc = mat_run(a, beta, source_zs, 1, mask, proj_fp64 != 0, nullptr, nullptr))) {
		return rc;
	}
	HIP_TRY(hipMemcpy(image_xyz, dimg.p, nvox * sizeof(double), hipMemcpyDeviceToHost));
	return FDKHIP_OK;
}

// ---------------------------------------------------------------------------
// autotuning: launch shape and views per LDS stage, chosen by timing

namespace {

// launch shapes (consumer waves, slices per lane) the library is built with;
// plans also have shapes of several interleaved 16-slice chunks per wave
// (4 x 48, 4 x 64, 8 x 32), the batch kernel does not
const int kShapes[][2] = {{8, 16}, {8, 8}, {4, 16}, {4, 32}, {4, 8}};
const int kPlanShapes[][2] = {{8, 16}, {8, 8}, {4, 16}, {4, 32}, {4, 8}, {4, 48}, {4, 64},
	{8, 32}};
const int kStages[] = {0, 2, 3, 4, 6, 8, 12, 16};
const int kTuneViews = 128;   // views per timed launch while tuning
const int kTuneFinalists = 4; // fastest candidates timed again over all views

struct Tuned {
	int nw, zl, vstage;   // zl: slices per consumer wave (the configuration's zper)
	int minw = 4;         // kernel build: minimum waves per SIMD
};

std::mutex g_tune_mu;
std::vector<std::pair<std::string, Tuned>> g_tuned;   // geometry key -> choice

std::string tune_key(const fdkhip_plan *p)
{
	char buf[512];
	const fdkhip_geom &g = p->geom;
	snprintf(buf, sizeof buf,
		"%d %d %d %a %a %a %a %a %a %a %a %a %d %d %a %a %a %a %d %d %d helix %d %d %a",
		g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.offset_x, g.offset_y, g.offset_z, g.dso, g.dsd,
		g.dfs, g.ns, g.nt, g.ds, g.dt, g.offset_s, g.offset_t, p->na, p->fmt, p->math,
		(int) p->helix, p->na1_half, (double) p->pitch);
	return buf;
}

bool tuned_lookup(const std::string &key, Tuned &t)
{
	std::lock_guard<std::mutex> lk(g_tune_mu);
	for (auto &kv : g_tuned)
		if (kv.first == key) {
			t = kv.second;
			return true;
		}
	return false;
}

void tuned_store(const std::string &key, const Tuned &t)
{
	std::lock_guard<std::mutex> lk(g_tune_mu);
	for (auto &kv : g_tuned)
		if (kv.first == key) {
			kv.second = t;
			return;
		}
	g_tuned.emplace_back(key, t);
}

} // namespace

// average milliseconds of `reps` calls of `run` on `s` after one warm-up call
static int time_launches(const std::function<int(hipStream_t)> &run, hipStream_t s, int reps,
	float *ms)
{
	hipEvent_t e0 = nullptr, e1 = nullptr;
	HIP_TRY(hipEventCreate(&e0));
	hipError_t e = hipEventCreate(&e1);
	if (e != hipSuccess) {
		(void) hipEventDestroy(e0);
		return fail(FDKHIP_EHIP, "hipEventCreate: %s", hipGetErrorString(e));
	}
	int rc = run(s);
	if (!rc && hipEventRecord(e0, s) != hipSuccess)
		rc = fail(FDKHIP_EHIP, "hipEventRecord");
	for (int r = 0; r < reps && !rc; ++r)
		rc = run(s);
	if (!rc && hipEventRecord(e1, s) != hipSuccess)
		rc = fail(FDKHIP_EHIP, "hipEventRecord");
	if (!rc && hipEventSynchronize(e1) != hipSuccess)
		rc = fail(FDKHIP_EHIP, "hipEventSynchronize");
	float t = 0.f;
	if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess)
		rc = fail(FDKHIP_EHIP, "hipEventElapsedTime");
	(void) hipEventDestroy(e0);
	(void) hipEventDestroy(e1);
	if (!rc)
		*ms = t / (float) std::max(reps, 1);
	return rc;
}

extern "C" int fdkhip_plan_autotune(fdkhip_plan *p, const void *proj, const float *scale,
	float *image, void *stream, int32_t reps, float *best_ms)
{
	int rc = check_exec(p, proj, scale, image);
	if (rc)
		return rc;
	if (reps < 1)
		return fail(FDKHIP_EARG, "reps must be >= 1");
	hipStream_t s = (hipStream_t) stream;
	Tuned best{p->nw, p->zl * p->zr, p->vstage, plan_minw(p)};
	float best_t = 3.4e38f;
	// the padded views once; then the backprojection kernel alone over the
	// first kTuneViews views (its time is proportional to the view count)
	rc = plan_launch(p, proj, scale, image, s);
	if (rc)
		return rc;
	std::vector<std::pair<float, Tuned>> cand;
	for (const auto &sh : kPlanShapes) {
		int last_vstage = -1;
		for (int vs : kStages) {
			if (configure(p, sh[0], sh[1], vs, 0) != FDKHIP_OK)
				continue;
			if (p->vstage == last_vstage)
				continue;   // the LDS budget clamped it to a depth already timed
			last_vstage = p->vstage;
			if (p->ngroups == 0 || p->na == 0)
				continue;
			// each kernel build this shape has (4 waves per SIMD; 8 for one-chunk
			// FP16/BF16 shapes)
			for (int mw : {4, 8}) {
				if (!back_minw_supported(mw, p->zr, p->fmt))
					continue;
				p->minw = mw;
				float t = 0.f;
				rc = time_launches([&](hipStream_t st) {
					BackArgs a = make_args(p, proj, scale, image);
					a.g.na = std::min(p->na, kTuneViews);
					HIP_TRY(launch_back(a, p->nw, p->zl, p->zr, p->fmt,
						p->math == FDKHIP_MATH_FAST, plan_minw(p), st));
					return FDKHIP_OK;
				}, s, reps, &t);
				if (rc)
					return rc;
				cand.emplace_back(t, Tuned{sh[0], sh[1], p->vstage, mw});
				if (t < best_t) {
					best_t = t;
					best = Tuned{sh[0], sh[1], p->vstage, mw};
				}
			}
		}
	}
	if (p->na > kTuneViews && cand.size() > 1) {
		// the fastest few again over all views: a prefix can rank shapes
		// whose start-up and per-view costs differ in another order
		std::sort(cand.begin(), cand.end(),
			[](const std::pair<float, Tuned> &x, const std::pair<float, Tuned> &y) {
				return x.first < y.first;
			});
		cand.resize(std::min(cand.size(), (size_t) kTuneFinalists));
		best_t = 3.4e38f;
		for (const auto &c : cand) {
			rc = configure(p, c.second.nw, c.second.zl, c.second.vstage, 0);
			if (rc)
				return rc;
			p->minw = c.second.minw;
			float t = 0.f;
			rc = time_launches([&](hipStream_t st) {
				const BackArgs a = make_args(p, proj, scale, image);
				HIP_TRY(launch_back(a, p->nw, p->zl, p->zr, p->fmt,
					p->math == FDKHIP_MATH_FAST, plan_minw(p), st));
				return FDKHIP_OK;
			}, s, reps, &t);
			if (rc)
				return rc;
			if (t < best_t) {
				best_t = t;
				best = c.second;
			}
		}
	}
	rc = configure(p, best.nw, best.zl, best.vstage, 0);
	if (rc)
		return rc;
	p->minw = best.minw;
	tuned_store(tune_key(p), best);
	if (best_ms)
		*best_ms = best_t;
	return FDKHIP_OK;
}

extern "C" int fdkhip_plan_profile(fdkhip_plan *p, const void *proj, const float *scale,
	float *image, void *stream, int32_t reps, float ms[3])
{
	int rc = check_exec(p, proj, scale, image);
	if (rc)
		return rc;
	if (reps < 1 || !ms)
		return fail(FDKHIP_EARG, "reps must be >= 1 and ms non-NULL");
	hipStream_t s = (hipStream_t) stream;
	hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
	for (auto &e : ev)
		if (hipEventCreate(&e) != hipSuccess) {
			rc = fail(FDKHIP_EHIP, "hipEventCreate");
			break;
		}
	double sum[3] = {0, 0, 0};
	for (int r = 0; r < reps && !rc; ++r) {
		rc = plan_launch(p, proj, scale, image, s, ev);
		if (!rc && hipEventSynchronize(ev[3]) != hipSuccess)
			rc = fail(FDKHIP_EHIP, "hipEventSynchronize");
		for (int k = 0; k < 3 && !rc; ++k) {
			float t = 0.f;
			if (hipEventElapsedTime(&t, ev[k], ev[k + 1]) != hipSuccess)
				rc = fail(FDKHIP_EHIP, "hipEventElapsedTime");
			sum[k] += t;
		}
	}
	for (auto &e : ev)
		if (e)
			(void) hipEventDestroy(e);
	if (!rc)
		for (int k = 0; k < 3; ++k)
			ms[k] = (float) (sum[k] / reps);
	return rc;
}


// apply a choice made earlier in this process for the same geometry
static int apply_tuned(fdkhip_plan *p)
{
	Tuned t;
	if (!tuned_lookup(tune_key(p), t))
		return FDKHIP_OK;
	p->minw = t.minw;
	return configure(p, t.nw, t.zl, t.vstage, 0);
}

// a plan's own launch shape: 8 waves x 16 slices = 128-slice z-groups (4 x 16
// when the support is short), or the autotuner's choice made earlier in this
// process for the same geometry
static int plan_default_config(fdkhip_plan *p)
{
	const int nw = (p->zs_hi - p->zs_lo <= 64) ? 4 : 8;
	const int rc = configure(p, nw, 16, 0, 0);
	return rc ? rc : apply_tuned(p);
}

// ---------------------------------------------------------------------------
// batches of scans: one persistent launch driven by a device work queue, or
// one launch per scan

struct fdkhip_batch {
	int nscan = 0, fmt = FMT_F32, math = FDKHIP_MATH_FAST;
	std::vector<fdkhip_plan *> plans;
	// one back_kernel launch per scan, each plan with its own launch shape,
	// instead of the persistent kernel (fdkhip_batch_set_per_scan); the
	// launches go to two streams (the caller's and `side`), joined again
	bool per_scan = false;
	hipStream_t side = nullptr;
	hipEvent_t fork_ev = nullptr, join_ev = nullptr;
	int nw = 8, zl = 16, vstage = 8, slot = 64, nwg = 256;
	int vstage_req = 0;   // requested views per stage (0: largest that fits)
	// per-scan slices per consumer wave: chosen by batch_configure
	// (zper_req empty) or fixed by fdkhip_batch_set_zper
	std::vector<int> zper_req;
	long long nitems = 0;
	std::vector<long long> item_begin;
	std::vector<BackArgs> host_args;
	BackArgs *d_scans = nullptr;
	long long *d_item_begin = nullptr;
	unsigned long long *d_counter = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t gexec = nullptr;
	hipStream_t cap_stream = nullptr;
	std::vector<const void *> g_proj;
	std::vector<const float *> g_scale;
	std::vector<float *> g_image;
};

static int batch_device(const fdkhip_batch *b)
{
	int dev = 0;
	if (!b->plans.empty() && b->plans[0])
		return b->plans[0]->device;
	(void) hipGetDevice(&dev);
	return dev;
}

static void batch_drop_graph(fdkhip_batch *b)
{
	if (b->gexec || b->graph)
		quiesce(batch_device(b));
	if (b->gexec) (void) hipGraphExecDestroy(b->gexec);
	if (b->graph) (void) hipGraphDestroy(b->graph);
	b->gexec = nullptr;
	b->graph = nullptr;
	b->g_proj.clear();
	b->g_scale.clear();
	b->g_image.clear();
}

static int batch_free(fdkhip_batch *b)
{
	if (!b)
		return FDKHIP_OK;
	quiesce(batch_device(b));
	batch_drop_graph(b);
	if (b->cap_stream) (void) hipStreamDestroy(b->cap_stream);
	if (b->side) (void) hipStreamDestroy(b->side);
	if (b->fork_ev) (void) hipEventDestroy(b->fork_ev);
	if (b->join_ev) (void) hipEventDestroy(b->join_ev);
	for (auto *p : b->plans)
		plan_free(p);
	(void) hipFree(b->d_scans);
	(void) hipFree(b->d_item_begin);
	(void) hipFree(b->d_counter);
	delete b;
	return FDKHIP_OK;
}

// Half the slices per consumer wave, the batch kernel's second height for a
// launch of zl slices per wave (zl itself where no half height is built).
static int batch_half_zper(int zl)
{
	return zl >= 16 ? zl / 2 : zl;
}

// Launch shape and LDS layout for the batch, and each scan's z-group height.
// Without zper_req, every scan starts at zl slices per wave; the LDS layout
// is chosen for the largest footprint, and a scan whose footprint at zl
// exceeds that layout's slot moves to zl/2 (its z-groups half as tall, its
// footprints smaller), until no scan at zl overflows.  The layout is then
// chosen again for the heights kept.
static int batch_configure(fdkhip_batch *b, int nw, int zl, int vstage_req)
{
	if (!back_config_supported(nw, zl, 1, b->fmt))
		return fail(FDKHIP_EUNSUPPORTED, "configuration waves=%d zper=%d not built", nw, zl);
	const int zh = batch_half_zper(zl);
	if (!b->zper_req.empty())
		for (int i = 0; i < b->nscan; ++i)
			if (b->zper_req[i] != zl && b->zper_req[i] != zh)
				return fail(FDKHIP_EARG, "scan %d: zper %d is neither %d nor %d", i,
					b->zper_req[i], zl, zh);
	const int n = b->nscan;
	std::vector<int> area_hi(n, 0), area_lo(n, 0), zper(n, zl);
	for (int i = 0; i < n; ++i) {
		fdkhip_plan *p = b->plans[i];
		int rc = configure(p, nw, zl, 0, 0);
		if (rc)
			return rc;
		area_hi[i] = p->maxarea;
		area_lo[i] = p->maxarea;
		if (zh != zl) {
			rc = configure(p, nw, zh, 0, 0);
			if (rc)
				return rc;
			area_lo[i] = p->maxarea;
		}
		if (!b->zper_req.empty())
			zper[i] = b->zper_req[i];
	}
	const size_t lds_max = b->plans.empty() ? 65536 : b->plans[0]->lds_max;
	const size_t lds_limit = (nw >= 8) ? lds_max : lds_max / 2;
	int vstage = 0, slot = 0;
	for (;;) {
		int maxarea = 64;
		for (int i = 0; i < n; ++i)
			maxarea = std::max(maxarea, zper[i] == zl ? area_hi[i] : area_lo[i]);
		if (!choose_layout(nw, b->fmt, maxarea, vstage_req, 0, lds_limit, vstage, slot))
			return fail(FDKHIP_EUNSUPPORTED, "LDS budget too small for waves=%d", nw);
		bool moved = false;
		for (int i = 0; b->zper_req.empty() && zh != zl && i < n; ++i)
			if (zper[i] == zl && area_hi[i] > slot) {
				zper[i] = zh;
				moved = true;
			}
		if (!moved)
			break;
	}
	b->nw = nw;
	b->zl = zl;
	b->vstage_req = vstage_req;
	b->vstage = vstage;
	b->slot = slot;
	b->item_begin.assign(1, 0);
	for (int i = 0; i < n; ++i) {
		fdkhip_plan *p = b->plans[i];
		int rc = configure(p, nw, zper[i], 0, 0);   // survey cached above
		if (rc)
			return rc;
		p->vstage = vstage;
		p->slot = slot;
		const long long items = (long long) p->tiles_x * p->tiles_y
			* ((p->na > 0) ? p->ngroups : 0);
		b->item_begin.push_back(b->item_begin.back() + items);
	}
	b->nitems = b->item_begin.back();
	int cus = 0, dev = 0;
	HIP_TRY(hipGetDevice(&dev));
	if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess
		|| cus <= 0)
		cus = 256;
	b->nwg = cus * batch_blocks_per_cu(nw, zl, vstage, slot, b->fmt,
		b->math == FDKHIP_MATH_FAST);
	b->per_scan = false;
	(void) hipFree(b->d_item_begin);
	b->d_item_begin = nullptr;
	HIP_TRY(hipMalloc(&b->d_item_begin, sizeof(long long) * b->item_begin.size()));
	HIP_TRY(hipMemcpy(b->d_item_begin, b->item_begin.data(),
		sizeof(long long) * b->item_begin.size(), hipMemcpyHostToDevice));
	batch_drop_graph(b);
	return FDKHIP_OK;
}

extern "C" int fdkhip_batch_create(fdkhip_batch **out, int32_t nscan, const fdkhip_geom *geoms,
	const int32_t *na, const double *const *beta, const uint8_t *const *mask, int32_t fmt,
	int32_t math)
{
	if (!out)
		return fail(FDKHIP_EARG, "batch output pointer is NULL");
	*out = nullptr;
	if (nscan < 1 || !geoms || !na || !beta)
		return fail(FDKHIP_EARG, "a batch needs nscan >= 1 and per-scan geometry, na, beta");
	fdkhip_batch *b = new fdkhip_batch;
	b->nscan = nscan;
	b->fmt = fmt;
	b->math = (fmt == FDKHIP_F32) ? math : FDKHIP_MATH_FAST;
	for (int i = 0; i < nscan; ++i) {
		fdkhip_plan *p = nullptr;
		int rc = fdkhip_plan_create(&p, &geoms[i], na[i], beta[i], mask ? mask[i] : nullptr,
			fmt, math);
		if (rc) {
			batch_free(b);
			return rc;
		}
		b->plans.push_back(p);
	}
	int rc = batch_configure(b, 8, 16, 0);
	if (rc) {
		batch_free(b);
		return rc;
	}
	hipError_t e = hipMalloc(&b->d_scans, sizeof(BackArgs) * nscan);
	if (e == hipSuccess)
		e = hipMalloc(&b->d_counter, sizeof(unsigned long long));
	if (e != hipSuccess) {
		batch_free(b);
		return fail(FDKHIP_ENOMEM, "device allocation failed: %s", hipGetErrorString(e));
	}
	*out = b;
	return FDKHIP_OK;
}

extern "C" int fdkhip_batch_destroy(fdkhip_batch *b)
{
	return batch_free(b);
}

extern "C" int fdkhip_batch_config(const fdkhip_batch *b, int32_t cfg[6])
{
	if (!b || !cfg)
		return fail(FDKHIP_EARG, "NULL argument");
	if (b->per_scan) {
		long long wg = 0;
		for (const auto *p : b->plans)
			wg += (p->na > 0) ? (long long) p->tiles_x * p->tiles_y * p->ngroups : 0;
		cfg[0] = cfg[1] = cfg[2] = cfg[3] = 0;
		cfg[4] = cfg[5] = (int32_t) std::min(wg, (long long) 0x7fffffff);
		return FDKHIP_OK;
	}
	cfg[0] = b->nw;
	cfg[1] = b->zl;
	cfg[2] = b->vstage;
	cfg[3] = b->slot;
	cfg[4] = b->nwg;
	cfg[5] = (int32_t) std::min(b->nitems, (long long) 0x7fffffff);
	return FDKHIP_OK;
}

extern "C" int fdkhip_batch_set_per_scan(fdkhip_batch *b)
{
	if (!b)
		return fail(FDKHIP_EARG, "batch is NULL");
	for (auto *p : b->plans) {
		const int rc = plan_default_config(p);
		if (rc)
			return rc;
	}
	b->per_scan = true;
	batch_drop_graph(b);
	return FDKHIP_OK;
}

extern "C" int fdkhip_batch_set_config(fdkhip_batch *b, int32_t waves, int32_t zper,
	int32_t vstage)
{
	if (!b)
		return fail(FDKHIP_EARG, "batch is NULL");
	if (vstage < 0 || vstage > 64)
		return fail(FDKHIP_EARG, "views per stage out of range (%d)", vstage);
	b->zper_req.clear();
	return batch_configure(b, waves, zper, vstage);
}

extern "C" int fdkhip_batch_set_zper(fdkhip_batch *b, const int32_t *zper)
{
	if (!b)
		return fail(FDKHIP_EARG, "batch is NULL");
	const std::vector<int> keep = b->zper_req;
	if (zper)
		b->zper_req.assign(zper, zper + b->nscan);
	else
		b->zper_req.clear();
	int rc = batch_configure(b, b->nw, b->zl, b->vstage_req);
	if (rc) {
		b->zper_req = keep;
		(void) batch_configure(b, b->nw, b->zl, b->vstage_req);
	}
	return rc;
}

extern "C" int fdkhip_batch_scan_zper(const fdkhip_batch *b, int32_t *zper)
{
	if (!b || !zper)
		return fail(FDKHIP_EARG, "NULL argument");
	for (int i = 0; i < b->nscan; ++i)
		zper[i] = b->plans[i]->zl * b->plans[i]->zr;
	return FDKHIP_OK;
}

extern "C" int fdkhip_batch_count_updates(fdkhip_batch *b, int64_t *n)
{
	if (!b || !n)
		return fail(FDKHIP_EARG, "NULL argument");
	long long s = 0;
	for (auto *p : b->plans)
		s += p->n_updates;
	*n = s;
	return FDKHIP_OK;
}

// the backprojection alone on the uploaded arguments: the persistent kernel
// (with its queue reset), or each scan's own launch.  The per-scan launches
// go to two streams, `s` and the batch's side stream (forked from `s` and
// joined back into it), the scans in descending order of update count each
// to the stream with fewer updates queued: the next scan's workgroups then
// fill the CUs that a launch's last round leaves idle.
static int batch_kernel_launch(fdkhip_batch *b, hipStream_t s)
{
	if (b->per_scan) {
		if (!b->side) {
			HIP_TRY(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
			HIP_TRY(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming));
			HIP_TRY(hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming));
		}
		std::vector<int> order;
		for (int i = 0; i < b->nscan; ++i)
			if (b->plans[i]->ngroups > 0 && b->plans[i]->na > 0)
				order.push_back(i);
		std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
			return b->plans[x]->n_updates > b->plans[y]->n_updates; });
		HIP_TRY(hipEventRecord(b->fork_ev, s));
		HIP_TRY(hipStreamWaitEvent(b->side, b->fork_ev, 0));
		long long queued[2] = {0, 0};
		for (int i : order) {
			const fdkhip_plan *p = b->plans[i];
			const int k = queued[1] < queued[0] ? 1 : 0;
			queued[k] += p->n_updates;
			HIP_TRY(launch_back(b->host_args[i], p->nw, p->zl, p->zr, p->fmt,
				p->math == FDKHIP_MATH_FAST, plan_minw(p), k ? b->side : s));
		}
		HIP_TRY(hipEventRecord(b->join_ev, b->side));
		HIP_TRY(hipStreamWaitEvent(s, b->join_ev, 0));
		return FDKHIP_OK;
	}
	if (b->nitems == 0)
		return FDKHIP_OK;
	BatchArgs ba;
	ba.scans = b->d_scans;
	ba.item_begin = b->d_item_begin;
	ba.nscan = b->nscan;
	ba.nitems = b->nitems;
	ba.counter = b->d_counter;
	HIP_TRY(launch_batch(ba, b->nwg, b->nw, b->zl, b->vstage, b->slot, b->fmt,
		b->math == FDKHIP_MATH_FAST, s));
	return FDKHIP_OK;
}

// Per-scan zero fills and padded copies, then the backprojection
// (batch_kernel_launch).  With `ev`, ev[0..2] are recorded before the
// per-scan steps, between them and the backprojection, and after it.
static int batch_launch(fdkhip_batch *b, const void *const *proj, const float *const *scale,
	float *const *image, hipStream_t s, hipEvent_t *ev = nullptr)
{
	if (ev)
		HIP_TRY(hipEventRecord(ev[0], s));
	for (int i = 0; i < b->nscan; ++i) {
		fdkhip_plan *p = b->plans[i];
		const long ncol = (long) p->geom.nx * p->geom.ny;
		if (ncol == 0 || p->geom.nz == 0)
			continue;
		const bool work = p->ngroups > 0 && p->na > 0;
		const int zend = std::min(p->z0 + p->ngroups * p->nw * p->zl * p->zr, p->geom.nz);
		HIP_TRY(launch_fill(image[i], ncol, p->geom.nz, work ? p->z0 : 0, work ? zend : 0, s));
		if (work && p->fmt == FDKHIP_F32)
			HIP_TRY(launch_pad(static_cast<const float *>(proj[i]),
				static_cast<float *>(p->d_padded), p->geom.ns, p->geom.nt, p->na, s));
	}
	if (ev)
		HIP_TRY(hipEventRecord(ev[1], s));
	int rc = batch_kernel_launch(b, s);
	if (rc)
		return rc;
	if (ev)
		HIP_TRY(hipEventRecord(ev[2], s));
	return FDKHIP_OK;
}

static int batch_check(fdkhip_batch *b, const void *const *proj, const float *const *scale,
	float *const *image)
{
	if (!b)
		return fail(FDKHIP_EARG, "batch is NULL");
	if (!proj || !image || (b->fmt != FDKHIP_F32 && !scale))
		return fail(FDKHIP_EARG, "per-scan proj/image (and scale for F16/BF16) arrays required");
	for (int i = 0; i < b->nscan; ++i) {
		int rc = check_exec(b->plans[i], proj[i], scale ? scale[i] : nullptr, image[i]);
		if (rc)
			return rc;
	}
	return FDKHIP_OK;
}

// per-scan kernel arguments for these buffers -> device (ordered on `s`)
static int batch_upload_args(fdkhip_batch *b, const void *const *proj,
	const float *const *scale, float *const *image, hipStream_t s, int na_cap = 0)
{
	b->host_args.resize(b->nscan);
	for (int i = 0; i < b->nscan; ++i) {
		b->host_args[i] = make_args(b->plans[i], proj[i], scale ? scale[i] : nullptr, image[i]);
		if (na_cap > 0)
			b->host_args[i].g.na = std::min(b->host_args[i].g.na, na_cap);
	}
	HIP_TRY(hipMemcpyAsync(b->d_scans, b->host_args.data(), sizeof(BackArgs) * b->nscan,
		hipMemcpyHostToDevice, s));
	HIP_TRY(hipStreamSynchronize(s));   // host_args may be rewritten by the next call
	return FDKHIP_OK;
}

extern "C" int fdkhip_batch_execute_direct(fdkhip_batch *b, const void *const *proj,
	const float *const *scale, float *const *image, void *stream)
{
	int rc = batch_check(b, proj, scale, image);
	if (rc)
		return rc;
	hipStream_t s = (hipStream_t) stream;
	rc = batch_upload_args(b, proj, scale, image, s);
	if (rc)
		return rc;
	batch_drop_graph(b);   // the uploaded arguments may differ from the graph's
	return batch_launch(b, proj, scale, image, s);
}

extern "C" int fdkhip_batch_execute(fdkhip_batch *b, const void *const *proj,
	const float *const *scale, float *const *image, void *stream)
{
	int rc = batch_check(b, proj, scale, image);
	if (rc)
		return rc;
	hipStream_t s = (hipStream_t) stream;
	bool same = b->gexec && (int) b->g_proj.size() == b->nscan;
	for (int i = 0; same && i < b->nscan; ++i)
		same = b->g_proj[i] == proj[i] && b->g_image[i] == image[i]
			&& b->g_scale[i] == (scale ? scale[i] : nullptr);
	if (!same) {
		batch_drop_graph(b);
		rc = batch_upload_args(b, proj, scale, image, s);
		if (rc)
			return rc;
		if (!b->cap_stream)
			HIP_TRY(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
		HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
		rc = batch_launch(b, proj, scale, image, b->cap_stream);
		hipGraph_t graph = nullptr;
		hipError_t e = hipStreamEndCapture(b->cap_stream, &graph);
		if (rc)
			return rc;
		if (e != hipSuccess)
			return fail(FDKHIP_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
		b->graph = graph;
		HIP_TRY(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
		for (int i = 0; i < b->nscan; ++i) {
			b->g_proj.push_back(proj[i]);
			b->g_scale.push_back(scale ? scale[i] : nullptr);
			b->g_image.push_back(image[i]);
		}
	}
	HIP_TRY(hipGraphLaunch(b->gexec, s));
	return FDKHIP_OK;
}

extern "C" int fdkhip_batch_autotune(fdkhip_batch *b, const void *const *proj,
	const float *const *scale, float *const *image, void *stream, int32_t reps, float *best_ms)
{
	int rc = batch_check(b, proj, scale, image);
	if (rc)
		return rc;
	if (reps < 1)
		return fail(FDKHIP_EARG, "reps must be >= 1");
	hipStream_t s = (hipStream_t) stream;
	int bn = b->nw, bz = b->zl, bv = b->vstage;
	std::vector<int> breq = b->zper_req;
	float best_t = 3.4e38f;
	// zero fills and padded views once; then the persistent kernel alone over
	// the first kTuneViews views of every scan
	rc = batch_upload_args(b, proj, scale, image, s);
	if (!rc)
		rc = batch_launch(b, proj, scale, image, s);
	if (rc)
		return rc;
	// per shape: the chosen per-scan heights, then (where that choice moved
	// some scan to the half height) every scan at the shape's own height
	for (const auto &sh : kShapes) {
		bool moved = false;
		for (int pass = 0; pass < 2; ++pass) {
			if (pass == 1 && !moved)
				break;
			if (pass == 0)
				b->zper_req.clear();
			else
				b->zper_req.assign(b->nscan, sh[1]);
			int last_vstage = -1;
			for (int vs : kStages) {
				if (batch_configure(b, sh[0], sh[1], vs) != FDKHIP_OK)
					continue;
				for (auto *p : b->plans)
					moved = moved || p->zl != sh[1];
				if (b->vstage == last_vstage)
					continue;
				last_vstage = b->vstage;
				rc = batch_upload_args(b, proj, scale, image, s, kTuneViews);
				if (rc)
					return rc;
				float t = 0.f;
				rc = time_launches([&](hipStream_t st) {
					return batch_kernel_launch(b, st);
				}, s, reps, &t);
				if (rc)
					return rc;
				if (t < best_t) {
					best_t = t;
					bn = sh[0];
					bz = sh[1];
					bv = b->vstage;
					breq = b->zper_req;
				}
			}
		}
	}
	// one launch per scan, each plan tuned alone on its own buffers
	for (int i = 0; i < b->nscan; ++i) {
		rc = fdkhip_plan_autotune(b->plans[i], proj[i], scale ? scale[i] : nullptr, image[i],
			stream, reps, nullptr);
		if (rc)
			return rc;
	}
	b->per_scan = true;
	rc = batch_upload_args(b, proj, scale, image, s, kTuneViews);
	if (rc)
		return rc;
	float t_scan = 0.f;
	rc = time_launches([&](hipStream_t st) {
		return batch_kernel_launch(b, st);
	}, s, reps, &t_scan);
	if (rc)
		return rc;
	if (t_scan < best_t) {
		best_t = t_scan;
		batch_drop_graph(b);
	} else {
		b->zper_req = breq;
		rc = batch_configure(b, bn, bz, bv);
		if (rc)
			return rc;
	}
	if (best_ms)
		*best_ms = best_t;
	return FDKHIP_OK;
}

extern "C" int fdkhip_batch_profile(fdkhip_batch *b, const void *const *proj,
	const float *const *scale, float *const *image, void *stream, int32_t reps, float ms[2])
{
	int rc = batch_check(b, proj, scale, image);
	if (rc)
		return rc;
	if (reps < 1 || !ms)
		return fail(FDKHIP_EARG, "reps must be >= 1 and ms non-NULL");
	hipStream_t s = (hipStream_t) stream;
	rc = batch_upload_args(b, proj, scale, image, s);
	if (rc)
		return rc;
	batch_drop_graph(b);
	hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
	for (auto &e : ev)
		if (hipEventCreate(&e) != hipSuccess) {
			rc = fail(FDKHIP_EHIP, "hipEventCreate");
			break;
		}
	double sum[2] = {0, 0};
	for (int r = 0; r < reps && !rc; ++r) {
		rc = batch_launch(b, proj, scale, image, s, ev);
		if (!rc && hipEventSynchronize(ev[2]) != hipSuccess)
			rc = fail(FDKHIP_EHIP, "hipEventSynchronize");
		for (int k = 0; k < 2 && !rc; ++k) {
			float t = 0.f;
			if (hipEventElapsedTime(&t, ev[k], ev[k + 1]) != hipSuccess)
				rc = fail(FDKHIP_EHIP, "hipEventElapsedTime");
			sum[k] += t;
		}
	}
	for (auto &e : ev)
		if (e)
			(void) hipEventDestroy(e);
	if (!rc)
		for (int k = 0; k < 2; ++k)
			ms[k] = (float) (sum[k] / reps);
	return rc;
}
