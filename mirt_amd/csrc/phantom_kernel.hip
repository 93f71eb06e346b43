// phantom_kernel.hip -- analytic cone-beam line integrals of ellipsoids on
// gfx950: ellipsoid_proj (fbp/ellipsoid_proj.m:35-125) with the
// detector-to-parallel coordinate maps of fbp/ir_coord_cb_flat_to_par.m and
// fbp/ir_coord_cb_arc_to_par.m, in FP64 as the MATLAB computes it.
//
// One lane per detector sample (s fastest, so the [ns nt na] writes are
// coalesced), each lane walking a block of views: the sample's parallel-beam
// coordinates (uu, vv, azimuth, polar angle) do not depend on the view and
// are computed once per block; per view one sincos of the azimuth; per
// ellipsoid the chord 2 val sqrt(B^2 - A C) / A of ellipsoid_proj.m:110-118,
// 0 where the ray misses (the "real part" trick of ellipsoid_proj.m:123).
// The ellipsoid table sits in LDS.  With oversampling (ellipsoid_proj.m:44-58)
// each lane evaluates its sample's No x No sub-samples and averages them in
// downsample3's order (along s, then along t).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "fdkhip.h"

namespace fdk {

// per ellipsoid: centre, 1/r^2 per axis, cos/sin of the xy rotation, 2*val
struct EllRow {
	double cx, cy, cz, irx2, iry2, irz2, cxa, sxa, val2;
};

enum { ELL_PAR = 0, ELL_FLAT = 1, ELL_ARC = 2 };
constexpr int ELL_MAX = 512;        // ellipsoids per launch (LDS table)
constexpr int ELL_VIEWS = 8;        // views per lane

struct EllArgs {
	const double *s, *t;             // [ns], [nt] detector sample positions
	const double *beta, *zs;         // [na] source angle (radians), source height
	const EllRow *ell;               // [nell]
	int ns, nt, na, nell, no;        // no: oversampling factor
	double ds, dt, dso, dod;         // sample spacing (oversampling), distances
	int kind;
	double *proj;                    // [na][nt][ns]
};

// (uu, vv, azim, polar) of a detector sample at beta = 0
__device__ __forceinline__ void par_coords(const EllArgs &a, double ss, double tt,
	double &uu, double &vv, double &azim, double &polar)
{
	if (a.kind == ELL_PAR) {
		uu = ss;
		vv = tt;
		azim = 0.0;
		polar = 0.0;
		return;
	}
	const double Ds = a.dso, Dc = a.dso + a.dod;
	if (a.kind == ELL_FLAT) {        // ir_coord_cb_flat_to_par.m
		const double r2 = Dc * Dc + ss * ss;
		const double r = sqrt(r2);
		uu = Ds * ss / r;
		vv = Ds * tt / sqrt(r2 + tt * tt) * Dc / r;
		polar = -atan(tt / r);
		azim = atan(ss / Dc);
		return;
	}
	// arc (dfs = 0), ir_coord_cb_arc_to_par.m: focal radius Dc, source at (0, dso, 0)
	const double Rf = Dc;
	const double p1 = Rf * sin(ss / Rf);
	const double p2 = Ds - Rf * cos(ss / Rf);
	const double e1 = p1, e2 = p2 - Ds, e3 = tt;
	const double en = sqrt(e1 * e1 + e2 * e2 + e3 * e3);
	polar = -asin(e3 / en);
	azim = -atan((e1 / en) / (e2 / en));
	uu = sin(azim) * Ds;
	vv = (-cos(azim) * Ds) * sin(polar);
}

// sum over ellipsoids of the chord lengths along one ray
__device__ __forceinline__ double ray_sum(const EllRow *E, int nell, double uu, double vv,
	double caz0, double saz0, double cpol, double spol, double zs)
{
	double acc = 0.0;
	for (int k = 0; k < nell; ++k) {
		const EllRow e = E[k];
		const double cz_eff = e.cz - zs;
		const double ushift = e.cx * caz0 + e.cy * saz0;
		const double vshift = (e.cx * saz0 - e.cy * caz0) * spol + cz_eff * cpol;
		// az - xang by the angle-difference identity
		const double caz = caz0 * e.cxa + saz0 * e.sxa;
		const double saz = saz0 * e.cxa - caz0 * e.sxa;
		const double du = uu - ushift, dv = vv - vshift;
		const double p1 = du * caz + dv * saz * spol;
		const double p2 = du * saz - dv * caz * spol;
		const double p3 = dv * cpol;
		const double e1 = -saz * cpol, e2 = caz * cpol, e3 = spol;
		const double A = e1 * e1 * e.irx2 + e2 * e2 * e.iry2 + e3 * e3 * e.irz2;
		const double B = p1 * e1 * e.irx2 + p2 * e2 * e.iry2 + p3 * e3 * e.irz2;
		const double C = p1 * p1 * e.irx2 + p2 * p2 * e.iry2 + p3 * p3 * e.irz2 - 1.0;
		const double d = B * B - A * C;
		if (d > 0.0)
			acc += e.val2 * sqrt(d) / A;
	}
	return acc;
}

__global__ void __launch_bounds__(256)
ellipsoid_proj_kernel(const EllArgs a)
{
	__shared__ EllRow E[ELL_MAX];
	for (int k = threadIdx.x; k < a.nell; k += blockDim.x)
		E[k] = a.ell[k];
	__syncthreads();
	const long nst = (long) a.ns * a.nt;
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nst)
		return;
	const int is = (int) (i % a.ns), it = (int) (i / a.ns);
	const int v0 = blockIdx.y * ELL_VIEWS;
	const int nv = min(ELL_VIEWS, a.na - v0);
	double sum[ELL_VIEWS];
#pragma unroll
	for (int v = 0; v < ELL_VIEWS; ++v)
		sum[v] = 0.0;
	const int no = a.no;
	for (int jt = 0; jt < no; ++jt) {
		double row[ELL_VIEWS];       // this t sub-row, summed over the s sub-samples
#pragma unroll
		for (int v = 0; v < ELL_VIEWS; ++v)
			row[v] = 0.0;
		// sub-sample offsets of ellipsoid_proj.m:50-51: [-(No-1):2:(No-1)]/(2 No) * d
		const double tt = (no > 1) ? ((double) (2 * jt - (no - 1)) / (2.0 * no)) * a.dt + a.t[it]
			: a.t[it];
		for (int js = 0; js < no; ++js) {
			const double ss = (no > 1)
				? ((double) (2 * js - (no - 1)) / (2.0 * no)) * a.ds + a.s[is] : a.s[is];
			double uu, vv, azim, polar;
			par_coords(a, ss, tt, uu, vv, azim, polar);
			double spol, cpol;
			sincos(polar, &spol, &cpol);
#pragma unroll
			for (int v = 0; v < ELL_VIEWS; ++v) {
				if (v < nv) {
					double saz0, caz0;
					sincos(a.beta[v0 + v] + azim, &saz0, &caz0);
					row[v] += ray_sum(E, a.nell, uu, vv, caz0, saz0, cpol, spol, a.zs[v0 + v]);
				}
			}
		}
#pragma unroll
		for (int v = 0; v < ELL_VIEWS; ++v)
			sum[v] += (no > 1) ? row[v] / no : row[v];
	}
#pragma unroll
	for (int v = 0; v < ELL_VIEWS; ++v)
		if (v < nv)
			a.proj[(long) (v0 + v) * nst + i] = (no > 1) ? sum[v] / no : sum[v];
}

} // namespace fdk

using namespace fdk;

int fdkhip_fail_msg(int code, const char *msg);   // fdk_api.hip

#define ELL_TRY(expr)                                                             \
	do {                                                                          \
		hipError_t e_ = (expr);                                                   \
		if (e_ != hipSuccess) {                                                   \
			rc = fdkhip_fail_msg(FDKHIP_EHIP, hipGetErrorString(e_));             \
			goto done;                                                            \
		}                                                                         \
	} while (0)

extern "C" int fdkhip_ellipsoid_proj(const double *s, int32_t ns, const double *t, int32_t nt,
	const double *beta, const double *source_zs, int32_t na, double dso, double dod,
	double dfs, const double *params, int32_t nell, int32_t oversample, double *proj_dev,
	void *stream)
{
	if (ns < 0 || nt < 0 || na < 0 || nell < 0)
		return fdkhip_fail_msg(FDKHIP_EARG, "negative size");
	if (nell > ELL_MAX)
		return fdkhip_fail_msg(FDKHIP_EUNSUPPORTED, "at most 512 ellipsoids per call");
	if (oversample < 1)
		return fdkhip_fail_msg(FDKHIP_EARG, "oversample must be >= 1");
	if ((long) ns * nt * na == 0)
		return FDKHIP_OK;
	if (!s || !t || !beta || !proj_dev || (nell > 0 && !params))
		return fdkhip_fail_msg(FDKHIP_EARG, "NULL argument");
	int kind;
	if (isinf(dso))
		kind = ELL_PAR;
	else if (isinf(dfs))
		kind = ELL_FLAT;
	else if (dfs == 0)
		kind = ELL_ARC;
	else
		return fdkhip_fail_msg(FDKHIP_EUNSUPPORTED, "not done");   // ellipsoid_proj.m:81
	if (oversample > 1 && ((ns > 1 && !(s[1] != s[0])) || (nt > 1 && !(t[1] != t[0]))))
		return fdkhip_fail_msg(FDKHIP_EARG, "uniform spacing required for oversampling");
	if (oversample > 1) {   // ellipsoid_proj.m:44-48
		const double ds = ns > 1 ? s[1] - s[0] : 0, dt = nt > 1 ? t[1] - t[0] : 0;
		for (int i = 1; i < ns; ++i)
			if (fabs((s[i] - s[i - 1]) / ds - 1) > 1e-6)
				return fdkhip_fail_msg(FDKHIP_EARG, "uniform spacing required for oversampling");
		for (int i = 1; i < nt; ++i)
			if (fabs((t[i] - t[i - 1]) / dt - 1) > 1e-6)
				return fdkhip_fail_msg(FDKHIP_EARG, "uniform spacing required for oversampling");
	}
	std::vector<EllRow> rows((size_t) nell);
	for (int k = 0; k < nell; ++k) {
		const double *p = params + 9 * (size_t) k;
		if (p[7] != 0)
			return fdkhip_fail_msg(FDKHIP_EUNSUPPORTED, "z rotation not done");   // :94
		const double xang = p[6] * (M_PI / 180);
		rows[k] = EllRow{p[0], p[1], p[2], 1 / (p[3] * p[3]), 1 / (p[4] * p[4]),
			1 / (p[5] * p[5]), cos(xang), sin(xang), 2 * p[8]};
	}
	hipStream_t st = (hipStream_t) stream;
	int rc = FDKHIP_OK;
	double *d_s = nullptr, *d_t = nullptr, *d_b = nullptr, *d_z = nullptr;
	EllRow *d_e = nullptr;
	{
		std::vector<double> zs(source_zs ? source_zs : beta, (source_zs ? source_zs : beta) + na);
		if (!source_zs)
			std::fill(zs.begin(), zs.end(), 0.0);
		ELL_TRY(hipMalloc(&d_s, sizeof(double) * ns));
		ELL_TRY(hipMalloc(&d_t, sizeof(double) * nt));
		ELL_TRY(hipMalloc(&d_b, sizeof(double) * na));
		ELL_TRY(hipMalloc(&d_z, sizeof(double) * na));
		ELL_TRY(hipMalloc(&d_e, sizeof(EllRow) * std::max(nell, 1)));
		ELL_TRY(hipMemcpyAsync(d_s, s, sizeof(double) * ns, hipMemcpyHostToDevice, st));
		ELL_TRY(hipMemcpyAsync(d_t, t, sizeof(double) * nt, hipMemcpyHostToDevice, st));
		ELL_TRY(hipMemcpyAsync(d_b, beta, sizeof(double) * na, hipMemcpyHostToDevice, st));
		ELL_TRY(hipMemcpyAsync(d_z, zs.data(), sizeof(double) * na, hipMemcpyHostToDevice, st));
		if (nell > 0)
			ELL_TRY(hipMemcpyAsync(d_e, rows.data(), sizeof(EllRow) * nell,
				hipMemcpyHostToDevice, st));
		EllArgs a;
		a.s = d_s;
		a.t = d_t;
		a.beta = d_b;
		a.zs = d_z;
		a.ell = d_e;
		a.ns = ns;
		a.nt = nt;
		a.na = na;
		a.nell = nell;
		a.no = oversample;
		a.ds = ns > 1 ? s[1] - s[0] : 0;
		a.dt = nt > 1 ? t[1] - t[0] : 0;
		a.dso = dso;
		a.dod = dod;
		a.kind = kind;
		a.proj = proj_dev;
		const long nst = (long) ns * nt;
		const dim3 grid((unsigned) ((nst + 255) / 256), (unsigned) ((na + ELL_VIEWS - 1) / ELL_VIEWS));
		hipLaunchKernelGGL(ellipsoid_proj_kernel, grid, dim3(256), 0, st, a);
		ELL_TRY(hipGetLastError());
		// the host tables above are released at return: wait for the copies
		ELL_TRY(hipStreamSynchronize(st));
	}
done:
	(void) hipFree(d_s);
	(void) hipFree(d_t);
	(void) hipFree(d_b);
	(void) hipFree(d_z);
	(void) hipFree(d_e);
	return rc;
}
