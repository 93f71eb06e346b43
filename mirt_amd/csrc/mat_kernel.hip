// mat_kernel.hip -- cbct_back_mat (fbp/cbct_back.m:142-270) on gfx950.
//
// The reference's MATLAB backprojector, the use_mex = 0 route of cbct_back:
// double precision throughout, every voxel of the mask against every view
// 1:ia_skip:na, with its own edge rules (cbct_back.m:221-233):
//   * s: a sample position bs outside [1, ns] (1-based) reads the extra zero
//     column ns+1, so it adds nothing (no one-bin fall-off as in the mex);
//   * t: bt is clamped to [1, nt] (replication of the edge rows) and bt = nt
//     reads rows nt-1 and nt with weight 1 on the last;
// and its own weights (cbct_back.m:244-249): mag^2 for flat and parallel
// detectors, dsd^2 / (r_loop^2 + y_betas^2) for the arc, r_loop = x_beta -
// offset_source.  Helical source positions (source_zs, one per view) enter
// tprime = mag * (zc - source_zs) as in cbct_back.m:219.
//
// Every product, quotient and sum is one IEEE double operation in the order
// the MATLAB expression evaluates it (the library is built with
// -ffp-contract=off, so nothing is fused), views are added to each voxel in
// view order starting from 0, as img2 = img2 + p0 does.  The per-view cos and
// sin come from the host's libm.
//
// The helical version (mex/src/fdk-helix/cbct_back_helix.m:185-350) adds,
// when MatArgs::zrange is set: per-slice view ranges (slice iz uses views
// ia_min:ia_skip:ia_max around the view whose source height is nearest, the
// host computes them), the cone-parallel geometry of rebinned views
// (cone_par: mag = dsd / (sqrt(dso^2 - x_beta^2) - (dso - y_betas)),
// sprime = x_beta, a source height offset asin(-x_beta / dso) *
// source_dz_per_rad, no distance weight) and a per-slice scale.
//
// Launch: one wave per 8 x 8 tile of image columns and MAT_ZC slices; lane =
// column, the lane's MAT_ZC accumulators stay in registers over all views.
// Column geometry (x_beta, y_betas, mag, bs and its weights) is computed once
// per (column, view) and reused for the slices.  The 1-D grid is remapped so
// that neighbouring tiles, whose detector footprints overlap, run on the same
// XCD and share its L2.

#include <hip/hip_runtime.h>

#include "fdk_kernels.h"

namespace fdk {

namespace {

__device__ __forceinline__ int mat_xcd_block(int bid, int nblocks)
{
	const int xcd = bid & 7, local = bid >> 3;
	const int per = nblocks >> 3, rem = nblocks & 7;
	return xcd * per + min(xcd, rem) + local;
}

__device__ __forceinline__ double mat_sample(const float *p, long i)
{
	return (double) p[i];
}

__device__ __forceinline__ double mat_sample(const double *p, long i)
{
	return p[i];
}

// HELIX: the helical backprojector's per-slice view ranges and scales and
// cone-parallel geometry (compiled out of the axial route)
template <typename TP, bool HELIX>
__global__ void __launch_bounds__(64) mat_kernel(const MatArgs a)
{
	const int tx = (a.nx + 7) >> 3, ty = (a.ny + 7) >> 3;
	const int nblk = tx * ty * a.nzc;
	const int b = mat_xcd_block(blockIdx.x, nblk);
	const int bxy = b % (tx * ty), bz = b / (tx * ty);
	const int ix = (bxy % tx) * 8 + (threadIdx.x & 7);
	const int iy = (bxy / tx) * 8 + (threadIdx.x >> 3);
	if (ix >= a.nx || iy >= a.ny)
		return;
	const long nxy = (long) a.nx * a.ny;
	const long col = ix + (long) iy * a.nx;
	const int z0 = bz * MAT_ZC;
	const int nk = min(MAT_ZC, a.nz - z0);
	double *out = a.image + col + (long) z0 * nxy;

	// embed(img2, mask): columns outside the mask are 0 (cbct_back.m:182-183, 261)
	if (a.mask && !a.mask[col]) {
		for (int k = 0; k < nk; ++k)
			out[(long) k * nxy] = 0.0;
		return;
	}

	// cbct_back.m:160-165: xc = ([0:nx-1] - wx) * dx, likewise yc and zc
	const double xc = ((double) ix - a.wx) * a.dx;
	const double yc = ((double) iy - a.wy) * a.dy;
	double zc[MAT_ZC], acc[MAT_ZC];
#pragma unroll
	for (int k = 0; k < MAT_ZC; ++k) {
		zc[k] = ((double) (z0 + k) - a.wz) * a.dz;
		acc[k] = 0.0;
	}
	const double ns1 = (double) a.ns + 1.0, ntd = (double) a.nt;
	const double dsd2 = a.dsd * a.dsd;
	// the views each slice takes (helical: its own range, cbct_back_helix.m:224-236)
	int vlo = 0, vhi = a.nv - 1;
	int lo[MAT_ZC], hi[MAT_ZC];
	if (HELIX) {
		vlo = a.nv;
		vhi = -1;
#pragma unroll
		for (int k = 0; k < MAT_ZC; ++k) {
			lo[k] = k < nk ? a.zrange[z0 + k].x : 0;
			hi[k] = k < nk ? a.zrange[z0 + k].y : -1;
			if (k < nk) {
				vlo = min(vlo, lo[k]);
				vhi = max(vhi, hi[k]);
			}
		}
	}

	for (int v = max(vlo, 0); v <= min(vhi, a.nv - 1); ++v) {
		bool use[MAT_ZC], any = false;
#pragma unroll
		for (int k = 0; k < MAT_ZC; ++k) {
			use[k] = k < nk && (!HELIX || (v >= lo[k] && v <= hi[k] &&
				(v - lo[k]) % a.vstep == 0));
			any = any || use[k];
		}
		if (HELIX && !any)
			continue;
		const double4 vw = a.views[v];              // {cos beta, sin beta, source_zs}
		const TP *pv = (const TP *) a.proj + (long) v * a.view_stride;
		// cbct_back.m:204-205
		const double x_beta = xc * vw.x + yc * vw.y;
		const double y_betas = a.dso - (-xc * vw.y + yc * vw.x);
		// cbct_back.m:208-220, cbct_back_helix.m:245-265
		double mag, sprime, r_loop = 0.0, zoff = 0.0;
		if (HELIX && a.cone_par) {
			mag = a.dsd / (sqrt(a.dso * a.dso - x_beta * x_beta) - (a.dso - y_betas));
			sprime = x_beta;
			zoff = asin(-x_beta / a.dso) * a.dz_per_rad;
		} else if (a.kind == KIND_PAR) {
			mag = 1.0;
			sprime = mag * x_beta;
		} else {
			mag = a.dsd / y_betas;
			if (a.kind == KIND_ARC) {
				r_loop = x_beta - a.offset_source;
				sprime = a.dsd * atan2(r_loop, y_betas);
			} else {
				sprime = mag * x_beta;
			}
		}
		// cbct_back.m:222-233 for s: out-of-range positions read column ns+1 (zero)
		double bs = sprime / a.ds + a.ws;
		if (!(bs >= 1.0 && bs <= (double) a.ns))
			bs = ns1;
		double is = floor(bs);
		if (is == ns1)
			is = (double) a.ns;
		const double wr = bs - is;
		const double wl = 1.0 - wr;
		const int is0 = (int) is - 1;                // 0-based left column
		const bool right_zero = is0 + 1 >= a.ns;      // column ns+1 is the zero column
		// cbct_back.m:244-249 weights
		const double wmag = mag * mag;
		const double wden = r_loop * r_loop + y_betas * y_betas;

#pragma unroll
		for (int k = 0; k < MAT_ZC; ++k) {
			if (use[k]) {
				// cbct_back.m:219-233 for t: clamp to [1, nt], bt = nt uses rows nt-1, nt
				const double tprime = (HELIX && a.cone_par) ? mag * (zc[k] - (vw.z + zoff))
					: mag * (zc[k] - vw.z);
				double bt = tprime / a.dt + a.wt;
				bt = fmax(bt, 1.0);                  // NaN -> 1, as MATLAB's max
				bt = fmin(bt, ntd);
				double it = floor(bt);
				if (it == ntd)
					it = ntd - 1.0;
				const double wu = bt - it;
				const double wd = 1.0 - wu;
				const long r0 = (long) ((int) it - 1) * a.ns + is0;
				const long r1 = r0 + a.ns;
				const double a00 = mat_sample(pv, r0);
				const double a01 = mat_sample(pv, r1);
				const double a10 = right_zero ? 0.0 : mat_sample(pv, r0 + 1);
				const double a11 = right_zero ? 0.0 : mat_sample(pv, r1 + 1);
				// cbct_back.m:236-241
				const double p1 = wl * a00 + wr * a10;
				const double p2 = wl * a01 + wr * a11;
				double p0 = wd * p1 + wu * p2;
				if (!(HELIX && a.cone_par)) {        // cone_par: no weight without w3d
					if (a.kind == KIND_ARC)
						p0 = p0 * dsd2 / wden;
					else
						p0 = p0 * wmag;
				}
				acc[k] = acc[k] + p0;                // cbct_back.m:251
			}
		}
	}
	// cbct_back.m:264-266: (scale) * img, one double product per voxel
	// (cbct_back_helix.m:341-345: the slice's own scale)
#pragma unroll
	for (int k = 0; k < MAT_ZC; ++k)
		if (k < nk)
			out[(long) k * nxy] = ((HELIX && a.zscale) ? a.zscale[z0 + k] : a.scale) * acc[k];
}

} // namespace

hipError_t launch_mat(const MatArgs &a, bool proj_fp64, hipStream_t s)
{
	const long tiles = (long) ((a.nx + 7) / 8) * ((a.ny + 7) / 8) * a.nzc;
	if (tiles <= 0)
		return hipSuccess;
	if (tiles > 0x7fffffffL)
		return hipErrorInvalidValue;
	const bool hx = a.zrange != nullptr;
	if (proj_fp64 && hx)
		hipLaunchKernelGGL((mat_kernel<double, true>), dim3((unsigned) tiles), dim3(64), 0, s, a);
	else if (proj_fp64)
		hipLaunchKernelGGL((mat_kernel<double, false>), dim3((unsigned) tiles), dim3(64), 0, s, a);
	else if (hx)
		hipLaunchKernelGGL((mat_kernel<float, true>), dim3((unsigned) tiles), dim3(64), 0, s, a);
	else
		hipLaunchKernelGGL((mat_kernel<float, false>), dim3((unsigned) tiles), dim3(64), 0, s, a);
	return hipGetLastError();
}

} // namespace fdk
