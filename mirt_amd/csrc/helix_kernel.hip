// helix_kernel.hip -- the helical / cone-parallel 'fdk,ts,back' of the
// fdk-helix kernel on gfx950.
//
// fdk_ts_back1 (mex/src/fdk-helix/mex/fdk,ts.c:55-251) with every branch the
// axial kernel leaves out: a source height per view (wzbeta = wz +
// source_zs/dz), the view window [iz_min_beta, iz_max_beta) of a helical
// scan (fdk,ts,t.c:78-85, computed on the host), the cone-parallel geometry
// of rebinned views (magnification dsd / (sqrt(dso^2 - xbeta^2) - (dso -
// ybetas)), s = xbeta, a height that follows asin(-xbeta/dso)), and the w3d
// weights of the conjugate ray (fdk,ts.c:196-226 with the table lookups of
// fdk,w3d,pow,tab.c:33-56).  Same float operations in the same order as the
// reference, doubles where it computes in double (atan2, asin, the square
// root its pow(., 0.5) takes, the cone-parallel magnification), so the
// result is the reference's bit for bit; the t_bin running sum is replayed
// from iz_min up to the lane's first slice.
//
// Launch: one wave per 8 x 8 tile of columns and HX_ZC slices, lane =
// column, HX_ZC accumulators per lane in LDS over the views of the chunk's
// window (chunk_views, from the host).  Per view the wave stages the
// rectangle of detector columns and rows its lanes read (with the
// reference's zero padding where it leaves the view) into LDS with reads
// that are contiguous along t, then gathers the bilinear samples from LDS;
// a footprint larger than the buffer, or a sample outside it, is read from
// the view directly.
// Deviations, as in the axial kernel: samples whose row is outside
// [-1, nt-1] (which the reference reads from beyond its padded view buffer)
// and voxels behind the source add nothing; w3d table indices past the
// reference's table read 0 (a zero tail, sized on the host).

#include <hip/hip_runtime.h>

#include "fdk_kernels.h"

namespace fdk {

namespace {

__device__ __forceinline__ int hx_xcd_block(int bid, int nblocks)
{
	const int xcd = bid & 7, local = bid >> 3;
	const int per = nblocks >> 3, rem = nblocks & 7;
	return xcd * per + min(xcd, rem) + local;
}

// (int) conversion with the x86-64 result (INT_MIN) for NaN and out-of-range
// values, as the reference C computes it on its host
__device__ __forceinline__ int hx_to_int(float f)
{
	return (f >= -2147483648.0f && f < 2147483648.0f) ? (int) f : (int) 0x80000000u;
}

// fdk_w3d_pow_tab (fdk,w3d,pow,tab.c:33-56)
__device__ __forceinline__ float hx_w3d(const HelixArgs &a, float t_bin, float t_bin_c,
	int iz)
{
	const GeomF &g = a.g;
	const int id = (a.pitch != 0.f) ? 0 : hx_to_int(fabsf((float) iz - g.wz));
	const float half = (float) a.nt2;
	const float d = fabsf(t_bin - g.wt), dc = fabsf(t_bin_c - g.wt);
	int i0 = hx_to_int(floorf(d < half ? d : half) + half * (float) id);
	int i1 = hx_to_int(floorf(dc < half ? dc : half) + half * (float) id);
	i0 = min(max(i0, 0), a.tab_len - 1);
	i1 = min(max(i1, 0), a.tab_len - 1);
	const float s = a.tab[i0], sc = a.tab[i1];
	const float eps = 0.001f;
	float r = sc / (sc + s);
	r = (r > eps) ? r : eps;
	r = (r < 1.f - eps) ? r : 1.f - eps;
	return 2.f * r;
}

__device__ __forceinline__ float hx_sample(const float *view, int ns, int nt, int s, int t)
{
	return (s >= 0 && s < ns && t >= 0 && t < nt) ? view[(long) s * nt + t] : 0.f;
}

__device__ __forceinline__ int hx_wave_min(int x)
{
	for (int o = 32; o > 0; o >>= 1)
		x = min(x, __shfl_xor(x, o));
	return x;
}

__device__ __forceinline__ int hx_wave_max(int x)
{
	for (int o = 32; o > 0; o >>= 1)
		x = max(x, __shfl_xor(x, o));
	return x;
}

// LDS per wave: HX_ZC accumulators per lane and one view footprint
#ifndef HX_FP_FLOATS
#define HX_FP_FLOATS 768
#endif
constexpr int HX_FP = HX_FP_FLOATS;   // floats of the staged footprint

#ifndef HX_MIN_WAVES
#define HX_MIN_WAVES 4
#endif

__global__ void __launch_bounds__(64, HX_MIN_WAVES) helix_kernel(const HelixArgs a)
{
	const GeomF &g = a.g;
	const int lane = threadIdx.x;
	const int tx = (g.nx + 7) >> 3, ty = (g.ny + 7) >> 3;
	const int b = hx_xcd_block(blockIdx.x, tx * ty * a.nzc);
	const int bxy = b % (tx * ty), bz = b / (tx * ty);
	const int ix = (bxy % tx) * 8 + (lane & 7);
	const int iy = (bxy / tx) * 8 + (lane >> 3);
	// lanes outside the image stay (inactive) for the wave-wide steps
	const bool inimg = ix < g.nx && iy < g.ny;
	const long col = ix + (long) iy * g.nx;
	const int z0 = bz * HX_ZC;
	const int nk = min(HX_ZC, g.nz - z0);
	// the lane's accumulators, slice-major so that lanes at different slices
	// hit different banks; each lane touches only its own entries
	__shared__ float sacc[HX_ZC * 64];
	__shared__ float sbuf[HX_FP];
	float *acc = sacc + lane;
	for (int k = 0; k < HX_ZC; ++k)
		acc[k * 64] = 0.f;

	const bool live = inimg && (!a.mask || a.mask[col]);
	const float yy = g.dy * ((float) iy - g.wy);
	const float xx = g.dx * ((float) ix - g.wx);
	const int2 vr = a.chunk_views[bz];
	const bool weigh = a.cone_par && a.w3d;
	for (int v = vr.x; v < vr.y; ++v) {
		const HelixView hv = a.views[v];
		const float *view = a.proj + (long) v * g.ns * g.nt;
		// fdk,ts.c:120-131
		const float xbeta = xx * hv.cosb + yy * hv.sinb;
		const float ybetas = g.dso - (-xx * hv.sinb + yy * hv.cosb);
		float mag;
		if (a.cone_par)
			mag = (float) ((double) g.dsd / (sqrt((double) (g.dso * g.dso - xbeta * xbeta))
				- (double) (g.dso - ybetas)));
		else
			mag = (g.kind == KIND_PAR) ? 1.f : g.dsd / ybetas;
		float ss;
		if (a.cone_par)
			ss = xbeta;
		else if (g.kind == KIND_ARC)
			ss = (float) ((double) g.dsd * atan2((double) xbeta, (double) ybetas));
		else
			ss = mag * xbeta;
		const float ss_bin = ss / g.ds + g.ws;
		const int is = hx_to_int(floorf(ss_bin));
		const float wr = ss_bin - (float) is;
		const float wl = 1.f - wr;
		const float t_inc = mag * g.dz / g.dt;
		// out of the detector in s, behind the source (no samples), masked
		bool act = live && is >= -1 && is < g.ns && t_inc > 0.f && t_inc != INFINITY;
		int lo = 0, hi = 0;
		float t_bin = 0.f;
		float wzbeta = 0.f;
		if (act) {
			// fdk,ts.c:150-158
			const float wzb0 = g.wz + hv.zs / g.dz;
			wzbeta = a.cone_par
				? (float) ((double) wzb0 + asin((double) (-xbeta / g.dso)) * a.dz_per_rad
					/ (double) g.dz)
				: wzb0;
			int iz_min = hx_to_int(floorf(wzbeta - (g.wt + 1.f) / t_inc) + 1.f);
			int iz_max = hx_to_int(ceilf(wzbeta + ((float) g.nt - g.wt) / t_inc) - 1.f + 1.f);
			iz_min = max(max(iz_min, 0), hv.izb0);
			iz_max = min(min(iz_max, g.nz), hv.izb1);
			t_bin = t_inc * ((float) iz_min - wzbeta) + g.wt;
			if (t_bin == -1.f) {
				++iz_min;
				t_bin += t_inc;
			}
			lo = max(iz_min, z0);
			hi = min(iz_max, z0 + nk);
			act = lo < hi;
			for (int iz = iz_min; iz < lo && act; ++iz)   // the running sum up to this chunk
				t_bin += t_inc;
			// from here t_bin is the value at slice lo and advances once per
			// slice of [lo, hi), in slice order
		}
		// the wave's footprint on the view: columns [s0, s1], rows [r0, r1]
		// (one row of margin each way over the rows the running sum reaches)
		const int s0 = hx_wave_min(act ? is : 0x7fffffff);
		if (s0 == 0x7fffffff)
			continue;                        // no lane of the wave reads this view
		const int s1 = min(hx_wave_max(act ? is + 1 : -0x7fffffff), g.ns);
		const int r0 = max(hx_wave_min(act ? hx_to_int(floorf(t_bin)) - 1 : 0x7fffffff), -1);
		const int r1 = min(hx_wave_max(act ? hx_to_int(floorf(t_bin + (float) (hi - 1 - lo)
			* t_inc)) + 2 : -0x7fffffff), g.nt);
		const int nrow = r1 - r0 + 1, ncol = s1 - s0 + 1;
		const bool staged = r1 >= r0 && nrow * ncol <= HX_FP;
		if (staged) {
			// coalesced along t; samples outside the view are the zero padding
			const float inv = 1.f / (float) nrow;
			for (int idx = lane; idx < nrow * ncol; idx += 64) {
				const int c = (int) (((float) idx + 0.5f) * inv);
				const int r = idx - c * nrow;
				const int s = s0 + c, t = r0 + r;
				sbuf[idx] = (s >= 0 && s < g.ns && t >= 0 && t < g.nt)
					? view[(long) s * g.nt + t] : 0.f;
			}
			__syncthreads();
		}
		if (act) {
			// conjugate ray (fdk,ts.c:169-175) and its heights, for w3d
			float t_inc_c = 0.f, wzc_hi = 0.f, wzc_lo = 0.f, zsc_hi = 0.f, zsc_lo = 0.f;
			if (weigh) {
				const float xbeta_c = -xbeta;
				const float ybetas_c = 2.f * g.dso - ybetas;
				const float mag_c = (float) ((double) g.dsd
					/ (sqrt((double) (g.dso * g.dso - xbeta_c * xbeta_c))
						- (double) (g.dso - ybetas_c)));
				t_inc_c = mag_c * g.dz / g.dt;
				const double ac = asin((double) (-xbeta_c / g.dso)) * a.dz_per_rad
					/ (double) g.dz;
				if (a.pitch != 0.f) {
					zsc_hi = (float) ((double) hv.zs + M_PI * a.dz_per_rad);
					zsc_lo = (float) ((double) hv.zs - M_PI * a.dz_per_rad);
				}
				wzc_hi = (float) ((double) (g.wz + zsc_hi / g.dz) + ac);
				wzc_lo = (float) ((double) (g.wz + zsc_lo / g.dz) + ac);
			}
			const float w2f = (g.kind == KIND_ARC)
				? (g.dsd * g.dsd) / (ybetas * ybetas + xbeta * xbeta)
				: mag * mag;
			const int cb = (is - s0) * nrow - r0;   // staged index of (is, row 0)
			for (int iz = lo; iz < hi; ++iz) {
				// fdk,ts.c:177-236; t_bin lies within a few rows of [-1, nt)
				// here, so the plain conversion equals the reference's
				const float fl = floorf(t_bin);
				const int it = (int) fl;
				const float wu = t_bin - fl;
				const float wd = 1.f - wu;
				float a00 = 0.f, a10 = 0.f, a01 = 0.f, a11 = 0.f;
				// [r0, r1] lies within [-1, nt]: a staged pair of rows is in range
				if (staged && it >= r0 && it < r1) {
					a00 = sbuf[cb + it];
					a01 = sbuf[cb + it + 1];
					a10 = sbuf[cb + nrow + it];
					a11 = sbuf[cb + nrow + it + 1];
				} else if (it >= -1 && it < g.nt) {
					a00 = hx_sample(view, g.ns, g.nt, is, it);
					a10 = hx_sample(view, g.ns, g.nt, is + 1, it);
					a01 = hx_sample(view, g.ns, g.nt, is, it + 1);
					a11 = hx_sample(view, g.ns, g.nt, is + 1, it + 1);
				}
				// rows outside [-1, nt-1] give p0 = wd*0 + wu*0 = 0
				const float p1 = wl * a00 + wr * a10;
				const float p2 = wl * a01 + wr * a11;
				const float p0 = wd * p1 + wu * p2;
				float w2 = w2f;
				if (a.cone_par) {
					float ww = 1.f;
					if (weigh) {
						const bool up = t_bin >= g.wt;
						const float zs_c = up ? zsc_hi : zsc_lo;
						const float wzbeta_c = up ? wzc_hi : wzc_lo;
						const float t_bin_c = t_inc_c * ((float) iz - wzbeta_c) + g.wt;
						if (zs_c >= a.zs_min && zs_c <= a.zs_max)
							ww = hx_w3d(a, t_bin, t_bin_c, iz);
					}
					w2 = ww;
				}
				acc[(iz - z0) * 64] += w2 * p0;
				t_bin += t_inc;
			}
		}
		if (staged)
			__syncthreads();                 // before the next view's footprint
	}
	if (inimg) {
		float *out = a.image + col * g.nz + z0;
		for (int k = 0; k < nk; ++k)
			out[k] = acc[k * 64];
	}
}

} // namespace

hipError_t launch_helix(const HelixArgs &a, hipStream_t s)
{
	const long tiles = (long) ((a.g.nx + 7) / 8) * ((a.g.ny + 7) / 8) * a.nzc;
	if (tiles <= 0)
		return hipSuccess;
	if (tiles > 0x7fffffffL)
		return hipErrorInvalidValue;
	hipLaunchKernelGGL(helix_kernel, dim3((unsigned) tiles), dim3(64), 0, s, a);
	return hipGetLastError();
}

} // namespace fdk
