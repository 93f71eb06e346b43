// fdk_kernels.hip -- gfx950 kernels of the FDK cone-beam backprojector.
//
// The arithmetic follows fdk_ts_back1 of the reference
// (mex/src/fdk-helix/mex/fdk,ts.c:55-251, axial branch) and fdk_st_back1
// (mex/src/fdk/fdk-orig/fdk,st.c:38-141).  This file is compiled with
// -ffp-contract=off: every multiply-add below is a separate multiply and add,
// as in the reference C built without fast-math, unless it is written as an
// explicit fmaf() (only in the FAST paths).
//
// Work decomposition of the main kernel (back_kernel):
//   * a workgroup owns an 8x8 tile of (x,y) columns and a z-group of NW*ZL
//     slices; consumer wave w owns slices [zg0 + w*ZL, zg0 + (w+1)*ZL) of
//     all 64 columns, lane = column, and keeps those ZL voxels in FP32
//     registers for the whole view loop; NP producer waves work ahead of
//     them, one view per producer wave;
//   * views go in stages of `vstage` views through two LDS buffers:
//       setup   -- one lane per (view, column) computes the per-column
//                  quantities (s bin, weights, slice range, t_bin), packs
//                  them in a 32-byte record, and a DPP wave reduction gives
//                  the view's detector footprint over the tile;
//       stage   -- the footprint of the zero-padded view (the padded copy
//                  plays fdk_ts_put_view, mex/src/fdk/fdk-orig/fdk,ts,t.c:13-21)
//                  is copied from HBM into LDS with asynchronous
//                  global-to-LDS loads (FP32 samples, or FP16/BF16 column
//                  pairs);
//       update  -- every consumer lane runs its ZL slices: bilinear
//                  interpolation from LDS, FP32 accumulation in registers;
//   * one write per voxel at the end, into the [nz nx ny] image.
// Arithmetic modes: EXACT (FP32) reproduces the reference's t_bin sequence
// (t_bin += t_inc per slice) and operation order, which makes it bit-for-bit
// equal to the reference C; FAST (FP32) folds w2 into the weights and uses
// FMAs; DIRECT (FP16/BF16 storage) evaluates t_bin in closed form and is held
// to the FP64 result.
//
// No MFMA: the operation is a gather plus a few multiply-adds per voxel-view.

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <type_traits>

#include "fdk_kernels.h"

namespace fdk {

// ---------------------------------------------------------------------------
// helpers

// (int) conversion with the x86-64 cvttss2si result for NaN and out-of-range
// values (INT_MIN), which is what the reference C does on its host.
__device__ __forceinline__ int to_int_x86(float f)
{
	return (f >= -2147483648.0f && f < 2147483648.0f) ? (int) f : (int) 0x80000000u;
}

// FP32 value of an FP16 sample (exact); in an FMA with an FP32 operand the
// compiler folds the conversion into one v_fma_mix_f32
__device__ __forceinline__ float f16_to_f32(uint32_t bits16)
{
	return (float) __builtin_bit_cast(_Float16, (unsigned short) bits16);
}

// FP32 sample idx of the zero-padded copy of FP32 views
__device__ __forceinline__ float load_sample(const void *padded, long idx)
{
	return static_cast<const float *>(padded)[idx];
}

// Per-column, per-view quantities of fdk,ts.c:120-177 (axial branch,
// cone_par = 0, source_zs = 0), same float operation order.
struct ColSetup {
	int is;          // left s bin, in [-1, ns-1]
	float wl, wr;    // horizontal weights
	float w2;        // fan-beam weight
	float t_inc;     // t' step per slice
	float t_bin;     // t' at iz_min
	int iz_min, iz_max;
};

// Returns SETUP_OUT when the column misses the view's detector (fdk,ts.c:134),
// SETUP_IDLE when it meets it but reads no sample, SETUP_ACTIVE otherwise.
enum { SETUP_OUT = 0, SETUP_ACTIVE = 1, SETUP_IDLE = 2 };

// A view's geometry: the trig of its angle and, on a helix, its source
// height as wzbeta = wz + source_zs / dz (fdk,ts.c:149-151, cone_par = 0) and
// its slice window [izb0, izb1) (fdk,ts,t.c:78-85); an axial view has wz and
// the whole volume
struct ViewGeo {
	float cosb, sinb, wzb;
	int izb0, izb1;
};

__device__ __forceinline__ ViewGeo view_geo(const BackArgs &a, int v)
{
	const float2 cs = a.trig[v];
	ViewGeo vg{cs.x, cs.y, a.g.wz, 0, a.g.nz};
	if (a.vzs) {
		vg.wzb = a.g.wz + a.vzs[v] / a.g.dz;
		const int2 w = a.izb[v];
		vg.izb0 = w.x;
		vg.izb1 = w.y;
	}
	return vg;
}

__device__ __forceinline__ int column_setup(const GeomF &g, float xx, float yy,
	const ViewGeo &vg, ColSetup &c)
{
	const float cosb = vg.cosb, sinb = vg.sinb, wzb = vg.wzb;
	const float xbeta = xx * cosb + yy * sinb;
	const float ybetas = g.dso - (-xx * sinb + yy * cosb);
	const float mag = (g.kind == KIND_PAR) ? 1.f : g.dsd / ybetas;
	float ss;
	if (g.kind == KIND_ARC)
		ss = (float) ((double) g.dsd * atan2((double) xbeta, (double) ybetas));
	else
		ss = mag * xbeta;
	const float ss_bin = ss / g.ds + g.ws;
	const int is = to_int_x86(floorf(ss_bin));
	if (is < -1 || is >= g.ns)
		return SETUP_OUT;
	c.is = is;
	c.wr = ss_bin - (float) is;
	c.wl = 1.f - c.wr;
	c.t_inc = mag * g.dz / g.dt;
	// t_inc <= 0 means a voxel behind the source (or dz = 0, where the
	// reference's slice range is empty): no samples are read for it.  The
	// footprint bounds below rely on t_bin increasing along z.
	if (!(c.t_inc > 0.f) || c.t_inc == INFINITY)
		return SETUP_IDLE;
	int iz_min = to_int_x86(floorf(wzb - (g.wt + 1.f) / c.t_inc) + 1.f);
	int iz_max = to_int_x86(ceilf(wzb + ((float) g.nt - g.wt) / c.t_inc) - 1.f + 1.f);
	iz_min = max(max(iz_min, 0), vg.izb0);
	iz_max = min(min(iz_max, g.nz), vg.izb1);
	float t_bin = c.t_inc * ((float) iz_min - wzb) + g.wt;
	if (t_bin == -1.f) {
		++iz_min;
		t_bin += c.t_inc;
	}
	c.t_bin = t_bin;
	c.iz_min = iz_min;
	c.iz_max = iz_max;
	c.w2 = (g.kind == KIND_ARC)
		? (g.dsd * g.dsd) / (ybetas * ybetas + xbeta * xbeta)
		: mag * mag;
	return SETUP_ACTIVE;
}

// Per-column, per-view quantities for the FP16/BF16 paths, which are held to
// the FP64 result rather than to the reference's bits: reciprocals instead
// of divisions, float atan2 for the arc detector, FMA.  iz_min/iz_max are
// only a first guess here; setup_direct settles the slice range exactly.
__device__ __forceinline__ int column_setup_fast(const GeomF &g, float xx, float yy,
	const ViewGeo &vg, ColSetup &c)
{
	const float cosb = vg.cosb, sinb = vg.sinb, wzb = vg.wzb;
	const float xbeta = fmaf(xx, cosb, yy * sinb);
	const float ybetas = g.dso - fmaf(-xx, sinb, yy * cosb);
	const float rb = __builtin_amdgcn_rcpf(ybetas);
	const float mag = (g.kind == KIND_PAR) ? 1.f : g.dsd * rb;
	const float ss = (g.kind == KIND_ARC) ? g.dsd * atan2f(xbeta, ybetas) : mag * xbeta;
	const float ss_bin = fmaf(ss, g.inv_ds, g.ws);
	if (!(ss_bin >= -1.f && ss_bin < (float) g.ns))   // also rejects NaN
		return SETUP_OUT;
	const float fl = floorf(ss_bin);
	c.is = (int) fl;
	c.wr = ss_bin - fl;
	c.wl = 1.f - c.wr;
	c.t_inc = mag * g.dz_dt;
	if (!(c.t_inc > 0.f) || c.t_inc == INFINITY)
		return SETUP_IDLE;
	const float rti = __builtin_amdgcn_rcpf(c.t_inc);
	const float zmin = floorf(fmaf(-(g.wt + 1.f), rti, wzb)) + 1.f;
	const float zmax = ceilf(fmaf((float) g.nt - g.wt, rti, wzb));
	c.iz_min = max((int) fminf(fmaxf(zmin, -1.f), (float) g.nz + 1.f), vg.izb0);
	c.iz_max = min((int) fminf(fmaxf(zmax, -1.f), (float) g.nz + 1.f), vg.izb1);
	c.t_bin = fmaf(c.t_inc, (float) c.iz_min - wzb, g.wt);
	if (g.kind == KIND_ARC)
		c.w2 = (g.dsd * g.dsd) * __builtin_amdgcn_rcpf(fmaf(ybetas, ybetas, xbeta * xbeta));
	else
		c.w2 = mag * mag;
	return SETUP_ACTIVE;
}

// Walk of the reference's t_bin sequence over one z-group [zg0, zg1):
// t_bin starts at iz_min and grows by repeated addition of t_inc
// (fdk,ts.c:169,244), so the value at slice z is reproduced only by the same
// sequence of additions.  Records, for each of the NW chunks of ZL slices,
// t_bin at the chunk's first active slice, and returns the effective range
// [zlo, zhi) with t_bin at zlo (exact) and an upper bound of t_bin at zhi-1.
// Leading slices whose t_bin is <= -1 are dropped: every sample they would
// read is a zero pad, so the reference adds exactly 0 there; trailing ones
// at t_bin >= nt are kept and read zero rows (see the bound below).
template <int NW, int ZL>
__device__ __forceinline__ bool setup_walk(const GeomF &g, const ColSetup &c,
	int zg0, int zg1, float (&Tk)[NW], int &zlo, int &zhi,
	float &t_first, float &t_last)
{
	const int lo = max(c.iz_min, zg0), hi = min(c.iz_max, zg1);
	const float ti = c.t_inc;
	float t = c.t_bin;
	// z-groups above the first: walk up to zg0, eight additions a step --
	// plain while every lane has eight left, the last steps masked per lane
	int n = lo - c.iz_min;
	while (__all(n >= 8)) {
#pragma unroll
		for (int j = 0; j < 8; ++j)
			t += ti;
		n -= 8;
	}
	while (__any(n > 0)) {
#pragma unroll
		for (int j = 0; j < 8; ++j)
			t = (j < n) ? t + ti : t;
		n -= 8;
	}
	int z = lo;
	while (z < hi && t <= -1.f) { // rare: leading zero-contribution slices
		t += ti;
		++z;
	}
	zlo = z;
	zhi = hi;
	t_first = t;
#pragma unroll
	for (int k = 0; k < NW; ++k)
		Tk[k] = 0.f;
	if (zlo >= zhi) {
		t_last = t;
		return false;
	}
	// Tk[k] = t_bin at s_k = max(zlo, zg0 + k ZL), the lane's first slice in
	// chunk k: s_k - s_(k-1) more additions (0..ZL) from chunk k-1's value.
	// Uniform control flow: ZL plain additions where every lane needs them,
	// masked ones where some do (adding nothing leaves the value exact);
	// lanes past their range add harmlessly.
	float tt = t;
	int sp = zlo;
	Tk[0] = t;
#pragma unroll
	for (int k = 1; k < NW; ++k) {
		const int ck = zg0 + k * ZL;
		if (__all(zhi <= ck))
			break;                                 // no active lane reaches chunk k
		const int sk = max(zlo, ck);
		const int nk = sk - sp;
		if (__all(nk == ZL)) {
#pragma unroll
			for (int j = 0; j < ZL; ++j)
				tt += ti;
		} else if (__any(nk > 0)) {
#pragma unroll
			for (int j = 0; j < ZL; ++j)
				tt = (j < nk) ? tt + ti : tt;
		}
		Tk[k] = tt;
		sp = sk;
	}
	// upper bound of t_bin at zhi - 1: m additions of ti each round to within
	// half an ulp of the running value (|t| < 2^24 here).  The reference's
	// slice range keeps t < nt up to that rounding, so t_last < nt + 1: a
	// trailing slice at t >= nt reads padded rows nt + 1 and nt + 2, both
	// zero (row nt + 2 of a padded column is row 0 of the next, and the
	// padded copy ends in zeros), and adds exactly 0, as in the reference.
	const float m = (float) (zhi - 1 - zlo);
	const float est = fmaf(m, ti, t_first);
	const float mag = fmaxf(fmaxf(fabsf(t_first), fabsf(est)), 1.f);
	t_last = est + (m + 1.f) * mag * 1.2e-7f + 1e-6f;
	return zlo < zhi;
}

// Slice range of one z-group for the FP16/BF16 paths, in u = t_bin + 1 (the
// padded detector row coordinate, > 0 for every slice that reads a sample).
// u is evaluated in closed form, as the consumer evaluates it:
//   U0   = u(iz_min) + (zg0 - iz_min) * t_inc          (the group's first slice)
//   U[k] = U0 + k * ZL * t_inc                          (chunk k's first slice)
//   u    = U[k] + j * t_inc                             (slice zg0 + k*ZL + j)
// each step one fmaf (one rounding each; nearer the exact value than the
// reference's running sum).  [zlo, zhi) is [iz_min - 1, iz_max + 1) within
// the group and the view's window: column_setup_fast's bounds are within one
// slice of where u leaves (0, nt + 1), so the range holds every slice with
// 0 < u < nt + 1 in the window, and u <= 0 below it and u >= nt + 1 above
// it unless the window ends it (the consumer skips the slices outside the
// window); the few slices it holds outside (0, nt + 1) read zero rows with
// weight 0 under the consumer's clamp (ColRange).  first/last are u at zlo
// and zhi - 1.  Uk[0] receives U0.
template <int NW, int ZL>
__device__ __forceinline__ bool setup_direct(const GeomF &g, const ColSetup &c, const ViewGeo &vg,
	int zg0, int zg1, float (&Uk)[NW], int &zlo, int &zhi, float &first, float &last)
{
	const float ti = c.t_inc;
	const float U0 = fmaf((float) (zg0 - c.iz_min), ti, c.t_bin + 1.f);
	Uk[0] = U0;
	auto uz = [&](int z) {           // z in [zg0, zg1)
		const unsigned k = (unsigned) (z - zg0), j0 = k & ~(unsigned) (ZL - 1);
		return fmaf((float) (k - j0), ti, fmaf((float) j0, ti, U0));
	};
	// the extra slice at each end stays inside the view's window (a helix)
	const int lo = min(max(max(c.iz_min - 1, vg.izb0), zg0), zg1);
	const int hi = max(min(min(c.iz_max + 1, vg.izb1), zg1), lo);
	zlo = lo;
	zhi = hi;
	first = lo < hi ? uz(lo) : 0.f;
	last = lo < hi ? uz(hi - 1) : 0.f;
	return lo < hi;
}

// One column's view over one z-group, in either mode: the per-column
// quantities, the slice range [zlo, zhi), the vertical coordinate at each
// chunk's start (t_bin for the reference walk, u = t_bin + 1 for DIRECT) and
// the padded detector rows [row_lo, row_hi] its bilinear reads touch.
// DIRECT consumers compute every slice of a chunk they work on, with u
// clamped to [0, u_hi]: a slice below the range (u <= 0) then reads padded
// rows 0 and 1 with vertical weight 0, i.e. exactly the zero pad row; one
// above it (u >= nt + 1) reads rows nt + 1 and nt + 2 with weight 0, two
// zero rows (row nt + 2 is one more zero row of the staged footprint).  So
// the rows read include 0 when the range starts inside the group and
// nt + 2 when it ends inside it.  Where the range reaches the group's end,
// u_hi = u at zhi - 1 (within [0, nt + 1]) keeps the (unwritten) slices past
// the volume's last slice inside the footprint.
struct ColRange {
	ColSetup c;
	int zlo, zhi;
	float first, last;     // t (walk) or u (DIRECT) at zlo and zhi - 1
	int row_lo, row_hi;    // padded rows read
	float u_hi;            // DIRECT: upper clamp of u
};

template <int NW, int ZL, bool DIRECT>
__device__ __forceinline__ bool col_range(const GeomF &g, float xx, float yy, const ViewGeo &vg,
	int zg0, int zg1, ColRange &r, float (&Tk)[NW])
{
	r.zlo = r.zhi = 0;
	r.first = r.last = 0.f;
	r.u_hi = 0.f;
	if constexpr (DIRECT) {
		if (column_setup_fast(g, xx, yy, vg, r.c) != SETUP_ACTIVE)
			return false;
		if (!setup_direct<NW, ZL>(g, r.c, vg, zg0, zg1, Tk, r.zlo, r.zhi, r.first, r.last))
			return false;
		// every slice the consumer computes reads rows floor(uc), floor(uc) + 1
		// of uc = u clamped to [0, u_hi] (truncation is floor for uc >= 0)
		const float ntp1 = (float) g.nt + 1.f;
		r.u_hi = (r.zhi < zg1) ? ntp1 : fmaxf(fminf(r.last, ntp1), 0.f);
		r.row_lo = (r.zlo > zg0) ? 0 : (int) fminf(fmaxf(r.first, 0.f), r.u_hi);
		r.row_hi = (int) r.u_hi + 1;
	} else {
		if (column_setup(g, xx, yy, vg, r.c) != SETUP_ACTIVE)
			return false;
		if (!setup_walk<NW, ZL>(g, r.c, zg0, zg1, Tk, r.zlo, r.zhi, r.first, r.last))
			return false;
		r.row_lo = (int) floorf(r.first) + 1;
		r.row_hi = (int) floorf(r.last) + 2;
	}
	return true;
}

// Wave-wide reductions (wave64, every lane active) with DPP: quad swaps,
// row half-mirror and mirror give each row of 16 its result, row_bcast:15
// and :31 carry it across rows; lane 63 holds the wave's result.  The
// operation's identity as the DPP "old" value (what lanes outside the row
// mask receive) lets the compiler fold each step into one v_min/v_max_dpp.
template <bool MAX>
__device__ __forceinline__ int wave_reduce(int v)
{
	auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
	constexpr int id = MAX ? (int) 0x80000000 : 0x7fffffff;
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0xb1, 0xf, 0xf, false));   // quad_perm 1,0,3,2
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x4e, 0xf, 0xf, false));   // quad_perm 2,3,0,1
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xf, 0xf, false));  // row_half_mirror
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xf, 0xf, false));  // row_mirror
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
	v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
	return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min(int v)
{
	return wave_reduce<false>(v);
}
__device__ __forceinline__ int wave_max(int v)
{
	return wave_reduce<true>(v);
}

// Logical block index with consecutive indices on the same XCD: the hardware
// hands workgroup i to XCD i % 8, so neighbouring tiles (which share detector
// footprints) land in the same L2.
__device__ __forceinline__ int xcd_block(int bid, int nblocks)
{
	const int xcd = bid & 7, local = bid >> 3;
	const int per = nblocks >> 3, rem = nblocks & 7;
	return xcd * per + min(xcd, rem) + local;
}

// ---------------------------------------------------------------------------
// main kernel

// dynamic LDS of the main kernel; every access below indexes this array
// with a word offset so that the compiler emits LDS (ds_*) instructions
extern __shared__ float fdk_lds[];

// LDS layout of one stage buffer (4-byte words); the kernel holds two, and
// after them two zero words that DIRECT lanes without work in a view read.
//   slots [vstage][slot]        staged footprints: FP32 samples (walk modes),
//                               16-bit column pairs (DIRECT)
//   P     [vstage][64][8]       per (view, column) record, two 16-byte halves
//                               read with one ds_read_b128 each:
//     walk:   P0 = {cb, wl, wr, w2}      P1 = {t_inc, zlo | zhi << 16, is, t_last}
//     DIRECT: P0 = {cb, wl, wr, U0}      P1 = {t_inc, zlo | zhi << 16, u_hi, 0}
//             (DIRECT wl, wr: FP32 horizontal weights times w2 times the
//             view scale)
//           cb: LDS word, relative to the stage buffer, of the lane's left
//               column at row 1 (walk) / padded row 0 (DIRECT); for views
//               gathered from HBM: is
//   T     [vstage][NW][64]      walk only: t_bin at each consumer wave's first
//                               active slice (DIRECT evaluates it from U0)
//   vinfo [vstage][4]           mode | zlo << 2 | zhi << 12: the union of the
//                               tile's slice ranges relative to the z-group
//                               (at most 256 slices), and ilo | ihi << 10:
//                               their intersection (walk modes with several
//                               chunks per wave; with one: s_lo, t_lo),
//                               tw, (mode << 16) | staged columns
template <int NW, bool TTAB>
struct Lds {
	int P, T, vinfo, total, zero;
	__device__ __forceinline__ Lds(int vstage, int slot)
	{
		P = vstage * slot;
		T = P + vstage * 64 * 8;
		vinfo = T + (TTAB ? vstage * NW * 64 : 0);
		total = vinfo + vstage * 4;
		zero = 2 * total;
	}
};

enum { MODE_LDS = 0, MODE_GLOBAL = 1, MODE_NONE = 2 };

// Detector footprint of one view over the tile and z-group in padded
// coordinates, the union over the wave's active lanes: columns
// [s_lo, s_hi], rows [t_lo, t_hi] (DIRECT: up to row nt + 2, a zero row).
// Staged words: (s_hi - s_lo + 1) x tw FP32 samples (walk) or
// (s_hi - s_lo) x tw column pairs (DIRECT), tw = t_hi - t_lo + 1.
struct Footprint {
	int s_lo, s_hi, t_lo, t_hi;
};

template <bool DIRECT>
__device__ __forceinline__ Footprint view_footprint(const GeomF &g, bool act, const ColRange &cr)
{
	Footprint f;
	f.s_lo = max(wave_min(act ? cr.c.is + 1 : 0x7fffffff), 0);
	f.s_hi = min(wave_max(act ? cr.c.is + 2 : (int) 0x80000000), g.ns + 1);
	f.t_lo = max(wave_min(act ? cr.row_lo : 0x7fffffff), 0);
	f.t_hi = min(wave_max(act ? cr.row_hi : (int) 0x80000000), g.nt + 2);
	return f;
}

// Asynchronous copy of one FP32 sample per lane from HBM into the LDS word
// `ldst + lane` (global_load_lds: the LDS address is wave-uniform plus
// 4 bytes per lane).
__device__ __forceinline__ void stage_word(const char *src, float *ldst)
{
	const auto g = (const void __attribute__((address_space(1))) *) src;
	const auto l = (void __attribute__((address_space(3))) *) ldst;
	__builtin_amdgcn_global_load_lds(g, l, 4, 0, 0);
}

// DIRECT staging of one view's footprint: slot word (k, r), k in [0, sw),
// r in [0, tw), holds the padded samples (s_lo + k, t_lo + r) in its low and
// (s_lo + k + 1, t_lo + r) in its high half, read straight from the caller's
// [ns][nt] 16-bit view and zero outside [1, ns] x [1, nt] (the zero border
// that fdk_ts_put_view gives each view, fdk,ts,t.c:13-21).  Lane = detector
// row: the wave walks the columns keeping the previous column's sample, with
// sixteen 2-byte loads in flight (64 consecutive samples of a column per
// load, one 128-byte line).  Each sample is read from HBM once per footprint.
// Columns are tw words apart (the caller makes tw odd, so that lanes reading
// the same row of neighbouring columns fall in different LDS banks).
__device__ __forceinline__ void stage_pairs(const uint16_t *view, const GeomF &g,
	const Footprint &f, int sw, int tw, int slot, int lane)
{
	// zero rows of the footprint: padded rows 0, nt + 1 and nt + 2
	for (int q = 0; q < 3; ++q) {
		const int zr = (q == 0) ? 0 : g.nt + q;
		if (zr < f.t_lo || zr > f.t_hi)
			continue;
		for (int k = lane; k < sw; k += 64)
			fdk_lds[slot + k * tw + (zr - f.t_lo)] = 0.f;
	}
	const int ra = max(f.t_lo, 1), rb = min(f.t_hi, g.nt);   // detector rows
	for (int r0 = ra; r0 <= rb; r0 += 64) {
		const int r = r0 + lane;
		const bool rv = r <= rb;
		const int srow = (rv ? r : rb) - 1;    // sample row (a valid one past the end)
		const int lw = slot + (r - f.t_lo);
		uint32_t prev = 0u;                    // padded column s_lo (0: the zero pad)
		if (f.s_lo >= 1)
			prev = view[(long) (f.s_lo - 1) * g.nt + srow];
		for (int k0 = 0; k0 < sw; k0 += 16) {
			uint32_t cur[16];
#pragma unroll
			for (int q = 0; q < 16; ++q) {
				// padded column s_lo + k0 + q + 1 is sample column s_lo + k0 + q;
				// clamped so that the loads past the footprint stay in the view
				const int sc = min(f.s_lo + k0 + q, g.ns - 1);
				cur[q] = view[(long) sc * g.nt + srow];
			}
#pragma unroll
			for (int q = 0; q < 16; ++q) {
				const int k = k0 + q;
				const uint32_t hi = (f.s_lo + k + 1 <= g.ns) ? cur[q] : 0u;
				if (k < sw && rv)
					fdk_lds[lw + k * tw] = __uint_as_float(prev | (hi << 16));
				prev = hi;
			}
		}
	}
}

// Producer side: per-(column, view) setup of `nv` views into the stage
// buffer at word offset `bo`, and staging of their footprints, by producer
// wave `pw` of NP.  Walk modes (FP32) stage from the zero-padded copy with
// asynchronous global-to-LDS loads, which the caller waits for (vmcnt)
// before the stage barrier; DIRECT (FP16/BF16) stages column pairs from the
// caller's views (stage_pairs).  R: chunks per consumer wave (the walk
// modes' consumers test slice groups against the tile's ranges with R > 1).
template <int NW, int NP, int ZL, int FMT, bool FAST, int R>
__device__ __forceinline__ void produce(const BackArgs &a, int bo,
	const Lds<NW, FMT == FMT_F32> &L,
	int v0, int nv, int pw, int lane, bool colvalid, float xx, float yy, int zg0, int zg1)
{
	const GeomF &g = a.g;
	constexpr bool DIRECT = FMT != FMT_F32;
	for (int vv = pw; vv < nv; vv += NP) {
		const int v = v0 + vv;
		// zero-initialised: a lane whose column is invalid never runs
		// col_range, and its record must still describe an empty range
		ColRange cr{};
		float Tk[NW] = {};
		const bool act = colvalid
			&& col_range<NW, ZL, DIRECT>(g, xx, yy, view_geo(a, v), zg0, zg1, cr, Tk);
		const ColSetup &c = cr.c;
		const Footprint f = view_footprint<DIRECT>(g, act, cr);
		int mode = MODE_NONE, tw = 0, sw = 0;
		if (f.s_lo <= f.s_hi && f.t_lo <= f.t_hi) {
			// DIRECT: an odd column stride (bank spread, stage_pairs)
			tw = DIRECT ? ((f.t_hi - f.t_lo + 1) | 1) : f.t_hi - f.t_lo + 1;
			sw = f.s_hi - f.s_lo + (DIRECT ? 0 : 1);     // staged columns (DIRECT: pairs)
			mode = ((long) sw * tw <= a.slot) ? MODE_LDS : MODE_GLOBAL;
		}
		const int is_st = act ? c.is : 0;
		const int zhi_st = act ? cr.zhi : cr.zlo;
		float4 *P4 = reinterpret_cast<float4 *>(fdk_lds + bo + L.P) + (vv * 64 + lane) * 2;
		if constexpr (DIRECT) {
			// LDS word of padded row 0 in the lane's column pair; lanes without
			// work in this view read the two zero words with zero weights
			const int cb = !act ? L.zero - bo
				: (mode == MODE_LDS ? vv * a.slot + (c.is + 1 - f.s_lo) * tw - f.t_lo : c.is);
			const float w2s = act ? c.w2 * a.scale[v] : 0.f;
			P4[0] = make_float4(__int_as_float(cb), act ? w2s * c.wl : 0.f,
				act ? w2s * c.wr : 0.f, act ? Tk[0] : 0.f);
			P4[1] = make_float4(act ? c.t_inc : 0.f,
				__int_as_float(act ? cr.zlo | (zhi_st << 16) : 0), act ? cr.u_hi : 0.f, 0.f);
		} else {
			// LDS word of row 1, t = 0 in the lane's left column
			const int cb = act ? vv * a.slot + (c.is + 1 - f.s_lo) * tw + 1 - f.t_lo : 0;
			float wl = act ? c.wl : 0.f, wr = act ? c.wr : 0.f;
			const float w2 = act ? c.w2 : 0.f;
			if (FAST) {
				wl = w2 * wl;
				wr = w2 * wr;
			}
			P4[0] = make_float4(__int_as_float(mode == MODE_LDS ? cb : is_st), wl, wr, w2);
			P4[1] = make_float4(act ? c.t_inc : 0.f,
				__int_as_float(act ? cr.zlo | (zhi_st << 16) : 0), __int_as_float(is_st),
				act ? cr.last : 0.f);
			// t_bin at each chunk's first active slice; chunks the lane's range
			// misses get t_first, so that their (discarded) reads stay inside
			// the staged footprint
#pragma unroll
			for (int k = 0; k < NW; ++k) {
				const bool meets = cr.zlo < zg0 + (k + 1) * ZL && zg0 + k * ZL < cr.zhi;
				fdk_lds[bo + L.T + (vv * NW + k) * 64 + lane] =
					act ? (meets ? Tk[k] : cr.first) : 0.f;
			}
		}
		// The union of the active lanes' slice ranges (DIRECT, and the walk
		// modes with several chunks per wave, which also take the
		// intersection of all 64 lanes' ranges), so that the consumers find
		// the groups they work on with scalar tests.  A lane is active only
		// with a non-empty range (an inactive one empties the intersection);
		// with none active the view's mode is MODE_NONE and no range is read.
		int zrange = 0, zboth = 0;
		if constexpr (DIRECT || R > 1) {
			const int zl = wave_min(act ? cr.zlo - zg0 : 0x7fffffff);
			const int zh = wave_max(act ? cr.zhi - zg0 : (int) 0x80000000);
			zrange = (mode == MODE_NONE) ? 0 : (zl << 2) | (zh << 12);
			if constexpr (!DIRECT) {
				const int il = min(wave_max(act ? cr.zlo - zg0 : 1023), 1023);
				const int ih = max(wave_min(act ? cr.zhi - zg0 : 0), 0);
				zboth = il | (ih << 10);
			}
		}
		if (lane == 0) {
			int *vi = reinterpret_cast<int *>(fdk_lds) + bo + L.vinfo + vv * 4;
			vi[0] = (DIRECT || R > 1) ? mode | zrange : f.s_lo;
			vi[1] = (DIRECT || R > 1) ? zboth : f.t_lo;
			vi[2] = tw;
			vi[3] = (mode << 16) | (sw & 0xffff);
		}
		if (mode != MODE_LDS)
			continue;
		const int slot = bo + vv * a.slot;
		if constexpr (DIRECT) {
			stage_pairs(static_cast<const uint16_t *>(a.proj) + (size_t) v * g.ns * g.nt, g, f,
				sw, tw, slot, lane);
			continue;
		}
		// FP32: LDS word e of the slot <- padded sample (s_lo + e / tw, t_lo + e % tw)
		const int area = sw * tw;
		const int pnt = g.nt + 2;
		constexpr int esz = 4;
		const char *pv = static_cast<const char *>(a.padded)
			+ ((size_t) v * (g.ns + 2) + f.s_lo) * pnt * esz + (size_t) f.t_lo * esz;
		// byte offsets from the view's footprint corner stay below 2^32 (one
		// view), so the loads take a scalar base plus a 32-bit lane offset
		if (tw == pnt) {
			// full detector height: the footprint is contiguous in the padded view
			unsigned off = (unsigned) lane * esz;
			for (int e0 = 0; e0 < area; e0 += 64, off += 64 * esz)
				if (e0 + lane < area)
					stage_word(pv + off, fdk_lds + slot + e0);
		} else {
			// rows of tw samples: (r, tc) of element e, advanced by 64 per step;
			// off = (r * pnt + tc) * esz
			const unsigned magic = 0xffffffffu / (unsigned) tw + 1u;  // e/tw, exact (tw >= 2)
			int r = (int) __umulhi((unsigned) lane, magic);
			r -= (r * tw > lane) ? 1 : 0;
			int tc = lane - r * tw;
			const int dr = 64 / tw, dtc = 64 - dr * tw;
			unsigned off = (unsigned) (r * pnt + tc) * esz;
			const unsigned step = (unsigned) (dr * pnt + dtc) * esz;
			const unsigned wrap_step = (unsigned) (pnt - tw) * esz;
			for (int e0 = 0; e0 < area; e0 += 64) {
				if (e0 + lane < area)
					stage_word(pv + off, fdk_lds + slot + e0);
				tc += dtc;
				off += step;
				if (tc >= tw) {
					tc -= tw;
					off += wrap_step;
				}
			}
		}
	}
}

// One slice update of the FP32 paths: bilinear interpolation at (is, t) from
// the two staged columns c1 (s = is) and c2 (s = is+1), added to acc.  FAST
// folds w2 into wl/wr and uses fused multiply-adds; EXACT follows
// fdk,ts.c:199-241 operation by operation.
template <int FMT, bool FAST>
__device__ __forceinline__ float slice_update(float acc, float t, int c1, int c2,
	float wl, float wr, float w2)
{
	static_assert(FMT == FMT_F32, "FP16/BF16 views take slice_update_d");
	const float fl = floorf(t);
	const int it = (int) fl;
	const float wu = t - fl;
	const float wd = 1.f - wu;
	const float a0 = fdk_lds[c1 + it], a1 = fdk_lds[c1 + it + 1];
	const float b0 = fdk_lds[c2 + it], b1 = fdk_lds[c2 + it + 1];
	if (FAST) {
		const float q1 = fmaf(wr, b0, fmaf(wl, a0, 0.f));
		const float q2 = fmaf(wr, b1, fmaf(wl, a1, 0.f));
		acc = fmaf(wd, q1, acc);
		return fmaf(wu, q2, acc);
	} else {
		const float q1 = wl * a0 + wr * b0;
		const float q2 = wl * a1 + wr * b1;
		const float p0 = wd * q1 + wu * q2;
		return acc + w2 * p0;
	}
}

// FP32 value of a 16-bit sample (FP16 or BF16; exact)
template <int FMT>
__device__ __forceinline__ float sample16(uint32_t bits)
{
	if constexpr (FMT == FMT_F16)
		return f16_to_f32(bits);
	else
		return __uint_as_float(bits << 16);
}

// FP32 values of the two halves of a column-pair word (FP16 or BF16)
template <int FMT>
__device__ __forceinline__ float pair_lo(uint32_t w)
{
	if constexpr (FMT == FMT_F16)
		return f16_to_f32(w & 0xffffu);
	else
		return __uint_as_float(w << 16);
}
template <int FMT>
__device__ __forceinline__ float pair_hi(uint32_t w)
{
	if constexpr (FMT == FMT_F16)
		return f16_to_f32(w >> 16);
	else
		return __uint_as_float(w & 0xffff0000u);
}

// Bilinear interpolation of FP16/BF16 samples with FP32 weights, added to
// acc in FP32: a0, b0 on the lower row, a1, b1 on the upper one, wl, wr the
// horizontal weights (w2 and the view scale folded in), wu the vertical one.
// -q1 and q2 - q1 come from two chains of two FMAs (for FP16 each one
// v_fma_mix_f32, which widens the 16-bit sample inside the instruction),
// then acc + q1 + wu * (q2 - q1).
__device__ __forceinline__ float bilerp_d(float acc, float a0, float b0, float a1, float b1,
	float wl, float wr, float wu)
{
	const float nq1 = fmaf(-wr, b0, fmaf(-wl, a0, 0.f));
	const float d = fmaf(wr, b1, fmaf(wl, a1, nq1));
	return fmaf(wu, d, acc - nq1);
}

// DIRECT slice update at u = t_bin + 1 (> 0, or clamped: see ColRange): the
// padded row is the truncation of u, the vertical weight its fraction; c1
// addresses padded row 0 of the lane's staged column pair (is, is + 1), so
// one ds_read2_b32 fetches both rows' pairs.
template <int FMT>
__device__ __forceinline__ float slice_update_d(float acc, float u, int c1, float wl, float wr)
{
	const int iu = (int) u;
	const float wu = __builtin_amdgcn_fractf(u);
	const uint32_t w0 = __float_as_uint(fdk_lds[c1 + iu]);
	const uint32_t w1 = __float_as_uint(fdk_lds[c1 + iu + 1]);
	return bilerp_d(acc, pair_lo<FMT>(w0), pair_hi<FMT>(w0), pair_lo<FMT>(w1),
		pair_hi<FMT>(w1), wl, wr, wu);
}

// padded sample (c, r) of a 16-bit view [ns][nt]: zero outside [1, ns] x [1, nt]
// (the load itself is clamped into the view, so that it needs no branch)
__device__ __forceinline__ uint32_t padded16(const uint16_t *view, const GeomF &g, int c, int r)
{
	const bool in = c >= 1 && c <= g.ns && r >= 1 && r <= g.nt;
	const int cc = min(max(c, 1), g.ns) - 1, rr = min(max(r, 1), g.nt) - 1;
	const uint32_t v = view[cc * g.nt + rr];
	return in ? v : 0u;
}

// Consumer side of the walk modes (FP32): consumer wave `wave` adds the
// views of the stage buffer at `bo` to its ZL register accumulators (slices
// [zc, zc + ZL) of its column), in view order (EXACT reproduces the
// reference's sum order).  The next view's record is read while the current
// one is added, so the per-view LDS round trips overlap the slice arithmetic.
// R chunks per consumer wave: chunk c of the z-group (ZL slices from
// zg0 + c ZL) belongs to wave c % NW, as its accumulators acc[c / NW]; the
// chunks of a wave are NW ZL slices apart, so that a helical view, which
// covers only part of the group, still gives every wave work.
template <int NW, int R, int ZL, int FMT, bool FAST>
__device__ __forceinline__ void consume_walk(const BackArgs &a, int bo,
	const Lds<NW * R, true> &L, int v0, int nv, int wave, int lane, int zg0,
	float (&accs)[R][ZL])
{
	constexpr int NC = NW * R;
	const GeomF &g = a.g;
	const float4 *P4 = reinterpret_cast<const float4 *>(fdk_lds + bo + L.P) + lane * 2;
	const float *Tw = fdk_lds + bo + L.T + wave * 64 + lane;
	// R > 1: the view's four info words; R = 1: tw and mode (the last two)
	using Info = std::conditional_t<(R > 1), int4, int2>;
	constexpr int IW = R > 1 ? 1 : 2;   // Info entries per view
	const Info *VI = reinterpret_cast<const Info *>(fdk_lds + bo + L.vinfo + (R > 1 ? 0 : 2));
	float4 np0 = P4[0], np1 = P4[1];
	float nts[R];
#pragma unroll
	for (int r = 0; r < R; ++r)
		nts[r] = Tw[r * NW * 64];
	Info nvi = VI[0];
	for (int vv = 0; vv < nv; ++vv) {
		const float4 p0 = np0, p1 = np1;
		const Info vi = nvi;
		float ts[R];               // t_bin at the lane's first active slice of each chunk
#pragma unroll
		for (int r = 0; r < R; ++r)
			ts[r] = nts[r];
		if (vv + 1 < nv) {
			np0 = P4[(vv + 1) * 128];
			np1 = P4[(vv + 1) * 128 + 1];
#pragma unroll
			for (int r = 0; r < R; ++r)
				nts[r] = Tw[((vv + 1) * NC + r * NW) * 64];
			nvi = VI[(vv + 1) * IW];
		}
		int zu = 0, zi = 0, mode, tw;
		if constexpr (R > 1) {
			// the tile's union and intersection of slice ranges (scalar)
			zu = __builtin_amdgcn_readfirstlane(vi.x);
			zi = __builtin_amdgcn_readfirstlane(vi.y);
			mode = zu & 3;
			tw = vi.z;
		} else {
			mode = __builtin_amdgcn_readfirstlane(vi.y) >> 16;
			tw = vi.x;
		}
		if (mode == MODE_NONE)
			continue;
		const int zlohi = __float_as_int(p1.y);
		const int zlo = zlohi & 0xffff, zhi = (int) ((unsigned) zlohi >> 16);
#pragma unroll
		for (int r = 0; r < R; ++r) {
			float (&acc)[ZL] = accs[r];
			const float t0 = ts[r];
			const int zoff = (wave + r * NW) * ZL;
			const int zc = zg0 + zoff;
			const int jlo = max(zlo - zc, 0);
			const int jhi = min(zhi - zc, ZL);
			const int ulo = max(((zu >> 2) & 1023) - zoff, 0);
			const int uhi = min(((zu >> 12) & 1023) - zoff, ZL);
			const int ilo = max((zi & 1023) - zoff, 0);
			const int ihi = min(((zi >> 10) & 1023) - zoff, ZL);
			if (R > 1 ? ulo >= uhi : !__any(jlo < jhi))
				continue;                  // no lane of this wave meets the view here
			const int cb = __float_as_int(p0.x);
			const float wl = p0.y, wr = p0.z, w2 = p0.w;
			const float ti = p1.x;
			const unsigned nj = (unsigned) max(jhi - jlo, 0);
			if (mode == MODE_LDS) {
				const int c1 = bo + cb;
				const int c2 = c1 + __builtin_amdgcn_readfirstlane(tw);
				if (R > 1 ? ilo == 0 && ihi == ZL : __all(jlo == 0 && jhi == ZL)) {
					// every lane covers the whole chunk: straight-line slices
					float t = t0;
#pragma unroll
					for (int j = 0; j < ZL; ++j) {
						acc[j] = slice_update<FMT, FAST>(acc[j], t, c1, c2, wl, wr, w2);
						t += ti;
					}
				} else {
					// partial chunk, in groups of GS slices: a group every lane
					// adds runs straight (t of each lane is then at the group's
					// first slice); in a group some lanes add, slices outside
					// [jlo, jhi) read inside the staged footprint and are
					// discarded by a select; a group no lane adds is skipped (t
					// advances only on added slices).  Several chunks per wave
					// (helical windows cut them often) take finer groups.
					constexpr int GS = R > 1 ? 4 : 8;
					float t = t0;
#pragma unroll
					for (int g0 = 0; g0 < ZL; g0 += GS) {
						if (R > 1 ? g0 >= uhi || g0 + GS <= ulo
							: !__any(jlo < g0 + GS && jhi > g0 && jlo < jhi))
							continue;
						if (R > 1 && ilo <= g0 && ihi >= g0 + GS) {
#pragma unroll
							for (int j = g0; j < g0 + GS; ++j) {
								acc[j] = slice_update<FMT, FAST>(acc[j], t, c1, c2, wl, wr, w2);
								t += ti;
							}
							continue;
						}
#pragma unroll
						for (int j = g0; j < g0 + GS; ++j) {
							const bool rr = (unsigned) (j - jlo) < nj;
							const float tv = rr ? t : t0;
							const float up = slice_update<FMT, FAST>(acc[j], tv, c1, c2, wl, wr, w2);
							acc[j] = rr ? up : acc[j];
							t = rr ? t + ti : t;
						}
					}
				}
			} else {
				// footprint larger than the LDS slot: gather from the padded copy
				// in HBM, every slice computed (no per-lane branches, so the
				// accumulators stay in place), the ones outside [jlo, jhi) read at
				// t0, inside the padded view, and discarded by a select
				const int is = cb;                    // the record holds is for these views
				const int pnt = g.nt + 2;
				const size_t col = ((size_t) (v0 + vv) * (g.ns + 2) + is + 1) * pnt;
				const size_t c1 = col + 1, c2 = c1 + pnt;  // (s = is, t = 0), (is+1, 0)
				float t = t0;
#pragma unroll
				for (int j = 0; j < ZL; ++j) {
					const bool rr = (unsigned) (j - jlo) < nj;
					const float tv = rr ? t : t0;
					const int it = (int) floorf(tv);
					const float wu = tv - (float) it;
					const float wd = 1.f - wu;
					const float a0 = load_sample(a.padded, c1 + it);
					const float a1 = load_sample(a.padded, c1 + it + 1);
					const float b0 = load_sample(a.padded, c2 + it);
					const float b1 = load_sample(a.padded, c2 + it + 1);
					float up;
					if (FAST) {
						const float q1 = fmaf(wr, b0, wl * a0);
						const float q2 = fmaf(wr, b1, wl * a1);
						up = fmaf(wu, q2, fmaf(wd, q1, acc[j]));
					} else {
						const float q1 = wl * a0 + wr * b0;
						const float q2 = wl * a1 + wr * b1;
						const float p0v = wd * q1 + wu * q2;
						up = acc[j] + w2 * p0v;
					}
					acc[j] = rr ? up : acc[j];
					t = rr ? t + ti : t;
					// one slice's gathers at a time (register budget)
					__builtin_amdgcn_sched_barrier(0);
				}
			}
		}
	}
}

// The view's slice window [izb0, izb1) (a helix; wave-uniform), else all
// slices.
template <bool WINDOW>
__device__ __forceinline__ int2 view_window(const BackArgs &a, int v)
{
	if (!WINDOW)
		return make_int2((int) 0x80000000, 0x7fffffff);
	const int2 w = a.izb[v];
	return make_int2(__builtin_amdgcn_readfirstlane(w.x), __builtin_amdgcn_readfirstlane(w.y));
}

// Consumer side of DIRECT (FP16/BF16 storage): the views staged in LDS
// first, in view order, then -- after the stage's LDS views, which only
// changes the FP32 summation order -- any view whose footprint exceeded the
// slot, gathered from HBM.  Keeping the two apart keeps one code path over
// the accumulators in the main loop, so they stay in place in registers.
// Every slice of a chunk a wave works on inside the view's slice window is
// computed with u clamped to [0, u_hi] (ColRange): slices outside a lane's
// range read zero rows with vertical weight 0, and lanes without work read
// the zero words.  WINDOW (a helical plan): the view's slice window is the
// same for every lane; groups of 8 slices that it cuts are added in a
// second pass over the stage's views, with a select on the window, so that
// the main loop keeps one code path over the accumulators (and they stay in
// place in registers).  The slices outside the window
// are not on the clamp's zero rows when the window, not the cone, ends a
// lane's range, so they must not be added.
// The R chunks of consumer wave `wave` (of NCW) are chunks wave + r NCW of the
// z-group, ZL slices each, with accumulators accs[r]; one pass over the
// stage's views adds all of them (each view's record and window read once).
template <int NCW, int R, int ZL, int FMT, bool WINDOW>
__device__ __forceinline__ void consume_direct(const BackArgs &a, int bo,
	const Lds<NCW * R, false> &L, int v0, int nv, int wave, int lane, int zg0,
	float (&accs)[R][ZL])
{
	const GeomF &g = a.g;
	const float4 *P4 = reinterpret_cast<const float4 *>(fdk_lds + bo + L.P) + lane * 2;
	const int2 *VI = reinterpret_cast<const int2 *>(fdk_lds + bo + L.vinfo + 2);
	const int *ZR = reinterpret_cast<const int *>(fdk_lds + bo + L.vinfo);
	// relative to the z-group: chunk r starts at slice zoff(r)
	auto zoff = [&](int r) { return (wave + r * NCW) * ZL; };
	// window of chunk r relative to its first slice, clamped to [0, ZL]
	auto window = [&](int2 wa, int r) {
		const int zc = zg0 + zoff(r);
		return WINDOW ? make_int2(max(wa.x - zc, 0), min(wa.y - zc, ZL)) : make_int2(0, ZL);
	};
	bool hbm_views = false, cut_views = false;
	int nzr = ZR[0];
	for (int vv = 0; vv < nv; ++vv) {
		const int zr = __builtin_amdgcn_readfirstlane(nzr);
		if (vv + 1 < nv)
			nzr = ZR[(vv + 1) * 4];
		const int mode = zr & 3;
		hbm_views = hbm_views || mode == MODE_GLOBAL;
		if (mode != MODE_LDS)
			continue;
		// the union of the tile's slice ranges (scalar): a group inside it
		// where some lanes have no slices adds exactly 0 for those lanes
		// (their u is clamped onto zero rows, or the group is one the view's
		// window cuts, added with a select below)
		const int ulo = (zr >> 2) & 1023, uhi = (zr >> 12) & 1023;
		bool meets = false;
#pragma unroll
		for (int r = 0; r < R; ++r)
			meets = meets || (ulo < zoff(r) + ZL && uhi > zoff(r));
		if (!meets)
			continue;                  // no lane of this wave meets the view
		const float4 p0 = P4[vv * 128], p1 = P4[vv * 128 + 1];
		const int c1 = bo + __float_as_int(p0.x);
		const float wl = p0.y, wr = p0.z;
		const float ti = p1.x;
		const float uh = p1.z;
		const int2 wa = view_window<WINDOW>(a, v0 + vv);
#pragma unroll
		for (int r = 0; r < R; ++r) {
			float (&acc)[ZL] = accs[r];
			const int jlo = max(ulo - zoff(r), 0);
			const int jhi = min(uhi - zoff(r), ZL);
			if (jlo >= jhi)
				continue;
			const float t0 = fmaf((float) zoff(r), ti, p0.w);   // u at the chunk's first slice
			const int2 w = window(wa, r);
#pragma unroll
			for (int g0 = 0; g0 < ZL; g0 += 8) {
				// groups of 8 slices outside the union are skipped; groups the
				// view's window cuts (a helix) are left to the pass below
				if (g0 >= w.y || g0 + 8 <= w.x || g0 >= jhi || g0 + 8 <= jlo)
					continue;
				if (g0 < w.x || g0 + 8 > w.y) {
					cut_views = true;
					continue;
				}
				// axial: halves of 4 slices outside the union skipped as well
				// (finer tests cost more than they save where windows cut groups)
				constexpr int HS = WINDOW ? 8 : 4;
#pragma unroll
				for (int h0 = g0; h0 < g0 + 8; h0 += HS) {
					if (h0 + HS <= jlo || h0 >= jhi)
						continue;
					if constexpr (!WINDOW) {
						// the half's four LDS reads first, then its arithmetic
						// (a helical plan's 8-slice groups would need the
						// registers of eight reads in flight)
						uint32_t w0[HS], w1[HS];
						float wu[HS];
#pragma unroll
						for (int k = 0; k < HS; ++k) {
							const float u = __builtin_amdgcn_fmed3f(
								fmaf((float) (h0 + k), ti, t0), 0.f, uh);
							const int iu = (int) u;
							wu[k] = __builtin_amdgcn_fractf(u);
							w0[k] = __float_as_uint(fdk_lds[c1 + iu]);
							w1[k] = __float_as_uint(fdk_lds[c1 + iu + 1]);
						}
#pragma unroll
						for (int k = 0; k < HS; ++k)
							acc[h0 + k] = bilerp_d(acc[h0 + k], pair_lo<FMT>(w0[k]),
								pair_hi<FMT>(w0[k]), pair_lo<FMT>(w1[k]),
								pair_hi<FMT>(w1[k]), wl, wr, wu[k]);
					} else {
#pragma unroll
						for (int j = h0; j < h0 + HS; ++j) {
							const float u = __builtin_amdgcn_fmed3f(fmaf((float) j, ti, t0),
								0.f, uh);
							acc[j] = slice_update_d<FMT>(acc[j], u, c1, wl, wr);
						}
					}
				}
			}
		}
	}
	if (cut_views) {
		// groups of 8 slices that a view's window cuts: the slices outside the
		// window keep their sums (a select on a wave-uniform condition)
		for (int vv = 0; vv < nv; ++vv) {
			const int zr = __builtin_amdgcn_readfirstlane(ZR[vv * 4]);
			if ((zr & 3) != MODE_LDS)
				continue;
			const int2 wa = view_window<WINDOW>(a, v0 + vv);
			bool cut = false;
#pragma unroll
			for (int r = 0; r < R; ++r) {
				const int2 w = window(wa, r);
				cut = cut || !(w.x <= 0 && w.y >= ZL);
			}
			if (!cut)
				continue;
			const int ulo = (zr >> 2) & 1023, uhi = (zr >> 12) & 1023;
			const float4 p0 = P4[vv * 128], p1 = P4[vv * 128 + 1];
			const int c1 = bo + __float_as_int(p0.x);
			const float wl = p0.y, wr = p0.z;
			const float ti = p1.x;
			const float uh = p1.z;
#pragma unroll
			for (int r = 0; r < R; ++r) {
				float (&acc)[ZL] = accs[r];
				const int jlo = max(ulo - zoff(r), 0);
				const int jhi = min(uhi - zoff(r), ZL);
				const int2 w = window(wa, r);
				if (jlo >= jhi || (w.x <= 0 && w.y >= ZL))
					continue;
				const float t0 = fmaf((float) zoff(r), ti, p0.w);
#pragma unroll
				for (int g0 = 0; g0 < ZL; g0 += 8) {
					if (g0 >= w.y || g0 + 8 <= w.x || (w.x <= g0 && g0 + 8 <= w.y)
						|| g0 >= jhi || g0 + 8 <= jlo)
						continue;
#pragma unroll
					for (int j = g0; j < g0 + 8; ++j) {
						const float u = __builtin_amdgcn_fmed3f(fmaf((float) j, ti, t0), 0.f, uh);
						const float up = slice_update_d<FMT>(acc[j], u, c1, wl, wr);
						acc[j] = (j >= w.x && j < w.y) ? up : acc[j];
					}
				}
			}
		}
	}
	if (!hbm_views)
		return;
	// views whose footprint exceeded the LDS slot: gather from the caller's
	// views, zero outside the detector; lanes without slices in a view read
	// columns left of the detector (zeros)
	for (int vv = 0; vv < nv; ++vv) {
		const int2 vi = VI[vv * 2];
		if ((__builtin_amdgcn_readfirstlane(vi.y) >> 16) != MODE_GLOBAL)
			continue;
		const float4 p0 = P4[vv * 128], p1 = P4[vv * 128 + 1];
		const int zlohi = __float_as_int(p1.y);
		const int zlo = zlohi & 0xffff, zhi = (int) ((unsigned) zlohi >> 16);
		const float wl = p0.y, wr = p0.z;
		const float ti = p1.x;
		const float uh = p1.z;
		const uint16_t *pv = static_cast<const uint16_t *>(a.proj)
			+ (size_t) (v0 + vv) * g.ns * g.nt;
		const int2 wa = view_window<WINDOW>(a, v0 + vv);
#pragma unroll
		for (int r = 0; r < R; ++r) {
			float (&acc)[ZL] = accs[r];
			const int zc = zg0 + zoff(r);
			const int jlo = max(zlo - zc, 0);
			const int jhi = min(zhi - zc, ZL);
			if (!__any(jlo < jhi))
				continue;
			const int is = (jlo < jhi) ? __float_as_int(p0.x) : -3;   // the record holds is
			const float t0 = fmaf((float) zoff(r), ti, p0.w);
			const int2 w = window(wa, r);
#pragma unroll
			for (int j = 0; j < ZL; ++j) {
				const float u = __builtin_amdgcn_fmed3f(fmaf((float) j, ti, t0), 0.f, uh);
				const int iu = (int) u;
				const float wu = __builtin_amdgcn_fractf(u);
				const float up = bilerp_d(acc[j],
					sample16<FMT>(padded16(pv, g, is + 1, iu)),
					sample16<FMT>(padded16(pv, g, is + 2, iu)),
					sample16<FMT>(padded16(pv, g, is + 1, iu + 1)),
					sample16<FMT>(padded16(pv, g, is + 2, iu + 1)), wl, wr, wu);
				acc[j] = (j >= w.x && j < w.y) ? up : acc[j];
				// one slice's four gathers at a time (register budget)
				__builtin_amdgcn_sched_barrier(0);
			}
		}
	}
}

// Workgroup = NW consumer waves (z chunks of ZL slices, lane = column) and
// NP producer waves.  Views go in stages of `vstage` through two LDS
// buffers: while the consumers add stage s from one buffer, the producers
// set up and stage s+1 into the other; one barrier per stage.
template <int NW, int NP, int ZL, int FMT, bool FAST, int R = 1, bool WIN = false>
__device__ __forceinline__ void tile_pipeline(const BackArgs &a, int tile, int group)
{
	static_assert(!WIN || FMT != FMT_F32, "the walk modes keep to the windows by their ranges");
	static_assert(NW * R * ZL <= 1023, "DIRECT packs z-group-relative slice ranges in 10 bits");
	constexpr int NC = NW * R;   // ZL-slice chunks of the z-group
	const GeomF &g = a.g;
	const int tid = threadIdx.x;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int lane = tid & 63;
	const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
	const int zg0 = a.z0 + group * (NC * ZL);
	const int zg1 = min(zg0 + NC * ZL, g.nz);
	const int ix = tx * 8 + (lane & 7);
	const int iy = ty * 8 + (lane >> 3);

	const Lds<NC, FMT == FMT_F32> L(a.vstage, a.slot);
	// the views whose slice window meets this z-group (a helix), else all
	int va = 0, vb = g.na;
	if (a.gviews) {
		// (at most g.na of them: the autotuner times a prefix of each range)
		const int2 r = a.gviews[group];
		va = __builtin_amdgcn_readfirstlane(r.x);
		vb = min(__builtin_amdgcn_readfirstlane(r.y), va + g.na);
	}
	const int nstage = (vb - va + a.vstage - 1) / a.vstage;

	if (wave >= NW) {
		// ------------------------------------------------ producer waves
		const int pw = wave - NW;
		const bool colvalid = ix < g.nx && iy < g.ny &&
			(a.mask == nullptr || a.mask[(long) iy * g.nx + ix] != 0);
		const float yy = g.dy * ((float) iy - g.wy);   // fdk,ts.c:114,120
		const float xx = g.dx * ((float) ix - g.wx);
		if (FMT != FMT_F32 && pw == 0 && lane < 2)
			fdk_lds[L.zero + lane] = 0.f;   // read by DIRECT lanes without work
		if (nstage > 0)
			produce<NC, NP, ZL, FMT, FAST, R>(a, 0, L, va, min(a.vstage, vb - va), pw, lane,
				colvalid, xx, yy, zg0, zg1);
		__builtin_amdgcn_s_waitcnt(0);   // staged footprints have landed in LDS
		__syncthreads();
		for (int s = 0; s < nstage; ++s) {
			if (s + 1 < nstage) {
				const int v0 = va + (s + 1) * a.vstage;
				produce<NC, NP, ZL, FMT, FAST, R>(a, ((s + 1) & 1) * L.total, L, v0,
					min(a.vstage, vb - v0), pw, lane, colvalid, xx, yy, zg0, zg1);
			}
			__builtin_amdgcn_s_waitcnt(0);
			__syncthreads();
		}
	} else {
		// ------------------------------------------------ consumer waves
		// chunks wave, wave + NW, ... (R of them) of the z-group
		float accs[R][ZL];
#pragma unroll
		for (int r = 0; r < R; ++r)
#pragma unroll
			for (int j = 0; j < ZL; ++j)
				accs[r][j] = 0.f;
		__syncthreads();
		for (int s = 0; s < nstage; ++s) {
			const int v0 = va + s * a.vstage;
			if constexpr (FMT == FMT_F32)
				consume_walk<NW, R, ZL, FMT, FAST>(a, (s & 1) * L.total, L, v0,
					min(a.vstage, vb - v0), wave, lane, zg0, accs);
			else {
				// chunks wave + r NW of the z-group in one pass over the
				// stage's views (u is in closed form from the group's U0)
				consume_direct<NW, R, ZL, FMT, WIN>(a, (s & 1) * L.total, L, v0,
					min(a.vstage, vb - v0), wave, lane, zg0, accs);
			}
			__syncthreads();
		}
		// one write per voxel, [nz nx ny]
		if (ix < g.nx && iy < g.ny) {
			float *col = a.image + ((long) iy * g.nx + ix) * g.nz;
#pragma unroll
			for (int r = 0; r < R; ++r) {
				const int zc = zg0 + (wave + r * NW) * ZL;
#pragma unroll
				for (int j = 0; j < ZL; ++j) {
					const int z = zc + j;
					if (z < zg1)
						col[z] = accs[r][j];
				}
			}
		}
	}
}

// The second launch bound MW is the minimum waves per SIMD: 4 (at most 128
// VGPRs), or 8 (at most 64) for a second build of the one-chunk FP16/BF16
// kernels, whose small LDS stages can then let four 8-wave workgroups share
// a CU; which one runs is a plan's choice (its autotuner times both).
// WIN: FP16/BF16 views of a helical plan (slice windows per view)
template <int NW, int NP, int ZL, int FMT, bool FAST, int R, bool WIN, int MW>
__global__ void __launch_bounds__((NW + NP) * 64, MW)
back_kernel(const BackArgs a)
{
	const int nblocks = a.ntiles * a.ngroups;
	const int lb = xcd_block(blockIdx.x, nblocks);
	const int tile = lb / a.ngroups;
	tile_pipeline<NW, NP, ZL, FMT, FAST, R, WIN>(a, tile, lb - tile * a.ngroups);
}

// Persistent form for a batch of scans of uneven sizes: every workgroup takes
// (scan, tile, z-group) items from a device counter until the queue is empty.
// Item ranges are contiguous per scan (item_begin, a prefix sum), so the
// scan of an item is found by binary search.  All scans share the waves per
// workgroup and the LDS layout (vstage, slot); the slices per consumer wave
// are per scan, ZL or ZL/2 (a.zper): a steep-cone scan whose footprint over
// NW*ZL slices overflows the shared slot takes z-groups half as tall, whose
// footprints fit, instead of gathering from HBM.
template <int NW, int NP, int ZL, int FMT, bool FAST>
__global__ void __launch_bounds__((NW + NP) * 64, 4)
batch_kernel(const BatchArgs b)
{
	__shared__ long long next_item;
	for (;;) {
		if (threadIdx.x == 0)
			next_item = (long long) atomicAdd(b.counter, 1ull);
		__syncthreads();
		// the item is the same in every lane: keep it (and all that derives
		// from it) in scalar registers
		const long long it = next_item;
		const long long item = (long long) (((unsigned long long) (unsigned)
			__builtin_amdgcn_readfirstlane((int) (it >> 32)) << 32)
			| (unsigned) __builtin_amdgcn_readfirstlane((int) it));
		__syncthreads();
		if (item >= b.nitems)
			break;
		int lo = 0, hi = b.nscan - 1;
		while (lo < hi) {
			const int mid = (lo + hi + 1) >> 1;
			if (b.item_begin[mid] <= item)
				lo = mid;
			else
				hi = mid - 1;
		}
		// the scan table is read-only for the kernel's lifetime: read it
		// through the constant address space so that it lands in SGPRs
		BackArgs a;
		{
			constexpr int nw32 = sizeof(BackArgs) / 4;
			const auto *src = (const __attribute__((address_space(4))) int *) (b.scans + lo);
			int *dst = reinterpret_cast<int *>(&a);
#pragma unroll
			for (int w = 0; w < nw32; ++w)
				dst[w] = src[w];
		}
		const int local = (int) (item - b.item_begin[lo]);
		const int tile = local / a.ngroups;
		constexpr int ZH = ZL >= 16 ? ZL / 2 : ZL;
		if (ZH != ZL && a.zper == ZH)
			tile_pipeline<NW, NP, ZH, FMT, FAST>(a, tile, local - tile * a.ngroups);
		else
			tile_pipeline<NW, NP, ZL, FMT, FAST>(a, tile, local - tile * a.ngroups);
	}
}

__global__ void reset_counter_kernel(unsigned long long *c)
{
	if (threadIdx.x == 0)
		*c = 0ull;
}

// ---------------------------------------------------------------------------
// survey kernel: per plan, over the same tiling as back_kernel -- counts the
// reference loop's (voxel, view) iterations, the largest staged footprint and
// whether any column meets any view (for the dz/dt < 0 error).

template <int NW, int ZL, bool DIRECT>
__global__ void __launch_bounds__(64)
survey_kernel(const BackArgs a)
{
	const GeomF &g = a.g;
	const int lane = threadIdx.x;
	const int nblocks = a.ntiles * a.ngroups;
	const int lb = blockIdx.x;
	if (lb >= nblocks)
		return;
	const int tile = lb / a.ngroups;
	const int group = lb - tile * a.ngroups;
	const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
	const int zg0 = a.z0 + group * (NW * ZL);
	const int zg1 = min(zg0 + NW * ZL, g.nz);
	const int ix = tx * 8 + (lane & 7);
	const int iy = ty * 8 + (lane >> 3);
	const bool colvalid = ix < g.nx && iy < g.ny &&
		(a.mask == nullptr || a.mask[(long) iy * g.nx + ix] != 0);
	const float yy = g.dy * ((float) iy - g.wy);
	const float xx = g.dx * ((float) ix - g.wx);
	unsigned long long cnt = 0;
	int maxarea = 0, any = 0;
	for (int v = 0; v < g.na; ++v) {
		const ViewGeo vg = view_geo(a, v);
		ColSetup c;
		const int code = colvalid ? column_setup(g, xx, yy, vg, c) : SETUP_OUT;
		any |= (code != SETUP_OUT) ? 1 : 0;
		// reference loop iterations of this column-view (whole range, once)
		if (code == SETUP_ACTIVE && group == 0)
			cnt += (unsigned long long) max(c.iz_max - c.iz_min, 0);
		// the footprint exactly as the backprojection kernel stages it
		ColRange cr{};
		float Tk[NW];
		const bool act = colvalid
			&& col_range<NW, ZL, DIRECT>(g, xx, yy, vg, zg0, zg1, cr, Tk);
		const Footprint f = view_footprint<DIRECT>(g, act, cr);
		if (f.s_lo <= f.s_hi && f.t_lo <= f.t_hi) {
			const int rows = f.t_hi - f.t_lo + 1;
			const long area = (long) (f.s_hi - f.s_lo + (DIRECT ? 0 : 1)) * (DIRECT ? rows | 1 : rows);
			maxarea = max(maxarea, (int) min(area, (long) 0x7fffffff));
		}
	}
	// wave reduce the count
	for (int o = 32; o > 0; o >>= 1)
		cnt += __shfl_xor(cnt, o, 64);
	any = wave_max(any);
	if (lane == 0) {
		atomicAdd(a.count, cnt);
		atomicMax(a.maxarea, maxarea);
		if (any)
			atomicMax(a.fov_any, 1);
	}
}

// ---------------------------------------------------------------------------
// zero fill of the slices outside every z-group, [nz nx ny]

__global__ void fill_kernel(float *image, long ncol, int nz, int zlo, int zhi)
{
	// zero z in [0, zlo) and [zhi, nz) of every column
	const int nout = zlo + (nz - zhi);
	const long total = ncol * nout;
	for (long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < total;
		i += (long) gridDim.x * blockDim.x) {
		const long col = i / nout;
		const int r = (int) (i - col * nout);
		const int z = (r < zlo) ? r : zhi + (r - zlo);
		image[col * nz + z] = 0.f;
	}
}

// zero-padded copy of the views: [na][ns][nt] -> [na][ns+2][nt+2] (the
// one-sample zero border of fdk_ts_put_view, fdk,ts,t.c:13-21, for all views)
template <typename T>
__global__ void pad_kernel(const T *src, T *dst, int ns, int nt, long total)
{
	const int pnt = nt + 2, pns = ns + 2;
	for (long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < total;
		i += (long) gridDim.x * blockDim.x) {
		const int t = (int) (i % pnt);
		const long q = i / pnt;
		const int s = (int) (q % pns);
		const long v = q / pns;
		T val = T(0);
		if (s >= 1 && s <= ns && t >= 1 && t <= nt)
			val = src[(v * ns + (s - 1)) * nt + (t - 1)];
		dst[i] = val;
	}
}

// Tiled transpose with row shift/clamp and scale, the device form of the
// host-side permutes around the mex call (fbp/cbct_back.m:86-98,104,133-135):
//   out[b*out_b + c*out_c + r] = scale * in[b*in_b + clamp(r - rshift, 0, R-1)*in_r + c]
// for r in [0, Rout), c in [0, C), b in [0, B).  32 x 32 tiles through LDS so
// that reads (c fastest) and writes (r fastest) are both coalesced.
__global__ void __launch_bounds__(256)
transpose_kernel(const float *in, float *out, int B, int R, int C, long in_b, long in_r,
	long out_b, long out_c, int Rout, int rshift, float scale)
{
	__shared__ float tile[32][33];
	const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
	for (int b = blockIdx.z; b < B; b += gridDim.z) {
		const float *ib = in + (long) b * in_b;
		float *ob = out + (long) b * out_b;
#pragma unroll
		for (int k = 0; k < 32; k += 8) {
			const int r = r0 + ty + k, c = c0 + tx;
			float v = 0.f;
			if (r < Rout && c < C) {
				const int rs = min(max(r - rshift, 0), R - 1);
				v = ib[(long) rs * in_r + c];
			}
			tile[ty + k][tx] = v;
		}
		__syncthreads();
#pragma unroll
		for (int k = 0; k < 32; k += 8) {
			const int c = c0 + ty + k, r = r0 + tx;
			if (r < Rout && c < C)
				ob[(long) c * out_c + r] = scale * tile[tx][ty + k];
		}
		__syncthreads();
	}
}

__global__ void zero_u64_kernel(unsigned long long *p, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 0ull;
}

__global__ void zero_i32_kernel(int *p, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 0;
}

// ---------------------------------------------------------------------------
// 'fdk,st,back' (fdk,st.c:38-141): proj [ns nt na], image [nx ny nz].
// Thread per (x,y) column and a chunk of ST_Z slices; views inner.

constexpr int ST_Z = 16;

__global__ void __launch_bounds__(256)
st_kernel(const StArgs a)
{
	const GeomF &g = a.g;
	const long nxy = (long) g.nx * g.ny;
	const long j = (long) blockIdx.x * blockDim.x + threadIdx.x;
	const int zc = blockIdx.y * ST_Z;
	if (j >= nxy)
		return;
	const int ix = (int) (j % g.nx), iy = (int) (j / g.nx);
	float acc[ST_Z];
#pragma unroll
	for (int k = 0; k < ST_Z; ++k)
		acc[k] = 0.f;
	const bool valid = a.mask == nullptr || a.mask[j] != 0;
	if (valid) {
		const float yy = g.dy * ((float) iy - g.wy);
		const float xx = g.dx * ((float) ix - g.wx);
		const long view_elems = (long) g.ns * g.nt;
		for (int v = 0; v < g.na; ++v) {
			const float2 cs = a.trig[v];
			const float xbeta = xx * cs.x + yy * cs.y;
			const float ybetas = g.dso - (-xx * cs.y + yy * cs.x);
			const float mag = (g.kind == KIND_PAR) ? 1.f : g.dsd / ybetas;
			float ss;
			if (g.kind == KIND_ARC)
				ss = (float) ((double) g.dsd * atan2((double) xbeta, (double) ybetas));
			else
				ss = mag * xbeta;
			const float ss_bin = ss / g.ds + g.ws;
			const int is = to_int_x86(floorf(ss_bin));
			if (is < 0 || is >= g.ns - 1)
				continue;
			const float wr = ss_bin - (float) is, wl = 1.f - wr;
			const float w2 = (g.kind == KIND_ARC)
				? (g.dsd * g.dsd) / (ybetas * ybetas + xbeta * xbeta)
				: mag * mag;
			const float *view = a.proj + (long) v * view_elems;
#pragma unroll
			for (int k = 0; k < ST_Z; ++k) {
				const int iz = zc + k;
				if (iz >= g.nz)
					break;
				const float zz = g.dz * ((float) iz - g.wz);
				const float tt = mag * zz;
				const float tt_bin = tt / g.dt + g.wt;
				const int it = to_int_x86(floorf(tt_bin));
				if (it < 0 || it >= g.nt - 1)
					continue;
				const float wu = tt_bin - (float) it, wd = 1.f - wu;
				const float p1 = wl * view[is + it * g.ns] + wr * view[is + 1 + it * g.ns];
				const float p2 = wl * view[is + (it + 1) * g.ns]
					+ wr * view[is + 1 + (it + 1) * g.ns];
				acc[k] += w2 * (wu * p1 + wd * p2);
			}
		}
	}
#pragma unroll
	for (int k = 0; k < ST_Z; ++k) {
		const int iz = zc + k;
		if (iz < g.nz)
			a.image[j + (long) iz * nxy] = acc[k];
	}
}

// ---------------------------------------------------------------------------
// per-view quantisation (storage step of the F16 / BF16 paths)

__global__ void __launch_bounds__(256)
view_maxabs_kernel(const float *proj, long view_elems, int *maxbits)
{
	const int v = blockIdx.y;
	const float *p = proj + (long) v * view_elems;
	float m = 0.f;
	for (long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < view_elems;
		i += (long) gridDim.x * blockDim.x) {
		const float x = fabsf(p[i]);
		m = (x > m || x != x) ? x : m;   // NaN wins
	}
	for (int o = 32; o > 0; o >>= 1) {
		const float y = __shfl_xor(m, o, 64);
		m = (y > m || y != y) ? y : m;
	}
	__shared__ float red[4];
	if ((threadIdx.x & 63) == 0)
		red[threadIdx.x >> 6] = m;
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int w = 1; w < (int) (blockDim.x >> 6); ++w)
			m = (red[w] > m || red[w] != red[w]) ? red[w] : m;
		// non-negative floats (and NaN, 0x7fc..) order like their bit patterns
		atomicMax(maxbits + v, (int) __float_as_uint(m));
	}
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f)
{
	uint32_t u = __float_as_uint(f);
	if ((u & 0x7fffffffu) > 0x7f800000u)  // NaN: keep it quiet
		return (uint16_t) ((u >> 16) | 0x40u);
	u += 0x7fffu + ((u >> 16) & 1u);
	return (uint16_t) (u >> 16);
}

__global__ void __launch_bounds__(256)
view_quantize_kernel(const float *proj, long view_elems, int fmt,
	const int *maxbits, void *out, float *scale)
{
	const int v = blockIdx.y;
	const float m = __uint_as_float((uint32_t) maxbits[v]);
	int e = 0;
	if (m > 0.f && m <= 3.4e38f) {
		int ex;
		frexpf(m, &ex);   // m = f * 2^ex, f in [0.5, 1)
		e = ex - 15;      // m / 2^e in [2^14, 2^15)
	}
	if (blockIdx.x == 0 && threadIdx.x == 0)
		scale[v] = ldexpf(1.f, e);
	const float *p = proj + (long) v * view_elems;
	for (long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < view_elems;
		i += (long) gridDim.x * blockDim.x) {
		const float q = ldexpf(p[i], -e);
		if (fmt == FMT_F16)
			static_cast<__half *>(out)[(long) v * view_elems + i] = __float2half_rn(q);
		else
			static_cast<uint16_t *>(out)[(long) v * view_elems + i] = f32_to_bf16_rne(q);
	}
}

// ---------------------------------------------------------------------------
// launchers (host side)

template <int NW, int NP, int ZL, int FMT, bool FAST, int R = 1, int MW = 4>
static hipError_t launch_back_t(const BackArgs &a, size_t lds_bytes, hipStream_t s)
{
	const int nblocks = a.ntiles * a.ngroups;
	// FP16/BF16 views of a helical plan: the instantiation whose consumers
	// keep to the views' slice windows (a kernel of its own, so that the
	// axial kernel's code stays as small as it is)
	if constexpr (FMT != FMT_F32) {
		if (a.izb) {
			hipLaunchKernelGGL((back_kernel<NW, NP, ZL, FMT, FAST, R, true, MW>), dim3(nblocks),
				dim3((NW + NP) * 64), lds_bytes, s, a);
			return hipGetLastError();
		}
	}
	hipLaunchKernelGGL((back_kernel<NW, NP, ZL, FMT, FAST, R, false, MW>), dim3(nblocks),
		dim3((NW + NP) * 64), lds_bytes, s, a);
	return hipGetLastError();
}

// two stage buffers of: slots + 8-word (view, column) records + the walk
// modes' t_bin table + view headers (4-byte words)
size_t back_lds_bytes(int nw, int vstage, int slot, int fmt)
{
	const size_t words = (size_t) vstage * slot + (size_t) vstage * 64 * 8
		+ (fmt == FMT_F32 ? (size_t) vstage * nw * 64 : 0) + (size_t) vstage * 4;
	return 2 * words * 4 + 2 * 4;   // two stage buffers and the zero words
}

#define FDK_DISPATCH_FMT(NW, ZL)                                                               \
	do {                                                                                       \
		constexpr int NP_ = (NW) >= 8 ? 8 : 4;                                                 \
		if (fmt == FMT_F32 && !fast) return launch_back_t<NW, NP_, ZL, FMT_F32, false>(a, lds, s); \
		if (fmt == FMT_F32 && fast) return launch_back_t<NW, NP_, ZL, FMT_F32, true>(a, lds, s);   \
		if (fmt == FMT_F16) return minw == 8 ? launch_back_t<NW, NP_, ZL, FMT_F16, true, 1, 8>(a, lds, s) \
			: launch_back_t<NW, NP_, ZL, FMT_F16, true>(a, lds, s);                                \
		if (fmt == FMT_BF16) return minw == 8 ? launch_back_t<NW, NP_, ZL, FMT_BF16, true, 1, 8>(a, lds, s) \
			: launch_back_t<NW, NP_, ZL, FMT_BF16, true>(a, lds, s);                               \
	} while (0)

bool back_minw_supported(int minw, int zr, int fmt)
{
	return minw == 4 || (minw == 8 && zr == 1 && fmt != FMT_F32);
}

hipError_t launch_back(const BackArgs &a, int nw, int zl, int zr, int fmt, bool fast,
	int minw, hipStream_t s)
{
	if (!back_minw_supported(minw, zr, fmt))
		return hipErrorInvalidValue;
	const size_t lds = back_lds_bytes(nw * zr, a.vstage, a.slot, fmt);
	if (zr > 1) {
		// chunks interleaved over the waves
#define FDK_DISPATCH_R(NW, ZL, R)                                                            \
		if (nw == NW && zl == ZL && zr == R) {                                               \
			constexpr int NP_ = (NW) >= 8 ? 8 : 4;                                           \
			if (fmt == FMT_F16) return launch_back_t<NW, NP_, ZL, FMT_F16, true, R>(a, lds, s); \
			if (fmt == FMT_BF16) return launch_back_t<NW, NP_, ZL, FMT_BF16, true, R>(a, lds, s); \
			return fast ? launch_back_t<NW, NP_, ZL, FMT_F32, true, R>(a, lds, s)            \
				: launch_back_t<NW, NP_, ZL, FMT_F32, false, R>(a, lds, s);                  \
		}
		FDK_DISPATCH_R(4, 16, 3)
		FDK_DISPATCH_R(4, 16, 4)
		FDK_DISPATCH_R(8, 16, 2)
#undef FDK_DISPATCH_R
		return hipErrorInvalidValue;
	}
	if (zl == 16) {
		if (nw == 4) FDK_DISPATCH_FMT(4, 16);
		if (nw == 8) FDK_DISPATCH_FMT(8, 16);
	} else if (zl == 8) {
		if (nw == 4) FDK_DISPATCH_FMT(4, 8);
		if (nw == 8) FDK_DISPATCH_FMT(8, 8);
	} else if (zl == 32) {
		if (nw == 4) FDK_DISPATCH_FMT(4, 32);
	}
	return hipErrorInvalidValue;
}

template <int NW, int NP, int ZL, int FMT, bool FAST>
static hipError_t launch_batch_t(const BatchArgs &b, int nwg, size_t lds_bytes, hipStream_t s)
{
	hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, b.counter);
	hipLaunchKernelGGL((batch_kernel<NW, NP, ZL, FMT, FAST>), dim3(nwg), dim3((NW + NP) * 64),
		lds_bytes, s, b);
	return hipGetLastError();
}

#define FDK_DISPATCH_BATCH(NW, ZL)                                                          \
	do {                                                                                    \
		constexpr int NP_ = (NW) >= 8 ? 8 : 4;                                              \
		if (fmt == FMT_F32 && !fast) return launch_batch_t<NW, NP_, ZL, FMT_F32, false>(b, nwg, lds, s); \
		if (fmt == FMT_F32 && fast) return launch_batch_t<NW, NP_, ZL, FMT_F32, true>(b, nwg, lds, s);   \
		if (fmt == FMT_F16) return launch_batch_t<NW, NP_, ZL, FMT_F16, true>(b, nwg, lds, s);          \
		if (fmt == FMT_BF16) return launch_batch_t<NW, NP_, ZL, FMT_BF16, true>(b, nwg, lds, s);        \
	} while (0)

hipError_t launch_batch(const BatchArgs &b, int nwg, int nw, int zl, int vstage, int slot,
	int fmt, bool fast, hipStream_t s)
{
	const size_t lds = back_lds_bytes(nw, vstage, slot, fmt);
	if (zl == 16) {
		if (nw == 4) FDK_DISPATCH_BATCH(4, 16);
		if (nw == 8) FDK_DISPATCH_BATCH(8, 16);
	} else if (zl == 8) {
		if (nw == 4) FDK_DISPATCH_BATCH(4, 8);
		if (nw == 8) FDK_DISPATCH_BATCH(8, 8);
	} else if (zl == 32) {
		if (nw == 4) FDK_DISPATCH_BATCH(4, 32);
	}
	return hipErrorInvalidValue;
}

template <int NW, int NP, int ZL, int FMT, bool FAST>
static int batch_occupancy_t(size_t lds_bytes)
{
	int n = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_kernel<NW, NP, ZL, FMT, FAST>,
			(NW + NP) * 64, lds_bytes) != hipSuccess)
		return 1;
	return n > 0 ? n : 1;
}

#define FDK_OCC_BATCH(NW, ZL)                                                               \
	do {                                                                                    \
		constexpr int NP_ = (NW) >= 8 ? 8 : 4;                                              \
		if (fmt == FMT_F32 && !fast) return batch_occupancy_t<NW, NP_, ZL, FMT_F32, false>(lds); \
		if (fmt == FMT_F32 && fast) return batch_occupancy_t<NW, NP_, ZL, FMT_F32, true>(lds);   \
		if (fmt == FMT_F16) return batch_occupancy_t<NW, NP_, ZL, FMT_F16, true>(lds);          \
		if (fmt == FMT_BF16) return batch_occupancy_t<NW, NP_, ZL, FMT_BF16, true>(lds);        \
	} while (0)

int batch_blocks_per_cu(int nw, int zl, int vstage, int slot, int fmt, bool fast)
{
	const size_t lds = back_lds_bytes(nw, vstage, slot, fmt);
	if (zl == 16) {
		if (nw == 4) FDK_OCC_BATCH(4, 16);
		if (nw == 8) FDK_OCC_BATCH(8, 16);
	} else if (zl == 8) {
		if (nw == 4) FDK_OCC_BATCH(4, 8);
		if (nw == 8) FDK_OCC_BATCH(8, 8);
	} else if (zl == 32) {
		if (nw == 4) FDK_OCC_BATCH(4, 32);
	}
	return 1;
}

bool back_config_supported(int nw, int zl, int zr, int fmt)
{
	(void) fmt;
	if (zr > 1)
		return zl == 16 && ((nw == 4 && (zr == 3 || zr == 4)) || (nw == 8 && zr == 2));
	return zr == 1 && ((zl == 16 && (nw == 4 || nw == 8)) || (zl == 8 && (nw == 4 || nw == 8))
		|| (zl == 32 && nw == 4));
}

hipError_t launch_survey(const BackArgs &a, int nw, int zl, bool direct, hipStream_t s)
{
	const int nblocks = a.ntiles * a.ngroups;
#define FDK_SURVEY(NW, ZL) \
	if (nw == NW && zl == ZL) { \
		if (direct) \
			hipLaunchKernelGGL((survey_kernel<NW, ZL, true>), dim3(nblocks), dim3(64), 0, s, a); \
		else \
			hipLaunchKernelGGL((survey_kernel<NW, ZL, false>), dim3(nblocks), dim3(64), 0, s, a); \
		return hipGetLastError(); }
	FDK_SURVEY(4, 16)
	FDK_SURVEY(8, 16)
	FDK_SURVEY(4, 8)
	FDK_SURVEY(8, 8)
	FDK_SURVEY(4, 32)
	FDK_SURVEY(12, 16)   // 4 waves x 3 chunks
	FDK_SURVEY(16, 16)   // 4 x 4 and 8 x 2
#undef FDK_SURVEY
	return hipErrorInvalidValue;
}

hipError_t launch_fill(float *image, long ncol, int nz, int zlo, int zhi, hipStream_t s)
{
	if (zlo <= 0 && zhi >= nz)
		return hipSuccess;
	const long total = ncol * (long) (zlo + (nz - zhi));
	const int blocks = (int) min((total + 255) / 256, (long) 65536);
	hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, s, image, ncol, nz, zlo, zhi);
	return hipGetLastError();
}

hipError_t launch_pad(const float *src, float *dst, int ns, int nt, int na, hipStream_t s)
{
	const long total = (long) na * (ns + 2) * (nt + 2);
	const int blocks = (int) min((total + 255) / 256, (long) 8192);
	hipLaunchKernelGGL(pad_kernel<float>, dim3(blocks), dim3(256), 0, s, src, dst, ns, nt, total);
	return hipGetLastError();
}

hipError_t launch_transpose(const float *in, float *out, int B, int R, int C, long in_b,
	long in_r, long out_b, long out_c, int Rout, int rshift, float scale, hipStream_t s)
{
	if (B <= 0 || R <= 0 || C <= 0 || Rout <= 0)
		return hipSuccess;
	const dim3 grid((unsigned) ((C + 31) / 32), (unsigned) ((Rout + 31) / 32),
		(unsigned) min(B, 65535));
	hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, s, in, out, B, R, C, in_b, in_r,
		out_b, out_c, Rout, rshift, scale);
	return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long *p, int n, hipStream_t s)
{
	hipLaunchKernelGGL(zero_u64_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
	return hipGetLastError();
}

hipError_t launch_zero_i32(int *p, int n, hipStream_t s)
{
	hipLaunchKernelGGL(zero_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
	return hipGetLastError();
}

hipError_t launch_st(const StArgs &a, hipStream_t s)
{
	const long nxy = (long) a.g.nx * a.g.ny;
	dim3 grid((unsigned) ((nxy + 255) / 256), (unsigned) ((a.g.nz + ST_Z - 1) / ST_Z));
	hipLaunchKernelGGL(st_kernel, grid, dim3(256), 0, s, a);
	return hipGetLastError();
}

hipError_t launch_quantize(const float *proj, int ns, int nt, int na, int fmt,
	void *out, float *scale, int *maxbits, hipStream_t s)
{
	const long view_elems = (long) ns * nt;
	hipError_t e = launch_zero_i32(maxbits, na, s);
	if (e != hipSuccess)
		return e;
	const int bx = (int) min((view_elems + 255) / 256, (long) 64);
	hipLaunchKernelGGL(view_maxabs_kernel, dim3(bx, na), dim3(256), 0, s, proj, view_elems, maxbits);
	if ((e = hipGetLastError()) != hipSuccess)
		return e;
	hipLaunchKernelGGL(view_quantize_kernel, dim3(bx, na), dim3(256), 0, s, proj, view_elems,
		fmt, (const int *) maxbits, out, scale);
	return hipGetLastError();
}

} // namespace fdk
