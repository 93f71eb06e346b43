/*
 * ref_driver.c -- TEST INFRASTRUCTURE ONLY (CPU checker, never shipped).
 *
 * A small C driver that calls the reference C backprojector compiled from
 * /root/reference sources (see oracle/Makefile) with the exact layouts and
 * scalar narrowing of the reference mex gateway, so that the reference
 * algorithm can be run here without MATLAB:
 *
 *   - fdk_mex_back in mex/src/fdk/fdk-orig/fdk,mex.c:22-171 partitions the
 *     uint8 mask with cbct_mask_init (:116), copies the double geometry into
 *     the float cbct_ig / cbct_cg structs (:119-137), then calls
 *     fdk_ts_back_t (:161) or, for the 'fdk,st,back' order, fdk_st_back1 once
 *     per view (:144-155).  This driver does the same steps.
 *   - Built with -DREF_HELIX it drives the debugged kernel of
 *     mex/src/fdk-helix/mex/fdk,ts,t.c:107-167 on its axial path
 *     (source_zs = 0, na1_half = 0, cone_par = 0, w3d = 0,
 *     source_dz_per_rad = 0, pitch = 0) -- the "fixed" reference -- and,
 *     through ref_fdk_ts_back_helix, with all of its helical arguments.
 *   - Without it, it drives mex/src/fdk/fdk-orig/fdk,ts,t.c:72-110, the
 *     "as-shipped" reference with the swapped vertical weights.
 *
 * Only tests/, the smoke check and the CPU-baseline leg of the benchmark use
 * the library built from this file.
 */
#include "def,fdk.h"

#include <stdint.h>
#include <stdlib.h>

/* Mirrors fdk_mex_back (fdk,mex.c:22-171) for 'fdk,ts,back'.
 * proj is [nt ns na] (t fastest), image is [nz nx ny] (z fastest).
 * Returns 1 on success, 0 on failure (the reference's sof convention). */
int ref_fdk_ts_back(
	float *image,
	const int32_t *nxyz, const double *dxyz, const double *cxyz,
	const uint8_t *mask,
	double dso, double dsd, double dfs,
	const double *dst, const double *ost,
	const float *proj, int32_t ns, int32_t nt, int32_t na,
	const double *beta, int32_t nthread)
{
	cbct_ig ig_, *ig = &ig_;
	cbct_cg cg_, *cg = &cg_;
	const int nx = nxyz[0], ny = nxyz[1], nz = nxyz[2];
	int ok;

	ig->mask2 = (byte *) calloc((size_t) nx * ny, 1);
	if (!ig->mask2) return 0;
	if (!cbct_mask_init(ig->mask2, (cbyte *) mask, nx, ny, nthread, 0)) {
		free(ig->mask2);
		return 0;
	}

	ig->nx = nx; ig->ny = ny; ig->nz = nz;
	ig->dx = dxyz[0]; ig->dy = dxyz[1]; ig->dz = dxyz[2];
	ig->offset_x = cxyz[0]; ig->offset_y = cxyz[1]; ig->offset_z = cxyz[2];

	cg->dso = dso; cg->dsd = dsd; cg->dfs = dfs;
	cg->ns = ns; cg->nt = nt;
	cg->ds = dst[0]; cg->dt = dst[1];
	cg->offset_s = ost[0]; cg->offset_t = ost[1];

#ifdef REF_HELIX
	{
		double *source_zs = (double *) calloc((size_t) (na > 0 ? na : 1), sizeof(double));
		if (!source_zs) { free(ig->mask2); return 0; }
		ok = fdk_ts_back_t(image, ig, cg, na, proj, beta,
			source_zs, 0 /* na1_half */, 0 /* cone_par */,
			0 /* w3d */, 0. /* source_dz_per_rad */, 0.f /* pitch */,
			nthread, 0);
		free(source_zs);
	}
#else
	ok = fdk_ts_back_t(image, ig, cg, na, proj, beta, nthread, 0);
#endif
	free(ig->mask2);
	return ok;
}

#ifdef REF_HELIX
/* The 19-operand 'fdk,ts,back' of the fdk-helix gateway
 * (mex/src/fdk-helix/mex/fdk,mex.c:22-176, 227-236): as ref_fdk_ts_back plus
 * source_zs [na], na1_half, cone_par, w3d, source_dz_per_rad (narrowed to
 * float by the gateway, fdk,mex.c:128) and pitch (single). */
int ref_fdk_ts_back_helix(
	float *image,
	const int32_t *nxyz, const double *dxyz, const double *cxyz,
	const uint8_t *mask,
	double dso, double dsd, double dfs,
	const double *dst, const double *ost,
	const float *proj, int32_t ns, int32_t nt, int32_t na,
	const double *beta, const double *source_zs, int32_t na1_half,
	int32_t cone_par, int32_t w3d, double source_dz_per_rad, float pitch,
	int32_t nthread)
{
	cbct_ig ig_, *ig = &ig_;
	cbct_cg cg_, *cg = &cg_;
	const int nx = nxyz[0], ny = nxyz[1], nz = nxyz[2];
	const float dz_per_rad = (float) source_dz_per_rad;
	int ok;

	ig->mask2 = (byte *) calloc((size_t) nx * ny, 1);
	if (!ig->mask2) return 0;
	if (!cbct_mask_init(ig->mask2, (cbyte *) mask, nx, ny, nthread, 0)) {
		free(ig->mask2);
		return 0;
	}
	ig->nx = nx; ig->ny = ny; ig->nz = nz;
	ig->dx = dxyz[0]; ig->dy = dxyz[1]; ig->dz = dxyz[2];
	ig->offset_x = cxyz[0]; ig->offset_y = cxyz[1]; ig->offset_z = cxyz[2];
	cg->dso = dso; cg->dsd = dsd; cg->dfs = dfs;
	cg->ns = ns; cg->nt = nt;
	cg->ds = dst[0]; cg->dt = dst[1];
	cg->offset_s = ost[0]; cg->offset_t = ost[1];
	ok = fdk_ts_back_t(image, ig, cg, na, proj, beta, source_zs, na1_half,
		cone_par, w3d, dz_per_rad, pitch, nthread, 0);
	free(ig->mask2);
	return ok;
}
#endif

/* Mirrors the 'fdk,st,back' branch of fdk_mex_back (fdk,mex.c:93-100,139-156):
 * proj is [ns nt na] (s fastest), image is [nx ny nz] and is zeroed first
 * (mxCreateNumericArray returns zeros). Single threaded, as the reference. */
int ref_fdk_st_back(
	float *image,
	const int32_t *nxyz, const double *dxyz, const double *cxyz,
	const uint8_t *mask,
	double dso, double dsd, double dfs,
	const double *dst, const double *ost,
	const float *proj, int32_t ns, int32_t nt, int32_t na,
	const double *beta)
{
	const int nx = nxyz[0], ny = nxyz[1], nz = nxyz[2];
	byte *mask2 = (byte *) calloc((size_t) nx * ny, 1);
	if (!mask2) return 0;
	if (!cbct_mask_init(mask2, (cbyte *) mask, nx, ny, 1, 0)) {
		free(mask2);
		return 0;
	}
	for (long j = 0; j < (long) nx * ny * nz; ++j)
		image[j] = 0;
	for (int ia = 0; ia < na; ++ia)
		fdk_st_back1(image, nx, ny, nz,
			(float) dxyz[0], (float) dxyz[1], (float) dxyz[2],
			(float) cxyz[0], (float) cxyz[1], (float) cxyz[2],
			mask2, (float) dso, (float) dsd, (float) dfs, ns, nt,
			(float) dst[0], (float) dst[1], (float) ost[0], (float) ost[1],
			proj + (long) ia * nt * ns, (float) beta[ia]);
	free(mask2);
	return 1;
}
