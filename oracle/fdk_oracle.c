/*
 * fdk_oracle.c -- TEST INFRASTRUCTURE ONLY (the CPU checker).
 *
 * Clean-room CPU restatement of the reference FDK cone-beam voxel-driven
 * backprojector, used to check the HIP product path.  Nothing in the product
 * links, loads or calls this file; only tests/, the smoke check and the
 * CPU-baseline leg of the benchmark do.
 *
 * Functions and the reference code each one follows:
 *
 *   orc_back_ts_f32   fdk_ts_back1 in mex/src/fdk-helix/mex/fdk,ts.c:55-251 on
 *                     its axial branch (cone_par = w3d = 0, source_zs = 0),
 *                     same float operation order, no FMA contraction
 *                     (built with -ffp-contract=off), views visited in order
 *                     per voxel exactly as fdk,ts,t.c:74-100 does.
 *                     `shipped_weights` = 1 selects the vertical weights of
 *                     mex/src/fdk/fdk-orig/fdk,ts.c:169-175 (wu*p1 + wd*p2).
 *   orc_back_ts_f64   the same algorithm carried out entirely in double
 *                     precision (double beta, sin/cos/atan2/floor in double);
 *                     the error reference for every precision claim.
 *   orc_back_ts_helix_f32
 *                     fdk_ts_back_t / fdk_ts_back1 of mex/src/fdk-helix/mex
 *                     with every branch: per-view source heights, the view
 *                     window iz_min_beta/iz_max_beta (fdk,ts,t.c:76-85),
 *                     cone-parallel magnification and heights
 *                     (fdk,ts.c:115-158), the w3d weights of
 *                     fdk,w3d,pow,tab.c:9-56 with the table filled as
 *                     fdk,ts,t.c:136-141 fills it (its `dt` argument receives
 *                     wz).  Table entries past the reference's allocation
 *                     read 0 here (the reference reads past its calloc'ed
 *                     table there; glibc's zeroed slack gives 0 when the
 *                     table size is a multiple of 16 bytes).
 *   orc_back_st_f32   fdk_st_back1 in mex/src/fdk/fdk-orig/fdk,st.c:38-141
 *                     ([ns nt] view order, no zero padding, edge samples
 *                     skipped, vertical weights as in that file).
 *
 * Geometry narrowing follows the mex gateway (fdk,mex.c:119-137: every double
 * scalar is stored into a float struct field) and the kernel prologue
 * (fdk,ts.c:90-98: w* computed in double then narrowed, beta narrowed to
 * float before sinf/cosf).
 *
 * Pinning (tests/test_oracle.py, tests/test_oracle_helix.py): the f32
 * restatements are compared with the reference C compiled from its own
 * sources (oracle/_ref) on seeded flat, arc and parallel geometries (helical,
 * cone-parallel and w3d cases for the helix one) and on the golden fixtures
 * in tests/golden; the f64
 * restatement against closed forms (constant views on the parallel geometry,
 * linearity, masked columns, impulse footprints) and against the f32 one.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
	int32_t nx, ny, nz;
	double dx, dy, dz;
	double offset_x, offset_y, offset_z;
	double dso, dsd, dfs;
	int32_t ns, nt;
	double ds, dt;
	double offset_s, offset_t;
} orc_geom;

/* detector model, fdk,ts.c:100-108 */
enum { ORC_FLAT = 0, ORC_ARC = 1, ORC_PAR = 2 };

static int orc_kind(double dso, double dsd, double dfs)
{
	if (isinf(dsd) || isinf(dso)) return ORC_PAR;
	if (dfs == 0) return ORC_ARC;
	return ORC_FLAT; /* finite nonzero dfs: reference warns "junk", treats as flat */
}

/* ------------------------------------------------------------------ */
/* FP32 restatement                                                    */
/* ------------------------------------------------------------------ */

typedef struct {
	float dx, dy, dz, dso, dsd, ds, dt;
	float wx, wy, wz, ws, wt;
	int nx, ny, nz, ns, nt, kind;
} orc_f32;

static void orc_f32_init(orc_f32 *f, const orc_geom *g)
{
	/* narrowing as in fdk,mex.c:119-137 */
	const float ox = (float) g->offset_x, oy = (float) g->offset_y;
	const float oz = (float) g->offset_z;
	const float os = (float) g->offset_s, ot = (float) g->offset_t;
	f->nx = g->nx; f->ny = g->ny; f->nz = g->nz;
	f->ns = g->ns; f->nt = g->nt;
	f->dx = (float) g->dx; f->dy = (float) g->dy; f->dz = (float) g->dz;
	f->dso = (float) g->dso; f->dsd = (float) g->dsd;
	f->ds = (float) g->ds; f->dt = (float) g->dt;
	f->kind = orc_kind((float) g->dso, (float) g->dsd, (float) g->dfs);
	/* fdk,ts.c:90-94: double arithmetic, stored as float */
	f->wx = (float) ((g->nx - 1) / 2. + ox);
	f->wy = (float) ((g->ny - 1) / 2. + oy);
	f->wz = (float) ((g->nz - 1) / 2. + oz);
	f->ws = (float) ((g->ns - 1) / 2. + os);
	f->wt = (float) ((g->nt - 1) / 2. + ot);
}

/* One column, all views, fdk,ts.c:112-245 for a single (ix,iy).
 * col_out: [nz] accumulators (float), view_pad: scratch [(nt+2)(ns+2)]
 * Returns 0 on the dz/dt < 0 failure, 1 otherwise. */
static int orc_column_f32(const orc_f32 *f, int ix, int iy,
	const float *proj, int na, const float *sinb, const float *cosb,
	int shipped_weights, float *col_out, long long *n_upd)
{
	const int nt = f->nt, ns = f->ns, nz = f->nz;
	const float yy = f->dy * (iy - f->wy);
	const float xx = f->dx * (ix - f->wx);
	long long cnt = 0;

	for (int ia = 0; ia < na; ++ia) {
		const float *view = proj + (long) ia * ns * nt; /* [nt ns] */
		const float xbeta = xx * cosb[ia] + yy * sinb[ia];
		const float ybetas = f->dso - (-xx * sinb[ia] + yy * cosb[ia]);
		const float mag = (f->kind == ORC_PAR) ? 1.f : f->dsd / ybetas;
		float ss;
		if (f->kind == ORC_ARC)
			ss = (float) ((double) f->dsd * atan2((double) xbeta, (double) ybetas));
		else
			ss = mag * xbeta;
		const float ss_bin = ss / f->ds + f->ws;
		const int is = (int) floorf(ss_bin);
		if (is < -1 || is >= ns)
			continue;
		const float wr = ss_bin - is;
		const float wl = 1.f - wr;
		const float t_inc = mag * f->dz / f->dt;
		if (f->dz / f->dt < 0)
			return 0;
		int iz_min = (int) (floorf(f->wz - (f->wt + 1) / t_inc) + 1);
		int iz_max = (int) (ceilf(f->wz + (nt - f->wt) / t_inc) - 1 + 1);
		if (iz_min < 0) iz_min = 0;
		if (iz_max > nz) iz_max = nz;
		float t_bin = t_inc * (iz_min - f->wz) + f->wt;
		if (t_bin == -1) {
			++iz_min;
			t_bin += t_inc;
		}
		const float w2 = (f->kind == ORC_ARC)
			? (f->dsd * f->dsd) / (ybetas * ybetas + xbeta * xbeta)
			: mag * mag;
		/* padded view: padded (s', t') = (is+1.., it+1..); zero outside */
		for (int iz = iz_min; iz < iz_max; ++iz) {
			const int it = (int) floorf(t_bin);
			const float wu = t_bin - it;
			const float wd = 1.f - wu;
			float a[2][2]; /* a[ds][dt] = view at (is+ds, it+dt), 0 outside */
			for (int a_s = 0; a_s < 2; ++a_s)
				for (int a_t = 0; a_t < 2; ++a_t) {
					const int s = is + a_s, t = it + a_t;
					a[a_s][a_t] = (s >= 0 && s < ns && t >= 0 && t < nt)
						? view[t + (long) s * nt] : 0.f;
				}
			const float p1 = wl * a[0][0] + wr * a[1][0];
			const float p2 = wl * a[0][1] + wr * a[1][1];
			if (shipped_weights)
				col_out[iz] += w2 * (wu * p1 + wd * p2);
			else {
				const float p0 = wd * p1 + wu * p2;
				col_out[iz] += w2 * p0;
			}
			t_bin += t_inc;
			++cnt;
		}
	}
	if (n_upd) *n_upd += cnt;
	return 1;
}

/* Backproject na views [nt ns na] into columns.
 * cols == NULL: all nx*ny columns, image is [nz nx ny] (z fastest).
 * cols != NULL: ncol flat indices ix + iy*nx, image is [nz ncol].
 * Columns with mask == 0 (mask may be NULL = all ones) are left at zero.
 * Returns 1 on success, 0 on the reference's dz/dt < 0 failure. */
int orc_back_ts_f32(const orc_geom *g, const float *proj, int32_t na,
	const double *beta, const uint8_t *mask, int32_t shipped_weights,
	const int32_t *cols, int32_t ncol, float *image, long long *n_updates)
{
	orc_f32 f;
	orc_f32_init(&f, g);
	const long nxy = (long) g->nx * g->ny;
	const long nc = cols ? ncol : nxy;
	float *sinb = (float *) malloc(sizeof(float) * (na > 0 ? na : 1));
	float *cosb = (float *) malloc(sizeof(float) * (na > 0 ? na : 1));
	int ok = 1;
	long long total = 0;
	for (int ia = 0; ia < na; ++ia) {
		const float b = (float) beta[ia]; /* fdk,ts,t.c:94 passes beta[ia] to a float */
		sinb[ia] = sinf(b);
		cosb[ia] = cosf(b);
	}
	memset(image, 0, sizeof(float) * nc * g->nz);

	#pragma omp parallel for schedule(dynamic, 16) reduction(&& : ok) reduction(+ : total)
	for (long c = 0; c < nc; ++c) {
		const long j = cols ? cols[c] : c;
		if (mask && !mask[j]) continue;
		long long cnt = 0;
		const int r = orc_column_f32(&f, (int) (j % g->nx), (int) (j / g->nx),
			proj, na, sinb, cosb, shipped_weights, image + c * g->nz,
			&cnt);
		ok = ok && r;
		total += cnt;
	}
	free(sinb);
	free(cosb);
	if (n_updates) *n_updates = total;
	return ok;
}

/* ------------------------------------------------------------------ */
/* FP64 restatement (error reference)                                  */
/* ------------------------------------------------------------------ */

int orc_back_ts_f64(const orc_geom *g, const float *proj, int32_t na,
	const double *beta, const uint8_t *mask,
	const int32_t *cols, int32_t ncol, double *image)
{
	const int nx = g->nx, ny = g->ny, nz = g->nz, ns = g->ns, nt = g->nt;
	const long nxy = (long) nx * ny;
	const long nc = cols ? ncol : nxy;
	const int kind = orc_kind(g->dso, g->dsd, g->dfs);
	const double wx = (nx - 1) / 2. + g->offset_x;
	const double wy = (ny - 1) / 2. + g->offset_y;
	const double wz = (nz - 1) / 2. + g->offset_z;
	const double ws = (ns - 1) / 2. + g->offset_s;
	const double wt = (nt - 1) / 2. + g->offset_t;
	int ok = 1;

	memset(image, 0, sizeof(double) * nc * nz);
	if (g->dz / g->dt < 0) {
		/* the reference fails at the first in-FOV column; checked below */
	}

	#pragma omp parallel for schedule(dynamic, 16) reduction(&& : ok)
	for (long c = 0; c < nc; ++c) {
		const long j = cols ? cols[c] : c;
		if (mask && !mask[j]) continue;
		const int ix = (int) (j % nx), iy = (int) (j / nx);
		const double xx = g->dx * (ix - wx);
		const double yy = g->dy * (iy - wy);
		double *col = image + c * nz;
		for (int ia = 0; ia < na; ++ia) {
			const float *view = proj + (long) ia * ns * nt;
			const double sb = sin(beta[ia]), cb = cos(beta[ia]);
			const double xbeta = xx * cb + yy * sb;
			const double ybetas = g->dso - (-xx * sb + yy * cb);
			const double mag = (kind == ORC_PAR) ? 1. : g->dsd / ybetas;
			const double ss = (kind == ORC_ARC) ? g->dsd * atan2(xbeta, ybetas)
				: mag * xbeta;
			const double ss_bin = ss / g->ds + ws;
			const int is = (int) floor(ss_bin);
			if (is < -1 || is >= ns) continue;
			if (g->dz / g->dt < 0) { ok = 0; break; }
			const double wr = ss_bin - is, wl = 1. - wr;
			const double t_inc = mag * g->dz / g->dt;
			int iz_min = (int) (floor(wz - (wt + 1) / t_inc) + 1);
			int iz_max = (int) ceil(wz + (nt - wt) / t_inc);
			if (iz_min < 0) iz_min = 0;
			if (iz_max > nz) iz_max = nz;
			double t_bin = t_inc * (iz_min - wz) + wt;
			if (t_bin == -1) { ++iz_min; t_bin += t_inc; }
			const double w2 = (kind == ORC_ARC)
				? g->dsd * g->dsd / (ybetas * ybetas + xbeta * xbeta)
				: mag * mag;
			for (int iz = iz_min; iz < iz_max; ++iz) {
				const int it = (int) floor(t_bin);
				const double wu = t_bin - it, wd = 1. - wu;
				double a[2][2];
				for (int a_s = 0; a_s < 2; ++a_s)
					for (int a_t = 0; a_t < 2; ++a_t) {
						const int s = is + a_s, t = it + a_t;
						a[a_s][a_t] = (s >= 0 && s < ns && t >= 0 && t < nt)
							? (double) view[t + (long) s * nt] : 0.;
					}
				const double p1 = wl * a[0][0] + wr * a[1][0];
				const double p2 = wl * a[0][1] + wr * a[1][1];
				col[iz] += w2 * (wd * p1 + wu * p2);
				t_bin += t_inc;
			}
		}
	}
	return ok;
}

/* ------------------------------------------------------------------ */
/* helical / cone-parallel path of fdk-helix, FP32                     */
/* ------------------------------------------------------------------ */

typedef struct {
	int na1_half, cone_par, w3d, ntab, nt2;
	float pitch, zs_min, zs_max, wz;
	double dz_per_rad;       /* float-narrowed by the gateway (fdk,mex.c:128) */
	float *tab;              /* ntab entries, then a zero tail */
} orc_hx;

/* fdk_w3d_pow_tab (fdk,w3d,pow,tab.c:33-56) */
static float orc_w3d(const orc_hx *h, float t_bin, float t_bin_c, float wt, int iz)
{
	const int id = (int) (h->pitch ? 0.f : fabsf((float) iz - h->wz));
	const float half = (float) h->nt2;
	const float a = fabsf(t_bin - wt), a_c = fabsf(t_bin_c - wt);
	const int indx = (int) (floorf(a < half ? a : half) + ceilf(half) * id);
	const int indx_c = (int) (floorf(a_c < half ? a_c : half) + ceilf(half) * id);
	const float scale = h->tab[indx], scale_c = h->tab[indx_c];
	const float eps = 0.001f;
	float r = scale_c / (scale_c + scale);
	r = (r > eps) ? r : eps;
	r = (r < 1 - eps) ? r : 1 - eps;
	return 2 * r;
}

/* One column, all views: fdk,ts.c:112-245 (fdk-helix) for a single (ix,iy). */
static void orc_column_helix(const orc_f32 *f, const orc_hx *h, int ix, int iy,
	const float *proj, int na, const float *sinb, const float *cosb,
	const float *zs, const int *izb0, const int *izb1, float *col)
{
	const int nt = f->nt, ns = f->ns;
	const float yy = f->dy * (iy - f->wy);
	const float xx = f->dx * (ix - f->wx);
	const int is_par = f->kind == ORC_PAR, is_arc = f->kind == ORC_ARC;
	for (int ia = 0; ia < na; ++ia) {
		const float *view = proj + (long) ia * ns * nt; /* [nt ns] */
		const float xbeta = xx * cosb[ia] + yy * sinb[ia];
		const float ybetas = f->dso - (-xx * sinb[ia] + yy * cosb[ia]);
		const float mag = h->cone_par
			? (float) (f->dsd / (pow(f->dso * f->dso - xbeta * xbeta, 0.5)
				- (f->dso - ybetas)))
			: (is_par ? 1.f : f->dsd / ybetas);
		const float ss = h->cone_par ? xbeta
			: (is_arc ? (float) ((double) f->dsd * atan2((double) xbeta, (double) ybetas))
				: mag * xbeta);
		const float ss_bin = ss / f->ds + f->ws;
		const int is = (int) floorf(ss_bin);
		if (is < -1 || is >= ns)
			continue;
		const float wr = ss_bin - is;
		const float wl = 1.f - wr;
		const float t_inc = mag * f->dz / f->dt;
		const float wzb0 = f->wz + zs[ia] / f->dz;
		const float wzbeta = h->cone_par
			? (float) (wzb0 + asin(-xbeta / f->dso) * h->dz_per_rad / f->dz)
			: wzb0;
		int iz_min = (int) (floorf(wzbeta - (f->wt + 1) / t_inc) + 1);
		int iz_max = (int) (ceilf(wzbeta + (nt - f->wt) / t_inc) - 1 + 1);
		iz_min = iz_min > 0 ? iz_min : 0;
		iz_min = iz_min > izb0[ia] ? iz_min : izb0[ia];
		iz_max = iz_max < f->nz ? iz_max : f->nz;
		iz_max = iz_max < izb1[ia] ? iz_max : izb1[ia];
		float t_bin = t_inc * (iz_min - wzbeta) + f->wt;
		if (t_bin == -1) {
			++iz_min;
			t_bin += t_inc;
		}
		/* conjugate ray (fdk,ts.c:169-175) */
		const float xbeta_c = -xbeta;
		const float ybetas_c = 2 * f->dso - ybetas;
		const float mag_c = (float) (f->dsd / (pow(f->dso * f->dso - xbeta_c * xbeta_c, 0.5)
			- (f->dso - ybetas_c)));
		const float t_inc_c = mag_c * f->dz / f->dt;
		for (int iz = iz_min; iz < iz_max; ++iz) {
			const int it = (int) floorf(t_bin);
			const float wu = t_bin - it;
			const float wd = 1.f - wu;
			float a[2][2]; /* view at (is+ds, it+dt), 0 outside */
			for (int a_s = 0; a_s < 2; ++a_s)
				for (int a_t = 0; a_t < 2; ++a_t) {
					const int s = is + a_s, t = it + a_t;
					a[a_s][a_t] = (s >= 0 && s < ns && t >= 0 && t < nt)
						? view[t + (long) s * nt] : 0.f;
				}
			float p0 = 0.f;
			if (it >= -1 && it < nt) {
				const float p1 = wl * a[0][0] + wr * a[1][0];
				const float p2 = wl * a[0][1] + wr * a[1][1];
				p0 = wd * p1 + wu * p2;
			}
			float ww = 1.f;
			if (h->cone_par && h->w3d) {
				const float zs_c = (float) (h->pitch
					? ((t_bin >= f->wt) ? (zs[ia] + M_PI * h->dz_per_rad)
						: (zs[ia] - M_PI * h->dz_per_rad))
					: 0);
				const float wzbeta_c = (float) (f->wz + zs_c / f->dz
					+ asin(-xbeta_c / f->dso) * h->dz_per_rad / f->dz);
				const float t_bin_c = t_inc_c * (iz - wzbeta_c) + f->wt;
				if (zs_c >= h->zs_min && zs_c <= h->zs_max)
					ww = orc_w3d(h, t_bin, t_bin_c, f->wt, iz);
			}
			const float w2 = h->cone_par ? ww : (is_arc
				? (f->dsd * f->dsd) / (ybetas * ybetas + xbeta * xbeta)
				: mag * mag);
			col[iz] += w2 * p0;
			t_bin += t_inc;
		}
	}
}

/* fdk_ts_back_t (fdk,ts,t.c:107-167, fdk-helix) with all of its arguments:
 * proj [nt ns na], beta and source_zs [na] (double), image [nz nx ny].
 * source_dz_per_rad is narrowed to float as the gateway does.  Returns 0 on
 * the reference's dz/dt < 0 failure (checked up front here), 1 otherwise. */
int orc_back_ts_helix_f32(const orc_geom *g, const float *proj, int32_t na,
	const double *beta, const double *source_zs, int32_t na1_half, int32_t cone_par,
	int32_t w3d, double source_dz_per_rad, float pitch, const uint8_t *mask,
	float *image)
{
	orc_f32 f;
	orc_f32_init(&f, g);
	const long nxy = (long) g->nx * g->ny;
	memset(image, 0, sizeof(float) * nxy * g->nz);
	if (na <= 0)
		return 1;
	if (f.dz / f.dt < 0)
		return 0;
	orc_hx h;
	h.na1_half = na1_half;
	h.cone_par = cone_par;
	h.w3d = w3d;
	h.pitch = pitch;
	h.dz_per_rad = (float) source_dz_per_rad;
	h.zs_min = (float) source_zs[0];
	h.zs_max = (float) source_zs[na - 1];
	h.wz = (float) ((g->nz - 1) / 2. + (float) g->offset_z);
	h.nt2 = g->nt / 2;
	const int rows = (int) ceilf((float) (g->nz / 2));
	h.ntab = pitch ? (int) ceilf((float) h.nt2) : (int) (ceilf((float) h.nt2) * rows);
	/* largest index orc_w3d can form: nt/2 + (nt/2) * trunc(max |iz - wz|) */
	const float span = fabsf(h.wz) > fabsf((float) (g->nz - 1) - h.wz)
		? fabsf(h.wz) : fabsf((float) (g->nz - 1) - h.wz);
	const long tail = (long) h.nt2 * ((long) span + 2) + 2;
	h.tab = (float *) calloc((size_t) (h.ntab + tail), sizeof(float));
	/* fdk_w3d_pow_tab_init (fdk,w3d,pow,tab.c:9-31) called with dt := wz */
	const float dtw = h.wz, dsd = f.dsd;
	if (pitch) {
		for (int it = 0; it < ceilf((float) h.nt2); it++)
			h.tab[it] = (float) pow(it * dtw / dsd, w3d * pitch);
	} else {
		float *t = h.tab;
		for (int iz = 0; iz < rows; iz++)
			for (int it = 0; it < ceilf((float) h.nt2); it++, t++)
				*t = (float) pow(it * dtw / dsd, w3d * iz * f.dz);
	}
	float *sinb = (float *) malloc(sizeof(float) * na);
	float *cosb = (float *) malloc(sizeof(float) * na);
	float *zs = (float *) malloc(sizeof(float) * na);
	int *izb0 = (int *) malloc(sizeof(int) * na);
	int *izb1 = (int *) malloc(sizeof(int) * na);
	for (int ia = 0; ia < na; ++ia) {
		const float b = (float) beta[ia];
		sinb[ia] = sinf(b);
		cosb[ia] = cosf(b);
		zs[ia] = (float) source_zs[ia];
		izb0[ia] = 0;
		izb1[ia] = g->nz;
		if (pitch) { /* fdk,ts,t.c:80-85 */
			const int a0 = ia - na1_half > 0 ? ia - na1_half : 0;
			const int a1 = ia + na1_half < na - 1 ? ia + na1_half : na - 1;
			const float c0 = ceilf((float) (h.wz + source_zs[a0] / f.dz));
			const float c1 = ceilf((float) (h.wz + source_zs[a1] / f.dz));
			izb0[ia] = (int) (c0 > 0 ? c0 : 0);
			izb1[ia] = (int) (c1 < g->nz ? c1 : g->nz);
		}
	}
	#pragma omp parallel for schedule(dynamic, 16)
	for (long j = 0; j < nxy; ++j) {
		if (mask && !mask[j]) continue;
		orc_column_helix(&f, &h, (int) (j % g->nx), (int) (j / g->nx), proj, na,
			sinb, cosb, zs, izb0, izb1, image + j * g->nz);
	}
	free(sinb); free(cosb); free(zs); free(izb0); free(izb1); free(h.tab);
	return 1;
}

/* ------------------------------------------------------------------ */
/* 'fdk,st,back' order (use_mex = 3), FP32                             */
/* ------------------------------------------------------------------ */

/* proj [ns nt na] (s fastest); image [nx ny nz] (x fastest), zeroed here. */
int orc_back_st_f32(const orc_geom *g, const float *proj, int32_t na,
	const double *beta, const uint8_t *mask, float *image)
{
	orc_f32 f;
	orc_f32_init(&f, g);
	const int nx = g->nx, ny = g->ny, nz = g->nz, ns = g->ns, nt = g->nt;
	const long nxy = (long) nx * ny;
	memset(image, 0, sizeof(float) * nxy * nz);

	#pragma omp parallel for schedule(dynamic, 16)
	for (long j = 0; j < nxy; ++j) {
		if (mask && !mask[j]) continue;
		const int ix = (int) (j % nx), iy = (int) (j / nx);
		const float yy = f.dy * (iy - f.wy);
		const float xx = f.dx * (ix - f.wx);
		for (int ia = 0; ia < na; ++ia) {
			const float *view = proj + (long) ia * ns * nt; /* [ns nt] */
			const float b = (float) beta[ia];
			const float sb = sinf(b), cb = cosf(b);
			const float xbeta = xx * cb + yy * sb;
			const float ybetas = f.dso - (-xx * sb + yy * cb);
			const float mag = (f.kind == ORC_PAR) ? 1.f : f.dsd / ybetas;
			float ss;
			if (f.kind == ORC_ARC)
				ss = (float) ((double) f.dsd * atan2((double) xbeta, (double) ybetas));
			else
				ss = mag * xbeta;
			const float ss_bin = ss / f.ds + f.ws;
			const int is = (int) floorf(ss_bin);
			if (is < 0 || is >= ns - 1) continue;
			const float wr = ss_bin - is, wl = 1.f - wr;
			const float w2 = (f.kind == ORC_ARC)
				? (f.dsd * f.dsd) / (ybetas * ybetas + xbeta * xbeta)
				: mag * mag;
			for (int iz = 0; iz < nz; ++iz) {
				const float zz = f.dz * (iz - f.wz);
				const float tt = mag * zz;
				const float tt_bin = tt / f.dt + f.wt;
				const int it = (int) floorf(tt_bin);
				if (it < 0 || it >= nt - 1) continue;
				const float wu = tt_bin - it, wd = 1.f - wu;
				const float p1 = wl * view[is + it * ns] + wr * view[is + 1 + it * ns];
				const float p2 = wl * view[is + (it + 1) * ns]
					+ wr * view[is + 1 + (it + 1) * ns];
				image[j + iz * nxy] += w2 * (wu * p1 + wd * p2);
			}
		}
	}
	return 1;
}
