/*
 * fdk_mex.c -- MATLAB mex gateway of the MI355X cone-beam backprojector.
 *
 *   img = fdk_mex('fdk,ts,back', nxyz, dxyz, cxyz, mask, dso, dsd, dfs, dst, ost,
 *                 proj, beta, nthread)
 *   img = fdk_mex('fdk,st,back', ...the same operands...)
 *
 * The drop-in for the reference gateway fdk_mex (mex/src/fdk/fdk-orig/fdk,mex.c),
 * which fbp/cbct_back.m reaches through fdk_mex_call (cbct_back.m:274-283):
 * the same command strings, operands, checks and messages
 * (fdk,mex.c:39-91 operand checks, :181-209 command and count checks), and
 * the same results -- 'fdk,ts,back' gives the [nz nx ny] single image of
 * fdk_ts_back_t, 'fdk,st,back' the [nx ny nz] image of fdk_st_back1 -- with
 * the backprojection run by libfdkhip on the GPU (fdkhip_fdk_ts_back,
 * fdkhip_fdk_st_back in include/fdkhip.h).  nthread is checked as the
 * reference checks it (a scalar int32, at most 250) and otherwise unused.
 *
 * On a failed check the reference prints "FAIL <file> <line>: <message>" and
 * raises the MATLAB error "fdk_mex"; this gateway raises the error
 * "fdk_mex: <message>" (identifier fdkhip:fdk_mex) with the same message,
 * the library's own message (fdkhip_last_error) for failures past the
 * checks.
 *
 * Build (MATLAB):
 *   mex -I<repo>/include fdk_mex.c -L<repo>/mirt_amd/lib -lfdkhip
 * The repository's CPU tests build it against a small stand-in of the mx
 * API (tests/mex_stub) and call it with the reference's error cases.
 */
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "mex.h"
#include "fdkhip.h"

#define USAGE "usage error. see above"

static void fdk_mex_help(void)
{
	mexPrintf("Usage for FDK mex routines for\n"
		"cone-beam weighted 3d backprojection of 2d projection view(s).\n"
		"  img = fdk_mex('fdk,ts,back', nxyz, dxyz, cxyz, mask, dso, dsd, dfs,\n"
		"                dst, ost, proj, beta, nthread)\n"
		"    proj [nt ns] or [nt ns na] single, img [nz nx ny] single\n"
		"  img = fdk_mex('fdk,st,back', ... same operands ...)\n"
		"    proj [ns nt] or [ns nt na] single, img [nx ny nz] single\n"
		"  nxyz int32[3], dxyz cxyz double[3], mask uint8 [nx ny],\n"
		"  dso dsd dfs double scalars (dfs 0: arc, inf: flat), dst ost double[2],\n"
		"  beta double [na] (radians), nthread int32 scalar\n");
}

/* message buffer of the failed check */
static char g_msg[512];

static int failf(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_msg, sizeof g_msg, fmt, ap);
	va_end(ap);
	return 0;
}

/* the operand predicates of the reference's def,mexarg.h */
static int is_real_double(const mxArray *a)
{
	return mxIsDouble(a) && !mxIsComplex(a);
}

static int is_real_single(const mxArray *a)
{
	return mxIsSingle(a) && !mxIsComplex(a);
}

static int is_scalar(const mxArray *a)
{
	return mxGetNumberOfDimensions(a) == 2 && mxGetM(a) == 1 && mxGetN(a) == 1;
}

static int is_scalar_int32(const mxArray *a)
{
	return is_scalar(a) && mxIsInt32(a) && !mxIsComplex(a);
}

static int is_scalar_double(const mxArray *a)
{
	return is_scalar(a) && is_real_double(a);
}

/* element count as the reference counts it (rows times columns) */
static size_t count_mn(const mxArray *a)
{
	return mxGetM(a) * mxGetN(a);
}

/* operands in the order of the reference's fdk_mex_back (fdk,mex.c:22-36) */
enum { A_NXYZ, A_DXYZ, A_CXYZ, A_MASK, A_DSO, A_DSD, A_DFS, A_DST, A_OST, A_PROJ,
	A_BETA, A_NTHREAD, N_OPERANDS };

/* fdk_mex_back (fdk,mex.c:22-171): checks, the output array, the library call */
static int fdk_mex_back(mxArray *plhs[], const mxArray *const op[], int is_st)
{
	if (!mxIsInt32(op[A_NXYZ]))
		return failf("mxIsInt32");
	if (count_mn(op[A_NXYZ]) != 3)
		return failf("mx_nxyz must have 3 elements");
	const int32_t *nxyz = (const int32_t *) mxGetData(op[A_NXYZ]);
	const int nx = nxyz[0], ny = nxyz[1], nz = nxyz[2];

	if (!is_real_double(op[A_DXYZ]))
		return failf("mxIsRealDouble");
	if (count_mn(op[A_DXYZ]) != 3)
		return failf("mx_dxyz must have 3 elements");
	if (!is_real_double(op[A_CXYZ]))
		return failf("mxIsRealDouble");
	if (count_mn(op[A_CXYZ]) != 3)
		return failf("mx_cxyz must have 3 elements");

	if (!mxIsUint8(op[A_MASK]))
		return failf("mxIsUint8");
	if (mxGetNumberOfDimensions(op[A_MASK]) != 2)
		return failf("mask must be 2d");
	if (nx != (int) mxGetM(op[A_MASK]) || ny != (int) mxGetN(op[A_MASK]))
		return failf("mask size not nx by ny?");

	if (!is_scalar_double(op[A_DSO]) || !is_scalar_double(op[A_DSD])
		|| !is_scalar_double(op[A_DFS]))
		return failf("mxIsScalarDouble");

	if (!is_real_double(op[A_DST]))
		return failf("mxIsRealDouble");
	if (count_mn(op[A_DST]) != 2)
		return failf("mx_dst must have 2 elements");
	if (!is_real_double(op[A_OST]))
		return failf("mxIsRealDouble");
	if (count_mn(op[A_OST]) != 2)
		return failf("mx_ost must have 2 elements");

	if (!is_real_single(op[A_PROJ]))
		return failf("mxIsRealSingle");
	const mwSize pdim = mxGetNumberOfDimensions(op[A_PROJ]);
	const mwSize *pd = mxGetDimensions(op[A_PROJ]);
	int na;
	if (pdim == 2) {
		na = 1;
		if (!is_scalar_double(op[A_BETA]))
			return failf("mxIsScalarDouble");
	} else if (pdim == 3) {
		na = (int) pd[2];
		if (!is_real_double(op[A_BETA]))
			return failf("mxIsRealDouble");
		const int na_beta = (int) count_mn(op[A_BETA]);
		if (na != na_beta)
			return failf("beta length=%d vs na=size(proj,3)=%d", na_beta, na);
	} else {
		return failf("mx_proj must be 2d or 3d");
	}

	if (!is_scalar_int32(op[A_NTHREAD]))
		return failf("mxIsScalarInt32");
	const int32_t nthread = *(const int32_t *) mxGetData(op[A_NTHREAD]);

	/* 'fdk,st,back': [ns nt] views, [nx ny nz] image; 'fdk,ts,back': [nt ns]
	   views, [nz nx ny] image (the reference's "zxy" order) */
	const int32_t ns = (int32_t) (is_st ? pd[0] : pd[1]);
	const int32_t nt = (int32_t) (is_st ? pd[1] : pd[0]);
	mwSize odim[3];
	if (is_st) {
		odim[0] = (mwSize) nx;
		odim[1] = (mwSize) ny;
		odim[2] = (mwSize) nz;
	} else {
		odim[0] = (mwSize) nz;
		odim[1] = (mwSize) nx;
		odim[2] = (mwSize) ny;
	}
	plhs[0] = mxCreateNumericArray(3, odim, mxSINGLE_CLASS, mxREAL);
	if (!plhs[0])
		return failf("mxCreateNumericArray");

	const double *dst = (const double *) mxGetData(op[A_DST]);
	const double *ost = (const double *) mxGetData(op[A_OST]);
	const double dso = *(const double *) mxGetData(op[A_DSO]);
	const double dsd = *(const double *) mxGetData(op[A_DSD]);
	const double dfs = *(const double *) mxGetData(op[A_DFS]);
	const int rc = (is_st ? fdkhip_fdk_st_back : fdkhip_fdk_ts_back)(
		nxyz, (const double *) mxGetData(op[A_DXYZ]), (const double *) mxGetData(op[A_CXYZ]),
		(const uint8_t *) mxGetData(op[A_MASK]), dso, dsd, dfs, dst, ost,
		(const float *) mxGetData(op[A_PROJ]), ns, nt, na,
		(const double *) mxGetData(op[A_BETA]), nthread, (float *) mxGetData(plhs[0]));
	if (rc != FDKHIP_OK)
		return failf("%s", fdkhip_last_error());
	return 1;
}

/* fdk_mex (fdk,mex.c:175-213): the command and operand count */
static int fdk_mex(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
	if (nrhs <= 1 || !mxIsChar(prhs[0])) {   /* expect 'fdk,ts,back', ... */
		fdk_mex_help();
		return failf(USAGE);
	}
	char cmd[64];
	if (mxGetString(prhs[0], cmd, sizeof cmd) != 0)
		cmd[sizeof cmd - 1] = '\0';           /* longer than any command */
	if (!strcmp(cmd, "fdk,st,back") || !strcmp(cmd, "fdk,ts,back")) {
		const int is_st = !strncmp(cmd, "fdk,st", 6);
		if (nrhs != 1 + N_OPERANDS || nlhs != 1)
			return failf(USAGE);
		return fdk_mex_back(plhs, prhs + 1, is_st);
	}
	return failf("command '%s' unknown", cmd);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
	if (!nlhs && !nrhs) {
		fdk_mex_help();
		return;
	}
	if (!fdk_mex(nlhs, plhs, nrhs, prhs))
		mexErrMsgIdAndTxt("fdkhip:fdk_mex", "fdk_mex: %s", g_msg);
}
