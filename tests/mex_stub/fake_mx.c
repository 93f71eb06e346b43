/*
 * fake_mx.c -- TEST INFRASTRUCTURE: a small implementation of the mx API
 * declared in mex.h (this directory), so that the mex gateway
 * mirt_amd/mex/fdk_mex.c can be built into a shared library and driven from
 * Python (tests/test_mex_gateway.py) without MATLAB.
 *
 * Arrays wrap caller-owned buffers (fmx_wrap) or own what mxCreateNumericArray
 * allocates; an error raised with mexErrMsgIdAndTxt returns to fmx_call with
 * the message, as MATLAB returns to its prompt.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
	mxClassID cls;
	mxComplexity cplx;
	mwSize ndim;
	mwSize dims[8];
	void *data;
	int owned;
};

static jmp_buf g_jump;
static char g_err[1024];
static char g_out[4096];
static size_t g_out_len;

static size_t class_bytes(mxClassID c)
{
	switch (c) {
	case mxDOUBLE_CLASS: return 8;
	case mxSINGLE_CLASS: case mxINT32_CLASS: case mxUINT32_CLASS: return 4;
	case mxCHAR_CLASS: return 2;
	default: return 1;
	}
}

/* an array over a caller-owned buffer (MATLAB order: first index fastest);
   trailing singleton dimensions are dropped down to 2, as MATLAB does */
mxArray *fmx_wrap(int cls, int cplx, int ndim, const size_t *dims, void *data)
{
	mxArray *a = (mxArray *) calloc(1, sizeof *a);
	if (!a || ndim < 0 || ndim > 8)
		return NULL;
	a->cls = (mxClassID) cls;
	a->cplx = cplx ? mxCOMPLEX : mxREAL;
	a->ndim = (mwSize) (ndim < 2 ? 2 : ndim);
	for (int i = 0; i < 8; ++i)
		a->dims[i] = i < ndim ? dims[i] : 1;
	while (a->ndim > 2 && a->dims[a->ndim - 1] == 1)
		--a->ndim;
	a->data = data;
	return a;
}

/* a 1 x n char array holding s */
mxArray *fmx_string(const char *s)
{
	const size_t n = strlen(s);
	const size_t dims[2] = {1, n};
	char *copy = (char *) malloc(n + 1);
	memcpy(copy, s, n + 1);
	mxArray *a = fmx_wrap(mxCHAR_CLASS, 0, 2, dims, copy);
	a->owned = 1;
	return a;
}

void fmx_free(mxArray *a)
{
	if (!a)
		return;
	if (a->owned)
		free(a->data);
	free(a);
}

/* Call the gateway.  Returns 0 on success (outputs in plhs), 1 when it
   raised an error (message in err). */
int fmx_call(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs, char *err, int errlen)
{
	g_err[0] = '\0';
	if (setjmp(g_jump)) {
		snprintf(err, (size_t) errlen, "%s", g_err);
		return 1;
	}
	mexFunction(nlhs, plhs, nrhs, prhs);
	if (errlen > 0)
		err[0] = '\0';
	return 0;
}

/* what the gateway printed since the last call of this function */
const char *fmx_printed(void)
{
	static char copy[sizeof g_out];
	memcpy(copy, g_out, g_out_len);
	copy[g_out_len] = '\0';
	g_out_len = 0;
	return copy;
}

void *fmx_data(const mxArray *a) { return a->data; }
int fmx_ndim(const mxArray *a) { return (int) a->ndim; }
size_t fmx_dim(const mxArray *a, int i) { return a->dims[i]; }
int fmx_class(const mxArray *a) { return (int) a->cls; }

/* ---- the mx API of mex.h ---- */

int mxIsChar(const mxArray *a) { return a->cls == mxCHAR_CLASS; }
int mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsSingle(const mxArray *a) { return a->cls == mxSINGLE_CLASS; }
int mxIsInt32(const mxArray *a) { return a->cls == mxINT32_CLASS; }
int mxIsUint8(const mxArray *a) { return a->cls == mxUINT8_CLASS; }
int mxIsComplex(const mxArray *a) { return a->cplx == mxCOMPLEX; }
size_t mxGetM(const mxArray *a) { return a->dims[0]; }

size_t mxGetN(const mxArray *a)
{
	size_t n = 1;
	for (mwSize i = 1; i < a->ndim; ++i)
		n *= a->dims[i];
	return n;
}

mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->ndim; }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims; }
void *mxGetData(const mxArray *a) { return a->data; }

int mxGetString(const mxArray *a, char *buf, mwSize buflen)
{
	if (!mxIsChar(a) || buflen == 0)
		return 1;
	const size_t n = mxGetM(a) * mxGetN(a);
	const size_t k = n < buflen - 1 ? n : buflen - 1;
	memcpy(buf, a->data, k);
	buf[k] = '\0';
	return k < n;
}

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls,
	mxComplexity cplx)
{
	size_t n = 1;
	for (mwSize i = 0; i < ndim; ++i)
		n *= dims[i];
	void *data = calloc(n ? n : 1, class_bytes(cls));
	mxArray *a = fmx_wrap((int) cls, (int) cplx, (int) ndim, dims, data);
	a->owned = 1;
	return a;
}

int mexPrintf(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	const int n = vsnprintf(g_out + g_out_len, sizeof g_out - g_out_len, fmt, ap);
	va_end(ap);
	if (n > 0)
		g_out_len += (size_t) n < sizeof g_out - g_out_len ? (size_t) n
			: sizeof g_out - g_out_len - 1;
	return n;
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
	(void) id;
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof g_err, fmt, ap);
	va_end(ap);
	longjmp(g_jump, 1);
}
