/*
 * mex.h -- a stand-in for MATLAB's mex.h, for the repository's tests only.
 *
 * Declares the part of the MATLAB C API that mirt_amd/mex/fdk_mex.c uses,
 * with MATLAB's names and signatures (separate complex storage, 64-bit
 * sizes).  fake_mx.c in this directory implements it on plain C buffers so
 * that the gateway can be built and driven from Python without MATLAB.  A
 * MATLAB build of the gateway uses MATLAB's own mex.h instead.
 */
#ifndef MIRT_AMD_TEST_MEX_H
#define MIRT_AMD_TEST_MEX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;

typedef enum {
	mxUNKNOWN_CLASS = 0,
	mxCHAR_CLASS = 4,
	mxDOUBLE_CLASS = 6,
	mxSINGLE_CLASS = 7,
	mxINT8_CLASS = 8,
	mxUINT8_CLASS = 9,
	mxINT32_CLASS = 12,
	mxUINT32_CLASS = 13
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

int mxIsChar(const mxArray *a);
int mxIsDouble(const mxArray *a);
int mxIsSingle(const mxArray *a);
int mxIsInt32(const mxArray *a);
int mxIsUint8(const mxArray *a);
int mxIsComplex(const mxArray *a);
size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
void *mxGetData(const mxArray *a);
int mxGetString(const mxArray *a, char *buf, mwSize buflen);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls,
	mxComplexity cplx);
int mexPrintf(const char *fmt, ...);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

#ifdef __cplusplus
}
#endif

#endif
