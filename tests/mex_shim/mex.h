/*
 * mex.h -- test-only stand-in for MATLAB's MEX API, enough to compile and drive the
 * product's four MEX gateways (mb-rec-msot_amd/mex/) without MATLAB.
 *
 * Interleaved-complex API (-R2018a): mxGetDoubles / mxGetInt32s.  mxArrays are plain
 * host structs; mexErrMsgIdAndTxt long-jumps back to shim_call(), which reports the
 * error identifier.  Not a MATLAB substitute: only what the gateways use.
 */
#ifndef MBREC_TEST_MEX_H
#define MBREC_TEST_MEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;
typedef int32_t mxInt32;
typedef double mxDouble;
typedef enum { mxUNKNOWN_CLASS = 0, mxDOUBLE_CLASS = 6, mxSINGLE_CLASS = 7, mxINT32_CLASS = 12 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct mxArray_tag mxArray;

int mxIsDouble(const mxArray *a);
int mxIsComplex(const mxArray *a);
int mxIsInt32(const mxArray *a);
mxDouble *mxGetDoubles(const mxArray *a);
mxInt32 *mxGetInt32s(const mxArray *a);
double mxGetScalar(const mxArray *a);
size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
mwSize mxGetNumberOfDimensions(const mxArray *a);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity cplx);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity cplx);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexAtExit(void (*fn)(void));

/* the gateway's entry point (C linkage, as MATLAB's own mex.h declares it) */
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

/* test-side helpers */
typedef void (*mex_function_t)(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
mxArray *shim_array(int cls, int complex_flag, mwSize ndim, const mwSize *dims, const void *data);
void shim_destroy(mxArray *a);
void *shim_data(const mxArray *a);
int shim_call(mex_function_t fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs, char *errid,
              size_t errid_len, char *msg, size_t msg_len);
void shim_run_at_exit(void);

#ifdef __cplusplus
}
#endif
#endif
