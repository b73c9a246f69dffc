/* Test-only implementation of the mex.h stand-in (see mex.h). */
#include "mex.h"

#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct mxArray_tag {
    mxClassID cls;
    int cplx;
    mwSize ndim;
    mwSize dims[8];
    void *data;
};

static size_t esize(mxClassID c) { return c == mxINT32_CLASS ? 4 : c == mxSINGLE_CLASS ? 4 : 8; }

static size_t numel(const mxArray *a)
{
    size_t n = 1;
    for (mwSize i = 0; i < a->ndim; ++i) n *= a->dims[i];
    return n;
}

int mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsComplex(const mxArray *a) { return a->cplx; }
int mxIsInt32(const mxArray *a) { return a->cls == mxINT32_CLASS; }
mxDouble *mxGetDoubles(const mxArray *a) { return a->cls == mxDOUBLE_CLASS ? (mxDouble *)a->data : NULL; }
mxInt32 *mxGetInt32s(const mxArray *a) { return a->cls == mxINT32_CLASS ? (mxInt32 *)a->data : NULL; }

double mxGetScalar(const mxArray *a)
{
    if (numel(a) == 0) return 0.0;
    switch (a->cls) {
    case mxDOUBLE_CLASS: return ((double *)a->data)[0];
    case mxINT32_CLASS: return (double)((int32_t *)a->data)[0];
    case mxSINGLE_CLASS: return (double)((float *)a->data)[0];
    default: return 0.0;
    }
}

size_t mxGetM(const mxArray *a) { return a->ndim ? a->dims[0] : 0; }

size_t mxGetN(const mxArray *a)
{
    size_t n = 1;
    for (mwSize i = 1; i < a->ndim; ++i) n *= a->dims[i];
    return n;
}

const mwSize *mxGetDimensions(const mxArray *a) { return a->dims; }
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->ndim; }

mxArray *shim_array(int cls, int complex_flag, mwSize ndim, const mwSize *dims, const void *data)
{
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    a->cls = (mxClassID)cls;
    a->cplx = complex_flag;
    a->ndim = ndim < 2 ? 2 : ndim;
    a->dims[0] = a->dims[1] = 1;
    for (mwSize i = 0; i < ndim && i < 8; ++i) a->dims[i] = dims[i];
    const size_t n = numel(a);
    a->data = calloc(n ? n : 1, esize(a->cls));   /* zero-filled, as MATLAB's creators */
    if (data) memcpy(a->data, data, n * esize(a->cls));
    return a;
}

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity cplx)
{
    return shim_array(cls, cplx == mxCOMPLEX, ndim, dims, NULL);
}

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity cplx)
{
    const mwSize d[2] = {m, n};
    return shim_array(mxDOUBLE_CLASS, cplx == mxCOMPLEX, 2, d, NULL);
}

void shim_destroy(mxArray *a)
{
    if (!a) return;
    free(a->data);
    free(a);
}

void *shim_data(const mxArray *a) { return a->data; }

static jmp_buf *g_jmp;
static char g_id[256], g_msg[1024];

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    snprintf(g_id, sizeof g_id, "%s", id ? id : "");
    vsnprintf(g_msg, sizeof g_msg, fmt, ap);
    va_end(ap);
    if (g_jmp) longjmp(*g_jmp, 1);
    fprintf(stderr, "mexErrMsgIdAndTxt outside shim_call: %s %s\n", g_id, g_msg);
    abort();
}

static void (*g_exit_fns[16])(void);
static int g_nexit;

int mexAtExit(void (*fn)(void))
{
    if (g_nexit < 16) g_exit_fns[g_nexit++] = fn;
    return 0;
}

void shim_run_at_exit(void)
{
    while (g_nexit > 0) g_exit_fns[--g_nexit]();
}

int shim_call(mex_function_t fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs, char *errid,
              size_t errid_len, char *msg, size_t msg_len)
{
    jmp_buf env;
    jmp_buf *prev = g_jmp;
    g_id[0] = g_msg[0] = 0;
    if (setjmp(env)) {
        g_jmp = prev;
        if (errid && errid_len) snprintf(errid, errid_len, "%s", g_id);
        if (msg && msg_len) snprintf(msg, msg_len, "%s", g_msg);
        return 1;
    }
    g_jmp = &env;
    fn(nlhs, plhs, nrhs, prhs);
    g_jmp = prev;
    return 0;
}
