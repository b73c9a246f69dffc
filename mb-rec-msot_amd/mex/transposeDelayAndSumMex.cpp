// transposeDelayAndSumMex gateway: img = transposeDelayAndSumMex(sigMat, travelTimes, R)
// Replaces reconstruction_model_based/models/transposeDelayAndSumMex.cpp:15-98 (MSOTmodel.m:250-252).
#include "mbrec_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs; (void)nrhs;
    mbrec_gateway_register_exit();
    mbrec_need_double("transposeDelayAndSumMex", prhs[0], 0);
    mbrec_need_double("transposeDelayAndSumMex", prhs[1], 1);
    mbrec_need_double("transposeDelayAndSumMex", prhs[2], 2);
    const int W = (int)mxGetM(prhs[0]);
    const int nT = (int)mxGetDimensions(prhs[0])[1];
    const int P = (int)mxGetM(prhs[1]), D = (int)mxGetN(prhs[1]);
    plhs[0] = mxCreateDoubleMatrix(W, P, mxREAL);
    if (mbrec_transposeDelayAndSumMex(mxGetDoubles(prhs[0]), W, nT, D, mxGetDoubles(prhs[1]), P,
                                      mxGetScalar(prhs[2]), mxGetDoubles(plhs[0])) != MBREC_OK)
        mbrec_gateway_fail();
}
