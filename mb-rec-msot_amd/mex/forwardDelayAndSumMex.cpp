// forwardDelayAndSumMex gateway: sig = forwardDelayAndSumMex(image, travelTimes, numTimesteps, R)
// Replaces reconstruction_model_based/models/forwardDelayAndSumMex.cpp:15-104 (MSOTmodel.m:214-217).
#include "mbrec_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs; (void)nrhs;
    mbrec_gateway_register_exit();
    mbrec_need_double("forwardDelayAndSumMex", prhs[0], 0);
    mbrec_need_double("forwardDelayAndSumMex", prhs[1], 1);
    mbrec_need_int32("forwardDelayAndSumMex", prhs[2], 2);
    mbrec_need_double("forwardDelayAndSumMex", prhs[3], 3);
    const int W = (int)mxGetM(prhs[0]), P = (int)mxGetN(prhs[0]);
    const int D = (int)mxGetN(prhs[1]);
    const int nT = (int)mxGetScalar(prhs[2]);
    const mwSize dims[3] = {(mwSize)W, (mwSize)nT, (mwSize)D};
    plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
    if (mbrec_forwardDelayAndSumMex(mxGetDoubles(prhs[0]), W, P, mxGetDoubles(prhs[1]), D, nT,
                                    mxGetScalar(prhs[3]), mxGetDoubles(plhs[0])) != MBREC_OK)
        mbrec_gateway_fail();
}
