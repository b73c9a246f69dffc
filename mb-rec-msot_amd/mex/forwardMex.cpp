// forwardMex gateway: sig = forwardMex(image, sirIndices, travelTimes, sir, numTimesteps,
//                                      sirTimestepsStartAndEnd, R)
// Replaces reconstruction_model_based/models/forwardMex.cpp:16-162 (same call at MSOTmodel.m:195-201).
#include "mbrec_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs; (void)nrhs;
    mbrec_gateway_register_exit();
    mbrec_need_double("forwardMex", prhs[0], 0);
    mbrec_need_int32("forwardMex", prhs[1], 1);
    mbrec_need_double("forwardMex", prhs[2], 2);
    mbrec_need_double("forwardMex", prhs[3], 3);
    mbrec_need_int32("forwardMex", prhs[4], 4);
    mbrec_need_int32("forwardMex", prhs[5], 5);
    mbrec_need_double("forwardMex", prhs[6], 6);
    const int W = (int)mxGetM(prhs[0]), P = (int)mxGetN(prhs[0]);
    const int D = (int)mxGetN(prhs[1]);
    const int Ls = (int)mxGetM(prhs[3]), Nsir = (int)mxGetN(prhs[3]);
    const int nT = (int)mxGetScalar(prhs[4]);
    const mwSize dims[3] = {(mwSize)W, (mwSize)nT, (mwSize)D};
    plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
    if (mbrec_forwardMex(mxGetDoubles(prhs[0]), W, P, mxGetInt32s(prhs[1]), mxGetDoubles(prhs[2]), D,
                         mxGetDoubles(prhs[3]), Ls, Nsir, nT, mxGetInt32s(prhs[5]), mxGetScalar(prhs[6]),
                         mxGetDoubles(plhs[0])) != MBREC_OK)
        mbrec_gateway_fail();
}
