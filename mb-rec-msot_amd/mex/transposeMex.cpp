// transposeMex gateway: img = transposeMex(sigMat, sirIndices, travelTimes, sir,
//                                          sirTimestepsStartAndEnd, R)
// Replaces reconstruction_model_based/models/transposeMex.cpp:16-158 (MSOTmodel.m:231-236).
#include "mbrec_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs; (void)nrhs;
    mbrec_gateway_register_exit();
    mbrec_need_double("transposeMex", prhs[0], 0);
    mbrec_need_int32("transposeMex", prhs[1], 1);
    mbrec_need_double("transposeMex", prhs[2], 2);
    mbrec_need_double("transposeMex", prhs[3], 3);
    mbrec_need_int32("transposeMex", prhs[4], 4);
    mbrec_need_double("transposeMex", prhs[5], 5);
    const int W = (int)mxGetM(prhs[0]);
    const int nT = (int)mxGetDimensions(prhs[0])[1];
    const int P = (int)mxGetM(prhs[1]), D = (int)mxGetN(prhs[1]);
    const int Ls = (int)mxGetM(prhs[3]), Nsir = (int)mxGetN(prhs[3]);
    plhs[0] = mxCreateDoubleMatrix(W, P, mxREAL);
    if (mbrec_transposeMex(mxGetDoubles(prhs[0]), W, nT, D, mxGetInt32s(prhs[1]), mxGetDoubles(prhs[2]), P,
                           mxGetDoubles(prhs[3]), Ls, Nsir, mxGetInt32s(prhs[4]), mxGetScalar(prhs[5]),
                           mxGetDoubles(plhs[0])) != MBREC_OK)
        mbrec_gateway_fail();
}
