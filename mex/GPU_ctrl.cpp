// GPU_ctrl.cpp -- MEX gateway: GPU_ctrl('memory' | 'info' | 'reset' | 'peers' | 'createStreams' | 'destroyStreams' [, device])
// Call forms of gpuclass/GPU_ctrl.cu (GPUManager.isValidDeviceList,
// GPUManager.m:99-103; save/uploadDataArrays.m:49-54). One GPU per process:
// 'peers' reports no peer access, the stream calls have nothing to do (all
// work runs on the library's stream), 'reset' releases the library's cached
// temporaries.
#include <cstring>

#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs < 1 || !mxIsChar(prhs[0])) imogen_mex_usage("GPU_ctrl", "GPU_ctrl('memory' | 'info' | 'reset' | 'peers' | 'createStreams' | 'destroyStreams' [, device])");
    char what[32];
    mxGetString(prhs[0], what, sizeof(what));
    const int device = (nrhs >= 2) ? (int)mxGetScalar(prhs[1]) : 0;
    if (strcmp(what, "memory") == 0) {
        size_t fr = 0, tot = 0;
        imogen_mex_fail(GPU_ctrl_memory(device, &fr, &tot), "GPU_ctrl");
        if (nlhs > 0) {
            plhs[0] = mxCreateDoubleMatrix(1, 2, mxREAL);
            mxGetPr(plhs[0])[0] = (double)fr;
            mxGetPr(plhs[0])[1] = (double)tot;
        }
    } else if (strcmp(what, "info") == 0) {
        char text[512];
        imogen_mex_fail(GPU_ctrl_info(device, text, (int)sizeof(text)), "GPU_ctrl");
        if (nlhs > 0) plhs[0] = mxCreateString(text);
        else mexPrintf("%s\n", text);
    } else if (strcmp(what, "reset") == 0) {
        imogen_mex_fail(GPU_ctrl_reset(), "GPU_ctrl");
    } else if (strcmp(what, "peers") == 0) {
        if (nlhs > 0) plhs[0] = mxCreateDoubleMatrix(1, 1, mxREAL);
    } else if (strcmp(what, "createStreams") == 0 || strcmp(what, "destroyStreams") == 0) {
        if (nlhs > 0) plhs[0] = mxCreateNumericMatrix(0, 1, mxINT64_CLASS, mxREAL);
    } else {
        imogen_mex_fail(ERROR_INVALID_ARGS, "GPU_ctrl (unknown request)");
    }
}
