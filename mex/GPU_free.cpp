// GPU_free.cpp -- MEX gateway: GPU_free(tag, ...)
// Call form of gpuclass/GPU_free.cu:15-43 (GPU_Type.clearArray, GPU_Type.m:152-158):
// owners are freed, slab views ignored.
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)plhs;
    if (nlhs != 0 || nrhs < 1) imogen_mex_usage("GPU_free", "GPU_free(tag [, tag ...])");
    for (int k = 0; k < nrhs; k++) {
        int64_t tag[GPU_TAG_MAXLEN];
        imogen_mex_fail(imogen_mex_tag(prhs[k], 0, tag), "GPU_free");
        imogen_mex_fail(GPU_free(tag), "GPU_free");
    }
}
