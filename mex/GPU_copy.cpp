// GPU_copy.cpp -- MEX gateway: GPU_copy(destination, source)
// Call form of gpuclass/GPU_copy.cu:15-50: identical shapes, device to device.
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)plhs;
    if (nlhs != 0 || nrhs != 2) imogen_mex_usage("GPU_copy", "GPU_copy(destination, source)");
    int64_t dst[GPU_TAG_MAXLEN], src[GPU_TAG_MAXLEN];
    imogen_mex_fail(imogen_mex_tag(prhs[0], 0, dst), "GPU_copy");
    imogen_mex_fail(imogen_mex_tag(prhs[1], 0, src), "GPU_copy");
    imogen_mex_fail(GPU_copy(dst, src), "GPU_copy");
}
