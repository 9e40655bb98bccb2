// GPU_download.cpp -- MEX gateway: host = GPU_download(tag)
// Call form of gpuclass/GPU_download.cu:20-84 (GPU_Type.get.array, GPU_Type.m:68-73).
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nlhs != 1 || nrhs < 1 || nrhs > 2) imogen_mex_usage("GPU_download", "host = GPU_download(tag [, partition])");
    int64_t tag[GPU_TAG_MAXLEN];
    MGArray m;
    imogen_mex_fail(imogen_mex_tag(prhs[0], 0, tag), "GPU_download");
    imogen_mex_fail(deserializeTagToMGArray(tag, &m), "GPU_download");
    const mwSize dims[3] = {(mwSize)m.dim[0], (mwSize)m.dim[1], (mwSize)m.dim[2]};
    plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
    imogen_mex_fail(GPU_download(tag, mxGetPr(plhs[0])), "GPU_download");
}
