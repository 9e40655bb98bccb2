// GPU_setslab.cpp -- MEX gateway: view = GPU_setslab(tag, k, host array | GPU array)
// Call form of gpuclass/GPU_setslab.cu:15-81 (save/uploadDataArrays.m): slab k
// is filled from host data, or copied from another GPU array of its size.
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nlhs != 1 || nrhs != 3) imogen_mex_usage("GPU_setslab", "view = GPU_setslab(tag, slab index, host array | gpu array)");
    int64_t tag[GPU_TAG_MAXLEN], view[GPU_TAG_MAXLEN];
    imogen_mex_fail(imogen_mex_tag(prhs[0], 0, tag), "GPU_setslab");
    const int k = (int)mxGetScalar(prhs[1]);
    if (mxIsDouble(prhs[2])) {
        MGArray m;
        imogen_mex_fail(deserializeTagToMGArray(tag, &m), "GPU_setslab");
        if ((int64_t)mxGetNumberOfElements(prhs[2]) != m.numel)
            imogen_mex_fail(ERROR_INVALID_ARGS, "GPU_setslab (host array size differs from the slab)");
        imogen_mex_fail(GPU_setslab(tag, k, mxGetPr(prhs[2]), view), "GPU_setslab");
    } else {
        int64_t src[GPU_TAG_MAXLEN];
        imogen_mex_fail(imogen_mex_tag(prhs[2], 0, src), "GPU_setslab");
        imogen_mex_fail(GPU_getslab(tag, k, view), "GPU_setslab");
        imogen_mex_fail(GPU_copy(view, src), "GPU_setslab");
    }
    plhs[0] = imogen_mex_tag_array(view);
}
