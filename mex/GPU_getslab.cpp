// GPU_getslab.cpp -- MEX gateway: view = GPU_getslab(tag, k)
// Call form of gpuclass/GPU_getslab.cu:15-47: a view of slab k (numSlabs = -k).
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nlhs != 1 || nrhs != 2) imogen_mex_usage("GPU_getslab", "view = GPU_getslab(tag, slab index)");
    int64_t tag[GPU_TAG_MAXLEN], view[GPU_TAG_MAXLEN];
    imogen_mex_fail(imogen_mex_tag(prhs[0], 0, tag), "GPU_getslab");
    imogen_mex_fail(GPU_getslab(tag, (int)mxGetScalar(prhs[1]), view), "GPU_getslab");
    plhs[0] = imogen_mex_tag_array(view);
}
