// GPU_clone.cpp -- MEX gateway: newtag = GPU_clone(tag)
// Call form of gpuclass/GPU_clone.cu:15-50: a new owner with a copy of the data.
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nlhs != 1 || nrhs != 1) imogen_mex_usage("GPU_clone", "newtag = GPU_clone(tag)");
    int64_t tag[GPU_TAG_MAXLEN], nu[GPU_TAG_MAXLEN];
    imogen_mex_fail(imogen_mex_tag(prhs[0], 0, tag), "GPU_clone");
    imogen_mex_fail(GPU_clone(tag, nu), "GPU_clone");
    plhs[0] = imogen_mex_tag_array(nu);
}
