// GPU_makeslab.cpp -- MEX gateway: newtag = GPU_makeslab(tag, N)
// Call form of gpuclass/GPU_makeslab.cu:15-72 (GPU_Type.createSlabs, GPU_Type.m:166-169).
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nlhs != 1 || nrhs != 2) imogen_mex_usage("GPU_makeslab", "newtag = GPU_makeslab(tag, number of slabs)");
    int64_t tag[GPU_TAG_MAXLEN], nu[GPU_TAG_MAXLEN];
    imogen_mex_fail(imogen_mex_tag(prhs[0], 0, tag), "GPU_makeslab");
    imogen_mex_fail(GPU_makeslab(tag, (int)mxGetScalar(prhs[1]), nu), "GPU_makeslab");
    plhs[0] = imogen_mex_tag_array(nu);
}
