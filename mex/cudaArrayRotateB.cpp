// cudaArrayRotateB.cpp -- MEX gateway: cudaArrayRotateB(array, code) or new = cudaArrayRotateB(array, code)
// Call form of gpuclass/cudaArrayRotateB.cu:47-90 (ImogenArray.arrayIndexExchange,
// the flux.m path): code 2 = X<->Y, 3 = X<->Z, 4 = Y<->Z, 5 = XYZ -> YZX,
// 6 = XYZ -> ZXY. In place, the tag is rewritten where it lives: an int64 tag
// argument directly (as the reference does), a GPU_Type / ImogenArray through
// its tag property.
#include <cstdlib>
#include <cstring>

#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs != 2 || nlhs > 1) imogen_mex_usage("cudaArrayRotateB", "cudaArrayRotateB(GPU array, dir to transpose with X)");
    MGArray src;
    imogen_mex_fail(imogen_mex_array(prhs[0], 0, &src), "cudaArrayRotateB");
    const int code = (int)mxGetScalar(prhs[1]);
    int64_t tag[GPU_TAG_MAXLEN];
    if (nlhs == 1) {
        MGArray *nu = NULL;
        imogen_mex_fail(flipArrayIndices(&src, &nu, 1, code, NULL, NULL), "cudaArrayRotateB");
        serializeMGArrayToTag(nu, tag);
        free(nu);
        plhs[0] = imogen_mex_tag_array(tag);
        return;
    }
    imogen_mex_fail(flipArrayIndices(&src, NULL, 1, code, NULL, NULL), "cudaArrayRotateB");
    serializeMGArrayToTag(&src, tag);
    if (mxGetClassID(prhs[0]) == mxINT64_CLASS) {
        memcpy(mxGetData(prhs[0]), tag, mxGetNumberOfElements(prhs[0]) * sizeof(int64_t));
    } else {
        const char *cls = mxGetClassName(prhs[0]);
        mxSetProperty((mxArray *)prhs[0], 0, (cls && strcmp(cls, "GPU_Type") == 0) ? "GPU_MemPtr" : "gputag",
                      imogen_mex_tag_array(tag));
    }
}
