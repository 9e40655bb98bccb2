// GPU_upload.cpp -- MEX gateway: tag = GPU_upload(host array [, device list [, [halo, partitionDir, exteriorHalo, clone]]])
// Call form of gpuclass/GPU_upload.cu:19-130 (gpuclass/GPU_Type.m:193,225).
// One GPU: a device list must name exactly one device, the current one.
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nlhs != 1 || nrhs < 1 || nrhs > 3)
        imogen_mex_usage("GPU_upload", "result_tag = GPU_upload(host array [, device list [, (halo [,partition direct [, clone_partitions]])]])");
    if (!mxIsDouble(prhs[0]) || mxIsComplex(prhs[0])) imogen_mex_fail(ERROR_INVALID_ARGS, "GPU_upload (host array must be real double)");
    if (nrhs >= 2 && mxGetNumberOfElements(prhs[1]) != 1)
        imogen_mex_fail(ERROR_NOIMPLEMENT, "GPU_upload (one GPU per process: the device list must name one device)");
    int halo = 0, partDir = PARTITION_X;
    if (nrhs >= 3 && mxGetNumberOfElements(prhs[2]) >= 1) {
        const double *d = mxGetPr(prhs[2]);
        halo = (int)d[0] < 0 ? 0 : (int)d[0];
        if (mxGetNumberOfElements(prhs[2]) >= 2) {
            partDir = (int)d[1];
            if (partDir < 1 || partDir > 3) partDir = PARTITION_X;
        }
    }
    const mwSize nd = mxGetNumberOfDimensions(prhs[0]);
    if (nd > 3) imogen_mex_fail(ERROR_INVALID_ARGS, "GPU_upload (more than 3 dimensions)");
    const mwSize *idims = mxGetDimensions(prhs[0]);
    int dims[3] = {1, 1, 1};
    for (mwSize k = 0; k < nd; k++) dims[k] = (int)idims[k];
    int64_t tag[GPU_TAG_MAXLEN];
    imogen_mex_fail(GPU_upload(mxGetPr(prhs[0]), dims, halo, partDir, tag), "GPU_upload");
    plhs[0] = imogen_mex_tag_array(tag);
}
