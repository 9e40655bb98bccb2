// cudaStatics.cpp -- MEX gateway: cudaStatics(ImogenArray, blockdim, GeometryManager, direction)
// Call form of gpuclass/cudaStatics.cu:85-104 (ImogenArray.applyBoundaryConditions):
// both faces of the direction, minus then plus.
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)plhs;
    if (nlhs != 0 || nrhs != 4)
        imogen_mex_usage("cudaStatics", "cudaStatics(ImogenArray, blockdim, GeometryManager, direction)");
    MGArray phi;
    GeometryParams geo;
    imogen_mex_fail(imogen_mex_geometry(prhs[2], &geo), "cudaStatics");
    imogen_mex_fail(imogen_mex_array(prhs[0], 0, &phi), "cudaStatics");
    const int dir = (int)mxGetScalar(prhs[3]);
    imogen_mex_fail(setArrayBoundaryConditions(&phi, &geo, dir, 0), "cudaStatics");
    imogen_mex_fail(setArrayBoundaryConditions(&phi, &geo, dir, 1), "cudaStatics");
}
