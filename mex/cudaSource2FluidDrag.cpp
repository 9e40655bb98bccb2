// cudaSource2FluidDrag.cpp -- MEX gateway: cudaSource2FluidDrag(FluidManager[2], geometry, [dt, method])
// Call form of gpuclass/cudaSource2FluidDrag.cu:125-180 (sources/src0001.m:7,13):
// fluid 1 is the gas, fluid 2 the dust.
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)plhs;
    if (nrhs != 3 || nlhs != 0)
        imogen_mex_usage("cudaSource2FluidDrag", "cudaSource2FluidDrag(FluidManager[2], geometry, [dt, solverMethod])");
    if (mxGetNumberOfElements(prhs[0]) < 2) imogen_mex_fail(ERROR_INVALID_ARGS, "cudaSource2FluidDrag (needs 2 fluids)");
    if (!mxIsDouble(prhs[2]) || mxGetNumberOfElements(prhs[2]) != 2)
        imogen_mex_usage("cudaSource2FluidDrag", "a 2-element parameter vector [dt, method]");
    MGArray gas[5], dust[5];
    ThermoDetails tg, td;
    GeometryParams geo;
    imogen_mex_fail(imogen_mex_fluid(prhs[0], 0, gas), "cudaSource2FluidDrag");
    imogen_mex_fail(imogen_mex_thermo(prhs[0], 0, &tg), "cudaSource2FluidDrag");
    imogen_mex_fail(imogen_mex_fluid(prhs[0], 1, dust), "cudaSource2FluidDrag");
    imogen_mex_fail(imogen_mex_thermo(prhs[0], 1, &td), "cudaSource2FluidDrag");
    imogen_mex_fail(imogen_mex_geometry(prhs[1], &geo), "cudaSource2FluidDrag");
    const double *p = mxGetPr(prhs[2]);
    imogen_mex_fail(sourcefunction_2FluidDrag(gas, dust, &geo, &tg, &td, p[0], (int)p[1]), "cudaSource2FluidDrag");
}
