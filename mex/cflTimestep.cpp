// cflTimestep.cpp -- MEX gateway: tau = cflTimestep(FluidManager, soundspeed, GeometryManager, method)
// Call form of gpuclass/cflTimestep.cu:64-103 (time/TimeManager.m:150).
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nlhs != 1 || nrhs != 4)
        imogen_mex_usage("cflTimestep", "tau = cflTimestep(FluidManager, soundspeed gpu array, GeometryManager, cfd_method)");
    MGArray fluid[5], cs;
    GeometryParams geo;
    imogen_mex_fail(imogen_mex_fluid(prhs[0], 0, fluid), "cflTimestep");
    imogen_mex_fail(imogen_mex_array(prhs[1], 0, &cs), "cflTimestep");
    imogen_mex_fail(imogen_mex_geometry(prhs[2], &geo), "cflTimestep");
    if (!mxIsDouble(prhs[3]) || mxGetNumberOfElements(prhs[3]) < 1) imogen_mex_usage("cflTimestep", "a numeric method");
    const int method = (int)mxGetPr(prhs[3])[0];
    double rez[3] = {(double)geo.globalRez[0], (double)geo.globalRez[1], (double)geo.globalRez[2]};
    plhs[0] = mxCreateDoubleMatrix(1, 1, mxREAL);
    imogen_mex_fail(computeLocalCFLTimestep(fluid, &cs, &geo, method, rez, mxGetPr(plhs[0])), "cflTimestep");
}
