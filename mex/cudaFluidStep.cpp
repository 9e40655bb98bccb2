// cudaFluidStep.cpp -- MEX gateway: cudaFluidStep([dt, purehydro, gamma,
//     MINMASS, method, direction (, C_iso)], rho, E, px, py, pz, bx, by, bz, geometry)
// Call form of gpuclass/cudaFluidStep.cu:161-227 (fluid/relaxingFluid.m:27):
// one directional predictor + corrector. A 7th parameter makes the step
// isothermal with gamma = 2, as there.
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)plhs;
    if (nrhs != 10 || nlhs != 0)
        imogen_mex_usage("cudaFluidStep",
                         "cudaFluidStep([dt, purehydro, gamma, MINMASS, method, direction], rho, E, px, py, pz, bx, by, bz, geometry)");
    const size_t np = mxGetNumberOfElements(prhs[0]);
    if (!mxIsDouble(prhs[0]) || (np != 6 && np != 7))
        imogen_mex_usage("cudaFluidStep", "a 6- or 7-element parameter vector");
    const double *s = mxGetPr(prhs[0]);
    FluidStepParams p = {};
    p.dt = s[0];
    p.onlyHydro = (int)s[1];
    p.thermoGamma = s[2];
    p.minimumRho = s[3];
    const int method = (int)s[4];
    if (method < 1 || method > 3) imogen_mex_fail(ERROR_INVALID_ARGS, "cudaFluidStep (method must be 1, 2 or 3)");
    p.stepMethod = (FluidMethods)method;
    p.stepDirection = (int)s[5];
    p.Cisothermal = -1;
    if (np == 7) {
        p.Cisothermal = s[6];
        p.thermoGamma = 2;
    }
    p.slopeLimiter = LIMITER_OSPRE;
    if (!p.onlyHydro) imogen_mex_fail(ERROR_NOIMPLEMENT, "cudaFluidStep (magnetohydrodynamics)");
    MGArray fluid[5];
    for (int k = 0; k < 5; k++) imogen_mex_fail(imogen_mex_array(prhs[1 + k], 0, &fluid[k]), "cudaFluidStep");
    imogen_mex_fail(imogen_mex_geometry(prhs[9], &p.geometry), "cudaFluidStep");
    ParallelTopology topo;
    imogen_mex_topology(&topo);
    imogen_mex_fail(performFluidUpdate_1D(fluid, p, &topo, NULL), "cudaFluidStep");
}
