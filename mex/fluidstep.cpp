// fluidstep.cpp -- MEX gateway: fluidstep(FluidManager[], bx, by, bz,
//     [dt, purehydro, stepDirection, stepNumber, method], GeometryManager)
// Call form and checks of gpuclass/flux_ML_iface.cu:37-162 (imogen.m:122,125):
// every fluid of the FluidManager array gets one performFluidUpdate_3D with
// its own gamma, isothermal sound speed and density floor. An isothermal
// fluid runs with gamma = 2 there (flux_ML_iface.cu:128-133), kept here.
#include "imogen_mex.h"

static FluidMethods method_of(double m)
{
    return (m == 1) ? METHOD_HLL : ((m == 3) ? METHOD_XINJIN : METHOD_HLLC);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)plhs;
    if (nrhs != 6 || nlhs != 0)
        imogen_mex_usage("fluidstep", "fluidstep(fluid, bx, by, bz, [dt, purehydro, order, step #, method], geometry)");
    if (!mxIsDouble(prhs[4]) || mxGetNumberOfElements(prhs[4]) != 5)
        imogen_mex_usage("fluidstep", "a 5-element parameter vector [dt, purehydro?, order, step #, step method]");
    const double *s = mxGetPr(prhs[4]);
    const int method = (int)s[4];
    if (method < 1 || method > 3) imogen_mex_fail(ERROR_INVALID_ARGS, "fluidstep (method must be 1, 2 or 3)");
    FluidStepParams fsp = {};
    imogen_mex_fail(imogen_mex_geometry(prhs[5], &fsp.geometry), "fluidstep");
    fsp.dt = s[0];
    fsp.onlyHydro = (int)s[1];
    fsp.stepDirection = (int)s[2];
    fsp.stepNumber = (int)s[3];
    fsp.stepMethod = method_of(s[4]);
    fsp.slopeLimiter = LIMITER_OSPRE;
    ParallelTopology topo;
    imogen_mex_topology(&topo);
    const int nf = (int)mxGetNumberOfElements(prhs[0]);
    for (int f = 0; f < nf; f++) {
        MGArray fluid[5];
        ThermoDetails th;
        imogen_mex_fail(imogen_mex_fluid(prhs[0], f, fluid), "fluidstep");
        imogen_mex_fail(imogen_mex_thermo(prhs[0], f, &th), "fluidstep");
        int ok;
        const double rhomin = imogen_mex_scalar(prhs[0], f, "MINMASS", &ok);
        if (!ok) imogen_mex_fail(ERROR_NULL_POINTER, "fluidstep (fluid has no MINMASS property)");
        fsp.thermoGamma = (th.Cisothermal != -1) ? 2.0 : th.gamma;
        fsp.Cisothermal = th.Cisothermal;
        fsp.minimumRho = rhomin;
        imogen_mex_fail(performFluidUpdate_3D(fluid, &topo, fsp, NULL), "fluidstep");
    }
}
