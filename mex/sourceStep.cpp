// sourceStep.cpp -- MEX gateway: sourceStep(run, fluid, bx, by, bz, xyvector, [order, step #, method, tFraction])
// Call form of gpuclass/sourceStep.cu:71-200: dt = run.time.dTime * tFraction;
// the source types present in the run decide srcType (the reference's table,
// sourceStep.cu:228-250). This library builds the drag-only types (0, and 8
// for two fluids: one sourcefunction_2FluidDrag call with
// run.multifluidDragMethod); radiation, gravity, rotating-frame and
// cylindrical terms return ERROR_NOIMPLEMENT from performSourceFunctions.
#include <cstring>

#include "imogen_mex.h"

// run.(a).(b) as a double; the property copy is released here
static double field2(const mxArray *run, const char *a, const char *b, int *ok)
{
    mxArray *x = mxGetProperty(run, 0, a);
    if (!x) {
        *ok = 0;
        return 0.0;
    }
    const double v = imogen_mex_scalar(x, 0, b, ok);
    mxDestroyArray(x);
    return v;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)plhs;
    if (nrhs != 7 || nlhs != 0)
        imogen_mex_usage("sourceStep", "sourceStep(run, fluid, bx, by, bz, xyvector, [order, step #, step method, tFraction])");
    if (!mxIsDouble(prhs[6]) || mxGetNumberOfElements(prhs[6]) != 4)
        imogen_mex_usage("sourceStep", "a 4-element parameter vector [order, step #, step method, tFraction]");
    const double *s = mxGetPr(prhs[6]);
    const mxArray *run = prhs[0];
    int ok;
    FluidStepParams fsp = {};
    fsp.dt = field2(run, "time", "dTime", &ok) * s[3];
    if (!ok) imogen_mex_fail(ERROR_INVALID_ARGS, "sourceStep (run.time.dTime missing)");
    fsp.onlyHydro = (int)imogen_mex_scalar(run, 0, "pureHydro", &ok);
    fsp.stepDirection = (int)s[0];
    fsp.stepNumber = (int)s[1];
    fsp.stepMethod = (FluidMethods)(int)s[2];
    mxArray *geo = mxGetProperty(run, 0, "geometry");
    const int grc = imogen_mex_geometry(geo, &fsp.geometry);
    if (geo) mxDestroyArray(geo);
    imogen_mex_fail(grc, "sourceStep");
    const int nf = (int)mxGetNumberOfElements(prhs[1]);
    if (nf > 2) imogen_mex_fail(ERROR_INVALID_ARGS, "sourceStep (at most 2 fluids)");
    fsp.multifluidDragMethod = (nf > 1) ? (int)imogen_mex_scalar(run, 0, "multifluidDragMethod", &ok) : 0;
    ParametricRadiation prad = {};
    if ((int)field2(run, "radiation", "type", &ok)) {
        prad.exponent = field2(run, "radiation", "exponent", &ok);
        prad.minTemperature = 1.05;
        prad.prefactor = field2(run, "radiation", "strength", &ok) * fsp.dt;
    }
    // srcType bits: 1 cylindrical, 2 rotating frame, 4 scalar potential, 8 two fluids
    int srcType = 0;
    if (fsp.geometry.shape == CYLINDRICAL) srcType |= 1;
    if (fsp.geometry.frameOmega != 0) srcType |= 2;
    if ((int)field2(run, "potentialField", "ACTIVE", &ok)) srcType |= 4;
    if (nf > 1) srcType |= 8;
    GridFluid fluids[2];
    memset(fluids, 0, sizeof(fluids));
    for (int f = 0; f < nf; f++) {
        imogen_mex_fail(imogen_mex_fluid(prhs[1], f, fluids[f].data), "sourceStep");
        imogen_mex_fail(imogen_mex_thermo(prhs[1], f, &fluids[f].thermo), "sourceStep");
        fluids[f].DataHolder = fluids[f].data[0];
        fluids[f].rhoMin = imogen_mex_scalar(prhs[1], f, "MINMASS", &ok);
    }
    ParallelTopology topo;
    imogen_mex_topology(&topo);
    imogen_mex_fail(performSourceFunctions(srcType, fluids, nf, fsp, &topo, NULL, &prad, NULL), "sourceStep");
}
