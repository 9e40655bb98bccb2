// cudaSoundspeed.cpp -- MEX gateway: cs = cudaSoundspeed(rho, E, px, py, pz, gamma)
// Call form of gpuclass/cudaSoundspeed.cu:54-100 (time/TimeManager.m:143-146).
// The magnetised form (bx, by, bz before gamma) returns ERROR_NOIMPLEMENT.
#include <cstdlib>

#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nlhs != 1 || (nrhs != 6 && nrhs != 9))
        imogen_mex_usage("cudaSoundspeed", "c_s = cudaSoundspeed(mass, ener, momx, momy, momz, [bx, by, bz,] gamma)");
    if (nrhs == 9) imogen_mex_fail(ERROR_NOIMPLEMENT, "cudaSoundspeed (magnetised form)");
    MGArray fluid[5];
    for (int k = 0; k < 5; k++) imogen_mex_fail(imogen_mex_array(prhs[k], 0, &fluid[k]), "cudaSoundspeed");
    const double gamma = mxGetScalar(prhs[5]);
    MGArray *cs = NULL;
    imogen_mex_fail(MGA_allocArrays(&cs, 1, &fluid[0]), "cudaSoundspeed");
    const int rc = calculateSoundspeed(fluid, NULL, cs, gamma);
    if (rc != SUCCESSFUL) {
        MGA_delete(cs);
        free(cs);
        imogen_mex_fail(rc, "cudaSoundspeed");
    }
    int64_t tag[GPU_TAG_MAXLEN];
    serializeMGArrayToTag(cs, tag);
    free(cs);  // the device array now belongs to the returned tag (GPU_Type frees it)
    plhs[0] = imogen_mex_tag_array(tag);
}
