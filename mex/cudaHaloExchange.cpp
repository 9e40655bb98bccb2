// cudaHaloExchange.cpp -- MEX gateway: cudaHaloExchange(array, direction, topology, circularity)
// Call form of gpuclass/cudaHaloExchange.cu (the flux.m path). One rank has no
// halos to exchange: periodic lines wrap inside the sweeps.
#include "imogen_mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)plhs;
    if (nlhs != 0 || nrhs != 4) imogen_mex_usage("cudaHaloExchange", "cudaHaloExchange(array, direction, topology, circularity)");
    MGArray phi;
    imogen_mex_fail(imogen_mex_array(prhs[0], 0, &phi), "cudaHaloExchange");
    ParallelTopology topo;
    imogen_mex_topology(&topo);
    imogen_mex_fail(exchange_MPI_Halos(&phi, 1, &topo, (int)mxGetScalar(prhs[1])), "cudaHaloExchange");
}
