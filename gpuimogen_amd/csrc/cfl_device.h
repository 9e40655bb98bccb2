// cfl_device.h -- per-cell sound speed and CFL time-step limit (FP64).
// cukern_Soundspeed_hd (gpuclass/cudaSoundspeed.cu:178-195) and
// cukern_CFLtimestep (gpuclass/cflTimestep.cu:262-377, SQUARE geometry).
#pragma once
#include <hip/hip_runtime.h>

namespace imogen {

// The CFL arithmetic is evaluated exactly as written (no FMA contraction), so
// the limit of a cell is bitwise the same in the CFL kernel and in a sweep
// that folds the CFL in, whatever the surrounding code lets the compiler fuse.
__device__ __forceinline__ double cs_of(double rho, double E, double px, double py, double pz, double gg1)
{
#pragma clang fp contract(off)
    double rhoinv = 1 / rho;
    double csq = gg1 * (E - .5 * (px * px + py * py + pz * pz) * rhoinv) * rhoinv;
    if (csq < 0.0) csq = 0.0;
    return sqrt(csq);
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    return v;
}

// Adds the wave's total of the lanes' counts to *slot with one atomic (none
// when the total is 0). Must be reached by all 64 lanes; slot null: no-op.
__device__ __forceinline__ void wave_count(unsigned long long *slot, unsigned v)
{
    if (!slot) return;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(slot, (unsigned long long)v);
}

// dim = simulation dimension (1..3), xinjin selects the per-direction form.
// fused != 0: sound speed computed in-line from gamma instead of read from cs.
struct CflArgs {
    double gg1, hx, hy, hz, hmin;
    int fused, dim, xinjin;
};

// time-step limit of one cell (cflTimestep.cu:292-335)
__device__ __forceinline__ double cfl_cell(const CflArgs &a, double rho, double E, double px, double py,
                                           double pz, double cIn)
{
#pragma clang fp contract(off)
    const double c = a.fused ? cs_of(rho, E, px, py, pz, a.gg1) : cIn;
    double u;
    if (!a.xinjin) {
        double pm;
        if (a.dim == 1) {
            pm = fabs(px);
        } else {
            double p = px * px;
            p = p + py * py;
            if (a.dim > 2) p = p + pz * pz;
            pm = sqrt(p);
        }
        u = pm / rho + c;
        u = a.hmin / u;
    } else {
        u = a.hx / (fabs(px) / rho + c);
        double v = a.hy / (fabs(py) / rho + c);
        u = (u < v) ? u : v;
        if (a.dim > 2) {
            v = a.hz / (fabs(pz) / rho + c);
            u = (u < v) ? u : v;
        }
    }
    return u;
}

// CFL minimum folded into a sweep: each wave of the sweep writes the minimum
// time-step limit over the cells it stores whose value survives the
// direction's boundary conditions (cells [lo, hi) of the swept direction;
// the ghost cells the boundary then overwrites are copies or mirror images
// of kept cells, with the same limit). part = null: off.
struct CflFuse {
    double *part;
    CflArgs a;
    int lo, hi;
    int64_t cap;  // entries available in part
};

}  // namespace imogen
