// drag.hip -- gas-dust drag source (sourcefunction_2FluidDrag) for gfx950.
//
// Per cell, from the gas (rho, E, p) and dust (rho, p) states:
//   |dv| and u_0 (cukern_prepareForERK_h0, gpuclass/cudaSource2FluidDrag.cu:1883-1908),
//   a decay factor f of the relative velocity over dt, from
//     LogTrap2 / LogTrap3 (cukern_LogTrapSolve, :1532-1630), or
//     exponential Euler with the rate frozen at t = 0 (solveDragETDRK1,
//     :477-577), or
//     the explicit midpoint / RK4 integrators of |dv| (solveDragEMP,
//     solveDragRK4, :305-467; these rebuild the state with
//     cukern_applyFinalDeltaV, :1160-1237), or
//     the adaptive reference integrator (this library's "exact" method), or
//     the learned surrogate with per-cell fallback (mlp_drag.hip),
//   then new momenta about the barycentric velocity and the gas energy change
//   that conserves total energy (:1632-1676).
// The reference stages |dv|, u and k through three temporary slabs and
// several launches; here one thread carries a cell through the whole method
// in registers: 72 bytes read and 56 bytes written per cell.
#include <cstring>

#include "cfl_device.h"
#include "drag_device.h"
#include "imogen_internal.h"

namespace imogen {

// Cells the method cannot be applied to (drag_cell_defined) are updated as
// the arithmetic falls (NaN, as the reference's kernels do) and counted into
// err[ERRW_DRAG_INVALID]; the adaptive integrator's budget overruns into
// err[ERRW_DRAG_BUDGET]. `calls` (1 or 2) drag calls of dt each are applied to
// a cell in registers before it is stored: sources/src0001.m's two
// back-to-back half-interval calls in one pass over memory, with the same
// arithmetic as two launches (the second call reads what the first stored).
template <int METHOD>
__global__ __launch_bounds__(256) void drag_kernel(double *__restrict__ g, double *__restrict__ d, int64_t pitch,
                                                   int64_t n, DragConsts c, DragDt dtv, int calls,
                                                   unsigned long long *__restrict__ err)
{
    const double dt = drag_dt(dtv);
    unsigned invalid = 0, overrun = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        // the nine values the method reads and writes, as a pitch-1 cell
        double G[5], D[5];
#pragma unroll
        for (int k = 0; k < 5; k++) G[k] = g[i + k * pitch];
        D[0] = d[i];
#pragma unroll
        for (int k = 2; k < 5; k++) D[k] = d[i + k * pitch];
        D[1] = 0.0;  // the dust energy is neither read nor written
        for (int r = 0; r < calls; r++) {
            DragCell s = drag_load(G, D, 1);
            if (METHOD == DRAG_ETDRK1) {
                invalid += !drag_cell_defined(s.rg, s.rd, s.s0, s.u0);
                drag_etd1_apply(G, D, 1, dt, drag_k(c, s.rg, s.rd, s.s0, s.u0));
                continue;
            }
            double f;
            if (METHOD == DRAG_EXACT) {
                int k = 0;
                f = drag_exact_ratio(c, s.rg, s.rd, s.s0, s.u0, dt, &k);
                invalid += (k == kDragInvalid);
                overrun += (k < 0 && k != kDragInvalid);
            } else {
                invalid += !drag_cell_defined(s.rg, s.rd, s.s0, s.u0);
                f = exp(-dt * drag_logtrap_rate(c, s.rg, s.rd, s.s0, s.u0, dt, METHOD == DRAG_LOGTRAP3 ? 3 : 2));
            }
            drag_apply(G, D, 1, s, f);
        }
        g[i + pitch] = G[1];
#pragma unroll
        for (int k = 2; k < 5; k++) {
            g[i + k * pitch] = G[k];
            d[i + k * pitch] = D[k];
        }
    }
    wave_count(err + ERRW_DRAG_INVALID, invalid);
    if (METHOD == DRAG_EXACT) wave_count(err + ERRW_DRAG_BUDGET, overrun);
}

// explicit midpoint (method 0) / RK4 (method 1): 72 bytes read, 64 written per
// cell (the dust energy is reset); `calls` as drag_kernel
template <int METHOD>
__global__ __launch_bounds__(256) void drag_rk_kernel(double *__restrict__ g, double *__restrict__ d, int64_t pitch,
                                                      int64_t n, DragConsts c, DragDt dtv, int calls,
                                                      unsigned long long *__restrict__ err)
{
    const double dt = drag_dt(dtv);
    unsigned invalid = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double G[5], D[5];
#pragma unroll
        for (int k = 0; k < 5; k++) G[k] = g[i + k * pitch];
        D[0] = d[i];
        D[1] = 0.0;
#pragma unroll
        for (int k = 2; k < 5; k++) D[k] = d[i + k * pitch];
        for (int r = 0; r < calls; r++) {
            const DragCell s = drag_rk_load(G, D, 1);
            invalid += !drag_cell_defined(s.rg, s.rd, s.s0, s.u0);
            drag_rk_apply(G, D, 1, drag_rk_magdv<METHOD>(c, s.rg, s.rd, s.s0, s.u0, dt));
        }
#pragma unroll
        for (int k = 1; k < 5; k++) {
            g[i + k * pitch] = G[k];
            d[i + k * pitch] = D[k];
        }
    }
    wave_count(err + ERRW_DRAG_INVALID, invalid);
}

DragConsts drag_consts(const ThermoDetails *tg, const ThermoDetails *td)
{
    // hostDrag[0..8] of sourcefunction_2FluidDrag (cudaSource2FluidDrag.cu:210-220)
    const double gam = tg->gamma;
    const double ddust = sqrt(td->sigma0 / 3.141592653589793);
    DragConsts c;
    c.visc0 = tg->mu0;
    c.viscpow = -tg->muTindex;
    c.lampow = tg->sigmaTindex;
    c.alpha = tg->m * (gam - 1.0) / (298.15 * tg->kBolt);
    c.beta = sqrt(2.0) * tg->m / (tg->sigma0 * ddust);
    c.delta = ddust / tg->mu0;
    c.epsilon = td->sigma0 / (1.0 * td->m);
    c.gm1 = gam - 1.0;
    c.tol = kDragExactTol;
    return c;
}

int launch_drag(double *g, double *d, int64_t pitch, int64_t n, const DragConsts &c, DragDt dt, int method,
                hipStream_t s, int calls)
{
    if (calls < 1 || calls > 2) return IM_FAIL(ERROR_INVALID_ARGS, "drag calls per pass: %d", calls);
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    unsigned long long *err = st->err;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    switch (method) {
    case DRAG_LOGTRAP2:
        hipLaunchKernelGGL((drag_kernel<DRAG_LOGTRAP2>), dim3((unsigned)blocks), dim3(256), 0, s, g, d, pitch, n, c,
                           dt, calls, err);
        break;
    case DRAG_LOGTRAP3:
        hipLaunchKernelGGL((drag_kernel<DRAG_LOGTRAP3>), dim3((unsigned)blocks), dim3(256), 0, s, g, d, pitch, n, c,
                           dt, calls, err);
        break;
    case DRAG_EXACT:
        hipLaunchKernelGGL((drag_kernel<DRAG_EXACT>), dim3((unsigned)blocks), dim3(256), 0, s, g, d, pitch, n, c,
                           dt, calls, err);
        break;
    case DRAG_ETDRK1:
        hipLaunchKernelGGL((drag_kernel<DRAG_ETDRK1>), dim3((unsigned)blocks), dim3(256), 0, s, g, d, pitch, n, c,
                           dt, calls, err);
        break;
    case DRAG_EMP:
        hipLaunchKernelGGL((drag_rk_kernel<DRAG_EMP>), dim3((unsigned)blocks), dim3(256), 0, s, g, d, pitch, n, c, dt, calls, err);
        break;
    case DRAG_RK4:
        hipLaunchKernelGGL((drag_rk_kernel<DRAG_RK4>), dim3((unsigned)blocks), dim3(256), 0, s, g, d, pitch, n, c, dt, calls, err);
        break;
    default: return IM_FAIL(ERROR_INVALID_ARGS, "drag method %d", method);
    }
    return IM_LAUNCHED("drag_kernel");
}

}  // namespace imogen

using namespace imogen;

extern "C" {

int sourcefunction_2FluidDrag(MGArray *fluidA, MGArray *fluidB, GeometryParams *geo, ThermoDetails *thermogas,
                              ThermoDetails *thermodust, double dt, int method)
{
    (void)geo;
    double *g, *d;
    int64_t pg, pd;
    int rc = fluid_view(fluidA, &g, &pg);
    if (rc) return rc;
    rc = fluid_view(fluidB, &d, &pd);
    if (rc) return rc;
    if (!thermogas || !thermodust) return IM_FAIL(ERROR_NULL_POINTER, "null ThermoDetails");
    for (int i = 0; i < 3; i++)
        if (fluidA->dim[i] != fluidB->dim[i])
            return IM_FAIL(ERROR_INVALID_ARGS, "gas and dust grids differ in dimension %d", i + 1);
    if (is_xyz(fluidA) != is_xyz(fluidB) ||
        (!is_xyz(fluidA) && memcmp(fluidA->currentPermutation, fluidB->currentPermutation, sizeof(int) * 3)))
        return IM_FAIL(ERROR_INVALID_ARGS, "gas and dust arrays are in different memory orders");
    if (pg != pd) return IM_FAIL(ERROR_INVALID_ARGS, "gas and dust slab pitches differ");
    const int64_t n = (int64_t)fluidA->dim[0] * fluidA->dim[1] * fluidA->dim[2];
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    const DragConsts c = drag_consts(thermogas, thermodust);
    const int slot = prof_begin(PROF_DRAG, st->stream);
    switch (method) {
    case DRAG_LOGTRAP2:
    case DRAG_LOGTRAP3:
    case DRAG_EXACT:
    case DRAG_EMP:
    case DRAG_RK4:
    case DRAG_ETDRK1: rc = launch_drag(g, d, pg, n, c, host_dt(dt), method, st->stream); break;
    case DRAG_LEARNED: rc = launch_learned_drag(g, d, pg, n, c, host_dt(dt), nullptr, st->stream); break;
    case DRAG_ETDRK2:  // listed as not implemented by the reference itself (classes/ImogenManager.m:181)
        rc = IM_FAIL(ERROR_NOIMPLEMENT, "drag method %d is not built", method);
        break;
    default: rc = IM_FAIL(ERROR_INVALID_ARGS, "unknown drag method %d", method);
    }
    prof_end(slot, st->stream);
    if (rc) return rc;
    return take_device_errors(st);  // waits: reports cells the method could not be applied to
}

int performSourceFunctions(int srcType, GridFluid *fluids, int nFluids, FluidStepParams fsp, ParallelTopology *topo,
                           GravityData *gravdata, ParametricRadiation *prad, MGArray *tmpStorage)
{
    (void)gravdata;
    (void)tmpStorage;
    if (!prad) return IM_FAIL(ERROR_NULL_POINTER, "null ParametricRadiation");
    int rc = check_topology(topo);
    if (rc) return rc;
    if (srcType < 0 || srcType > 15) return IM_FAIL(ERROR_INVALID_ARGS, "source type %d", srcType);
    if (srcType != 0 && srcType != 8)
        return IM_FAIL(ERROR_NOIMPLEMENT, "source type %d (cylindrical / rotating-frame / gravity terms) is not built",
                       srcType);
    if (prad->prefactor > 0) return IM_FAIL(ERROR_NOIMPLEMENT, "optically thin radiation is not built");
    if (srcType == 0) return SUCCESSFUL;  // srcBlank, sourceStep.cu:293-297
    // src___D without radiation (sourceStep.cu:300-326): one full-dt drag call
    if (!fluids) return IM_FAIL(ERROR_NULL_POINTER, "null fluids");
    if (nFluids < 2) return IM_FAIL(ERROR_INVALID_ARGS, "two-fluid drag needs 2 fluids (got %d)", nFluids);
    return sourcefunction_2FluidDrag(&fluids[0].data[0], &fluids[1].data[0], &fsp.geometry, &fluids[0].thermo,
                                     &fluids[1].thermo, fsp.dt, fsp.multifluidDragMethod);
}

}  // extern "C"
