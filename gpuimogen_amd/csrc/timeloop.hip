// timeloop.hip -- Imogen iterations with the time step kept on the device.
//
// One iteration is what imogen.m:115-139 runs:
//     TimeManager.update (time/TimeManager.m:124-186): tau = min over fluids of
//         the CFL limit (cflTimestep.cu:107-377); dTime = CFL * tau, clamped so
//         that t + 2 dTime does not pass tMax; t += 2 dTime
//     source(1/2); fluidstep(+1) for every fluid; source(1); fluidstep(-1) for
//         every fluid; source(1/2)
//         (fluidstep: flux.cu:27-155, three directional sweeps, each followed
//         by its boundary conditions; source(f): sources/src0001.m, two
//         gas-dust drag calls over f dTime / 2 each, then the boundary
//         conditions of both fluids in every direction -- only when a drag
//         method is set, imogen_timeloop_set_drag)
// The host-driven form of this loop reads tau back after every CFL reduction
// (cflTimestep.cu:214 synchronises), so the GPU idles while the host issues
// the next sweeps. Here the CFL partial minima, dTime and the per-direction
// lambda = dTime / h stay in device memory: the sweep kernels read lambda
// from there, and each iteration is one replayed hipGraph (one per sweep
// order, iteration mod 6). The kernels, and so the results, are those of
// performFluidUpdate_3D and imogen_cflTimestepFused.
#include <cmath>
#include <cstring>

#include "cfl_device.h"
#include "drag_device.h"
#include "imogen_internal.h"

namespace imogen {

// device clock: [0] time, [1] dTime, [2..4] lambda x, y, z
constexpr int kClockDoubles = 8;
constexpr int kCflSlot = 2056;  // tau + 7 spare + up to 2048 block minima (launch_cfl)
constexpr int kMaxFluids = 4;

// A limit tau that is not positive and below 1e36 (every cell NaN, or no cell
// limiting: the reductions' 1e37 / 1e38 sentinels) takes no time step: dTime
// = 0 and the refusal is counted in the device error word.
__global__ void clock_advance(const double *__restrict__ cflOut, int nfluids, double cfl, double tmax,
                              double hx, double hy, double hz, double *__restrict__ clk,
                              unsigned long long *__restrict__ err)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double tau = cflOut[0];
    for (int f = 1; f < nfluids; f++) {
        const double t = cflOut[(int64_t)f * kCflSlot];
        tau = (t < tau) ? t : tau;
    }
    double dt = cfl * tau;
    const double t = clk[0];
    if (t + 2 * dt > tmax) dt = 0.5 * (tmax - t);
    if (!(tau > 0 && tau < 1e36)) {
        dt = 0.0;
        atomicAdd(err + ERRW_TIMESTEP, 1ull);
    }
    clk[1] = dt;
    clk[2] = dt / hx;
    clk[3] = dt / hy;
    clk[4] = dt / hz;
    clk[0] = t + 2 * dt;
}

struct TimeloopImpl {
    int nf = 0;
    MGArray fluid[kMaxFluids][5];
    FluidStepParams p[kMaxFluids];
    double cflGamma[kMaxFluids];  // gamma of the CFL sound speed (the fluid's own)
    GeometryParams geo;
    double cfl = 0.4, tmax = HUGE_VAL;
    double *dbuf = nullptr;  // clock, then one CFL slot per fluid
    hipGraphExec_t graph[6] = {nullptr};
    void *scratchAt = nullptr;  // scratch block the graphs were captured with
    size_t scratchBytes = 0;
    // CFL folded into the last sweep of each iteration: per fluid, whether its
    // method and boundaries allow it, the per-wave minima and their capacity
    bool fuse[kMaxFluids] = {false};
    CflFuse cf[kMaxFluids];
    // drag source coupling fluid 0 (gas) and fluid 1 (dust); -1 = none
    int dragMethod = -1;
    DragConsts dc{};
    uint64_t modelGen = 0;  // learned model the graphs were captured with
};

}  // namespace imogen

struct ImogenTimeloop {
    imogen::TimeloopImpl impl;
};

namespace imogen {

static void drop_graphs(TimeloopImpl *L)
{
    for (int k = 0; k < 6; k++)
        if (L->graph[k]) {
            (void)hipGraphExecDestroy(L->graph[k]);
            L->graph[k] = nullptr;
        }
}

// CFL limit of fluid f into its slot (cfl_common in fluidstep.hip, on the device)
static int cfl_of(TimeloopImpl *L, int f, hipStream_t s)
{
    double *base;
    int64_t pitch;
    int rc = fluid_view(L->fluid[f], &base, &pitch);
    if (rc) return rc;
    const GeometryParams &g = L->geo;
    int spacedim = 0;
    if (g.globalRez[1] > 1) spacedim = 1;
    if (g.globalRez[2] > 1) spacedim = 2;
    const int dim = spacedim + 1;
    double hmin = g.h[0];
    if (dim > 1 && g.h[1] < hmin) hmin = g.h[1];
    if (dim > 2 && g.h[2] < hmin) hmin = g.h[2];
    const int method = (int)L->p[f].stepMethod;
    const int xj = (method == METHOD_XINJIN) && dim > 1;
    const int halve = (method == METHOD_HLL || method == METHOD_HLLC);
    const int slot = prof_begin(PROF_CFL, s);
    rc = launch_cfl(base, pitch, L->fluid[f][0].partNumel[0], nullptr, L->cflGamma[f], 1, dim, xj, halve,
                    g.h[0], g.h[1], g.h[2], hmin, L->dbuf + kClockDoubles + (int64_t)f * kCflSlot, s);
    prof_end(slot, s);
    return rc;
}

// The CFL args launch_cfl would use for fluid f (cfl_common in fluidstep.hip)
static CflArgs cfl_args(const TimeloopImpl *L, int f)
{
    const GeometryParams &g = L->geo;
    int spacedim = 0;
    if (g.globalRez[1] > 1) spacedim = 1;
    if (g.globalRez[2] > 1) spacedim = 2;
    const int dim = spacedim + 1;
    double hmin = g.h[0];
    if (dim > 1 && g.h[1] < hmin) hmin = g.h[1];
    if (dim > 2 && g.h[2] < hmin) hmin = g.h[2];
    const double gamma = L->cflGamma[f];
    return CflArgs{gamma * (gamma - 1), g.h[0], g.h[1], g.h[2], hmin, 1, dim, 0};
}

// sources/src0001.m: the drag over fraction * dTime as two calls of half that
// interval (dt read from the device clock; one pass over memory applies both
// calls), then the boundary conditions of both fluids in the directions
// FluidManager.setBoundaries visits
static int source(TimeloopImpl *L, double fraction, hipStream_t s)
{
    if (L->dragMethod < 0) return SUCCESSFUL;
    double *g, *d;
    int64_t pg, pd;
    int rc = fluid_view(L->fluid[0], &g, &pg);
    if (!rc) rc = fluid_view(L->fluid[1], &d, &pd);
    if (rc) return rc;
    const int64_t n = (int64_t)L->fluid[0][0].dim[0] * L->fluid[0][0].dim[1] * L->fluid[0][0].dim[2];
    const DragDt dt{0.0, L->dbuf + 1, fraction / 2};
    {
        const int slot = prof_begin(PROF_DRAG, s);
        rc = (L->dragMethod == DRAG_LEARNED) ? launch_learned_drag(g, d, pg, n, L->dc, dt, nullptr, s, 2)
                                             : launch_drag(g, d, pg, n, L->dc, dt, L->dragMethod, s, 2);
        prof_end(slot, s);
        if (rc) return rc;
    }
    for (int f = 0; f < 2; f++)
        for (int dir = 1; dir <= 3; dir++)
            if (L->geo.globalRez[dir - 1] > 1 || dir == 1)
                if ((rc = setFluidBoundary(L->fluid[f], &L->geo, dir))) return rc;
    return SUCCESSFUL;
}

// fusedCfl: take this iteration's CFL from the minima the previous iteration's
// last sweep left (fluids that allow it); otherwise a full CFL pass.
static int iteration(TimeloopImpl *L, int it, bool fusedCfl, hipStream_t s)
{
    int rc;
    for (int f = 0; f < L->nf; f++) {
        if (fusedCfl && L->fuse[f]) {
            const MGArray *a = L->fluid[f];
            const int64_t n = sweep_waves(a->dim[0], a->dim[1], a->dim[2], last_sweep_dir(a, it - 1, -1));
            rc = launch_cfl_final(L->cf[f].part, n, 1, L->dbuf + kClockDoubles + (int64_t)f * kCflSlot, s);
        } else {
            rc = cfl_of(L, f, s);
        }
        if (rc) return rc;
    }
    hipLaunchKernelGGL(clock_advance, dim3(1), dim3(64), 0, s, L->dbuf + kClockDoubles, L->nf, L->cfl, L->tmax,
                       L->geo.h[0], L->geo.h[1], L->geo.h[2], L->dbuf, state()->err);
    if ((rc = IM_LAUNCHED("clock_advance"))) return rc;
    if ((rc = source(L, 0.5, s))) return rc;
    for (int sgn = 1; sgn >= -1; sgn -= 2) {
        for (int f = 0; f < L->nf; f++) {
            FluidStepParams p = L->p[f];
            p.stepNumber = it;
            p.stepDirection = sgn;
            p.dt = 0.0;  // unused: lambda comes from the device clock
            const CflFuse *cf = (sgn < 0 && L->fuse[f]) ? &L->cf[f] : nullptr;
            if ((rc = update_3d(L->fluid[f], p, L->dbuf + 2, cf))) return rc;
        }
        if ((rc = source(L, sgn > 0 ? 1.0 : 0.5, s))) return rc;
    }
    return SUCCESSFUL;
}

}  // namespace imogen

using namespace imogen;

extern "C" {

int imogen_timeloop_create(MGArray *fluids, int nfluids, const GeometryParams *geo, const FluidStepParams *params,
                           const double *cflGamma, double cfl, double tmax, ImogenTimeloop **out)
{
    if (!fluids || !geo || !params || !out) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    *out = nullptr;
    if (nfluids < 1 || nfluids > kMaxFluids)
        return IM_FAIL(ERROR_INVALID_ARGS, "1..%d fluids (got %d)", kMaxFluids, nfluids);
    if (!(cfl > 0)) return IM_FAIL(ERROR_INVALID_ARGS, "CFL factor %g", cfl);
    for (int f = 0; f < nfluids; f++) {
        double *base;
        int64_t pitch;
        int rc = fluid_view(fluids + 5 * f, &base, &pitch);
        if (rc) return rc;
        if (!is_xyz(fluids + 5 * f)) return IM_FAIL(ERROR_NOIMPLEMENT, "the device loop needs XYZ-ordered fluids");
        if (params[f].geometry.shape != SQUARE) return IM_FAIL(ERROR_NOIMPLEMENT, "only SQUARE geometry is built");
        if (params[f].onlyHydro != 1) return IM_FAIL(ERROR_NOIMPLEMENT, "MHD is not built");
        const int m = (int)params[f].stepMethod;
        if (m != METHOD_HLL && m != METHOD_HLLC && m != METHOD_XINJIN)
            return IM_FAIL(ERROR_INVALID_ARGS, "unknown fluid method %d", m);
        for (int d = 0; d < 3; d++)
            if (!(geo->h[d] > 0)) return IM_FAIL(ERROR_INVALID_ARGS, "cell size h[%d] = %g", d, geo->h[d]);
    }
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    ImogenTimeloop *T = new ImogenTimeloop();
    TimeloopImpl *L = &T->impl;
    L->nf = nfluids;
    for (int f = 0; f < nfluids; f++) {
        memcpy(L->fluid[f], fluids + 5 * f, sizeof(MGArray) * 5);
        L->p[f] = params[f];
        L->cflGamma[f] = cflGamma ? cflGamma[f] : params[f].thermoGamma;
    }
    L->geo = *geo;
    L->cfl = cfl;
    L->tmax = (tmax > 0) ? tmax : HUGE_VAL;
    for (int f = 0; f < nfluids; f++) {
        const int m = (int)L->p[f].stepMethod;
        MGArray *a = L->fluid[f];
        bool ok = (m == METHOD_HLLC || m == METHOD_HLL);
        int64_t cap = 0;
        for (int d = 1; d <= 3 && ok; d++) {
            const bool swept = (a->dim[1] == 1 && a->dim[2] == 1) ? (d == 1) : (a->dim[d - 1] > 3);
            if (!swept) continue;
            int lo, hi;
            if (a->dim[d - 1] <= 3 || cfl_keep_range(a, d, &lo, &hi)) ok = false;
            const int64_t w = sweep_waves(a->dim[0], a->dim[1], a->dim[2], d);
            cap = (w > cap) ? w : cap;
        }
        if (!ok || cap == 0) continue;
        CflFuse &c = L->cf[f];
        c = CflFuse{};
        if (IM_HIP(hipMalloc(&c.part, (size_t)cap * sizeof(double)))) {
            c.part = nullptr;
            imogen_timeloop_destroy(T);
            return ERROR_NOMEM;
        }
        c.a = cfl_args(L, f);
        c.cap = cap;
        L->fuse[f] = true;
    }
    const size_t bytes = sizeof(double) * (kClockDoubles + (size_t)nfluids * kCflSlot);
    if (IM_HIP(hipMalloc(&L->dbuf, bytes))) {
        L->dbuf = nullptr;
        imogen_timeloop_destroy(T);
        return ERROR_NOMEM;
    }
    // start at t = 0
    static const double zeros[kClockDoubles] = {0};
    int rc = IM_HIP(hipMemcpyAsync(L->dbuf, zeros, sizeof(zeros), hipMemcpyHostToDevice, st->stream));
    if (!rc) rc = IM_HIP(hipStreamSynchronize(st->stream));
    if (rc) {
        imogen_timeloop_destroy(T);
        return rc;
    }
    *out = T;
    return SUCCESSFUL;
}

// Runs `iterations` iterations numbered firstIteration, firstIteration + 1, ...
// (the number selects the sweep order, flux.cu:80-86). useGraph != 0 replays
// one captured graph per sweep order; the first iteration of each order is
// captured after an eager warm-up has sized the library's scratch block, and
// the graphs are re-captured if that block has moved since.
int imogen_timeloop_run(ImogenTimeloop *T, int firstIteration, int iterations, int useGraph)
{
    if (!T) return IM_FAIL(ERROR_NULL_POINTER, "null time loop");
    TimeloopImpl *L = &T->impl;
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    for (int k = 0; k < iterations; k++) {
        const int it = firstIteration + k;
        // The first iteration of a call takes a full CFL pass (the state may
        // have changed since the last call); later ones read the minima of
        // the previous iteration's last sweep.
        if (!useGraph || k == 0) {
            int rc = iteration(L, it, k > 0, st->stream);
            if (rc) return rc;
            continue;
        }
        if (L->scratchAt != st->scratch || L->scratchBytes != st->scratchBytes) {
            drop_graphs(L);
            L->scratchAt = st->scratch;
            L->scratchBytes = st->scratchBytes;
        }
        if (L->dragMethod == DRAG_LEARNED && L->modelGen != learned_model_generation()) {
            drop_graphs(L);  // the graphs hold the network's weights by value
            L->modelGen = learned_model_generation();
        }
        hipGraphExec_t &g = L->graph[((it % 6) + 6) % 6];
        if (!g) {
            // eager run of this iteration: sizes the scratch block (no allocation
            // may happen while capturing) and advances the state as the replay would
            int rc = iteration(L, it, true, st->stream);
            if (rc) return rc;
            if (L->scratchAt != st->scratch || L->scratchBytes != st->scratchBytes) {
                drop_graphs(L);
                L->scratchAt = st->scratch;
                L->scratchBytes = st->scratchBytes;
            }
            // record (not run) the same iteration as this order's graph
            hipGraph_t graph;
            prof_suspend(1);
            if (IM_HIP(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal))) {
                prof_suspend(0);
                return ERROR_CUDA_BLEW_UP;
            }
            rc = iteration(L, it, true, st->stream);
            hipError_t ec = hipStreamEndCapture(st->stream, &graph);
            prof_suspend(0);
            if (rc) return rc;
            if (IM_HIP(ec)) return ERROR_CUDA_BLEW_UP;
            ec = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (IM_HIP(ec)) return ERROR_CUDA_BLEW_UP;
            continue;  // this iteration already ran eagerly
        }
        if (IM_HIP(hipGraphLaunch(g, st->stream))) return ERROR_CUDA_BLEW_UP;
    }
    return take_device_errors(st);
}

int imogen_timeloop_set_drag(ImogenTimeloop *T, const ThermoDetails *gas, const ThermoDetails *dust, int method)
{
    if (!T) return IM_FAIL(ERROR_NULL_POINTER, "null time loop");
    TimeloopImpl *L = &T->impl;
    if (method < 0) {
        L->dragMethod = -1;
        drop_graphs(L);
        return SUCCESSFUL;
    }
    if (!gas || !dust) return IM_FAIL(ERROR_NULL_POINTER, "null ThermoDetails");
    if (L->nf != 2) return IM_FAIL(ERROR_INVALID_ARGS, "the drag source couples exactly two fluids (got %d)", L->nf);
    switch (method) {
    case DRAG_EMP:
    case DRAG_RK4:
    case DRAG_ETDRK1:
    case DRAG_LOGTRAP2:
    case DRAG_LOGTRAP3:
    case DRAG_EXACT:
    case DRAG_LEARNED: break;
    case DRAG_ETDRK2: return IM_FAIL(ERROR_NOIMPLEMENT, "drag method %d is not built", method);
    default: return IM_FAIL(ERROR_INVALID_ARGS, "unknown drag method %d", method);
    }
    double *g, *d;
    int64_t pg, pd;
    int rc = fluid_view(L->fluid[0], &g, &pg);
    if (!rc) rc = fluid_view(L->fluid[1], &d, &pd);
    if (rc) return rc;
    for (int i = 0; i < 3; i++)
        if (L->fluid[0][0].dim[i] != L->fluid[1][0].dim[i])
            return IM_FAIL(ERROR_INVALID_ARGS, "gas and dust grids differ in dimension %d", i + 1);
    if (pg != pd) return IM_FAIL(ERROR_INVALID_ARGS, "gas and dust slab pitches differ");
    if (learned_drag_prepare()) return ERROR_CUDA_BLEW_UP;  // no allocation once graphs capture
    L->dragMethod = method;
    L->dc = drag_consts(gas, dust);
    L->modelGen = learned_model_generation();
    // the last source changes the state after the last sweep, so the next
    // CFL limit needs its own pass
    for (int f = 0; f < L->nf; f++) L->fuse[f] = false;
    drop_graphs(L);
    return SUCCESSFUL;
}

int imogen_timeloop_clock(ImogenTimeloop *T, double *time, double *dt)
{
    if (!T) return IM_FAIL(ERROR_NULL_POINTER, "null time loop");
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    double c[2];
    int rc = IM_HIP(hipMemcpyAsync(c, T->impl.dbuf, sizeof(c), hipMemcpyDeviceToHost, st->stream));
    if (!rc) rc = IM_HIP(hipStreamSynchronize(st->stream));
    if (rc) return rc;
    if (time) *time = c[0];
    if (dt) *dt = c[1];
    return SUCCESSFUL;
}

int imogen_timeloop_set_time(ImogenTimeloop *T, double time)
{
    if (!T) return IM_FAIL(ERROR_NULL_POINTER, "null time loop");
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    int rc = IM_HIP(hipMemcpyAsync(T->impl.dbuf, &time, sizeof(double), hipMemcpyHostToDevice, st->stream));
    if (!rc) rc = IM_HIP(hipStreamSynchronize(st->stream));
    return rc;
}

int imogen_timeloop_destroy(ImogenTimeloop *T)
{
    if (!T) return SUCCESSFUL;
    DeviceState *st = state();
    if (st) (void)hipStreamSynchronize(st->stream);
    drop_graphs(&T->impl);
    if (T->impl.dbuf) (void)hipFree(T->impl.dbuf);
    for (int f = 0; f < T->impl.nf; f++)
        if (T->impl.cf[f].part) (void)hipFree(T->impl.cf[f].part);
    delete T;
    return SUCCESSFUL;
}

}  // extern "C"
