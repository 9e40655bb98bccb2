// fluidstep.hip -- host side of the fluid step: the reference's compiled-core
// entry points (performFluidUpdate_3D/_1D, setFluidBoundary, calculateSoundspeed,
// computeLocalCFLTimestep) over the kernels of this library.
//
// The sweep order and per-direction skip rules follow gpuclass/flux.cu:27-155.
// No array is transposed: each direction's kernel walks its own stride, so the
// permutation bookkeeping of the reference (cudaArrayRotateB.cu:551-640)
// always ends where it started (identity), which is what the caller observes.
#include <cstring>

#include "cfl_device.h"
#include "imogen_internal.h"

using namespace imogen;

int imogen::fluid_view(MGArray *fluid, double **base, int64_t *pitch)
{
    if (!fluid) return IM_FAIL(ERROR_NULL_POINTER, "null fluid");
    if (fluid->nGPUs != 1)
        return IM_FAIL(ERROR_NOIMPLEMENT, "only single-GPU fluids are supported (nGPUs = %d)", fluid->nGPUs);
    const int64_t p = fluid->slabPitch[0] / (int64_t)sizeof(double);
    for (int i = 1; i < 5; i++) {
        if (fluid[i].devicePtr[0] != fluid[0].devicePtr[0] + i * p)
            return IM_FAIL(ERROR_INVALID_ARGS, "fluid array %d is not slab %d of the canister", i, i);
        // the slabs are rotated one by one (flux.m xchgIndices): all must agree
        if (is_xyz(&fluid[i]) != is_xyz(fluid) ||
            (!is_xyz(fluid) && memcmp(fluid[i].currentPermutation, fluid->currentPermutation, sizeof(int) * 3)))
            return IM_FAIL(ERROR_INVALID_ARGS, "fluid array %d is in another memory order than array 0", i);
    }
    if (!fluid->devicePtr[0]) return IM_FAIL(ERROR_NULL_POINTER, "null device pointer");
    *base = fluid->devicePtr[0];
    *pitch = p;
    return SUCCESSFUL;
}

bool imogen::is_xyz(const MGArray *a)
{
    return a->permtag == 0 || a->permtag == 1;
}

int imogen::check_topology(const ParallelTopology *t)
{
    if (!t) return SUCCESSFUL;
    for (int i = 0; i < 3; i++)
        if (t->nproc[i] > 1)
            return IM_FAIL(ERROR_NOIMPLEMENT, "multi-rank topologies are not built (nproc[%d] = %d)", i, t->nproc[i]);
    return SUCCESSFUL;
}

namespace {

// fluxcall table of gpuclass/flux.cu:46: direction swept at position n of sweep s
const int kFluxcall[3][6] = {{1, 2, 1, 2, 3, 3}, {3, 1, 2, 3, 1, 2}, {2, 3, 3, 1, 2, 1}};

// sas code of one array for one face (cudaStatics.cu:226-250)
int sas_code(const MGArray *a, int direction, int side, int *code)
{
    const BCModeTypes m = a->boundaryConditions.mode[2 * (direction - 1) + side];
    switch (m) {
    case circular:
    case stationary: *code = -1; return SUCCESSFUL;
    case mirror: *code = (a->vectorComponent == direction) ? 1 : 0; return SUCCESSFUL;
    case extrapConstant: *code = 2; return SUCCESSFUL;
    case extrapLinear: *code = 3; return SUCCESSFUL;
    case wall: return IM_FAIL(ERROR_INVALID_ARGS, "Wall BC is not implemented (as in the reference)");
    default: return IM_FAIL(ERROR_NOIMPLEMENT, "boundary mode %d is not built", (int)m);
    }
}

}  // namespace

namespace imogen {

// Runs fn on an XYZ-ordered copy of a canister whose arrays were rotated
// (flipArrayIndices) and writes the result back in the rotated order. The
// sweeps walk each physical direction's own stride, so a rotated canister is
// stepped exactly as its unrotated image: the reference's 1D update of memory-x
// (always the swept direction once flux.m has rotated the arrays) is the sweep
// of the memory axis that holds the physical direction.
template <class Fn>
static int on_xyz_copy(MGArray *fluid, Fn &&fn)
{
    double *base;
    int64_t pitch;
    int rc = fluid_view(fluid, &base, &pitch);
    if (rc) return rc;
    const int *perm = fluid->currentPermutation;
    if (perm[0] + perm[1] + perm[2] != 6 || perm[0] * perm[1] * perm[2] != 6)
        return IM_FAIL(ERROR_INVALID_ARGS, "currentPermutation (%d %d %d) is not a permutation of 1 2 3", perm[0],
                       perm[1], perm[2]);
    int pdim[3], toXyz[3], back[3];
    for (int a = 0; a < 3; a++) {
        if (perm[a] < 1 || perm[a] > 3) return IM_FAIL(ERROR_INVALID_ARGS, "currentPermutation[%d] = %d", a, perm[a]);
        pdim[perm[a] - 1] = fluid->dim[a];
        toXyz[perm[a] - 1] = a;  // XYZ axis perm[a]-1 comes from memory axis a
        back[a] = perm[a] - 1;   // memory axis a comes from XYZ axis perm[a]-1
    }
    const int64_t numel = (int64_t)pdim[0] * pdim[1] * pdim[2];
    const int64_t xpitch = ((numel * 8 + 255) / 256) * 256 / 8;
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    double *x = nullptr;
    if (IM_HIP(hipMalloc(&x, (size_t)xpitch * 5 * sizeof(double)))) return ERROR_NOMEM;
    MGArray xf[5];
    for (int k = 0; k < 5 && !rc; k++) {
        rc = launch_permute_copy(base + k * pitch, x + k * xpitch, fluid->dim, toXyz, st->stream);
        xf[k] = fluid[k];
        for (int a = 0; a < 3; a++) {
            xf[k].dim[a] = pdim[a];
            xf[k].currentPermutation[a] = a + 1;
        }
        xf[k].permtag = 1;
        xf[k].numel = numel;
        xf[k].partNumel[0] = (int)numel;
        xf[k].slabPitch[0] = xpitch * (int64_t)sizeof(double);
        xf[k].devicePtr[0] = x + k * xpitch;
        xf[k].numSlabs = (k == 0) ? 5 : -k;
    }
    if (!rc) rc = fn(xf);
    for (int k = 0; k < 5 && !rc; k++) rc = launch_permute_copy(x + k * xpitch, base + k * pitch, pdim, back, st->stream);
    (void)hipStreamSynchronize(st->stream);
    (void)hipFree(x);
    return rc;
}

// Static cells of one array (setArrayStaticCells, cudaStatics.cu:288-350): the
// entries listed for its current orientation, 2 (perm[0]-1) + (perm[1] > perm[2]).
static int static_cells(MGArray *phi, MGArray *statics, const double *compOffset, hipStream_t s)
{
    if (!statics || !compOffset) return IM_FAIL(ERROR_NULL_POINTER, "null static-cell table");
    if (phi->nGPUs != 1 || statics->nGPUs != 1 || !statics->devicePtr[0])
        return IM_FAIL(ERROR_NOIMPLEMENT, "single-GPU static-cell tables only");
    const int *perm = phi->currentPermutation;
    const int oi = 2 * (perm[0] - 1) + (perm[1] > perm[2] ? 1 : 0);
    if (oi < 0 || oi > 5) return IM_FAIL(ERROR_INVALID_ARGS, "array permutation %d %d %d", perm[0], perm[1], perm[2]);
    const int64_t off = (int64_t)compOffset[2 * oi];
    const int64_t n = (int64_t)compOffset[2 * oi + 1];
    const int64_t blk = statics->dim[0];
    if (n <= 0) return SUCCESSFUL;
    if (off < 0 || off + n > blk || statics->numel < 3 * blk || n > INT32_MAX)
        return IM_FAIL(ERROR_INVALID_ARGS, "static cells %lld..%lld outside a table of %lld rows", (long long)off,
                       (long long)(off + n), (long long)blk);
    return launch_static_fade(phi->devicePtr[0], statics->devicePtr[0] + off, (int)n, blk, phi->numel, s);
}

// the static cells of an array whose externalData lists some
static int array_statics(MGArray *phi, hipStream_t s)
{
    const ImogenBoundaryData *bd = (const ImogenBoundaryData *)phi->boundaryConditions.externalData;
    if (!bd || !bd->statics) return SUCCESSFUL;
    return static_cells(phi, bd->statics, bd->compOffset, s);
}

// setFreeBalanceCondition's host constants (cudaStatics.cu:710-745) for one face
static FreeBalanceConsts freebalance_consts(const MGArray *fluid, const GeometryParams *geom, int side, int depth,
                                            double gamma)
{
    FreeBalanceConsts fc{};
    const int nx0 = fluid->dim[0];
    const double h0 = geom->h[0];
    const double rc = side ? geom->Rinner + (nx0 - 1 - depth) * h0 : geom->Rinner + depth * h0;
    fc.vFrame = geom->frameOmega * rc;
    for (int k = 0; k < depth; k++) {
        const double r = side ? geom->Rinner + (nx0 - depth + k) * h0 : geom->Rinner + k * h0;
        const double q = r / rc;
        fc.x0[k] = (q - 1) / q;
        fc.qh[k] = pow(q, -0.5);
        fc.qq[k] = q;
    }
    fc.invgm1 = 1 / (gamma - 1);
    return fc;
}

// Freebalance face of the fluid whose density array is `rho` (the reference
// reads only that array's struct and reaches the other four slabs through its
// slab pitch, cudaStatics.cu:780-842): the owner of the 5-slab block or its
// slab-0 view.
static int freebalance_face(MGArray *rho, const GeometryParams *geom, int side, int direction,
                            const MGArray *gravpot, double gamma, hipStream_t s)
{
    MGArray *fluid = rho;
    if (fluid->nGPUs != 1 || !fluid->devicePtr[0]) return IM_FAIL(ERROR_NOIMPLEMENT, "single-GPU arrays only");
    if (!(fluid->numSlabs == 0 || fluid->numSlabs >= 5))
        return IM_FAIL(ERROR_INVALID_ARGS, "freebalance needs the density array of a 5-slab fluid (numSlabs %d)",
                       fluid->numSlabs);
    double *base = fluid->devicePtr[0];
    const int64_t pitch = fluid->slabPitch[0] / (int64_t)sizeof(double);
    if (pitch < fluid->numel) return IM_FAIL(ERROR_INVALID_ARGS, "slab pitch below the array size");
    if (!geom) return IM_FAIL(ERROR_NULL_POINTER, "freebalance needs the geometry");
    if (!gravpot || gravpot->nGPUs != 1 || !gravpot->devicePtr[0])
        return IM_FAIL(ERROR_INVALID_ARGS, "freebalance needs a single-GPU gravity potential array");
    if (!is_xyz(fluid) || !is_xyz(gravpot))
        return IM_FAIL(ERROR_NOIMPLEMENT, "freebalance boundaries of rotated arrays are not built");
    for (int a = 0; a < 3; a++)
        if (gravpot->dim[a] != fluid->dim[a])
            return IM_FAIL(ERROR_INVALID_ARGS, "gravity potential %d x %d x %d, fluid %d x %d x %d", gravpot->dim[0],
                           gravpot->dim[1], gravpot->dim[2], fluid->dim[0], fluid->dim[1], fluid->dim[2]);
    if (direction < 1 || direction > 3 || side < 0 || side > 1)
        return IM_FAIL(ERROR_INVALID_ARGS, "direction %d side %d", direction, side);
    if (!(gamma > 1.0)) return IM_FAIL(ERROR_INVALID_ARGS, "freebalance needs gamma > 1 (got %g)", gamma);
    const int depth = fluid->haloSize > 0 ? fluid->haloSize : 4;
    if (depth > kMaxFreeBalanceDepth)
        return IM_FAIL(ERROR_NOIMPLEMENT, "freebalance halo depth %d (at most %d)", depth, kMaxFreeBalanceDepth);
    const FreeBalanceConsts fc = freebalance_consts(fluid, geom, side, depth, gamma);
    const int *d3 = fluid->dim;
    return launch_freebalance_bc(base, pitch, gravpot->devicePtr[0], d3[0], d3[1], d3[2], direction, side, depth, fc,
                                 s);
}

int fluid_boundary(MGArray *fluid, int direction, int xrules, const GeometryParams *geo)
{
    double *base;
    int64_t pitch;
    int rc = fluid_view(fluid, &base, &pitch);
    if (rc) return rc;
    if (direction < 0 || direction > 3)
        return IM_FAIL(ERROR_INVALID_ARGS, "direction %d is not one of 1 (X) 2 (Y) 3 (Z)", direction);
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    bool statics = false;
    for (int i = 0; i < 5; i++) {
        const ImogenBoundaryData *bd = (const ImogenBoundaryData *)fluid[i].boundaryConditions.externalData;
        statics = statics || (bd && bd->statics);
    }
    if (direction == 0) {  // statics only (cudaStatics.cu:209-211)
        for (int i = 0; i < 5 && !rc; i++) rc = array_statics(&fluid[i], st->stream);
        return rc;
    }
    int modes10[10];
    bool out[2] = {false, false}, fb[2] = {false, false};
    for (int side = 0; side < 2; side++) {
        // the density array's mode decides (cudaStatics.cu:140-152)
        const BCModeTypes bt = fluid[0].boundaryConditions.mode[2 * (direction - 1) + side];
        out[side] = (bt == outflow);
        fb[side] = (bt == freebalance);
        for (int i = 0; i < 5; i++) {
            if (out[side] || fb[side]) {
                modes10[2 * i + side] = -1;
                continue;
            }
            rc = sas_code(&fluid[i], direction, side, &modes10[2 * i + side]);
            if (rc) return rc;
        }
    }
    const ImogenBoundaryData *bd0 = (const ImogenBoundaryData *)fluid[0].boundaryConditions.externalData;
    if ((fb[0] || fb[1]) && (!bd0 || !bd0->gravpot))
        return IM_FAIL(ERROR_INVALID_ARGS, "freebalance needs the gravity potential: the density array's "
                                           "boundaryConditions.externalData must point to an ImogenBoundaryData");
    // modes and mirror signs follow the physical direction, the kernels the
    // memory axis that holds it (cudaStatics.cu:230)
    const int memdir = is_xyz(fluid) ? direction : MGA_dir2memdir(fluid->currentPermutation, direction);
    if (memdir < 1) return IM_FAIL(ERROR_INVALID_ARGS, "no memory axis holds direction %d", direction);
    if ((out[0] || out[1]) && !is_xyz(fluid))
        return IM_FAIL(ERROR_NOIMPLEMENT, "outflow boundaries of rotated arrays are not built");
    const int depth = fluid->haloSize > 0 ? fluid->haloSize : 4;
    const int *d3 = fluid->dim;
    const int slot = prof_begin(PROF_BC, st->stream);
    if (!statics && !fb[0] && !fb[1] && !out[0] && !out[1]) {
        rc = launch_fluid_bc(base, pitch, d3[0], d3[1], d3[2], memdir, modes10, depth, st->stream, xrules);
    } else if (!statics && !fb[0] && !fb[1] && out[0] && out[1] && d3[direction - 1] >= 2 * depth + 2) {
        rc = launch_outflow_bc(base, pitch, d3[0], d3[1], d3[2], direction, 0, 2, depth, st->stream);
    } else {
        // the minus face, then the plus face (setFluidBoundary's side loop),
        // each after the static cells of every array (:141-189)
        rc = SUCCESSFUL;
        for (int side = 0; side < 2 && !rc; side++) {
            for (int i = 0; i < 5 && !rc; i++) rc = array_statics(&fluid[i], st->stream);
            if (rc) break;
            if (fb[side]) {
                rc = freebalance_face(fluid, geo, side, direction, bd0->gravpot, bd0->gamma, st->stream);
            } else if (out[side]) {
                rc = launch_outflow_bc(base, pitch, d3[0], d3[1], d3[2], direction, side, 1, depth, st->stream);
            } else {
                int one[10];
                for (int k = 0; k < 10; k++) one[k] = ((k & 1) == side) ? modes10[k] : -1;
                rc = launch_fluid_bc(base, pitch, d3[0], d3[1], d3[2], memdir, one, depth, st->stream, xrules);
            }
        }
    }
    prof_end(slot, st->stream);
    return rc;
}

int update_1d(MGArray *fluid, const FluidStepParams &params, const double *lampBase, const CflFuse *cfl)
{
    double *base;
    int64_t pitch;
    int rc = fluid_view(fluid, &base, &pitch);
    if (rc) return rc;
    if (!is_xyz(fluid)) {
        if (lampBase || cfl) return IM_FAIL(ERROR_NOIMPLEMENT, "device-loop sweeps need XYZ-ordered arrays");
        return on_xyz_copy(fluid, [&](MGArray *xf) { return update_1d(xf, params, nullptr, nullptr); });
    }
    if (params.geometry.shape != SQUARE)
        return IM_FAIL(ERROR_NOIMPLEMENT, "only SQUARE (cartesian) geometry is built");
    if (params.onlyHydro != 1) return IM_FAIL(ERROR_NOIMPLEMENT, "MHD is not built (dead in the reference)");
    const int dir = params.stepDirection;
    if (dir < 1 || dir > 3) return IM_FAIL(ERROR_INVALID_ARGS, "stepDirection %d", dir);
    const double lambda = params.dt / params.geometry.h[dir - 1];
    const double *lamp = lampBase ? lampBase + (dir - 1) : nullptr;
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    switch (params.stepMethod) {
    case METHOD_HLLC:
    case METHOD_HLL:
        return launch_fluid_sweep(base, pitch, fluid->dim[0], fluid->dim[1], fluid->dim[2], dir,
                                  (int)params.stepMethod, lambda, params.thermoGamma, params.minimumRho,
                                  params.Cisothermal, params.slopeLimiter, st->stream, lamp, cfl);
    case METHOD_XINJIN:
        if (cfl) return IM_FAIL(ERROR_NOIMPLEMENT, "the CFL is not folded into Xin-Jin sweeps");
        return launch_xinjin_sweep(base, pitch, fluid->dim[0], fluid->dim[1], fluid->dim[2], dir, lambda,
                                   params.thermoGamma, params.minimumRho, st->stream, lamp);
    default: return IM_FAIL(ERROR_INVALID_ARGS, "unknown fluid method %d", (int)params.stepMethod);
    }
}

// flux.cu:27-155 on one rank: the sweep order of (stepNumber, stepDirection),
// each sweep followed by that direction's boundary conditions.
int last_sweep_dir(const MGArray *fluid, int stepNumber, int stepDirection)
{
    if (fluid->dim[1] == 1 && fluid->dim[2] == 1) return 1;
    int sweep = (stepNumber + 3 * (stepDirection > 0)) % 6;
    if (sweep < 0) sweep += 6;
    int last = 0;
    for (int n = 0; n < 3; n++) {
        const int d = kFluxcall[n][sweep];
        if (fluid->dim[d - 1] > 3) last = d;
    }
    return last;
}

int cfl_keep_range(MGArray *fluid, int dir, int *lo, int *hi)
{
    const int N = fluid->dim[dir - 1];
    const int depth = fluid->haloSize > 0 ? fluid->haloSize : 4;
    *lo = 0;
    *hi = N;
    // static cells change interior cells after the sweep
    for (int i = 0; i < 5; i++) {
        const ImogenBoundaryData *bd = (const ImogenBoundaryData *)fluid[i].boundaryConditions.externalData;
        if (bd && bd->statics) return 1;
    }
    for (int side = 0; side < 2; side++) {
        const BCModeTypes bt = fluid[0].boundaryConditions.mode[2 * (dir - 1) + side];
        if (bt == outflow || bt == freebalance) return 1;
        int code0 = -2;
        for (int i = 0; i < 5; i++) {
            int code;
            if (sas_code(&fluid[i], dir, side, &code)) return 1;
            if (code == 3) return 1;  // linear extrapolation makes new values
            if (code0 != -2 && ((code < 0) != (code0 < 0))) return 1;
            code0 = code;
        }
        if (code0 < 0) continue;  // circular: the sweep's values stay
        if (N < 2 * depth + 2) return 1;
        // a mirror face copies cells up to 2 depth from the face: they must lie
        // in the kept range
        if ((code0 == 0 || code0 == 1) && N <= 3 * depth) return 1;
        // the fluid step's boundaries (x conventions along every direction)
        // overwrite the ghosts [0, depth) and [N - depth, N) with copies or
        // mirror images of cells in [depth, N - depth)
        if (side == 0) *lo = depth;
        else *hi = N - depth;
    }
    return 0;
}

int update_3d(MGArray *fluid, const FluidStepParams &fsp, const double *lampBase, const CflFuse *lastCfl)
{
    int rc;
    if (!fluid) return IM_FAIL(ERROR_NULL_POINTER, "null fluid");
    if (!is_xyz(fluid)) {
        if (lampBase || lastCfl) return IM_FAIL(ERROR_NOIMPLEMENT, "device-loop sweeps need XYZ-ordered arrays");
        return on_xyz_copy(fluid, [&](MGArray *xf) { return update_3d(xf, fsp, nullptr, nullptr); });
    }
    FluidStepParams p = fsp;
    const int last = lastCfl ? last_sweep_dir(fluid, fsp.stepNumber, fsp.stepDirection) : 0;
    CflFuse cf;
    if (lastCfl) {
        cf = *lastCfl;
        if (last == 0 || cfl_keep_range(fluid, last, &cf.lo, &cf.hi))
            return IM_FAIL(ERROR_INVALID_ARGS, "this step's boundaries do not allow a folded CFL");
    }
    if (fluid->dim[1] == 1 && fluid->dim[2] == 1) {  // flux.cu:57-66
        p.stepDirection = 1;
        rc = update_1d(fluid, p, lampBase, lastCfl ? &cf : nullptr);
        if (rc) return rc;
        return fluid_boundary(fluid, 1, 1, &fsp.geometry);
    }
    int sweep = (fsp.stepNumber + 3 * (fsp.stepDirection > 0)) % 6;
    if (sweep < 0) sweep += 6;
    for (int n = 0; n < 3; n++) {
        const int d = kFluxcall[n][sweep];
        if (fluid->dim[d - 1] > 3) {  // flux.cu:102
            p.stepDirection = d;
            rc = update_1d(fluid, p, lampBase, (lastCfl && d == last) ? &cf : nullptr);
            if (rc) return rc;
            // the reference sets these boundaries with direction d rotated into
            // memory x (flux.cu:99-107): the x kernels' conventions along d
            rc = fluid_boundary(fluid, d, 1, &fsp.geometry);
            if (rc) return rc;
        }
    }
    return SUCCESSFUL;
}

}  // namespace imogen

extern "C" {

int setFluidBoundary(MGArray *fluid, GeometryParams *geo, int direction)
{
    return fluid_boundary(fluid, direction, 0, geo);
}

int setFreeBalanceCondition(MGArray *fluid, GeometryParams *geom, int rightside, int direction, MGArray *gravpot,
                            double fluidGamma)
{
    if (!fluid || !geom || !gravpot) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    return freebalance_face(fluid, geom, rightside, direction, gravpot, fluidGamma, st->stream);
}

int applyStaticCells(MGArray *phi, MGArray *statics, const double *compOffset)
{
    if (!phi || !statics || !compOffset) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    return static_cells(phi, statics, compOffset, st->stream);
}

int setArrayBoundaryConditions(MGArray *phi, GeometryParams *geo, int direction, int side)
{
    if (!phi) return IM_FAIL(ERROR_NULL_POINTER, "null array");
    if (phi->nGPUs != 1) return IM_FAIL(ERROR_NOIMPLEMENT, "single-GPU arrays only");
    if (direction < 0 || direction > 3 || side < 0 || side > 1)
        return IM_FAIL(ERROR_INVALID_ARGS, "direction %d side %d", direction, side);
    {
        DeviceState *st0 = state();
        if (!st0) return ERROR_CUDA_BLEW_UP;
        int rs = array_statics(phi, st0->stream);  // cudaStatics.cu:208
        if (rs || direction == 0) return rs;
    }
    if (phi->boundaryConditions.mode[2 * (direction - 1) + side] == freebalance) {
        // set through the density array alone (numSlabs 0, :259-281); the
        // reference leaves the face as it is without a potential (:264)
        const ImogenBoundaryData *bd = (const ImogenBoundaryData *)phi->boundaryConditions.externalData;
        if (phi->numSlabs != 0 || !bd || !bd->gravpot) return SUCCESSFUL;
        DeviceState *st1 = state();
        return freebalance_face(phi, geo, side, direction, bd->gravpot, bd->gamma, st1->stream);
    }
    int code;
    int rc = sas_code(phi, direction, side, &code);
    if (rc) return rc;
    int modes10[10];
    for (int i = 0; i < 10; i++) modes10[i] = -1;
    modes10[side] = code;  // slab 0 of a one-slab launch at phi's own pointer
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    const int memdir = is_xyz(phi) ? direction : MGA_dir2memdir(phi->currentPermutation, direction);
    if (memdir < 1) return IM_FAIL(ERROR_INVALID_ARGS, "no memory axis holds direction %d", direction);
    // A single array is addressed as slab 0; the other 4 table entries are inert.
    return launch_fluid_bc(phi->devicePtr[0], 0, phi->dim[0], phi->dim[1], phi->dim[2], memdir, modes10,
                           phi->haloSize > 0 ? phi->haloSize : 4, st->stream);
}

int performFluidUpdate_1D(MGArray *fluid, FluidStepParams params, ParallelTopology *topo, MGArray *tmpst)
{
    (void)tmpst;
    int rc = check_topology(topo);
    if (rc) return rc;
    return imogen::update_1d(fluid, params, nullptr);
}

int performFluidUpdate_3D(MGArray *fluid, ParallelTopology *parallelTopo, FluidStepParams fsp,
                          MGArray *tempStorage)
{
    (void)tempStorage;
    int rc = check_topology(parallelTopo);
    if (rc) return rc;
    return imogen::update_3d(fluid, fsp, nullptr);
}

int performFluidUpdate_3D_noalloc(MGArray *fluid, ParallelTopology *parallelTopo, FluidStepParams fsp)
{
    int rc = check_topology(parallelTopo);
    if (rc) return rc;
    double *base;
    int64_t pitch;
    rc = fluid_view(fluid, &base, &pitch);
    if (rc) return rc;
    if (!is_xyz(fluid))
        return IM_FAIL(ERROR_INVALID_ARGS,
                       "the allocation-free fluid step takes XYZ-ordered arrays (a rotated canister is stepped "
                       "on a transposed copy, which allocates)");
    const int method = (int)fsp.stepMethod;
    if (method != METHOD_HLLC && method != METHOD_HLL && method != METHOD_XINJIN)
        return IM_FAIL(ERROR_INVALID_ARGS, "unknown fluid method %d", method);
    // every sweep update_3d will launch must fit the fixed scratch array
    const int *d3 = fluid->dim;
    const bool line = d3[1] == 1 && d3[2] == 1;
    int64_t need = 0;
    for (int d = 1; d <= 3; d++) {
        if (line ? d != 1 : d3[d - 1] <= 3) continue;
        const int64_t b = (method == METHOD_XINJIN) ? xinjin_scratch_bytes(pitch, d3[0], d3[1], d3[2], d)
                                                    : sweep_scratch_bytes(d3[0], d3[1], d3[2], d);
        if (b > need) need = b;
    }
    if (need > (int64_t)kFixedScratchBytes)
        return IM_FAIL(ERROR_NOMEM,
                       "a %d x %d x %d fluid step needs %lld bytes of scratch; the allocation-free step has a fixed "
                       "%zu (use performFluidUpdate_3D)",
                       d3[0], d3[1], d3[2], (long long)need, kFixedScratchBytes);
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    st->fixedScratchOn = 1;
    rc = imogen::update_3d(fluid, fsp, nullptr);
    st->fixedScratchOn = 0;
    return rc;
}

int calculateSoundspeed(MGArray *fluid, MGArray *mag, MGArray *cs, double gamma)
{
    if (mag) return IM_FAIL(ERROR_NOIMPLEMENT, "MHD sound speed is not built");
    double *base;
    int64_t pitch;
    int rc = fluid_view(fluid, &base, &pitch);
    if (rc) return rc;
    if (!cs || cs->nGPUs != 1 || cs->numel != fluid->numel)
        return IM_FAIL(ERROR_INVALID_ARGS, "sound speed array does not match the fluid");
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    return launch_soundspeed(base, pitch, fluid->partNumel[0], gamma, cs->devicePtr[0], st->stream);
}

static int cfl_common(MGArray *fluid, MGArray *csound, GeometryParams *geom, int method, double gamma,
                      double *globalResolution, double *tstep)
{
    double *base;
    int64_t pitch;
    int rc = fluid_view(fluid, &base, &pitch);
    if (rc) return rc;
    if (!geom || !globalResolution || !tstep) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    if (geom->shape != SQUARE) return IM_FAIL(ERROR_NOIMPLEMENT, "only SQUARE geometry is built");
    if (method < 1 || method > 3) return IM_FAIL(ERROR_INVALID_ARGS, "method %d", method);
    int spacedim = 0;
    if (globalResolution[1] > 1) spacedim = 1;
    if (globalResolution[2] > 1) spacedim = 2;
    const int dim = spacedim + 1;
    double hmin = geom->h[0];
    if (dim > 1 && geom->h[1] < hmin) hmin = geom->h[1];
    if (dim > 2 && geom->h[2] < hmin) hmin = geom->h[2];
    const int xj = (method == METHOD_XINJIN) && dim > 1;  // the XJ-1D table entry is the HLLC form
    const int halve = (method == METHOD_HLL || method == METHOD_HLLC);
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    const int slot = prof_begin(PROF_CFL, st->stream);
    rc = launch_cfl(base, pitch, fluid->partNumel[0], csound ? csound->devicePtr[0] : nullptr, gamma,
                    csound ? 0 : 1, dim, xj, halve, geom->h[0], geom->h[1], geom->h[2], hmin, st->devScalars,
                    st->stream);
    prof_end(slot, st->stream);
    if (rc) return rc;
    rc = pinned_buffers(st);
    if (rc) return rc;
    rc = IM_HIP(hipMemcpyAsync(st->pinned, st->devScalars, sizeof(double), hipMemcpyDeviceToHost, st->stream));
    if (rc) return rc;
    rc = IM_HIP(hipStreamSynchronize(st->stream));
    if (rc) return rc;
    tstep[0] = st->pinned[0];
    return take_device_errors(st);
}

int computeLocalCFLTimestep(MGArray *fluid, MGArray *csound, GeometryParams *geom, int method,
                            double *globalResolution, double *tstep)
{
    if (!csound) return IM_FAIL(ERROR_NULL_POINTER, "null sound speed array");
    if (csound->nGPUs != 1 || csound->numel != fluid->numel)
        return IM_FAIL(ERROR_INVALID_ARGS, "sound speed array does not match the fluid");
    return cfl_common(fluid, csound, geom, method, 0.0, globalResolution, tstep);
}

int imogen_cflTimestepFused(MGArray *fluid, GeometryParams *geom, int method, double gamma,
                            double *globalResolution, double *tstep)
{
    return cfl_common(fluid, nullptr, geom, method, gamma, globalResolution, tstep);
}

}  // extern "C"
