// imogen_internal.h -- shared host-side plumbing of libimogen_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/imogen_abi.h"

namespace imogen {

struct CflFuse;  // cfl_device.h


// Records a HIP failure (with where it happened) and returns ERROR_CUDA_BLEW_UP,
// mirroring CHECK_CUDA_ERROR of the reference (gpuclass/cudaCommon.h:331-337).
int check_hip(hipError_t e, const char *what, const char *file, int line);
// Records an Imogen status (non-zero) with a message and passes it through.
int fail(int status, const char *file, int line, const char *fmt, ...);

#define IM_HIP(call) ::imogen::check_hip((call), #call, __FILE__, __LINE__)
#define IM_LAUNCHED(name) ::imogen::check_hip(hipGetLastError(), name, __FILE__, __LINE__)
#define IM_FAIL(status, ...) ::imogen::fail((status), __FILE__, __LINE__, __VA_ARGS__)

// Counters of invalid states that kernels meet (the device error word):
// kernels add to them, take_device_errors turns them into ERROR_CRASH.
enum DeviceErrorSlot {
    ERRW_DRAG_INVALID = 0,  // drag cells with a density, |dv| or (moving) internal energy not positive/finite
    ERRW_DRAG_BUDGET = 1,   // cells whose adaptive drag integration ran out of steps
    ERRW_CFL_NAN = 2,       // cells whose CFL limit is NaN (a standalone CFL pass, the ensemble's folded CFL)
    ERRW_TIMESTEP = 3,      // time steps not taken: tau not positive and finite (device clocks)
    ERRW_SLOTS = 8
};

// Size of the fixed per-device scratch array that the allocation-free entry
// points use in place of the growable scratch block.
constexpr size_t kFixedScratchBytes = (size_t)64 << 20;

// Per-device state: one stream all kernels go to, a growable scratch block,
// the device addresses of the library's static device arrays, and small
// pinned host buffers for scalar read-back.
struct DeviceState {
    int device = -1;
    hipStream_t stream = nullptr;
    void *scratch = nullptr;
    size_t scratchBytes = 0;
    double *fixedScratch = nullptr;  // this device's copy of the fixed scratch array (kFixedScratchBytes)
    int fixedScratchOn = 0;          // scratch_block hands out fixedScratch (never allocates) while set
    double *pinned = nullptr;   // 64 doubles of pinned host memory (pinned_buffers)
    double *devScalars = nullptr;  // 4096 doubles of device memory for reductions
    unsigned long long *err = nullptr;      // ERRW_SLOTS device counters, zero when no error is pending
    unsigned long long *errHost = nullptr;  // their pinned host copy (pinned_buffers)
};

// State of the current device, initialised on first use: the stream and the
// device addresses of the static device arrays. Allocates no memory.
DeviceState *state();
// Allocates the pinned host buffers (pinned, errHost) on first use.
int pinned_buffers(DeviceState *st);
// Waits for the stream, reads the device error counters and, when any is
// set, clears them and returns ERROR_CRASH naming what was counted. Never
// call it while the stream is being captured.
int take_device_errors(DeviceState *st);

// Optional per-kernel timing with HIP events on the launch stream
// (imogen_profile_enable / imogen_profile_read). prof_begin returns a slot
// (or -1 when timing is off or the event pool is exhausted).
enum ProfKernel { PROF_SWEEP_X = 0, PROF_SWEEP_Y = 1, PROF_SWEEP_Z = 2, PROF_BC = 3, PROF_CFL = 4,
                  PROF_DRAG = 5, PROF_MLP = 6, PROF_NKERNELS = 8 };
int prof_begin(int kernel, hipStream_t s);
void prof_end(int slot, hipStream_t s);
// Nested suspension of the event timing (kept off while a graph is captured).
void prof_suspend(int on);
// Device scratch of at least `bytes`; contents are undefined on return.
// While st->fixedScratchOn is set it returns the fixed array (null with an
// ERROR_NOMEM message when bytes is above kFixedScratchBytes) and never
// allocates.
void *scratch_block(size_t bytes);
// Scratch bytes launch_fluid_sweep / launch_xinjin_sweep take for one sweep
// (the launch shapes of fluid_sweep.hip and xinjin.hip).
int64_t sweep_scratch_bytes(int nx, int ny, int nz, int dir);
int64_t xinjin_scratch_bytes(int64_t pitch, int nx, int ny, int nz, int dir);

// Validates a 5-array fluid canister (slab views of one 5-slab owner, XYZ
// order, one GPU) and returns its base pointer and slab pitch in doubles.
int fluid_view(MGArray *fluid, double **base, int64_t *pitch);
// ERROR_NOIMPLEMENT for topologies of more than one rank (null: one rank).
int check_topology(const ParallelTopology *t);
// Whether an array is in XYZ memory order (permtag 0 or 1: never rotated, or
// rotated back). fluid_view accepts rotated canisters (all five arrays in the
// same order); the sweeps and boundaries map directions through
// currentPermutation.
bool is_xyz(const MGArray *a);
// Copies an sdim[0] x sdim[1] x sdim[2] array so that destination axis a walks
// source axis ax[a] (rotate.hip).
int launch_permute_copy(const double *src, double *dst, const int sdim[3], const int ax[3], hipStream_t s);

// performFluidUpdate_1D / _3D without the topology check. lampBase: when not
// null, the sweeps of direction d read lambda from lampBase[d - 1] (device).
// lastCfl (update_3d): when not null, the last sweep of the step also leaves
// per-wave CFL minima (its lo/hi are filled in from the boundary modes).
int update_1d(MGArray *fluid, const FluidStepParams &params, const double *lampBase,
              const CflFuse *cfl = nullptr);
int update_3d(MGArray *fluid, const FluidStepParams &fsp, const double *lampBase, const CflFuse *lastCfl = nullptr);
// setFluidBoundary with a choice of kernel conventions (launch_fluid_bc's
// xrules); update_3d uses xrules = 1 after every sweep.
int fluid_boundary(MGArray *fluid, int direction, int xrules, const GeometryParams *geo);
// Direction swept last by update_3d for (stepNumber, stepDirection), 0 if none.
int last_sweep_dir(const MGArray *fluid, int stepNumber, int stepDirection);
// Cells [lo, hi) of direction dir whose swept values survive the boundary
// conditions update_3d sets after the sweep, when every other ghost they write
// is a copy or mirror image of one of them (modes circular, mirror, const).
// Returns non-zero when the boundaries do not allow folding the CFL into the
// sweep.
int cfl_keep_range(MGArray *fluid, int dir, int *lo, int *hi);

// Launchers (fluid_sweep.hip, xinjin.hip, statics.hip, cfl.hip, drag.hip)
// method: METHOD_HLLC or METHOD_HLL. rhomin, ciso: HLL floors and isothermal
// sound speed (-1: adiabatic). lamp: when not null, the kernels read lambda
// = dt/h from this device address instead of the `lambda` argument.
// cfl: when not null, the sweep also writes the per-wave CFL minima of the
// cells it stores (cfl_device.h).
int launch_fluid_sweep(double *base, int64_t pitch, int nx, int ny, int nz, int dir, int method, double lambda,
                       double gamma, double rhomin, double ciso, int limiter, hipStream_t s,
                       const double *lamp = nullptr, const CflFuse *cfl = nullptr);
int launch_xinjin_sweep(double *base, int64_t pitch, int nx, int ny, int nz, int dir, double lambda,
                        double gamma, double rhomin, hipStream_t s, const double *lamp = nullptr);
// xrules: apply the memory-x kernels' conventions along direction dir (the
// boundaries the fluid step sets after each sweep); 0: each memory axis's own
// kernels (setFluidBoundary / setArrayBoundaryConditions).
int launch_fluid_bc(double *base, int64_t pitch, int nx, int ny, int nz, int dir, const int *modes10,
                    int depth, hipStream_t s, int xrules = 0);
// Outflow ghosts of faces side0 .. side0 + nsides - 1 of direction dir (all 5
// slabs; both faces together only when they are disjoint, N >= 2 depth + 2).
int launch_outflow_bc(double *base, int64_t pitch, int nx, int ny, int nz, int dir, int side0, int nsides,
                      int depth, hipStream_t s);
// Freebalance ghosts of one face (side 0/1) of direction dir: the host-side
// constants of setFreeBalanceCondition (cudaStatics.cu:710-745) per ghost k.
constexpr int kMaxFreeBalanceDepth = 8;
struct FreeBalanceConsts {
    double x0[kMaxFreeBalanceDepth];  // (q-1)/q, q = ghost radius / reference cell radius
    double qh[kMaxFreeBalanceDepth];  // q^-1/2
    double qq[kMaxFreeBalanceDepth];  // q
    double vFrame;                    // frame velocity at the reference cell
    double invgm1;                    // 1 / (gamma - 1)
};
int launch_freebalance_bc(double *base, int64_t pitch, const double *phi, int nx, int ny, int nz, int dir,
                          int side, int depth, const FreeBalanceConsts &fc, hipStream_t s);
// Static-cell fade of nspec entries of table (columns of blk doubles) on an
// array of numel cells.
int launch_static_fade(double *phi, const double *table, int nspec, int64_t blk, int64_t numel, hipStream_t s);
int launch_soundspeed(const double *base, int64_t pitch, int64_t n, double gamma, double *cs,
                      hipStream_t s);
int launch_cfl(const double *base, int64_t pitch, int64_t n, const double *cs, double gamma, int fused,
               int dim, int xinjin, int halve, double hx, double hy, double hz, double hmin, double *devOut,
               hipStream_t s);
// devOut[0] = min(part[0..n)) (x 1/2 when halve): the second pass over
// per-wave minima a fused sweep left
int launch_cfl_final(const double *part, int64_t n, int halve, double *devOut, hipStream_t s);
// Wave count of a sweep launch over an nx x ny x nz HLLC/HLL array in
// direction dir (the entries a fused CFL needs).
int64_t sweep_waves(int nx, int ny, int nz, int dir);

}  // namespace imogen
