/*
 * imogen_oracle.h -- CPU restatement (FP64, no FMA contraction) of the
 * Imogen fluid-step arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it. The product library (libimogen_hip.so) never links it and has
 * no CPU fallback.
 *
 * Array convention (identical to the reference's 5-slab fluid canister,
 * save/uploadDataArrays.m:93-111): one allocation of 5 slabs
 *   slab 0 = rho, 1 = E (total energy), 2 = px, 3 = py, 4 = pz
 * each slab nx*ny*nz doubles with x fastest, slabs `pitch` doubles apart.
 */
#ifndef IMOGEN_ORACLE_H
#define IMOGEN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Slope limiter selector (gpuclass/limiter_functions.h). 0 is the
 * reference's compiled-in default SLOPEFUNC = slopeLimiter_Ospre
 * (gpuclass/cudaFluidStep.cu:132). */
enum orc_limiter { ORC_LIM_OSPRE = 0, ORC_LIM_VANLEER = 1, ORC_LIM_MINMOD = 2,
                   ORC_LIM_SUPERBEE = 3, ORC_LIM_VANALBADA = 4, ORC_LIM_MC = 5,
                   ORC_LIM_ZERO = 6 };

/* Fluid methods, gpuclass/cudaFluidStep.h:6-8 */
enum orc_method { ORC_HLL = 1, ORC_HLLC = 2, ORC_XINJIN = 3 };

/* BC modes, gpuclass/cudaCommon.h:49 */
enum orc_bcmode { ORC_BC_CIRCULAR = 0, ORC_BC_MIRROR = 1, ORC_BC_WALL = 2, ORC_BC_STATIONARY = 3,
                  ORC_BC_CONST = 4, ORC_BC_LINEAR = 5, ORC_BC_OUTFLOW = 6, ORC_BC_FREEBALANCE = 7 };

typedef struct {
    double gamma;     /* thermoGamma (already forced to 2 by the caller for isothermal fluids) */
    double rhomin;    /* fluid MINMASS */
    double ciso;      /* isothermal sound speed, -1 = adiabatic */
    int limiter;      /* enum orc_limiter */
} orc_fluid_params;

typedef struct {
    int nx, ny, nz;
    long pitch;       /* doubles between slabs */
} orc_grid;

/* Thread count for the OpenMP loops over lines (<=0: leave default). */
void orc_set_threads(int n);
int  orc_get_threads(void);

/* One directional SSP-RK2 update (predictor + corrector) of all lines along
 * physical direction dir (1=x, 2=y, 3=z), periodic wrap along the line.
 * lambda = dt/h[dir]. Follows performFluidUpdate_1D (gpuclass/cudaFluidStep.cu:233-663).
 * Returns 0 or a negative status. */
int orc_fluid_update_1d(double *q, const orc_grid *g, int dir, int method, double lambda,
                        const orc_fluid_params *p);

/* Xin-Jin first stage only: freezing speed of Q, then the predictor
 * U = floor(Q - lambda/4 dF(Q)) (gpuclass/cudaFluidStep.cu:380-424), written
 * into q. Exported so the tests can pin the predictor on its own. */
int orc_xj_predict_1d(double *q, const orc_grid *g, int dir, double lambda, const orc_fluid_params *p);

/* Boundary conditions for direction dir on both faces. modes[6] = [-x,+x,-y,+y,-z,+z],
 * depth = halo depth (4). Follows setFluidBoundary (gpuclass/cudaStatics.cu:112-196). */
int orc_set_fluid_boundary(double *q, const orc_grid *g, int dir, const int *modes, int depth);
/* outflow ghosts of one face (side 0 = minus, 1 = plus) of all 5 slabs */
int orc_outflow_face(double *q, const orc_grid *g, int dir, int side, int depth);
/* freebalance ghosts of one face (side 0 = minus, 1 = plus) of all 5 slabs:
 * phi = the gravity potential (one slab, same grid), gamma the fluid's
 * adiabatic index, Rinner / h0 / frameOmega / nx0 the geometry the radial
 * (direction 1) balance reads (GeometryParams Rinner, h[0], frameOmega and the
 * fluid's x size) */
int orc_freebalance_face(double *q, const orc_grid *g, const double *phi, int dir, int side, int depth,
                         double gamma, double Rinner, double h0, double frameOmega, int nx0);
/* static cells of one array (setArrayStaticCells's fade): table holds nspec
 * entries per column block of blk doubles: cell index, value f0, coefficient c */
void orc_static_cells(double *phi, const double *table, int nspec, long blk);
/* the boundaries orc_fluidstep sets after sweeping dir: memory-x conventions
 * along dir (flux.cu:99-107 after the rotations of flux.cu:46-48) */
int orc_fluid_boundary_swept(double *q, const orc_grid *g, int dir, const int *modes, int depth);

/* One `fluidstep` call: 3 directional sweeps in the order picked by
 * (stepNumber + 3*(stepDirection>0)) % 6, each followed by its BC.
 * Follows performFluidUpdate_3D (gpuclass/flux.cu:27-155). */
int orc_fluidstep(double *q, const orc_grid *g, const double *h, double dt, int stepDirection,
                  int stepNumber, int method, const orc_fluid_params *p, const int *modes, int depth);

/* Returns the sweep order (3 physical directions) for a sweep index 0..5,
 * derived from the fluxcall table of gpuclass/flux.cu:46. */
void orc_sweep_order(int sweep, int *dirs3);

/* Sound speed, gpuclass/cudaSoundspeed.cu:178-195 */
void orc_soundspeed(const double *q, const orc_grid *g, double gamma, double *cs);

/* CFL time step limit, gpuclass/cflTimestep.cu:107-377 (SQUARE geometry).
 * globalRez selects the dimensionality as the reference does. Returns tau
 * (already halved for HLL/HLLC). */
double orc_cfl_timestep(const double *q, const double *cs, const orc_grid *g, const double *h,
                        int method, const int *globalRez);

/* ---- gas-dust drag (drag_oracle.c) ---- */

/* Microphysics of one fluid; same fields and order as the reference's
 * ThermoDetails (gpuclass/cudaCommon.h:146-158). */
typedef struct {
    double gamma, m, Cisothermal, kBolt, mu0, muTindex, sigma0, sigmaTindex;
} orc_thermo;

/* Drag integrators: the reference's method codes 4/5
 * (cudaSource2FluidDrag.cu:257-262) plus this build's adaptive reference
 * integrator. */
enum orc_drag_method {
    ORC_DRAG_EMP = 0,
    ORC_DRAG_RK4 = 1,
    ORC_DRAG_ETDRK1 = 2,
    ORC_DRAG_LOGTRAP2 = 4,
    ORC_DRAG_LOGTRAP3 = 5,
    ORC_DRAG_EXACT = 6
};

/* The 9 drag constants hostDrag[0..8] (cudaSource2FluidDrag.cu:210-220). */
void orc_drag_params(const orc_thermo *gas, const orc_thermo *dust, double *P9);

/* Drag rate k(|dv|, u) of one cell (:836-890); 0 below |dv| = 1e-9. */
double orc_drag_k(const double *P9, double rhog, double rhod, double magdv, double u);

/* |dv(t)|/|dv(0)| of the relaxation problem, adaptive Dormand-Prince 5(4) on
 * ln|dv| to absolute tolerance tol. steps: step count (negative if the budget
 * of 4096 steps ran out; ORC_DRAG_INVALID, with a NaN result returned at
 * once, for a cell the problem is not defined on). */
#define ORC_DRAG_INVALID (-(1 << 30))
double orc_drag_rk_magdv(const double *P, double rhog, double rhod, double s0, double u0, double dt, int method);
double orc_drag_exact_ratio(const double *P9, double rhog, double rhod, double s0, double u0, double t, double tol,
                            int *steps);

/* One drag source call over the grid: updates gas E, p and dust p in place.
 * Returns 0 or a negative status (-101 when the exact integrator met invalid
 * cells or ran out of steps; every cell is still updated); *work (optional)
 * receives the exact integrator's total step count. */
int orc_drag(double *gas, double *dust, const orc_grid *g, const orc_thermo *tgas, const orc_thermo *tdust,
             double dt, int method, double tol, long *work);

/* Slope limiter evaluation (exported for the limiter tests). */
double orc_slope_limiter(int which, double a, double b);

#ifdef __cplusplus
}
#endif
#endif
