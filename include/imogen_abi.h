/*
 * imogen_abi.h -- C ABI of libimogen_hip.so, the MI355X (gfx950) fluid step.
 *
 * Every entry point keeps the name, argument meaning, ownership rules and
 * status-code convention of the reference's compiled core (the functions its
 * MEX gateways call), so a MEX gateway, a cgo/ctypes binding or a C++ driver
 * binds it the way the reference's own gateways bind the CUDA build. Each
 * declaration cites the reference interface it replaces.
 *
 * Plain C: no HIP, torch or MATLAB types in any signature. Device pointers are
 * carried inside MGArray / int64 tags exactly as in the reference.
 */
#ifndef IMOGEN_ABI_H
#define IMOGEN_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: gpuclass/imogenChecks.h:12-24 ----------------------- */
#define SUCCESSFUL 0
#define ERROR_INVALID_ARGS -100
#define ERROR_CRASH -101
#define ERROR_NULL_POINTER -102
#define ERROR_GET_GPUTAG_FAILED -103
#define ERROR_DESERIALIZE_GPUTAG_FAILED -104
#define ERROR_NOIMPLEMENT -105
#define ERROR_NOMEM -106
#define ERROR_LIBFAILED -107
#define ERROR_NOFILE -108
#define ERROR_FILE_IO -109
#define ERROR_CUDA_BLEW_UP -1000

/* ---- array tag layout: gpuclass/cudaCommon.h:184-196 -------------------- */
#define MAX_GPUS_USED 4
#define GPU_TAG_LENGTH 11
#define GPU_TAG_DIM0 0
#define GPU_TAG_DIM1 1
#define GPU_TAG_DIM2 2
#define GPU_TAG_DIMSLAB 3
#define GPU_TAG_HALO 4
#define GPU_TAG_PARTDIR 5
#define GPU_TAG_NGPUS 6
#define GPU_TAG_EXTERIORHALO 7
#define GPU_TAG_DIMPERMUTATION 8
#define GPU_TAG_CIRCULARBITS 9
#define GPU_TAG_VECTOR_COMPONENT 10
#define GPU_TAG_MAXLEN (GPU_TAG_LENGTH + 2 * MAX_GPUS_USED)

#define PARTITION_X 1
#define PARTITION_Y 2
#define PARTITION_Z 3

/* ---- boundary modes: gpuclass/cudaCommon.h:49 --------------------------- */
typedef enum BCModeTypes {
    circular, mirror, wall, stationary, extrapConstant, extrapLinear, outflow, freebalance
} BCModeTypes;

/* gpuclass/cudaCommon.h:51-60 */
typedef struct BCSettings {
    BCModeTypes mode[6]; /* [-X, +X, -Y, +Y, -Z, +Z] */
    void *externalData;  /* caller-owned (MATLAB class handle in the MEX build); unused here */
    int extIndex;
    double *staticCells[MAX_GPUS_USED];
    int nStatics[4];
} BCSettings;

/* gpuclass/cudaCommon.h:66-118 */
typedef struct MGArray {
    int dim[3];                        /* size excluding partition halos */
    int64_t numel;
    int numSlabs;                      /* >0: owner of N slabs; <=0: view of slab -numSlabs */
    int64_t slabPitch[MAX_GPUS_USED];  /* bytes per slab, rounded up to 256 */
    int haloSize;
    int partitionDir;
    int addExteriorHalo;
    int permtag;
    int currentPermutation[3];
    int nGPUs;
    int deviceID[MAX_GPUS_USED];
    double *devicePtr[MAX_GPUS_USED];
    int partNumel[MAX_GPUS_USED];
    int mpiCircularBoundaryBits;
    BCSettings boundaryConditions;
    int vectorComponent;               /* 0 scalar, 1/2/3 = x/y/z component */
} MGArray;

/* gpuclass/cudaCommon.h:122 */
typedef enum geometryType_t { SQUARE, CYLINDRICAL, RZSQUARE, RZCYLINDRICAL } geometryType_t;

/* gpuclass/cudaCommon.h:122-144 */
typedef struct GeometryParams {
    geometryType_t shape;
    double h[3];
    double x0, y0, z0;
    double Rinner;
    double frameRotateCenter[3];
    double frameOmega;
    int globalRez[3];
    int localRez[3];
    int gridAffine[3];
    MGArray *XYVector;
} GeometryParams;

/* gpuclass/cudaCommon.h:146-158 */
typedef struct ThermoDetails {
    double gamma;
    double m;
    double Cisothermal; /* -1 = adiabatic */
    double kBolt;
    double mu0;
    double muTindex;
    double sigma0;
    double sigmaTindex;
} ThermoDetails;

/* gpuclass/cudaFluidStep.h:6-8 */
typedef enum FluidMethods { METHOD_HLL = 1, METHOD_HLLC = 2, METHOD_XINJIN = 3 } FluidMethods;

/* Slope limiter of the MUSCL reconstruction. The reference selects it at
 * compile time (SLOPEFUNC, gpuclass/cudaFluidStep.cu:127-133; default Ospre);
 * here it is a run-time field of FluidStepParams, 0 = the reference default. */
typedef enum SlopeLimiters {
    LIMITER_OSPRE = 0, LIMITER_VANLEER = 1, LIMITER_MINMOD = 2, LIMITER_SUPERBEE = 3,
    LIMITER_VANALBADA = 4, LIMITER_MC = 5, LIMITER_ZERO = 6
} SlopeLimiters;

/* gpuclass/cudaFluidStep.h:10-26, plus slopeLimiter appended (see above) */
typedef struct FluidStepParams {
    double dt;
    double cflPrefactor;
    int onlyHydro;
    double thermoGamma;
    double minimumRho;
    double Cisothermal;
    FluidMethods stepMethod;
    int stepDirection;  /* performFluidUpdate_1D: 1/2/3 = X/Y/Z; _3D: +1 forward, -1 backward */
    int stepNumber;
    GeometryParams geometry;
    int multifluidDragMethod;
    int slopeLimiter;   /* SlopeLimiters */
} FluidStepParams;

/* mpi/mpi_common.h:14-29 without MPI types. One rank: nproc = {1,1,1}. */
typedef struct ParallelTopology {
    int ndim;
    int comm;
    int nproc[3];
    int coord[3];
    int neighbor_left[3];
    int neighbor_right[3];
    int dimcomm[3];
} ParallelTopology;

/* ======================================================================== */
/* Array-tag runtime (gpuclass/cudaCommon.cu, gpuclass/cudaCommon.h:257-296) */
/* ======================================================================== */

/* cudaCommon.cu:50-91 */
void calcPartitionExtent(MGArray *m, int P, int *sub);
/* cudaCommon.cu:117-172 */
int deserializeTagToMGArray(int64_t *tag, MGArray *mg);
/* cudaCommon.cu:175-195 */
void serializeMGArrayToTag(MGArray *mg, int64_t *tag);
/* cudaCommon.cu:198-212 */
void MGA_permtagToNums(int permtag, int *p);
/* cudaCommon.cu:216-235 */
int MGA_numsToPermtag(int *nums);
/* cudaCommon.cu:237-246 */
int MGA_dir2memdir(int *perm, int dir);
/* cudaCommon.cu:256-335: *ret == NULL -> malloc'd array of N MGArrays */
int MGA_allocArrays(MGArray **ret, int N, MGArray *skeleton);
/* cudaCommon.cu:337-405 */
int MGA_allocSlab(MGArray *skeleton, MGArray *nu, int Nslabs);
/* cudaCommon.cu:407-468: frees owners (numSlabs > 0), ignores slab views */
int MGA_delete(MGArray *victim);
/* cudaCommon.cu:1867-2046: host array (dim[0]*dim[1]*dim[2] doubles) -> device */
int MGA_uploadArrayToGPU(double *p, MGArray *g, int partitionTo);
/* cudaCommon.cu:2048-2217: device -> host; *p == NULL -> malloc'd for the caller */
int MGA_downloadArrayToCPU(MGArray *g, double **p, int partitionFrom);

/* ---- index exchanges and halos (gpuclass/cudaArrayRotateB.cu, cudaHaloExchange.h) */
/* cudaArrayRotateB.cu:583-640: the metadata of an index exchange (dims,
 * currentPermutation, permtag, partitionDir, partNumel); exchangeCode 2 = X<->Y,
 * 3 = X<->Z, 4 = Y<->Z, 5 = rotate left (XYZ -> YZX), 6 = rotate right
 * (XYZ -> ZXY); other codes ERROR_INVALID_ARGS. src may equal dst. */
int alterArrayMetadata(MGArray *src, MGArray *dst, int exchangeCode);
/* cudaArrayRotateB.cu:103-300 (the work of the cudaArrayRotateB MEX,
 * :56-88): transposes each of nArrays arrays in memory by exchangeCode and
 * rewrites its metadata. newArrays == NULL: in place (the arrays keep their
 * pointers; slab views of a fluid canister may be rotated one by one, as
 * flux.m does). newArrays != NULL: new owner arrays are allocated (*newArrays
 * == NULL: a malloc'd vector of nArrays MGArrays for the caller to free).
 * Exchanges that do not change the memory order (a unit dimension) only
 * rewrite the metadata (:551-579). streamPtrs and tempStorage are accepted
 * for the reference's signature and not used. A rotated fluid canister can
 * be passed to every fluid entry point except the device loop: the fluid
 * steps sweep the memory axis that holds the requested physical direction,
 * the boundaries map it through currentPermutation (cudaStatics.cu:230). */
int flipArrayIndices(MGArray *phi, MGArray **newArrays, int nArrays, int exchangeCode, void *streamPtrs,
                     MGArray *tempStorage);
/* cudaHaloExchange.h:13 (cudaHaloExchange.cu:140): one rank has no halos to
 * exchange (periodic lines wrap inside the sweeps); more ranks return
 * ERROR_NOIMPLEMENT. */
int exchange_MPI_Halos(MGArray *phi, int nArrays, ParallelTopology *topo, int xchgDir);

/* ======================================================================== */
/* MEX-level tag primitives (gpuclass/GPU_*.cu) with host buffers and int64  */
/* tags in place of mxArrays. Tags are GPU_TAG_LENGTH + 2 int64.             */
/* ======================================================================== */

/* GPU_upload.cu:20-130: tag = GPU_upload(host, devList, [halo partDir extHalo]) */
int GPU_upload(const double *host, const int *dims3, int haloSize, int partitionDir, int64_t *tagOut);
/* GPU_download.cu:20-84 */
int GPU_download(const int64_t *tag, double *host);
/* GPU_free.cu:15-43 */
int GPU_free(const int64_t *tag);
/* GPU_makeslab.cu:15-72: new owner with N slabs; slab 0 copied from the old array, old freed */
int GPU_makeslab(const int64_t *tag, int N, int64_t *newTag);
/* GPU_getslab.cu:15-47: view of slab k (numSlabs = -k) */
int GPU_getslab(const int64_t *tag, int k, int64_t *viewTag);
/* GPU_setslab.cu:15-81: upload host data into slab k and return its view */
int GPU_setslab(const int64_t *tag, int k, const double *host, int64_t *viewTag);
/* GPU_clone.cu:15-50 */
int GPU_clone(const int64_t *tag, int64_t *newTag);
/* GPU_copy.cu:15-50: dst <- src, identical shapes */
int GPU_copy(const int64_t *dstTag, const int64_t *srcTag);
/* GPU_ctrl.cu 'memory' query */
int GPU_ctrl_memory(int device, size_t *freeBytes, size_t *totalBytes);
/* GPU_ctrl.cu:154-165 'reset'. The reference resets the device; here the
 * library releases its cached temporary storage (nothing cached survives a
 * reset, as the reference requires of temporaries) and caller-owned arrays
 * stay valid. Device loops re-capture their graphs when the temporary block
 * moves. */
int GPU_ctrl_reset(void);
/* GPU_ctrl.cu:19-35, 69-79 'info': a one-line description of the device. */
int GPU_ctrl_info(int device, char *text, int len);

/* ======================================================================== */
/* Fluid step core                                                           */
/* ======================================================================== */

/* gpuclass/flux.h:17, flux.cu:27-155. fluid -> 5 MGArrays (rho E px py pz), the
 * slab views of one 5-slab owner. One `fluidstep`: 3 directional SSP-RK2 sweeps
 * in the order (stepNumber + 3*(stepDirection>0)) % 6, each followed by its
 * boundary conditions. tempStorage may be NULL (scratch is library-cached). */
int performFluidUpdate_3D(MGArray *fluid, ParallelTopology *parallelTopo, FluidStepParams fsp,
                          MGArray *tempStorage);

/* No counterpart in the reference: performFluidUpdate_3D for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). It takes its sweeps' scratch from a fixed
 * 64 MiB device array that each GPU the library runs on holds from the
 * library's first call until the process ends, and calls no hipMalloc/hipFree.
 * Limit: every sweep's scratch must fit 64 MiB (HLLC/HLL: 40 B per X-sweep
 * thread and 200 B per Y/Z marching lane, which fits every cube up to 256^3,
 * e.g. 29.5 MB at 192^3; Xin-Jin: 5 slabs, up to about 1.6 M cells); larger
 * steps return ERROR_NOMEM before any launch. XYZ-ordered canisters only (ERROR_INVALID_ARGS for rotated ones).
 * All calls on a GPU share the array: they are ordered by the library's
 * stream; two executions that overlap on one GPU (e.g. a captured graph of
 * this call replayed on another stream while the library's stream runs one)
 * race. Results are bitwise those of performFluidUpdate_3D. */
int performFluidUpdate_3D_noalloc(MGArray *fluid, ParallelTopology *parallelTopo, FluidStepParams fsp);

/* gpuclass/cudaFluidStep.h:56-58, cudaFluidStep.cu:233-663: one directional
 * predictor+corrector along params.stepDirection (1/2/3 = X/Y/Z). */
int performFluidUpdate_1D(MGArray *fluid, FluidStepParams params, ParallelTopology *topo,
                          MGArray *tmpst);

/* What the reference's boundary code reads from the MATLAB class's
 * boundaryData struct through BCSettings.externalData (cudaStatics.cu:165-176,
 * 262-272, 293-321) for the parts that need more than the arrays: the
 * gravity potential and adiabatic index of the freebalance mode, and an
 * array's static cells. Here BCSettings.externalData points to one of these
 * (or is NULL: no statics, and a freebalance face is an error); the MEX
 * gateways fill it from boundaryData. The freebalance fields are read from
 * the density array's, the statics of each array from its own. The struct
 * and the arrays it names must stay valid while a call that received the
 * array, or a device time loop created from it, can still run. */
typedef struct ImogenBoundaryData {
    MGArray *gravpot;  /* boundaryData.gravpot: one slab of the fluid's dims, or NULL */
    double gamma;      /* boundaryData.gamma */
    MGArray *statics;  /* boundaryData.staticsData (table columns: cell index, value, coefficient), or NULL */
    double compOffset[12 * MAX_GPUS_USED]; /* boundaryData.compOffset: per GPU, 6 orientations x (offset, count) */
} ImogenBoundaryData;

/* gpuclass/cudaStatics.h: setFluidBoundary (cudaStatics.cu:112-196) and
 * setArrayBoundaryConditions (cudaStatics.cu:199-286). direction 1/2/3,
 * side 0 = minus face, 1 = plus face; direction 0 applies only the static
 * cells. Each array's static cells (externalData) are applied first. */
int setFluidBoundary(MGArray *fluid, GeometryParams *geo, int direction);
int setArrayBoundaryConditions(MGArray *phi, GeometryParams *geo, int direction, int side);

/* cudaStatics.cu:700-857: freebalance ghosts of one face (rightside 0/1) of
 * the 5-array fluid along direction 1/2/3, balanced against the gravity
 * potential gravpot (same dims) at the interior cell's temperature; direction
 * 1 adds the centripetal potential of a Kepler-extrapolated azimuthal
 * velocity (geom Rinner, h[0], frameOmega). XYZ-ordered arrays. */
int setFreeBalanceCondition(MGArray *fluid, GeometryParams *geom, int rightside, int direction, MGArray *gravpot,
                            double fluidGamma);

/* cudaStatics.cu:288-350 (setArrayStaticCells) with the boundaryData fields
 * passed explicitly: the fade of phi's static cells listed in statics for
 * phi's current orientation (compOffset as in ImogenBoundaryData). */
int applyStaticCells(MGArray *phi, MGArray *statics, const double *compOffset);

/* gpuclass/cudaSoundspeed.h, cudaSoundspeed.cu:107-152 (mag must be NULL: hydro only) */
int calculateSoundspeed(MGArray *fluid, MGArray *mag, MGArray *cs, double gamma);

/* gpuclass/cflTimestep.h:11, cflTimestep.cu:107-243: tstep[0] = the CFL time
 * step limit (already halved for HLL/HLLC). Synchronises the device. */
int computeLocalCFLTimestep(MGArray *fluid, MGArray *csound, GeometryParams *geom, int method,
                            double *globalResolution, double *tstep);

/* Same result as calculateSoundspeed + computeLocalCFLTimestep, in one pass:
 * the sound speed is formed in-line from gamma (no counterpart in the
 * reference, which reads a separately computed sound-speed array). */
int imogen_cflTimestepFused(MGArray *fluid, GeometryParams *geom, int method, double gamma,
                            double *globalResolution, double *tstep);

/* ======================================================================== */
/* Gas-dust drag source                                                     */
/* ======================================================================== */

/* Drag integrators. 0..5 are the reference's codes (cudaSource2FluidDrag.cu:
 * 244-263; 4 = LogTrap2 is its default, experiment/Initializer.m:126).
 * 6 and 7 have no counterpart in the reference: 6 integrates the relaxation
 * problem to a tight tolerance (adaptive Dormand-Prince 5(4) on ln|dv|,
 * absolute tolerance 1e-12), 7 is the learned surrogate with per-cell fallback
 * to 6 (needs imogen_drag_set_model first). */
typedef enum DragMethods {
    DRAG_EMP = 0, DRAG_RK4 = 1, DRAG_ETDRK1 = 2, DRAG_ETDRK2 = 3, DRAG_LOGTRAP2 = 4, DRAG_LOGTRAP3 = 5,
    DRAG_EXACT = 6, DRAG_LEARNED = 7
} DragMethods;

/* cudaSource2FluidDrag.cu:189 (header cudaSource2FluidDrag.h). fluidA = gas,
 * fluidB = dust canisters (5 slab views each, same grid). Updates gas E and
 * momenta and dust momenta in place over dt. Built: methods 0-2 and 4-7;
 * method 3 (ETDRK2, listed as not implemented by the reference itself,
 * classes/ImogenManager.m:181) returns ERROR_NOIMPLEMENT. */
int sourcefunction_2FluidDrag(MGArray *fluidA, MGArray *fluidB, GeometryParams *geo, ThermoDetails *thermogas,
                              ThermoDetails *thermodust, double dt, int method);

/* ---- compiled source step: gpuclass/sourceStep.h, cudaCommon.h:216-234,
 *      cudaFreeRadiation.h:11-15 ------------------------------------------- */
typedef struct GravityData {
    MGArray *phi;  /* gravity potential */
    int spaceOrder, timeOrder;
} GravityData;

typedef struct GridFluid {
    MGArray DataHolder;  /* the 5-slab allocation */
    MGArray data[5];     /* rho, E, px, py, pz slab views */
    ThermoDetails thermo;
    BCSettings bcond[5];
    double rhoMin;
} GridFluid;

typedef struct ParametricRadiation {
    double exponent;
    double prefactor;  /* 0 = no radiation */
    double minTemperature;
} ParametricRadiation;

/* sourceStep.cu:200-226. srcType = cyl + 2 rotatingFrame + 4 potential
 * + 8 twoFluid (the table at sourceStep.cu:228-250). Built: 0 (nothing
 * without radiation) and 8 (src___D, sourceStep.cu:300-326: one
 * sourcefunction_2FluidDrag call of fluids[0] (gas) and fluids[1] (dust) over
 * fsp.dt with fsp.multifluidDragMethod; no boundary update, as there).
 * Radiation (prad->prefactor > 0) and srcType 1-7, 9-15 (cylindrical,
 * rotating-frame and gravity terms) return ERROR_NOIMPLEMENT; other values
 * ERROR_INVALID_ARGS. gravdata and tmpStorage are not used by the built
 * types and may be NULL; prad must not be NULL (the reference reads it). */
int performSourceFunctions(int srcType, GridFluid *fluids, int nFluids, FluidStepParams fsp, ParallelTopology *topo,
                           GravityData *gravdata, ParametricRadiation *prad, MGArray *tmpStorage);

/* ---- learned drag surrogate (no counterpart in the reference) -------------
 * Network 5 -> 64 -> 64 -> 3, GELU (tanh form), parameters in FP32 as
 * W1[5][64] b1[64] W2[64][64] b2[64] W3[64][3] b3[3] (row = input unit),
 * IMOGEN_MLP_NPARAMS floats. Features of a cell (natural logs): ln(dt k0),
 * ln(rho_d/rho_g), ln(|dv|/c_s), ln Re_0, ln Kn_0, standardised with
 * fmean/fstd. With tau = dt k0 and g = d ln k/d tau at 0 (analytic), the
 * rates along the exact path are k/k0 = 1 + tau (c g + tau r), c = 1/2 for the
 * mean rate and the rate at dt/2, c = 1 for the rate at dt; the targets are
 * the three remainders r, standardised with tmean/tstd. */
#define IMOGEN_MLP_NPARAMS 4739
typedef struct DragModelNorm {
    float fmean[5], fstd[5];   /* feature standardisation */
    float tmean[3], tstd[3];   /* target standardisation */
    float flo[5], fhi[5];      /* raw feature box the network was trained on */
    float residTol;            /* fallback threshold on the end-rate residual */
} DragModelNorm;

/* Uploads a trained network (converted to BF16 for inference). */
int imogen_drag_set_model(const float *params, const DragModelNorm *norm);
/* Cells and fallbacks of the learned drag since the last call (resets). */
int imogen_drag_learned_stats(int64_t *cells, int64_t *fallbacks);
/* Training set from cell states on the device: cells[n][5] = (rho_g, rho_d,
 * |dv|, u_0, dt k0) in FP64 (the step is given through its stiffness; dt =
 * (dt k0)/k0); writes raw (unstandardised) features[n][5] and targets[n][3]
 * in FP32 from the exact integrator (NaN rows where |dv| < 1e-9). */
int imogen_drag_training_samples(const ThermoDetails *gas, const ThermoDetails *dust, int64_t n,
                                 const double *cells, float *features, float *targets);
/* FP32 training primitives on device buffers (batch rows, row-major).
 * act holds batch x 256 floats (pre-activations z1, z2 and activations
 * h1, h2). grads receives IMOGEN_MLP_NPARAMS floats (overwritten). */
int imogen_mlp_forward(const float *params, int64_t batch, const float *x, float *act, float *y);
int imogen_mlp_backward(const float *params, int64_t batch, const float *x, const float *act, const float *dy,
                        float *grads);
/* Mean-squared-error gradient dy = 2 (y - t) / (3 batch) over batch x 3
 * outputs; adds sum((y - t)^2) / (3 batch) into *lossAccum (device double). */
int imogen_mlp_mse_grad(int64_t batch, const float *y, const float *t, float *dy, double *lossAccum);
/* dst[r][:] = src[idx[r]][:] for r < nrows (ncols floats per row). */
int imogen_gather_rows(const float *src, const int64_t *idx, int64_t nrows, int ncols, float *dst);
/* One Adam step (Kingma & Ba 2015) over n parameters; step counts from 1. */
int imogen_mlp_adam(float *params, const float *grads, float *m, float *v, int64_t n, float lr, float beta1,
                    float beta2, float eps, int step);
/* BF16-MFMA inference of standardised features x[batch][5] -> y[batch][3]
 * (standardised outputs), the path the learned drag uses. */
int imogen_mlp_infer_bf16(int64_t batch, const float *x, float *y);

/* ======================================================================== */
/* Ensembles of independent 2D problems (no counterpart in the reference)   */
/* ======================================================================== */

/* Many independent 2D HLLC problems of uneven size (each 8..512 cells per
 * direction, constant-extrapolation boundaries, the same gamma, limiter and
 * CFL number), advanced together: each iteration computes every problem's own
 * CFL time step on the device and applies fluidstep(+1), fluidstep(-1) with
 * batched sweeps, optionally replayed from HIP graphs. Problem p's state is
 * 5 x ny[p] x nx[p] doubles (rho, E, px, py, pz; x fastest), grid spacing
 * hx[p], hy[p]. */
typedef struct ImogenEnsemble ImogenEnsemble;
int imogen_ensemble_create(int nprob, const int *nx, const int *ny, const double *hx, const double *hy,
                           double gamma, int limiter, double cfl, ImogenEnsemble **out);
int imogen_ensemble_set_state(ImogenEnsemble *e, int p, const double *host);
int imogen_ensemble_get_state(ImogenEnsemble *e, int p, double *host);
/* Runs iterations firstIteration .. firstIteration+iterations-1 (the
 * iteration number picks the sweep order as stepNumber does); useGraph != 0
 * captures one HIP graph per sweep-order residue and replays it. The first
 * iteration of a call reduces every problem's CFL limit over its whole state;
 * later ones take, for problems without dust, the minimum the previous
 * iteration's last sweep folded in (the same value, bitwise). Waits for
 * the last iteration and returns ERROR_CRASH when the kernels met invalid
 * states (imogen_check_errors); a problem whose CFL limit is not positive and
 * finite takes no time step in that iteration. */
int imogen_ensemble_run(ImogenEnsemble *e, int firstIteration, int iterations, int useGraph);
/* Simulation time reached by every problem (nprob doubles; waits). */
int imogen_ensemble_times(ImogenEnsemble *e, double *t);
/* Times `trials` iterations per candidate and keeps, per size class (ny <= 128,
 * <= 256, <= 512), the fastest segment count S in {1, 2, 4} of the Y-sweep
 * march (chosen[3], optional, receives them), then per line-length group of
 * the X sweep (nx <= 64, <= 128, <= 192, <= 256, <= 368, <= 496, <= 512) the
 * fastest cells per lane K among the least K that fits and up to two more. A
 * candidate replaces the default (S = 1, 1, 2; the least K) only when it is 1 %
 * faster. The ensemble state advances during tuning and the clocks are reset
 * to 0 afterwards. */
int imogen_ensemble_autotune(ImogenEnsemble *e, int trials, int *chosen);
/* The X sweep's cells per lane and split mode (1: lines swept as two halves)
 * per line-length group as above (7 ints each; 0 for groups with no problem). */
int imogen_ensemble_x_lanes(ImogenEnsemble *e, int *cellsPerLane, int *split);
/* Gives the problems with hasDust[p] != 0 a second fluid, dust (HLLC with
 * gamma dustGamma, same grid and boundaries), coupled to the gas by the
 * two-fluid drag source of sourcefunction_2FluidDrag (method 4, 5, 6 or 7;
 * gas and dust microphysics as there) applied as sources/src0001.m does
 * around the fluid steps of imogen.m:115-139; the time step of a dusty
 * problem is the minimum over both fluids. Call once, before running. */
int imogen_ensemble_set_dust(ImogenEnsemble *e, const int *hasDust, double dustGamma, const ThermoDetails *gas,
                             const ThermoDetails *dust, int dragMethod);
int imogen_ensemble_set_dust_state(ImogenEnsemble *e, int p, const double *host);
int imogen_ensemble_get_dust_state(ImogenEnsemble *e, int p, double *host);
int imogen_ensemble_destroy(ImogenEnsemble *e);

/* ======================================================================== */
/* Time loop on the device                                                  */
/* ======================================================================== */

/* The iteration of imogen.m:115-139 (TimeManager.update,
 * time/TimeManager.m:124-186, then fluidstep(+1) and fluidstep(-1) of every
 * fluid, flux.cu:27-155, and the drag source when imogen_timeloop_set_drag
 * adds it) with dTime kept in device memory: no host round trip
 * per CFL reduction (cflTimestep.cu:214), one replayed HIP graph per
 * iteration. The host-driven equivalent is imogen_cflTimestepFused on every
 * fluid, dTime = cfl * min(tau) (clamped so t + 2 dTime <= tmax when
 * tmax > 0), then performFluidUpdate_3D(+1), (-1) per fluid with that dTime:
 * the results are bitwise the same. fluids holds nfluids x 5 MGArray (each
 * fluid a 5-slab canister), params one FluidStepParams per fluid (method,
 * gamma, floors, C_iso, limiter; dt, stepNumber, stepDirection are ignored),
 * cflGamma (optional) the gamma of each fluid's CFL sound speed, i.e. the
 * gamma imogen_cflTimestepFused would get (default: thermoGamma; the two
 * differ for isothermal fluids, whose fluid step uses gamma = 2,
 * flux_ML_iface.cu:128-133). The arrays must stay allocated while the loop
 * exists. */
typedef struct ImogenTimeloop ImogenTimeloop;
int imogen_timeloop_create(MGArray *fluids, int nfluids, const GeometryParams *geo, const FluidStepParams *params,
                           const double *cflGamma, double cfl, double tmax, ImogenTimeloop **out);
/* Iterations firstIteration .. firstIteration + iterations - 1 (the number
 * picks the sweep order as stepNumber does). Waits for the last iteration and
 * returns ERROR_CRASH when the kernels met invalid states (see
 * imogen_check_errors); an iteration whose CFL limit is not positive and
 * finite takes no time step (dTime = 0). */
int imogen_timeloop_run(ImogenTimeloop *loop, int firstIteration, int iterations, int useGraph);
/* Simulation time and the last dTime (waits). */
/* Adds the gas-dust drag source of sources/src0001.m to every iteration:
 * source(1/2), fluidstep(+1), source(1), fluidstep(-1), source(1/2)
 * (imogen.m:118-127), each source two sourcefunction_2FluidDrag calls over
 * half its interval followed by setFluidBoundary of both fluids. Fluid 0 is
 * the gas, fluid 1 the dust; method as in sourcefunction_2FluidDrag (the
 * learned method needs imogen_drag_set_model first; graphs are re-captured
 * when the model changes). method < 0 removes the source. The results are
 * bitwise those of the host-driven loop with the same calls. */
int imogen_timeloop_set_drag(ImogenTimeloop *loop, const ThermoDetails *gas, const ThermoDetails *dust, int method);
int imogen_timeloop_clock(ImogenTimeloop *loop, double *time, double *dt);
int imogen_timeloop_set_time(ImogenTimeloop *loop, double time);
int imogen_timeloop_destroy(ImogenTimeloop *loop);

/* ======================================================================== */
/* Runtime utilities (no counterpart in the reference: it used the legacy   */
/* default stream; this library issues all work on one stream per device).  */
/* ======================================================================== */

/* The hipStream_t all kernels of this library are launched on (device 0). */
void *imogen_stream(void);
/* Blocks until all work of this library on the current device is done. */
int imogen_synchronize(void);
/* Waits for the library stream and reports the invalid states its kernels
 * have met since the last report: drag cells with a density, |dv| or (moving
 * cells) gas internal energy that is not positive and finite, cells whose
 * adaptive drag integration ran out of steps, NaN CFL limits, and time steps
 * a device clock refused (limit not positive and finite). Returns
 * ERROR_CRASH (and clears the record) when any was met. The entry points that
 * wait anyway -- sourcefunction_2FluidDrag, performSourceFunctions,
 * computeLocalCFLTimestep, imogen_timeloop_run, imogen_ensemble_run --
 * return the same status themselves. No counterpart in the reference, whose
 * kernels let such states run on silently. */
int imogen_check_errors(void);
/* Plain device buffers for callers without their own GPU runtime (the learned
 * drag's training data). Copies are ordered on the library stream and return
 * when complete; free waits for the stream first. */
int imogen_device_alloc(size_t bytes, void **ptr);
int imogen_device_free(void *ptr);
int imogen_memcpy_htod(void *dst, const void *src, size_t bytes);
int imogen_memcpy_dtoh(void *dst, const void *src, size_t bytes);
/* Per-kernel timing with HIP events on the library stream. enable(1) clears
 * the record and starts timing every launch; read() returns the number of
 * timed launches and their summed duration for a kernel id:
 * 0/1/2 = X/Y/Z fluid sweep, 3 = boundary conditions, 4 = CFL reduction,
 * 5 = drag source, 6 = learned drag. */
int imogen_profile_enable(int on);
int imogen_profile_read(int kernel, int64_t *count, double *totalMs);
/* Text of the last error recorded by the library (never NULL). */
const char *imogen_last_error(void);
/* Library version string. */
const char *imogen_version(void);

#ifdef __cplusplus
}
#endif
#endif
