// ensemble.hip -- many independent 2D HLLC problems of uneven sizes, advanced
// together on one GPU.
//
// Every problem keeps the reference's fluid layout (5 slabs, x fastest) in one
// pooled allocation. One Imogen iteration of the whole ensemble is
//   per-problem CFL (two batched reduction passes: each problem's time step,
//   lambda = dt/h per direction and its clock stay on the device)
//   fluidstep(+1), fluidstep(-1): per direction one batched sweep launch over
//   the lines of all problems, then one batched boundary launch,
// with the sweep order of gpuclass/flux.cu:46 (z skipped: the problems are 2D).
// Every kernel reads its problem's lambda from device memory, so an iteration
// needs no host round trip and is captured once per sweep-order residue
// (iteration mod 6) as a HIP graph and replayed.
//
// Work decomposition: X sweeps group problems by line length into the
// register-blocking classes of the single-array X kernel (K cells per lane,
// one launch per class, one wave per group of whole lines); Y sweeps use the
// column march with a per-problem segment count S (columns split in S
// segments held by lanes of one wave). The autotuner picks S per size class
// by timing the candidates on the ensemble itself.
#include <algorithm>
#include <cstring>
#include <vector>

#include "cfl_device.h"
#include "drag_device.h"
#include "imogen_internal.h"
#include "sweep_kernels.h"

namespace imogen {

// One fluid array of one problem (a table entry). Entries 0..P-1 are the
// problems' gas; dusty problems add an entry for their dust.
struct EnsProblem {
    double *q;
    int64_t pitch;  // doubles between slabs
    int nx, ny;
    double hx, hy;
    double gamma;
    int owner;  // problem whose clock and time step this entry follows
    int dust;   // gas entries: the entry of the problem's dust, -1 if none
    int fold;   // 1: the iteration's last sweep folds this entry's CFL minimum (fold_eligible)
    int pad;
};

struct EnsItem {
    int p;   // problem
    int a;   // X: first line; Y: first column
    int b;   // Y: segments S
    int pad;
};

constexpr int kEnsCflBlocks = 32;  // CFL partial minima per problem
constexpr int kEnsGhost = 4;       // ghost cells per face set by the boundary kernel
constexpr double kEnsCflNone = 1e37;  // CFL minimum of an entry with no limiting cell

// CFL minimum folded into the last sweep of an iteration. A problem without
// dust whose sides both exceed 2 kEnsGhost cells ends its iteration with that
// sweep and the const boundaries of the swept direction (the source step's
// boundary calls copy cells too), and every boundary cell then holds a copy of
// a cell the sweep kept (cells [kEnsGhost, N - kEnsGhost) of the swept
// direction, every cell of the other): the minimum over the kept cells is the
// minimum over the whole state. Problems with dust end the iteration with a
// drag source, and take the two-pass reduction. gas: the entry is a problem's
// gas (not a dust entry).
static bool fold_eligible(const EnsProblem &P, bool gas)
{
    return gas && P.dust < 0 && P.nx > 2 * kEnsGhost && P.ny > 2 * kEnsGhost;
}

// Order-preserving map of doubles (NaN aside) to unsigned 64-bit keys, so that
// waves of any problem can take a problem's minimum with one integer atomic.
__host__ __device__ inline unsigned long long ens_order_key(double v)
{
    unsigned long long b;
    memcpy(&b, &v, sizeof b);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double ens_order_value(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

// Per-wave end of a folding sweep: the wave's minimum into its problem's slot
// (lane values past the kept cells stay kEnsCflNone), NaN limits counted as
// the reduction pass counts them. All 64 lanes reach it.
__device__ __forceinline__ void ens_fold_wave(double cmin, unsigned nan, unsigned long long *slot,
                                              unsigned long long *err)
{
    wave_count(err + ERRW_CFL_NAN, nan);
    cmin = wave_min(cmin);
    if ((threadIdx.x & 63) == 0 && cmin < kEnsCflNone) atomicMin(slot, ens_order_key(cmin));
}
// X line classes: K cells per lane; lines of up to 256 cells are swept whole
// (several short lines per wave), longer ones split into two halves swept by
// a pair of waves (ens_sweep_xh)
constexpr int kXClasses = 7;
static const int kXClassK[kXClasses] = {1, 2, 3, 4, 3, 4, 5};
static const bool kXClassSplit[kXClasses] = {false, false, false, false, true, true, true};

static int x_class_of(int N)
{
    if (N <= 64) return 0;
    if (N <= 128) return 1;
    if (N <= 192) return 2;
    if (N <= 256) return 3;
    if (N <= 3 * 64 * 2 - 16) return 4;  // halves of <= 184 cells + 8 halo cells in 3-cell lanes
    if (N <= 4 * 64 * 2 - 16) return 5;  // halves of <= 248 cells + 8 halo cells in 4-cell lanes
    if (N <= 512) return 6;              // halves of <= 256 cells + 8 halo cells in 5-cell lanes
    return -1;
}

// Class of lines of N cells swept with K cells per lane, whole or split into
// halves (lines above 256 cells only); -1 when the line or half window does
// not fit 64 lanes.
static int x_class_for(int N, int K, bool split)
{
    if (split) {
        if (N <= 256 || N > 512 || (N + 1) / 2 + 8 > 64 * K) return -1;
        return K == 3 ? 4 : (K == 4 ? 5 : (K == 5 ? 6 : -1));
    }
    if (K < 1 || K > 4 || N > 64 * K) return -1;
    return K - 1;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// the CFL arguments of a 2D entry (cflTimestep.cu:292-335, fused sound speed)
__device__ __forceinline__ CflArgs ens_cfl_args(const EnsProblem &P)
{
    return CflArgs{P.gamma * (P.gamma - 1), P.hx, P.hy, 1.0, P.hx < P.hy ? P.hx : P.hy, 1, 2, 0};
}

// partial minima of the entries listed in `entries` (workgroups (block, list index))
__global__ __launch_bounds__(256) void ens_cfl_pass1(const EnsProblem *__restrict__ probs,
                                                     const int *__restrict__ entries, double *__restrict__ part,
                                                     unsigned long long *__restrict__ err)
{
    const int ent = entries[blockIdx.y];
    const EnsProblem P = probs[ent];
    const int64_t n = (int64_t)P.nx * P.ny;
    const CflArgs a = ens_cfl_args(P);
    double m = kEnsCflNone;
    unsigned nan = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double *c = P.q + i;
        const double u = cfl_cell(a, c[0], c[P.pitch], c[2 * P.pitch], c[3 * P.pitch], c[4 * P.pitch], 0.0);
        m = (u < m) ? u : m;
        nan += (u != u);
    }
    wave_count(err + ERRW_CFL_NAN, nan);
    m = wave_min(m);
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = w[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) r = (w[k] < r) ? w[k] : r;
        part[(int64_t)ent * gridDim.x + blockIdx.x] = r;
    }
}

// one wave per problem: tau (halved, HLLC rule cflTimestep.cu:232-237), dt,
// lambda per direction, and the problem's clock t += 2 dt. A limit that is
// not positive and below 1e36 (every cell NaN, or no cell limiting: the 1e37
// sentinel) takes no time step (dt = 0), counted in the device error word.
// useFold: the gas limit of problems whose entry folds comes from the previous
// iteration's last sweep (fold[p]); every problem's fold slot is then reset for
// this iteration's last sweep.
__global__ __launch_bounds__(64) void ens_cfl_pass2(const EnsProblem *__restrict__ probs, const double *part,
                                                    int nb, double cfl, double *__restrict__ lam,
                                                    double *__restrict__ clock, double *__restrict__ dtp,
                                                    unsigned long long *__restrict__ fold, int useFold,
                                                    unsigned long long *__restrict__ err)
{
    const int p = blockIdx.x;
    const int de = probs[p].dust;
    // min over fluids of each fluid's halved limit, as TimeManager takes it
    double m = 1e38;
    if (useFold && probs[p].fold) {
        m = ens_order_value(fold[p]);
    } else {
        for (int k = threadIdx.x; k < nb; k += 64)
            m = (part[(int64_t)p * nb + k] < m) ? part[(int64_t)p * nb + k] : m;
        m = wave_min(m);
    }
    double md = 1e38;
    if (de >= 0) {
        for (int k = threadIdx.x; k < nb; k += 64)
            md = (part[(int64_t)de * nb + k] < md) ? part[(int64_t)de * nb + k] : md;
        md = wave_min(md);
    }
    if (threadIdx.x == 0) {
        const double tg = m / 2, td = md / 2;
        const double tau = (de >= 0 && td < tg) ? td : tg;
        double dt = cfl * tau;
        if (!(tau > 0 && tau < 1e36)) {
            dt = 0.0;
            atomicAdd(err + ERRW_TIMESTEP, 1ull);
        }
        lam[2 * p] = dt / probs[p].hx;
        lam[2 * p + 1] = dt / probs[p].hy;
        if (de >= 0) {
            lam[2 * de] = lam[2 * p];
            lam[2 * de + 1] = lam[2 * p + 1];
        }
        dtp[p] = dt;
        clock[p] += 2 * dt;
        fold[p] = ens_order_key(kEnsCflNone);
    }
}

// the HLLC constants of an entry (launch_fluid_sweep's, from the entry's gamma)
__device__ __forceinline__ HydroConsts ens_consts(double gamma)
{
    HydroConsts hc;
    hc.gamma = gamma;
    hc.gm1 = gamma - 1.0;
    hc.govergm1 = gamma / (gamma - 1.0);
    hc.rhomin = 0.0;
    hc.emin = 1e-5 / (gamma - 1.0);
    hc.ciso = -1.0;
    hc.cisosq = 1.0;
    return hc;
}

// X sweep of whole lines (K <= 4 cells per lane, two waves per SIMD); one
// work item (problem, first line) per wave. CFL: fold the CFL minimum of the
// output cells into fold[problem] (entries with P.fold set).
template <int K, int LIM, bool CFL>
__global__ __launch_bounds__(256, 2) void ens_sweep_x(const EnsProblem *__restrict__ probs,
                                                   const EnsItem *__restrict__ work, int nwork,
                                                   const double *__restrict__ lam, int rt,
                                                   double *__restrict__ dummy, unsigned long long *__restrict__ fold,
                                                   unsigned long long *__restrict__ err)
{
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave >= nwork) return;
    const int lane = threadIdx.x & 63;
    const EnsItem it = work[wave];
    const EnsProblem P = probs[it.p];
    const HydroConsts hc = ens_consts(P.gamma);
    const double lambda = lam[2 * it.p];
    const int N = P.nx;
    const int nl = (N + K - 1) / K;
    const int lpw = 64 / nl;
    const int sub = lane / nl;
    const int j = lane - sub * nl;
    const int line = it.a + sub;
    const bool own = sub < lpw && line < P.ny;
    const int lbase = sub * nl;
    const bool laneUsed = sub < lpw;
    const int srcL = laneUsed ? lbase + (j == 0 ? nl - 1 : j - 1) : lane;
    const int srcR = laneUsed ? lbase + (j == nl - 1 ? 0 : j + 1) : lane;
    const int x0 = laneUsed ? j * K : 0;
    int cnt = N - x0;
    cnt = cnt < K ? cnt : K;
    if (!laneUsed) cnt = 0;
    const int lc = line < P.ny ? line : P.ny - 1;
    double *b = P.q + (int64_t)lc * N + x0;
    V5 Q[K];
#pragma unroll
    for (int k = 0; k < K; k++) Q[k] = (k < cnt) ? load_cell<Hllc, 1>(b + k, P.pitch) : V5{1.0, 1.0, 0.0, 0.0, 0.0};
    V5 F[K], Fm;
    x_fluxes<Hllc, K, LIM>(hc, rt, Q, cnt, srcL, srcR, F, Fm);
    V5 U1[K];
#pragma unroll
    for (int k = 0; k < K; k++) U1[k] = axpy_update(Q[k], lambda, F[k], k == 0 ? Fm : F[k > 0 ? k - 1 : 0]);
    x_fluxes<Hllc, K, LIM>(hc, rt, U1, cnt, srcL, srcR, F, Fm);
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
    double *dst = own ? b : dummy + gtid;
    const int64_t st = own ? P.pitch : nthr;
    const bool fl = CFL && own && P.fold;
    const CflArgs ca = ens_cfl_args(P);
    double cmin = kEnsCflNone;
    unsigned nan = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (k < cnt) {
            const V5 o = avg_update(add5(U1[k], Q[k]), lambda, F[k], k == 0 ? Fm : F[k > 0 ? k - 1 : 0]);
            store_cell<Hllc, 1>(own ? dst + k : dst, st, o);
            if (fl && x0 + k >= kEnsGhost && x0 + k < N - kEnsGhost) {
                const double u = cfl_of_cell<Hllc, 1>(ca, o);
                cmin = (u < cmin) ? u : cmin;
                nan += (u != u);
            }
        }
    }
    if (CFL) ens_fold_wave(cmin, nan, fold + P.owner, err);
}

// X sweep of lines split into two halves, one wave per half. A wave holds its
// half and the 4 cells beyond each end (the two stages' stencil, periodic
// within the line, as the whole-line sweep reads them), so it forms its half's
// outputs alone; the window's outer cells only feed discarded values. The two
// waves of a line are in one workgroup (4 waves: 2 lines), and every load of
// the workgroup precedes the barrier while every store follows it, so no wave
// overwrites a cell its partner has still to read. One work item (problem,
// line) per wave pair; lines of at most 2 (64 K - 8) cells.
template <int K, int LIM, bool CFL>
__global__ __launch_bounds__(256, 2) void ens_sweep_xh(const EnsProblem *__restrict__ probs,
                                                    const EnsItem *__restrict__ work, int nwork,
                                                    const double *__restrict__ lam, int rt,
                                                    double *__restrict__ dummy,
                                                    unsigned long long *__restrict__ fold,
                                                    unsigned long long *__restrict__ err)
{
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    int item = (int)blockIdx.x * 2 + (wib >> 1);
    const bool itemValid = item < nwork;
    item = itemValid ? item : nwork - 1;  // waves past the list still meet the barrier
    const int part = wib & 1;
    const EnsItem it = work[item];
    const EnsProblem P = probs[it.p];
    const HydroConsts hc = ens_consts(P.gamma);
    const double lambda = lam[2 * it.p];
    const int N = P.nx;
    const int h = (N + 1) >> 1;
    const int s0 = part ? h : 0;          // first output cell of this half
    const int Wn = (part ? N - h : h) + 8;  // window: the half and 4 cells beyond each end
    const int nl = (Wn + K - 1) / K;
    const bool laneUsed = lane < nl;
    const int srcL = laneUsed ? (lane == 0 ? nl - 1 : lane - 1) : lane;
    const int srcR = laneUsed ? (lane == nl - 1 ? 0 : lane + 1) : lane;
    const int w0 = lane * K;
    int cnt = Wn - w0;
    cnt = cnt < K ? cnt : K;
    if (!laneUsed) cnt = 0;
    double *row = P.q + (int64_t)it.a * N;
    V5 Q[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        int g = s0 - 4 + w0 + k;  // window cell -> line cell, periodic
        g = (g < 0) ? g + N : g;
        g = (g >= N) ? g - N : g;
        g = (g < N) ? g : N - 1;
        Q[k] = (k < cnt) ? load_cell<Hllc, 1>(row + g, P.pitch) : V5{1.0, 1.0, 0.0, 0.0, 0.0};
    }
    __syncthreads();
    V5 F[K], Fm;
    x_fluxes<Hllc, K, LIM>(hc, rt, Q, cnt, srcL, srcR, F, Fm);
    V5 U1[K];
#pragma unroll
    for (int k = 0; k < K; k++) U1[k] = axpy_update(Q[k], lambda, F[k], k == 0 ? Fm : F[k > 0 ? k - 1 : 0]);
    x_fluxes<Hllc, K, LIM>(hc, rt, U1, cnt, srcL, srcR, F, Fm);
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
    const bool fl = CFL && P.fold;
    const CflArgs ca = ens_cfl_args(P);
    double cmin = kEnsCflNone;
    unsigned nan = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (k < cnt) {
            const int w = w0 + k;
            const bool own = itemValid && w >= 4 && w < Wn - 4;
            const int x = s0 + w - 4;
            const V5 o = avg_update(add5(U1[k], Q[k]), lambda, F[k], k == 0 ? Fm : F[k > 0 ? k - 1 : 0]);
            store_cell<Hllc, 1>(own ? row + x : dummy + gtid, own ? P.pitch : nthr, o);
            if (fl && own && x >= kEnsGhost && x < N - kEnsGhost) {
                const double u = cfl_of_cell<Hllc, 1>(ca, o);
                cmin = (u < cmin) ? u : cmin;
                nan += (u != u);
            }
        }
    }
    if (CFL) ens_fold_wave(cmin, nan, fold + P.owner, err);
}

// workgroups of an X class launch over n work items
static int64_t x_blocks(int cls, int64_t n)
{
    return kXClassSplit[cls] ? (n + 1) / 2 : (n + 3) / 4;
}

// Y sweep: column march, one work item (problem, first column, S) per wave
template <int LIM, bool CFL>
__global__ __launch_bounds__(256) void ens_sweep_y(const EnsProblem *__restrict__ probs,
                                                   const EnsItem *__restrict__ work, int nwork,
                                                   const double *__restrict__ lam, int rt,
                                                   double *__restrict__ defer, unsigned long long *__restrict__ fold,
                                                   unsigned long long *__restrict__ err)
{
    __shared__ double sring[4][4][5][64];
    const int wib = threadIdx.x >> 6;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave >= nwork) return;
    const int lane = threadIdx.x & 63;
    const EnsItem it = work[wave];
    const EnsProblem P = probs[it.p];
    const HydroConsts hc = ens_consts(P.gamma);
    const double lambda = lam[2 * it.p + 1];
    const int S = it.b;
    const int W = 64 / S;
    const int s = lane / W;
    const int c = lane - s * W;
    const int x = it.a + c;
    const bool colValid = (x < P.nx) && (s < S);
    const int N = P.ny;
    const int seglen = (N + S - 1) / S;
    const int j0 = colValid ? s * seglen : 0;  // lanes of no column: zero-length segment at row 0
    int L = N - j0;
    L = L < seglen ? L : seglen;
    if (!colValid) L = 0;
    double *col = P.q + (colValid ? x : 0);
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
    // cells [kEnsGhost, N - kEnsGhost) of the column fold their CFL limits
    const bool fl = CFL && P.fold;
    const CflFuse cf{nullptr, ens_cfl_args(P), fl ? kEnsGhost : 0, fl ? N - kEnsGhost : 0, 0};
    double cmin = kEnsCflNone;
    unsigned nan = 0;
    march_body<Hllc, 2, LIM, CFL>(col, P.pitch, (int64_t)P.nx, N, seglen, j0, L, lambda, hc, rt, defer, gtid,
                                  nthr, sring[wib], lane, cf, cmin, nan);
    if (CFL) ens_fold_wave(cmin, nan, fold + P.owner, err);
}

// constant-extrapolation ghost cells (cudaStatics.cu:1339-1366, 1440-1466 with
// the +y innermost-ghost convention), depth d, both faces of one direction;
// workgroups (entry, 16): on the 512-problem ensemble 16 per entry took the x
// faces from 64 to 46 us per launch and the y faces from 21 to 16 (4 per
// entry), 32 gained nothing more
// xrules: the +y face copies the last interior cell, as the fluid step sets
// its boundaries (flux.cu:99-107, the swept direction rotated into memory x);
// 0: the y kernel's innermost-ghost copy (cudaStatics.cu:1461-1466), as
// src0001.m's boundary calls on unrotated arrays
__global__ __launch_bounds__(256) void ens_bc_const(const EnsProblem *__restrict__ probs, int dir, int d, int xrules)
{
    const EnsProblem P = probs[blockIdx.x];
    const int nx = P.nx, ny = P.ny;
    const int N = (dir == 1) ? nx : ny;
    if (N < d + 2) return;
    // Both faces in one pass only when the cells one face reads are disjoint
    // from those the other writes (N >= 3 d + 2); otherwise the minus face
    // first, then the plus face, as the reference's two calls run, by the
    // entry's first workgroup alone (the passes are ordered by its barrier).
    const bool together = N >= 3 * d + 2;
    if (!together && blockIdx.y > 0) return;
    const int nsides = together ? 2 : 1;
    const int t0 = together ? (int)(blockIdx.y * blockDim.x + threadIdx.x) : (int)threadIdx.x;
    const int tstride = together ? (int)(gridDim.y * blockDim.x) : (int)blockDim.x;
    for (int pass = 0; pass < (together ? 1 : 2); pass++) {
        if (pass) __syncthreads();
        if (dir == 1) {
            // a thread per (ghost, side, slab, row): the d ghosts of one row end
            // are adjacent lanes, so a store instruction writes whole row-end
            // segments (a thread per row end writing its d ghosts one by one:
            // 104 against 64 us per launch, 4 workgroups per entry)
            const int total = ny * 5 * nsides * d;
            for (int t = t0; t < total; t += tstride) {
                const int g = t % d;
                int r = t / d;
                const int side = together ? (r & 1) : pass;
                r = together ? (r >> 1) : r;
                const int slab = r % 5;
                const int row = r / 5;
                double *a = P.q + slab * P.pitch + (int64_t)row * nx;
                if (side == 0) a[g] = a[d];
                else a[nx - d + g] = a[nx - d - 1];
            }
        } else {
            // a thread per (column, slab, side)
            const int total = nx * 5 * nsides;
            for (int t = t0; t < total; t += tstride) {
                const int x = t % nx;
                const int r = t / nx;
                const int side = together ? (r & 1) : pass;
                const int slab = together ? (r >> 1) : r;
                double *a = P.q + slab * P.pitch + x;
                if (side == 0) {
                    const double f = a[(int64_t)d * nx];
                    for (int g = 0; g < d; g++) a[(int64_t)g * nx] = f;
                } else {
                    const double f = a[(int64_t)(xrules ? ny - d - 1 : ny - d) * nx];
                    for (int g = 0; g < d; g++) a[(int64_t)(ny - d + g) * nx] = f;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct EnsembleImpl {
    int P = 0;   // problems (gas entries 0..P-1)
    int NE = 0;  // table entries: gas, then the dusty problems' dust
    double gamma = 1.4, cfl = 0.4;
    int limiter = 0;
    std::vector<EnsProblem> host;
    std::vector<int> segS;  // chosen Y segments per problem
    std::vector<int> xcls;  // X line class per problem (x_class_of unless tuned)
    double *pool = nullptr;
    EnsProblem *dprobs = nullptr;
    double *lam = nullptr, *clock = nullptr, *cflPart = nullptr;
    // folded CFL minima (one key per problem) and the entries the reduction
    // pass reads: every entry (cflList[0..NE)), then those that do not fold
    unsigned long long *fold = nullptr;
    int *cflList = nullptr;
    int nNoFold = 0;
    EnsItem *xwork[kXClasses] = {nullptr};
    int nx_work[kXClasses] = {0};
    EnsItem *ywork = nullptr;
    int ny_work = 0;
    double *xdummy = nullptr, *ydefer = nullptr;
    int64_t xThreads = 0, yThreads = 0;
    int64_t xcells[kXClasses] = {0};  // cells of each X line class (the stream balance)
    // The X line classes sweep disjoint lines: their launches fork onto these
    // streams (default priority) and join back, parallel branches of the graph
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
    hipGraphExec_t graph[2][6] = {{nullptr}};  // [reduction reads the folded minima][sweep order]
    int64_t cells = 0;
    // dust (imogen_ensemble_set_dust)
    int Pd = 0;
    double *dpool = nullptr, *dtp = nullptr;
    EnsDragTable *dragTab = nullptr;
    EnsDragItem *dragItems = nullptr;
    int nDragItems = 0;
    int64_t dragCells = 0;  // cells of the dusty problems
    int64_t *queue = nullptr;
    unsigned long long *qcount = nullptr;
    DragConsts dc{};
    int dragMethod = DRAG_LEARNED;
    uint64_t modelGen = 0;
};

static void free_graphs(EnsembleImpl *e)
{
    for (int f = 0; f < 2; f++)
        for (int k = 0; k < 6; k++) {
            if (e->graph[f][k]) (void)hipGraphExecDestroy(e->graph[f][k]);
            e->graph[f][k] = nullptr;
        }
}

static int build_work(EnsembleImpl *e)
{
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    // the tables below are replaced while no kernel can still read them
    if (IM_HIP(hipStreamSynchronize(st->stream))) return ERROR_CUDA_BLEW_UP;
    // which entries fold their CFL minimum into the last sweep; the reduction
    // pass lists every entry, then the entries that do not fold
    std::vector<int> cl;
    for (int p = 0; p < e->NE; p++) {
        e->host[p].fold = fold_eligible(e->host[p], p < e->P) ? 1 : 0;
        cl.push_back(p);
    }
    for (int p = 0; p < e->NE; p++)
        if (!e->host[p].fold) cl.push_back(p);
    e->nNoFold = (int)cl.size() - e->NE;
    if (IM_HIP(hipMemcpy(e->dprobs, e->host.data(), e->NE * sizeof(EnsProblem), hipMemcpyHostToDevice)))
        return ERROR_CUDA_BLEW_UP;
    if (e->cflList) (void)hipFree(e->cflList);
    e->cflList = nullptr;
    if (IM_HIP(hipMalloc(&e->cflList, cl.size() * sizeof(int)))) return ERROR_CUDA_BLEW_UP;
    if (IM_HIP(hipMemcpy(e->cflList, cl.data(), cl.size() * sizeof(int), hipMemcpyHostToDevice)))
        return ERROR_CUDA_BLEW_UP;
    std::vector<EnsItem> xw[kXClasses], yw;
    for (int c = 0; c < kXClasses; c++) e->xcells[c] = 0;
    for (int p = 0; p < e->NE; p++) {
        const EnsProblem &P = e->host[p];
        const int cls = e->xcls[P.owner];
        const int K = kXClassK[cls];
        const int nl = (P.nx + K - 1) / K;
        const int lpw = kXClassSplit[cls] ? 1 : 64 / nl;  // split classes: one item per line
        for (int l = 0; l < P.ny; l += lpw) xw[cls].push_back(EnsItem{p, l, 0, 0});
        e->xcells[cls] += (int64_t)P.nx * P.ny;
        const int S = e->segS[e->host[p].owner];
        const int W = 64 / S;
        for (int x = 0; x < P.nx; x += W) yw.push_back(EnsItem{p, x, S, 0});
    }
    for (int c = 0; c < kXClasses; c++) {
        std::stable_sort(xw[c].begin(), xw[c].end(),
                         [e](const EnsItem &a, const EnsItem &b) { return e->host[a.p].nx > e->host[b.p].nx; });
        if (e->xwork[c]) (void)hipFree(e->xwork[c]);
        e->xwork[c] = nullptr;
        e->nx_work[c] = (int)xw[c].size();
        if (xw[c].empty()) continue;
        if (IM_HIP(hipMalloc(&e->xwork[c], xw[c].size() * sizeof(EnsItem)))) return ERROR_CUDA_BLEW_UP;
        if (IM_HIP(hipMemcpy(e->xwork[c], xw[c].data(), xw[c].size() * sizeof(EnsItem), hipMemcpyHostToDevice)))
            return ERROR_CUDA_BLEW_UP;
    }
    // longest marches first, so that the short ones fill in behind them
    std::stable_sort(yw.begin(), yw.end(), [e](const EnsItem &a, const EnsItem &b) {
        return (e->host[a.p].ny + a.b - 1) / a.b > (e->host[b.p].ny + b.b - 1) / b.b;
    });
    if (e->ywork) (void)hipFree(e->ywork);
    e->ny_work = (int)yw.size();
    if (IM_HIP(hipMalloc(&e->ywork, yw.size() * sizeof(EnsItem)))) return ERROR_CUDA_BLEW_UP;
    if (IM_HIP(hipMemcpy(e->ywork, yw.data(), yw.size() * sizeof(EnsItem), hipMemcpyHostToDevice)))
        return ERROR_CUDA_BLEW_UP;
    // per-lane scratch: X dummy slots (5 doubles), Y deferred outputs (25 doubles)
    int64_t xt = 0;
    for (int c = 0; c < kXClasses; c++) xt = std::max<int64_t>(xt, x_blocks(c, e->nx_work[c]) * 256);
    const int64_t yt = (((int64_t)e->ny_work + 3) / 4) * 256;
    if (xt > e->xThreads) {
        if (e->xdummy) (void)hipFree(e->xdummy);
        if (IM_HIP(hipMalloc(&e->xdummy, (size_t)xt * 5 * sizeof(double)))) return ERROR_CUDA_BLEW_UP;
        e->xThreads = xt;
    }
    if (yt > e->yThreads) {
        if (e->ydefer) (void)hipFree(e->ydefer);
        if (IM_HIP(hipMalloc(&e->ydefer, (size_t)yt * 25 * sizeof(double)))) return ERROR_CUDA_BLEW_UP;
        e->yThreads = yt;
    }
    free_graphs(e);
    return SUCCESSFUL;
}

static int ens_bc(EnsembleImpl *e, int dir, hipStream_t s, int xrules)
{
    hipLaunchKernelGGL(ens_bc_const, dim3((unsigned)e->NE, 16), dim3(256), 0, s, e->dprobs, dir, 4, xrules);
    return IM_LAUNCHED("ens_bc_const");
}

// sources/src0001.m for the dusty problems: two drag calls over dTime*fraction/2
// each, then every fluid's boundaries (idempotent for the problems without dust)
static int ens_source(EnsembleImpl *e, double fraction, hipStream_t s)
{
    if (!e->Pd) return SUCCESSFUL;
    // both half-interval calls in one pass over memory
    int rc = launch_ens_drag(e->dragTab, e->dragItems, e->nDragItems, e->dragCells, e->dtp, fraction / 2, e->dc,
                             e->dragMethod, e->queue, e->qcount, s, 2);
    if (rc) return rc;
    for (int d = 1; d <= 2; d++) {
        rc = ens_bc(e, d, s, 0);
        if (rc) return rc;
    }
    return SUCCESSFUL;
}

// one sweep of every entry in direction dir, then that direction's boundaries;
// CFL: the sweep folds the CFL minima of the entries that fold
template <int LIM, bool CFL>
static int ens_sweep(EnsembleImpl *e, int dir, int rt, hipStream_t s)
{
    unsigned long long *err = state()->err;
    if (dir == 1) {
        // classes to the three streams, largest first to the least loaded
        int order[kXClasses], nc = 0;
        for (int c = 0; c < kXClasses; c++)
            if (e->nx_work[c]) order[nc++] = c;
        std::sort(order, order + nc, [e](int a, int b) { return e->xcells[a] > e->xcells[b]; });
        int64_t load[3] = {0, 0, 0};
        if (IM_HIP(hipEventRecord(e->fork, s))) return ERROR_CUDA_BLEW_UP;
        for (int k = 0; k < 2; k++)
            if (IM_HIP(hipStreamWaitEvent(e->side[k], e->fork, 0))) return ERROR_CUDA_BLEW_UP;
        for (int i = 0; i < nc; i++) {
            const int c = order[i];
            const int sl = (int)(std::min_element(load, load + 3) - load);
            load[sl] += e->xcells[c];
            hipStream_t sx = sl == 0 ? s : e->side[sl - 1];
            const unsigned blocks = (unsigned)x_blocks(c, e->nx_work[c]);
#define IM_EX(KK)                                                                                              \
    hipLaunchKernelGGL((ens_sweep_x<KK, LIM, CFL>), dim3(blocks), dim3(256), 0, sx, e->dprobs, e->xwork[c],     \
                       e->nx_work[c], e->lam, rt, e->xdummy, e->fold, err)
#define IM_EXH(KK)                                                                                             \
    hipLaunchKernelGGL((ens_sweep_xh<KK, LIM, CFL>), dim3(blocks), dim3(256), 0, sx, e->dprobs, e->xwork[c],    \
                       e->nx_work[c], e->lam, rt, e->xdummy, e->fold, err)
            if (kXClassSplit[c]) {
                switch (kXClassK[c]) {
                case 3: IM_EXH(3); break;
                case 4: IM_EXH(4); break;
                case 5: IM_EXH(5); break;
                }
            } else {
                switch (kXClassK[c]) {
                case 1: IM_EX(1); break;
                case 2: IM_EX(2); break;
                case 3: IM_EX(3); break;
                case 4: IM_EX(4); break;
                }
            }
#undef IM_EX
#undef IM_EXH
            int rc = IM_LAUNCHED("ens_sweep_x");
            if (rc) return rc;
        }
        for (int k = 0; k < 2; k++) {
            if (IM_HIP(hipEventRecord(e->join[k], e->side[k]))) return ERROR_CUDA_BLEW_UP;
            if (IM_HIP(hipStreamWaitEvent(s, e->join[k], 0))) return ERROR_CUDA_BLEW_UP;
        }
    } else {
        const unsigned blocks = (unsigned)((e->ny_work + 3) / 4);
        hipLaunchKernelGGL((ens_sweep_y<LIM, CFL>), dim3(blocks), dim3(256), 0, s, e->dprobs, e->ywork,
                           e->ny_work, e->lam, rt, e->ydefer, e->fold, err);
        int rc = IM_LAUNCHED("ens_sweep_y");
        if (rc) return rc;
    }
    return ens_bc(e, dir, s, 1);
}

template <bool CFL>
static int ens_sweep_lim(EnsembleImpl *e, int dir, hipStream_t s)
{
    switch (e->limiter) {
    case LIMITER_OSPRE: return ens_sweep<0, CFL>(e, dir, 0, s);
    case LIMITER_VANLEER: return ens_sweep<1, CFL>(e, dir, 1, s);
    default: return ens_sweep<-1, CFL>(e, dir, e->limiter, s);
    }
}

// one iteration (imogen.m:115-139): CFL over all fluids, source(1/2),
// fluidstep(+1), source(1), fluidstep(-1), source(1/2); the sources act only
// when some problems carry dust. The last sweep folds the CFL minima of the
// entries that fold; useFold: this iteration's CFL takes those minima (left by
// the previous iteration, whose state this one starts from) instead of
// reducing the entries again.
static int ens_iteration(EnsembleImpl *e, int it, hipStream_t s, bool useFold)
{
    static const int kFluxcall[3][6] = {{1, 2, 1, 2, 3, 3}, {3, 1, 2, 3, 1, 2}, {2, 3, 3, 1, 2, 1}};
    unsigned long long *err = state()->err;
    const int npass = useFold ? e->nNoFold : e->NE;
    const int *list = e->cflList + (useFold ? e->NE : 0);
    int rc;
    if (npass > 0) {
        hipLaunchKernelGGL(ens_cfl_pass1, dim3(kEnsCflBlocks, (unsigned)npass), dim3(256), 0, s, e->dprobs, list,
                           e->cflPart, err);
        if ((rc = IM_LAUNCHED("ens_cfl_pass1"))) return rc;
    }
    hipLaunchKernelGGL(ens_cfl_pass2, dim3((unsigned)e->P), dim3(64), 0, s, e->dprobs, e->cflPart, kEnsCflBlocks,
                       e->cfl, e->lam, e->clock, e->dtp, e->fold, useFold ? 1 : 0, err);
    if ((rc = IM_LAUNCHED("ens_cfl_pass2"))) return rc;
    if ((rc = ens_source(e, 0.5, s))) return rc;
    for (int dirn = 1; dirn >= -1; dirn -= 2) {
        const int sweep = (it + 3 * (dirn > 0)) % 6;
        int last = 0;
        for (int n = 0; n < 3; n++)
            if (kFluxcall[n][sweep] != 3) last = n;
        for (int n = 0; n < 3; n++) {
            const int d = kFluxcall[n][sweep];
            if (d == 3) continue;  // 2D problems
            rc = (dirn < 0 && n == last) ? ens_sweep_lim<true>(e, d, s) : ens_sweep_lim<false>(e, d, s);
            if (rc) return rc;
        }
        if ((rc = ens_source(e, dirn > 0 ? 1.0 : 0.5, s))) return rc;
    }
    return SUCCESSFUL;
}

}  // namespace imogen

using namespace imogen;

struct ImogenEnsemble {
    EnsembleImpl impl;
};

extern "C" {

int imogen_ensemble_create(int nprob, const int *nx, const int *ny, const double *hx, const double *hy, double gamma,
                           int limiter, double cfl, ImogenEnsemble **out)
{
    if (!out || !nx || !ny || !hx || !hy) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    *out = nullptr;
    if (nprob <= 0) return IM_FAIL(ERROR_INVALID_ARGS, "need at least one problem");
    if (limiter < 0 || limiter > LIMITER_ZERO) return IM_FAIL(ERROR_INVALID_ARGS, "unknown slope limiter %d", limiter);
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    ImogenEnsemble *E = new ImogenEnsemble();
    EnsembleImpl *e = &E->impl;
    e->P = nprob;
    e->NE = nprob;
    e->gamma = gamma;
    e->cfl = cfl;
    e->limiter = limiter;
    int64_t total = 0;
    std::vector<int64_t> off(nprob);
    for (int p = 0; p < nprob; p++) {
        if (nx[p] < 8 || ny[p] < 8 || x_class_of(nx[p]) < 0) {
            delete E;
            return IM_FAIL(ERROR_INVALID_ARGS, "problem %d: %d x %d (need 8..512 cells per direction)", p, nx[p],
                           ny[p]);
        }
        const int64_t n = (int64_t)nx[p] * ny[p];
        const int64_t pitch = (n + 31) / 32 * 32;  // 256-byte slabs
        off[p] = total;
        total += 5 * pitch;
        e->host.push_back(EnsProblem{nullptr, pitch, nx[p], ny[p], hx[p], hy[p], gamma, p, -1, 0, 0});
        e->cells += n;
        e->segS.push_back(ny[p] >= 256 ? 2 : 1);
        e->xcls.push_back(x_class_of(nx[p]));
    }
    int rc = SUCCESSFUL;
    if (IM_HIP(hipMalloc(&e->pool, (size_t)total * sizeof(double)))) rc = ERROR_CUDA_BLEW_UP;
    for (int p = 0; p < nprob && !rc; p++) e->host[p].q = e->pool + off[p];
    if (!rc && IM_HIP(hipMalloc(&e->dprobs, nprob * sizeof(EnsProblem)))) rc = ERROR_CUDA_BLEW_UP;
    if (!rc && IM_HIP(hipMemcpy(e->dprobs, e->host.data(), nprob * sizeof(EnsProblem), hipMemcpyHostToDevice)))
        rc = ERROR_CUDA_BLEW_UP;
    if (!rc && IM_HIP(hipMalloc(&e->lam, 2 * nprob * sizeof(double)))) rc = ERROR_CUDA_BLEW_UP;
    for (int k = 0; k < 2 && !rc; k++) {
        if (IM_HIP(hipStreamCreateWithFlags(&e->side[k], hipStreamNonBlocking))) rc = ERROR_CUDA_BLEW_UP;
        if (!rc && IM_HIP(hipEventCreateWithFlags(&e->join[k], hipEventDisableTiming))) rc = ERROR_CUDA_BLEW_UP;
    }
    if (!rc && IM_HIP(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming))) rc = ERROR_CUDA_BLEW_UP;
    if (!rc && IM_HIP(hipMalloc(&e->clock, nprob * sizeof(double)))) rc = ERROR_CUDA_BLEW_UP;
    if (!rc && IM_HIP(hipMemset(e->clock, 0, nprob * sizeof(double)))) rc = ERROR_CUDA_BLEW_UP;
    if (!rc && IM_HIP(hipMalloc(&e->dtp, nprob * sizeof(double)))) rc = ERROR_CUDA_BLEW_UP;
    if (!rc && IM_HIP(hipMalloc(&e->cflPart, (size_t)nprob * kEnsCflBlocks * sizeof(double)))) rc = ERROR_CUDA_BLEW_UP;
    if (!rc && IM_HIP(hipMalloc(&e->fold, nprob * sizeof(unsigned long long)))) rc = ERROR_CUDA_BLEW_UP;
    if (!rc) {
        const std::vector<unsigned long long> none(nprob, ens_order_key(kEnsCflNone));
        if (IM_HIP(hipMemcpy(e->fold, none.data(), nprob * sizeof(unsigned long long), hipMemcpyHostToDevice)))
            rc = ERROR_CUDA_BLEW_UP;
    }
    if (!rc) rc = build_work(e);
    if (rc) {
        imogen_ensemble_destroy(E);
        return rc;
    }
    *out = E;
    return SUCCESSFUL;
}

int imogen_ensemble_set_state(ImogenEnsemble *E, int p, const double *host)
{
    if (!E || !host) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    EnsembleImpl *e = &E->impl;
    if (p < 0 || p >= e->P) return IM_FAIL(ERROR_INVALID_ARGS, "problem %d", p);
    DeviceState *st = state();
    const EnsProblem &P = e->host[p];
    const int64_t n = (int64_t)P.nx * P.ny;
    for (int k = 0; k < 5; k++)
        if (IM_HIP(hipMemcpyAsync(P.q + k * P.pitch, host + k * n, n * sizeof(double), hipMemcpyHostToDevice,
                                  st->stream)))
            return ERROR_CUDA_BLEW_UP;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

int imogen_ensemble_get_state(ImogenEnsemble *E, int p, double *host)
{
    if (!E || !host) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    EnsembleImpl *e = &E->impl;
    if (p < 0 || p >= e->P) return IM_FAIL(ERROR_INVALID_ARGS, "problem %d", p);
    DeviceState *st = state();
    const EnsProblem &P = e->host[p];
    const int64_t n = (int64_t)P.nx * P.ny;
    for (int k = 0; k < 5; k++)
        if (IM_HIP(hipMemcpyAsync(host + k * n, P.q + k * P.pitch, n * sizeof(double), hipMemcpyDeviceToHost,
                                  st->stream)))
            return ERROR_CUDA_BLEW_UP;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

int imogen_ensemble_set_dust(ImogenEnsemble *E, const int *hasDust, double dustGamma, const ThermoDetails *gas,
                             const ThermoDetails *dust, int dragMethod)
{
    if (!E || !hasDust || !gas || !dust) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    EnsembleImpl *e = &E->impl;
    if (e->Pd) return IM_FAIL(ERROR_INVALID_ARGS, "the ensemble already carries dust");
    if (dragMethod != DRAG_LOGTRAP2 && dragMethod != DRAG_LOGTRAP3 && dragMethod != DRAG_EXACT &&
        dragMethod != DRAG_LEARNED)
        return IM_FAIL(ERROR_NOIMPLEMENT, "ensemble drag method %d", dragMethod);
    if (!(dustGamma > 1.0)) return IM_FAIL(ERROR_INVALID_ARGS, "dust gamma %g", dustGamma);
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    if (IM_HIP(hipStreamSynchronize(st->stream))) return ERROR_CUDA_BLEW_UP;
    if (learned_drag_prepare()) return ERROR_CUDA_BLEW_UP;  // no allocation once graphs capture
    std::vector<int> dusty;
    int64_t total = 0, cells = 0;
    for (int p = 0; p < e->P; p++)
        if (hasDust[p]) {
            dusty.push_back(p);
            total += 5 * e->host[p].pitch;
            cells += (int64_t)e->host[p].nx * e->host[p].ny;
        }
    if (dusty.empty()) return SUCCESSFUL;
    const int NE = e->P + (int)dusty.size();
    double *dpool = nullptr, *lam = nullptr, *part = nullptr;
    EnsProblem *dprobs = nullptr;
    EnsDragTable *tab = nullptr;
    EnsDragItem *items = nullptr;
    int64_t *queue = nullptr;
    unsigned long long *qcount = nullptr;
    std::vector<EnsProblem> host = e->host;
    std::vector<EnsDragTable> htab;
    std::vector<EnsDragItem> hitems;
    int64_t off = 0;
    int rc = IM_HIP(hipMalloc(&dpool, (size_t)total * sizeof(double)));
    for (size_t k = 0; k < dusty.size() && !rc; k++) {
        const int p = dusty[k];
        EnsProblem D = host[p];
        D.q = dpool + off;
        D.gamma = dustGamma;
        D.owner = p;
        D.dust = -1;
        off += 5 * D.pitch;
        host[p].dust = e->P + (int)k;
        host.push_back(D);
        const int64_t n = (int64_t)D.nx * D.ny;
        htab.push_back(EnsDragTable{host[p].q, D.q, D.pitch, n, p, 0});
        for (int64_t b = 0; b < n; b += 64) hitems.push_back(EnsDragItem{(int)k, 0, b});
    }
    if (!rc) rc = IM_HIP(hipMalloc(&dprobs, NE * sizeof(EnsProblem)));
    if (!rc) rc = IM_HIP(hipMemcpy(dprobs, host.data(), NE * sizeof(EnsProblem), hipMemcpyHostToDevice));
    if (!rc) rc = IM_HIP(hipMalloc(&lam, 2 * NE * sizeof(double)));
    if (!rc) rc = IM_HIP(hipMalloc(&part, (size_t)NE * kEnsCflBlocks * sizeof(double)));
    if (!rc) rc = IM_HIP(hipMalloc(&tab, htab.size() * sizeof(EnsDragTable)));
    if (!rc) rc = IM_HIP(hipMemcpy(tab, htab.data(), htab.size() * sizeof(EnsDragTable), hipMemcpyHostToDevice));
    if (!rc) rc = IM_HIP(hipMalloc(&items, hitems.size() * sizeof(EnsDragItem)));
    if (!rc) rc = IM_HIP(hipMemcpy(items, hitems.data(), hitems.size() * sizeof(EnsDragItem), hipMemcpyHostToDevice));
    if (!rc) rc = IM_HIP(hipMalloc(&queue, (size_t)cells * 2 * sizeof(int64_t)));  // queue + second-call list
    if (!rc) rc = IM_HIP(hipMalloc(&qcount, 3 * sizeof(unsigned long long)));
    if (rc) {
        void *bufs[] = {dpool, dprobs, lam, part, tab, items, queue, qcount};
        for (void *b : bufs)
            if (b) (void)hipFree(b);
        return ERROR_CUDA_BLEW_UP;
    }
    void *old[] = {e->dprobs, e->lam, e->cflPart};
    for (void *b : old)
        if (b) (void)hipFree(b);
    e->host = host;
    e->NE = NE;
    e->Pd = (int)dusty.size();
    e->dpool = dpool;
    e->dprobs = dprobs;
    e->lam = lam;
    e->cflPart = part;
    e->dragTab = tab;
    e->dragItems = items;
    e->nDragItems = (int)hitems.size();
    e->dragCells = cells;
    e->queue = queue;
    e->qcount = qcount;
    e->dc = drag_consts(gas, dust);
    e->dragMethod = dragMethod;
    e->modelGen = learned_model_generation();
    return build_work(e);
}

static int dust_entry(EnsembleImpl *e, int p)
{
    if (p < 0 || p >= e->P) return -1;
    return e->host[p].dust;
}

int imogen_ensemble_set_dust_state(ImogenEnsemble *E, int p, const double *host)
{
    if (!E || !host) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    EnsembleImpl *e = &E->impl;
    const int de = dust_entry(e, p);
    if (de < 0) return IM_FAIL(ERROR_INVALID_ARGS, "problem %d carries no dust", p);
    DeviceState *st = state();
    const EnsProblem &P = e->host[de];
    const int64_t n = (int64_t)P.nx * P.ny;
    for (int k = 0; k < 5; k++)
        if (IM_HIP(hipMemcpyAsync(P.q + k * P.pitch, host + k * n, n * sizeof(double), hipMemcpyHostToDevice,
                                  st->stream)))
            return ERROR_CUDA_BLEW_UP;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

int imogen_ensemble_get_dust_state(ImogenEnsemble *E, int p, double *host)
{
    if (!E || !host) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    EnsembleImpl *e = &E->impl;
    const int de = dust_entry(e, p);
    if (de < 0) return IM_FAIL(ERROR_INVALID_ARGS, "problem %d carries no dust", p);
    DeviceState *st = state();
    const EnsProblem &P = e->host[de];
    const int64_t n = (int64_t)P.nx * P.ny;
    for (int k = 0; k < 5; k++)
        if (IM_HIP(hipMemcpyAsync(host + k * n, P.q + k * P.pitch, n * sizeof(double), hipMemcpyDeviceToHost,
                                  st->stream)))
            return ERROR_CUDA_BLEW_UP;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

int imogen_ensemble_run(ImogenEnsemble *E, int firstIteration, int iterations, int useGraph)
{
    if (E && E->impl.Pd && E->impl.dragMethod == DRAG_LEARNED && E->impl.modelGen != learned_model_generation()) {
        free_graphs(&E->impl);  // the graphs hold the previous model by value
        E->impl.modelGen = learned_model_generation();
    }
    if (!E) return IM_FAIL(ERROR_NULL_POINTER, "null ensemble");
    EnsembleImpl *e = &E->impl;
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    for (int k = 0; k < iterations; k++) {
        const int it = firstIteration + k;
        // the first iteration of a call reduces every entry (the state may
        // have been set since the last call); later ones take folded minima
        const bool useFold = k > 0;
        if (!useGraph) {
            int rc = ens_iteration(e, it, st->stream, useFold);
            if (rc) return rc;
            continue;
        }
        hipGraphExec_t &g = e->graph[useFold][it % 6];
        if (!g) {
            hipGraph_t graph;
            if (IM_HIP(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal))) return ERROR_CUDA_BLEW_UP;
            int rc = ens_iteration(e, it, st->stream, useFold);
            hipError_t ec = hipStreamEndCapture(st->stream, &graph);
            if (rc) return rc;
            if (IM_HIP(ec)) return ERROR_CUDA_BLEW_UP;
            ec = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (IM_HIP(ec)) return ERROR_CUDA_BLEW_UP;
        }
        if (IM_HIP(hipGraphLaunch(g, st->stream))) return ERROR_CUDA_BLEW_UP;
    }
    return take_device_errors(st);
}

int imogen_ensemble_times(ImogenEnsemble *E, double *t)
{
    if (!E || !t) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    EnsembleImpl *e = &E->impl;
    DeviceState *st = state();
    if (IM_HIP(hipMemcpyAsync(t, e->clock, e->P * sizeof(double), hipMemcpyDeviceToHost, st->stream)))
        return ERROR_CUDA_BLEW_UP;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

int imogen_ensemble_autotune(ImogenEnsemble *E, int trials, int *chosen)
{
    if (!E) return IM_FAIL(ERROR_NULL_POINTER, "null ensemble");
    EnsembleImpl *e = &E->impl;
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    if (trials < 1) trials = 1;
    // Candidates are timed over whole iterations, one size class at a time with
    // the others fixed. The state advances during tuning; callers tune before
    // loading their data or reload it afterwards.
    hipEvent_t a, b;
    if (IM_HIP(hipEventCreate(&a)) || IM_HIP(hipEventCreate(&b))) return ERROR_CUDA_BLEW_UP;
    auto timed = [&](float *ms) {
        int rc = build_work(e);
        if (!rc) rc = ens_iteration(e, 0, st->stream, false);  // warm
        if (!rc) rc = IM_HIP(hipEventRecord(a, st->stream));
        for (int k = 0; k < trials && !rc; k++) rc = ens_iteration(e, k, st->stream, true);
        if (!rc) rc = IM_HIP(hipEventRecord(b, st->stream));
        if (!rc) rc = IM_HIP(hipEventSynchronize(b));
        if (!rc) rc = IM_HIP(hipEventElapsedTime(ms, a, b));
        return rc;
    };
    int rc = SUCCESSFUL;
    // Y-march segments per size class (ny <= 128, <= 256, <= 512) among {1, 2, 4}:
    // the class's default first, another count only when it is 1 % faster (two
    // trial iterations differ by about that)
    static const int cand[3][3] = {{1, 2, 4}, {1, 2, 4}, {2, 1, 4}};
    auto cls_of = [](int ny) { return ny <= 128 ? 0 : (ny <= 256 ? 1 : 2); };
    int best[3] = {1, 1, 2};
    for (int cls = 0; cls < 3 && !rc; cls++) {
        bool any = false;
        for (int p = 0; p < e->P; p++) any = any || cls_of(e->host[p].ny) == cls;
        if (!any) continue;
        float bestMs = 1e30f;
        for (int ci = 0; ci < 3 && !rc; ci++) {
            for (int p = 0; p < e->P; p++)
                if (cls_of(e->host[p].ny) == cls) e->segS[p] = cand[cls][ci];
            float ms = 0;
            rc = timed(&ms);
            if (!rc && (ci == 0 || ms < 0.99f * bestMs)) {
                bestMs = ms;
                best[cls] = cand[cls][ci];
            }
        }
        for (int p = 0; p < e->P; p++)
            if (cls_of(e->host[p].ny) == cls) e->segS[p] = best[cls];
    }
    // X cells per lane per line-length group (the groups of x_class_of): the
    // group's least K and up to two more where the lines (whole lines: several
    // per wave) or half windows still fit 64 lanes
    for (int g = 0; g < kXClasses && !rc; g++) {
        std::vector<int> members;
        for (int p = 0; p < e->P; p++)
            if (x_class_of(e->host[p].nx) == g) members.push_back(p);
        if (members.empty()) continue;
        std::vector<int> cands;
        for (int K = kXClassK[g]; K <= kXClassK[g] + 2; K++) {
            bool ok = true;
            for (int p : members) ok = ok && x_class_for(e->host[p].nx, K, kXClassSplit[g]) >= 0;
            if (ok) cands.push_back(x_class_for(e->host[members[0]].nx, K, kXClassSplit[g]));
        }
        if (cands.size() < 2) continue;
        // cands[0] is the least K (x_class_of's); another K must beat it by 1 %
        // (the whole-iteration timings of two trials differ by about that)
        float bestMs = 1e30f;
        int bestCls = g;
        for (size_t ci = 0; ci < cands.size() && !rc; ci++) {
            for (int p : members) e->xcls[p] = cands[ci];
            float ms = 0;
            rc = timed(&ms);
            if (!rc && (ci == 0 ? true : ms < 0.99f * bestMs)) {
                bestMs = ms;
                bestCls = cands[ci];
            }
        }
        for (int p : members) e->xcls[p] = bestCls;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (chosen)
        for (int c = 0; c < 3; c++) chosen[c] = best[c];
    if (IM_HIP(hipMemset(e->clock, 0, e->P * sizeof(double)))) return ERROR_CUDA_BLEW_UP;
    return build_work(e);
}

int imogen_ensemble_x_lanes(ImogenEnsemble *E, int *cellsPerLane, int *split)
{
    if (!E || !cellsPerLane || !split) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    EnsembleImpl *e = &E->impl;
    for (int g = 0; g < kXClasses; g++) {
        cellsPerLane[g] = 0;
        split[g] = 0;
        for (int p = 0; p < e->P; p++)
            if (x_class_of(e->host[p].nx) == g) {
                cellsPerLane[g] = kXClassK[e->xcls[p]];
                split[g] = kXClassSplit[e->xcls[p]] ? 1 : 0;
                break;
            }
    }
    return SUCCESSFUL;
}

int imogen_ensemble_destroy(ImogenEnsemble *E)
{
    if (!E) return SUCCESSFUL;
    EnsembleImpl *e = &E->impl;
    DeviceState *st = state();
    if (st) (void)hipStreamSynchronize(st->stream);
    free_graphs(e);
    for (int c = 0; c < kXClasses; c++)
        if (e->xwork[c]) (void)hipFree(e->xwork[c]);
    void *bufs[] = {e->pool,   e->dprobs,   e->lam,       e->clock, e->cflPart, e->fold, e->cflList, e->ywork, e->xdummy,
                    e->ydefer, e->dpool,    e->dtp,       e->dragTab, e->dragItems, e->queue, e->qcount};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    for (int k = 0; k < 2; k++) {
        if (e->side[k]) (void)hipStreamDestroy(e->side[k]);
        if (e->join[k]) (void)hipEventDestroy(e->join[k]);
    }
    if (e->fork) (void)hipEventDestroy(e->fork);
    delete E;
    return SUCCESSFUL;
}

}  // extern "C"
