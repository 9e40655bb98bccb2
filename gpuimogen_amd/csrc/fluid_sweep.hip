// fluid_sweep.hip -- fused HLLC / HLL SSP-RK2 directional sweeps for gfx950.
//
// One kernel launch performs both SSP-RK2 stages of one directional sweep
// (the reference runs them as two or more launches through a 5-slab temporary
// plus a physical transpose, gpuclass/cudaFluidStep.cu:485-627 and :630-700,
// gpuclass/flux.cu:92-116). For HLL the pressure solver's floors, which the
// reference applies in place before each stage (:646-690), are applied to
// each cell as it is read: Q before stage A, U1 before stage B; the floored
// values feed both the fluxes and the SSP-RK2 combination. The fluid is updated in place in the reference's
// 5-slab layout (rho, E, px, py, pz; x fastest), with the reference's periodic
// wrap along the swept line (cudaFluidStep.cu:1159-1162).
//
// Two data-movement schemes, both free of transposes:
//  * X (contiguous) lines: a wave owns whole lines. Each lane holds K
//    consecutive cells in registers; the one-cell neighbours it needs from the
//    adjacent lanes (primitives, left face, flux) come by cross-lane shuffle.
//    No LDS, no barrier, no halo re-read.
//  * Y/Z (strided) lines: each lane marches along one column (coalesced over
//    x across the wave) keeping a register window of the two-stage stencil.
//    Columns may be split into S segments held by lanes of the same wave, for
//    occupancy; a segment reads its 4-cell halos directly, and keeps its first
//    4 outputs in a scratch buffer until the whole wave has finished reading,
//    so the in-place update (and the periodic wrap) is race-free.
#include "sweep_kernels.h"
#include "imogen_internal.h"

namespace imogen {


// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

// Segments per column for the Y/Z march, S in {1, 2, 4} (W = 64/S columns per
// wave, so every lane has a column slot). The kernel's time follows the
// busiest SIMD: with w waves over 1024 SIMDs (256 CUs x 4, up to 2 resident
// waves each) it runs ceil(w / 1024) waves of seglen + 9 steps, and a launch
// with fewer waves than SIMDs runs about 1.8x slower per step (idle SIMDs,
// no latency hiding). Segments shorter than 16 cells are not used. Measured
// Y / Z times (HLLC, ms) and the rule's pick:
//   96^3  S=1 0.140  S=2 0.077  S=4 0.051  -> 4     320^3 S=1 1.04  S=2 0.79  S=4 0.79  -> 4
//   160^3 S=1 0.216  S=2 0.201  S=4 0.125  -> 4     384^3 S=1 1.37  S=2 1.24  S=4 1.26  -> 4
//   192^3 S=1 0.249  S=2 0.236  S=4 0.203  -> 4     256^3 S=1 0.315 S=2 0.309..0.329 S=4 0.324 -> 1
// Counts that do not divide 64 leave a wave's row segment off the 128-byte
// lines: S = 3, 6 (21, 10 columns) ran 1.5-2x slower than S = 2 or 4 (smooth
// state, 160^3-384^3; 192^3: S=3 0.374, S=6 0.395 against S=4 0.197), and
// S = 8 (8 columns, 64 bytes) 1.3-1.5x slower.
static int march_segments(int64_t columns, int N)
{
    int best = 1;
    double bestCost = 0.0;
    for (int S = 1; S <= 4; S *= 2) {
        if (S > 1 && N < 16 * S) break;
        const int64_t waves = (columns * S + 63) / 64;
        const int64_t rounds = (waves + 1023) / 1024;
        const double cost = (double)rounds * ((N + S - 1) / S + 9) * (waves < 1024 ? 1.8 : 1.0);
        if (S == 1 || cost < bestCost) {
            best = S;
            bestCost = cost;
        }
    }
    return best;
}

// Launch shape of an X sweep over nlines lines of N cells: K cells per lane,
// nl lanes per line, 256-thread blocks. K = 0 when N is above 1024.
struct XGeom {
    int K = 0, nl = 0;
    int64_t blocks = 0;
    bool split = false;  // sweep_xh: halves of the line, one wave each, 2 lines per 256-thread block
};

static XGeom x_geom(int N, int64_t nlines)
{
    XGeom g;
    if (N > 256 && N <= 512 && N != 384 && N != 512) {
        // halves of <= 184 / 248 / 256 cells plus 8 window cells in 3 / 4 / 5-cell lanes.
        // Lines of exactly 384 or 512 cells fill every lane of a whole-line wave
        // (6 or 8 cells each) and stay whole: 384^3 X sweep 1.13 ms whole,
        // 1.43 split; 512^3 3.33 whole, 3.29 split; 320^3 1.29 whole (partly
        // filled lanes), 0.78 split.
        g.split = true;
        g.K = (N <= 3 * 64 * 2 - 16) ? 3 : (N <= 4 * 64 * 2 - 16 ? 4 : 5);
        g.nl = 64;
        g.blocks = (nlines + 1) / 2;
        return g;
    }
    if (N <= 64) g.K = 1;
    else if (N <= 128) g.K = 2;
    else if (N <= 192) g.K = 3;
    else if (N <= 256) g.K = 4;
    else if (N == 384) g.K = 6;
    else if (N == 512) g.K = 8;
    else if (N <= 768) g.K = 12;
    else if (N <= 1024) g.K = 16;
    else return g;
    g.nl = (N + g.K - 1) / g.K;
    const int lpw = 64 / g.nl;
    int64_t waves = (nlines + lpw - 1) / lpw;
    // Resident-wave budget: 2 waves per SIMD on 1024 SIMDs; beyond it a wave
    // takes several line groups and prefetches the next one.
    if (waves > 2048) waves = 2048;
    g.blocks = (waves + 3) / 4;
    return g;
}

// Launch shape of a Y (dir 2) or Z (dir 3) march: S segments per column, a
// grid of blocksPerRow x nother blocks of 256 threads (nthr in all).
struct MarchGeom {
    int N = 0, nother = 0, S = 1, seglen = 0, wavesPerRow = 0, blocksPerRow = 0;
    int64_t nthr = 0;
};

static MarchGeom march_geom(int nx, int ny, int nz, int dir)
{
    MarchGeom g;
    g.N = (dir == 2) ? ny : nz;
    g.nother = (dir == 2) ? nz : ny;
    g.S = march_segments((int64_t)nx * g.nother, g.N);
    g.seglen = (g.N + g.S - 1) / g.S;
    const int W = 64 / g.S;
    g.wavesPerRow = (nx + W - 1) / W;
    g.blocksPerRow = (g.wavesPerRow + 3) / 4;
    g.nthr = (int64_t)g.blocksPerRow * g.nother * 256;
    return g;
}

template <class M, int LIM>
static int launch_x(double *q, int64_t pitch, int N, int64_t nlines, double lambda, const double *lamp,
                    const HydroConsts &hc, int rt, const CflFuse &cf, hipStream_t s)
{
    const XGeom g = x_geom(N, nlines);
    if (g.K == 0) return IM_FAIL(ERROR_NOIMPLEMENT, "X sweep supports lines of at most 1024 cells (got %d)", N);
    const int K = g.K, nl = g.nl;
    const int64_t blocks = g.blocks;
    if (cf.part && blocks * 4 > cf.cap) return IM_FAIL(ERROR_INVALID_ARGS, "CFL partial buffer too small");
    if (blocks > INT32_MAX) return IM_FAIL(ERROR_INVALID_ARGS, "too many lines for one X launch");
    dim3 grid((unsigned)blocks), block(256);
    const bool full = (nl * K == N) && (64 % nl == 0);
    double *dummy = (double *)scratch_block((size_t)blocks * 256 * 5 * sizeof(double));
    if (!dummy) return IM_FAIL(ERROR_NOMEM, "no scratch for the X sweep");
    if (g.split) {
#define IM_XH(KK)                                                                                        \
    case KK:                                                                                             \
        if (cf.part)                                                                                     \
            hipLaunchKernelGGL((sweep_xh<M, KK, LIM, true>), grid, block, 0, s, q, pitch, N, nlines, lambda, \
                               lamp, hc, rt, dummy, cf);                                                 \
        else                                                                                             \
            hipLaunchKernelGGL((sweep_xh<M, KK, LIM, false>), grid, block, 0, s, q, pitch, N, nlines,     \
                               lambda, lamp, hc, rt, dummy, cf);                                         \
        break;
        switch (K) {
            IM_XH(3)
            IM_XH(4)
            IM_XH(5)
        }
#undef IM_XH
        return IM_LAUNCHED("sweep_xh");
    }
#define IM_XK(KK)                                                                                       \
    case KK:                                                                                            \
        if (cf.part) {                                                                                  \
            if (full)                                                                                   \
                hipLaunchKernelGGL((sweep_x<M, KK, LIM, true, true>), grid, block, 0, s, q, pitch, N,    \
                                   nlines, nl, lambda, lamp, hc, rt, dummy, cf);                        \
            else                                                                                        \
                hipLaunchKernelGGL((sweep_x<M, KK, LIM, false, true>), grid, block, 0, s, q, pitch, N,   \
                                   nlines, nl, lambda, lamp, hc, rt, dummy, cf);                        \
        } else if (full) {                                                                              \
            hipLaunchKernelGGL((sweep_x<M, KK, LIM, true, false>), grid, block, 0, s, q, pitch, N,       \
                               nlines, nl, lambda, lamp, hc, rt, dummy, cf);                            \
        } else {                                                                                        \
            hipLaunchKernelGGL((sweep_x<M, KK, LIM, false, false>), grid, block, 0, s, q, pitch, N,      \
                               nlines, nl, lambda, lamp, hc, rt, dummy, cf);                            \
        }                                                                                               \
        break;
#define IM_XK_FULL(KK)                                                                                  \
    case KK:                                                                                            \
        if (!full) return IM_FAIL(ERROR_INVALID_ARGS, "X lines of %d cells in %d-cell lanes", N, KK);     \
        if (cf.part)                                                                                    \
            hipLaunchKernelGGL((sweep_x<M, KK, LIM, true, true>), grid, block, 0, s, q, pitch, N, nlines, \
                               nl, lambda, lamp, hc, rt, dummy, cf);                                    \
        else                                                                                            \
            hipLaunchKernelGGL((sweep_x<M, KK, LIM, true, false>), grid, block, 0, s, q, pitch, N, nlines, \
                               nl, lambda, lamp, hc, rt, dummy, cf);                                    \
        break;
    switch (K) {
        IM_XK(1)
        IM_XK(2)
        IM_XK(3)
        IM_XK(4)
        IM_XK_FULL(6)
        IM_XK_FULL(8)
        IM_XK(12)
        IM_XK(16)
    }
#undef IM_XK
#undef IM_XK_FULL
    return IM_LAUNCHED("sweep_x");
}

template <class M, int LIM>
static int launch_march(double *q, int64_t pitch, int nx, int ny, int nz, int dir, double lambda,
                        const double *lamp, const HydroConsts &hc, int rt, const CflFuse &cf, hipStream_t s)
{
    const int64_t sstride = (dir == 2) ? (int64_t)nx : (int64_t)nx * ny;
    const int64_t ostride = (dir == 2) ? (int64_t)nx * ny : (int64_t)nx;
    const MarchGeom g = march_geom(nx, ny, nz, dir);
    const int N = g.N, S = g.S, seglen = g.seglen, wavesPerRow = g.wavesPerRow;
    dim3 grid((unsigned)g.blocksPerRow, (unsigned)g.nother), block(256);
    const int64_t nthr = g.nthr;
    if (cf.part && nthr / 64 > cf.cap) return IM_FAIL(ERROR_INVALID_ARGS, "CFL partial buffer too small");
    // 4 deferred output slots + 1 dummy slot of 5 doubles per lane
    double *defer = (double *)scratch_block((size_t)nthr * 25 * sizeof(double));
    if (!defer) return IM_FAIL(ERROR_NOMEM, "no scratch for deferred outputs");
    // rows of a multiple of 16 cells in a 256-byte-aligned slab: every segment
    // row (16, 32 or 64 columns) is whole 128-byte lines -> nontemporal hint
    const bool lines = (nx % 16) == 0 && (pitch % 32) == 0 && ((uintptr_t)q % 256) == 0;
#define IM_MARCH(D, C, T)                                                                                  \
    hipLaunchKernelGGL((sweep_march<M, D, LIM, C, T>), grid, block, 0, s, q, pitch, nx, sstride, ostride, N, S, \
                       seglen, wavesPerRow, lambda, lamp, hc, rt, defer, cf)
#define IM_MARCH_NT(D, C)             \
    do {                              \
        if (lines) IM_MARCH(D, C, true);  \
        else IM_MARCH(D, C, false);       \
    } while (0)
    if (dir == 2) {
        if (cf.part) IM_MARCH_NT(2, true);
        else IM_MARCH_NT(2, false);
    } else {
        if (cf.part) IM_MARCH_NT(3, true);
        else IM_MARCH_NT(3, false);
    }
#undef IM_MARCH_NT
#undef IM_MARCH
    return IM_LAUNCHED("sweep_march");
}

template <class M, int LIM>
static int launch_dir(double *q, int64_t pitch, int nx, int ny, int nz, int dir, double lambda,
                      const double *lamp, const HydroConsts &hc, int rt, const CflFuse &cf, hipStream_t s)
{
    const int slot = prof_begin(PROF_SWEEP_X + dir - 1, s);
    int rc;
    if (dir == 1) rc = launch_x<M, LIM>(q, pitch, nx, (int64_t)ny * nz, lambda, lamp, hc, rt, cf, s);
    else rc = launch_march<M, LIM>(q, pitch, nx, ny, nz, dir, lambda, lamp, hc, rt, cf, s);
    prof_end(slot, s);
    return rc;
}

template <class M>
static int launch_method(double *base, int64_t pitch, int nx, int ny, int nz, int dir, double lambda,
                         const double *lamp, const HydroConsts &hc, int limiter, const CflFuse &cf, hipStream_t s)
{
    switch (limiter) {
    case LIMITER_OSPRE: return launch_dir<M, 0>(base, pitch, nx, ny, nz, dir, lambda, lamp, hc, 0, cf, s);
    case LIMITER_VANLEER: return launch_dir<M, 1>(base, pitch, nx, ny, nz, dir, lambda, lamp, hc, 1, cf, s);
    case LIMITER_MINMOD:
    case LIMITER_SUPERBEE:
    case LIMITER_VANALBADA:
    case LIMITER_MC:
    case LIMITER_ZERO: return launch_dir<M, -1>(base, pitch, nx, ny, nz, dir, lambda, lamp, hc, limiter, cf, s);
    default: return IM_FAIL(ERROR_INVALID_ARGS, "unknown slope limiter %d", limiter);
    }
}

// The grid arithmetic of launch_x / launch_march, for sizing fused-CFL buffers.
int64_t sweep_waves(int nx, int ny, int nz, int dir)
{
    if (dir == 1) return x_geom(nx, (int64_t)ny * nz).blocks * 4;
    const MarchGeom g = march_geom(nx, ny, nz, dir);
    return (int64_t)g.blocksPerRow * g.nother * 4;
}

int64_t sweep_scratch_bytes(int nx, int ny, int nz, int dir)
{
    const int N = (dir == 1) ? nx : (dir == 2 ? ny : nz);
    if (dir < 1 || dir > 3 || N < 3) return 0;
    if (dir == 1) return x_geom(nx, (int64_t)ny * nz).blocks * 256 * 5 * (int64_t)sizeof(double);
    return march_geom(nx, ny, nz, dir).nthr * 25 * (int64_t)sizeof(double);
}

int launch_fluid_sweep(double *base, int64_t pitch, int nx, int ny, int nz, int dir, int method, double lambda,
                       double gamma, double rhomin, double ciso, int limiter, hipStream_t s, const double *lamp,
                       const CflFuse *cfl)
{
    if (dir < 1 || dir > 3) return IM_FAIL(ERROR_INVALID_ARGS, "sweep direction %d", dir);
    const int N = (dir == 1) ? nx : (dir == 2 ? ny : nz);
    if (N < 3) return SUCCESSFUL;  // cudaFluidStep.cu:301
    HydroConsts hc;
    hc.gamma = gamma;
    hc.gm1 = gamma - 1.0;
    hc.govergm1 = gamma / (gamma - 1.0);
    hc.rhomin = rhomin;
    hc.emin = 1e-5 / (gamma - 1.0);  // cudaFluidStep.cu:298
    hc.ciso = ciso;
    hc.cisosq = ciso * ciso;
    const CflFuse cf = cfl ? *cfl : CflFuse{};
    if (method == METHOD_HLLC)
        return launch_method<Hllc>(base, pitch, nx, ny, nz, dir, lambda, lamp, hc, limiter, cf, s);
    if (method == METHOD_HLL) return launch_method<Hll>(base, pitch, nx, ny, nz, dir, lambda, lamp, hc, limiter, cf, s);
    return IM_FAIL(ERROR_INVALID_ARGS, "sweep kernel for method %d", method);
}

}  // namespace imogen
