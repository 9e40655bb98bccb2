// xinjin.hip -- Xin-Jin relaxing TVD sweeps (second order) for gfx950.
//
// One directional sweep is four launches, in the reference's order
// (gpuclass/cudaFluidStep.cu:357-475):
//   freeze(Q) -> predict U = floor(Q - lambda/4 dF(Q))       into a temporary
//   freeze(U) -> correct Q = floor(Q - lambda/2 dF_tvd(U))   in place
// The freezing speed of a line is max |p_x|/rho + sqrt(gamma P/rho) over that
// line (cukern_PressureFreezeSolverHydro, :2868-2941); the stage arithmetic,
// including the Ospre flux limiter of the corrector and the density/energy
// floors, follows cukern_XinJinHydro_step (:2427-2654) for cartesian geometry.
// The data movement is this library's: X lines are handled by waves whose
// lanes hold consecutive cells (neighbours by cross-lane shuffle, 60 outputs
// per 64 lanes); Y/Z lines by lanes marching along their column with a
// register window. The corrector reads U and writes Q, and each lane reads Q
// only at the cell it writes, so it runs in place without hazards.
#include "sweep_device.h"
#include "imogen_internal.h"

namespace imogen {

struct XjConsts {
    double gamma, gm1, rhomin, emin;
};

// Ospre flux limiter, gpuclass/limiter_functions.h:36-42
__device__ __forceinline__ double xj_ospre(double a, double b)
{
    // a^2 + ab + b^2 below the normal range (|a|, |b| < 1.5e-154, numerator
    // already +-0) is raised to the smallest normal: its hardware reciprocal
    // would be infinite and the quotient NaN instead of +-0
    const double r = a * b;
    const double v = div_nr(1.5 * r * (a + b), fmax(a * a + r + b * b, 0x1p-1022));
    return (r <= 0.0) ? 0.0 : v;
}

// Left- and right-going fluxes (C_f q - w, C_f q + w) of one cell for every
// slab; NS is the slab of the normal momentum (2, 3, 4).
template <int NS>
__device__ __forceinline__ void xj_cell(const XjConsts &k, double cf, const double (&q)[5], double (&wl)[5],
                                        double (&wr)[5])
{
    const double ir = rcp_nr(q[0]);
    const double P = k.gm1 * (q[1] - .5 * (q[4] * q[4] + q[3] * q[3] + q[2] * q[2]) * ir);
    const double vx = q[NS] * ir;
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double w;
        if (v == 0) w = q[NS];
        else if (v == 1) w = vx * (q[1] + P);
        else if (v == NS) w = vx * q[v] + P;
        else w = vx * q[v];
        wl[v] = cf * q[v] - w;
        wr[v] = cf * q[v] + w;
    }
}

// density and internal-energy floors of a new state (:2620-2628)
__device__ __forceinline__ void xj_floor(const XjConsts &k, double (&p)[5])
{
    p[0] = (p[0] < k.rhomin) ? k.rhomin : p[0];
    const double w = .5 * (p[2] * p[2] + p[3] * p[3] + p[4] * p[4]) * rcp_nr(p[0]);
    if ((p[1] - w) < p[0] * k.emin) p[1] = w + p[0] * k.emin;
}

__device__ __forceinline__ double xj_wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o > v) ? o : v;
    }
    return v;
}

// freezing speed contribution of one cell (:2887-2896)
__device__ __forceinline__ double xj_cf_cell(const XjConsts &k, const double *__restrict__ c, int64_t pitch)
{
    const double invrho = rcp_nr(c[0]);
    double psq = c[3 * pitch];
    double px = c[4 * pitch];
    psq = psq * psq + px * px;
    px = c[2 * pitch];
    psq += px * px;
    const double P = (c[pitch] - .5 * psq * invrho) * k.gm1;
    return fabs(px) * invrho + sqrt_nr(k.gamma * P * invrho);
}

// X lines: one wave per line, lanes strided along x
__global__ __launch_bounds__(256) void xj_freeze_x(const double *__restrict__ q, int64_t pitch, int N,
                                                   int64_t nlines, XjConsts k, double *__restrict__ cf)
{
    const int64_t line = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (line >= nlines) return;
    const int lane = threadIdx.x & 63;
    const double *row = q + line * (int64_t)N;
    double cmax = 0.0;
    for (int x = lane; x < N; x += 64) {
        const double loc = xj_cf_cell(k, row + x, pitch);
        cmax = (loc > cmax) ? loc : cmax;
    }
    cmax = xj_wave_max(cmax);
    if (lane == 0) cf[line] = cmax;
}

// Y/Z lines: one lane per column, marching along the line
__global__ __launch_bounds__(256) void xj_freeze_march(const double *__restrict__ q, int64_t pitch, int nx,
                                                       int64_t sstride, int64_t ostride, int N, XjConsts k,
                                                       double *__restrict__ cf)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx) return;
    const double *col = q + x + (int64_t)blockIdx.y * ostride;
    double cmax = 0.0;
    for (int j = 0; j < N; j++) {
        const double loc = xj_cf_cell(k, col + (int64_t)j * sstride, pitch);
        cmax = (loc > cmax) ? loc : cmax;
    }
    cf[x + (int64_t)nx * blockIdx.y] = cmax;
}

// One stage along X. Lane l of a wave holds cell x0 - 2 + l (periodic); lanes
// 2..61 produce outputs. src: state the fluxes are taken of (Q or U); base:
// state the update starts from; dst: where the new state goes.
template <bool CORRECT>
__global__ __launch_bounds__(256) void xj_stage_x(const double *__restrict__ src, const double *base,
                                                  double *dst, int64_t pitch, int N, int64_t nlines, int nseg,
                                                  double lambda, const double *__restrict__ lamp, double lscale,
                                                  XjConsts k, const double *__restrict__ cf)
{
    if (lamp) lambda = lscale * lamp[0];  // time step kept on the device
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= nlines * nseg) return;
    const int lane = threadIdx.x & 63;
    const int64_t line = wave / nseg;
    const int seg = (int)(wave - line * nseg);
    const int xr = seg * 60 - 2 + lane;  // >= -2
    int x = xr % N;
    if (x < 0) x += N;
    const int64_t off = line * (int64_t)N + x;
    const double cfl = cf[line];
    double q[5];
#pragma unroll
    for (int v = 0; v < 5; v++) q[v] = src[off + v * pitch];
    double wl[5], wr[5];
    xj_cell<2>(k, cfl, q, wl, wr);
    const bool delivers = lane >= 2 && lane <= 61 && xr < N;
    double p[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double WL = wl[v], WR = wr[v];
        if (CORRECT) {
            const double dl = .5 * (wl[v] - __shfl(wl[v], lane - 1, 64));
            const double dr = .5 * (wr[v] - __shfl(wr[v], lane - 1, 64));
            WL -= xj_ospre(dl, __shfl(dl, lane + 1, 64));
            WR += xj_ospre(dr, __shfl(dr, lane + 1, 64));
        }
        const double WRm = __shfl(WR, lane - 1, 64);
        const double WLp = __shfl(WL, lane + 1, 64);
        const double b = CORRECT ? base[off + v * pitch] : q[v];
        p[v] = b - lambda * (WR - WRm - WLp + WL);
    }
    xj_floor(k, p);
    if (delivers) {
#pragma unroll
        for (int v = 0; v < 5; v++) dst[off + v * pitch] = p[v];
    }
}

// One stage along Y (DIR 2) or Z (DIR 3): each lane marches one segment of a
// column. At load k the window gains cell m = j0 - 2 + k; the corrected
// fluxes of cell m-1 complete and cell m-2 is written (from k = 4 on).
template <int DIR, bool CORRECT>
__global__ __launch_bounds__(256) void xj_stage_march(const double *__restrict__ src, const double *base,
                                                      double *dst, int64_t pitch, int nx, int64_t sstride,
                                                      int64_t ostride, int N, int S, int seglen, double lambda,
                                                      const double *__restrict__ lamp, double lscale, XjConsts k,
                                                      const double *__restrict__ cf)
{
    if (lamp) lambda = lscale * lamp[0];  // time step kept on the device
    constexpr int NS = DIR + 1;
    const int lane = threadIdx.x & 63;
    const int W = 64 / S;
    const int s = lane / W;
    const int64_t x = (int64_t)blockIdx.x * (blockDim.x / S) + (threadIdx.x >> 6) * W + (lane - s * W);
    if (x >= nx || s >= S) return;
    const int j0 = s * seglen;
    int L = N - j0;
    L = L < seglen ? L : seglen;
    if (L <= 0) return;
    const int64_t colOff = x + (int64_t)blockIdx.y * ostride;
    const double cfl = cf[x + (int64_t)nx * blockIdx.y];
    int mw = j0 - 2;
    if (mw < 0) mw += N;
    int ow = j0;  // wrapped index of the next output cell (never wraps: j0 + L <= N)

    double wlp[5], wrp[5], dlp[5], drp[5], WLp[5], WRp[5], WRpp[5];
#pragma unroll
    for (int v = 0; v < 5; v++) wlp[v] = wrp[v] = dlp[v] = drp[v] = WLp[v] = WRp[v] = WRpp[v] = 0.0;

    for (int t = 0; t < L + 4; t++) {
        double q[5];
        const int64_t o = colOff + (int64_t)mw * sstride;
#pragma unroll
        for (int v = 0; v < 5; v++) q[v] = src[o + v * pitch];
        mw = (mw + 1 == N) ? 0 : mw + 1;
        double wl[5], wr[5];
        xj_cell<NS>(k, cfl, q, wl, wr);
        double p[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            const double dl = .5 * (wl[v] - wlp[v]);
            const double dr = .5 * (wr[v] - wrp[v]);
            // corrected fluxes of cell m-1
            double WL = wlp[v], WR = wrp[v];
            if (CORRECT) {
                WL -= xj_ospre(dlp[v], dl);
                WR += xj_ospre(drp[v], dr);
            }
            // flux difference at cell m-2
            p[v] = WRp[v] - WRpp[v] - WL + WLp[v];
            wlp[v] = wl[v];
            wrp[v] = wr[v];
            dlp[v] = dl;
            drp[v] = dr;
            WRpp[v] = WRp[v];
            WRp[v] = WR;
            WLp[v] = WL;
        }
        if (t >= 4) {
            const int64_t oo = colOff + (int64_t)ow * sstride;
            ow++;
            double r[5];
#pragma unroll
            for (int v = 0; v < 5; v++) r[v] = base[oo + v * pitch] - lambda * p[v];
            xj_floor(k, r);
#pragma unroll
            for (int v = 0; v < 5; v++) dst[oo + v * pitch] = r[v];
        }
    }
}

static int xj_freeze(const double *q, int64_t pitch, int nx, int ny, int nz, int dir, const XjConsts &k,
                     double *cf, hipStream_t s)
{
    if (dir == 1) {
        const int64_t nlines = (int64_t)ny * nz;
        hipLaunchKernelGGL(xj_freeze_x, dim3((unsigned)((nlines + 3) / 4)), dim3(256), 0, s, q, pitch, nx, nlines,
                           k, cf);
    } else {
        const int N = (dir == 2) ? ny : nz;
        const int64_t sstride = (dir == 2) ? (int64_t)nx : (int64_t)nx * ny;
        const int64_t ostride = (dir == 2) ? (int64_t)nx * ny : (int64_t)nx;
        const int nother = (dir == 2) ? nz : ny;
        hipLaunchKernelGGL(xj_freeze_march, dim3((unsigned)((nx + 255) / 256), (unsigned)nother), dim3(256), 0, s,
                           q, pitch, nx, sstride, ostride, N, k, cf);
    }
    return IM_LAUNCHED("xj_freeze");
}

template <bool CORRECT>
static int xj_stage(const double *src, const double *base, double *dst, int64_t pitch, int nx, int ny, int nz,
                    int dir, double lambda, const double *lamp, double lscale, const XjConsts &k, const double *cf,
                    hipStream_t s)
{
    if (dir == 1) {
        const int64_t nlines = (int64_t)ny * nz;
        const int nseg = (nx + 59) / 60;
        const int64_t waves = nlines * nseg;
        hipLaunchKernelGGL((xj_stage_x<CORRECT>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, src, base,
                           dst, pitch, nx, nlines, nseg, lambda, lamp, lscale, k, cf);
        return IM_LAUNCHED("xj_stage_x");
    }
    const int N = (dir == 2) ? ny : nz;
    const int64_t sstride = (dir == 2) ? (int64_t)nx : (int64_t)nx * ny;
    const int64_t ostride = (dir == 2) ? (int64_t)nx * ny : (int64_t)nx;
    const int nother = (dir == 2) ? nz : ny;
    const int64_t columns = (int64_t)nx * nother;
    int S = 1;
    while (S < 4 && columns * S < 2048 * 64 && N >= 32 * S) S *= 2;
    const int seglen = (N + S - 1) / S;
    const int colsPerBlock = 256 / S;
    dim3 grid((unsigned)((nx + colsPerBlock - 1) / colsPerBlock), (unsigned)nother), block(256);
    if (dir == 2)
        hipLaunchKernelGGL((xj_stage_march<2, CORRECT>), grid, block, 0, s, src, base, dst, pitch, nx, sstride,
                           ostride, N, S, seglen, lambda, lamp, lscale, k, cf);
    else
        hipLaunchKernelGGL((xj_stage_march<3, CORRECT>), grid, block, 0, s, src, base, dst, pitch, nx, sstride,
                           ostride, N, S, seglen, lambda, lamp, lscale, k, cf);
    return IM_LAUNCHED("xj_stage_march");
}

int64_t xinjin_scratch_bytes(int64_t pitch, int nx, int ny, int nz, int dir)
{
    const int N = (dir == 1) ? nx : (dir == 2 ? ny : nz);
    if (dir < 1 || dir > 3 || N < 3) return 0;
    const int64_t nlines = (dir == 1) ? (int64_t)ny * nz : (dir == 2 ? (int64_t)nx * nz : (int64_t)nx * ny);
    return (5 * pitch + nlines + 32) * (int64_t)sizeof(double);
}

int launch_xinjin_sweep(double *base, int64_t pitch, int nx, int ny, int nz, int dir, double lambda,
                        double gamma, double rhomin, hipStream_t s, const double *lamp)
{
    if (dir < 1 || dir > 3) return IM_FAIL(ERROR_INVALID_ARGS, "sweep direction %d", dir);
    const int N = (dir == 1) ? nx : (dir == 2 ? ny : nz);
    if (N < 3) return SUCCESSFUL;  // cudaFluidStep.cu:301
    if ((int64_t)nx * ny * nz > pitch) return IM_FAIL(ERROR_INVALID_ARGS, "slab pitch below the cell count");
    XjConsts k{gamma, gamma - 1.0, rhomin, 1e-5 / (gamma - 1.0)};
    const int64_t nlines = (dir == 1) ? (int64_t)ny * nz : (dir == 2 ? (int64_t)nx * nz : (int64_t)nx * ny);
    // temporary: 5 slabs of U at the fluid's pitch, then the per-line freezing speeds
    double *U = (double *)scratch_block((size_t)xinjin_scratch_bytes(pitch, nx, ny, nz, dir));
    if (!U) return IM_FAIL(ERROR_NOMEM, "no scratch for the Xin-Jin stages");
    double *cf = U + 5 * pitch;
    const int slot = prof_begin(PROF_SWEEP_X + dir - 1, s);
    int rc = xj_freeze(base, pitch, nx, ny, nz, dir, k, cf, s);
    if (!rc) rc = xj_stage<false>(base, base, U, pitch, nx, ny, nz, dir, 0.25 * lambda, lamp, 0.25, k, cf, s);
    if (!rc) rc = xj_freeze(U, pitch, nx, ny, nz, dir, k, cf, s);
    if (!rc) rc = xj_stage<true>(U, base, base, pitch, nx, ny, nz, dir, 0.5 * lambda, lamp, 0.5, k, cf, s);
    prof_end(slot, s);
    return rc;
}

}  // namespace imogen
