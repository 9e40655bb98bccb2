// cfl.hip -- sound speed and the CFL time-step reduction.
//
// Per-cell formulas follow cukern_Soundspeed_hd (gpuclass/cudaSoundspeed.cu:178-195)
// and cukern_CFLtimestep (gpuclass/cflTimestep.cu:262-377, SQUARE geometry).
// The reduction is this library's: each lane keeps a running minimum over a
// grid-stride loop, the wave reduces with 64-lane xor shuffles, waves combine
// through LDS, and a one-block second pass reduces the per-block minima, so
// the result is deterministic and no host round trip per block is needed.
// Comparisons use `u < m ? u : m` exactly as the reference does, so NaN cells
// are skipped as they are there; the standalone pass also counts them into the
// device error word, so the entry points can report the invalid state.
#include "cfl_device.h"
#include "imogen_internal.h"

namespace imogen {

__global__ void soundspeed_kernel(const double *__restrict__ q, int64_t pitch, int64_t n, double gg1,
                                  double *__restrict__ cs)
{
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x)
        cs[x] = cs_of(q[x], q[pitch + x], q[2 * pitch + x], q[3 * pitch + x], q[4 * pitch + x], gg1);
}

// Two cells per lane per load (16-byte loads) and two such pairs in flight
// per iteration. Requires an even slab pitch and 16-byte aligned slabs (the
// 256-byte slab pitch guarantees both).
__global__ __launch_bounds__(256) void cfl_pass1(const double *__restrict__ q, int64_t pitch, int64_t n,
                                                 const double *__restrict__ cs, CflArgs a,
                                                 double *__restrict__ blockMin, unsigned long long *__restrict__ err)
{
    double m = 1e37;
    unsigned nan = 0;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    const int64_t n2 = n / 2;
    const double2 *q2 = (const double2 *)q;
    const double2 *c2 = (const double2 *)cs;
    const int64_t p2 = pitch / 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += 2 * nth) {
        const int64_t i2 = (i + nth < n2) ? i + nth : i;
        double2 r[2], e[2], x[2], y[2], z[2], c[2];
        r[0] = q2[i];
        e[0] = q2[p2 + i];
        x[0] = q2[2 * p2 + i];
        y[0] = q2[3 * p2 + i];
        z[0] = q2[4 * p2 + i];
        r[1] = q2[i2];
        e[1] = q2[p2 + i2];
        x[1] = q2[2 * p2 + i2];
        y[1] = q2[3 * p2 + i2];
        z[1] = q2[4 * p2 + i2];
        if (!a.fused) {
            c[0] = c2[i];
            c[1] = c2[i2];
        } else {
            c[0] = c[1] = make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const bool fresh = (h == 0) || (i2 != i);  // i2 == i revisits the first pair
            double u = cfl_cell(a, r[h].x, e[h].x, x[h].x, y[h].x, z[h].x, c[h].x);
            m = (u < m) ? u : m;
            nan += (u != u) && fresh;
            u = cfl_cell(a, r[h].y, e[h].y, x[h].y, y[h].y, z[h].y, c[h].y);
            m = (u < m) ? u : m;
            nan += (u != u) && fresh;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) {
        const int64_t t = n - 1;
        const double u = cfl_cell(a, q[t], q[pitch + t], q[2 * pitch + t], q[3 * pitch + t], q[4 * pitch + t],
                                  a.fused ? 0.0 : cs[t]);
        m = (u < m) ? u : m;
        nan += (u != u);
    }
    wave_count(err + ERRW_CFL_NAN, nan);
    m = wave_min(m);
    __shared__ double part[4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) part[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = part[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) r = (part[k] < r) ? part[k] : r;
        blockMin[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(256) void cfl_pass2(const double *__restrict__ blockMin, int nblocks,
                                                 double halve, double *__restrict__ out)
{
    double m = 1e38;
    for (int k = threadIdx.x; k < nblocks; k += blockDim.x) m = (blockMin[k] < m) ? blockMin[k] : m;
    m = wave_min(m);
    __shared__ double part[4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) part[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = part[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) r = (part[k] < r) ? part[k] : r;
        out[0] = halve ? r / 2 : r;  // cflTimestep.cu:232-237
    }
}

int launch_soundspeed(const double *base, int64_t pitch, int64_t n, double gamma, double *cs, hipStream_t s)
{
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(soundspeed_kernel, dim3((unsigned)blocks), dim3(256), 0, s, base, pitch, n,
                       gamma * (gamma - 1), cs);
    return IM_LAUNCHED("soundspeed_kernel");
}

// devOut[0] receives tau; devOut[1..] is used as block scratch (<= 2048 blocks).
int launch_cfl(const double *base, int64_t pitch, int64_t n, const double *cs, double gamma, int fused,
               int dim, int xinjin, int halve, double hx, double hy, double hz, double hmin, double *devOut,
               hipStream_t s)
{
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    unsigned long long *err = st->err;
    int64_t blocks = (n / 4 + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    if ((pitch & 1) || (((uintptr_t)base | (uintptr_t)cs) & 15))
        return IM_FAIL(ERROR_INVALID_ARGS, "CFL needs 16-byte aligned slabs with an even pitch");
    double *blockMin = devOut + 8;
    CflArgs a{gamma * (gamma - 1), hx, hy, hz, hmin, fused, dim, xinjin};
    hipLaunchKernelGGL(cfl_pass1, dim3((unsigned)blocks), dim3(256), 0, s, base, pitch, n, cs, a, blockMin, err);
    int rc = IM_LAUNCHED("cfl_pass1");
    if (rc) return rc;
    hipLaunchKernelGGL(cfl_pass2, dim3(1), dim3(256), 0, s, blockMin, (int)blocks, halve ? 1.0 : 0.0, devOut);
    return IM_LAUNCHED("cfl_pass2");
}

int launch_cfl_final(const double *part, int64_t n, int halve, double *devOut, hipStream_t s)
{
    if (n < 1 || n > (int64_t)1 << 30) return IM_FAIL(ERROR_INVALID_ARGS, "CFL partial count %lld", (long long)n);
    hipLaunchKernelGGL(cfl_pass2, dim3(1), dim3(256), 0, s, part, (int)n, halve ? 1.0 : 0.0, devOut);
    return IM_LAUNCHED("cfl_pass2");
}

}  // namespace imogen
