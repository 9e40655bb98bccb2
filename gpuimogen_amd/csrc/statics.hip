// statics.hip -- ghost-cell boundary conditions of the fluid arrays.
//
// Restates the kernel table of setBoundarySAS / callBCKernel
// (gpuclass/cudaStatics.cu:1086-1240) with its per-direction index
// conventions (cudaStatics.cu:1290-1532), including the reference's quirks:
//   * z-minus mirror reflects about cell 4.5 (index 2*depth+1-q), x/y about 4;
//   * +y and +z 'const' copy the innermost ghost cell (n-depth), +x copies the
//     last interior cell (n-depth-1).
// The fluid step sets each sweep's boundaries with the memory-x kernels'
// conventions along the swept direction (xrules: the reference rotates that
// direction into memory x first, flux.cu:99-107), so these quirks act only in
// setFluidBoundary / setArrayBoundaryConditions calls on unrotated arrays.
// One launch handles both faces of one direction for all 5 slabs when neither
// face reads a cell the other writes (lines of at least 3*depth+2 cells);
// otherwise the minus face runs first, then the plus face, the reference's
// order.
#include "imogen_internal.h"

namespace imogen {

// sas codes per (slab, side) at [2*slab + side]: -1 none, 0 symmetric,
// 1 antisymmetric, 2 constant, 3 linear
struct SasTable {
    int s[10];
};

__global__ void bc_face_x(double *__restrict__ q, int64_t pitch, int nx, int64_t rows, int side0, int nsides,
                          int d, SasTable t)
{
    // one thread per (ghost cell, slab, side, row)
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = rows * 5 * d * nsides;
    if (tid >= total) return;
    const int g = (int)(tid % d);
    int64_t rest = tid / d;
    const int slab = (int)(rest % 5);
    rest /= 5;
    const int side = side0 + (int)(rest % nsides);
    const int64_t row = rest / nsides;
    const int sas = t.s[2 * slab + side];
    if (sas < 0) return;
    double *r = q + slab * pitch + row * nx;
    if (side == 0) {
        switch (sas) {
        case 0: r[d - 1 - g] = r[d + 1 + g]; break;            // cudaStatics.cu:1308-1313
        case 1: r[d - 1 - g] = -r[d + 1 + g]; break;           // :1315-1320
        case 2: r[g] = r[d]; break;                            // :1339-1344
        case 3: r[g] = r[d] + (d - g) * (r[d] - r[d + 1]); break;  // :1372-1380
        }
    } else {
        switch (sas) {
        case 0: r[nx - d + g] = r[nx - d - 2 - g]; break;      // :1322-1327
        case 1: r[nx - d + g] = -r[nx - d - 2 - g]; break;     // :1329-1334
        case 2: r[nx - d + g] = r[nx - d - 1]; break;          // :1361-1366
        }
    }
}

// memdir 2 (y) and 3 (z): one thread per (face cell, slab), looping over depth.
// xrules: the memory-x kernels' index conventions along this axis (what the
// reference's fluid step applies: it rotates the swept direction into memory
// x before setting its boundaries, flux.cu:99-107), i.e. no z-minus mirror
// offset, +side const from the last interior cell, and -side linear.
__global__ void bc_face_yz(double *__restrict__ q, int64_t pitch, int nx, int64_t n1, int64_t n2,
                           int64_t stride, int64_t ostride, int N, int zdir, int side0, int d, SasTable t,
                           int xrules)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t face = (int64_t)nx * n2;
    const int slab = (int)(blockIdx.y % 5);
    const int side = side0 + (int)(blockIdx.y / 5);
    if (tid >= face) return;
    const int sas = t.s[2 * slab + side];
    if (sas < 0) return;
    (void)n1;
    const int64_t x = tid % nx;
    const int64_t o = tid / nx;
    double *c = q + slab * pitch + x + o * ostride;
    if (side == 0) {
        const int off = (zdir && !xrules) ? 2 * d + 1 : 2 * d;  // z-minus mirror quirk (:1496-1500)
        switch (sas) {
        case 0:
            for (int g = 0; g < d; g++) c[g * stride] = c[(off - g) * stride];
            break;
        case 1:
            for (int g = 0; g < d; g++) c[g * stride] = -c[(off - g) * stride];
            break;
        case 2: {
            const double f = c[d * stride];
            for (int g = 0; g < d; g++) c[g * stride] = f;
        } break;
        case 3: {  // -x linear extrapolation (:1372-1380), xrules only
            const double f0 = c[d * stride], f1 = c[(d + 1) * stride];
            for (int g = 0; g < d; g++) c[g * stride] = f0 + (d - g) * (f0 - f1);
        } break;
        }
    } else {
        switch (sas) {
        case 0:
            for (int g = 0; g < d; g++) c[(N - 1 - g) * stride] = c[(N - 1 - 2 * d + g) * stride];
            break;
        case 1:
            for (int g = 0; g < d; g++) c[(N - 1 - g) * stride] = -c[(N - 1 - 2 * d + g) * stride];
            break;
        case 2: {
            // y/z kernels copy the innermost ghost (:1461-1466, :1524-1532), the
            // x kernel the last interior cell (:1361-1366)
            const double f = c[(xrules ? N - d - 1 : N - d) * stride];
            for (int g = 0; g < d; g++) c[(N - d + g) * stride] = f;
        } break;
        }
    }
}

// Outflow ghosts (BCModeTypes outflow, gpuclass/cudaCommon.h:46) of the faces
// side0 .. side0 + nsides - 1 of direction dir, all five slabs: one thread per
// (line along dir, side), the line's decision taken once from the normal
// momentum of the interior cell next to the face (cell d or N-1-d). Outward
// flow: the ghosts copy that cell. Otherwise the rules of cukern_SetOutflowX's
// +x branch (cudaStatics.cu:486-531), which the fluid step applies along every
// direction after rotating it into memory x: density copied, slab 2 zeroed,
// slab 3 set to rho times the rest-frame speed on x faces (1.05 on -x,
// -1.198 - 0.0047 i on +x, :382-385) or to rho times the cell's own velocity
// elsewhere, slab 4 zeroed, and the total energy changed by the kinetic
// energy of each replaced momentum in that order. Reads and writes stay
// inside the line (the reference's -x kernel starts depth cells before it
// and its y/z kernels read the wrong rows, cudaStatics.cu:539-560, 581, 612).
__global__ void bc_outflow(double *__restrict__ q, int64_t pitch, int nx, int64_t n2, int64_t stride,
                           int64_t ostride, int N, int dir, int side0, int nsides, int d)
{
    // evaluated as written (no FMA contraction): the energy update uses the
    // stored momenta p2 = rho v, as the reference's kernel does
#pragma clang fp contract(off)
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t lines = (int64_t)((dir == 1) ? 1 : nx) * n2;  // dir 1: rows (y, z); else (x, other)
    if (tid >= lines * nsides) return;
    const int side = side0 + (int)(tid / lines);
    const int64_t l = tid % lines;
    double *r = (dir == 1) ? q + l * ostride : q + (l % nx) + (l / nx) * ostride;
    const int64_t c = (side ? N - 1 - d : d) * stride;
    const double a = r[(1 + dir) * pitch + c];
    const double dirsign = side ? 1.0 : -1.0;
    if (a * dirsign > 0) {
        for (int sl = 0; sl < 5; sl++) {
            const double v = r[sl * pitch + c];
            for (int k = 0; k < d; k++) r[sl * pitch + (side ? N - d + k : k) * stride] = v;
        }
        return;
    }
    const double rho = r[c], E0 = r[pitch + c];
    const double px = r[2 * pitch + c], py = r[3 * pitch + c], pz = r[4 * pitch + c];
    for (int k = 0; k < d; k++) {
        const int64_t gi = (side ? N - d + k : k) * stride;
        double E = E0;
        double p2 = rho * 0.0;
        r[gi] = rho;
        r[2 * pitch + gi] = p2;
        E += .5 * (p2 - px) * (p2 + px) / rho;
        const double v = (dir == 1) ? (side ? -1.198 - .0047 * k : 1.05) : py / rho;
        p2 = rho * v;
        r[3 * pitch + gi] = p2;
        E += .5 * (p2 - py) * (p2 + py) / rho;
        p2 = rho * 0.0;
        r[4 * pitch + gi] = p2;
        E += .5 * (p2 - pz) * (p2 + pz) / rho;
        r[pitch + gi] = E;
    }
}

int launch_outflow_bc(double *base, int64_t pitch, int nx, int ny, int nz, int dir, int side0, int nsides,
                      int depth, hipStream_t s)
{
    const int N = (dir == 1) ? nx : (dir == 2 ? ny : nz);
    if (N < depth + 2) return IM_FAIL(ERROR_INVALID_ARGS, "outflow BC needs >= %d cells (got %d)", depth + 2, N);
    if (nsides == 2 && N < 2 * depth + 2) return IM_FAIL(ERROR_INVALID_ARGS, "outflow faces overlap");
    int64_t n2, stride, ostride;
    if (dir == 1) {
        n2 = (int64_t)ny * nz;
        stride = 1;
        ostride = nx;
    } else {
        n2 = (dir == 2) ? nz : ny;
        stride = (dir == 2) ? (int64_t)nx : (int64_t)nx * ny;
        ostride = (dir == 2) ? (int64_t)nx * ny : (int64_t)nx;
    }
    const int64_t threads = ((dir == 1) ? 1 : (int64_t)nx) * n2 * nsides;
    hipLaunchKernelGGL(bc_outflow, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, base, pitch, nx, n2,
                       stride, ostride, N, dir, side0, nsides, depth);
    return IM_LAUNCHED("bc_outflow");
}

// Freebalance ghosts of one face (cukern_SetFreeBalance{X,Y,Z}, cudaStatics.cu:859-1003):
// one thread per face cell computes all its ghosts from the interior cell next
// to the face, in the reference's operation order (no FMA contraction).
__global__ void bc_freebalance(double *__restrict__ q, int64_t pitch, const double *__restrict__ phi, int nx,
                               int64_t n2, int64_t stride, int64_t ostride, int N, int dir, int side, int d,
                               FreeBalanceConsts fc)
{
#pragma clang fp contract(off)
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t lines = (int64_t)((dir == 1) ? 1 : nx) * n2;
    if (tid >= lines) return;
    const int64_t off = (dir == 1) ? tid * ostride : (tid % nx) + (tid / nx) * ostride;
    double *r = q + off;
    const double *ph = phi + off;
    const int64_t c = (int64_t)(side ? N - 1 - d : d) * stride;
    const double rho = r[c];
    const double vx = r[2 * pitch + c] / rho;
    const double vy = r[3 * pitch + c] / rho;
    double vz = r[4 * pitch + c] / rho;
    const double E = r[pitch + c] - .5 * rho * (vx * vx + vy * vy + vz * vz);
    const double T = E / rho;
    const double phi0 = ph[c];
    const double direct = side ? 1.0 : -1.0;
    for (int k = 0; k < d; k++) {
        const int64_t gi = (int64_t)(side ? N - d + k : k) * stride;
        double phi1 = ph[gi], rho2, th;
        if (rho < 1e-7) {  // the tenuous (solar-wind) side of the disk setups, :908-910
            phi1 = phi0;
            vz = (vz * direct < 0) ? 0 : vz;
        }
        if (dir != 1) {
            rho2 = rho * exp(-fc.invgm1 * (phi1 - phi0) / T);
            th = vy;
        } else {
            th = vy + fc.vFrame;
            const double cp = th * th * fc.x0[k];
            th = th * fc.qh[k] - fc.vFrame * fc.qq[k];
            rho2 = rho * exp(-fc.invgm1 * (phi1 - phi0 - cp) / T);
        }
        r[gi] = rho2;
        r[2 * pitch + gi] = rho2 * vx;
        r[3 * pitch + gi] = rho2 * th;
        r[4 * pitch + gi] = rho2 * vz;
        r[pitch + gi] = rho2 * T + .5 * rho2 * (vx * vx + th * th + vz * vz);
    }
}

int launch_freebalance_bc(double *base, int64_t pitch, const double *phi, int nx, int ny, int nz, int dir,
                          int side, int depth, const FreeBalanceConsts &fc, hipStream_t s)
{
    const int N = (dir == 1) ? nx : (dir == 2 ? ny : nz);
    if (depth < 1 || depth > kMaxFreeBalanceDepth)
        return IM_FAIL(ERROR_INVALID_ARGS, "freebalance depth %d (1..%d)", depth, kMaxFreeBalanceDepth);
    if (N < depth + 2) return IM_FAIL(ERROR_INVALID_ARGS, "freebalance BC needs >= %d cells (got %d)", depth + 2, N);
    int64_t n2, stride, ostride;
    if (dir == 1) {
        n2 = (int64_t)ny * nz;
        stride = 1;
        ostride = nx;
    } else {
        n2 = (dir == 2) ? nz : ny;
        stride = (dir == 2) ? (int64_t)nx : (int64_t)nx * ny;
        ostride = (dir == 2) ? (int64_t)nx * ny : (int64_t)nx;
    }
    const int64_t threads = ((dir == 1) ? 1 : (int64_t)nx) * n2;
    hipLaunchKernelGGL(bc_freebalance, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, base, pitch, phi,
                       nx, n2, stride, ostride, N, dir, side, depth, fc);
    return IM_LAUNCHED("bc_freebalance");
}

// Static cells of one array (cukern_applySpecial_fade, cudaStatics.cu:1251-1281):
// entry s holds the cell index table[s], the value f0 = table[s + blk] and the
// coefficient c = table[s + 2 blk]. Indices outside [0, numel) are skipped
// (the reference does not check them).
__global__ void static_fade(double *__restrict__ phi, const double *__restrict__ table, int nspec, int64_t blk,
                            int64_t numel)
{
#pragma clang fp contract(off)
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= nspec) return;
    const double xi = table[s];
    if (!(xi >= 0.0 && xi < (double)numel)) return;
    const int64_t x = (int64_t)xi;
    const double f0 = table[s + blk], c = table[s + 2 * blk];
    if (c != 1.0) {
        double w = f0 * c + (1.0 - c) * phi[x];
        if (w != w) w = f0 * c;
        phi[x] = w;
    } else {
        phi[x] = f0;
    }
}

int launch_static_fade(double *phi, const double *table, int nspec, int64_t blk, int64_t numel, hipStream_t s)
{
    if (nspec <= 0) return SUCCESSFUL;
    hipLaunchKernelGGL(static_fade, dim3((unsigned)((nspec + 255) / 256)), dim3(256), 0, s, phi, table, nspec, blk,
                       numel);
    return IM_LAUNCHED("static_fade");
}

int launch_fluid_bc(double *base, int64_t pitch, int nx, int ny, int nz, int dir, const int *modes10,
                    int depth, hipStream_t s, int xrules)
{
    const int N = (dir == 1) ? nx : (dir == 2 ? ny : nz);
    SasTable t;
    bool any[2] = {false, false};
    for (int k = 0; k < 10; k++) {
        const int side = k & 1;
        t.s[k] = modes10[k];
        if (t.s[k] < 0) continue;
        any[side] = true;
        if ((t.s[k] == 0 || t.s[k] == 1) && N < 2 * depth + 2)
            return IM_FAIL(ERROR_INVALID_ARGS, "mirror BC needs >= %d cells (got %d)", 2 * depth + 2, N);
        if (N < depth + 2) return IM_FAIL(ERROR_INVALID_ARGS, "BC needs >= %d cells (got %d)", depth + 2, N);
        if (t.s[k] == 3 && ((dir != 1 && !xrules) || side != 0))
            return IM_FAIL(ERROR_NOIMPLEMENT, "linear extrapolation BC exists only on the -x face");
    }
    // Both faces in one launch when neither reads a cell the other writes:
    // a mirror face reads up to 2 depth (z-minus: 2 depth + 1) cells in from
    // its edge, so the faces are disjoint for N >= 3 depth + 2. Otherwise
    // side 0, then side 1, the reference's order.
    const bool together = any[0] && any[1] && N >= 3 * depth + 2;
    for (int side = 0; side < 2; side++) {
        if (!any[side]) continue;
        const int nsides = together ? 2 : 1;
        if (dir == 1) {
            const int64_t rows = (int64_t)ny * nz;
            const int64_t total = rows * 5 * depth * nsides;
            hipLaunchKernelGGL(bc_face_x, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, base, pitch,
                               nx, rows, side, nsides, depth, t);
        } else {
            const int64_t stride = (dir == 2) ? (int64_t)nx : (int64_t)nx * ny;
            const int64_t ostride = (dir == 2) ? (int64_t)nx * ny : (int64_t)nx;
            const int64_t n2 = (dir == 2) ? nz : ny;
            const int64_t face = (int64_t)nx * n2;
            hipLaunchKernelGGL(bc_face_yz, dim3((unsigned)((face + 255) / 256), 5 * nsides), dim3(256), 0, s, base,
                               pitch, nx, (int64_t)N, n2, stride, ostride, N, dir == 3 ? 1 : 0, side, depth, t,
                               xrules);
        }
        int rc = IM_LAUNCHED("boundary condition kernel");
        if (rc) return rc;
        if (together) break;
    }
    return SUCCESSFUL;
}

}  // namespace imogen
