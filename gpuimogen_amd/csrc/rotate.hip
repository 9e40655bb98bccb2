// rotate.hip -- array index exchanges (cudaArrayRotateB) and the one-rank
// halo exchange.
//
// The reference's 1D fluid update always sweeps memory-x (the direction
// argument only picks the normal momentum slab and h), so its flux.m debug
// path rotates every fluid array before each sweep (flux/flux.m:55-62, codes
// from its preperm / permcall tables) with flipArrayIndices
// (gpuclass/cudaArrayRotateB.cu:103-300): the data is transposed in memory and
// the array's metadata (dims, currentPermutation, permtag, partitionDir,
// partNumel: alterArrayMetadata, :583-640) says which physical axis each
// memory axis now holds. Here one gather kernel performs every exchange code
// (one thread per destination element: coalesced writes, strided reads; this
// path runs outside the time-critical loop, whose sweeps walk each direction's
// own stride instead of transposing).
#include <cstdlib>
#include <cstring>

#include "imogen_internal.h"

namespace imogen {

// dst[i0 + d0 (i1 + d1 i2)] = src[i0 st0 + i1 st1 + i2 st2]: destination axis
// a walks the source axis whose stride is st_a
__global__ __launch_bounds__(256) void permute_copy_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                                           int64_t d0, int64_t d1, int64_t st0, int64_t st1,
                                                           int64_t st2, int64_t n)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i0 = t % d0;
        const int64_t r = t / d0;
        const int64_t i1 = r % d1;
        const int64_t i2 = r / d1;
        dst[t] = src[i0 * st0 + i1 * st1 + i2 * st2];
    }
}

// source axis held by destination axis a after exchange `code`
// (alterArrayMetadata, cudaArrayRotateB.cu:583-630)
static int code_axes(int code, int ax[3])
{
    switch (code) {
    case 2: ax[0] = 1; ax[1] = 0; ax[2] = 2; return 0;  // XYZ -> YXZ
    case 3: ax[0] = 2; ax[1] = 1; ax[2] = 0; return 0;  // XYZ -> ZYX
    case 4: ax[0] = 0; ax[1] = 2; ax[2] = 1; return 0;  // XYZ -> XZY
    case 5: ax[0] = 1; ax[1] = 2; ax[2] = 0; return 0;  // rotate left, XYZ -> YZX
    case 6: ax[0] = 2; ax[1] = 0; ax[2] = 1; return 0;  // rotate right, XYZ -> ZXY
    default: return 1;
    }
}

int launch_permute_copy(const double *src, double *dst, const int sdim[3], const int ax[3], hipStream_t s)
{
    const int64_t n = (int64_t)sdim[0] * sdim[1] * sdim[2];
    if (n <= 0) return SUCCESSFUL;
    const int64_t stride[3] = {1, (int64_t)sdim[0], (int64_t)sdim[0] * sdim[1]};
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(permute_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, (int64_t)sdim[ax[0]],
                       (int64_t)sdim[ax[1]], stride[ax[0]], stride[ax[1]], stride[ax[2]], n);
    return IM_LAUNCHED("permute_copy_kernel");
}

// Exchanges that leave the memory order unchanged (cudaArrayRotateB.cu:551-579)
static int needs_reorder(const int *dims, int code)
{
    switch (code) {
    case 2:
        if (dims[0] == 1 || dims[1] == 1) return 0;
        break;
    case 3:
        if (dims[1] == 1 && (dims[0] == 1 || dims[2] == 1)) return 0;
        break;
    case 4:
        if (dims[1] == 1 || dims[2] == 1) return 0;
        break;
    case 5:
        if (dims[0] == 1) return 0;
        if (dims[1] == 1 && dims[2] == 1) return 0;
        break;
    case 6:
        if (dims[2] == 1) return 0;
        if (dims[0] == 1 && dims[1] == 1) return 0;
        break;
    }
    return 1;
}

}  // namespace imogen

using namespace imogen;

extern "C" {

int alterArrayMetadata(MGArray *src, MGArray *dst, int exchangeCode)
{
    if (!src || !dst) return IM_FAIL(ERROR_NULL_POINTER, "null array");
    int ax[3];
    if (code_axes(exchangeCode, ax))
        return IM_FAIL(ERROR_INVALID_ARGS, "index exchange code %d is not in 2-6", exchangeCode);
    const MGArray o = *src;  // src may be dst
    for (int a = 0; a < 3; a++) {
        dst->dim[a] = o.dim[ax[a]];
        dst->currentPermutation[a] = o.currentPermutation[ax[a]];
    }
    // the partition direction follows its axis
    for (int a = 0; a < 3; a++)
        if (o.partitionDir == ax[a] + 1) dst->partitionDir = a + 1;
    dst->permtag = MGA_numsToPermtag(dst->currentPermutation);
    for (int i = 0; i < dst->nGPUs; i++) {
        int sub[6];
        calcPartitionExtent(dst, i, sub);
        dst->partNumel[i] = sub[3] * sub[4] * sub[5];
    }
    return SUCCESSFUL;
}

int flipArrayIndices(MGArray *phi, MGArray **newArrays, int nArrays, int exchangeCode, void *streamPtrs,
                     MGArray *tempStorage)
{
    (void)streamPtrs;
    (void)tempStorage;
    if (!phi || nArrays < 1) return IM_FAIL(ERROR_NULL_POINTER, "no arrays");
    int ax[3];
    if (code_axes(exchangeCode, ax))
        return IM_FAIL(ERROR_INVALID_ARGS, "index exchange code %d is not in 2-6", exchangeCode);
    for (int j = 0; j < nArrays; j++)
        if (phi[j].nGPUs != 1) return IM_FAIL(ERROR_NOIMPLEMENT, "single-GPU arrays only");
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    MGArray *psi = nullptr;
    if (newArrays) {
        if (!newArrays[0]) {
            newArrays[0] = (MGArray *)calloc((size_t)nArrays, sizeof(MGArray));
            if (!newArrays[0]) return ERROR_NOMEM;
        }
        psi = newArrays[0];
    }
    double *tmp = nullptr;
    size_t tmpBytes = 0;
    int rc = SUCCESSFUL;
    for (int j = 0; j < nArrays && !rc; j++) {
        MGArray *a = &phi[j];
        MGArray trans = *a;
        if ((rc = alterArrayMetadata(a, &trans, exchangeCode))) break;
        const size_t bytes = (size_t)a->partNumel[0] * sizeof(double);
        const bool reorder = needs_reorder(a->dim, exchangeCode) != 0;
        if (psi) {
            // a new array of the exchanged shape (cudaArrayRotateB.cu:168-175)
            MGArray *nu = &psi[j];
            if ((rc = MGA_allocArrays(&nu, 1, &trans))) break;
            if (reorder) rc = launch_permute_copy(a->devicePtr[0], nu->devicePtr[0], a->dim, ax, st->stream);
            else rc = IM_HIP(hipMemcpyAsync(nu->devicePtr[0], a->devicePtr[0], bytes, hipMemcpyDeviceToDevice,
                                            st->stream));
            continue;
        }
        if (reorder) {
            // in place through a staging buffer (cudaArrayRotateB.cu:263-284)
            if (bytes > tmpBytes) {
                if (tmp) {
                    (void)hipStreamSynchronize(st->stream);
                    (void)hipFree(tmp);
                    tmp = nullptr;
                }
                if ((rc = IM_HIP(hipMalloc(&tmp, bytes)))) {
                    tmp = nullptr;
                    break;
                }
                tmpBytes = bytes;
            }
            if ((rc = launch_permute_copy(a->devicePtr[0], tmp, a->dim, ax, st->stream))) break;
            if ((rc = IM_HIP(hipMemcpyAsync(a->devicePtr[0], tmp, bytes, hipMemcpyDeviceToDevice, st->stream))))
                break;
        }
        // keep the data pointer and ownership, take the exchanged metadata
        trans.devicePtr[0] = a->devicePtr[0];
        *a = trans;
    }
    if (tmp) {
        (void)hipStreamSynchronize(st->stream);
        (void)hipFree(tmp);
    }
    return rc;
}

// gpuclass/cudaHaloExchange.cu:140: on one rank there is nothing to exchange
// (the sweeps wrap periodic lines themselves and the boundary conditions fill
// the ghost cells).
int exchange_MPI_Halos(MGArray *phi, int nArrays, ParallelTopology *topo, int xchgDir)
{
    if (!phi || nArrays < 0) return IM_FAIL(ERROR_NULL_POINTER, "null arrays");
    if (xchgDir < 1 || xchgDir > 3) return IM_FAIL(ERROR_INVALID_ARGS, "exchange direction %d", xchgDir);
    return check_topology(topo);
}

}  // extern "C"
