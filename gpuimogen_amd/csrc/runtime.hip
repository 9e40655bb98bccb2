// runtime.hip -- MGArray / int64 tag runtime, allocation, host<->device copies
// and the MEX-level GPU_* primitives, for one GPU per process.
//
// Semantics follow gpuclass/cudaCommon.cu (tag serdes :117-195, partition
// extent :50-91, alloc/slab/delete :256-468) and the GPU_*.cu gateways. The
// multi-GPU partitioning of the reference (nGPUs up to 4 in one process) is
// not built: arrays with nGPUs != 1 are rejected with ERROR_NOIMPLEMENT.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "imogen_internal.h"

namespace imogen {

static char g_lastError[1024] = "no error";
static DeviceState g_state[16];

int check_hip(hipError_t e, const char *what, const char *file, int line)
{
    if (e == hipSuccess) return SUCCESSFUL;
    snprintf(g_lastError, sizeof(g_lastError), "HIP error %d (%s) at %s:%d in %s", (int)e,
             hipGetErrorString(e), file, line, what);
    fprintf(stderr, "========== FAULT IN COMPILED CODE: %s\n", g_lastError);
    return ERROR_CUDA_BLEW_UP;
}

int fail(int status, const char *file, int line, const char *fmt, ...)
{
    char msg[768];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    snprintf(g_lastError, sizeof(g_lastError), "status %d at %s:%d: %s", status, file, line, msg);
    fprintf(stderr, "========== FAULT IN COMPILED CODE: %s\n", g_lastError);
    return status;
}

// Device memory every call on a GPU can use without allocating: the
// reductions' scalars, the error counters (zero at load, as a static array) and
// the fixed scratch of the allocation-free entry points. Each GPU holds its own
// copy of these arrays from the library's first call on it until the process
// ends; host code addresses them only through the per-device pointers that
// hipGetSymbolAddress returns in state().
static __device__ __align__(256) double scratch[kFixedScratchBytes / sizeof(double)];
static __device__ __align__(256) double g_devScalars[4096];
static __device__ __align__(256) unsigned long long g_errWord[ERRW_SLOTS];

DeviceState *state()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if (dev < 0 || dev >= 16) return nullptr;
    DeviceState *st = &g_state[dev];
    if (st->device < 0) {
        if (IM_HIP(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking))) return nullptr;
        if (IM_HIP(hipGetSymbolAddress((void **)&st->fixedScratch, HIP_SYMBOL(scratch)))) return nullptr;
        if (IM_HIP(hipGetSymbolAddress((void **)&st->devScalars, HIP_SYMBOL(g_devScalars)))) return nullptr;
        if (IM_HIP(hipGetSymbolAddress((void **)&st->err, HIP_SYMBOL(g_errWord)))) return nullptr;
        st->device = dev;
    }
    return st;
}

int pinned_buffers(DeviceState *st)
{
    if (!st) return ERROR_CUDA_BLEW_UP;
    if (!st->pinned && IM_HIP(hipHostMalloc((void **)&st->pinned, 64 * sizeof(double), hipHostMallocDefault)))
        return ERROR_CUDA_BLEW_UP;
    if (!st->errHost && IM_HIP(hipHostMalloc((void **)&st->errHost, ERRW_SLOTS * sizeof(unsigned long long),
                                             hipHostMallocDefault)))
        return ERROR_CUDA_BLEW_UP;
    return SUCCESSFUL;
}

int take_device_errors(DeviceState *st)
{
    if (!st) return ERROR_CUDA_BLEW_UP;
    if (pinned_buffers(st)) return ERROR_CUDA_BLEW_UP;
    const size_t bytes = ERRW_SLOTS * sizeof(unsigned long long);
    if (IM_HIP(hipMemcpyAsync(st->errHost, st->err, bytes, hipMemcpyDeviceToHost, st->stream))) return ERROR_CUDA_BLEW_UP;
    if (IM_HIP(hipStreamSynchronize(st->stream))) return ERROR_CUDA_BLEW_UP;
    const unsigned long long *e = st->errHost;
    bool any = false;
    for (int k = 0; k < ERRW_SLOTS; k++) any = any || e[k];
    if (!any) return SUCCESSFUL;
    const unsigned long long c0 = e[ERRW_DRAG_INVALID], c1 = e[ERRW_DRAG_BUDGET], c2 = e[ERRW_CFL_NAN],
                             c3 = e[ERRW_TIMESTEP];
    if (IM_HIP(hipMemsetAsync(st->err, 0, bytes, st->stream))) return ERROR_CUDA_BLEW_UP;
    if (IM_HIP(hipStreamSynchronize(st->stream))) return ERROR_CUDA_BLEW_UP;
    return IM_FAIL(ERROR_CRASH,
                   "invalid fluid state: %llu drag cells with non-positive or non-finite density, |dv| or internal "
                   "energy; %llu cells exhausted the drag integrator's step budget; %llu cells with a NaN CFL limit; "
                   "%llu time steps not taken (CFL limit not positive and finite)",
                   c0, c1, c2, c3);
}

void *scratch_block(size_t bytes)
{
    DeviceState *st = state();
    if (!st) return nullptr;
    if (st->fixedScratchOn) {
        if (bytes <= kFixedScratchBytes) return st->fixedScratch;
        IM_FAIL(ERROR_NOMEM, "%zu bytes of scratch asked for inside an allocation-free call (fixed scratch: %zu)",
                bytes, kFixedScratchBytes);
        return nullptr;
    }
    if (bytes <= st->scratchBytes) return st->scratch;
    if (st->scratch) {
        if (IM_HIP(hipStreamSynchronize(st->stream))) return nullptr;
        if (IM_HIP(hipFree(st->scratch))) return nullptr;
        st->scratch = nullptr;
        st->scratchBytes = 0;
    }
    size_t want = bytes + bytes / 4;
    if (IM_HIP(hipMalloc(&st->scratch, want))) return nullptr;
    st->scratchBytes = want;
    return st->scratch;
}

// ---- event timing -----------------------------------------------------------
struct ProfSlot {
    hipEvent_t a = nullptr, b = nullptr;
    int kernel = -1;
};
static const int kProfSlots = 16384;
static ProfSlot g_prof[kProfSlots];
static int g_profUsed = 0;
static bool g_profOn = false;
static int g_profSuspended = 0;

void prof_suspend(int on)
{
    g_profSuspended += on ? 1 : -1;
}

int prof_begin(int kernel, hipStream_t s)
{
    if (!g_profOn || g_profSuspended > 0 || g_profUsed >= kProfSlots) return -1;
    ProfSlot &p = g_prof[g_profUsed];
    if (!p.a) {
        if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess) return -1;
    }
    p.kernel = kernel;
    if (hipEventRecord(p.a, s) != hipSuccess) return -1;
    return g_profUsed++;
}

void prof_end(int slot, hipStream_t s)
{
    if (slot < 0) return;
    if (hipEventRecord(g_prof[slot].b, s) != hipSuccess) g_prof[slot].kernel = -1;
}

// Plain device copy (16 B per lane per iteration, grid-stride). Used by the
// GPU_copy / GPU_clone / GPU_makeslab primitives, and as the known-size kernel
// that calibrates the HBM byte counters.
__global__ void copy_kernel(double2 *__restrict__ dst, const double2 *__restrict__ src, int64_t n2,
                            double *__restrict__ dtail, const double *__restrict__ stail, int tail)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) dst[i] = src[i];
    if (blockIdx.x == 0 && threadIdx.x < tail) dtail[threadIdx.x] = stail[threadIdx.x];
}

int device_copy(double *dst, const double *src, int64_t ndoubles, hipStream_t s)
{
    if (ndoubles <= 0) return SUCCESSFUL;
    if (((uintptr_t)dst | (uintptr_t)src) & 15)
        return IM_HIP(hipMemcpyAsync(dst, src, (size_t)ndoubles * sizeof(double), hipMemcpyDeviceToDevice, s));
    const int64_t n2 = ndoubles / 2;
    const int tail = (int)(ndoubles - 2 * n2);
    int64_t blocks = (n2 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (double2 *)dst, (const double2 *)src,
                       n2, dst + 2 * n2, src + 2 * n2, tail);
    return IM_LAUNCHED("copy_kernel");
}

static int64_t slab_pitch_bytes(int64_t numel)
{
    int64_t b = numel * (int64_t)sizeof(double);
    return ((b + 255) / 256) * 256;
}

}  // namespace imogen

using namespace imogen;

extern "C" {

const char *imogen_last_error(void) { return g_lastError; }

const char *imogen_version(void) { return "imogen-mi355x 0.1 (gfx950)"; }

void *imogen_stream(void)
{
    DeviceState *st = state();
    return st ? (void *)st->stream : nullptr;
}

int imogen_synchronize(void)
{
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

int imogen_check_errors(void)
{
    return take_device_errors(state());
}

int imogen_device_alloc(size_t bytes, void **ptr)
{
    if (!ptr) return IM_FAIL(ERROR_NULL_POINTER, "null output pointer");
    *ptr = nullptr;
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    if (bytes == 0) return SUCCESSFUL;
    return IM_HIP(hipMalloc(ptr, bytes));
}

int imogen_device_free(void *ptr)
{
    if (!ptr) return SUCCESSFUL;
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    int rc = IM_HIP(hipStreamSynchronize(st->stream));
    if (rc) return rc;
    return IM_HIP(hipFree(ptr));
}

int imogen_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return SUCCESSFUL;
    if (!dst || !src) return IM_FAIL(ERROR_NULL_POINTER, "null pointer");
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    int rc = IM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st->stream));
    if (rc) return rc;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

int imogen_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return SUCCESSFUL;
    if (!dst || !src) return IM_FAIL(ERROR_NULL_POINTER, "null pointer");
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    int rc = IM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st->stream));
    if (rc) return rc;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

int imogen_profile_enable(int on)
{
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    int rc = IM_HIP(hipStreamSynchronize(st->stream));
    g_profUsed = 0;
    g_profOn = (on != 0);
    return rc;
}

int imogen_profile_read(int kernel, int64_t *count, double *totalMs)
{
    if (!count || !totalMs) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    int rc = IM_HIP(hipStreamSynchronize(st->stream));
    if (rc) return rc;
    int64_t n = 0;
    double tot = 0.0;
    for (int i = 0; i < g_profUsed; i++) {
        if (g_prof[i].kernel != kernel) continue;
        float ms = 0.f;
        rc = IM_HIP(hipEventElapsedTime(&ms, g_prof[i].a, g_prof[i].b));
        if (rc) return rc;
        tot += ms;
        n++;
    }
    *count = n;
    *totalMs = tot;
    return SUCCESSFUL;
}

void calcPartitionExtent(MGArray *m, int P, int *sub)
{
    int i;
    if (P >= m->nGPUs) {
        for (i = 0; i < 6; i++) sub[i] = -1;
        return;
    }
    int direct = m->partitionDir - 1;
    if (direct < 0 || direct > 2) direct = 0;
    for (i = 0; i < 3; i++) {
        if (i == direct) continue;
        sub[i] = 0;
        sub[3 + i] = m->dim[i];
    }
    int alpha = m->dim[direct] / m->nGPUs;
    sub[direct] = P * alpha;
    sub[3 + direct] = alpha;
    if (P == m->nGPUs - 1) sub[3 + direct] = m->dim[direct] - P * alpha;
    if (m->nGPUs > 1) {
        if (m->addExteriorHalo != 0 || P > 0) {
            sub[direct] -= m->haloSize;
            sub[3 + direct] += m->haloSize;
        }
        if (m->addExteriorHalo != 0 || P < m->nGPUs - 1) sub[3 + direct] += m->haloSize;
    }
}

void MGA_permtagToNums(int permtag, int *p)
{
    if (!p) return;
    switch (permtag) {
    case 1: p[0] = 1; p[1] = 2; p[2] = 3; break;
    case 2: p[0] = 1; p[1] = 3; p[2] = 2; break;
    case 3: p[0] = 2; p[1] = 1; p[2] = 3; break;
    case 4: p[0] = 2; p[1] = 3; p[2] = 1; break;
    case 5: p[0] = 3; p[1] = 1; p[2] = 2; break;
    case 6: p[0] = 3; p[1] = 2; p[2] = 1; break;
    default: p[0] = 1; p[1] = 2; p[2] = 3; break;
    }
}

int MGA_numsToPermtag(int *n)
{
    if (!n) return -1;
    switch (n[0]) {
    case 1:
        if (n[1] == 2 && n[2] == 3) return 1;
        if (n[1] == 3 && n[2] == 2) return 2;
        break;
    case 2:
        if (n[1] == 1 && n[2] == 3) return 3;
        if (n[1] == 3 && n[2] == 1) return 4;
        break;
    case 3:
        if (n[1] == 1 && n[2] == 2) return 5;
        if (n[1] == 2 && n[2] == 1) return 6;
        break;
    }
    return 0;
}

int MGA_dir2memdir(int *perm, int dir)
{
    if (!perm) return -1;
    if (perm[0] == dir) return 1;
    if (perm[1] == dir) return 2;
    if (perm[2] == dir) return 3;
    return -1;
}

int deserializeTagToMGArray(int64_t *tag, MGArray *mg)
{
    if (!tag || !mg) return IM_FAIL(ERROR_NULL_POINTER, "null tag or MGArray");
    memset(mg, 0, sizeof(*mg));
    mg->dim[0] = (int)tag[GPU_TAG_DIM0];
    mg->dim[1] = (int)tag[GPU_TAG_DIM1];
    mg->dim[2] = (int)tag[GPU_TAG_DIM2];
    mg->numel = (int64_t)mg->dim[0] * mg->dim[1] * mg->dim[2];
    mg->numSlabs = (int)tag[GPU_TAG_DIMSLAB];
    mg->haloSize = (int)tag[GPU_TAG_HALO];
    mg->partitionDir = (int)tag[GPU_TAG_PARTDIR];
    mg->nGPUs = (int)tag[GPU_TAG_NGPUS];
    mg->addExteriorHalo = (int)tag[GPU_TAG_EXTERIORHALO];
    mg->permtag = (int)tag[GPU_TAG_DIMPERMUTATION];
    MGA_permtagToNums(mg->permtag, mg->currentPermutation);
    mg->vectorComponent = (int)tag[GPU_TAG_VECTOR_COMPONENT];
    mg->mpiCircularBoundaryBits = (int)tag[GPU_TAG_CIRCULARBITS];
    if (mg->dim[0] < 1 || mg->dim[1] < 1 || mg->dim[2] < 1 || mg->nGPUs < 1 || mg->nGPUs > MAX_GPUS_USED)
        return IM_FAIL(ERROR_DESERIALIZE_GPUTAG_FAILED, "malformed tag (dims %d %d %d, nGPUs %d)", mg->dim[0],
                       mg->dim[1], mg->dim[2], mg->nGPUs);
    int sub[6];
    const int64_t *t = tag + GPU_TAG_LENGTH;
    for (int i = 0; i < mg->nGPUs; i++) {
        mg->deviceID[i] = (int)t[2 * i];
        mg->devicePtr[i] = (double *)(intptr_t)t[2 * i + 1];
        calcPartitionExtent(mg, i, sub);
        mg->partNumel[i] = sub[3] * sub[4] * sub[5];
        mg->slabPitch[i] = slab_pitch_bytes(mg->partNumel[i]);
    }
    for (int i = mg->nGPUs; i < MAX_GPUS_USED; i++) {
        mg->deviceID[i] = -1;
        mg->devicePtr[i] = nullptr;
    }
    return SUCCESSFUL;
}

void serializeMGArrayToTag(MGArray *mg, int64_t *tag)
{
    tag[GPU_TAG_DIM0] = mg->dim[0];
    tag[GPU_TAG_DIM1] = mg->dim[1];
    tag[GPU_TAG_DIM2] = mg->dim[2];
    tag[GPU_TAG_DIMSLAB] = mg->numSlabs;
    tag[GPU_TAG_HALO] = mg->haloSize;
    tag[GPU_TAG_PARTDIR] = mg->partitionDir;
    tag[GPU_TAG_NGPUS] = mg->nGPUs;
    tag[GPU_TAG_EXTERIORHALO] = mg->addExteriorHalo;
    tag[GPU_TAG_DIMPERMUTATION] = mg->permtag;
    tag[GPU_TAG_CIRCULARBITS] = mg->mpiCircularBoundaryBits;
    tag[GPU_TAG_VECTOR_COMPONENT] = mg->vectorComponent;
    for (int i = 0; i < mg->nGPUs; i++) {
        tag[GPU_TAG_LENGTH + 2 * i] = mg->deviceID[i];
        tag[GPU_TAG_LENGTH + 2 * i + 1] = (int64_t)(intptr_t)mg->devicePtr[i];
    }
}

static int check_single_gpu(const MGArray *m)
{
    if (m->nGPUs != 1)
        return IM_FAIL(ERROR_NOIMPLEMENT, "only single-GPU arrays are supported (nGPUs = %d)", m->nGPUs);
    return SUCCESSFUL;
}

int MGA_allocSlab(MGArray *skeleton, MGArray *nu, int Nslabs)
{
    if (!skeleton || !nu) return IM_FAIL(ERROR_NULL_POINTER, "null MGArray");
    int rc = check_single_gpu(skeleton);
    if (rc) return rc;
    if (Nslabs < 1) return IM_FAIL(ERROR_INVALID_ARGS, "Nslabs = %d", Nslabs);
    *nu = *skeleton;
    for (int i = 0; i < MAX_GPUS_USED; i++) nu->boundaryConditions.staticCells[i] = nullptr;
    int sub[6];
    calcPartitionExtent(skeleton, 0, sub);
    int64_t numel = (int64_t)sub[3] * sub[4] * sub[5];
    nu->partNumel[0] = (int)numel;
    nu->slabPitch[0] = slab_pitch_bytes(numel);
    if (state() == nullptr) return ERROR_CUDA_BLEW_UP;
    rc = IM_HIP(hipMalloc((void **)&nu->devicePtr[0], (size_t)nu->slabPitch[0] * Nslabs));
    if (rc) {
        nu->devicePtr[0] = nullptr;
        return ERROR_NOMEM;
    }
    nu->numSlabs = Nslabs;
    return SUCCESSFUL;
}

int MGA_allocArrays(MGArray **ret, int N, MGArray *skeleton)
{
    if (!ret || !skeleton || N < 1) return IM_FAIL(ERROR_NULL_POINTER, "bad arguments");
    MGArray *m = *ret;
    if (!m) {
        m = (MGArray *)calloc((size_t)N, sizeof(MGArray));
        if (!m) return ERROR_NOMEM;
        *ret = m;
    }
    for (int i = 0; i < N; i++) {
        int rc = MGA_allocSlab(skeleton, &m[i], 1);
        if (rc) return rc;
    }
    return SUCCESSFUL;
}

int MGA_delete(MGArray *victim)
{
    if (!victim) return IM_FAIL(ERROR_NULL_POINTER, "MGA_delete passed a null MGArray");
    if (victim->numSlabs < 1) return SUCCESSFUL;  // slab views are never freed (cudaCommon.cu:427)
    int rc = SUCCESSFUL;
    for (int j = 0; j < victim->nGPUs; j++) {
        if (victim->devicePtr[j]) {
            DeviceState *st = state();
            if (st) (void)hipStreamSynchronize(st->stream);
            rc = IM_HIP(hipFree(victim->devicePtr[j]));
            victim->devicePtr[j] = nullptr;
        }
    }
    return rc;
}

int MGA_uploadArrayToGPU(double *p, MGArray *g, int partitionTo)
{
    (void)partitionTo;
    if (!p || !g) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    int rc = check_single_gpu(g);
    if (rc) return rc;
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    rc = IM_HIP(hipMemcpyAsync(g->devicePtr[0], p, (size_t)g->partNumel[0] * sizeof(double),
                               hipMemcpyHostToDevice, st->stream));
    if (rc) return rc;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

int MGA_downloadArrayToCPU(MGArray *g, double **p, int partitionFrom)
{
    (void)partitionFrom;
    if (!p || !g) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    int rc = check_single_gpu(g);
    if (rc) return rc;
    if (*p == nullptr) {
        *p = (double *)malloc((size_t)g->numel * sizeof(double));
        if (!*p) return ERROR_NOMEM;
    }
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    rc = IM_HIP(hipMemcpyAsync(*p, g->devicePtr[0], (size_t)g->partNumel[0] * sizeof(double),
                               hipMemcpyDeviceToHost, st->stream));
    if (rc) return rc;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

/* ---- MEX-level primitives ------------------------------------------------ */

static void init_array(MGArray *m, const int *dims3, int halo, int partDir)
{
    memset(m, 0, sizeof(*m));
    m->dim[0] = dims3[0];
    m->dim[1] = dims3[1];
    m->dim[2] = dims3[2];
    m->numel = (int64_t)dims3[0] * dims3[1] * dims3[2];
    m->numSlabs = 1;
    m->haloSize = halo;
    m->partitionDir = (partDir >= 1 && partDir <= 3) ? partDir : PARTITION_X;
    m->nGPUs = 1;
    m->permtag = 1;
    m->currentPermutation[0] = 1;
    m->currentPermutation[1] = 2;
    m->currentPermutation[2] = 3;
    int dev = 0;
    (void)hipGetDevice(&dev);
    m->deviceID[0] = dev;
    for (int i = 1; i < MAX_GPUS_USED; i++) m->deviceID[i] = -1;
    m->partNumel[0] = (int)m->numel;
    m->slabPitch[0] = slab_pitch_bytes(m->numel);
}

int GPU_upload(const double *host, const int *dims3, int haloSize, int partitionDir, int64_t *tagOut)
{
    if (!dims3 || !tagOut) return IM_FAIL(ERROR_NULL_POINTER, "null argument");
    if (dims3[0] < 1 || dims3[1] < 1 || dims3[2] < 1) return IM_FAIL(ERROR_INVALID_ARGS, "bad dims");
    MGArray skel, m;
    init_array(&skel, dims3, haloSize, partitionDir);
    int rc = MGA_allocSlab(&skel, &m, 1);
    if (rc) return rc;
    if (host) {
        rc = MGA_uploadArrayToGPU((double *)host, &m, -1);
        if (rc) {
            MGA_delete(&m);
            return rc;
        }
    }
    serializeMGArrayToTag(&m, tagOut);
    return SUCCESSFUL;
}

int GPU_download(const int64_t *tag, double *host)
{
    MGArray m;
    int rc = deserializeTagToMGArray((int64_t *)tag, &m);
    if (rc) return rc;
    return MGA_downloadArrayToCPU(&m, &host, -1);
}

int GPU_free(const int64_t *tag)
{
    MGArray m;
    int rc = deserializeTagToMGArray((int64_t *)tag, &m);
    if (rc) return rc;
    return MGA_delete(&m);
}

int GPU_makeslab(const int64_t *tag, int N, int64_t *newTag)
{
    MGArray m, nu;
    int rc = deserializeTagToMGArray((int64_t *)tag, &m);
    if (rc) return rc;
    rc = check_single_gpu(&m);
    if (rc) return rc;
    rc = MGA_allocSlab(&m, &nu, N);
    if (rc) return rc;
    DeviceState *st = state();
    rc = device_copy(nu.devicePtr[0], m.devicePtr[0], m.partNumel[0], st->stream);
    if (rc) return rc;
    if (m.numSlabs > 0) {
        rc = MGA_delete(&m);
        if (rc) return rc;
    }
    serializeMGArrayToTag(&nu, newTag);
    return SUCCESSFUL;
}

int GPU_getslab(const int64_t *tag, int k, int64_t *viewTag)
{
    MGArray m;
    int rc = deserializeTagToMGArray((int64_t *)tag, &m);
    if (rc) return rc;
    if (k < 0 || (m.numSlabs > 0 && k >= m.numSlabs))
        return IM_FAIL(ERROR_INVALID_ARGS, "slab %d out of range (array has %d)", k, m.numSlabs);
    MGArray v = m;
    v.devicePtr[0] = m.devicePtr[0] + k * (m.slabPitch[0] / (int64_t)sizeof(double));
    v.numSlabs = -k;
    serializeMGArrayToTag(&v, viewTag);
    return SUCCESSFUL;
}

int GPU_setslab(const int64_t *tag, int k, const double *host, int64_t *viewTag)
{
    int rc = GPU_getslab(tag, k, viewTag);
    if (rc) return rc;
    if (!host) return SUCCESSFUL;
    MGArray v;
    rc = deserializeTagToMGArray(viewTag, &v);
    if (rc) return rc;
    return MGA_uploadArrayToGPU((double *)host, &v, -1);
}

int GPU_clone(const int64_t *tag, int64_t *newTag)
{
    MGArray m, nu;
    int rc = deserializeTagToMGArray((int64_t *)tag, &m);
    if (rc) return rc;
    rc = MGA_allocSlab(&m, &nu, 1);
    if (rc) return rc;
    DeviceState *st = state();
    rc = device_copy(nu.devicePtr[0], m.devicePtr[0], m.partNumel[0], st->stream);
    if (rc) return rc;
    rc = IM_HIP(hipStreamSynchronize(st->stream));
    if (rc) return rc;
    serializeMGArrayToTag(&nu, newTag);
    return SUCCESSFUL;
}

int GPU_copy(const int64_t *dstTag, const int64_t *srcTag)
{
    MGArray d, s;
    int rc = deserializeTagToMGArray((int64_t *)dstTag, &d);
    if (rc) return rc;
    rc = deserializeTagToMGArray((int64_t *)srcTag, &s);
    if (rc) return rc;
    if (d.numel != s.numel) return IM_FAIL(ERROR_INVALID_ARGS, "GPU_copy: shapes differ");
    DeviceState *st = state();
    rc = device_copy(d.devicePtr[0], s.devicePtr[0], s.partNumel[0], st->stream);
    if (rc) return rc;
    return IM_HIP(hipStreamSynchronize(st->stream));
}

int GPU_ctrl_reset(void)
{
    DeviceState *st = state();
    if (!st) return ERROR_CUDA_BLEW_UP;
    int rc = IM_HIP(hipStreamSynchronize(st->stream));
    if (rc) return rc;
    if (st->scratch) {
        rc = IM_HIP(hipFree(st->scratch));
        st->scratch = nullptr;
        st->scratchBytes = 0;
    }
    return rc;
}

int GPU_ctrl_info(int device, char *text, int len)
{
    if (!text || len < 1) return IM_FAIL(ERROR_NULL_POINTER, "no text buffer");
    hipDeviceProp_t p;
    int rc = IM_HIP(hipGetDeviceProperties(&p, device));
    if (rc) return rc;
    snprintf(text, (size_t)len,
             "device %d: %s (%s), %d compute units, %zu MiB global memory, %zu KiB LDS per block, "
             "wavefront %d, %d MHz",
             device, p.name, p.gcnArchName, p.multiProcessorCount, (size_t)(p.totalGlobalMem >> 20),
             (size_t)(p.sharedMemPerBlock >> 10), p.warpSize, p.clockRate / 1000);
    return SUCCESSFUL;
}

int GPU_ctrl_memory(int device, size_t *freeBytes, size_t *totalBytes)
{
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (device != cur) {
        int rc = IM_HIP(hipSetDevice(device));
        if (rc) return rc;
    }
    int rc = IM_HIP(hipMemGetInfo(freeBytes, totalBytes));
    if (device != cur) (void)hipSetDevice(cur);
    return rc;
}

}  // extern "C"
