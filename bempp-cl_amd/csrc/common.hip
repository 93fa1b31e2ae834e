// Error reporting, argument checks and the zero-fill kernel of libbempp_hip.so.
#include "bem_device.h"
#include "bem_internal.h"

namespace bem {

static thread_local char g_error[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

void clear_error() { g_error[0] = '\0'; }

int check_space(const bem_space_t *s, const char *which) {
    if (s->n_elements < 0) {
        set_error("%s space: negative element count", which);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (s->shapeset < BEM_SHAPE_P0 || s->shapeset > BEM_SHAPE_RWG) {
        set_error("%s space: unknown shapeset %d", which, s->shapeset);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (s->n_elements > 0 && (!s->elements || !s->geometry || !s->vertex_ids || !s->local2global || !s->multipliers)) {
        set_error("%s space: null device array", which);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    return BEM_OK;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s launch failed: %s", what, hipGetErrorString(e));
        return BEM_ERR_DEVICE;
    }
    return BEM_OK;
}

__global__ void fill_zero_kernel(uint4 *p, size_t n16, unsigned char *tail, size_t ntail) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

}  // namespace bem

using namespace bem;

extern "C" int bem_fill_zero(void *ptr, size_t bytes, void *stream) {
    clear_error();
    if (bytes == 0) return BEM_OK;
    if (!ptr || ((uintptr_t)ptr & 15)) {
        set_error("bem_fill_zero needs a non-null 16-byte aligned pointer");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    size_t n16 = bytes / 16, ntail = bytes % 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (uint4 *)ptr, n16,
                       (unsigned char *)ptr + 16 * n16, ntail);
    return check_launch("fill_zero_kernel");
}

extern "C" const char *bem_last_error(void) { return g_error; }

extern "C" const char *bem_version(void) { return "bempp-hip 0.1 (gfx950)"; }
