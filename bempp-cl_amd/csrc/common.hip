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

bool orbit_rule(RuleArgs &r) {
    if (r.n != 6) return false;
    double lam[6][3];
    for (int q = 0; q < 6; ++q) {
        lam[q][0] = 1.0 - r.x[q] - r.y[q];
        lam[q][1] = r.x[q];
        lam[q][2] = r.y[q];
    }
    auto close = [](double u, double v) { return fabs(u - v) <= 1e-12 * (fabs(u) + fabs(v) + 1e-300); };
    // special index of each point: the one barycentric coordinate that differs from the other two
    int special[6];
    double av[6], bv[6];
    for (int q = 0; q < 6; ++q) {
        special[q] = -1;
        for (int k = 0; k < 3; ++k) {
            const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
            if (close(lam[q][k1], lam[q][k2]) && !close(lam[q][k], lam[q][k1])) special[q] = k;
        }
        if (special[q] < 0) return false;
        av[q] = lam[q][special[q]];
        bv[q] = lam[q][(special[q] + 1) % 3];
    }
    // two orbits: same (a, b) and weight, special indices 0, 1, 2
    int order[6], used[6] = {0, 0, 0, 0, 0, 0}, n = 0;
    for (int o = 0; o < 2; ++o) {
        int first = -1;
        for (int q = 0; q < 6 && first < 0; ++q)
            if (!used[q]) first = q;
        if (first < 0) return false;
        for (int k = 0; k < 3; ++k) {
            int found = -1;
            for (int q = 0; q < 6; ++q)
                if (!used[q] && special[q] == k && close(av[q], av[first]) && close(bv[q], bv[first]) &&
                    close(r.w[q], r.w[first]))
                    found = q;
            if (found < 0) return false;
            used[found] = 1;
            order[n++] = found;
        }
        r.orb[2 * o] = (float)bv[first];
        r.orb[2 * o + 1] = (float)(av[first] - bv[first]);
    }
    double x[6], y[6], w[6];
    for (int q = 0; q < 6; ++q) {
        x[q] = r.x[order[q]];
        y[q] = r.y[order[q]];
        w[q] = r.w[order[q]];
    }
    for (int q = 0; q < 6; ++q) {
        r.x[q] = x[q];
        r.y[q] = y[q];
        r.w[q] = w[q];
    }
    return true;
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
