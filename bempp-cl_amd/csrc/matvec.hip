// Dense matrix-vector products with an assembled operator held in HBM, for gfx950.
//
// y = op(A) x with op = identity, transpose or conjugate transpose; A is row-major
// (rows x cols, leading dimension ld) of float32/float64/complex64/complex128; x and y are
// complex128; sums are accumulated in FP64. This is the product the reference's
// DenseDiscreteBoundaryOperator performs with numpy (bempp_cl/api/assembly/
// discrete_boundary_operator.py:85-110, `self.A @ x`) and, for the adjoint, `A^H x`.
//
// The products stream the matrix once (HBM bound):
//  * op = N, 16-byte aligned rows: workgroups of 32 rows x 1024 columns; the x chunk is staged in
//    LDS once per workgroup, each wavefront streams 8 rows with 16-byte loads (8 independent loads
//    in flight per lane), and the column chunks' FP64 sums are combined with FP64 atomics.
//  * op = N otherwise: one wavefront per row; lanes read consecutive columns, x from L2.
//  * op = T/C: each workgroup owns 256 consecutive columns and a band of rows; every lane
//    walks its column down the band, and the band sums are added to y with FP64 atomics.
#include "bem_device.h"
#include "bem_internal.h"

namespace bem {
namespace mv {

template <typename AT>
struct Elem;
template <>
struct Elem<float> {
    static __device__ __forceinline__ void get(const void *A, int64_t i, double &re, double &im) {
        re = reinterpret_cast<const float *>(A)[i];
        im = 0.0;
    }
};
template <>
struct Elem<double> {
    static __device__ __forceinline__ void get(const void *A, int64_t i, double &re, double &im) {
        re = reinterpret_cast<const double *>(A)[i];
        im = 0.0;
    }
};
template <>
struct Elem<float2> {
    static __device__ __forceinline__ void get(const void *A, int64_t i, double &re, double &im) {
        const float2 v = reinterpret_cast<const float2 *>(A)[i];
        re = v.x;
        im = v.y;
    }
};
template <>
struct Elem<double2> {
    static __device__ __forceinline__ void get(const void *A, int64_t i, double &re, double &im) {
        const double2 v = reinterpret_cast<const double2 *>(A)[i];
        re = v.x;
        im = v.y;
    }
};

template <typename T>
__device__ __forceinline__ T wsum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// 16-byte loads of the matrix: E elements per load
template <typename AT>
struct Pack16;
template <>
struct Pack16<float2> {
    static constexpr int E = 2;
    using V = float4;
    static __device__ __forceinline__ void get(const V &v, int e, double &re, double &im) {
        re = e ? v.z : v.x;
        im = e ? v.w : v.y;
    }
};
template <>
struct Pack16<double2> {
    static constexpr int E = 1;
    using V = double2;
    static __device__ __forceinline__ void get(const V &v, int, double &re, double &im) {
        re = v.x;
        im = v.y;
    }
};
template <>
struct Pack16<float> {
    static constexpr int E = 4;
    using V = float4;
    static __device__ __forceinline__ void get(const V &v, int e, double &re, double &im) {
        re = e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
        im = 0.0;
    }
};
template <>
struct Pack16<double> {
    static constexpr int E = 2;
    using V = double2;
    static __device__ __forceinline__ void get(const V &v, int e, double &re, double &im) {
        re = e ? v.y : v.x;
        im = 0.0;
    }
};

constexpr int kMvRows = 32;          // rows per workgroup
constexpr int kMvRowsPerWave = 8;    // rows per wavefront
constexpr int kMvCols = 1024;        // columns per workgroup (x chunk: 16 KB of LDS)

// y_i += sum_{j in the chunk} A_ij x_j (y zeroed beforehand); rows start 16-byte aligned
template <typename AT>
__global__ __launch_bounds__(256, sizeof(AT) == 4 ? 3 : 4) void matvec_n_tiled_kernel(const void *A, int64_t rows, int64_t cols, int64_t ld,
                                                             const double *x, double *y) {
    using P = Pack16<AT>;
    constexpr int E = P::E;
    __shared__ double2 xs[kMvCols];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t c0 = (int64_t)blockIdx.y * kMvCols;
    const int nc = (int)min((int64_t)kMvCols, cols - c0);
    for (int j = tid; j < kMvCols; j += 256)
        xs[j] = j < nc ? make_double2(x[2 * (c0 + j)], x[2 * (c0 + j) + 1]) : make_double2(0.0, 0.0);
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * kMvRows + wave * kMvRowsPerWave;
    const int nr = (int)max((int64_t)0, min((int64_t)kMvRowsPerWave, rows - r0));  // uniform per wavefront
    double sr[kMvRowsPerWave], si[kMvRowsPerWave];
#pragma unroll
    for (int r = 0; r < kMvRowsPerWave; ++r) sr[r] = si[r] = 0.0;
    if (nr == 0) return;  // uniform per wavefront, after the only barrier
    // rows past the end read the last row again (in bounds); their sums are never stored, so
    // every step issues its kMvRowsPerWave loads back to back with no branches
    const AT *rowp[kMvRowsPerWave];
#pragma unroll
    for (int r = 0; r < kMvRowsPerWave; ++r)
        rowp[r] = reinterpret_cast<const AT *>(A) + (r0 + min(r, nr - 1)) * ld + c0;
    const int nfull = nc - nc % E;  // columns covered by whole 16-byte loads
    for (int j = lane * E; j < nfull; j += 64 * E) {
        double2 xv[E];
#pragma unroll
        for (int e = 0; e < E; ++e) xv[e] = xs[j + e];
        typename P::V v[kMvRowsPerWave];
#pragma unroll
        for (int r = 0; r < kMvRowsPerWave; ++r) v[r] = *reinterpret_cast<const typename P::V *>(rowp[r] + j);
#pragma unroll
        for (int r = 0; r < kMvRowsPerWave; ++r) {
            double tr = 0.0, ti = 0.0;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                double ar, ai;
                P::get(v[r], e, ar, ai);
                tr += ar * xv[e].x - ai * xv[e].y;
                ti += ar * xv[e].y + ai * xv[e].x;
            }
            sr[r] += tr;
            si[r] += ti;
        }
    }
    // the chunk's last nc % E columns (the row's last columns): element loads
    for (int j = nfull + lane; j < nc; j += 64) {
        const double2 xv = xs[j];
#pragma unroll
        for (int r = 0; r < kMvRowsPerWave; ++r) {
            double ar, ai;
            Elem<AT>::get(rowp[r], j, ar, ai);
            sr[r] += ar * xv.x - ai * xv.y;
            si[r] += ar * xv.y + ai * xv.x;
        }
    }
#pragma unroll
    for (int r = 0; r < kMvRowsPerWave; ++r) {
        if (r >= nr) break;
        const double a = wsum(sr[r]), b = wsum(si[r]);
        if (lane == 0) {
            atomic_add(y + 2 * (r0 + r), a);
            atomic_add(y + 2 * (r0 + r) + 1, b);
        }
    }
}

// y_i = sum_j A_ij x_j, one wavefront per row
template <typename AT>
__global__ __launch_bounds__(256) void matvec_n_kernel(const void *A, int64_t rows, int64_t cols, int64_t ld,
                                                       const double *x, double *y) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= rows) return;  // uniform per wavefront
    double sr = 0.0, si = 0.0;
    const int64_t base = i * ld;
#pragma unroll 4
    for (int64_t j = lane; j < cols; j += 64) {
        double ar, ai;
        Elem<AT>::get(A, base + j, ar, ai);
        const double xr = x[2 * j], xi = x[2 * j + 1];
        sr += ar * xr - ai * xi;
        si += ar * xi + ai * xr;
    }
    sr = wsum(sr);
    si = wsum(si);
    if (lane == 0) {
        y[2 * i] = sr;
        y[2 * i + 1] = si;
    }
}

// y_j += sum_{i in band} op(A_ij) x_i, one lane per column
template <typename AT, bool CONJ>
__global__ __launch_bounds__(256) void matvec_t_kernel(const void *A, int64_t rows, int64_t cols, int64_t ld,
                                                       int64_t band, const double *x, double *y) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.y * band, i1 = min(i0 + band, rows);
    if (j >= cols) return;
    double sr = 0.0, si = 0.0;
#pragma unroll 4
    for (int64_t i = i0; i < i1; ++i) {
        double ar, ai;
        Elem<AT>::get(A, i * ld + j, ar, ai);
        if (CONJ) ai = -ai;
        const double xr = x[2 * i], xi = x[2 * i + 1];
        sr += ar * xr - ai * xi;
        si += ar * xi + ai * xr;
    }
    atomic_add(y + 2 * j, sr);
    atomic_add(y + 2 * j + 1, si);
}

__global__ void zero_kernel(double *y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = 0.0;
}

template <typename AT>
int launch(int trans, const void *A, int64_t rows, int64_t cols, int64_t ld, const double *x, double *y,
           hipStream_t s) {
    const bool aligned = reinterpret_cast<uintptr_t>(A) % 16 == 0 && (ld * (int64_t)sizeof(AT)) % 16 == 0;
    const int64_t nchunks = (cols + kMvCols - 1) / kMvCols;
    if (trans == 0 && aligned && nchunks <= 65535) {
        hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((2 * rows + 255) / 256)), dim3(256), 0, s, y, 2 * rows);
        const dim3 grid((unsigned)((rows + kMvRows - 1) / kMvRows), (unsigned)nchunks);
        hipLaunchKernelGGL(matvec_n_tiled_kernel<AT>, grid, dim3(256), 0, s, A, rows, cols, ld, x, y);
        return check_launch("matvec_n_tiled_kernel");
    }
    if (trans == 0) {
        hipLaunchKernelGGL(matvec_n_kernel<AT>, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s, A, rows, cols, ld, x,
                           y);
        return check_launch("matvec_n_kernel");
    }
    const int64_t n = 2 * cols;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y, n);
    // row bands sized for >= ~4096 workgroups
    const int64_t ncb = (cols + 255) / 256;
    int64_t nbands = (4096 + ncb - 1) / ncb;
    if (nbands > rows) nbands = rows;
    if (nbands > 65535) nbands = 65535;
    if (nbands < 1) nbands = 1;
    const int64_t band = (rows + nbands - 1) / nbands;
    const dim3 grid((unsigned)ncb, (unsigned)((rows + band - 1) / band));
    if (trans == 1) hipLaunchKernelGGL((matvec_t_kernel<AT, false>), grid, dim3(256), 0, s, A, rows, cols, ld, band, x, y);
    else hipLaunchKernelGGL((matvec_t_kernel<AT, true>), grid, dim3(256), 0, s, A, rows, cols, ld, band, x, y);
    return check_launch("matvec_t_kernel");
}

// ---------------------------------------------------------------- square-matrix helpers
// A <- A + A^T (n x n, leading dimension ld), in place: the workgroup of tile pair (bi, bj),
// bi <= bj, reads both 32 x 32 tiles into LDS before writing either.
template <typename E>
__device__ __forceinline__ E eadd(E a, E b) {
    return a + b;
}
template <>
__device__ __forceinline__ float2 eadd(float2 a, float2 b) {
    return make_float2(a.x + b.x, a.y + b.y);
}
template <>
__device__ __forceinline__ double2 eadd(double2 a, double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
}

template <typename E>
__global__ __launch_bounds__(256) void add_transpose_kernel(E *A, int64_t n, int64_t ld) {
    const int64_t bi = blockIdx.y, bj = blockIdx.x;
    if (bi > bj) return;  // uniform per workgroup, before any barrier
    __shared__ E X[32][33], Y[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t r0 = bi * 32, c0 = bj * 32;
    const bool off = bi != bj;
    for (int r = ty; r < 32; r += 8) {
        if (r0 + r < n && c0 + tx < n) X[r][tx] = A[(r0 + r) * ld + c0 + tx];
        if (off && c0 + r < n && r0 + tx < n) Y[r][tx] = A[(c0 + r) * ld + r0 + tx];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        // A[r0 + r][c0 + tx] += A[c0 + tx][r0 + r]
        if (r0 + r < n && c0 + tx < n) A[(r0 + r) * ld + c0 + tx] = eadd(X[r][tx], off ? Y[tx][r] : X[tx][r]);
        if (off && c0 + r < n && r0 + tx < n) A[(c0 + r) * ld + r0 + tx] = eadd(Y[r][tx], X[tx][r]);
    }
}

// Lower triangle from the upper one, zero diagonal: A[i][j] = A[j][i] for i > j, A[i][i] = 0. The
// workgroup of destination tile (bi, bj), bi >= bj, reads source tile (bj, bi) into LDS first.
template <typename E>
__global__ __launch_bounds__(256) void mirror_upper_kernel(E *A, int64_t n, int64_t ld) {
    const int64_t bi = blockIdx.y, bj = blockIdx.x;
    if (bi < bj) return;  // uniform per workgroup, before the barrier
    __shared__ E X[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t r0 = bi * 32, c0 = bj * 32;
    for (int r = ty; r < 32; r += 8)
        if (c0 + r < n && r0 + tx < n) X[r][tx] = A[(c0 + r) * ld + r0 + tx];
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = r0 + r, j = c0 + tx;  // A[i][j] = A[j][i] = X[tx][r]
        if (i < n && j < n) {
            if (i > j) A[i * ld + j] = X[tx][r];
            else if (i == j) A[i * ld + j] = E{};
        }
    }
}

template <typename E>
__global__ __launch_bounds__(256) void zero_square_kernel(E *A, int64_t n, int64_t ld) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    for (int64_t r = blockIdx.y; r < n; r += gridDim.y) A[r * ld + c] = E{};
}

template <typename E>
static int square_launch(int what, void *A, int64_t n, int64_t ld, hipStream_t s) {
    if (what == 0) {
        // contiguous rows: one 16-byte store per lane over the whole matrix
        if (ld == n && reinterpret_cast<uintptr_t>(A) % 16 == 0) return bem_fill_zero(A, (size_t)n * n * sizeof(E), s);
        const unsigned rows = (unsigned)(n < 4096 ? n : 4096);
        hipLaunchKernelGGL((zero_square_kernel<E>), dim3((unsigned)((n + 255) / 256), rows), dim3(256), 0, s, (E *)A, n,
                           ld);
        return check_launch("zero_square_kernel");
    }
    const unsigned nb = (unsigned)((n + 31) / 32);
    if (what == 2) {
        hipLaunchKernelGGL((mirror_upper_kernel<E>), dim3(nb, nb), dim3(256), 0, s, (E *)A, n, ld);
        return check_launch("mirror_upper_kernel");
    }
    hipLaunchKernelGGL((add_transpose_kernel<E>), dim3(nb, nb), dim3(256), 0, s, (E *)A, n, ld);
    return check_launch("add_transpose_kernel");
}

}  // namespace mv

int square_matrix_op(int what, void *A, int dtype, int64_t n, int64_t ld, hipStream_t s) {
    if (!A || n < 0 || ld < n || n > 65535 * 32) {
        set_error("invalid square matrix (n = %lld, ld = %lld)", (long long)n, (long long)ld);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (n == 0) return BEM_OK;
    switch (dtype) {
    case BEM_F32: return mv::square_launch<float>(what, A, n, ld, s);
    case BEM_F64: return mv::square_launch<double>(what, A, n, ld, s);
    case BEM_C64: return mv::square_launch<float2>(what, A, n, ld, s);
    case BEM_C128: return mv::square_launch<double2>(what, A, n, ld, s);
    default: set_error("unknown matrix dtype %d", dtype); return BEM_ERR_INVALID_ARGUMENT;
    }
}

// ------------------------------------------------------------------ GMRES orthogonalisation
// One classical Gram-Schmidt step with reorthogonalisation (CGS2) of w against the k rows of V
// (complex128, row stride ldv), three passes that each read V once:
//   1: c1 = conj(V) w and |w|^2;
//   2: w <- w - V^T c1 and, from the updated elements (the update is elementwise), c2 = conj(V) w;
//   3: w <- w - V^T c2 and |w|^2.
// A thread owns elements j of w (grid-stride). For the k dot products each wave sums its lanes
// with shuffles into an LDS row per dot product; after one barrier per 256 dot products the
// block adds the waves' sums to out with one FP64 atomic per value.
// out (complex128): c1[0, k), |w_in|^2 at k, c2[k + 1, 2k + 1), |w_out|^2 at 2k + 1.
namespace gs {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kDotChunk = 256;  // dot products staged in LDS per barrier

// (re, im) summed over the wavefront (every lane gets the total)
__device__ __forceinline__ void wave_sum2(double &re, double &im) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
    }
}

// (re, im) summed over the block; every thread gets the total
__device__ __forceinline__ void block_sum2(double &re, double &im, double (*red)[2]) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = re;
        red[wave][1] = im;
    }
    __syncthreads();
    re = red[0][0] + red[1][0] + red[2][0] + red[3][0];
    im = red[0][1] + red[1][1] + red[2][1] + red[3][1];
    __syncthreads();  // red is reused by the next sum
}

// PASS 1: dots of w; PASS 2: update by c1 (out[0, k)), then dots; PASS 3: update by c2, norm
template <int PASS>
__global__ __launch_bounds__(kThreads) void cgs2_kernel(const double2 *V, int64_t ldv, int k, int64_t n, double2 *w,
                                                        double *out) {
    __shared__ double red[kThreads / 64][2];
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    const int64_t j0 = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const double2 *cin = reinterpret_cast<const double2 *>(out) + (PASS == 2 ? 0 : k + 1);
    // update: w_j <- w_j - sum_i V_ij c_i (the thread's elements, kept in place)
    if (PASS >= 2) {
        for (int64_t j = j0; j < n; j += stride) {
            double2 v = w[j];
            for (int i = 0; i < k; ++i) {
                const double2 a = V[(int64_t)i * ldv + j], c = cin[i];
                v.x -= a.x * c.x - a.y * c.y;
                v.y -= a.x * c.y + a.y * c.x;
            }
            w[j] = v;
        }
    }
    if (PASS <= 2) {
        // dots conj(V_i) . w into out[0, k) (pass 1) or out[k + 1, 2k + 1) (pass 2)
        __shared__ double part[kWaves][kDotChunk][2];
        double *dst = out + 2 * (PASS == 1 ? 0 : k + 1);
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int i0 = 0; i0 < k; i0 += kDotChunk) {
            const int m = min(kDotChunk, k - i0);
            for (int i = 0; i < m; ++i) {
                double re = 0.0, im = 0.0;
                for (int64_t j = j0; j < n; j += stride) {
                    const double2 a = V[(int64_t)(i0 + i) * ldv + j], v = w[j];
                    re += a.x * v.x + a.y * v.y;
                    im += a.x * v.y - a.y * v.x;
                }
                wave_sum2(re, im);
                if (lane == 0) {
                    part[wave][i][0] = re;
                    part[wave][i][1] = im;
                }
            }
            __syncthreads();
            for (int i = threadIdx.x; i < m; i += kThreads) {
                double re = 0.0, im = 0.0;
#pragma unroll
                for (int wv = 0; wv < kWaves; ++wv) {
                    re += part[wv][i][0];
                    im += part[wv][i][1];
                }
                atomic_add(dst + 2 * (i0 + i), re);
                atomic_add(dst + 2 * (i0 + i) + 1, im);
            }
            __syncthreads();  // part is reused by the next chunk
        }
    }
    if (PASS != 2) {
        // |w|^2: the input (pass 1, at k) or the result (pass 3, at 2k + 1)
        double re = 0.0, im = 0.0;
        for (int64_t j = j0; j < n; j += stride) {
            const double2 v = w[j];
            re += v.x * v.x + v.y * v.y;
        }
        block_sum2(re, im, red);
        if (threadIdx.x == 0) atomic_add(out + 2 * (PASS == 1 ? k : 2 * k + 1), re);
    }
}

// next basis vector: v = w / |w| (or w itself when |w| <= eps |w_in|, the breakdown test of GMRES)
__global__ __launch_bounds__(kThreads) void cgs2_scale_kernel(const double2 *w, int64_t n, int k, const double *out,
                                                              double eps, double2 *v) {
    const double h0 = sqrt(out[2 * k]), h1 = sqrt(out[2 * (2 * k + 1)]);
    const bool keep = h1 <= eps * h0;
    for (int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x; j < n; j += (int64_t)gridDim.x * kThreads) {
        double2 x = w[j];
        if (!keep) {
            x.x /= h1;
            x.y /= h1;
        }
        v[j] = x;
    }
}

}  // namespace gs

}  // namespace bem

using namespace bem;

extern "C" int bem_gmres_cgs2(const void *V, int64_t ldv, int k, int64_t n, void *w, void *out, void *v_next,
                              double eps, void *stream) {
    clear_error();
    if (!V || !w || !out || k < 1 || n < 1 || ldv < n) {
        set_error("invalid argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (reinterpret_cast<uintptr_t>(V) % 16 || reinterpret_cast<uintptr_t>(w) % 16 || reinterpret_cast<uintptr_t>(out) % 8 ||
        reinterpret_cast<uintptr_t>(v_next) % 16) {
        set_error("V, w and v_next must be 16-byte aligned complex128 arrays");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    hipStream_t s = (hipStream_t)stream;
    const int64_t nout = 2 * (2 * (int64_t)k + 2);
    hipLaunchKernelGGL(mv::zero_kernel, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, s, (double *)out, nout);
    int64_t blocks = (n + gs::kThreads - 1) / gs::kThreads;
    if (blocks > 512) blocks = 512;
    const dim3 grid((unsigned)blocks);
    const double2 *Vd = (const double2 *)V;
    double2 *wd = (double2 *)w;
    double *od = (double *)out;
    hipLaunchKernelGGL(gs::cgs2_kernel<1>, grid, dim3(gs::kThreads), 0, s, Vd, ldv, k, n, wd, od);
    hipLaunchKernelGGL(gs::cgs2_kernel<2>, grid, dim3(gs::kThreads), 0, s, Vd, ldv, k, n, wd, od);
    hipLaunchKernelGGL(gs::cgs2_kernel<3>, grid, dim3(gs::kThreads), 0, s, Vd, ldv, k, n, wd, od);
    if (v_next)
        hipLaunchKernelGGL(gs::cgs2_scale_kernel, grid, dim3(gs::kThreads), 0, s, wd, n, k, od, eps, (double2 *)v_next);
    return check_launch("cgs2_kernel");
}

extern "C" int bem_dense_add_transpose(void *A, int dtype, int64_t n, int64_t ld, void *stream) {
    clear_error();
    return square_matrix_op(1, A, dtype, n, ld, (hipStream_t)stream);
}

extern "C" int bem_dense_zero(void *A, int dtype, int64_t n, int64_t ld, void *stream) {
    clear_error();
    return square_matrix_op(0, A, dtype, n, ld, (hipStream_t)stream);
}

extern "C" int bem_dense_mirror_upper(void *A, int dtype, int64_t n, int64_t ld, void *stream) {
    clear_error();
    return square_matrix_op(2, A, dtype, n, ld, (hipStream_t)stream);
}

extern "C" int bem_dense_matvec(int trans, const void *A, int a_dtype, int64_t rows, int64_t cols, int64_t ld,
                                const void *x, void *y, void *stream) {
    clear_error();
    if (!A || !x || !y) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (trans < 0 || trans > 2) {
        set_error("trans must be 0 (A x), 1 (A^T x) or 2 (A^H x)");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (rows < 0 || cols < 0 || ld < cols || rows > (int64_t)4 * 2147483647) {
        set_error("invalid matrix shape");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (rows == 0 || cols == 0) {
        // empty sums
        const int64_t n = 2 * (trans == 0 ? rows : cols);
        if (n > 0)
            hipLaunchKernelGGL(mv::zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                               (double *)y, n);
        return check_launch("zero_kernel");
    }
    const double *xd = (const double *)x;
    double *yd = (double *)y;
    hipStream_t s = (hipStream_t)stream;
    switch (a_dtype) {
    case BEM_F32: return mv::launch<float>(trans, A, rows, cols, ld, xd, yd, s);
    case BEM_F64: return mv::launch<double>(trans, A, rows, cols, ld, xd, yd, s);
    case BEM_C64: return mv::launch<float2>(trans, A, rows, cols, ld, xd, yd, s);
    case BEM_C128: return mv::launch<double2>(trans, A, rows, cols, ld, xd, yd, s);
    default: set_error("unknown matrix dtype %d", a_dtype); return BEM_ERR_INVALID_ARGUMENT;
    }
}
