// Elementwise kernels of the host-composed decomposition plans
// (Planned.transformDecompose, Planned.hs:114-119 and 255-293, for lengths
// with a prime factor that no single device plan serves):
//
//  * Good-Thomas (Planned.hs:584-617): the input gathered through the
//    skew-grid map, two batches of row transforms around a transpose, the
//    output gathered through the CRT map; no twiddles;
//  * Cooley-Tukey (Planned.hs:543-557): transposes around two batches of row
//    transforms and the twiddles W_n^(j2 k1), with j2 k1 reduced mod n exactly;
//  * Rader (Planned.hs:650-692) whose length p - 1 convolution is itself a
//    decomposition: the input gathered at g^-q, the spectrum product, and the
//    output X[0] = sum x, X[g^r] = x[0] + c[r].
//
// Complex arrays are planar (re, im) or interleaved (im null) at the plan's
// boundary; intermediates are interleaved.
#include <hip/hip_runtime.h>

#include "fft_device.hpp"
#include "kernels.hpp"

namespace af {

namespace {

template <typename T>
__device__ __forceinline__ typename Cx<T>::C dc_ld(const void *re, const void *im, int64_t i) {
    if (im) return mk(((const T *)re)[i], ((const T *)im)[i]);
    return ((const typename Cx<T>::C *)re)[i];
}
template <typename T>
__device__ __forceinline__ void dc_st(void *re, void *im, int64_t i, typename Cx<T>::C v) {
    if (im) {
        ((T *)re)[i] = v.x;
        ((T *)im)[i] = v.y;
    } else {
        ((typename Cx<T>::C *)re)[i] = v;
    }
}

unsigned dc_grid(int64_t count) {
    const int64_t g = (count + 255) / 256;
    return (unsigned)(g < 65536 ? g : 65536);
}

}  // namespace

// dst[r m + i] = src[r n + map[i]], i < m, r < rows
template <typename T>
__global__ void __launch_bounds__(256)
k_dc_gather(const void *src_re, const void *src_im, int64_t n, void *dst_re, void *dst_im, int64_t m,
            const int32_t *__restrict__ map, int64_t rows) {
    const int64_t total = rows * m, step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += step) {
        const int64_t r = t / m, i = t - r * m;
        dc_st<T>(dst_re, dst_im, t, dc_ld<T>(src_re, src_im, r * n + map[i]));
    }
}

// rows of b x a elements (row-major): element (j2, k1) times W_n^(sign j2 k1)
template <typename T>
__global__ void __launch_bounds__(256)
k_dc_twiddle(typename Cx<T>::C *__restrict__ x, int64_t a, int64_t b, int64_t n, int sign, int64_t rows) {
    const int64_t per = a * b, total = rows * per, step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += step) {
        const int64_t i = t % per, j2 = i / a, k1 = i - j2 * a;
        const int64_t e = (int64_t)(((unsigned __int128)j2 * (unsigned __int128)k1) % (unsigned __int128)n);
        double sn, cs;
        sincospi(2.0 * (double)e / (double)n, &sn, &cs);
        const double2 w = make_double2(cs, sign * sn);
        const typename Cx<T>::C v = x[t];
        x[t] = mk(T((double)v.x * w.x - (double)v.y * w.y), T((double)v.x * w.y + (double)v.y * w.x));
    }
}

// x[r L + i] *= spec[i]
template <typename T>
__global__ void __launch_bounds__(256)
k_dc_spec_mul(typename Cx<T>::C *__restrict__ x, const typename Cx<T>::C *__restrict__ spec, int64_t L, int64_t rows) {
    const int64_t total = rows * L, step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += step) x[t] = cmul(x[t], spec[t % L]);
}

// Rader's output for one row per workgroup: X[0] = sum_j x[j] (accumulated in
// FP64), X[pw[i]] = x[0] + c[i] for i < p - 1.
template <typename T>
__global__ void __launch_bounds__(256)
k_dc_rader_out(const void *x_re, const void *x_im, const typename Cx<T>::C *__restrict__ c,
               const int32_t *__restrict__ pw, int64_t p, void *out_re, void *out_im) {
    const int64_t r = blockIdx.x;
    __shared__ double sre[256], sim[256];
    // every read of the row precedes the first write (in place: out == x)
    const typename Cx<T>::C x0 = dc_ld<T>(x_re, x_im, r * p);
    double ar = 0.0, ai = 0.0;
    for (int64_t j = threadIdx.x; j < p; j += blockDim.x) {
        const typename Cx<T>::C v = dc_ld<T>(x_re, x_im, r * p + j);
        ar += (double)v.x;
        ai += (double)v.y;
    }
    sre[threadIdx.x] = ar;
    sim[threadIdx.x] = ai;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sre[threadIdx.x] += sre[threadIdx.x + h];
            sim[threadIdx.x] += sim[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) dc_st<T>(out_re, out_im, r * p, mk(T(sre[0]), T(sim[0])));
    for (int64_t i = threadIdx.x; i < p - 1; i += blockDim.x)
        dc_st<T>(out_re, out_im, r * p + pw[i], cadd(x0, c[r * (p - 1) + i]));
}

hipError_t launch_dc_gather(int prec, const void *src_re, const void *src_im, int64_t n, void *dst_re, void *dst_im,
                            int64_t m, const int32_t *map, int64_t rows, hipStream_t s) {
    if (rows * m == 0) return hipSuccess;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_dc_gather<double>, dim3(dc_grid(rows * m)), dim3(256), 0, s, src_re, src_im, n, dst_re,
                           dst_im, m, map, rows);
    else
        hipLaunchKernelGGL(k_dc_gather<float>, dim3(dc_grid(rows * m)), dim3(256), 0, s, src_re, src_im, n, dst_re,
                           dst_im, m, map, rows);
    return hipGetLastError();
}

hipError_t launch_dc_twiddle(int prec, void *x, int64_t a, int64_t b, int64_t n, int sign, int64_t rows,
                             hipStream_t s) {
    if (rows * a * b == 0) return hipSuccess;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_dc_twiddle<double>, dim3(dc_grid(rows * a * b)), dim3(256), 0, s, (double2 *)x, a, b, n,
                           sign, rows);
    else
        hipLaunchKernelGGL(k_dc_twiddle<float>, dim3(dc_grid(rows * a * b)), dim3(256), 0, s, (float2 *)x, a, b, n,
                           sign, rows);
    return hipGetLastError();
}

hipError_t launch_dc_spec_mul(int prec, void *x, const void *spec, int64_t L, int64_t rows, hipStream_t s) {
    if (rows * L == 0) return hipSuccess;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_dc_spec_mul<double>, dim3(dc_grid(rows * L)), dim3(256), 0, s, (double2 *)x,
                           (const double2 *)spec, L, rows);
    else
        hipLaunchKernelGGL(k_dc_spec_mul<float>, dim3(dc_grid(rows * L)), dim3(256), 0, s, (float2 *)x,
                           (const float2 *)spec, L, rows);
    return hipGetLastError();
}

hipError_t launch_dc_rader_out(int prec, const void *x_re, const void *x_im, const void *c, const int32_t *pw,
                               int64_t p, int64_t rows, void *out_re, void *out_im, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (rows > 0x7fffffff) return hipErrorInvalidValue;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_dc_rader_out<double>, dim3((unsigned)rows), dim3(256), 0, s, x_re, x_im,
                           (const double2 *)c, pw, p, out_re, out_im);
    else
        hipLaunchKernelGGL(k_dc_rader_out<float>, dim3((unsigned)rows), dim3(256), 0, s, x_re, x_im,
                           (const float2 *)c, pw, p, out_re, out_im);
    return hipGetLastError();
}

}  // namespace af
