// HIP kernels of the transform path for gfx950 and their host launchers.
//
//  k_pass          one pass (tiles of B lines in LDS) of a transform plan
//  k_pass_grouped  the same for many pass descriptors in one launch (uneven batches)
//  k_real_conv     fused real FFT convolution y = fromSpectrum(toSpectrum x * toSpectrum k)
//  k_real_conv_vjp its adjoint (dx, dk from the cotangent w) in one launch
//  k_gen_*         FP64 generation of twiddles, chirps and Rader kernels on the GPU
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fft_device.hpp"
#include "kernels.hpp"

namespace af {

// ---------------------------------------------------------------------------
// Pass kernels
// ---------------------------------------------------------------------------

// All tiles of one workgroup, inlined into the kernel.  The descriptor is read
// in place from the kernel-argument segment through a pointer the compiler
// cannot see through: its fields are then loaded where they are used instead
// of all being held in registers (which spills), and no thread copies the
// descriptor (as a reference to the parameter would make it do).
template <typename T, bool FULL>
__device__ __forceinline__ void pass_tiles_body(const AfPass &d, typename Cx<T>::C *lds, TileLines &tl,
                                                typename Cx<T>::C *s_x0, typename Cx<T>::C *s_dc) {
    const bool one = (int64_t)gridDim.x == d.ntiles;
    for (int64_t t = blockIdx.x; t < d.ntiles; t += gridDim.x)
        pass_tile<T, FULL>(d, one ? xcd_tile(t, d.ntiles) : t, lds, tl, s_x0, s_dc);
}

template <typename T>
__device__ __forceinline__ const T *opaque(const T *p) {
    asm volatile("" : "+v"(p));
    return p;
}

template <typename T, bool FULL>
__global__ void __launch_bounds__(1024) k_pass(const AfPass d) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TileLines tl;
    __shared__ typename Cx<T>::C s_x0[64], s_dc[64];
    (void)d;
    const AfPass *dp = opaque((const AfPass *)(const void *)__builtin_amdgcn_kernarg_segment_ptr());
    pass_tiles_body<T, FULL>(*dp, (typename Cx<T>::C *)smem, tl, s_x0, s_dc);
}

// tile_prefix[i] = first global tile of descriptor i; tile_prefix[ndesc] = total.
// The descriptor in LDS is read directly.  Inlined: as a call, its register
// saves took 196 bytes of scratch per lane (FP32) in every workgroup, about as
// much memory traffic as the tile data of a workgroup with one or two tiles.
template <typename T, bool FULL>
__device__ __forceinline__ void pass_tiles_grouped(const AfPass *__restrict__ descs, const int64_t *__restrict__ tile_prefix,
                                                const int ndesc, typename Cx<T>::C *lds, TileLines &tl,
                                                typename Cx<T>::C *s_x0, typename Cx<T>::C *s_dc) {
    // the tile's descriptor is copied into LDS, where the tile reads it
    __shared__ AfPass sd;
    const int64_t total = tile_prefix[ndesc];
    int cur = -1;
    for (int64_t g = blockIdx.x; g < total; g += gridDim.x) {
        int lo = 0, hi = ndesc - 1;
        while (lo < hi) {  // last descriptor whose first tile is <= g
            const int mid = (lo + hi + 1) >> 1;
            if (tile_prefix[mid] <= g) lo = mid;
            else hi = mid - 1;
        }
        if (lo != cur) {
            __syncthreads();  // the previous tile is done with sd
            const uint32_t *src = (const uint32_t *)&descs[lo];
            uint32_t *dst = (uint32_t *)&sd;
            for (int i = threadIdx.x; i < (int)(sizeof(AfPass) / 4); i += blockDim.x) dst[i] = src[i];
            __syncthreads();
            cur = lo;
        }
        pass_tile<T, FULL>(sd, g - tile_prefix[lo], lds, tl, s_x0, s_dc);
    }
}

template <typename T, bool FULL>
__global__ void __launch_bounds__(1024)
k_pass_grouped(const AfPass *__restrict__ descs, const int64_t *__restrict__ tile_prefix, const int ndesc) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TileLines tl;
    __shared__ typename Cx<T>::C s_x0[64], s_dc[64];
    pass_tiles_grouped<T, FULL>(descs, tile_prefix, ndesc, (typename Cx<T>::C *)smem, tl, s_x0, s_dc);
}

// ---------------------------------------------------------------------------
// Fused real FFT convolution, one row per workgroup (Real.hs:49-107 composed
// with the pointwise product of Planned.hs:816-821).
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ void load_real_row(typename Cx<T>::C *buf, const AfPass &d, const T *src) {
    const auto *s2 = (const typename Cx<T>::C *)src;
    for (int e = threadIdx.x; e < d.L; e += blockDim.x) buf[pidx(e, d.padshift)] = s2[e];
}
template <typename T>
__device__ __forceinline__ void store_real_row(const typename Cx<T>::C *buf, const AfPass &d, T *dst) {
    auto *d2 = (typename Cx<T>::C *)dst;
    for (int e = threadIdx.x; e < d.L; e += blockDim.x) d2[e] = buf[pidx(e, d.padshift)];
}

// conj_b: multiply by conj of the second spectrum (correlation, for the adjoint).
// Pairs (e, L-e) are handled by one thread, so the update is in place.
template <typename T>
__device__ __forceinline__ void spectral_product(typename Cx<T>::C *A, typename Cx<T>::C *Bb,
                                                 typename Cx<T>::C *dst, const AfPass &d, bool conj_b) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    const int L = d.L, ps = d.padshift;
    const bool plain = (d.flags & AFP_PLAIN_TW) != 0;
    const W *aux = (const W *)d.aux_tw;
    for (int e = threadIdx.x; e <= L / 2; e += blockDim.x) {
        const int er = (L - e) % L;
        C z0, z1;
        rconv_pair<T>(A[pidx(e, ps)], A[pidx(er, ps)], Bb[pidx(e, ps)], Bb[pidx(er, ps)], aux[e], plain, conj_b, z0,
                      z1);
        dst[pidx(e, ps)] = z0;
        if (er != e) dst[pidx(er, ps)] = z1;
    }
}

template <typename T>
__global__ void __launch_bounds__(1024)
k_real_conv(const AfPass dv, const T *__restrict__ x, const T *__restrict__ k, T *__restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem[];
    using C = typename Cx<T>::C;
    // the descriptor read in place from the kernel arguments (as in k_pass):
    // taken by value, all of its fields were held in registers and spilled
    (void)dv;
    const AfPass &d = *opaque((const AfPass *)(const void *)__builtin_amdgcn_kernarg_segment_ptr());
    const int64_t N = 2 * (int64_t)d.L;
    C *A = (C *)smem;
    C *Bb = A + pidx(d.ld, d.padshift);
    for (int64_t r = blockIdx.x; r < d.nrows; r += gridDim.x) {
        load_real_row<T>(A, d, x + r * N);
        load_real_row<T>(Bb, d, k + r * N);
        __syncthreads();
        run_stages<T>(A, d, T(-1));
        run_stages<T>(Bb, d, T(-1));
        spectral_product<T>(A, Bb, A, d, false);
        __syncthreads();
        run_stages<T>(A, d, T(1));
        store_real_row<T>(A, d, y + r * N);
        __syncthreads();
    }
}

// dk = fromSpectrum(S_w conj S_x), dx = fromSpectrum(S_w conj S_k)
template <typename T>
__global__ void __launch_bounds__(1024)
k_real_conv_vjp(const AfPass dv, const T *__restrict__ x, const T *__restrict__ k, const T *__restrict__ w,
                T *__restrict__ dx, T *__restrict__ dk) {
    extern __shared__ __align__(16) unsigned char smem[];
    using C = typename Cx<T>::C;
    (void)dv;  // read in place, as in k_real_conv
    const AfPass &d = *opaque((const AfPass *)(const void *)__builtin_amdgcn_kernarg_segment_ptr());
    const int64_t N = 2 * (int64_t)d.L;
    C *A = (C *)smem;
    C *Bb = A + pidx(d.ld, d.padshift);
    for (int64_t r = blockIdx.x; r < d.nrows; r += gridDim.x) {
        load_real_row<T>(A, d, w + r * N);
        load_real_row<T>(Bb, d, x + r * N);
        __syncthreads();
        run_stages<T>(A, d, T(-1));
        run_stages<T>(Bb, d, T(-1));
        spectral_product<T>(A, Bb, Bb, d, true);
        __syncthreads();
        run_stages<T>(Bb, d, T(1));
        store_real_row<T>(Bb, d, dk + r * N);
        __syncthreads();
        load_real_row<T>(Bb, d, k + r * N);
        __syncthreads();
        run_stages<T>(Bb, d, T(-1));
        spectral_product<T>(A, Bb, Bb, d, true);
        __syncthreads();
        run_stages<T>(Bb, d, T(1));
        store_real_row<T>(Bb, d, dx + r * N);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Real packing in global memory, for half lengths the fused pass cannot take.
// Z holds rows of L interleaved half-length values; full spectra have 2L.
// ---------------------------------------------------------------------------

template <typename T>
__global__ void k_untangle(const typename Cx<T>::C *__restrict__ Z, void *out_re, void *out_im,
                           const typename Cx<T>::W *__restrict__ aux, int64_t L, int64_t rows, int plain) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= L * rows) return;
    const int64_t r = g / L, k = g - r * L;
    typename Cx<T>::C lo, hi;
    untangle<T>(Z[r * L + k], Z[r * L + (L - k) % L], aux[k], plain != 0, lo, hi);
    st_c<T>(out_re, out_im, r * 2 * L + k, lo);
    st_c<T>(out_re, out_im, r * 2 * L + k + L, hi);
}

// mode 0: fromSpectrum's entangling of S; mode 1: the same after the
// Hermitian projection (G_k + conj G_-k) / 2 (the adjoint of toSpectrum).
template <typename T>
__global__ void k_entangle(const void *s_re, const void *s_im, typename Cx<T>::C *__restrict__ z,
                           const typename Cx<T>::W *__restrict__ aux, int64_t L, int64_t rows, int plain, int mode) {
    using C = typename Cx<T>::C;
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= L * rows) return;
    const int64_t r = g / L, k = g - r * L, b = r * 2 * L;
    C p0 = ld_c<T>(s_re, s_im, b + k), p1 = ld_c<T>(s_re, s_im, b + k + L);
    if (mode == 1) {
        const int64_t kr = (L - k) % L;
        if (k == 0) {
            p0 = mk(p0.x, T(0));
            p1 = mk(p1.x, T(0));
        } else {
            p0 = cscale(cadd(p0, cconj(ld_c<T>(s_re, s_im, b + kr + L))), T(0.5));
            p1 = cscale(cadd(p1, cconj(ld_c<T>(s_re, s_im, b + kr))), T(0.5));
        }
    }
    z[r * L + k] = entangle<T>(p0, p1, aux[k], plain != 0);
}

// adjoint of the entangling: dS_k = U_k + c, dS_{k+L} = U_k - c, c = -i W_{2L}^k U_k
template <typename T>
__global__ void k_entangle_adj(const typename Cx<T>::C *__restrict__ U, void *out_re, void *out_im,
                               const typename Cx<T>::W *__restrict__ aux, int64_t L, int64_t rows, int plain) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= L * rows) return;
    const int64_t r = g / L, k = g - r * L;
    const auto u = U[r * L + k];
    const auto t = twmul(u, aux[k], T(-1), plain != 0);
    const auto c = mk(t.y, -t.x);
    st_c<T>(out_re, out_im, r * 2 * L + k, cadd(u, c));
    st_c<T>(out_re, out_im, r * 2 * L + k + L, csub(u, c));
}

// Pointwise product of two untangled spectra given by their half-length
// transforms, entangled again: the middle of the real convolution.
template <typename T>
__global__ void k_real_product(const typename Cx<T>::C *__restrict__ A, const typename Cx<T>::C *__restrict__ Bz,
                               typename Cx<T>::C *__restrict__ z, const typename Cx<T>::W *__restrict__ aux,
                               int64_t L, int64_t rows, int plain, int conj_b) {
    using C = typename Cx<T>::C;
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= L * rows) return;
    const int64_t r = g / L, k = g - r * L, kr = (L - k) % L, b = r * L;
    C sa_lo, sa_hi, sb_lo, sb_hi;
    untangle<T>(A[b + k], A[b + kr], aux[k], plain != 0, sa_lo, sa_hi);
    untangle<T>(Bz[b + k], Bz[b + kr], aux[k], plain != 0, sb_lo, sb_hi);
    if (conj_b) { sb_lo = cconj(sb_lo); sb_hi = cconj(sb_hi); }
    z[b + k] = entangle<T>(cmul(sa_lo, sb_lo), cmul(sa_hi, sb_hi), aux[k], plain != 0);
}

// ---------------------------------------------------------------------------
// Table generation in FP64 (sincospi of exactly reduced rational angles)
// ---------------------------------------------------------------------------

// out[i] = exp(-2 pi i (i*mul mod N) / N)
template <typename WT>
__global__ void k_gen_twiddles(WT *__restrict__ out, int64_t N, int64_t count, int64_t mul) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int64_t a = (int64_t)(((unsigned __int128)i * (uint64_t)mul) % (uint64_t)N);
    double s, c;
    sincospi(2.0 * (double)a / (double)N, &s, &c);
    to_w(make_double2(c, -s), out[i]);
}

// Bluestein chirp: out[k] = exp(sign i pi (kt^2 mod 2n) / n), kt = k or k - M
// (wrapped, as Private.chirp with padLen M, Private.hs:428-436), optionally
// conjugated.  kt^2 mod 2n is exact in 64-bit integers.
__global__ void k_gen_chirp(double2 *__restrict__ out, int64_t n, int64_t count, int sign, int64_t M, int conj) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= count) return;
    int64_t kt = k;
    if (M > 0 && 2 * k >= M) kt = k - M;
    if (kt < 0) kt = -kt;
    const uint64_t r = (uint64_t)(((unsigned __int128)kt * (uint64_t)kt) % (uint64_t)(2 * n));
    double s, c;
    sincospi((double)r / (double)n, &s, &c);
    double im = sign * s;
    if (conj) im = -im;
    out[k] = make_double2(c, im);
}

// Rader kernel: out[i] = exp(sign 2 pi i powg[i] / p)  (levelCachePrime, Planned.hs:655-664)
__global__ void k_gen_rader_kernel(double2 *__restrict__ out, const int32_t *__restrict__ powg, int64_t p,
                                   int64_t count, int sign) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    double s, c;
    sincospi(2.0 * (double)powg[i] / (double)p, &s, &c);
    out[i] = make_double2(c, sign * s);
}

// out[i] = scale * in[src(i)] in the table format; with N1 > 0 the source is the
// four-step row order: i = k1*N2 + k2 reads in[k1 + N1*k2].
template <typename WT>
__global__ void k_convert(WT *__restrict__ out, const double2 *__restrict__ in, int64_t count, double scale,
                          int64_t N1, int64_t N2) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t src = i;
    if (N1 > 0) src = (i / N2) + N1 * (i % N2);
    const double2 v = in[src];
    to_w(make_double2(v.x * scale, v.y * scale), out[i]);
}

// ---------------------------------------------------------------------------
// Two-for-one real packing (Real.hs:129-236): elementwise over rows of n,
// x over the elements of a row, y over the rows (grid-strided).
// ---------------------------------------------------------------------------

// untangleSpectra (Real.hs:168-187): for the spectrum z of x + i y,
// a_k = (z_k + conj z_-k) / 2 and b_k = (z_k - conj z_-k) (-i / 2).
// With h > 1 the rows form h x n matrices and the partner of (y, k) is
// ((-y) mod h, (-k) mod n): untangleSpectra2d with Cyclic.reverse2d
// (Real.hs:175-180, Cyclic.hs:30-39).
template <typename T>
__global__ void k_two_untangle(const void *z_re, const void *z_im, void *a_re, void *a_im, void *b_re, void *b_im,
                               int64_t n, int64_t rows, int64_t h) {
    using C = typename Cx<T>::C;
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t kr = k ? n - k : 0;
    for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) {
        const int64_t slab = r / h, y = r - slab * h;
        const int64_t rr = slab * h + (y ? h - y : 0);
        const C z = ld_c<T>(z_re, z_im, r * n + k), w = ld_c<T>(z_re, z_im, rr * n + kr);
        const T h = T(0.5);
        // (z - conj w) (-i / 2) = ((z.im + w.im) / 2, -(z.re - w.re) / 2)
        st_c<T>(a_re, a_im, r * n + k, mk((z.x + w.x) * h, (z.y - w.y) * h));
        st_c<T>(b_re, b_im, r * n + k, mk((z.y + w.y) * h, (w.x - z.x) * h));
    }
}

// entangleSpectra (Real.hs:208-236): z = c + d i.
template <typename T>
__global__ void k_two_entangle(const void *c_re, const void *c_im, const void *d_re, const void *d_im, void *z_re,
                               void *z_im, int64_t n, int64_t rows) {
    using C = typename Cx<T>::C;
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) {
        const C c = ld_c<T>(c_re, c_im, r * n + k), d = ld_c<T>(d_re, d_im, r * n + k);
        st_c<T>(z_re, z_im, r * n + k, mk(c.x - d.y, c.y + d.x));
    }
}

// ---------------------------------------------------------------------------
// Batched transpose of R x C complex matrices (the reference's
// LinAlg.transpose over the last two axes; with R = K M and C = N it is
// cycleDim3, Private.hs:414-425).  32 x 32 tiles through LDS (one pad column
// against bank conflicts), 32 x 8 threads, matrices of the batch over z.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_transpose(const void *in_re, const void *in_im, void *out_re, void *out_im, int64_t R, int64_t C, int64_t nb) {
    using C2 = typename Cx<T>::C;
    __shared__ C2 tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t c0 = (int64_t)blockIdx.x * 32, r0 = (int64_t)blockIdx.y * 32;
    for (int64_t b = blockIdx.z; b < nb; b += gridDim.z) {
        const int64_t base = b * R * C;
#pragma unroll
        for (int k = 0; k < 32; k += 8) {
            const int64_t r = r0 + ty + k, c = c0 + tx;
            if (r < R && c < C) tile[ty + k][tx] = ld_c<T>(in_re, in_im, base + r * C + c);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 32; k += 8) {
            const int64_t c = c0 + ty + k, r = r0 + tx;
            if (r < R && c < C) st_c<T>(out_re, out_im, base + c * R + r, tile[tx][ty + k]);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

void init_kernel_attributes() {
    const int lds = 160 * 1024;
    (void)hipFuncSetAttribute((const void *)k_pass<float, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_pass<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_pass<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_pass<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_pass_grouped<float, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_pass_grouped<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_pass_grouped<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_pass_grouped<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_real_conv<float>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_real_conv<double>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_real_conv_vjp<float>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_real_conv_vjp<double>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    init_karatsuba_attributes();
    init_fast_attributes();
    init_fast_ext_attributes();
    (void)hipGetLastError();
}

static inline unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// LDS of one workgroup of k_pass / k_pass_grouped: the dynamic tile plus the
// static line table and Rader slots.  A launch above the 160 KiB a workgroup
// can hold is refused here instead of reaching the GPU.
static bool pass_lds_fits(int prec, int64_t dynamic_bytes) {
    const int64_t stat = (int64_t)sizeof(TileLines) + 2 * 64 * (prec == PREC_F64 ? 16 : 8);
    return dynamic_bytes >= 0 && dynamic_bytes + stat <= 160 * 1024;
}

template <typename T>
static hipError_t launch_pass_t(const AfPass &d, hipStream_t s) {
    if (d.ntiles <= 0) return hipSuccess;
    if (!pass_lds_fits(sizeof(T) == 8 ? PREC_F64 : PREC_F32, d.lds_bytes) || d.B > 64) return hipErrorInvalidValue;
    // a workgroup loops over tiles once there are more than four rounds of the
    // 256 CUs at the tile's LDS occupancy
    const int per_cu = std::max(1, (int)((160 * 1024) / std::max(d.lds_bytes + 4096, 1)));
    const int64_t cap = (int64_t)256 * per_cu * 4;
    const int64_t g = d.ntiles < cap ? d.ntiles : cap;
    if (d.flags & kRarePassFlags) hipLaunchKernelGGL((k_pass<T, true>), dim3((unsigned)g), dim3(d.nthreads), d.lds_bytes, s, d);
    else hipLaunchKernelGGL((k_pass<T, false>), dim3((unsigned)g), dim3(d.nthreads), d.lds_bytes, s, d);
    return hipGetLastError();
}

hipError_t launch_pass(int prec, const AfPass &d, hipStream_t s) {
    if (d.ntiles <= 0) return hipSuccess;
    bool done = false;
    hipError_t e = launch_fast(prec, d, s, done);
    if (done || e != hipSuccess) return e;
    e = launch_fast_ext(prec, d, s, done);
    if (done || e != hipSuccess) return e;
    return prec == PREC_F64 ? launch_pass_t<double>(d, s) : launch_pass_t<float>(d, s);
}

hipError_t launch_pass_grouped(int prec, const AfPass *descs, const int64_t *prefix, int ndesc,
                               int64_t total_tiles, int nthreads, int lds_bytes, bool full, hipStream_t s) {
    if (total_tiles <= 0) return hipSuccess;
    if (!pass_lds_fits(prec, lds_bytes)) return hipErrorInvalidValue;
    const int per_cu = std::max(1, (int)((160 * 1024) / std::max(lds_bytes + 4096, 1)));
    const int64_t cap = (int64_t)256 * per_cu * 4;
    const int64_t g = total_tiles < cap ? total_tiles : cap;
    const dim3 gd((unsigned)g), bd(nthreads);
    if (prec == PREC_F64 && full)
        hipLaunchKernelGGL((k_pass_grouped<double, true>), gd, bd, lds_bytes, s, descs, prefix, ndesc);
    else if (prec == PREC_F64)
        hipLaunchKernelGGL((k_pass_grouped<double, false>), gd, bd, lds_bytes, s, descs, prefix, ndesc);
    else if (full)
        hipLaunchKernelGGL((k_pass_grouped<float, true>), gd, bd, lds_bytes, s, descs, prefix, ndesc);
    else
        hipLaunchKernelGGL((k_pass_grouped<float, false>), gd, bd, lds_bytes, s, descs, prefix, ndesc);
    return hipGetLastError();
}

hipError_t launch_real_conv(int prec, const AfPass &d, const void *x, const void *k, void *y, hipStream_t s) {
    if (d.nrows <= 0) return hipSuccess;
    bool done = false;
    const hipError_t e = launch_fast_rconv(prec, d, x, k, nullptr, y, nullptr, false, s, done);
    if (done || e != hipSuccess) return e;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_real_conv<double>, dim3((unsigned)d.nrows), dim3(d.nthreads), d.lds_bytes, s, d,
                           (const double *)x, (const double *)k, (double *)y);
    else
        hipLaunchKernelGGL(k_real_conv<float>, dim3((unsigned)d.nrows), dim3(d.nthreads), d.lds_bytes, s, d,
                           (const float *)x, (const float *)k, (float *)y);
    return hipGetLastError();
}

hipError_t launch_real_conv_vjp(int prec, const AfPass &d, const void *x, const void *k, const void *w, void *dx,
                                void *dk, hipStream_t s) {
    if (d.nrows <= 0) return hipSuccess;
    bool done = false;
    const hipError_t e = launch_fast_rconv(prec, d, w, x, k, dk, dx, true, s, done);
    if (done || e != hipSuccess) return e;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_real_conv_vjp<double>, dim3((unsigned)d.nrows), dim3(d.nthreads), d.lds_bytes, s, d,
                           (const double *)x, (const double *)k, (const double *)w, (double *)dx, (double *)dk);
    else
        hipLaunchKernelGGL(k_real_conv_vjp<float>, dim3((unsigned)d.nrows), dim3(d.nthreads), d.lds_bytes, s, d,
                           (const float *)x, (const float *)k, (const float *)w, (float *)dx, (float *)dk);
    return hipGetLastError();
}

hipError_t launch_two_pack(int prec, bool untangle, const void *p_re, const void *p_im, const void *q_re,
                           const void *q_im, void *o1_re, void *o1_im, void *o2_re, void *o2_im, int64_t n,
                           int64_t rows, hipStream_t s, int64_t h) {
    if (n <= 0 || rows <= 0) return hipSuccess;
    if (h < 1 || rows % h) return hipErrorInvalidValue;
    const dim3 g(grid_for(n, 256), (unsigned)std::min<int64_t>(rows, 65535)), b(256);
    if (untangle) {
        if (prec == PREC_F64)
            hipLaunchKernelGGL(k_two_untangle<double>, g, b, 0, s, p_re, p_im, o1_re, o1_im, o2_re, o2_im, n, rows, h);
        else
            hipLaunchKernelGGL(k_two_untangle<float>, g, b, 0, s, p_re, p_im, o1_re, o1_im, o2_re, o2_im, n, rows, h);
    } else {
        if (prec == PREC_F64)
            hipLaunchKernelGGL(k_two_entangle<double>, g, b, 0, s, p_re, p_im, q_re, q_im, o1_re, o1_im, n, rows);
        else
            hipLaunchKernelGGL(k_two_entangle<float>, g, b, 0, s, p_re, p_im, q_re, q_im, o1_re, o1_im, n, rows);
    }
    return hipGetLastError();
}

hipError_t launch_transpose(int prec, const void *in_re, const void *in_im, void *out_re, void *out_im, int64_t R,
                            int64_t C, int64_t nb, hipStream_t s) {
    if (R <= 0 || C <= 0 || nb <= 0) return hipSuccess;
    if ((C + 31) / 32 > INT32_MAX || (R + 31) / 32 > 65535) return hipErrorInvalidValue;
    const dim3 g((unsigned)((C + 31) / 32), (unsigned)((R + 31) / 32), (unsigned)std::min<int64_t>(nb, 65535));
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_transpose<double>, g, dim3(256), 0, s, in_re, in_im, out_re, out_im, R, C, nb);
    else
        hipLaunchKernelGGL(k_transpose<float>, g, dim3(256), 0, s, in_re, in_im, out_re, out_im, R, C, nb);
    return hipGetLastError();
}

hipError_t gen_twiddles(int prec, void *out, int64_t N, int64_t count, int64_t mul, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_gen_twiddles<double2>, dim3(grid_for(count, 256)), dim3(256), 0, s, (double2 *)out, N,
                           count, mul);
    else
        hipLaunchKernelGGL(k_gen_twiddles<float4>, dim3(grid_for(count, 256)), dim3(256), 0, s, (float4 *)out, N,
                           count, mul);
    return hipGetLastError();
}

hipError_t gen_twiddles_d(void *out, int64_t N, int64_t count, int64_t mul, hipStream_t s) {
    return gen_twiddles(PREC_F64, out, N, count, mul, s);
}

hipError_t gen_chirp_d(void *out, int64_t n, int64_t count, int sign, int64_t M, int conj, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_chirp, dim3(grid_for(count, 256)), dim3(256), 0, s, (double2 *)out, n, count, sign, M,
                       conj);
    return hipGetLastError();
}

hipError_t gen_rader_kernel_d(void *out, const int32_t *powg, int64_t p, int64_t count, int sign, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_rader_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, (double2 *)out, powg, p,
                       count, sign);
    return hipGetLastError();
}

hipError_t convert_to_w(int prec, void *out, const void *in, int64_t count, double scale, int64_t N1, int64_t N2,
                        hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_convert<double2>, dim3(grid_for(count, 256)), dim3(256), 0, s, (double2 *)out,
                           (const double2 *)in, count, scale, N1, N2);
    else
        hipLaunchKernelGGL(k_convert<float4>, dim3(grid_for(count, 256)), dim3(256), 0, s, (float4 *)out,
                           (const double2 *)in, count, scale, N1, N2);
    return hipGetLastError();
}

hipError_t convert_to_plain(int prec, void *out, const void *in, int64_t count, double scale, int64_t N1, int64_t N2,
                            hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_convert<double2>, dim3(grid_for(count, 256)), dim3(256), 0, s, (double2 *)out,
                           (const double2 *)in, count, scale, N1, N2);
    else
        hipLaunchKernelGGL(k_convert<float2>, dim3(grid_for(count, 256)), dim3(256), 0, s, (float2 *)out,
                           (const double2 *)in, count, scale, N1, N2);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_pack_t(int op, const void *a, const void *a_im, const void *b, void *o_re, void *o_im,
                                const void *aux, int64_t L, int64_t rows, int plain, int arg, hipStream_t s) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    const int64_t n = L * rows;
    if (n <= 0) return hipSuccess;
    const dim3 g(grid_for(n, 256)), bl(256);
    switch (op) {
    case PACK_UNTANGLE:
        hipLaunchKernelGGL(k_untangle<T>, g, bl, 0, s, (const C *)a, o_re, o_im, (const W *)aux, L, rows, plain);
        break;
    case PACK_ENTANGLE:
        hipLaunchKernelGGL(k_entangle<T>, g, bl, 0, s, a, a_im, (C *)o_re, (const W *)aux, L, rows, plain, arg);
        break;
    case PACK_ENTANGLE_ADJ:
        hipLaunchKernelGGL(k_entangle_adj<T>, g, bl, 0, s, (const C *)a, o_re, o_im, (const W *)aux, L, rows, plain);
        break;
    default:
        hipLaunchKernelGGL(k_real_product<T>, g, bl, 0, s, (const C *)a, (const C *)b, (C *)o_re, (const W *)aux, L,
                           rows, plain, arg);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_pack(int prec, int op, const void *a, const void *a_im, const void *b, void *o_re, void *o_im,
                       const void *aux, int64_t L, int64_t rows, int plain, int arg, hipStream_t s) {
    return prec == PREC_F64 ? launch_pack_t<double>(op, a, a_im, b, o_re, o_im, aux, L, rows, plain, arg, s)
                            : launch_pack_t<float>(op, a, a_im, b, o_re, o_im, aux, L, rows, plain, arg, s);
}

}  // namespace af
