// Specialised pass kernel for power-of-two line lengths (256 ... 8192).
//
// The generic pass (fft_device.hpp) loads a tile into LDS, runs a runtime list
// of stages and stores it from LDS.  For power-of-two lines the stage list is
// known at compile time: radix-16 stages with a radix 2, 4 or 8 stage last.
// Here every thread owns sixteen elements per stage; the first stage reads its
// inputs straight from HBM into registers and the last stage writes its
// outputs straight back, so a two-stage line (L = 256) touches LDS exactly
// once with a single barrier.  The thread-to-butterfly mapping of the first
// and last stages follows the memory layout: consecutive lanes take
// consecutive lines when the lines are columns (four-step passes) and
// consecutive elements when they are rows, so each wavefront access covers
// whole 128-byte segments.
#include <hip/hip_runtime.h>

#include "fft_device.hpp"
#include "kernels.hpp"

namespace af {

template <int L> struct P2Plan {
    static constexpr int LOG2 = L == 256 ? 8 : L == 512 ? 9 : L == 1024 ? 10 : L == 2048 ? 11 : L == 4096 ? 12 : 13;
    static constexpr int NST = (LOG2 + 3) / 4;               // stages
    static constexpr int LASTR = 1 << (LOG2 - 4 * (NST - 1)); // radix of the last stage (16 if LOG2 % 4 == 0)
    static constexpr int radix(int s) { return s == NST - 1 ? LASTR : 16; }
    static constexpr int p(int s) { return 1 << (4 * s); }
};

// Load / store of one complex element at a 64-bit element address.
template <typename T>
__device__ __forceinline__ typename Cx<T>::C gload(const AfPass &d, int64_t a) {
    if (d.in_im) return mk(__builtin_nontemporal_load(&((const T *)d.in_re)[a]),
                           __builtin_nontemporal_load(&((const T *)d.in_im)[a]));
    return ((const typename Cx<T>::C *)d.in_re)[a];
}
template <typename T>
__device__ __forceinline__ void gstore(const AfPass &d, int64_t a, typename Cx<T>::C v) {
    if (d.out_im) {
        ((T *)d.out_re)[a] = v.x;
        ((T *)d.out_im)[a] = v.y;
    } else {
        ((typename Cx<T>::C *)d.out_re)[a] = v;
    }
}

// Butterfly index bf in [0, B*L/R) to (line, i): lines fastest or elements fastest.
template <int B, int NB>
__device__ __forceinline__ void bf_map(int bf, bool lines_fast, int &b, int &i) {
    if (lines_fast) { b = bf % B; i = bf / B; }
    else { b = bf / NB; i = bf % NB; }
}

// LDS layout: one spare slot every 16 elements inside a line (stride-16
// Stockham writes then spread over the banks) and an odd line pitch.
__device__ __host__ constexpr int fpad(int e) { return e + (e >> 4); }
__device__ __host__ constexpr int fast_ld(int L, int B) { return B > 1 ? fpad(L) + 1 : fpad(L); }

template <typename T, int L, int B>
__global__ void __launch_bounds__(B * L / 16)
k_fast(const AfPass d) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    using PL = P2Plan<L>;
    constexpr int NT = B * L / 16;
    constexpr int LD = fast_ld(L, B);
    extern __shared__ __align__(16) unsigned char smem_fast[];
    C *lds = (C *)smem_fast;
    const int t = threadIdx.x;
    const bool plain = (d.flags & AFP_PLAIN_TW) != 0;
    const T s = T(d.sign);
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t gl0 = tile * B;
    // all lines of a tile lie in one row (B divides the lines per row)
    const int64_t row = gl0 / d.nlines;
    const int64_t line0 = gl0 - row * d.nlines;
    const int64_t ib = row * d.in_bs, ob = row * d.out_bs;

    C v[16];
#pragma unroll
    for (int st = 0; st < PL::NST; ++st) {
        const int R = PL::radix(st);
        const int p = PL::p(st);
        const int NB = L / R;          // butterflies per line in this stage
        const int PER = 16 / R;        // butterflies per thread
        const bool first = st == 0, last = st == PL::NST - 1;
        const bool lf = first ? d.in_lines_fast != 0 : (last ? d.out_lines_fast != 0 : false);
        // ---- gather inputs ----
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int b, i;
            if (R == 16) bf_map<B, L / 16>(t + q * NT, lf, b, i);
            else if (R == 8) bf_map<B, L / 8>(t + q * NT, lf, b, i);
            else if (R == 4) bf_map<B, L / 4>(t + q * NT, lf, b, i);
            else bf_map<B, L / 2>(t + q * NT, lf, b, i);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int e = i + j * NB;
                if (first) {
                    const int64_t l = line0 + b;
                    v[q * R + j] = gload<T>(d, ib + l * d.in_ls + (int64_t)e * d.in_es);
                } else {
                    v[q * R + j] = lds[b * LD + fpad(e)];
                }
            }
        }
        if (!first) __syncthreads();
        // ---- twiddle, butterfly, scatter ----
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int b, i;
            if (R == 16) bf_map<B, L / 16>(t + q * NT, lf, b, i);
            else if (R == 8) bf_map<B, L / 8>(t + q * NT, lf, b, i);
            else if (R == 4) bf_map<B, L / 4>(t + q * NT, lf, b, i);
            else bf_map<B, L / 2>(t + q * NT, lf, b, i);
            const int k = i % p;
            if (p > 1) {
                // w^j for j = 1..R-1 from one FP64 load: W_L^(k L / (p R)) and its powers
                const int tws = L / (p * R);
                double2 w1 = ((const double2 *)d.tw_d)[k * tws];
                if (d.sign > 0) w1.y = -w1.y;
                double2 wj = w1;
#pragma unroll
                for (int j = 1; j < R; ++j) {
                    v[q * R + j] = dmul(v[q * R + j], wj, plain);
                    if (j + 1 < R) wj = cmul(wj, w1);
                }
            }
            if (R == 16) dft<16>(&v[q * R], s);
            else if (R == 8) dft<8>(&v[q * R], s);
            else if (R == 4) dft<4>(&v[q * R], s);
            else dft<2>(&v[q * R], s);
            if (!last) {
                const int o = (i - k) * R + k;
#pragma unroll
                for (int j = 0; j < R; ++j) lds[b * LD + fpad(o + j * p)] = v[q * R + j];
            } else {
                const int64_t l = line0 + b;
                // four-step twiddles W_N^(l e), e = i + j NB: base W_N^(l i) times step^j,
                // step = W_N^(l NB), in FP64 (l * e < N <= 2^30 for every four-step length)
                double2 tw_e = make_double2(1.0, 0.0), tw_step = tw_e;
                if (d.flags & AFP_POSTTW) {
                    const uint32_t mask = (uint32_t)(d.ftN - 1);
                    tw_e = ft_twiddle_d(d, ((uint32_t)l * (uint32_t)i) & mask);
                    tw_step = ft_twiddle_d(d, ((uint32_t)l * (uint32_t)NB) & mask);
                }
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const int e = i + j * NB;
                    C x = v[q * R + j];
                    if (d.flags & AFP_POSTTW) {
                        x = dmul(x, tw_e, plain);
                        if (j + 1 < R) tw_e = cmul(tw_e, tw_step);
                    }
                    if (d.flags & AFP_SCALE) x = cscale(x, T(d.scale));
                    gstore<T>(d, ob + l * d.out_ls + (int64_t)e * d.out_es, x);
                }
            }
        }
        if (!last) __syncthreads();
    }
}

template <typename T, int L, int B>
static bool try_launch(const AfPass &d, hipStream_t s, hipError_t &err) {
    if (d.L != L || d.B != B) return false;
    const size_t lds = (size_t)fast_ld(L, B) * B * sizeof(typename Cx<T>::C);
    hipLaunchKernelGGL((k_fast<T, L, B>), dim3((unsigned)d.ntiles), dim3(B * L / 16), lds, s, d);
    err = hipGetLastError();
    return true;
}

// Flags the fast kernel implements; anything else goes to the generic pass.
static const int32_t kFastFlags = AFP_POSTTW | AFP_SCALE | AFP_PLAIN_TW;

bool fast_pass_eligible(int prec, const AfPass &d) {
    if (d.flags & ~kFastFlags) return false;
    if (d.in_map || d.out_map || d.in_mul || d.out_mul) return false;
    if (d.in_len != INT64_MAX || d.out_len != INT64_MAX) return false;
    if (d.L < 256 || d.L > 4096 || (d.L & (d.L - 1))) return false;
    if (!d.tw_d || ((d.flags & AFP_POSTTW) && (!d.ft_coarse_d || !d.ft_fine_d))) return false;
    if ((d.flags & AFP_POSTTW) && (d.ftN & (d.ftN - 1))) return false;
    if (d.nlines % d.B) return false;
    if (d.ntiles >= ((int64_t)1 << 31)) return false;
    const int lds = fast_ld(d.L, d.B) * d.B * (prec == PREC_F64 ? 16 : 8);
    if (lds > 160 * 1024) return false;
    return (int64_t)d.B * d.L / 16 <= 1024;
}

hipError_t launch_fast(int prec, const AfPass &d, hipStream_t s, bool &done) {
    hipError_t err = hipSuccess;
    done = false;
    if (!fast_pass_eligible(prec, d)) return hipSuccess;
    if (prec == PREC_F32) {
        done = try_launch<float, 256, 16>(d, s, err) || try_launch<float, 256, 32>(d, s, err) || try_launch<float, 256, 4>(d, s, err) ||
               try_launch<float, 256, 1>(d, s, err) || try_launch<float, 512, 8>(d, s, err) ||
               try_launch<float, 1024, 4>(d, s, err) || try_launch<float, 1024, 1>(d, s, err) ||
               try_launch<float, 2048, 2>(d, s, err) || try_launch<float, 4096, 4>(d, s, err) ||
               try_launch<float, 4096, 1>(d, s, err);
    } else {
        done = try_launch<double, 256, 8>(d, s, err) || try_launch<double, 256, 2>(d, s, err) ||
               try_launch<double, 256, 1>(d, s, err) || try_launch<double, 512, 4>(d, s, err) ||
               try_launch<double, 1024, 2>(d, s, err) || try_launch<double, 1024, 1>(d, s, err) ||
               try_launch<double, 2048, 1>(d, s, err) || try_launch<double, 4096, 2>(d, s, err);
    }
    return err;
}

template <typename T, int L, int B>
static void set_attr() {
    (void)hipFuncSetAttribute((const void *)k_fast<T, L, B>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

void init_fast_attributes() {
    set_attr<float, 256, 16>(); set_attr<float, 256, 32>(); set_attr<float, 256, 4>(); set_attr<float, 256, 1>(); set_attr<float, 512, 8>();
    set_attr<float, 1024, 4>(); set_attr<float, 1024, 1>(); set_attr<float, 2048, 2>(); set_attr<float, 4096, 4>();
    set_attr<float, 4096, 1>();
    set_attr<double, 256, 8>(); set_attr<double, 256, 2>(); set_attr<double, 256, 1>(); set_attr<double, 512, 4>();
    set_attr<double, 1024, 2>(); set_attr<double, 1024, 1>(); set_attr<double, 2048, 1>(); set_attr<double, 4096, 2>();
}

}  // namespace af
