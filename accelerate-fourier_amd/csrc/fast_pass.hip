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

// Stage twiddles come from an LDS copy of the W_L table when it is small
// (L <= 1024: at most 16 KiB), otherwise from one FP64 load per butterfly and
// FP64 powers.  Four-step twiddles come from FP64 recurrences (an LDS-table
// variant exists below, disabled by kFtLdsEntries).
template <int L> struct FastTables {
    static constexpr bool STAGE_LDS = L <= 1024;
    static constexpr int STAGE_ENTRIES = STAGE_LDS ? L : 0;
};
// Measured on MI355X: LDS copies of the four-step tables (two 16-byte LDS reads
// and a compensated product per element) ran slower than one FP64 load pair per
// butterfly plus FP64 recurrences, so the LDS variant is kept off.
__host__ __device__ constexpr int kFtLdsEntries() { return 0; }

__host__ __device__ inline bool ft_in_lds(int64_t ftN, int ftShift) {
    const int64_t T = (int64_t)1 << ftShift, nc = (ftN + T - 1) / T;
    return nc + T <= kFtLdsEntries();
}

// W_N^(sign a) = coarse[a >> shift] * fine[a & mask]: FP32 product of two hi/lo
// pairs with first-order lo corrections (about one ulp), FP64 plain product.
__device__ __forceinline__ float2 ft_prod(const float4 *co, const float4 *fi, uint32_t a, int shift, int sign) {
    const float4 c = co[a >> shift], f = fi[a & ((1u << shift) - 1)];
    float re = fmaf(c.x, f.z, fmaf(-c.y, f.w, fmaf(c.z, f.x, -c.w * f.y)));
    float im = fmaf(c.x, f.w, fmaf(c.y, f.z, fmaf(c.z, f.y, c.w * f.x)));
    re = fmaf(c.x, f.x, fmaf(-c.y, f.y, re));
    im = fmaf(c.x, f.y, fmaf(c.y, f.x, im));
    return make_float2(re, sign > 0 ? -im : im);
}
__device__ __forceinline__ double2 ft_prod(const double2 *co, const double2 *fi, uint32_t a, int shift, int sign) {
    double2 w = cmul(co[a >> shift], fi[a & ((1u << shift) - 1)]);
    if (sign > 0) w.y = -w.y;
    return w;
}

template <typename T, int L, int B>
__host__ __device__ constexpr size_t fast_lds_bytes(bool ft_lds, int ft_entries) {
    return (size_t)fast_ld(L, B) * B * sizeof(typename Cx<T>::C) +
           (size_t)(FastTables<L>::STAGE_ENTRIES + (ft_lds ? ft_entries : 0)) * 16;
}

template <typename T, int L, int B>
__global__ void __launch_bounds__(B * L / 16)
k_fast(const AfPass d) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    using PL = P2Plan<L>;
    using FT = FastTables<L>;
    constexpr int NT = B * L / 16;
    constexpr int LD = fast_ld(L, B);
    extern __shared__ __align__(16) unsigned char smem_fast[];
    C *lds = (C *)smem_fast;
    W *stab = (W *)(smem_fast + (size_t)LD * B * sizeof(C));
    W *ftc = stab + FT::STAGE_ENTRIES;
    const int t = threadIdx.x;
    const bool plain = (d.flags & AFP_PLAIN_TW) != 0;
    const bool posttw = (d.flags & AFP_POSTTW) != 0;
    const bool ftl = posttw && ft_in_lds(d.ftN, d.ftShift);
    const int ftT = 1 << d.ftShift;
    const int ftNc = (int)((d.ftN + ftT - 1) >> d.ftShift);
    W *ftf = ftc + (ftl ? ftNc : 0);
    const T s = T(d.sign);
    // tables into LDS; the first barrier below orders them before any use
    if (FT::STAGE_LDS)
        for (int i = t; i < L; i += NT) stab[i] = ((const W *)d.tw)[i];
    if (ftl) {
        for (int i = t; i < ftNc; i += NT) ftc[i] = ((const W *)d.ft_coarse)[i];
        for (int i = t; i < ftT; i += NT) ftf[i] = ((const W *)d.ft_fine)[i];
    }
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t gl0 = tile * B;
    // all lines of a tile lie in one row (B divides the lines per row)
    const int64_t row = gl0 / d.nlines;
    const int64_t line0 = gl0 - row * d.nlines;
    const int64_t ib = row * d.in_bs + line0 * d.in_ls, ob = row * d.out_bs + line0 * d.out_ls;
    const int ies = (int)d.in_es, oes = (int)d.out_es, ils = (int)d.in_ls, ols = (int)d.out_ls;
    // row base pointers; element offsets below stay 32-bit (b * ls + e * es < 2^31)
    const T *in_re = (const T *)d.in_re, *in_im = (const T *)d.in_im;
    const C *in_c = (const C *)d.in_re;
    T *out_re = (T *)d.out_re, *out_im = (T *)d.out_im;
    C *out_c = (C *)d.out_re;
    if (in_im) { in_re += ib; in_im += ib; } else { in_c += ib; }
    if (out_im) { out_re += ob; out_im += ob; } else { out_c += ob; }

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
                    const int off = b * ils + e * ies;
                    if (in_im) v[q * R + j] = mk(__builtin_nontemporal_load(in_re + off),
                                                 __builtin_nontemporal_load(in_im + off));
                    else v[q * R + j] = in_c[off];
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
                const int tws = L / (p * R);
                if (FT::STAGE_LDS) {
#pragma unroll
                    for (int j = 1; j < R; ++j) v[q * R + j] = twmul(v[q * R + j], stab[j * k * tws], s, plain);
                } else {
                    // w^j for j = 1..R-1 from one FP64 load and FP64 powers
                    double2 w1 = ((const double2 *)d.tw_d)[k * tws];
                    if (d.sign > 0) w1.y = -w1.y;
                    double2 wj = w1;
#pragma unroll
                    for (int j = 1; j < R; ++j) {
                        v[q * R + j] = dmul(v[q * R + j], wj, plain);
                        if (j + 1 < R) wj = cmul(wj, w1);
                    }
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
                const uint32_t l = (uint32_t)(line0 + b);
                // four-step twiddles W_N^(l e), e = i + j NB (l * e < N <= 2^30)
                double2 tw_e = make_double2(1.0, 0.0), tw_step = tw_e;
                const uint32_t mask = (uint32_t)(d.ftN - 1);
                if (posttw && !ftl) {
                    tw_e = ft_twiddle_d(d, (l * (uint32_t)i) & mask);
                    tw_step = ft_twiddle_d(d, (l * (uint32_t)NB) & mask);
                }
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const int e = i + j * NB;
                    C x = v[q * R + j];
                    if (posttw) {
                        if (ftl) {
                            x = cmul(x, ft_prod(ftc, ftf, (l * (uint32_t)e) & mask, d.ftShift, d.ftSign));
                        } else {
                            // the FP64 value rounded once to the working precision
                            x = cmul(x, mk(T(tw_e.x), T(tw_e.y)));
                            if (j + 1 < R) tw_e = cmul(tw_e, tw_step);
                        }
                    }
                    if (d.flags & AFP_SCALE) x = cscale(x, T(d.scale));
                    const int off = b * ols + e * oes;
                    if (out_im) {
                        out_re[off] = x.x;
                        out_im[off] = x.y;
                    } else {
                        out_c[off] = x;
                    }
                }
            }
        }
        if (!last) __syncthreads();
    }
}

template <typename T, int L, int B>
static bool try_launch(const AfPass &d, hipStream_t s, hipError_t &err) {
    if (d.L != L || d.B != B) return false;
    const bool ftl = (d.flags & AFP_POSTTW) && ft_in_lds(d.ftN, d.ftShift);
    const int T_ = 1 << d.ftShift;
    const int ft_entries = ftl ? (int)((d.ftN + T_ - 1) / T_) + T_ : 0;
    const size_t lds = fast_lds_bytes<T, L, B>(ftl, ft_entries);
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
    const int64_t row_span = (int64_t)(d.B - 1) * d.in_ls + (int64_t)(d.L - 1) * d.in_es;
    const int64_t row_span_o = (int64_t)(d.B - 1) * d.out_ls + (int64_t)(d.L - 1) * d.out_es;
    if (row_span >= ((int64_t)1 << 31) || row_span_o >= ((int64_t)1 << 31)) return false;
    const int T_ = (d.flags & AFP_POSTTW) ? 1 << d.ftShift : 0;
    const bool ftl = (d.flags & AFP_POSTTW) && ft_in_lds(d.ftN, d.ftShift);
    const int64_t ft_entries = ftl ? (d.ftN + T_ - 1) / T_ + T_ : 0;
    const int64_t lds = (int64_t)fast_ld(d.L, d.B) * d.B * (prec == PREC_F64 ? 16 : 8) +
                        ((d.L <= 1024 ? d.L : 0) + ft_entries) * 16;
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
