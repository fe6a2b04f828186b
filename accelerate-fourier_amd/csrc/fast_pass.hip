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
// FP64 powers.  Four-step twiddles come from FP64 recurrences: one coarse x
// fine table product per butterfly and a running FP64 product over its
// outputs, each value rounded once to the working precision.  (LDS copies of
// the four-step tables, two 16-byte LDS reads and a compensated product per
// element, measured slower on MI355X.)
template <int L> struct FastTables {
    static constexpr bool STAGE_LDS = L <= 1024;
    static constexpr int STAGE_ENTRIES = STAGE_LDS ? L : 0;
};

template <typename T, int L, int B>
__host__ __device__ constexpr size_t fast_lds_bytes() {
    return (size_t)fast_ld(L, B) * B * sizeof(typename Cx<T>::C) + (size_t)FastTables<L>::STAGE_ENTRIES * 16;
}

// Barrier for the LDS exchanges between stages: it orders this wave's LDS
// operations only and leaves its HBM stores in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// HBM accesses of the power-of-two kernel go through buffer resources: a
// uniform base and byte range in SGPRs and a 32-bit byte offset per lane.
// The hardware checks every offset against the range (loads outside read 0,
// stores outside are dropped), and the layouts keep distinct access widths:
// interleaved elements move as one 8- or 16-byte access, planar ones as two.
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc mk_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}
constexpr int kNT = 2;  // cache policy: nontemporal

// The byte offset is split into a per-lane part (voff, a VGPR) and a part that
// is uniform over the wavefront (soff, an SGPR): the per-element steps of a
// butterfly go into soff and cost no vector instruction.
template <typename V, int AUX> __device__ __forceinline__ V bld(Rsrc r, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(V) == 4) return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
    else if constexpr (sizeof(V) == 8) return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
    else return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, AUX));
}
constexpr int kStoreAux = 0;  // stores: default policy (streaming stores measured no faster end to end)
// 16-byte stores keep the whole offset in the VGPR: with an SGPR soffset the
// compiler inserts no wait state between such a store and a vector write of its
// data registers, and on gfx950 the store then wrote partly overwritten data
// under load (wrong FP64 interleaved results, seen at 32+ rows of 65536).
template <typename V> __device__ __forceinline__ void bst(Rsrc r, uint32_t voff, uint32_t soff, V v) {
    if constexpr (sizeof(V) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, kStoreAux);
    else if constexpr (sizeof(V) == 8) {
        typedef uint32_t u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, kStoreAux);
    } else {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, voff + soff, 0, kStoreAux);
    }
}

// One tile (all its lines lie in one row: B divides the lines per row):
// element (b, e) sits at byte offset b * ls + e * es from the tile's base,
// both parts for interleaved data, the real part under re and the imaginary
// part under im for planar data.
struct FastRow {
    Rsrc in_re, in_im, out_re, out_im;
    bool in_planar, out_planar;
    uint32_t line0;
};

// byte strides and byte span of one tile on the input (in = true) or output side
template <typename T>
__device__ __forceinline__ void fast_strides(const AfPass &d, bool in, uint32_t &es, uint32_t &ls, uint32_t &span) {
    const bool planar = (in ? d.in_im : d.out_im) != nullptr;
    const uint32_t S = planar ? (uint32_t)sizeof(T) : (uint32_t)(2 * sizeof(T));
    es = (uint32_t)(in ? d.in_es : d.out_es) * S;
    ls = (uint32_t)(in ? d.in_ls : d.out_ls) * S;
    span = (uint32_t)(d.B - 1) * ls + (uint32_t)(d.L - 1) * es + S;
}

template <typename T, int B>
__device__ __forceinline__ FastRow fast_row(const AfPass &d, int64_t tile) {
    FastRow r;
    const int64_t gl0 = tile * B;
    const int64_t row = gl0 / d.nlines;
    const int64_t line0 = gl0 - row * d.nlines;
    const int64_t ib = row * d.in_bs + line0 * d.in_ls, ob = row * d.out_bs + line0 * d.out_ls;
    uint32_t es, ls, span;
    fast_strides<T>(d, true, es, ls, span);
    r.in_planar = d.in_im != nullptr;
    if (r.in_planar) {
        r.in_re = mk_rsrc((const T *)d.in_re + ib, span);
        r.in_im = mk_rsrc((const T *)d.in_im + ib, span);
    } else {
        r.in_re = r.in_im = mk_rsrc((const T *)d.in_re + 2 * ib, span);
    }
    fast_strides<T>(d, false, es, ls, span);
    r.out_planar = d.out_im != nullptr;
    if (r.out_planar) {
        r.out_re = mk_rsrc((const T *)d.out_re + ob, span);
        r.out_im = mk_rsrc((const T *)d.out_im + ob, span);
    } else {
        r.out_re = r.out_im = mk_rsrc((const T *)d.out_re + 2 * ob, span);
    }
    r.line0 = (uint32_t)line0;
    return r;
}

// First stage inputs: one radix-16 butterfly per thread, straight from HBM.
template <typename T, int L, int B>
__device__ __forceinline__ void fast_load(const AfPass &d, const FastRow &r, typename Cx<T>::C (&v)[16]) {
    using C = typename Cx<T>::C;
    constexpr int NB = L / 16;
    int b, i;
    bf_map<B, NB>(threadIdx.x, d.in_lines_fast != 0, b, i);
    uint32_t es, ls, span;
    fast_strides<T>(d, true, es, ls, span);
    const uint32_t o0 = (uint32_t)b * ls + (uint32_t)i * es, step = (uint32_t)NB * es;
    // nontemporal loads only where the contiguous run of a tile row spans whole
    // 128-byte lines (B lines side by side, or the elements of a row): shorter
    // runs leave the rest of each line to neighbouring tiles, which then find it
    // in L2
    const bool nt = !d.in_lines_fast || B * (int)sizeof(T) >= 128;
    if (r.in_planar && nt) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t so = (uint32_t)j * step;
            v[j] = mk(bld<T, kNT>(r.in_re, o0, so), bld<T, kNT>(r.in_im, o0, so));
        }
    } else if (r.in_planar) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint32_t so = (uint32_t)j * step;
            v[j] = mk(bld<T, 0>(r.in_re, o0, so), bld<T, 0>(r.in_im, o0, so));
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = bld<C, 0>(r.in_re, o0, (uint32_t)j * step);
    }
}

// Last-stage outputs of one butterfly: element j goes to byte offset
// o0 + j * step, after the four-step twiddle (mode 2: tw_e * tw_step^j, FP64,
// rounded once), the scale (mode 1) or as they are (mode 0).
template <typename T, bool PLANAR>
__device__ __forceinline__ void fast_store(const FastRow &r, const int R, const typename Cx<T>::C *v, int mode, double2 tw_e,
                                           const double2 tw_step, const T sc, const uint32_t o0, const uint32_t step) {
    using C = typename Cx<T>::C;
    auto put = [&](int j, C x) {
        const uint32_t so = (uint32_t)j * step;
        if (PLANAR) {
            bst<T>(r.out_re, o0, so, x.x);
            bst<T>(r.out_im, o0, so, x.y);
        } else {
            bst<C>(r.out_re, o0, so, x);
        }
    };
    if (mode == 2) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            put(j, cmul(v[j], mk(T(tw_e.x), T(tw_e.y))));
            if (j + 1 < R) tw_e = cmul(tw_e, tw_step);
        }
    } else if (mode == 1) {
#pragma unroll
        for (int j = 0; j < R; ++j) put(j, cscale(v[j], sc));
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) put(j, v[j]);
    }
}

// W_N^(ftSign a) in FP64 from the coarse x fine tables, 32-bit indexing
__device__ __forceinline__ double2 fast_ft(const AfPass &d, uint32_t a) {
    const uint32_t mask = (1u << d.ftShift) - 1;
    const double2 c = *(const double2 *)((const char *)d.ft_coarse_d + (size_t)((a >> d.ftShift) * 16u));
    const double2 f = *(const double2 *)((const char *)d.ft_fine_d + (size_t)((a & mask) * 16u));
    double2 w = cmul(c, f);
    if (d.ftSign > 0) w.y = -w.y;
    return w;
}

// All stages of one tile whose first-stage inputs are in v; the last stage
// stores to HBM.  SG is the transform sign, a
// compile-time constant so that the products by +-i and the conjugations in
// the butterflies and twiddles fold away.
template <typename T, int L, int B, int SG>
__device__ __forceinline__ void fast_stages(const AfPass &d, const FastRow &r, typename Cx<T>::C (&v)[16],
                                            typename Cx<T>::C *lds, const typename Cx<T>::W *stab) {
    using C = typename Cx<T>::C;
    using PL = P2Plan<L>;
    using FT = FastTables<L>;
    constexpr int NT = B * L / 16;
    constexpr int LD = fast_ld(L, B);
    const int t = threadIdx.x;
    const bool posttw = (d.flags & AFP_POSTTW) != 0;
    constexpr T s = T(SG);
    uint32_t oes, ols, ospan;
    fast_strides<T>(d, false, oes, ols, ospan);
#pragma unroll
    for (int st = 0; st < PL::NST; ++st) {
        const int R = PL::radix(st);
        const int p = PL::p(st);
        const int NB = L / R;          // butterflies per line in this stage
        const int PER = 16 / R;        // butterflies per thread
        const bool first = st == 0, last = st == PL::NST - 1;
        const bool lf = first ? d.in_lines_fast != 0 : (last ? d.out_lines_fast != 0 : false);
        // ---- gather inputs (the first stage's are already in v) ----
        if (!first) {
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                int b, i;
                if (R == 16) bf_map<B, L / 16>(t + q * NT, lf, b, i);
                else if (R == 8) bf_map<B, L / 8>(t + q * NT, lf, b, i);
                else if (R == 4) bf_map<B, L / 4>(t + q * NT, lf, b, i);
                else bf_map<B, L / 2>(t + q * NT, lf, b, i);
#pragma unroll
                for (int j = 0; j < R; ++j) v[q * R + j] = lds[b * LD + fpad(i + j * NB)];
            }
            lds_barrier();
        }
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
                    for (int j = 1; j < R; ++j) v[q * R + j] = twmul(v[q * R + j], stab[j * k * tws], s, false);
                } else {
                    // w^j for j = 1..R-1 from one FP64 load and FP64 powers
                    double2 w1 = ((const double2 *)d.tw_d)[k * tws];
                    if (SG > 0) w1.y = -w1.y;
                    double2 wj = w1;
#pragma unroll
                    for (int j = 1; j < R; ++j) {
                        v[q * R + j] = dmul(v[q * R + j], wj, false);
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
                const uint32_t l = r.line0 + (uint32_t)b;
                const uint32_t o0 = (uint32_t)b * ols + (uint32_t)i * oes, ostep = (uint32_t)NB * oes;
                double2 tw_e = make_double2(1.0, 0.0), tw_step = tw_e;
                int mode = (d.flags & AFP_SCALE) ? 1 : 0;
                if (posttw) {
                    // four-step twiddles W_N^(l e), e = i + j NB (l * e < N <= 2^30), with the
                    // output scale folded into the FP64 start value
                    const uint32_t mask = (uint32_t)(d.ftN - 1);
                    tw_e = fast_ft(d, (l * (uint32_t)i) & mask);
                    tw_step = fast_ft(d, (l * (uint32_t)NB) & mask);
                    if (mode) tw_e = make_double2(tw_e.x * d.scale, tw_e.y * d.scale);
                    mode = 2;
                }
                if (r.out_planar) fast_store<T, true>(r, R, &v[q * R], mode, tw_e, tw_step, T(d.scale), o0, ostep);
                else fast_store<T, false>(r, R, &v[q * R], mode, tw_e, tw_step, T(d.scale), o0, ostep);
            }
        }
        if (!last) lds_barrier();
    }
}

// One launch of a power-of-two pass: one workgroup per tile.
template <typename T, int L, int B, int SG>
__global__ void __launch_bounds__(B * L / 16)
k_fast(const AfPass d) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    using FT = FastTables<L>;
    constexpr int NT = B * L / 16;
    constexpr int LD = fast_ld(L, B);
    extern __shared__ __align__(16) unsigned char smem_fast[];
    C *lds = (C *)smem_fast;
    W *stab = (W *)(smem_fast + (size_t)LD * B * sizeof(C));
    // stage table into LDS; the first barrier of the stages orders it before any use
    if (FT::STAGE_LDS)
        for (int i = threadIdx.x; i < L; i += NT) stab[i] = ((const W *)d.tw)[i];
    const FastRow r = fast_row<T, B>(d, xcd_tile(blockIdx.x, gridDim.x));
    C v[16];
    fast_load<T, L, B>(d, r, v);
    fast_stages<T, L, B, SG>(d, r, v, lds, stab);
}

template <typename T, int L, int B>
static bool try_launch(const AfPass &d, hipStream_t s, hipError_t &err) {
    if (d.L != L || d.B != B) return false;
    const unsigned g = (unsigned)d.ntiles;
    const size_t lds = fast_lds_bytes<T, L, B>();
    if (d.sign < 0) hipLaunchKernelGGL((k_fast<T, L, B, -1>), dim3(g), dim3(B * L / 16), lds, s, d);
    else hipLaunchKernelGGL((k_fast<T, L, B, 1>), dim3(g), dim3(B * L / 16), lds, s, d);
    err = hipGetLastError();
    return true;
}

// ---------------------------------------------------------------------------
// Fused real convolution and its adjoint for power-of-two half lengths
// L = 256 ... 4096 (Real.hs:49-107 around the pointwise product of
// Planned.hs:816-821): a workgroup takes B rows of 2L reals; both signals go
// from HBM straight into first-stage registers as L complex pairs, their
// transforms end in LDS, the untangle -> product -> entangle step runs in LDS
// and the inverse transform of the entangled spectrum goes back to HBM as
// real pairs.
// ---------------------------------------------------------------------------

// All stages of B lines (one radix-16 first-stage butterfly per thread, its
// inputs in v); each line ends in lds in natural order.  Consecutive threads
// take consecutive elements of a line in every stage.  Ends with a barrier.
template <typename T, int L, int B, int SG>
__device__ __forceinline__ void fast_fft_lds(const AfPass &d, typename Cx<T>::C (&v)[16], typename Cx<T>::C *lds,
                                             const typename Cx<T>::W *stab) {
    using PL = P2Plan<L>;
    using FT = FastTables<L>;
    constexpr int NT = B * L / 16;
    constexpr int LD = fast_ld(L, B);
    constexpr T s = T(SG);
    const int t = threadIdx.x;
#pragma unroll
    for (int st = 0; st < PL::NST; ++st) {
        const int R = PL::radix(st);
        const int p = PL::p(st);
        const int NB = L / R;
        const int PER = 16 / R;
        if (st > 0) {
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                int b, i;
                if (R == 16) bf_map<B, L / 16>(t + q * NT, false, b, i);
                else if (R == 8) bf_map<B, L / 8>(t + q * NT, false, b, i);
                else if (R == 4) bf_map<B, L / 4>(t + q * NT, false, b, i);
                else bf_map<B, L / 2>(t + q * NT, false, b, i);
#pragma unroll
                for (int j = 0; j < R; ++j) v[q * R + j] = lds[b * LD + fpad(i + j * NB)];
            }
            lds_barrier();
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            int b, i;
            if (R == 16) bf_map<B, L / 16>(t + q * NT, false, b, i);
            else if (R == 8) bf_map<B, L / 8>(t + q * NT, false, b, i);
            else if (R == 4) bf_map<B, L / 4>(t + q * NT, false, b, i);
            else bf_map<B, L / 2>(t + q * NT, false, b, i);
            const int k = i % p;
            if (p > 1) {
                const int tws = L / (p * R);
                if (FT::STAGE_LDS) {
#pragma unroll
                    for (int j = 1; j < R; ++j) v[q * R + j] = twmul(v[q * R + j], stab[j * k * tws], s, false);
                } else {
                    double2 w1 = ((const double2 *)d.tw_d)[k * tws];
                    if (SG > 0) w1.y = -w1.y;
                    double2 wj = w1;
#pragma unroll
                    for (int j = 1; j < R; ++j) {
                        v[q * R + j] = dmul(v[q * R + j], wj, false);
                        if (j + 1 < R) wj = cmul(wj, w1);
                    }
                }
            }
            if (R == 16) dft<16>(&v[q * R], s);
            else if (R == 8) dft<8>(&v[q * R], s);
            else if (R == 4) dft<4>(&v[q * R], s);
            else dft<2>(&v[q * R], s);
            const int o = (i - k) * R + k;  // the last stage (p = L / R) writes o + j p = i + j L/R
#pragma unroll
            for (int j = 0; j < R; ++j) lds[b * LD + fpad(o + j * p)] = v[q * R + j];
        }
        lds_barrier();
    }
}

// B rows of 2L reals from row0 on, as L complex pairs, into first-stage registers.
template <typename T, int L, int B>
__device__ __forceinline__ void fast_rows_load(const T *src, int64_t row0, typename Cx<T>::C (&v)[16]) {
    using C = typename Cx<T>::C;
    constexpr int NB = L / 16;
    int b, i;
    bf_map<B, NB>(threadIdx.x, false, b, i);
    const Rsrc r = mk_rsrc(src + row0 * 2 * L, (uint32_t)(B * L * sizeof(C)));
    const uint32_t o0 = (uint32_t)((b * L + i) * sizeof(C));
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = bld<C, kNT>(r, o0, (uint32_t)(j * NB * sizeof(C)));
}

// Inverse transform of the B lines in buf (natural order), stored to HBM as
// real pairs from row0 on.  buf is free again when it returns.
template <typename T, int L, int B>
__device__ __forceinline__ void fast_inverse_store(const AfPass &d, typename Cx<T>::C *buf,
                                                   const typename Cx<T>::W *stab, T *dst, int64_t row0) {
    using C = typename Cx<T>::C;
    constexpr int NB = L / 16;
    constexpr int LD = fast_ld(L, B);
    int b, i;
    bf_map<B, NB>(threadIdx.x, false, b, i);
    C v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = buf[b * LD + fpad(i + j * NB)];
    lds_barrier();
    fast_fft_lds<T, L, B, 1>(d, v, buf, stab);
    const Rsrc r = mk_rsrc(dst + row0 * 2 * L, (uint32_t)(B * L * sizeof(C)));
    const uint32_t o0 = (uint32_t)((b * L + i) * sizeof(C));
#pragma unroll
    for (int j = 0; j < 16; ++j) bst<C>(r, o0, (uint32_t)(j * NB * sizeof(C)), buf[b * LD + fpad(i + j * NB)]);
    lds_barrier();
}

// Per line: untangle both half-length spectra to full spectra, multiply (by
// the conjugate of the second for the adjoint's correlations) and entangle the
// product back to the half-length spectrum of the inverse.  One thread takes a
// pair (e, L - e), so dst may be A or Bb.
template <typename T, int L, int B>
__device__ __forceinline__ void fast_spec_product(const typename Cx<T>::C *A, const typename Cx<T>::C *Bb,
                                                  typename Cx<T>::C *dst, const typename Cx<T>::W *aux, bool conj_b) {
    using C = typename Cx<T>::C;
    constexpr int NT = B * L / 16;
    constexpr int LD = fast_ld(L, B);
    constexpr int H = L / 2 + 1;
    for (int idx = threadIdx.x; idx < B * H; idx += NT) {
        const int b = idx / H, e = idx - b * H;
        const int er = (L - e) & (L - 1);
        const int ie = b * LD + fpad(e), ir = b * LD + fpad(er);
        const C a0 = A[ie], a1 = A[ir], b0 = Bb[ie], b1 = Bb[ir];
        C sa_lo, sa_hi, sb_lo, sb_hi;
        untangle<T>(a0, a1, aux[e], false, sa_lo, sa_hi);
        untangle<T>(b0, b1, aux[e], false, sb_lo, sb_hi);
        if (conj_b) { sb_lo = cconj(sb_lo); sb_hi = cconj(sb_hi); }
        const C z0 = entangle<T>(cmul(sa_lo, sb_lo), cmul(sa_hi, sb_hi), aux[e], false);
        C z1 = z0;
        if (er != e) {
            C ta_lo, ta_hi, tb_lo, tb_hi;
            untangle<T>(a1, a0, aux[er], false, ta_lo, ta_hi);
            untangle<T>(b1, b0, aux[er], false, tb_lo, tb_hi);
            if (conj_b) { tb_lo = cconj(tb_lo); tb_hi = cconj(tb_hi); }
            z1 = entangle<T>(cmul(ta_lo, tb_lo), cmul(ta_hi, tb_hi), aux[er], false);
        }
        dst[ie] = z0;
        if (er != e) dst[ir] = z1;
    }
}

template <typename T, int L, int B>
__host__ __device__ constexpr size_t fast_rconv_lds_bytes() {
    return 2 * (size_t)fast_ld(L, B) * B * sizeof(typename Cx<T>::C) + (size_t)FastTables<L>::STAGE_ENTRIES * 16;
}

// Forward (VJP = false): out0 = fromSpectrum(toSpectrum in0 * toSpectrum in1), in0 = x, in1 = k.
// Adjoint (VJP = true) with cotangent in0 = w: out0 = dk from in1 = x and out1 = dx from
// in2 = k, each fromSpectrum(S_w conj S_other).
template <typename T, int L, int B, bool VJP>
__global__ void __launch_bounds__(B * L / 16)
k_fast_rconv(const AfPass d, const T *__restrict__ in0, const T *__restrict__ in1, const T *__restrict__ in2,
             T *__restrict__ out0, T *__restrict__ out1) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    using FT = FastTables<L>;
    constexpr int NT = B * L / 16;
    constexpr int LD = fast_ld(L, B);
    extern __shared__ __align__(16) unsigned char smem_rc[];
    C *A = (C *)smem_rc;
    C *Bb = A + LD * B;
    W *stab = (W *)(Bb + LD * B);
    if (FT::STAGE_LDS)
        for (int i = threadIdx.x; i < L; i += NT) stab[i] = ((const W *)d.tw)[i];
    const W *aux = (const W *)d.aux_tw;
    const int64_t row0 = (int64_t)blockIdx.x * B;
    C v[16];
    fast_rows_load<T, L, B>(in0, row0, v);
    fast_fft_lds<T, L, B, -1>(d, v, A, stab);
    fast_rows_load<T, L, B>(in1, row0, v);
    fast_fft_lds<T, L, B, -1>(d, v, Bb, stab);
    fast_spec_product<T, L, B>(A, Bb, Bb, aux, VJP);
    lds_barrier();
    fast_inverse_store<T, L, B>(d, Bb, stab, out0, row0);
    if (VJP) {
        fast_rows_load<T, L, B>(in2, row0, v);
        fast_fft_lds<T, L, B, -1>(d, v, Bb, stab);
        fast_spec_product<T, L, B>(A, Bb, Bb, aux, true);
        lds_barrier();
        fast_inverse_store<T, L, B>(d, Bb, stab, out1, row0);
    }
}

// X(type, L, rows per workgroup)
#define AF_RCONV_SHAPES(X)                                                                           \
    X(float, 256, 16) X(float, 512, 8) X(float, 1024, 4) X(float, 2048, 2) X(float, 4096, 1)        \
    X(double, 256, 8) X(double, 512, 4) X(double, 1024, 2) X(double, 2048, 1) X(double, 4096, 1)

template <typename T, int L, int B>
static bool try_rconv(const AfPass &d, const void *in0, const void *in1, const void *in2, void *out0, void *out1,
                      bool vjp, hipStream_t s, hipError_t &err) {
    if (d.L != L || d.nrows % B) return false;
    const size_t lds = fast_rconv_lds_bytes<T, L, B>();
    const dim3 g((unsigned)(d.nrows / B)), bl(B * L / 16);
    if (vjp)
        hipLaunchKernelGGL((k_fast_rconv<T, L, B, true>), g, bl, lds, s, d, (const T *)in0, (const T *)in1,
                           (const T *)in2, (T *)out0, (T *)out1);
    else
        hipLaunchKernelGGL((k_fast_rconv<T, L, B, false>), g, bl, lds, s, d, (const T *)in0, (const T *)in1,
                           (const T *)nullptr, (T *)out0, (T *)nullptr);
    err = hipGetLastError();
    return true;
}

hipError_t launch_fast_rconv(int prec, const AfPass &d, const void *in0, const void *in1, const void *in2, void *out0,
                             void *out1, bool vjp, hipStream_t s, bool &done) {
    hipError_t err = hipSuccess;
    done = false;
    if (d.flags & AFP_PLAIN_TW) return hipSuccess;
    if (!d.tw || !d.tw_d || !d.aux_tw || d.nrows <= 0 || d.nrows >= ((int64_t)1 << 31)) return hipSuccess;
#define AF_TRY_RC(T, L_, B_)                                                  \
    if (!done && (prec == PREC_F64) == (sizeof(T) == 8))                      \
        done = try_rconv<T, L_, B_>(d, in0, in1, in2, out0, out1, vjp, s, err);
    AF_RCONV_SHAPES(AF_TRY_RC)
#undef AF_TRY_RC
    return err;
}

// Flags the fast kernel implements; anything else (plain FP32 twiddles
// included) goes to the generic pass.
static const int32_t kFastFlags = AFP_POSTTW | AFP_SCALE;

bool fast_pass_eligible(int prec, const AfPass &d) {
    if (d.flags & ~kFastFlags) return false;
    if (d.in_map || d.out_map || d.in_mul || d.out_mul) return false;
    if (d.in_len != INT64_MAX || d.out_len != INT64_MAX) return false;
    if (d.L < 256 || d.L > 4096 || (d.L & (d.L - 1))) return false;
    if (!d.tw_d || ((d.flags & AFP_POSTTW) && (!d.ft_coarse_d || !d.ft_fine_d))) return false;
    if ((d.flags & AFP_POSTTW) && (d.ftN & (d.ftN - 1))) return false;
    if (d.nlines % d.B) return false;
    if (d.ntiles >= ((int64_t)1 << 31)) return false;
    // byte offsets inside a tile and into the FP64 four-step tables stay 32-bit
    const int64_t eb = prec == PREC_F64 ? 8 : 4;
    const int64_t span_i = ((int64_t)(d.B - 1) * d.in_ls + (int64_t)(d.L - 1) * d.in_es) * (d.in_im ? eb : 2 * eb) + eb;
    const int64_t span_o = ((int64_t)(d.B - 1) * d.out_ls + (int64_t)(d.L - 1) * d.out_es) * (d.out_im ? eb : 2 * eb) + eb;
    if (span_i >= ((int64_t)1 << 32) || span_o >= ((int64_t)1 << 32)) return false;
    if ((d.flags & AFP_POSTTW) && (d.ftN >> d.ftShift) * 16 + 16 >= ((int64_t)1 << 32)) return false;
    const int64_t lds = (int64_t)fast_ld(d.L, d.B) * d.B * (prec == PREC_F64 ? 16 : 8) + (d.L <= 1024 ? d.L : 0) * 16;
    if (lds > 160 * 1024) return false;
    return (int64_t)d.B * d.L / 16 <= 1024;
}

// Tile shapes with a power-of-two kernel: lengths 256 ... 4096 and line
// counts B from one wavefront (B * L / 16 >= 64 threads) up to what one tile
// holds in LDS.  X(type, L, B).
#define AF_FAST_SHAPES(X)                                                                          \
    X(float, 256, 4) X(float, 256, 8) X(float, 256, 16) X(float, 256, 32)                          \
    X(float, 512, 2) X(float, 512, 4) X(float, 512, 8) X(float, 512, 16) X(float, 512, 32)        \
    X(float, 1024, 1) X(float, 1024, 2) X(float, 1024, 4) X(float, 1024, 8) X(float, 1024, 16)     \
    X(float, 2048, 1) X(float, 2048, 2) X(float, 2048, 4) X(float, 2048, 8)                        \
    X(float, 4096, 1) X(float, 4096, 2) X(float, 4096, 4)                                          \
    X(double, 256, 4) X(double, 256, 8) X(double, 256, 16)                                         \
    X(double, 512, 2) X(double, 512, 4) X(double, 512, 8) X(double, 512, 16)                       \
    X(double, 1024, 1) X(double, 1024, 2) X(double, 1024, 4) X(double, 1024, 8)                    \
    X(double, 2048, 1) X(double, 2048, 2) X(double, 2048, 4)                                       \
    X(double, 4096, 1) X(double, 4096, 2)

static bool fast_has_shape(int prec, int L, int B) {
#define AF_HAS(T, L_, B_) if ((prec == PREC_F64) == (sizeof(T) == 8) && L == L_ && B == B_) return true;
    AF_FAST_SHAPES(AF_HAS)
#undef AF_HAS
    return false;
}

int fast_tile_lines(int prec, int64_t L, int B) {
    if (L < 256 || L > 4096 || (L & (L - 1))) return B;
    for (int b = B; b >= 1; b /= 2)
        if (fast_has_shape(prec, (int)L, b)) return b;
    return B;
}

hipError_t launch_fast(int prec, const AfPass &d, hipStream_t s, bool &done) {
    hipError_t err = hipSuccess;
    done = false;
    if (!fast_pass_eligible(prec, d)) return hipSuccess;
#define AF_TRY(T, L_, B_) if (!done && (prec == PREC_F64) == (sizeof(T) == 8)) done = try_launch<T, L_, B_>(d, s, err);
    AF_FAST_SHAPES(AF_TRY)
#undef AF_TRY
    return err;
}

template <typename T, int L, int B>
static void set_attr() {
    (void)hipFuncSetAttribute((const void *)k_fast<T, L, B, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute((const void *)k_fast<T, L, B, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

template <typename T, int L, int B>
static void set_attr_rc() {
    (void)hipFuncSetAttribute((const void *)k_fast_rconv<T, L, B, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    (void)hipFuncSetAttribute((const void *)k_fast_rconv<T, L, B, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
}

void init_fast_attributes() {
#define AF_ATTR(T, L_, B_) set_attr<T, L_, B_>();
    AF_FAST_SHAPES(AF_ATTR)
#undef AF_ATTR
#define AF_ATTR_RC(T, L_, B_) set_attr_rc<T, L_, B_>();
    AF_RCONV_SHAPES(AF_ATTR_RC)
#undef AF_ATTR_RC
}

}  // namespace af
