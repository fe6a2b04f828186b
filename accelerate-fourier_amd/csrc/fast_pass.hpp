// Device building blocks of the power-of-two pass kernels (fast_pass.hip,
// fast_ext.hip): compile-time stage plans, tile addressing through buffer
// resources, first-stage loads, the stage sequence with its HBM stores, and
// the in-LDS transform used by the fused convolutions.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_device.hpp"

namespace af {

template <int L> struct P2Plan {
    static constexpr int LOG2 = L == 256 ? 8 : L == 512 ? 9 : L == 1024 ? 10 : L == 2048 ? 11 : L == 4096 ? 12
                                : L == 8192 ? 13 : 14;
    static constexpr int NST = (LOG2 + 3) / 4;               // stages
    static constexpr int LASTR = 1 << (LOG2 - 4 * (NST - 1)); // radix of the last stage (16 if LOG2 % 4 == 0)
    static constexpr int radix(int s) { return s == NST - 1 ? LASTR : 16; }
    static constexpr int p(int s) { return 1 << (4 * s); }
};

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
// (L <= 1024: at most 16 KiB).  Longer lines (2048 ... 16384: three or four
// stages) take the second stage's W_L^(k j L/256) = W_256^(k j) from an LDS
// copy of W_256 (4 KiB) and the later stages' from one FP64 load per butterfly
// and FP64 powers (or, in the fused convolution, from registers: FastRegTw).  Four-step twiddles come from FP64 recurrences: one coarse x
// fine table product per butterfly and a running FP64 product over its
// outputs, each value rounded once to the working precision.  (LDS copies of
// the four-step tables, two 16-byte LDS reads and a compensated product per
// element, measured slower on MI355X.)
template <int L> struct FastTables {
    static constexpr bool STAGE_LDS = L <= 1024;
    static constexpr int STAGE_ENTRIES = STAGE_LDS ? L : 256;
};

// The LDS stage table of the line length (W_L, or W_256 beyond 1024) from
// the W_L table of the working precision; read after the first stage's
// barrier.
template <typename T, int L>
__device__ __forceinline__ void fast_tables_fill(const AfPass &d, typename Cx<T>::W *stab) {
    using W = typename Cx<T>::W;
    constexpr int step = FastTables<L>::STAGE_LDS ? 1 : L / 256;
    for (int i = threadIdx.x; i < FastTables<L>::STAGE_ENTRIES; i += blockDim.x) stab[i] = ((const W *)d.tw)[i * step];
}

// The last-stage twiddles of the fused convolutions' transforms at lengths
// 2048 and 4096 held in registers: each thread forms the 15 powers w^j of its
// butterflies' twiddle bases once, in FP64, and every transform of the
// workgroup reuses them (conjugated for the inverse) instead of a table load
// and a chain of FP64 products per transform.
template <int L> struct FastRegTw {
    double2 w[15];  // butterfly q, power j at q (R - 1) + j - 1
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

// First-stage load of the extended kernel: as fast_load, and element e of
// line l (row position pos = l in_ls + e in_es) reads as zero from
// pos >= in_len on and is multiplied by in_mul[pos] (plain complex values)
// when in_mul is set -- the Bluestein chirp and zero padding.
template <typename T, int L, int B>
__device__ __forceinline__ void fast_load_ext(const AfPass &d, const FastRow &r, typename Cx<T>::C (&v)[16]) {
    using C = typename Cx<T>::C;
    constexpr int NB = L / 16;
    int b, i;
    bf_map<B, NB>(threadIdx.x, d.in_lines_fast != 0, b, i);
    uint32_t es, ls, span;
    fast_strides<T>(d, true, es, ls, span);
    const uint32_t o0 = (uint32_t)b * ls + (uint32_t)i * es, step = (uint32_t)NB * es;
    const uint32_t p0 = (r.line0 + (uint32_t)b) * (uint32_t)d.in_ls + (uint32_t)i * (uint32_t)d.in_es;
    const uint32_t pstep = (uint32_t)NB * (uint32_t)d.in_es;
    const uint32_t plen = d.in_len < (int64_t)UINT32_MAX ? (uint32_t)d.in_len : UINT32_MAX;
    const bool mul = d.in_mul != nullptr;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint32_t pos = p0 + (uint32_t)j * pstep;
        C x = mk(T(0), T(0));
        if (pos < plen) {
            const uint32_t off = o0 + (uint32_t)j * step;
            if (r.in_planar) x = mk(bld<T, 0>(r.in_re, off, 0), bld<T, 0>(r.in_im, off, 0));
            else x = bld<C, 0>(r.in_re, off, 0);
            if (mul) x = cmul(x, gld<C>(d.in_mul, pos));
        }
        v[j] = x;
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

// Last-stage outputs of the extended kernel: as fast_store, and output
// position o = l out_ls + e out_es is written only below out_len and after a
// multiplication by out_mul[o] when out_mul is set.
template <typename T, bool PLANAR>
__device__ __forceinline__ void fast_store_ext(const AfPass &d, const FastRow &r, const int R,
                                               const typename Cx<T>::C *v, int mode, double2 tw_e,
                                               const double2 tw_step, const T sc, const uint32_t o0,
                                               const uint32_t step, const uint32_t p0, const uint32_t pstep) {
    using C = typename Cx<T>::C;
    const uint32_t plen = d.out_len < (int64_t)UINT32_MAX ? (uint32_t)d.out_len : UINT32_MAX;
    const bool mul = d.out_mul != nullptr;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        C x = v[j];
        if (mode == 2) {
            x = cmul(x, mk(T(tw_e.x), T(tw_e.y)));
            if (j + 1 < R) tw_e = cmul(tw_e, tw_step);
        } else if (mode == 1) {
            x = cscale(x, sc);
        }
        const uint32_t pos = p0 + (uint32_t)j * pstep;
        if (pos < plen) {
            if (mul) x = cmul(x, gld<C>(d.out_mul, pos));
            const uint32_t off = o0 + (uint32_t)j * step;
            if (PLANAR) {
                bst<T>(r.out_re, off, 0, x.x);
                bst<T>(r.out_im, off, 0, x.y);
            } else {
                bst<C>(r.out_re, off, 0, x);
            }
        }
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
template <typename T, int L, int B, int SG, bool EXT = false>
__device__ __forceinline__ void fast_stages(const AfPass &d, const FastRow &r, typename Cx<T>::C (&v)[16],
                                            typename Cx<T>::C *lds, const typename Cx<T>::W *stab) {
    using PL = P2Plan<L>;
    using FT = FastTables<L>;
    constexpr int NT = B * L / 16;
    constexpr int LD = fast_ld(L, B);
    const int t = threadIdx.x;
    const bool posttw = (d.flags & AFP_POSTTW) != 0;
    constexpr T s = T(SG);
    uint32_t oes, ols, ospan;
    fast_strides<T>(d, false, oes, ols, ospan);
#pragma clang loop unroll(full)  // (plain `unroll` left four-stage lines rolled, with scratch)
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
                } else if (st == 1) {  // the second of three stages: p R = 256, W_256 in LDS
#pragma unroll
                    for (int j = 1; j < R; ++j) v[q * R + j] = twmul(v[q * R + j], stab[k * j], s, false);
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
                if constexpr (EXT) {
                    const uint32_t p0 = l * (uint32_t)d.out_ls + (uint32_t)i * (uint32_t)d.out_es;
                    const uint32_t pstep = (uint32_t)NB * (uint32_t)d.out_es;
                    if (r.out_planar)
                        fast_store_ext<T, true>(d, r, R, &v[q * R], mode, tw_e, tw_step, T(d.scale), o0, ostep, p0, pstep);
                    else
                        fast_store_ext<T, false>(d, r, R, &v[q * R], mode, tw_e, tw_step, T(d.scale), o0, ostep, p0, pstep);
                } else {
                    if (r.out_planar) fast_store<T, true>(r, R, &v[q * R], mode, tw_e, tw_step, T(d.scale), o0, ostep);
                    else fast_store<T, false>(r, R, &v[q * R], mode, tw_e, tw_step, T(d.scale), o0, ostep);
                }
            }
        }
        if (!last) lds_barrier();
    }
}

// The register twiddles of the thread's last-stage butterflies, from the
// FP64 W_L table.
template <int L, int B>
__device__ __forceinline__ void fast_regtw_init(const AfPass &d, FastRegTw<L> &rt) {
    using PL = P2Plan<L>;
    static_assert(PL::NST == 3 && L > 1024, "register twiddles serve the three-stage lengths 2048 and 4096");
    constexpr int NT = B * L / 16;
    constexpr int st = PL::NST - 1;
    constexpr int R = PL::radix(st);
    constexpr int p = PL::p(st);
    constexpr int PER = 16 / R;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        int b, i;
        bf_map<B, L / R>((int)threadIdx.x + q * NT, false, b, i);
        const int k = i % p;
        const double2 w1 = ((const double2 *)d.tw_d)[k * (L / (p * R))];
        double2 wj = w1;
#pragma unroll
        for (int j = 1; j < R; ++j) {
            rt.w[q * (R - 1) + j - 1] = wj;
            if (j + 1 < R) wj = cmul(wj, w1);
        }
    }
}

// All stages of B lines (one radix-16 first-stage butterfly per thread, its
// inputs in v); each line ends in lds in natural order.  Consecutive threads
// take consecutive elements of a line in every stage.  Ends with a barrier.
// RT: the stage twiddles come from rt (fast_regtw_init) for lengths whose
// W_L table is not in LDS.
template <typename T, int L, int B, int SG, bool RT = false>
__device__ __forceinline__ void fast_fft_lds(const AfPass &d, typename Cx<T>::C (&v)[16], typename Cx<T>::C *lds,
                                             const typename Cx<T>::W *stab, const FastRegTw<L> *rt = nullptr) {
    using PL = P2Plan<L>;
    using FT = FastTables<L>;
    constexpr int NT = B * L / 16;
    constexpr int LD = fast_ld(L, B);
    constexpr T s = T(SG);
    const int t = threadIdx.x;
#pragma clang loop unroll(full)  // (plain `unroll` left four-stage lines rolled, with scratch)
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
                } else if (st == 1) {  // the second of three stages: p R = 256, W_256 in LDS
#pragma unroll
                    for (int j = 1; j < R; ++j) v[q * R + j] = twmul(v[q * R + j], stab[k * j], s, false);
                } else if (RT) {
#pragma unroll
                    for (int j = 1; j < R; ++j) {
                        double2 w = rt->w[q * (R - 1) + j - 1];
                        if (SG > 0) w.y = -w.y;
                        v[q * R + j] = dmul(v[q * R + j], w, false);
                    }
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

}  // namespace af
