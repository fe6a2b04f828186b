// Specialised pass kernel for power-of-two line lengths (256 ... 16384).
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

#include "fast_pass.hpp"
#include "kernels.hpp"

namespace af {

// One launch of a power-of-two pass: one workgroup per tile.
template <typename T, int L, int B, int SG>
__global__ void __launch_bounds__(B * L / 16)
k_fast(const AfPass d) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    constexpr int LD = fast_ld(L, B);
    extern __shared__ __align__(16) unsigned char smem_fast[];
    C *lds = (C *)smem_fast;
    W *stab = (W *)(smem_fast + (size_t)LD * B * sizeof(C));
    // stage table into LDS; the first barrier of the stages orders it before any use
    fast_tables_fill<T, L>(d, stab);
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
template <typename T, int L, int B, bool RT = false>
__device__ __forceinline__ void fast_inverse_store(const AfPass &d, typename Cx<T>::C *buf,
                                                   const typename Cx<T>::W *stab, T *dst, int64_t row0,
                                                   const FastRegTw<L> *rt = nullptr) {
    using C = typename Cx<T>::C;
    constexpr int NB = L / 16;
    constexpr int LD = fast_ld(L, B);
    int b, i;
    bf_map<B, NB>(threadIdx.x, false, b, i);
    C v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = buf[b * LD + fpad(i + j * NB)];
    lds_barrier();
    fast_fft_lds<T, L, B, 1, RT>(d, v, buf, stab, rt);
    const Rsrc r = mk_rsrc(dst + row0 * 2 * L, (uint32_t)(B * L * sizeof(C)));
    const uint32_t o0 = (uint32_t)((b * L + i) * sizeof(C));
#pragma unroll
    for (int j = 0; j < 16; ++j) bst<C>(r, o0, (uint32_t)(j * NB * sizeof(C)), buf[b * LD + fpad(i + j * NB)]);
    lds_barrier();
}

// The product pairs (e, L - e) of the B lines, H = L/2 + 1 per line: pair
// idx = tid + q NT (q < kRconvPairs) belongs to one thread throughout.
template <int L, int B> struct RconvPairs {
    static constexpr int NT = B * L / 16;
    static constexpr int H = L / 2 + 1;
    static constexpr int NP = (B * H + NT - 1) / NT;  // 9 for every shape
};

// The half-length spectrum values a thread's pairs need, from LDS into registers.
template <typename T, int L, int B>
__device__ __forceinline__ void fast_pairs_read(const typename Cx<T>::C *A, typename Cx<T>::C (&p0)[RconvPairs<L, B>::NP],
                                                typename Cx<T>::C (&p1)[RconvPairs<L, B>::NP]) {
    using RP = RconvPairs<L, B>;
    constexpr int LD = fast_ld(L, B);
#pragma unroll
    for (int q = 0; q < RP::NP; ++q) {
        const int idx = threadIdx.x + q * RP::NT;
        if (idx < B * RP::H) {
            const int b = idx / RP::H, e = idx - b * RP::H;
            const int er = (L - e) & (L - 1);
            p0[q] = A[b * LD + fpad(e)];
            p1[q] = A[b * LD + fpad(er)];
        }
    }
}

// Per pair: the product of the two full spectra (by the conjugate of the
// second for the adjoint's correlations), entangled back into Bb as the
// half-length spectrum of the inverse, in the closed form of rconv_pair (the
// first spectrum from registers, the second from Bb).  Each pair has one owner
// thread.
template <typename T, int L, int B>
__device__ __forceinline__ void fast_spec_product(const typename Cx<T>::C (&p0)[RconvPairs<L, B>::NP],
                                                  const typename Cx<T>::C (&p1)[RconvPairs<L, B>::NP],
                                                  typename Cx<T>::C *Bb, const typename Cx<T>::W *aux, bool conj_b) {
    using C = typename Cx<T>::C;
    using RP = RconvPairs<L, B>;
    constexpr int LD = fast_ld(L, B);
#pragma unroll
    for (int q = 0; q < RP::NP; ++q) {
        const int idx = threadIdx.x + q * RP::NT;
        if (idx >= B * RP::H) continue;
        const int b = idx / RP::H, e = idx - b * RP::H;
        const int er = (L - e) & (L - 1);
        const int ie = b * LD + fpad(e), ir = b * LD + fpad(er);
        C z0, z1;
        rconv_pair<T>(p0[q], p1[q], Bb[ie], Bb[ir], aux[e], false, conj_b, z0, z1);
        Bb[ie] = z0;
        if (er != e) Bb[ir] = z1;
    }
}

// The same with the first spectrum in a second LDS buffer (FP64, whose pair
// registers would take the whole register file); dst may be A or Bb.
template <typename T, int L, int B>
__device__ __forceinline__ void fast_spec_product2(const typename Cx<T>::C *A, const typename Cx<T>::C *Bb,
                                                   typename Cx<T>::C *dst, const typename Cx<T>::W *aux, bool conj_b) {
    using C = typename Cx<T>::C;
    constexpr int NT = B * L / 16;
    constexpr int LD = fast_ld(L, B);
    constexpr int H = L / 2 + 1;
    for (int idx = threadIdx.x; idx < B * H; idx += NT) {
        const int b = idx / H, e = idx - b * H;
        const int er = (L - e) & (L - 1);
        const int ie = b * LD + fpad(e), ir = b * LD + fpad(er);
        C z0, z1;
        rconv_pair<T>(A[ie], A[ir], Bb[ie], Bb[ir], aux[e], false, conj_b, z0, z1);
        dst[ie] = z0;
        if (er != e) dst[ir] = z1;
    }
}

// FP32: one line buffer; the first signal's spectrum is kept in the registers of the
// threads that own its pairs, so the buffer serves every transform in turn and
// twice as many workgroups fit per CU as with two buffers.
// FP32 at lengths 2048 and 4096 takes the last-stage register twiddles
// (FastRegTw); FP64, already at one workgroup per CU there, keeps the table
// loads.
template <typename T, int L> __host__ __device__ constexpr bool rconv_regtw() {
    return !FastTables<L>::STAGE_LDS && sizeof(T) == 4;
}

template <typename T, int L, int B>
__host__ __device__ constexpr size_t fast_rconv_lds_bytes() {
    return (sizeof(T) == 4 ? 1 : 2) * (size_t)fast_ld(L, B) * B * sizeof(typename Cx<T>::C) +
           (size_t)FastTables<L>::STAGE_ENTRIES * 16;
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
    using RP = RconvPairs<L, B>;
    constexpr int LD = fast_ld(L, B);
    constexpr bool ONE = sizeof(T) == 4;  // one line buffer (FP32) or two (FP64)
    extern __shared__ __align__(16) unsigned char smem_rc[];
    C *A = (C *)smem_rc;
    C *Bb = ONE ? A : A + LD * B;
    W *stab = (W *)(Bb + LD * B);
    fast_tables_fill<T, L>(d, stab);
    const W *aux = (const W *)d.aux_tw;
    const int64_t row0 = (int64_t)blockIdx.x * B;
    constexpr bool RT = rconv_regtw<T, L>();
    FastRegTw<L> rt;
    if constexpr (RT) fast_regtw_init<L, B>(d, rt);
    C v[16];
    if constexpr (ONE) {
        C p0[RP::NP], p1[RP::NP];
        fast_rows_load<T, L, B>(in0, row0, v);
        fast_fft_lds<T, L, B, -1, RT>(d, v, A, stab, &rt);
        fast_pairs_read<T, L, B>(A, p0, p1);
        // (loading in1 before the first transform, to overlap it, measured slower)
        fast_rows_load<T, L, B>(in1, row0, v);
        lds_barrier();  // every pair is read before the second transform writes A
        fast_fft_lds<T, L, B, -1, RT>(d, v, A, stab, &rt);
        fast_spec_product<T, L, B>(p0, p1, A, aux, VJP);
        lds_barrier();
        fast_inverse_store<T, L, B, RT>(d, A, stab, out0, row0, &rt);
        if (VJP) {
            fast_rows_load<T, L, B>(in2, row0, v);
            fast_fft_lds<T, L, B, -1, RT>(d, v, A, stab, &rt);
            fast_spec_product<T, L, B>(p0, p1, A, aux, true);
            lds_barrier();
            fast_inverse_store<T, L, B, RT>(d, A, stab, out1, row0, &rt);
        }
    } else {
        fast_rows_load<T, L, B>(in0, row0, v);
        fast_fft_lds<T, L, B, -1, RT>(d, v, A, stab, &rt);
        fast_rows_load<T, L, B>(in1, row0, v);
        fast_fft_lds<T, L, B, -1, RT>(d, v, Bb, stab, &rt);
        fast_spec_product2<T, L, B>(A, Bb, Bb, aux, VJP);
        lds_barrier();
        fast_inverse_store<T, L, B, RT>(d, Bb, stab, out0, row0, &rt);
        if (VJP) {
            fast_rows_load<T, L, B>(in2, row0, v);
            fast_fft_lds<T, L, B, -1, RT>(d, v, Bb, stab, &rt);
            fast_spec_product2<T, L, B>(A, Bb, Bb, aux, true);
            lds_barrier();
            fast_inverse_store<T, L, B, RT>(d, Bb, stab, out1, row0, &rt);
        }
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
    if (d.in_map || d.out_map || d.in_mul || d.out_mul || d.gt_in || d.gt_out || d.grp_mask) return false;
    if (d.in_len != INT64_MAX || d.out_len != INT64_MAX) return false;
    if (d.L < 256 || d.L > 16384 || (d.L & (d.L - 1))) return false;
    // the W_L table of the working precision feeds the LDS stage table at every length
    if (!d.tw || !d.tw_d || ((d.flags & AFP_POSTTW) && (!d.ft_coarse_d || !d.ft_fine_d))) return false;
    if ((d.flags & AFP_POSTTW) && (d.ftN & (d.ftN - 1))) return false;
    if (d.nlines % d.B) return false;
    if (d.ntiles >= ((int64_t)1 << 31)) return false;
    // byte offsets inside a tile and into the FP64 four-step tables stay 32-bit
    const int64_t eb = prec == PREC_F64 ? 8 : 4;
    const int64_t span_i = ((int64_t)(d.B - 1) * d.in_ls + (int64_t)(d.L - 1) * d.in_es) * (d.in_im ? eb : 2 * eb) + eb;
    const int64_t span_o = ((int64_t)(d.B - 1) * d.out_ls + (int64_t)(d.L - 1) * d.out_es) * (d.out_im ? eb : 2 * eb) + eb;
    if (span_i >= ((int64_t)1 << 32) || span_o >= ((int64_t)1 << 32)) return false;
    if ((d.flags & AFP_POSTTW) && (d.ftN >> d.ftShift) * 16 + 16 >= ((int64_t)1 << 32)) return false;
    const int64_t lds = (int64_t)fast_ld(d.L, d.B) * d.B * (prec == PREC_F64 ? 16 : 8) + (d.L <= 1024 ? d.L : 256) * 16;
    if (lds > 160 * 1024) return false;
    return (int64_t)d.B * d.L / 16 <= 1024;
}

// Tile shapes with a power-of-two kernel: lengths 256 ... 16384 and line
// counts B from one wavefront (B * L / 16 >= 64 threads) up to what one tile
// holds in LDS.  X(type, L, B).
#define AF_FAST_SHAPES(X)                                                                          \
    X(float, 256, 4) X(float, 256, 8) X(float, 256, 16) X(float, 256, 32)                          \
    X(float, 512, 2) X(float, 512, 4) X(float, 512, 8) X(float, 512, 16) X(float, 512, 32)        \
    X(float, 1024, 1) X(float, 1024, 2) X(float, 1024, 4) X(float, 1024, 8) X(float, 1024, 16)     \
    X(float, 2048, 1) X(float, 2048, 2) X(float, 2048, 4) X(float, 2048, 8)                        \
    X(float, 4096, 1) X(float, 4096, 2) X(float, 4096, 4) X(float, 8192, 1) X(float, 8192, 2)      \
    X(float, 16384, 1)                                                                             \
    X(double, 256, 4) X(double, 256, 8) X(double, 256, 16)                                         \
    X(double, 512, 2) X(double, 512, 4) X(double, 512, 8) X(double, 512, 16)                       \
    X(double, 1024, 1) X(double, 1024, 2) X(double, 1024, 4) X(double, 1024, 8)                    \
    X(double, 2048, 1) X(double, 2048, 2) X(double, 2048, 4)                                       \
    X(double, 4096, 1) X(double, 4096, 2) X(double, 8192, 1)

static bool fast_has_shape(int prec, int L, int B) {
#define AF_HAS(T, L_, B_) if ((prec == PREC_F64) == (sizeof(T) == 8) && L == L_ && B == B_) return true;
    AF_FAST_SHAPES(AF_HAS)
#undef AF_HAS
    return false;
}

int fast_tile_lines(int prec, int64_t L, int B) {
    if (L < 256 || L > 16384 || (L & (L - 1))) return B;
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
