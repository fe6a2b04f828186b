// Power-of-two passes of convolution plans with compile-time stages.
//
// The three passes of a four-step convolution (Bluestein: Planned.hs:704-737
// with the chirp of Private.hs:428-436; the convolution of Planned.hs:816-821)
// carry work around the transform that the plain fast kernel (fast_pass.hip)
// does not: the first multiplies its inputs by the chirp and zero-pads them
// past the signal length, the middle one multiplies the spectrum of each line
// by the kernel's spectrum and transforms back, the last multiplies its
// outputs by the chirp and writes only the signal's length.  k_fast_ext does
// all of it on the fast kernel's compile-time radix-16 stages:
//
//  * loads: element (l, e) at row position pos = l in_ls + e in_es reads as
//    zero from pos >= in_len on and is multiplied by in_mul[pos];
//  * CONV: forward stages into LDS (natural order), each first-stage input of
//    the inverse taken from LDS times spec[row spec_bs + l spec_ls + e], the
//    inverse stages then run as in the plain kernel;
//  * stores: four-step twiddle and scale as in the plain kernel, then
//    out_mul[o] and only positions o < out_len.
//
// k_fast_ext_grouped runs the tiles of many such descriptors (one per row of
// an uneven batch, all with the same L, B, sign and CONV) in one launch: the
// workgroup finds its descriptor by a binary search of the tile prefix sums.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fast_pass.hpp"
#include "kernels.hpp"

namespace af {

template <typename T, int L, int B, int SG, bool CONV>
__device__ __forceinline__ void ext_tile(const AfPass &d, int64_t tile, typename Cx<T>::C *lds,
                                         const typename Cx<T>::W *stab) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    const FastRow r = fast_row<T, B>(d, tile);
    C v[16];
    fast_load_ext<T, L, B>(d, r, v);
    if constexpr (CONV) {
        constexpr int NB = L / 16;
        constexpr int LD = fast_ld(L, B);
        fast_fft_lds<T, L, B, SG>(d, v, lds, stab);
        // the lines of a convolution pass are rows (elements fastest), so the
        // inverse's first stage takes element (b, i + j NB) in v[j]
        int b, i;
        bf_map<B, NB>(threadIdx.x, false, b, i);
        const int64_t row = tile * B / d.nlines;
        const int64_t sb = row * d.spec_bs + (int64_t)(r.line0 + (uint32_t)b) * d.spec_ls + i;
        if (d.flags & AFP_SPEC_PLAIN) {
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = cmul(lds[b * LD + fpad(i + j * NB)], gld<C>(d.spec, sb + j * NB));
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j)
                v[j] = vmul(lds[b * LD + fpad(i + j * NB)], gld<W>(d.spec, sb + j * NB), false);
        }
        lds_barrier();
        fast_stages<T, L, B, -SG, true>(d, r, v, lds, stab);
    } else {
        fast_stages<T, L, B, SG, true>(d, r, v, lds, stab);
    }
}

template <typename T, int L>
__device__ __forceinline__ void ext_stage_table(const AfPass &d, typename Cx<T>::W *stab) {
    // read before use: the first barrier of the stages orders it
    fast_tables_fill<T, L>(d, stab);
}

// Waves per SIMD that the LDS of one workgroup allows (at most 8): the
// register budget is held to it, so that registers never lower the occupancy
// below what LDS permits.
template <typename T, int L, int B>
constexpr int ext_waves_per_simd() {
    constexpr int lds = fast_ld(L, B) * B * (int)sizeof(typename Cx<T>::C) + FastTables<L>::STAGE_ENTRIES * 16;
    constexpr int wg = (160 * 1024) / lds;
    constexpr int w = wg * (B * L / 16 / 64) / 4;
    return w < 1 ? 1 : w > 8 ? 8 : w;
}

template <typename T, int L, int B, int SG, bool CONV>
__global__ void __launch_bounds__(B * L / 16) __attribute__((amdgpu_waves_per_eu(ext_waves_per_simd<T, L, B>())))
k_fast_ext(const AfPass d) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    constexpr int LD = fast_ld(L, B);
    extern __shared__ __align__(16) unsigned char smem_ext[];
    C *lds = (C *)smem_ext;
    W *stab = (W *)(smem_ext + (size_t)LD * B * sizeof(C));
    ext_stage_table<T, L>(d, stab);
    ext_tile<T, L, B, SG, CONV>(d, xcd_tile(blockIdx.x, gridDim.x), lds, stab);
}

template <typename T, int L, int B, int SG, bool CONV>
__global__ void __launch_bounds__(B * L / 16) __attribute__((amdgpu_waves_per_eu(ext_waves_per_simd<T, L, B>())))
k_fast_ext_grouped(const AfPass *__restrict__ descs, const int64_t *__restrict__ prefix, const int ndesc) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    constexpr int LD = fast_ld(L, B);
    extern __shared__ __align__(16) unsigned char smem_ext[];
    C *lds = (C *)smem_ext;
    W *stab = (W *)(smem_ext + (size_t)LD * B * sizeof(C));
    const int64_t g = blockIdx.x;
    int lo = 0, hi = ndesc - 1;
    while (lo < hi) {  // the last descriptor whose first tile is <= g
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    const AfPass &d = descs[lo];
    ext_stage_table<T, L>(d, stab);
    ext_tile<T, L, B, SG, CONV>(d, g - prefix[lo], lds, stab);
}

// X(type, L, B, CONV): the column passes (first and last) take the column tile
// lines of the four-step planner, the middle convolution pass its row tiles.
#define AF_EXT_SHAPES(X)                                                                               \
    X(float, 256, 16, false) X(float, 512, 16, false) X(float, 1024, 16, false)                        \
    X(float, 2048, 8, false) X(float, 4096, 4, false) X(float, 8192, 1, false)                         \
    X(float, 16384, 1, false)                                                                          \
    X(float, 256, 16, true) X(float, 512, 8, true) X(float, 1024, 4, true) X(float, 2048, 2, true)     \
    X(float, 4096, 1, true)                                                                            \
    X(double, 256, 8, false) X(double, 512, 8, false) X(double, 1024, 8, false)                        \
    X(double, 2048, 4, false) X(double, 4096, 2, false) X(double, 8192, 1, false)                      \
    X(double, 256, 8, true) X(double, 512, 4, true) X(double, 1024, 2, true) X(double, 2048, 1, true)  \
    X(double, 4096, 1, true)

static bool ext_has_shape(int prec, int L, int B, bool conv) {
#define AF_HAS_EXT(T, L_, B_, C_) \
    if ((prec == PREC_F64) == (sizeof(T) == 8) && L == L_ && B == B_ && conv == C_) return true;
    AF_EXT_SHAPES(AF_HAS_EXT)
#undef AF_HAS_EXT
    return false;
}

int fast_conv_lines(int prec, int64_t L) {
    if (L < 256 || L > 4096 || (L & (L - 1))) return 0;
    const int B = (int)((prec == PREC_F64 ? 2048 : 4096) / L);
    return ext_has_shape(prec, (int)L, B, true) ? B : 0;
}

static const int32_t kExtFlags = AFP_POSTTW | AFP_SCALE | AFP_CONV | AFP_SPEC_PLAIN;

bool fast_ext_eligible(int prec, const AfPass &d) {
    if (d.flags & ~kExtFlags) return false;
    if (d.in_map || d.out_map || d.gt_in || d.gt_out || d.grp_mask) return false;
    const bool conv = (d.flags & AFP_CONV) != 0;
    if (conv && (d.in_lines_fast || !d.spec)) return false;
    if (!ext_has_shape(prec, d.L, d.B, conv)) return false;
    if (!d.tw_d) return false;
    if (!d.tw) return false;  // the LDS stage table (W_L, or W_256 from it) at every length
    if ((d.flags & AFP_POSTTW) && (!d.ft_coarse_d || !d.ft_fine_d || (d.ftN & (d.ftN - 1)))) return false;
    if (d.nlines % d.B) return false;
    if (d.ntiles >= ((int64_t)1 << 31)) return false;
    const int64_t eb = prec == PREC_F64 ? 8 : 4;
    const int64_t span_i = ((int64_t)(d.B - 1) * d.in_ls + (int64_t)(d.L - 1) * d.in_es) * (d.in_im ? eb : 2 * eb) + eb;
    const int64_t span_o = ((int64_t)(d.B - 1) * d.out_ls + (int64_t)(d.L - 1) * d.out_es) * (d.out_im ? eb : 2 * eb) + eb;
    if (span_i >= ((int64_t)1 << 32) || span_o >= ((int64_t)1 << 32)) return false;
    // row positions (for in_len / in_mul, out_len / out_mul) stay 32-bit
    const int64_t pin = (d.nlines - 1) * d.in_ls + (int64_t)(d.L - 1) * d.in_es;
    const int64_t pout = (d.nlines - 1) * d.out_ls + (int64_t)(d.L - 1) * d.out_es;
    if (pin >= ((int64_t)1 << 32) || pout >= ((int64_t)1 << 32)) return false;
    if ((d.flags & AFP_POSTTW) && (d.ftN >> d.ftShift) * 16 + 16 >= ((int64_t)1 << 32)) return false;
    const int64_t lds = (int64_t)fast_ld(d.L, d.B) * d.B * (prec == PREC_F64 ? 16 : 8) + (d.L <= 1024 ? d.L : 256) * 16;
    return lds <= 160 * 1024;
}

template <typename T, int L, int B>
__host__ __device__ constexpr size_t ext_lds_bytes() {
    return (size_t)fast_ld(L, B) * B * sizeof(typename Cx<T>::C) + (size_t)FastTables<L>::STAGE_ENTRIES * 16;
}

template <typename T, int L, int B, bool CONV>
static bool try_ext(const AfPass &d, hipStream_t s, hipError_t &err) {
    if (d.L != L || d.B != B || ((d.flags & AFP_CONV) != 0) != CONV) return false;
    const unsigned g = (unsigned)d.ntiles;
    const size_t lds = ext_lds_bytes<T, L, B>();
    if (d.sign < 0) hipLaunchKernelGGL((k_fast_ext<T, L, B, -1, CONV>), dim3(g), dim3(B * L / 16), lds, s, d);
    else hipLaunchKernelGGL((k_fast_ext<T, L, B, 1, CONV>), dim3(g), dim3(B * L / 16), lds, s, d);
    err = hipGetLastError();
    return true;
}

hipError_t launch_fast_ext(int prec, const AfPass &d, hipStream_t s, bool &done) {
    hipError_t err = hipSuccess;
    done = false;
    if (!fast_ext_eligible(prec, d)) return hipSuccess;
#define AF_TRY_EXT(T, L_, B_, C_) \
    if (!done && (prec == PREC_F64) == (sizeof(T) == 8)) done = try_ext<T, L_, B_, C_>(d, s, err);
    AF_EXT_SHAPES(AF_TRY_EXT)
#undef AF_TRY_EXT
    return err;
}

template <typename T, int L, int B, bool CONV>
static bool try_ext_grouped(int Lq, int Bq, bool conv, int sign, const AfPass *descs, const int64_t *prefix, int ndesc,
                            int64_t total, hipStream_t s, hipError_t &err) {
    if (Lq != L || Bq != B || conv != CONV) return false;
    const size_t lds = ext_lds_bytes<T, L, B>();
    const dim3 g((unsigned)total), b(B * L / 16);
    if (sign < 0) hipLaunchKernelGGL((k_fast_ext_grouped<T, L, B, -1, CONV>), g, b, lds, s, descs, prefix, ndesc);
    else hipLaunchKernelGGL((k_fast_ext_grouped<T, L, B, 1, CONV>), g, b, lds, s, descs, prefix, ndesc);
    err = hipGetLastError();
    return true;
}

hipError_t launch_fast_ext_grouped(int prec, int L, int B, bool conv, int sign, const AfPass *descs,
                                   const int64_t *prefix, int ndesc, int64_t total_tiles, hipStream_t s) {
    if (total_tiles <= 0 || ndesc <= 0) return hipSuccess;
    if (total_tiles >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    hipError_t err = hipSuccess;
    bool done = false;
#define AF_TRY_EXTG(T, L_, B_, C_)                                                                         \
    if (!done && (prec == PREC_F64) == (sizeof(T) == 8))                                                   \
        done = try_ext_grouped<T, L_, B_, C_>(L, B, conv, sign, descs, prefix, ndesc, total_tiles, s, err);
    AF_EXT_SHAPES(AF_TRY_EXTG)
#undef AF_TRY_EXTG
    return done ? err : hipErrorInvalidValue;
}

template <typename T, int L, int B, bool CONV>
static void set_attr_ext() {
    (void)hipFuncSetAttribute((const void *)k_fast_ext<T, L, B, -1, CONV>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    (void)hipFuncSetAttribute((const void *)k_fast_ext<T, L, B, 1, CONV>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    (void)hipFuncSetAttribute((const void *)k_fast_ext_grouped<T, L, B, -1, CONV>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute((const void *)k_fast_ext_grouped<T, L, B, 1, CONV>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

void init_fast_ext_attributes() {
#define AF_ATTR_EXT(T, L_, B_, C_) set_attr_ext<T, L_, B_, C_>();
    AF_EXT_SHAPES(AF_ATTR_EXT)
#undef AF_ATTR_EXT
}

}  // namespace af
