// Host-side declarations of the kernel launchers (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "af_pass.hpp"

namespace af {

enum { PREC_F64 = 0, PREC_F32 = 1 };

hipError_t launch_pass(int prec, const AfPass &d, hipStream_t s);
hipError_t launch_pass_grouped(int prec, const AfPass *descs, const int64_t *prefix, int ndesc, int64_t total_tiles,
                               int nthreads, int lds_bytes, bool full, hipStream_t s);
hipError_t launch_real_conv(int prec, const AfPass &d, const void *x, const void *k, void *y, hipStream_t s);
hipError_t launch_real_conv_vjp(int prec, const AfPass &d, const void *x, const void *k, const void *w,
                                void *dx, void *dk, hipStream_t s);

// Twiddle tables W_N^(i*mul), i < count, forward sign, in the table format of prec.
hipError_t gen_twiddles(int prec, void *out, int64_t N, int64_t count, int64_t mul, hipStream_t s);
hipError_t gen_twiddles_d(void *out, int64_t N, int64_t count, int64_t mul, hipStream_t s);
hipError_t gen_chirp_d(void *out, int64_t n, int64_t count, int sign, int64_t M, int conj, hipStream_t s);
hipError_t gen_rader_kernel_d(void *out, const int32_t *powg, int64_t p, int64_t count, int sign, hipStream_t s);
hipError_t convert_to_w(int prec, void *out, const void *in, int64_t count, double scale, int64_t N1, int64_t N2,
                        hipStream_t s);
// the same as plain complex values of the precision (convolution spectra)
hipError_t convert_to_plain(int prec, void *out, const void *in, int64_t count, double scale, int64_t N1, int64_t N2,
                            hipStream_t s);

// Real packing in global memory (fft_kernels.hip)
enum { PACK_UNTANGLE = 0, PACK_ENTANGLE = 1, PACK_ENTANGLE_ADJ = 2, PACK_PRODUCT = 3 };
hipError_t launch_pack(int prec, int op, const void *a, const void *a_im, const void *b, void *o_re, void *o_im,
                       const void *aux, int64_t L, int64_t rows, int plain, int arg, hipStream_t s);

// Two-for-one real packing, elementwise over rows of n (fft_kernels.hip):
// untangle = true: (o1, o2) = untangleSpectra p; false: o1 = entangleSpectra (p, q).
// Each complex array is planar (re, im) or interleaved (im == nullptr); with
// h > 1 the rows form h x n matrices (the 2-D forms, untangling against
// reverse2d).
hipError_t launch_two_pack(int prec, bool untangle, const void *p_re, const void *p_im, const void *q_re,
                           const void *q_im, void *o1_re, void *o1_im, void *o2_re, void *o2_im, int64_t n,
                           int64_t rows, hipStream_t s, int64_t h = 1);

// out[b][c][r] = in[b][r][c] for nb matrices of R x C complex elements (planar
// or interleaved endpoints; out must not share memory with in).
hipError_t launch_transpose(int prec, const void *in_re, const void *in_im, void *out_re, void *out_im, int64_t R,
                            int64_t C, int64_t nb, hipStream_t s);

// Allow every kernel the full 160 KiB of LDS (called once).
void init_kernel_attributes();
void init_karatsuba_attributes();
void init_fast_attributes();
// Power-of-two passes with compile-time stages (fast_pass.hip); done = false when not eligible.
hipError_t launch_fast(int prec, const AfPass &d, hipStream_t s, bool &done);
// the descriptor meets the power-of-two kernel's conditions (its (L, B) may
// still lack an instantiation: see fast_tile_lines)
bool fast_pass_eligible(int prec, const AfPass &d);
// the largest line count <= B with a power-of-two kernel for lines of length L
// (B itself when L has none)
int fast_tile_lines(int prec, int64_t L, int B);
// Power-of-two passes of convolution plans (fast_ext.hip): chirp / zero-pad
// loads, spectrum product between forward and inverse stages, chirp / length
// limited stores.  done = false when the descriptor has no such kernel.
bool fast_ext_eligible(int prec, const AfPass &d);
hipError_t launch_fast_ext(int prec, const AfPass &d, hipStream_t s, bool &done);
// the same over the tiles of many descriptors with equal L, B, CONV and sign
hipError_t launch_fast_ext_grouped(int prec, int L, int B, bool conv, int sign, const AfPass *descs,
                                   const int64_t *prefix, int ndesc, int64_t total_tiles, hipStream_t s);
// lines per tile of the middle convolution pass for lines of length L (0: none)
int fast_conv_lines(int prec, int64_t L);
void init_fast_ext_attributes();
// fused real convolution (vjp = false: out0 = conv(in0, in1)) or its adjoint
// (vjp = true: cotangent in0, out0 = dk from in1 = x, out1 = dx from in2 = k)
// for power-of-two half lengths; done = false when the shape has no kernel
hipError_t launch_fast_rconv(int prec, const AfPass &d, const void *in0, const void *in1, const void *in2, void *out0,
                             void *out1, bool vjp, hipStream_t s, bool &done);

// Decomposition plans (decomp.hip): dst[r m + i] = src[r n + map[i]];
// W_n^(sign j2 k1) on rows of b x a interleaved elements; x[r L + i] *= spec[i];
// Rader's output X[0] = sum x, X[pw[i]] = x[0] + c[i] (one row per workgroup).
hipError_t launch_dc_gather(int prec, const void *src_re, const void *src_im, int64_t n, void *dst_re, void *dst_im,
                            int64_t m, const int32_t *map, int64_t rows, hipStream_t s);
hipError_t launch_dc_twiddle(int prec, void *x, int64_t a, int64_t b, int64_t n, int sign, int64_t rows,
                             hipStream_t s);
hipError_t launch_dc_spec_mul(int prec, void *x, const void *spec, int64_t L, int64_t rows, hipStream_t s);
hipError_t launch_dc_rader_out(int prec, const void *x_re, const void *x_im, const void *c, const int32_t *pw,
                               int64_t p, int64_t rows, void *out_re, void *out_im, hipStream_t s);

// Karatsuba (karatsuba.hip): the pairs whose values keep every partial sum
// exact in FP64 on the matrix cores, the rest in 64-bit integers; dev_taken
// holds one int32 per pair (device scratch, written by the first kernel)
hipError_t launch_karatsuba(const int64_t *x, const int64_t *y, int64_t *z, const int64_t *dev_meta, int64_t npairs,
                            int max_P, int cyclic, int32_t *dev_taken, hipStream_t s);
hipError_t launch_gauss_split(const int64_t *xr, const int64_t *xi, const int64_t *yr, const int64_t *yi, int64_t *x3,
                              int64_t *y3, int64_t count, hipStream_t s);
hipError_t launch_gauss_join(const int64_t *p3, int64_t *zr, int64_t *zi, int64_t count, hipStream_t s);
// Global-memory top levels for pairs whose padded length exceeds the LDS kernel's:
// a split task reads x, y (length L) and writes the three child operand pairs
// (low, high, low + high; h = ceil(L/2) words each, consecutive) to cx, cy; a
// join task combines the three child products (2h - 1 words each, consecutive
// at zc) into z (2L - 1 words, or L when cyclic).
struct KarSplitTask {
    const int64_t *x, *y;
    int64_t L;
    int64_t *cx, *cy;
};
struct KarJoinTask {
    const int64_t *zc;
    int64_t L;
    int64_t *z;
    int64_t cyclic;
};
hipError_t launch_kar_split(const KarSplitTask *dev_tasks, int64_t ntasks, hipStream_t s);
hipError_t launch_kar_join(const KarJoinTask *dev_tasks, int64_t ntasks, hipStream_t s);
// the largest padded pair length the LDS kernel takes
constexpr int kKaratsubaMaxP = 4608;

}  // namespace af
