// Pass descriptor shared by the host planner and the device kernels.
//
// One "pass" is one launch of the generic LDS transform kernel over a set of
// tiles.  A tile holds B lines of length L in LDS; each line is transformed by
// a mixed-radix Stockham sequence of register butterflies.  The descriptor says
// where a line's elements come from (strides, gathers, zero padding, a
// pre-multiplied vector), what happens between and after the stage sequences
// (pointwise spectrum product and an inverse stage sequence for convolutions,
// four-step twiddles, real-FFT packing) and where results go.
#pragma once
#include <stdint.h>

#define AF_MAXST 16

// Generic pass kernels: elements of a tile each thread holds in registers per
// stage (16 complex floats or 8 complex doubles), so a tile of B lines of
// length L takes ceil(B L / EPT) threads, at most 1024.
__host__ __device__ constexpr int pass_ept(bool f64) { return f64 ? 8 : 16; }

enum AfPassFlags : int32_t {
    AFP_CONV = 1 << 0,          // stages(sign) -> * spec -> stages(-sign)
    AFP_POSTTW = 1 << 1,        // multiply element e of line l by W_N^{ftSign * (l*e mod N)}
    AFP_RADER = 1 << 2,         // Rader single pass: x0 from the row, DC from the spectrum
    AFP_DC_CAPTURE = 1 << 3,    // write element (line 0, e 0) of the first spectrum to dc_out
    AFP_RADER_OUT = 1 << 4,     // add x0 (from x0 source) to every output, write X[0] = x0 + dc_in
    AFP_REAL_UNTANGLE = 1 << 5, // store: toSpectrum untangling of the half-length spectrum
    AFP_REAL_ENTANGLE = 1 << 6, // load: fromSpectrum entangling of the full spectrum
    AFP_SCALE = 1 << 7,         // multiply outputs by `scale`
    AFP_PLAIN_TW = 1 << 8,      // FP32: use the hi parts of twiddle tables only
    AFP_IN_CONJ_SYM = 1 << 9,   // load: Hermitian part (g[k] + conj g[-k]) / 2 of the input
    AFP_UNTANGLE_ADJ = 1 << 10, // load: transpose of the untangling (adjoint of toSpectrum's packing)
    AFP_ENTANGLE_ADJ = 1 << 11, // store: transpose of the entangling (adjoint of fromSpectrum's packing)
    AFP_SPEC_PLAIN = 1 << 12,   // the convolution spectrum holds plain complex values of the precision
    AFP_X0_CAPTURE = 1 << 13,   // copy row element 0 of the input to x0_out (Rader, before any store)
};

// Flags only the real-signal packings and Rader passes use.  The generic pass
// kernel has a lean instantiation without their code (small enough to inline
// without spills) and a full one for passes that set any of them.
constexpr int32_t kRarePassFlags = AFP_RADER | AFP_X0_CAPTURE | AFP_RADER_OUT | AFP_REAL_UNTANGLE | AFP_REAL_ENTANGLE |
                                   AFP_UNTANGLE_ADJ | AFP_ENTANGLE_ADJ;

struct AfPass {
    // geometry
    int32_t L;          // line (transform) length
    int32_t nst;        // number of Stockham stages
    int32_t B;          // lines per tile
    int32_t ld;         // LDS elements per line (L plus padding)
    int32_t radix[AF_MAXST];
    int32_t pst[AF_MAXST];     // product of the radices before each stage
    int64_t nlines;     // lines per row
    int64_t nrows;      // rows (batch items)
    int64_t ntiles;     // ceil(nrows * nlines / B)
    int32_t in_lines_fast;     // load phase: consecutive threads walk lines (1) or elements (0)
    int32_t out_lines_fast;    // store phase: idem
    int32_t sign;       // sign of the (first) stage sequence
    int32_t flags;      // AfPassFlags
    int32_t padshift;   // LDS index padding: idx + (idx >> padshift) (0 = none)
    int32_t nthreads;   // threads of one tile
    int32_t lds_bytes;  // dynamic LDS of one tile
    int32_t _pad0;
    // input: element (line l, e) of row r sits at logical position
    //   pos = l * in_ls + e * in_es within the row, address r * in_bs + pos
    const void *in_re;
    const void *in_im;  // nullptr: interleaved complex at in_re
    int64_t in_es, in_ls, in_bs;
    int64_t in_len;               // positions >= in_len read as zero
    const int32_t *in_map;        // gather: address r * in_bs + in_map[pos]
    const void *in_mul;           // multiply by in_mul[pos] (plain complex values of the precision)
    // stage twiddles W_L^m, m < L, forward sign, twiddle-table format
    const void *tw;
    // convolution spectrum, element (l, e) of row r at spec[r*spec_bs + l*spec_ls + e]
    const void *spec;
    int64_t spec_ls, spec_bs;
    void *dc_out;                 // complex (interleaved) per row
    // four-step twiddle W_N^(ftSign * a), a = l*e mod N, from coarse[a >> ftShift] * fine[a & mask]
    int64_t ftN;
    int32_t ftShift, ftSign;
    const void *ft_coarse;
    const void *ft_fine;
    // real packing twiddles: W_{2L}^k (forward sign), k < L
    const void *aux_tw;
    // output: position o = l * out_ls + e * out_es, address r * out_bs + o
    void *out_re;
    void *out_im;
    int64_t out_es, out_ls, out_bs;
    int64_t out_len;              // positions >= out_len are not written
    const int32_t *out_map;       // scatter: address r * out_bs + out_map[o]
    const void *out_mul;          // multiply by out_mul[o]
    // Rader: x0 source (row element 0) and DC from an earlier pass
    const void *x0_re;
    const void *x0_im;
    int64_t x0_bs;
    const void *dc_in;
    void *x0_out;                 // complex (interleaved) per row
    // FP64 copies of the stage and four-step twiddle tables (double2), read by
    // the power-of-two kernel once per butterfly
    const void *tw_d;
    const void *ft_coarse_d;
    const void *ft_fine_d;
    // Prime-factor (Good-Thomas) order of a generic pass (Planned.hs:584-617,
    // for any number of pairwise coprime factors): with every radix of the
    // stage list a power of a different prime, each stage is a factor of its
    // own and no stage has twiddles (grp_mask: bit s set for every stage
    // s > 0); element e of a line enters LDS at Stockham position gt_in[e]
    // (the skew-grid input map) and output e leaves from position gt_out[e]
    // (the Chinese-remainder output map).  nullptr / 0: Cooley-Tukey order.
    const int32_t *gt_in;
    const int32_t *gt_out;
    int32_t grp_mask;
    int32_t _pad1;
    double scale;
};
