/*
 * accelerate-fourier on MI355X (gfx950): the C ABI.
 *
 * This is the drop-in boundary for the transform path of the Haskell library
 * accelerate-fourier.  Each entry point replaces one pure `Acc -> Acc`
 * function of the reference; a Haskell binding calls it through `foreignAcc`
 * and a `foreign import ccall` (see INTEGRATION.md for the binding source).
 *
 * Conventions
 *  - Plain C types only; device buffers are raw pointers the caller owns.
 *  - Complex arrays are planar by default, as Accelerate (< 1.3) stores
 *    `Array sh (Complex a)`: one array of real parts and one of imaginary
 *    parts.  Passing `im == NULL` selects the interleaved layout (re, im pairs)
 *    for the same entry point.
 *  - Arrays are row-major with the transform axis innermost and contiguous;
 *    the leading `sh` extents are flattened into `batch`.
 *  - `stream` is a hipStream_t (NULL = the default stream); calls are
 *    asynchronous with respect to the host.
 *  - Every precondition that the reference documents but does not check
 *    returns AF_INVALID_ARG instead of computing garbage.  Nothing throws.
 *  - Plans are immutable after creation: concurrent af_execute calls on one
 *    plan from several host threads or streams are allowed.
 */
#ifndef ACCELERATE_FOURIER_H
#define ACCELERATE_FOURIER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    AF_SUCCESS = 0,
    AF_INVALID_ARG = 1,   /* a documented precondition of the reference is violated */
    AF_UNSUPPORTED = 2,   /* a valid request this library cannot serve */
    AF_NO_DEVICE = 3,     /* no gfx950 device is visible */
    AF_HIP_ERROR = 4,     /* a HIP runtime call failed */
    AF_OUT_OF_MEMORY = 5,
    AF_INTERNAL = 6
} af_status;

typedef enum { AF_F64 = 0, AF_F32 = 1 } af_precision;

/* Which reference function a plan stands for. */
typedef enum {
    AF_C2C = 0,          /* Planned.transform sign n                         (Planned.hs:103-106) */
    AF_R2C_FULL = 1,     /* Real.toSpectrum (Planned.transform sign (n/2))   (Real.hs:49-70)      */
    AF_C2R_FULL = 2,     /* Real.fromSpectrum (Planned.transform sign (n/2)) (Real.hs:92-107)     */
    AF_CONV_CYCLIC = 3,  /* Planned.convolveCyclic n                         (Planned.hs:779-792) */
    AF_REAL_CONV = 4     /* fromSpectrum t' (toSpectrum t x * toSpectrum t k), t/t' = forward/inverse
                            transform (n/2): the real FFT convolution (Real.hs:49-107 composed) */
} af_kind;

/* Algorithm selection, mirroring the reference's entry points. */
typedef enum {
    AF_ALG_AUTO = 0,       /* Planned.transform: decompose, Bluestein when Rader nests (Planned.hs:215-220) */
    AF_ALG_DECOMPOSE = 1,  /* Planned.transformDecompose: never Bluestein; Rader nests (Planned.hs:114-119) */
    AF_ALG_CHIRP2 = 2,     /* Planned.transformChirp2: Bluestein, power-of-two pad (Planned.hs:704-708) */
    AF_ALG_CHIRP235 = 3    /* Planned.transformChirp235: Bluestein, 5-smooth pad   (Planned.hs:716-720) */
} af_algorithm;

typedef struct {
    int64_t length;      /* n: the inner extent (for real kinds: the real length, even) */
    int32_t sign;        /* -1 = Sign.forward, +1 = Sign.inverse (Sign.hs:52-54); for the
                            real kinds: the sign of the half-length sub-transform */
    int32_t kind;        /* af_kind */
    int32_t precision;   /* af_precision */
    int32_t algorithm;   /* af_algorithm */
    uint32_t flags;      /* AF_FLAG_* */
} af_plan_desc;

/* FP32 stage twiddles and vectors as compensated hi+lo float pairs (default on). */
#define AF_FLAG_PLAIN_TWIDDLES 0x1u   /* use hi parts only (for error studies) */
/* Single-pass transforms whose length is a product of three or more pairwise
 * coprime butterfly radices (2, 4, 8 or 16 and distinct primes 3 ... 13, e.g.
 * 30, 60, 15015) run in prime-factor (Good-Thomas, Planned.hs:584-617) order by
 * default: skew-grid input map, Chinese-remainder output map, no twiddles.
 * This flag keeps the twiddled Cooley-Tukey order (Planned.hs:543-557)
 * instead; the results agree to rounding. */
#define AF_FLAG_COOLEY_TUKEY 0x2u

typedef struct af_plan_s *af_plan;
typedef struct af_batch_s *af_batch;

const char *af_status_string(af_status s);
const char *af_version(void);
/* The failed HIP call behind the last AF_HIP_ERROR / AF_OUT_OF_MEMORY returned on
 * the calling thread ("" if none): the call, its source line and HIP's error
 * string.  The reference reports errors as Accelerate runtime exceptions; this
 * is the ABI's text for them. */
const char *af_last_error(void);

/*
 * Plan creation: host-side planning (Planned.plan, Planned.hs:201-220) plus
 * device tables generated on the GPU in FP64 (twiddles, chirps, Rader and
 * Bluestein kernel spectra).  Replaces the reference's `cache sign len`
 * (Planned.hs:322-328), which is what partial application of
 * `transform sign len` shares between calls.
 */
af_status af_plan_create(const af_plan_desc *desc, af_plan *plan);
af_status af_plan_destroy(af_plan plan);
/* Text of the device plan (pass structure); host-only. */
af_status af_plan_describe(const af_plan_desc *desc, char *buf, size_t buflen);
/* Bytes of device scratch one call with `batch` rows needs (0 for most plans). */
af_status af_plan_workspace_size(af_plan plan, int64_t batch, size_t *bytes);

/*
 * Complex transform of `batch` rows of length n (AF_C2C plans).
 * Replaces: Planned.transform sign n :: Transform (sh:.Int) (Complex a)
 *           (Planned.hs:103-106).  Unnormalised, as the reference.
 * in_im == NULL (and then out_im == NULL) selects interleaved complex.
 * In-place (in == out) is allowed.  workspace may be NULL (stream-ordered
 * allocation inside the call) or a buffer of af_plan_workspace_size bytes.
 */
af_status af_execute(af_plan plan, int64_t batch,
                     const void *in_re, const void *in_im,
                     void *out_re, void *out_im,
                     void *workspace, void *stream);

/*
 * af_execute without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * scratch a plan needs (af_plan_workspace_size) comes from one fixed buffer
 * of 64 MiB per GPU, set up when the library first runs on that GPU and held
 * until the process ends; a call whose scratch exceeds it returns
 * AF_UNSUPPORTED (af_last_error names the sizes) and runs nothing.  Calls
 * that overlap in time on one GPU share the buffer and race: keep them on
 * one stream, or order them.  The caller passes inputs and outputs as for
 * af_execute; the call makes no hipMalloc/hipFree.
 */
af_status af_execute_noalloc(af_plan plan, int64_t batch, const void *in_re, const void *in_im,
                             void *out_re, void *out_im, void *stream);

/*
 * The launches of one af_execute call, one at a time, for profiling and for
 * callers that interleave other work between them.  af_plan_num_passes gives
 * the count for `batch` rows; pass i of af_execute_pass runs launch i of the
 * same sequence (workspace required when af_plan_workspace_size is not 0).
 * No reference counterpart: the reference has no notion of launches.
 */
af_status af_plan_num_passes(af_plan plan, int64_t batch, int32_t *npasses);
af_status af_execute_pass(af_plan plan, int32_t index, int64_t batch,
                          const void *in_re, const void *in_im,
                          void *out_re, void *out_im, void *workspace, void *stream);

/*
 * Autotuner.  No reference counterpart: the reference's plans have fixed
 * shapes.  Times, for `batch` rows in the given layout (planar = 1: re/im
 * arrays, 0: interleaved) on `stream`, every tile line count 1..64 that some
 * kernel can run for each launch of the transform, and for launches on the
 * mixed-radix kernel also three orders of their radices (as planned, reversed,
 * odd radices first); keeps what beats the incumbent (timed again after each
 * candidate) by more than 3% in the plan, where
 * af_execute, af_vjp_execute and af_execute_pass with the same batch and
 * layout use it, and in the process's wisdom.  Allocates and frees its own
 * zero-filled buffers and synchronizes the stream.  `report` (may be NULL)
 * receives one line per launch: chosen and planned shape with their times.
 */
af_status af_plan_tune(af_plan plan, int64_t batch, int32_t planar, void *stream, char *report, size_t report_len);

/*
 * Wisdom: the autotuner's choices as text, one line per (precision,
 * algorithm, length, batch, layout).  Export writes it to buf and the size it
 * needs, terminator included, to *needed (buf NULL with buflen 0 only
 * queries).  Import adds the lines to the process's wisdom, which plans
 * without a tuning of their own for a (batch, layout) use.
 */
af_status af_wisdom_export(char *buf, size_t buflen, size_t *needed);
af_status af_wisdom_import(const char *text);

/*
 * Two-for-one real packing (AF_C2C plans of length n, any n).
 * af_two_to_spectrum replaces Real.twoToSpectrum (Planned.transform sign n)
 * (Real.hs:129-135): the spectra a and b of the real rows x and y from one
 * complex transform of x + i y, untangled as untangleSpectra (168-187).
 * af_two_from_spectrum replaces Real.twoFromSpectrum (190-196): x and y are the
 * real and imaginary parts of the transform of the entangled a + i b
 * (entangleSpectra, 208-236).  x, y: batch rows of n reals each (the
 * reference's array of pairs as its two component arrays); a, b: planar
 * (re, im) or interleaved (im NULL).  Scratch is stream-ordered.
 */
af_status af_two_to_spectrum(af_plan plan, int64_t batch, const void *x, const void *y,
                             void *a_re, void *a_im, void *b_re, void *b_im, void *stream);
af_status af_two_from_spectrum(af_plan plan, int64_t batch, const void *a_re, const void *a_im,
                               const void *b_re, const void *b_im, void *x, void *y, void *stream);
/*
 * The packing steps alone, elementwise over batch rows of n: Real.untangleSpectra
 * (Real.hs:168-187; z must not share memory with a or b) and
 * Real.entangleSpectra (Real.hs:208-236).
 */
af_status af_untangle_spectra(int32_t precision, int64_t batch, int64_t n, const void *z_re, const void *z_im,
                              void *a_re, void *a_im, void *b_re, void *b_im, void *stream);
af_status af_entangle_spectra(int32_t precision, int64_t batch, int64_t n, const void *c_re, const void *c_im,
                              const void *d_re, const void *d_im, void *z_re, void *z_im, void *stream);

/*
 * Multi-dimensional transforms over the last two / three axes, each axis with
 * its own C2C plan (same precision), batch leading slabs.
 * af_execute_2d replaces Preprocessed.transform2d (SubTransformPair t0 t1)
 * (Preprocessed.hs:141-148; Adhoc.transform2d, Adhoc.hs:365-371, with one
 * sub-transform for both axes): t0 = `inner` along the last axis (length W),
 * t1 = `outer` along the one before (length H); arrays (batch, H, W).
 * af_execute_3d replaces Preprocessed.transform3d (SubTransformTriple t0 t1 t2)
 * (Preprocessed.hs:150-161; Adhoc.hs:373-380): arrays (batch, K, M, N) with
 * p0, p1, p2 of lengths N, M, K.  The reference transposes between axes
 * (LinAlg.transpose, cycleDim3); here axes other than the last run as strided
 * passes of plain single-pass plans, else through tiled transposes.  In place
 * allowed; scratch is stream-ordered.
 */
af_status af_execute_2d(af_plan inner, af_plan outer, int64_t batch, const void *in_re, const void *in_im,
                        void *out_re, void *out_im, void *stream);
af_status af_execute_3d(af_plan p0, af_plan p1, af_plan p2, int64_t batch, const void *in_re, const void *in_im,
                        void *out_re, void *out_im, void *stream);
/*
 * out[b][c][r] = in[b][r][c] for batch matrices of rows x cols complex values:
 * LinAlg.transpose over the last two axes; with rows = K M and cols = N it is
 * Fourier.cycleDim3 (Private.hs:414-425).  out must not share memory with in.
 */
af_status af_transpose(int32_t precision, int64_t batch, int64_t rows, int64_t cols, const void *in_re,
                       const void *in_im, void *out_re, void *out_im, void *stream);
/*
 * 2-D two-for-one packing: Real.twoToSpectrum2d / twoFromSpectrum2d
 * (Real.hs:138-145, 199-206) with the 2-D transform above, and
 * untangleSpectra2d (Real.hs:175-180) against Cyclic.reverse2d (Cyclic.hs:30-39).
 */
af_status af_two_to_spectrum_2d(af_plan inner, af_plan outer, int64_t batch, const void *x, const void *y,
                                void *a_re, void *a_im, void *b_re, void *b_im, void *stream);
af_status af_two_from_spectrum_2d(af_plan inner, af_plan outer, int64_t batch, const void *a_re, const void *a_im,
                                  const void *b_re, const void *b_im, void *x, void *y, void *stream);
af_status af_untangle_spectra_2d(int32_t precision, int64_t batch, int64_t height, int64_t width, const void *z_re,
                                 const void *z_im, void *a_re, void *a_im, void *b_re, void *b_im, void *stream);

/*
 * Real signal to full spectrum (AF_R2C_FULL plans).
 * Replaces: Real.toSpectrum (Planned.transform sign (n/2)) (Real.hs:49-70).
 * x: batch x n reals; out: batch x n complex (the full, not halved, spectrum).
 */
af_status af_to_spectrum(af_plan plan, int64_t batch, const void *x,
                         void *out_re, void *out_im, void *workspace, void *stream);

/*
 * Full self-adjoint spectrum to real signal (AF_C2R_FULL plans), unnormalised
 * (n times the signal).  Replaces: Real.fromSpectrum (Planned.transform sign
 * (n/2)) (Real.hs:92-107).
 */
af_status af_from_spectrum(af_plan plan, int64_t batch, const void *in_re, const void *in_im,
                           void *y, void *workspace, void *stream);

/*
 * Cyclic convolution of complex rows (AF_CONV_CYCLIC plans), normalised.
 * Replaces: Planned.convolveCyclic n (Planned.hs:779-792).  n >= 1.
 */
af_status af_convolve_cyclic(af_plan plan, int64_t batch,
                             const void *x_re, const void *x_im,
                             const void *y_re, const void *y_im,
                             void *out_re, void *out_im, void *workspace, void *stream);

/*
 * Real FFT convolution (AF_REAL_CONV plans):
 *   y = fromSpectrum (transform inverse (n/2)) (toSpectrum (transform forward (n/2)) x
 *                                               * toSpectrum (transform forward (n/2)) k)
 * = n * (x cyclic-convolved with k), real, batch x n.  Composes Real.hs:49-107
 * with the pointwise product of convolveSpectrumCyclicCache (Planned.hs:816-821).
 */
af_status af_real_convolve(af_plan plan, int64_t batch, const void *x, const void *k,
                           void *y, void *stream);

/* ---- Vector-Jacobian products (no reference counterpart: the adjoints of the
 * functions above, for callers that differentiate through them) ---- */

/* Adjoint of af_execute: the conjugate-transposed DFT, i.e. the transform with
 * the opposite sign applied to the cotangent (gradient convention dL/dconj(z)). */
af_status af_vjp_execute(af_plan plan, int64_t batch,
                         const void *g_re, const void *g_im,
                         void *out_re, void *out_im, void *workspace, void *stream);
/* Adjoint of af_to_spectrum: real gradient from a complex cotangent
 * (the transpose of the real-FFT untangling). */
af_status af_vjp_to_spectrum(af_plan plan, int64_t batch, const void *g_re, const void *g_im,
                             void *dx, void *workspace, void *stream);
/* Adjoint of af_from_spectrum: complex cotangent of the spectrum from a real
 * cotangent (the transpose of the entangling). */
af_status af_vjp_from_spectrum(af_plan plan, int64_t batch, const void *g,
                               void *out_re, void *out_im, void *workspace, void *stream);
/* Adjoint of af_real_convolve with cotangent w: dx = n (w correlated with k),
 * dk = n (w correlated with x), both in one fused launch. */
af_status af_vjp_real_convolve(af_plan plan, int64_t batch, const void *x, const void *k,
                               const void *w, void *dx, void *dk, void *stream);

/* ---- Karatsuba (Int64, bit-exact) ---- */

/*
 * Linear convolution of `batch` pairs of equal-length Int64 rows, wrap-around
 * arithmetic.  x, y: batch x len; z: batch x (2 len - 1) (nothing for len = 0).
 * Replaces: Convolution.Adhoc.karatsuba (Convolution/Adhoc.hs:36-64).
 */
af_status af_karatsuba_i64(int64_t batch, int64_t len, const int64_t *x, const int64_t *y,
                           int64_t *z, void *stream);
/*
 * Ragged batch of pairs: pair p has length offsets[p+1]-offsets[p] >= 1 and
 * its inputs start at offsets[p]; its 2L-1 outputs start at 2*offsets[p]-p.
 * `offsets` is a host array of npairs+1 entries.  Same function as above.
 */
af_status af_karatsuba_ragged_i64(int64_t npairs, const int64_t *offsets,
                                  const int64_t *x, const int64_t *y, int64_t *z, void *stream);
/* Convolution.Adhoc.cyclic karatsuba (Convolution/Adhoc.hs:104-111): batch x len outputs. */
af_status af_karatsuba_cyclic_i64(int64_t batch, int64_t len, const int64_t *x, const int64_t *y,
                                  int64_t *z, void *stream);
/*
 * Complex Int64 convolution, planar operands (re and im arrays of batch x len
 * each), in the Gauss form of three real Karatsuba products:
 * re = xr*yr - xi*yi, im = (xr + xi)*(yr + yi) - xr*yr - xi*yi, wrap-around
 * arithmetic.  cyclic = 0: linear, batch x (2 len - 1) outputs;
 * cyclic = 1: Convolution.cyclic of it, batch x len outputs.
 * Replaces: Convolution.Adhoc.complex karatsuba and
 * Convolution.Adhoc.complex (cyclic karatsuba) (Convolution/Adhoc.hs:114-133).
 */
af_status af_karatsuba_complex_i64(int64_t batch, int64_t len, int32_t cyclic,
                                   const int64_t *xr, const int64_t *xi, const int64_t *yr,
                                   const int64_t *yi, int64_t *zr, int64_t *zi, void *stream);

/* ---- Uneven batches: many lengths in one graph launch ---- */

/*
 * A batch of `count` complex transforms of lengths lengths[i] >= 0, all with
 * one sign and precision, packed back to back (row i starts at the prefix sum
 * of the lengths).  Equivalent to Planned.transform sign lengths[i] on each
 * row.  The work of all rows runs as one captured hipGraph, the k-th passes
 * of all rows grouped by kernel.  af_batch_tune may instead cut the rows whose
 * plans need intermediates (four-step, Rader and Bluestein cores) into chunks
 * of a bounded intermediate size, all passes of a chunk before the next, so
 * that intermediates can stay in the memory-side cache between passes and the
 * scratch memory is the chunk size instead of the sum over all rows.
 */
af_status af_batch_create(int64_t count, const int64_t *lengths, int32_t sign,
                          int32_t precision, af_batch *batch);
af_status af_batch_destroy(af_batch batch);
/*
 * Schedule of a batch's graph (no reference counterpart): rows with
 * intermediates in chunks of at most chunk_bytes of them (0: no chunking; a
 * row needing more is a chunk of its own), and lanes = 2 to run the rows
 * without intermediates on a branch of the graph beside the others (their
 * small launches then overlap the large ones) or 1 for one sequence; lanes = 3
 * (with chunk_bytes = 0 only) also moves the rows with intermediates that use
 * the generic kernel (mixed radix, Rader) to a third branch beside the
 * power-of-two Bluestein rows.  Recaptures the graph.  The default is (0, 2).
 */
af_status af_batch_set_schedule(af_batch batch, int64_t chunk_bytes, int32_t lanes);
/*
 * Autotuner of a batch (no reference counterpart): times the batch's graph
 * for several schedules (one lane unchunked; three lanes unchunked; two lanes unchunked and with
 * chunks of 32 MiB ... 384 MiB of intermediates) on zero-filled interleaved
 * buffers it allocates itself, keeps the fastest (it must beat the incumbent
 * by 3%) and recaptures the graph with it.  Then it times every grouped
 * launch of the generic kernel alone with its stages as planned, reversed and
 * odd radices first, and keeps the fastest order per launch (again by 3%);
 * the order lives in the launch's descriptors, so the graph stays as
 * captured.  af_batch_set_schedule resets the orders to the planned ones.
 * Synchronizes `stream`.  `report` (may be NULL) receives one line per
 * schedule, the choice, and one line per timed launch.
 */
af_status af_batch_tune(af_batch batch, void *stream, char *report, size_t report_len);
/*
 * One call runs every row.  A batch owns one set of launch descriptors and
 * intermediates, so calls on one batch are serialised: concurrent calls from
 * several threads are allowed, and a call on another stream is ordered after
 * the previous call's work (an event it waits on); a call with new buffers
 * waits on the host until the previous call's work is done before it rewrites
 * the descriptors.  Distinct batches run concurrently.
 */
af_status af_batch_execute(af_batch batch, const void *in_re, const void *in_im,
                           void *out_re, void *out_im, void *stream);
af_status af_batch_describe(af_batch batch, char *buf, size_t buflen);

/* ---- Host-side index maps, for bit-exact checks against the reference ---- */

/* Rader permutations for prime p (Permutation.hs:45-80): perm[i] = g^i mod p - 1,
 * rev = Permutation.reverse perm, inv = Permutation.inverse perm; each p-1 long. */
af_status af_rader_permutations(int64_t p, int64_t *perm, int64_t *rev, int64_t *inv);
/* Good-Thomas maps of the device kernels for n*m with gcd(n,m) = 1:
 * in_map[k*n + j] = (n*k + m*j) mod nm (permuteSkewGrid m n, Planned.hs:597-606),
 * out_map[q] = (q mod n)*m + (q mod m) (permuteSkewGridInv, Planned.hs:608-617). */
af_status af_good_thomas_maps(int64_t n, int64_t m, int64_t *in_map, int64_t *out_map);
/* NumberTheory.multiplicativeGenerator / ceiling5Smooth / ceilingPowerOfTwo */
int64_t af_multiplicative_generator(int64_t p);
int64_t af_ceiling_5smooth(int64_t n);
int64_t af_ceiling_power_of_two(int64_t n);

#ifdef __cplusplus
}
#endif
#endif
