/*
 * CPU oracle for the accelerate-fourier transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference's algorithms (Haskell, xc-jp/accelerate-fourier).  It is linked or
 * loaded only by the test-suite, by __graft_entry__.smoke() as the checker, and
 * by bench.py's cpu_baseline leg.  The product path (the HIP library) never
 * calls it.
 *
 * Every function cites the reference file:line it restates.  Paths are relative
 * to the reference root; `src/.../` stands for src/Data/Array/Accelerate/.
 *
 * Arrays follow Accelerate's own representation of `Array (sh:.n) (Complex a)`:
 * two planar component arrays (real parts, imaginary parts), rows contiguous,
 * the transform axis innermost.  `rows` is the product of the leading extents.
 */
#ifndef FOURIER_ORACLE_H
#define FOURIER_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Algorithm selectors mirroring the reference's entry points
 * (src/.../Fourier/Planned.hs:103-129, 704-737). */
enum { ORC_AUTO = 0, ORC_DECOMPOSE = 1, ORC_CHIRP2 = 2, ORC_CHIRP235 = 3 };

/* ---- number theory (src/.../NumberTheory.hs) ---- */
int64_t orc_fermat_factors(int64_t n, int64_t *pairs, int64_t max_pairs);
int64_t orc_multiplicative_generator(int64_t p);
int64_t orc_ceiling_power_of_two(int64_t n);
int64_t orc_ceiling_5smooth(int64_t n);

/* ---- permutations (src/.../Permutation.hs) ---- */
void orc_perm_multiplicative(int64_t p, int64_t *out);       /* length p-1 */
void orc_perm_inverse(int64_t len, const int64_t *perm, int64_t *out);
void orc_perm_reverse(int64_t len, const int64_t *perm, int64_t *out);

/* ---- Good-Thomas index maps (src/.../Fourier/Planned.hs:597-617) ---- */
/* skew grid of shape (m, n): out[k*n + j] = (n*k + m*j) mod (n*m) */
void orc_skew_grid(int64_t m, int64_t n, int64_t *out);
/* inverse map from a (n, m) matrix: out[q] = (q mod n)*m + (q mod m) */
void orc_skew_grid_inv(int64_t n, int64_t m, int64_t *out);

/* ---- planner (src/.../Fourier/Planned.hs:201-293, 731-737) ---- */
int orc_plan_string(int64_t n, int algorithm, char *buf, int64_t buflen);
int orc_plan_count_primes(int64_t n);

/* ---- complex transforms ---- */
int orc_transform_f64(int algorithm, int sign, int64_t n, int64_t rows,
                      const double *re, const double *im, double *ore, double *oim);
int orc_transform_f32(int algorithm, int sign, int64_t n, int64_t rows,
                      const float *re, const float *im, float *ore, float *oim);
/* brute-force DFT in long double with exact index reduction (closed form) */
void orc_dft_bruteforce(int sign, int64_t n, int64_t rows,
                        const double *re, const double *im, double *ore, double *oim);

/* ---- real-signal transforms (src/.../Fourier/Real.hs:43-121) ---- */
int orc_to_spectrum_f64(int sub_sign, int64_t n, int64_t rows, const double *x,
                        double *ore, double *oim);
int orc_from_spectrum_f64(int sub_sign, int64_t n, int64_t rows, const double *re,
                          const double *im, double *y);
int orc_to_spectrum_f32(int sub_sign, int64_t n, int64_t rows, const float *x,
                        float *ore, float *oim);
int orc_from_spectrum_f32(int sub_sign, int64_t n, int64_t rows, const float *re,
                          const float *im, float *y);

/* ---- two-for-one real packing (src/.../Fourier/Real.hs:124-236) ---- */
int orc_untangle_spectra_f64(int64_t n, int64_t rows, const double *zre, const double *zim, double *are,
                             double *aim, double *bre, double *bim);
int orc_entangle_spectra_f64(int64_t n, int64_t rows, const double *cre, const double *cim, const double *dre,
                             const double *dim, double *zre, double *zim);
int orc_untangle_spectra_f32(int64_t n, int64_t rows, const float *zre, const float *zim, float *are,
                             float *aim, float *bre, float *bim);
int orc_entangle_spectra_f32(int64_t n, int64_t rows, const float *cre, const float *cim, const float *dre,
                             const float *dim, float *zre, float *zim);

/* ---- FFT cyclic convolution (src/.../Fourier/Planned.hs:779-821) ---- */
int orc_convolve_cyclic_f64(int64_t n, int64_t rows, const double *xre, const double *xim,
                            const double *yre, const double *yim, double *ore, double *oim);

/* ---- Karatsuba (src/.../Convolution/Adhoc.hs, Preprocessed.hs) ---- */
void orc_karatsuba_i64(int64_t len, int64_t rows, const int64_t *x, const int64_t *y, int64_t *z);
void orc_karatsuba_rec_i64(int64_t len, int64_t rows, const int64_t *x, const int64_t *y, int64_t *z);
void orc_karatsuba_f64(int64_t len, int64_t rows, const double *x, const double *y, double *z);
void orc_schoolbook_i64(int64_t len, int64_t rows, const int64_t *x, const int64_t *y, int64_t *z);
void orc_karatsuba_ragged_i64(int64_t npairs, const int64_t *offsets, const int64_t *x,
                              const int64_t *y, int64_t *z);
void orc_cyclic_karatsuba_i64(int64_t len, int64_t rows, const int64_t *x, const int64_t *y, int64_t *z);
/* Gauss three-multiplication complex wrapper (Adhoc.hs:114-133) around the linear
 * (cyclic=0) or cyclic (cyclic=1) Karatsuba on doubles */
void orc_complex_karatsuba_f64(int cyclic, int64_t len, int64_t rows,
                               const double *xre, const double *xim,
                               const double *yre, const double *yim,
                               double *ore, double *oim);

/* the same wrapper around the Int64 Karatsuba, wrap-around arithmetic */
void orc_complex_karatsuba_i64(int cyclic, int64_t len, int64_t rows,
                               const int64_t *xre, const int64_t *xim,
                               const int64_t *yre, const int64_t *yim,
                               int64_t *ore, int64_t *oim);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
