/*
 * CPU oracle for the accelerate-fourier transform path -- TEST INFRASTRUCTURE ONLY.
 * Loaded by tests/, __graft_entry__.smoke() (as the checker) and bench.py's
 * cpu_baseline leg; never by the product path.  See fourier_oracle.h.
 *
 * This file holds the precision-independent parts: number theory, permutations,
 * Good-Thomas maps, the planner and the Karatsuba convolutions.  The transforms
 * are instantiated for double and float from fourier_oracle_body.h.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/build_oracle.sh).
 * -ffp-contract=off keeps every complex product as the two-rounding formula
 * that the reference's `Complex` instance spells out, with no fused multiply-add.
 */
#include "fourier_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Number theory: src/.../NumberTheory.hs                                    */
/* ------------------------------------------------------------------------ */

static int64_t isqrt64(int64_t v) {
    int64_t r = (int64_t)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

static int64_t gcd64(int64_t a, int64_t b) {
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a < 0 ? -a : a;
}

/* fermatFactors n (NumberTheory.hs:15-21): every factorisation n = (b-a)(b+a)
 * of an odd n, found where n + a^2 is a perfect square b^2, listed in the order
 * the reference's merge produces them (increasing a, i.e. the pair closest to
 * sqrt n first).  Returns the number of pairs; writes up to max_pairs pairs. */
int64_t orc_fermat_factors(int64_t n, int64_t *pairs, int64_t max_pairs) {
    int64_t count = 0;
    for (int64_t a = 0; a <= (n - 1) / 2; ++a) {
        int64_t v = n + a * a;
        int64_t b = isqrt64(v);
        if (b * b == v) {
            if (count < max_pairs) {
                pairs[2 * count] = b - a;
                pairs[2 * count + 1] = b + a;
            }
            ++count;
        }
    }
    return count;
}

/* modularPower (NumberTheory.hs:53-63) */
static int64_t modpow(int64_t modu, int64_t expo, int64_t base) {
    __int128 result = 1, b = base % modu;
    while (expo > 0) {
        if (expo & 1) result = (result * b) % modu;
        b = (b * b) % modu;
        expo >>= 1;
    }
    return (int64_t)result;
}

/* uniquePrimeFactors (NumberTheory.hs:65-81) by trial division */
static int unique_prime_factors(int64_t n, int64_t *out) {
    int cnt = 0;
    if (n % 2 == 0) {
        out[cnt++] = 2;
        while (n % 2 == 0) n /= 2;
    }
    for (int64_t p = 3; p * p <= n; p += 2) {
        if (n % p == 0) {
            out[cnt++] = p;
            while (n % p == 0) n /= p;
        }
    }
    if (n > 1) out[cnt++] = n;
    return cnt;
}

/* multiplicativeGenerator p (NumberTheory.hs:32-47): the smallest n in 1..p-1
 * coprime to p with n^(p-1) = 1 and n^((p-1)/q) != 1 for each prime q | p-1. */
int64_t orc_multiplicative_generator(int64_t p) {
    int64_t order = p - 1, facs[64];
    int nf = unique_prime_factors(order, facs);
    for (int64_t n = 1; n <= p - 1; ++n) {
        if (gcd64(n, p) != 1) continue;
        if (modpow(p, order, n) != 1) continue;
        int ok = 1;
        for (int i = 0; i < nf; ++i)
            if (modpow(p, order / facs[i], n) == 1) { ok = 0; break; }
        if (ok) return n;
    }
    return -1;
}

/* ceilingPowerOfTwo (NumberTheory.hs:94-101) */
int64_t orc_ceiling_power_of_two(int64_t n) {
    if (n <= 1) return 1;
    int64_t m = 1;
    while (m < n) m <<= 1;
    return m;
}

/* ceilingSmooths a b n m (NumberTheory.hs:118-124): successively trade one
 * factor a for one factor b (m*b, then divide out a's while staying >= n). */
static int ceiling_smooths(int64_t a, int64_t b, int64_t n, int64_t m, int64_t *out, int cap) {
    int cnt = 0;
    for (;;) {
        if (cnt < cap) out[cnt++] = m;
        if (m % a != 0) break;
        int64_t k = m * b;
        for (;;) {
            int64_t q = k / a, r = k % a;
            if (r == 0 && q >= n) k = q; else break;
        }
        m = k;
    }
    return cnt;
}

/* ceiling5Smooth (NumberTheory.hs:108-111) */
int64_t orc_ceiling_5smooth(int64_t n) {
    int64_t l3[128], l5[128];
    int n3 = ceiling_smooths(2, 3, n, orc_ceiling_power_of_two(n), l3, 128);
    int64_t best = -1;
    for (int i = 0; i < n3; ++i) {
        int n5 = ceiling_smooths(2, 5, n, l3[i], l5, 128);
        for (int j = 0; j < n5; ++j)
            if (best < 0 || l5[j] < best) best = l5[j];
    }
    return best;
}

/* ------------------------------------------------------------------------ */
/* Permutations: src/.../Permutation.hs                                      */
/* ------------------------------------------------------------------------ */

/* multiplicative (Permutation.hs:45-54): perm[i] = g^i mod p - 1 */
void orc_perm_multiplicative(int64_t p, int64_t *out) {
    int64_t g = orc_multiplicative_generator(p), x = 1;
    for (int64_t i = 0; i < p - 1; ++i) {
        out[i] = x - 1;
        x = (int64_t)(((__int128)g * x) % p);
    }
}

/* inverse (Permutation.hs:57-73): scatter i to position perm[i] */
void orc_perm_inverse(int64_t len, const int64_t *perm, int64_t *out) {
    for (int64_t i = 0; i < len; ++i) out[i] = 0;
    for (int64_t i = 0; i < len; ++i) out[perm[i]] += i;
}

/* reverse (Permutation.hs:75-80): out[i] = perm[(-i) mod len] */
void orc_perm_reverse(int64_t len, const int64_t *perm, int64_t *out) {
    for (int64_t i = 0; i < len; ++i) out[i] = perm[(len - i) % len];
}

/* permuteSkewGrid m n (Planned.hs:597-606) */
void orc_skew_grid(int64_t m, int64_t n, int64_t *out) {
    int64_t nm = n * m;
    for (int64_t k = 0; k < m; ++k)
        for (int64_t j = 0; j < n; ++j) out[k * n + j] = (n * k + m * j) % nm;
}

/* permuteSkewGridInv (Planned.hs:608-617) applied to an (n, m) matrix */
void orc_skew_grid_inv(int64_t n, int64_t m, int64_t *out) {
    for (int64_t q = 0; q < n * m; ++q) out[q] = (q % n) * m + (q % m);
}

/* ------------------------------------------------------------------------ */
/* Planner: src/.../Fourier/Planned.hs:183-293, 731-737                      */
/* ------------------------------------------------------------------------ */

enum { PL_IDENTITY, PL_SMALL, PL_SPLITRADIX, PL_PRIME, PL_COPRIME, PL_COMPOSITE, PL_CHIRP };

typedef struct oplan {
    int64_t n;
    int kind;
    struct oplan *a, *b;
} oplan;

static oplan *plan_new(int64_t n, int kind, oplan *a, oplan *b) {
    oplan *p = (oplan *)malloc(sizeof(oplan));
    p->n = n; p->kind = kind; p->a = a; p->b = b;
    return p;
}

static void plan_free(oplan *p) {
    if (!p) return;
    plan_free(p->a);
    plan_free(p->b);
    free(p);
}

/* planDecomposeWithMap (Planned.hs:255-277) together with the memo's initial
 * content smallPlanMap (Planned.hs:241-250).  Sharing of sub-plans does not
 * change results, so the memo is not restated. */
static oplan *plan_decompose(int64_t n) {
    if (n <= 1) return plan_new(n, PL_IDENTITY, NULL, NULL);
    if (n <= 5) return plan_new(n, PL_SMALL, NULL, NULL);
    if (n % 2 == 0) return plan_new(n, PL_SPLITRADIX, plan_decompose(n / 2), NULL);
    /* odd: the first coprime Fermat pair with a > 1, else the first pair */
    int64_t a = -1, b = -1, fa = -1, fb = -1;
    for (int64_t t = 0; t <= (n - 1) / 2; ++t) {
        int64_t v = n + t * t, r = isqrt64(v);
        if (r * r != v) continue;
        int64_t x = r - t, y = r + t;
        if (fa < 0) { fa = x; fb = y; }
        if (x > 1 && gcd64(x, y) == 1) { a = x; b = y; break; }
    }
    if (a > 0) return plan_new(n, PL_COPRIME, plan_decompose(a), plan_decompose(b));
    if (fa == 1) return plan_new(n, PL_PRIME, plan_decompose(n - 1), NULL);
    return plan_new(n, PL_COMPOSITE, plan_decompose(fa), plan_decompose(fb));
}

/* planCountPrimes (Planned.hs:222-232) */
static int plan_count_primes(const oplan *p) {
    switch (p->kind) {
    case PL_IDENTITY: case PL_SMALL: return 0;
    case PL_SPLITRADIX: case PL_CHIRP: return plan_count_primes(p->a);
    case PL_PRIME: return 1 + plan_count_primes(p->a);
    default: {
        int x = plan_count_primes(p->a), y = plan_count_primes(p->b);
        return x > y ? x : y;
    }
    }
}

/* planChirpWithMapUpdate (Planned.hs:731-737) */
static oplan *plan_chirp(int64_t n, int pow2) {
    if (n < 2) return plan_new(n, PL_IDENTITY, NULL, NULL);
    int64_t pad = pow2 ? orc_ceiling_power_of_two(2 * n - 1) : orc_ceiling_5smooth(2 * n - 1);
    return plan_new(n, PL_CHIRP, plan_decompose(pad), NULL);
}

/* plan / planWithMapUpdate (Planned.hs:201-220): Bluestein on ceiling5Smooth
 * when the Rader nesting depth reaches three. */
static oplan *plan_make(int64_t n, int algorithm) {
    switch (algorithm) {
    case ORC_DECOMPOSE: return plan_decompose(n);
    case ORC_CHIRP2: return plan_chirp(n, 1);
    case ORC_CHIRP235: return plan_chirp(n, 0);
    default: {
        oplan *p = plan_decompose(n);
        if (plan_count_primes(p) < 3) return p;
        plan_free(p);
        return plan_chirp(n, 0);
    }
    }
}

int orc_plan_count_primes(int64_t n) {
    oplan *p = plan_decompose(n);
    int c = plan_count_primes(p);
    plan_free(p);
    return c;
}

static void plan_print(const oplan *p, char *buf, int64_t len, int64_t *pos) {
#define EMIT(...) do { int64_t room = len - *pos; int w = snprintf(buf + *pos, room > 0 ? (size_t)room : 0, __VA_ARGS__); *pos += w; } while (0)
    switch (p->kind) {
    case PL_IDENTITY: EMIT("id%lld", (long long)p->n); break;
    case PL_SMALL: EMIT("c%lld", (long long)p->n); break;
    case PL_SPLITRADIX: EMIT("sr%lld(", (long long)p->n); plan_print(p->a, buf, len, pos); EMIT(")"); break;
    case PL_PRIME: EMIT("rader%lld(", (long long)p->n); plan_print(p->a, buf, len, pos); EMIT(")"); break;
    case PL_CHIRP: EMIT("chirp%lld(", (long long)p->n); plan_print(p->a, buf, len, pos); EMIT(")"); break;
    case PL_COPRIME: EMIT("gt%lld(", (long long)p->n); plan_print(p->a, buf, len, pos); EMIT(","); plan_print(p->b, buf, len, pos); EMIT(")"); break;
    case PL_COMPOSITE: EMIT("ct%lld(", (long long)p->n); plan_print(p->a, buf, len, pos); EMIT(","); plan_print(p->b, buf, len, pos); EMIT(")"); break;
    }
#undef EMIT
}

int orc_plan_string(int64_t n, int algorithm, char *buf, int64_t buflen) {
    oplan *p = plan_make(n, algorithm);
    int64_t pos = 0;
    if (buflen > 0) buf[0] = 0;
    plan_print(p, buf, buflen, &pos);
    plan_free(p);
    return pos < buflen ? 0 : -1;
}

/* ------------------------------------------------------------------------ */
/* Karatsuba: src/.../Convolution/Adhoc.hs:36-111, Preprocessed.hs:29-56     */
/* ------------------------------------------------------------------------ */

#define KARATSUBA_BODY(T, NAME)                                                          \
/* karatsuba (Adhoc.hs:36-64): breadth-first.  The first loop splits all rows \
 * (karatsubaReorder, 66-79) until the row length is 1 and records the lengths; \
 * then a pointwise product; then the second loop recombines (karatsubaGo, 81-98). */ \
static void NAME##_one(int64_t len, const T *x, const T *y, T *z) {                      \
    if (len == 0) return;                                                                \
    int64_t lens[70]; int nl = 0;                                                        \
    lens[nl++] = len;                                                                    \
    int64_t c = 1, l = len;                                                              \
    T *xs = (T *)malloc(sizeof(T) * (size_t)len), *ys = (T *)malloc(sizeof(T) * (size_t)len); \
    memcpy(xs, x, sizeof(T) * (size_t)len); memcpy(ys, y, sizeof(T) * (size_t)len);      \
    while (l > 1) {                                                                      \
        int64_t l2 = (l + 1) / 2;                                                        \
        T *nx = (T *)malloc(sizeof(T) * (size_t)(3 * c * l2));                           \
        T *ny = (T *)malloc(sizeof(T) * (size_t)(3 * c * l2));                           \
        for (int64_t r = 0; r < c; ++r)                                                  \
            for (int64_t k = 0; k < l2; ++k) {                                           \
                T xl = xs[r * l + k], yl = ys[r * l + k];                                \
                T xr = (l2 + k < l) ? xs[r * l + l2 + k] : (T)0;                         \
                T yr = (l2 + k < l) ? ys[r * l + l2 + k] : (T)0;                         \
                nx[r * l2 + k] = xl; ny[r * l2 + k] = yl;                                \
                nx[(c + r) * l2 + k] = xl + xr; ny[(c + r) * l2 + k] = yl + yr;          \
                nx[(2 * c + r) * l2 + k] = xr; ny[(2 * c + r) * l2 + k] = yr;            \
            }                                                                            \
        free(xs); free(ys); xs = nx; ys = ny;                                            \
        c *= 3; l = l2;                                                                  \
        lens[nl++] = l;                                                                  \
    }                                                                                    \
    /* pointwise product of the 1-element rows */                                        \
    T *zs = (T *)malloc(sizeof(T) * (size_t)c);                                          \
    for (int64_t r = 0; r < c; ++r) zs[r] = xs[r] * ys[r];                               \
    free(xs); free(ys);                                                                  \
    int64_t zl = 1; /* current row length of zs */                                       \
    /* lens, most recent first: lens[nl-1] = 1 ... lens[0] = len */                      \
    for (int i = nl - 1; i >= 1; --i) {                                                  \
        int64_t xlen = lens[i], zlen = 2 * lens[i - 1] - 1, n3 = c / 3;                  \
        T *nz = (T *)malloc(sizeof(T) * (size_t)(n3 * zlen));                            \
        for (int64_t r = 0; r < n3; ++r) {                                               \
            const T *zlo = zs + r * zl, *zmi = zs + (n3 + r) * zl, *zhi = zs + (2 * n3 + r) * zl; \
            for (int64_t k = 0; k < zlen; ++k) {                                         \
                T a = (k < zl) ? zlo[k] : (T)0;                                          \
                int64_t k1 = k - xlen, k2 = k - 2 * xlen;                                \
                T b = (k1 >= 0 && k1 < zl) ? zmi[k1] - (zlo[k1] + zhi[k1]) : (T)0;     \
                T d = (k2 >= 0 && k2 < zl) ? zhi[k2] : (T)0;                             \
                nz[r * zlen + k] = a + b + d;                                            \
            }                                                                            \
        }                                                                                \
        free(zs); zs = nz; c = n3; zl = zlen;                                            \
    }                                                                                    \
    memcpy(z, zs, sizeof(T) * (size_t)(2 * len - 1));                                    \
    free(zs);                                                                            \
}

/* Signed wrap-around arithmetic in two's complement, as Int64 behaves. */
typedef uint64_t u64w;
KARATSUBA_BODY(u64w, kara_u64)
KARATSUBA_BODY(double, kara_f64)

void orc_karatsuba_i64(int64_t len, int64_t rows, const int64_t *x, const int64_t *y, int64_t *z) {
    int64_t zlen = len > 0 ? 2 * len - 1 : 0;
#pragma omp parallel for schedule(dynamic)
    for (int64_t r = 0; r < rows; ++r)
        kara_u64_one(len, (const u64w *)(x + r * len), (const u64w *)(y + r * len), (u64w *)(z + r * zlen));
}

void orc_karatsuba_f64(int64_t len, int64_t rows, const double *x, const double *y, double *z) {
    int64_t zlen = len > 0 ? 2 * len - 1 : 0;
    for (int64_t r = 0; r < rows; ++r) kara_f64_one(len, x + r * len, y + r * len, z + r * zlen);
}

/* karatsuba len (Preprocessed.hs:29-56): depth-first recursion */
static void kara_rec(int64_t len, const u64w *x, const u64w *y, u64w *z) {
    if (len <= 1) {
        if (len == 1) z[0] = x[0] * y[0];
        return;
    }
    int64_t h = (len + 1) / 2;
    u64w *buf = (u64w *)calloc((size_t)(6 * h + 3 * (2 * h - 1)), sizeof(u64w));
    u64w *xl = buf, *xm = buf + h, *xr = buf + 2 * h, *yl = buf + 3 * h, *ym = buf + 4 * h, *yr = buf + 5 * h;
    u64w *zl = buf + 6 * h, *zm = zl + (2 * h - 1), *zr = zm + (2 * h - 1);
    for (int64_t k = 0; k < h; ++k) {
        xl[k] = x[k]; yl[k] = y[k];
        xr[k] = (h + k < len) ? x[h + k] : 0; yr[k] = (h + k < len) ? y[h + k] : 0;
        xm[k] = xl[k] + xr[k]; ym[k] = yl[k] + yr[k];
    }
    kara_rec(h, xl, yl, zl);
    kara_rec(h, xm, ym, zm);
    kara_rec(h, xr, yr, zr);
    int64_t zlen = 2 * h - 1;
    for (int64_t k = 0; k < 2 * len - 1; ++k) {
        u64w a = (k < zlen) ? zl[k] : 0;
        int64_t k1 = k - h, k2 = k - 2 * h;
        u64w b = (k1 >= 0 && k1 < zlen) ? zm[k1] - (zl[k1] + zr[k1]) : 0;
        u64w d = (k2 >= 0 && k2 < zlen) ? zr[k2] : 0;
        z[k] = a + b + d;
    }
    free(buf);
}

void orc_karatsuba_rec_i64(int64_t len, int64_t rows, const int64_t *x, const int64_t *y, int64_t *z) {
    int64_t zlen = len > 0 ? 2 * len - 1 : 0;
#pragma omp parallel for schedule(dynamic)
    for (int64_t r = 0; r < rows; ++r)
        kara_rec(len, (const u64w *)(x + r * len), (const u64w *)(y + r * len), (u64w *)(z + r * zlen));
}

/* Schoolbook convolution, z[k] = sum_i x[i] y[k-i], in wrap-around Int64. */
void orc_schoolbook_i64(int64_t len, int64_t rows, const int64_t *x, const int64_t *y, int64_t *z) {
    int64_t zlen = len > 0 ? 2 * len - 1 : 0;
#pragma omp parallel for schedule(dynamic)
    for (int64_t r = 0; r < rows; ++r) {
        const u64w *xr = (const u64w *)(x + r * len), *yr = (const u64w *)(y + r * len);
        u64w *zr = (u64w *)(z + r * zlen);
        for (int64_t k = 0; k < zlen; ++k) zr[k] = 0;
        for (int64_t i = 0; i < len; ++i)
            for (int64_t j = 0; j < len; ++j) zr[i + j] += xr[i] * yr[j];
    }
}

/* Ragged batch: pair p has length offsets[p+1]-offsets[p]; inputs are packed
 * at offsets[p], outputs at 2*offsets[p]-p (each output holds 2L-1 entries,
 * none for L = 0 would break the packing, so lengths must be >= 1). */
void orc_karatsuba_ragged_i64(int64_t npairs, const int64_t *offsets, const int64_t *x,
                              const int64_t *y, int64_t *z) {
#pragma omp parallel for schedule(dynamic)
    for (int64_t p = 0; p < npairs; ++p) {
        int64_t o = offsets[p], len = offsets[p + 1] - offsets[p];
        kara_u64_one(len, (const u64w *)(x + o), (const u64w *)(y + o), (u64w *)(z + 2 * o - p));
    }
}

/* cyclic (Adhoc.hs:104-111): wrap the tail of the linear product */
void orc_cyclic_karatsuba_i64(int64_t len, int64_t rows, const int64_t *x, const int64_t *y, int64_t *z) {
    if (len == 0) return;
    int64_t zlen = 2 * len - 1;
    u64w *tmp = (u64w *)malloc(sizeof(u64w) * (size_t)zlen);
    for (int64_t r = 0; r < rows; ++r) {
        kara_u64_one(len, (const u64w *)(x + r * len), (const u64w *)(y + r * len), tmp);
        for (int64_t k = 0; k < len; ++k)
            ((u64w *)z)[r * len + k] = tmp[k] + ((k + len < zlen) ? tmp[k + len] : 0);
    }
    free(tmp);
}

/* complex (Adhoc.hs:114-133): Gauss' three real convolutions */
void orc_complex_karatsuba_f64(int cyclic, int64_t len, int64_t rows,
                               const double *xre, const double *xim,
                               const double *yre, const double *yim,
                               double *ore, double *oim) {
    if (len == 0) return;
    int64_t zlen = 2 * len - 1, olen = cyclic ? len : zlen;
    double *xm = (double *)malloc(sizeof(double) * (size_t)len), *ym = (double *)malloc(sizeof(double) * (size_t)len);
    double *a = (double *)malloc(sizeof(double) * (size_t)zlen), *b = (double *)malloc(sizeof(double) * (size_t)zlen),
           *c = (double *)malloc(sizeof(double) * (size_t)zlen);
    for (int64_t r = 0; r < rows; ++r) {
        const double *xr = xre + r * len, *xi = xim + r * len, *yr = yre + r * len, *yi = yim + r * len;
        for (int64_t k = 0; k < len; ++k) { xm[k] = xr[k] + xi[k]; ym[k] = yr[k] + yi[k]; }
        kara_f64_one(len, xr, yr, a);
        kara_f64_one(len, xi, yi, b);
        kara_f64_one(len, xm, ym, c);
        if (cyclic) {
            for (int64_t k = 0; k < len; ++k) {
                a[k] = a[k] + ((k + len < zlen) ? a[k + len] : 0.0);
                b[k] = b[k] + ((k + len < zlen) ? b[k + len] : 0.0);
                c[k] = c[k] + ((k + len < zlen) ? c[k + len] : 0.0);
            }
        }
        for (int64_t k = 0; k < olen; ++k) {
            ore[r * olen + k] = a[k] - b[k];
            oim[r * olen + k] = c[k] - (a[k] + b[k]);
        }
    }
    free(xm); free(ym); free(a); free(b); free(c);
}

/* complex (Adhoc.hs:114-133) over the Int64 Karatsuba: the same three real
 * products, in wrap-around arithmetic */
void orc_complex_karatsuba_i64(int cyclic, int64_t len, int64_t rows,
                               const int64_t *xre, const int64_t *xim,
                               const int64_t *yre, const int64_t *yim,
                               int64_t *ore, int64_t *oim) {
    if (len == 0) return;
    int64_t zlen = 2 * len - 1, olen = cyclic ? len : zlen;
    u64w *xm = (u64w *)malloc(sizeof(u64w) * (size_t)len), *ym = (u64w *)malloc(sizeof(u64w) * (size_t)len);
    u64w *a = (u64w *)malloc(sizeof(u64w) * (size_t)zlen), *b = (u64w *)malloc(sizeof(u64w) * (size_t)zlen),
         *c = (u64w *)malloc(sizeof(u64w) * (size_t)zlen);
    for (int64_t r = 0; r < rows; ++r) {
        const u64w *xr = (const u64w *)(xre + r * len), *xi = (const u64w *)(xim + r * len);
        const u64w *yr = (const u64w *)(yre + r * len), *yi = (const u64w *)(yim + r * len);
        for (int64_t k = 0; k < len; ++k) { xm[k] = xr[k] + xi[k]; ym[k] = yr[k] + yi[k]; }
        kara_u64_one(len, xr, yr, a);
        kara_u64_one(len, xi, yi, b);
        kara_u64_one(len, xm, ym, c);
        if (cyclic) {
            for (int64_t k = 0; k < len; ++k) {
                a[k] = a[k] + ((k + len < zlen) ? a[k + len] : 0);
                b[k] = b[k] + ((k + len < zlen) ? b[k + len] : 0);
                c[k] = c[k] + ((k + len < zlen) ? c[k + len] : 0);
            }
        }
        for (int64_t k = 0; k < olen; ++k) {
            ((u64w *)ore)[r * olen + k] = a[k] - b[k];
            ((u64w *)oim)[r * olen + k] = c[k] - (a[k] + b[k]);
        }
    }
    free(xm); free(ym); free(a); free(b); free(c);
}

/* ------------------------------------------------------------------------ */
/* Transforms, instantiated twice                                            */
/* ------------------------------------------------------------------------ */

#define REAL double
#define SFX(name) name##_f64
#define COS cos
#define SIN sin
#define SQRT sqrt
#define PI_R 3.14159265358979323846
#include "fourier_oracle_body.h"
#undef REAL
#undef SFX
#undef COS
#undef SIN
#undef SQRT
#undef PI_R

#define REAL float
#define SFX(name) name##_f32
#define COS cosf
#define SIN sinf
#define SQRT sqrtf
#define PI_R 3.14159265358979323846f
#include "fourier_oracle_body.h"
#undef REAL
#undef SFX
#undef COS
#undef SIN
#undef SQRT
#undef PI_R

/* Brute-force DFT: X_k = sum_j x_j exp(s 2 pi i (j k mod n) / n) in long double. */
void orc_dft_bruteforce(int sign, int64_t n, int64_t rows,
                        const double *re, const double *im, double *ore, double *oim) {
    const long double twopi = 6.283185307179586476925286766559005768L;
    long double *c = (long double *)malloc(sizeof(long double) * (size_t)(n > 0 ? n : 1));
    long double *s = (long double *)malloc(sizeof(long double) * (size_t)(n > 0 ? n : 1));
    for (int64_t m = 0; m < n; ++m) {
        c[m] = cosl(twopi * (long double)m / (long double)n);
        s[m] = (long double)sign * sinl(twopi * (long double)m / (long double)n);
    }
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t k = 0; k < n; ++k) {
            long double ar = 0, ai = 0;
            for (int64_t j = 0; j < n; ++j) {
                int64_t m = (int64_t)(((__int128)j * k) % n);
                long double xr = re[r * n + j], xi = im[r * n + j];
                ar += xr * c[m] - xi * s[m];
                ai += xr * s[m] + xi * c[m];
            }
            ore[r * n + k] = (double)ar;
            oim[r * n + k] = (double)ai;
        }
    free(c);
    free(s);
}
