/*
 * Precision-generic body of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
 * Included twice by fourier_oracle.c with REAL = double and REAL = float.
 *
 * Each function restates one piece of the reference (paths relative to
 * src/Data/Array/Accelerate/).  The arithmetic follows the reference's
 * expressions operand by operand (the order of operations in the twiddle
 * arguments, the complex product formula, the Karatsuba-based 3- and 5-point
 * codelets), so that the only remaining differences are the libm cos/sin and
 * the association order of `fold1 (+)`, which the reference leaves to the
 * Accelerate backend.
 */

typedef struct { REAL re, im; } SFX(cx);
#define CX SFX(cx)

static inline CX SFX(mk)(REAL re, REAL im) { CX c; c.re = re; c.im = im; return c; }
static inline CX SFX(add)(CX a, CX b) { return SFX(mk)(a.re + b.re, a.im + b.im); }
static inline CX SFX(sub)(CX a, CX b) { return SFX(mk)(a.re - b.re, a.im - b.im); }
/* Complex product of Accelerate's Num (Complex a) instance. */
static inline CX SFX(mul)(CX a, CX b) {
    return SFX(mk)(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline CX SFX(conj)(CX a) { return SFX(mk)(a.re, -a.im); }
static inline CX SFX(scale)(REAL s, CX a) { return SFX(mk)(s * a.re, s * a.im); }

/* Sign.cis (Sign.hs:63-66): cos w :+ s * sin w */
static inline CX SFX(cis)(int sign, REAL w) { return SFX(mk)(COS(w), (REAL)sign * SIN(w)); }
/* Sign.cisRat (Sign.hs:68-72): cis (2*pi * numer / denom) */
static inline CX SFX(cisrat)(int sign, int64_t denom, int64_t numer) {
    REAL twopi = (REAL)2 * PI_R;
    return SFX(cis)(sign, (twopi * (REAL)numer) / (REAL)denom);
}

static void SFX(tr)(const oplan *p, int sign, const CX *in, CX *out);

/* Private.transform2..5 (Private.hs:95-218) with cache2..5 (56-82) */
static void SFX(small)(int64_t n, int sign, const CX *x, CX *o) {
    REAL s = (REAL)sign;
    if (n == 2) {
        CX z = SFX(mk)(-1, 0);
        o[0] = SFX(add)(x[0], x[1]);
        o[1] = SFX(add)(x[0], SFX(mul)(z, x[1]));
    } else if (n == 3) {
        REAL sqrt3d2 = SQRT((REAL)3) / 2, mhalf = (REAL)(-1) / 2;
        CX z = SFX(mk)(mhalf, s * sqrt3d2), z2 = SFX(mk)(mhalf, (-s) * sqrt3d2);
        /* Convolution.Small.sumAndConvolvePair (Small.hs:19-26) */
        CX sa01 = SFX(add)(x[1], x[2]), sb01 = SFX(add)(z, z2);
        CX ab0ab1 = SFX(add)(SFX(mul)(x[1], z), SFX(mul)(x[2], z2));
        CX c1 = SFX(sub)(SFX(mul)(sa01, sb01), ab0ab1);
        o[0] = SFX(add)(x[0], sa01);
        o[1] = SFX(add)(x[0], ab0ab1);
        o[2] = SFX(add)(x[0], c1);
    } else if (n == 4) {
        CX z = SFX(mk)(0, s), z2 = SFX(mk)(-1, 0), z3 = SFX(mk)(0, -s);
        CX x02a = SFX(add)(x[0], x[2]), x02b = SFX(add)(x[0], SFX(mul)(z2, x[2]));
        CX x13a = SFX(add)(x[1], x[3]), x13b = SFX(add)(x[1], SFX(mul)(z2, x[3]));
        o[0] = SFX(add)(x02a, x13a);
        o[1] = SFX(add)(x02b, SFX(mul)(z, x13b));
        o[2] = SFX(add)(x02a, SFX(mul)(z2, x13a));
        o[3] = SFX(add)(x02b, SFX(mul)(z3, x13b));
    } else { /* n == 5 */
        CX z1 = SFX(cisrat)(sign, 5, 1), z2 = SFX(cisrat)(sign, 5, 2);
        CX z3 = SFX(cisrat)(sign, 5, 3), z4 = SFX(cisrat)(sign, 5, 4);
        /* sumAndConvolveQuadruple (Small.hs:62-85) on (x1,x3,x4,x2) (z1,z2,z4,z3) */
        CX a0 = x[1], a1 = x[3], a2 = x[4], a3 = x[2];
        CX b0 = z1, b1 = z2, b2 = z4, b3 = z3;
        CX ab0 = SFX(mul)(a0, b0), ab1 = SFX(mul)(a1, b1);
        CX sa01 = SFX(add)(a0, a1), sb01 = SFX(add)(b0, b1);
        CX ab01 = SFX(sub)(SFX(mul)(sa01, sb01), SFX(add)(ab0, ab1));
        CX ab2 = SFX(mul)(a2, b2), ab3 = SFX(mul)(a3, b3);
        CX sa23 = SFX(add)(a2, a3), sb23 = SFX(add)(b2, b3);
        CX ab23 = SFX(sub)(SFX(mul)(sa23, sb23), SFX(add)(ab2, ab3));
        CX c0 = SFX(sub)(SFX(add)(ab0, ab2), SFX(add)(ab1, ab3));
        CX c1 = SFX(add)(ab01, ab23);
        CX ab02 = SFX(mul)(SFX(add)(a0, a2), SFX(add)(b0, b2));
        CX ab13 = SFX(mul)(SFX(add)(a1, a3), SFX(add)(b1, b3));
        CX sa0123 = SFX(add)(sa01, sa23), sb0123 = SFX(add)(sb01, sb23);
        CX ab0123 = SFX(sub)(SFX(mul)(sa0123, sb0123), SFX(add)(ab02, ab13));
        CX d0 = SFX(add)(ab13, c0), d1 = c1, d2 = SFX(sub)(ab02, c0), d3 = SFX(sub)(ab0123, c1);
        (void)sb0123;
        /* transform5 binds (d0,d1,d2,d3) to the outputs for indices 1,2,4,3 */
        o[0] = SFX(add)(x[0], sa0123);
        o[1] = SFX(add)(x[0], d0);
        o[2] = SFX(add)(x[0], d1);
        o[3] = SFX(add)(x[0], d3);
        o[4] = SFX(add)(x[0], d2);
    }
}

/* Private.transformRadix2InterleavedTime (Private.hs:294-313) with
 * twiddleFactors2 (264-276): evens +- cis(pi k / len2) * odds */
static void SFX(radix2)(const oplan *sub, int sign, int64_t n, const CX *x, CX *o) {
    int64_t h = n / 2;
    CX *buf = (CX *)malloc(sizeof(CX) * (size_t)(2 * n));
    CX *ev = buf, *od = buf + h, *E = buf + n, *O = buf + n + h;
    for (int64_t k = 0; k < h; ++k) { ev[k] = x[2 * k]; od[k] = x[2 * k + 1]; }
    SFX(tr)(sub, sign, ev, E);
    SFX(tr)(sub, sign, od, O);
    for (int64_t k = 0; k < h; ++k) {
        CX tw = SFX(cis)(sign, (PI_R * (REAL)k) / (REAL)h);
        CX t = SFX(mul)(tw, O[k]);
        o[k] = SFX(add)(E[k], t);
        o[k + h] = SFX(sub)(E[k], t);
    }
    free(buf);
}

/*
 * Breadth-first split-radix chain (Planned.hs:512-522, Private.hs:316-387).
 * `node` is a PlanSplitRadix node of length L2 whose sub-plan is split-radix
 * too, i.e. one CacheSplitRadixCons level (cacheFromPlanWithMap,
 * Planned.hs:376-391).  The pair (arr2: c2 rows of length L2, arr1: c1 rows of
 * length L2/2) is transformed in place.  When the sub-sub-plan is not
 * split-radix the chain ends in CacheSplitRadixEnd (cache sub) (cache sub.sub).
 */
static void SFX(sr_pair)(const oplan *node, int sign, CX *arr2, int64_t c2, int64_t L2,
                         CX *arr1, int64_t c1) {
    const oplan *sub = node->a; /* length L2/2 */
    /* ditSplitRadixReorder (Private.hs:358-364) */
    int64_t h = L2 / 2, q = L2 / 4;
    int64_t nc2 = c2 + c1, nc1 = 2 * c2;
    CX *u = (CX *)malloc(sizeof(CX) * (size_t)(nc2 * h + nc1 * q + h + 1));
    CX *z = u + nc2 * h, *tmp = z + nc1 * q;
    for (int64_t r = 0; r < c2; ++r)
        for (int64_t j = 0; j < h; ++j) u[r * h + j] = arr2[r * L2 + 2 * j];
    if (c1 > 0) memcpy(u + c2 * h, arr1, sizeof(CX) * (size_t)(c1 * h));
    for (int64_t r = 0; r < c2; ++r)
        for (int64_t j = 0; j < q; ++j) {
            z[(2 * r) * q + j] = arr2[r * L2 + 4 * j + 1];
            z[(2 * r + 1) * q + j] = arr2[r * L2 + 4 * j + 3];
        }
    if (sub->a->kind == PL_SPLITRADIX) {
        SFX(sr_pair)(sub, sign, u, nc2, h, z, nc1);
    } else {
        for (int64_t r = 0; r < nc2; ++r) {
            SFX(tr)(sub, sign, u + r * h, tmp);
            memcpy(u + r * h, tmp, sizeof(CX) * (size_t)h);
        }
        for (int64_t r = 0; r < nc1; ++r) {
            SFX(tr)(sub->a, sign, z + r * q, tmp);
            memcpy(z + r * q, tmp, sizeof(CX) * (size_t)q);
        }
    }
    /* ditSplitRadixStep (Private.hs:371-387) with twiddleFactorsSRPair for
     * len4 = L2/4 (levelCacheSplitRadix, Planned.hs:494-499) */
    CX imag = SFX(mk)(0, (REAL)sign);
    for (int64_t r = 0; r < c2; ++r) {
        for (int64_t j = 0; j < q; ++j) {
            REAL kj1 = (REAL)1 * (REAL)j, kj3 = (REAL)3 * (REAL)j, n4 = (REAL)q;
            CX tw1 = SFX(cis)(sign, (PI_R * kj1) / ((REAL)2 * n4));
            CX tw3 = SFX(cis)(sign, (PI_R * kj3) / ((REAL)2 * n4));
            CX ze = SFX(mul)(tw1, z[(2 * r) * q + j]);
            CX zo = SFX(mul)(tw3, z[(2 * r + 1) * q + j]);
            CX zsum = SFX(add)(ze, zo);
            CX zdiff = SFX(mul)(imag, SFX(sub)(ze, zo));
            CX ua = u[r * h + j], ub = u[r * h + q + j];
            arr2[r * L2 + j] = SFX(add)(ua, zsum);
            arr2[r * L2 + q + j] = SFX(add)(ub, zdiff);
            arr2[r * L2 + h + j] = SFX(sub)(ua, zsum);
            arr2[r * L2 + h + q + j] = SFX(sub)(ub, zdiff);
        }
    }
    /* the remaining c1 rows of u become arr1 (Sliced1.drop) */
    if (c1 > 0) memcpy(arr1, u + c2 * h, sizeof(CX) * (size_t)(c1 * h));
    free(u);
}

/* Cooley-Tukey transformComposite (Planned.hs:543-557) */
static void SFX(composite)(const oplan *p, int sign, const CX *x, CX *o) {
    int64_t n = p->a->n, m = p->b->n, N = p->n;
    CX *buf = (CX *)malloc(sizeof(CX) * (size_t)(2 * N + n + m));
    CX *A = buf, *B = buf + N, *t1 = buf + 2 * N, *t2 = t1 + n;
    for (int64_t k = 0; k < n; ++k) {
        for (int64_t j = 0; j < m; ++j) t2[j] = x[k + n * j];      /* sliceHorizontal */
        SFX(tr)(p->b, sign, t2, A + k * m);                        /* sub_m */
        for (int64_t j = 0; j < m; ++j)                            /* twiddleFactors */
            A[k * m + j] = SFX(mul)(SFX(cisrat)(sign, N, (k * j) % N), A[k * m + j]);
    }
    for (int64_t j = 0; j < m; ++j) {                               /* transpose, sub_n */
        for (int64_t k = 0; k < n; ++k) t1[k] = A[k * m + j];
        SFX(tr)(p->a, sign, t1, B + j * n);
    }
    for (int64_t q = 0; q < N; ++q) o[q] = B[(q % m) * n + q / m]; /* merge */
    free(buf);
}

/* Good-Thomas transformCoprime (Planned.hs:584-617) */
static void SFX(coprime)(const oplan *p, int sign, const CX *x, CX *o) {
    int64_t n = p->a->n, m = p->b->n, N = p->n;
    CX *buf = (CX *)malloc(sizeof(CX) * (size_t)(2 * N + n + m));
    CX *A = buf, *C = buf + N, *t1 = buf + 2 * N, *t2 = t1 + n;
    for (int64_t k = 0; k < m; ++k) {                               /* permuteSkewGrid m n */
        for (int64_t j = 0; j < n; ++j) t1[j] = x[(n * k + m * j) % N];
        SFX(tr)(p->a, sign, t1, A + k * n);                        /* sub_n */
    }
    for (int64_t a = 0; a < n; ++a) {                               /* transpose, sub_m */
        for (int64_t k = 0; k < m; ++k) t2[k] = A[k * n + a];
        SFX(tr)(p->b, sign, t2, C + a * m);
    }
    for (int64_t q = 0; q < N; ++q) o[q] = C[(q % n) * m + (q % m)]; /* permuteSkewGridInv */
    free(buf);
}

/* Rader: levelCachePrime (Planned.hs:650-664) and transformPrime (669-692).
 * The sub-transform pair is (inverse sign, forward sign) (directionModes,
 * Planned.hs:450-454). */
static void SFX(prime)(const oplan *p, int sign, const CX *x, CX *o) {
    int64_t P = p->n, L = P - 1;
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(3 * L));
    int64_t *rev = perm + L, *inv = rev + L;
    orc_perm_multiplicative(P, perm);
    orc_perm_reverse(L, perm, rev);
    orc_perm_inverse(L, perm, inv);
    CX *buf = (CX *)malloc(sizeof(CX) * (size_t)(4 * L));
    CX *kern = buf, *zs = buf + L, *a = buf + 2 * L, *t = buf + 3 * L;
    for (int64_t i = 0; i < L; ++i) kern[i] = SFX(cisrat)(sign, P, perm[i] + 1);
    SFX(tr)(p->a, +1, kern, zs);
    REAL r = (REAL)1 / (REAL)L;                                     /* scaleDown */
    for (int64_t i = 0; i < L; ++i) zs[i] = SFX(scale)(r, zs[i]);
    for (int64_t k = 0; k < L; ++k) a[k] = x[1 + rev[k]];           /* apply rev */
    SFX(tr)(p->a, +1, a, t);                                         /* trans */
    for (int64_t k = 0; k < L; ++k) t[k] = SFX(mul)(zs[k], t[k]);
    SFX(tr)(p->a, -1, t, a);                                         /* transInv */
    CX s = x[1];                                                     /* fold1 (+) */
    for (int64_t k = 1; k < L; ++k) s = SFX(add)(s, x[1 + k]);
    o[0] = SFX(add)(x[0], s);
    for (int64_t k = 0; k < L; ++k) o[1 + k] = SFX(add)(x[0], a[inv[k]]); /* apply inv */
    free(perm);
    free(buf);
}

/* Bluestein: levelCacheChirp (Planned.hs:744-754), Private.chirp
 * (Private.hs:428-436) and transformChirp (Planned.hs:760-773) */
static void SFX(chirp)(const oplan *p, int sign, const CX *x, CX *o) {
    int64_t n = p->n, M = p->a->n;
    CX *buf = (CX *)malloc(sizeof(CX) * (size_t)(4 * M));
    CX *ch = buf, *spec = buf + M, *y = buf + 2 * M, *t = buf + 3 * M;
    for (int64_t k = 0; k < M; ++k) {
        REAL sk = (REAL)((M > 2 * k) ? k : k - M);
        ch[k] = SFX(cis)(sign, ((PI_R * sk) * sk) / (REAL)n);
        y[k] = SFX(conj)(ch[k]);
    }
    SFX(tr)(p->a, +1, y, spec);
    REAL r = (REAL)1 / (REAL)M;
    for (int64_t k = 0; k < M; ++k) spec[k] = SFX(scale)(r, spec[k]);
    for (int64_t k = 0; k < M; ++k) y[k] = (k < n) ? SFX(mul)(ch[k], x[k]) : SFX(mk)(0, 0);
    SFX(tr)(p->a, +1, y, t);
    for (int64_t k = 0; k < M; ++k) t[k] = SFX(mul)(spec[k], t[k]);
    SFX(tr)(p->a, -1, t, y);
    for (int64_t k = 0; k < n; ++k) o[k] = SFX(mul)(ch[k], y[k]);
    free(buf);
}

/* transformWithCache (Planned.hs:137-167) for one row; in and out distinct. */
static void SFX(tr)(const oplan *p, int sign, const CX *in, CX *out) {
    switch (p->kind) {
    case PL_IDENTITY:
        if (p->n > 0) memcpy(out, in, sizeof(CX) * (size_t)p->n);
        break;
    case PL_SMALL:
        SFX(small)(p->n, sign, in, out);
        break;
    case PL_SPLITRADIX:
        if (p->a->kind != PL_SPLITRADIX) {
            SFX(radix2)(p->a, sign, p->n, in, out);            /* CacheRadix2 */
        } else {                                                /* CacheSplitRadix */
            /* initSplitRadix / finishSplitRadix (Private.hs:316-329) */
            memcpy(out, in, sizeof(CX) * (size_t)p->n);
            SFX(sr_pair)(p, sign, out, 1, p->n, NULL, 0);
        }
        break;
    case PL_PRIME: SFX(prime)(p, sign, in, out); break;
    case PL_COPRIME: SFX(coprime)(p, sign, in, out); break;
    case PL_COMPOSITE: SFX(composite)(p, sign, in, out); break;
    case PL_CHIRP: SFX(chirp)(p, sign, in, out); break;
    }
}

int SFX(orc_transform)(int algorithm, int sign, int64_t n, int64_t rows,
                       const REAL *re, const REAL *im, REAL *ore, REAL *oim) {
    if (n < 0 || rows < 0 || (sign != 1 && sign != -1)) return -1;
    if (n == 0 || rows == 0) return 0;
    oplan *p = plan_make(n, algorithm);
#pragma omp parallel
    {
        CX *a = (CX *)malloc(sizeof(CX) * (size_t)n), *b = (CX *)malloc(sizeof(CX) * (size_t)n);
#pragma omp for schedule(dynamic)
        for (int64_t r = 0; r < rows; ++r) {
            for (int64_t k = 0; k < n; ++k) a[k] = SFX(mk)(re[r * n + k], im[r * n + k]);
            SFX(tr)(p, sign, a, b);
            for (int64_t k = 0; k < n; ++k) { ore[r * n + k] = b[k].re; oim[r * n + k] = b[k].im; }
        }
        free(a);
        free(b);
    }
    plan_free(p);
    return 0;
}

/* toSpectrum (Real.hs:53-70) with complexDeinterleave (72-83) and
 * Cyclic.reverse (Cyclic.hs:20-28); the sub-transform is Planned.transform
 * sub_sign (N/2). */
int SFX(orc_to_spectrum)(int sub_sign, int64_t N, int64_t rows, const REAL *x,
                         REAL *ore, REAL *oim) {
    if (N < 0 || N % 2 != 0 || (sub_sign != 1 && sub_sign != -1)) return -1;
    if (N == 0 || rows == 0) return 0;
    int64_t n2 = N / 2;
    oplan *p = plan_make(n2, ORC_AUTO);
#pragma omp parallel
    {
        CX *z = (CX *)malloc(sizeof(CX) * (size_t)n2), *Z = (CX *)malloc(sizeof(CX) * (size_t)n2);
#pragma omp for schedule(dynamic)
        for (int64_t r = 0; r < rows; ++r) {
            const REAL *xr = x + r * N;
            for (int64_t j = 0; j < n2; ++j) z[j] = SFX(mk)(xr[2 * j], xr[2 * j + 1]);
            SFX(tr)(p, sub_sign, z, Z);
            for (int64_t k = 0; k < n2; ++k) {
                CX zk = Z[k], zr = Z[(n2 - k) % n2];
                CX xp = SFX(add)(zk, zr), xm = SFX(sub)(zk, zr);
                xp = SFX(mk)(xp.re / 2, xp.im / 2);
                xm = SFX(mk)(xm.re / 2, xm.im / 2);
                CX tw = SFX(mul)(SFX(mk)(0, 1), SFX(cis)(-1, (PI_R * (REAL)k) / (REAL)n2));
                CX ev = SFX(mk)(xp.re, xm.im);
                CX od = SFX(mul)(tw, SFX(mk)(xm.re, xp.im));
                CX lo = SFX(sub)(ev, od), hi = SFX(add)(ev, od);
                ore[r * N + k] = lo.re; oim[r * N + k] = lo.im;
                ore[r * N + n2 + k] = hi.re; oim[r * N + n2 + k] = hi.im;
            }
        }
        free(z);
        free(Z);
    }
    plan_free(p);
    return 0;
}

/* fromSpectrum (Real.hs:96-107) with complexInterleave (109-121) */
int SFX(orc_from_spectrum)(int sub_sign, int64_t N, int64_t rows, const REAL *re,
                           const REAL *im, REAL *y) {
    if (N < 0 || N % 2 != 0 || (sub_sign != 1 && sub_sign != -1)) return -1;
    if (N == 0 || rows == 0) return 0;
    int64_t n2 = N / 2;
    oplan *p = plan_make(n2, ORC_AUTO);
#pragma omp parallel
    {
        CX *z = (CX *)malloc(sizeof(CX) * (size_t)n2), *Z = (CX *)malloc(sizeof(CX) * (size_t)n2);
#pragma omp for schedule(dynamic)
        for (int64_t r = 0; r < rows; ++r) {
            for (int64_t k = 0; k < n2; ++k) {
                CX p0 = SFX(mk)(re[r * N + k], im[r * N + k]);
                CX p1 = SFX(mk)(re[r * N + n2 + k], im[r * N + n2 + k]);
                CX tw = SFX(mul)(SFX(mk)(0, 1), SFX(cis)(+1, (PI_R * (REAL)k) / (REAL)n2));
                CX fe = SFX(add)(p0, p1);
                CX fo = SFX(mul)(tw, SFX(sub)(p0, p1));
                z[k] = SFX(add)(fe, fo);
            }
            SFX(tr)(p, sub_sign, z, Z);
            for (int64_t j = 0; j < n2; ++j) { y[r * N + 2 * j] = Z[j].re; y[r * N + 2 * j + 1] = Z[j].im; }
        }
        free(z);
        free(Z);
    }
    plan_free(p);
    return 0;
}

/* untangleSpectra (Real.hs:168-173) with untangleCoefficient (182-187) and
 * Cyclic.reverse (Cyclic.hs:20-28): for the spectrum z of x + i y,
 * a_k = (z_k + conj z_-k) / 2 and b_k = (z_k - conj z_-k) (-i / 2). */
int SFX(orc_untangle_spectra)(int64_t n, int64_t rows, const REAL *zre, const REAL *zim, REAL *are,
                              REAL *aim, REAL *bre, REAL *bim) {
    if (n < 0 || rows < 0) return -1;
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t k = 0; k < n; ++k) {
            const int64_t kr = (n - k) % n;
            CX a = SFX(mk)(zre[r * n + k], zim[r * n + k]);
            CX bc = SFX(mk)(zre[r * n + kr], -zim[r * n + kr]);
            CX s = SFX(add)(a, bc), d = SFX(sub)(a, bc);
            CX half_s = SFX(mk)(s.re / 2, s.im / 2);
            CX mi2 = SFX(mul)(d, SFX(mk)(0, (REAL)-1 / 2));
            are[r * n + k] = half_s.re; aim[r * n + k] = half_s.im;
            bre[r * n + k] = mi2.re; bim[r * n + k] = mi2.im;
        }
    return 0;
}

/* entangleSpectra (Real.hs:208-224): the first component of
 * entangleCoefficient c d (231-236), c + d i. */
int SFX(orc_entangle_spectra)(int64_t n, int64_t rows, const REAL *cre, const REAL *cim, const REAL *dre,
                              const REAL *dim, REAL *zre, REAL *zim) {
    if (n < 0 || rows < 0) return -1;
    for (int64_t i = 0; i < n * rows; ++i) {
        CX di = SFX(mul)(SFX(mk)(dre[i], dim[i]), SFX(mk)(0, 1));
        CX z = SFX(add)(SFX(mk)(cre[i], cim[i]), di);
        zre[i] = z.re; zim[i] = z.im;
    }
    return 0;
}

/* convolveCyclic (Planned.hs:779-821): transInv (scaleDown (trans x) * trans y)
 * with (trans, transInv) = (inverse sign, forward sign). */
int SFX(orc_convolve_cyclic)(int64_t n, int64_t rows, const REAL *xre, const REAL *xim,
                             const REAL *yre, const REAL *yim, REAL *ore, REAL *oim) {
    if (n < 0 || rows < 0) return -1;
    if (n == 0 || rows == 0) return 0;
    oplan *p = plan_make(n, ORC_AUTO);
    REAL r = (REAL)1 / (REAL)n;
#pragma omp parallel
    {
        CX *a = (CX *)malloc(sizeof(CX) * (size_t)n), *b = (CX *)malloc(sizeof(CX) * (size_t)n);
        CX *c = (CX *)malloc(sizeof(CX) * (size_t)n);
#pragma omp for schedule(dynamic)
        for (int64_t row = 0; row < rows; ++row) {
            for (int64_t k = 0; k < n; ++k) a[k] = SFX(mk)(xre[row * n + k], xim[row * n + k]);
            SFX(tr)(p, +1, a, b);
            for (int64_t k = 0; k < n; ++k) a[k] = SFX(mk)(yre[row * n + k], yim[row * n + k]);
            SFX(tr)(p, +1, a, c);
            for (int64_t k = 0; k < n; ++k) a[k] = SFX(mul)(SFX(scale)(r, b[k]), c[k]);
            SFX(tr)(p, -1, a, b);
            for (int64_t k = 0; k < n; ++k) { ore[row * n + k] = b[k].re; oim[row * n + k] = b[k].im; }
        }
        free(a); free(b); free(c);
    }
    plan_free(p);
    return 0;
}

#undef CX
