/*
 * Replays, in plain C, the exact call sequence of the Haskell binding in
 * INTEGRATION.md against libaccelerate_fourier_amd.so:
 *
 *   transform / toSpectrum / fromSpectrum / karatsuba, for Double (AF_F64)
 *   and Float (AF_F32):
 *     af_plan_create once per partial application (the reference's `cache`);
 *     per call: device staging buffers kept with the plan and grown only
 *     when a call needs more, hipMemcpy of the planar component arrays host
 *     -> device, af_execute (or af_to_spectrum / af_from_spectrum /
 *     af_karatsuba_i64) on the default stream with no workspace, hipMemcpy
 *     device -> host, AF_UNSUPPORTED routed to the fallback (the reference
 *     function; here: a direct DFT on the host);
 *     af_plan_destroy when the plan's finaliser runs;
 *   and the device-resident path: a forward transform whose device output is
 *   the inverse transform's device input, no host copies in between.
 *
 * Results are checked against direct DFTs in long double and an exact
 * schoolbook convolution (wrap-around Int64 products), so the replay is a
 * parity test of the binding's marshalling: planar components, row-major
 * batches, the unnormalised convention, the sign of the sub-transform.
 *
 * Built by accelerate-fourier_amd/build.py with gcc (the HIP runtime's C
 * API, no HIP kernels here); run by tests/test_gpu_c_binding.py.  Prints one
 * line per check and "replay ok" at the end; exits non-zero on a failure.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "accelerate_fourier.h"

static int g_fail = 0;

#define HIP_OK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            exit(2);                                                                   \
        }                                                                              \
    } while (0)

static void check(int ok, const char *what, double err, double tol) {
    printf("%-58s err %.3e tol %.1e %s\n", what, err, tol, ok ? "ok" : "FAILED");
    fflush(stdout);
    if (!ok) g_fail = 1;
}

/* ---- deterministic inputs (splitmix64) ---- */
static uint64_t g_state;
static uint64_t next_u64(void) {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double next_unit(void) { return (double)(next_u64() >> 11) * (2.0 / 9007199254740992.0) - 1.0; }

/* ---- host references ---- */
/* X[k] = sum_j x[j] e^(sign 2 pi i j k / n) for k = 0, kstep, 2 kstep, ... (the
 * other bins are left 0), long double, exact index reduction */
static void dft(int sign, int64_t n, int64_t rows, const double *re, const double *im, double *ore, double *oim,
                int64_t kstep) {
    long double *c = malloc(sizeof(long double) * (size_t)n), *s = malloc(sizeof(long double) * (size_t)n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int64_t m = 0; m < n; ++m) {
        c[m] = cosl(2.0L * pi * (long double)m / (long double)n);
        s[m] = sign * sinl(2.0L * pi * (long double)m / (long double)n);
    }
    for (int64_t i = 0; i < rows * n; ++i) ore[i] = oim[i] = 0.0;
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t k = 0; k < n; k += kstep) {
            long double ar = 0, ai = 0;
            int64_t m = 0;
            for (int64_t j = 0; j < n; ++j) {
                const long double xr = re[r * n + j], xi = im[r * n + j];
                ar += xr * c[m] - xi * s[m];
                ai += xr * s[m] + xi * c[m];
                m += k;
                if (m >= n) m -= n;
            }
            ore[r * n + k] = (double)ar;
            oim[r * n + k] = (double)ai;
        }
    free(c);
    free(s);
}

/* relative L2 distance of a from b over the entries i with i % step == 0 */
static double rel_l2s(const double *ar, const double *ai, const double *br, const double *bi, int64_t count,
                      int64_t step) {
    long double num = 0, den = 0;
    for (int64_t i = 0; i < count; i += step) {
        const long double dr = (long double)ar[i] - br[i], di = ai ? (long double)ai[i] - bi[i] : 0;
        num += dr * dr + di * di;
        den += (long double)br[i] * br[i] + (bi ? (long double)bi[i] * bi[i] : 0);
    }
    return den > 0 ? (double)sqrtl(num / den) : (double)sqrtl(num);
}
static double rel_l2(const double *ar, const double *ai, const double *br, const double *bi, int64_t count) {
    return rel_l2s(ar, ai, br, bi, count, 1);
}

/* ---- the binding's plan object: the af_plan and its grow-only staging buffers ---- */
typedef struct {
    af_plan plan;
    int32_t precision;
    void *dev[4];
    size_t dev_bytes;
} Bound;

static af_status bind_plan(Bound *b, int64_t n, int sign, int kind, int precision) {
    af_plan_desc d;
    memset(&d, 0, sizeof d);
    d.length = n;
    d.sign = sign;
    d.kind = kind;
    d.precision = precision;
    d.algorithm = AF_ALG_AUTO;
    memset(b, 0, sizeof *b);
    b->precision = precision;
    return af_plan_create(&d, &b->plan);
}

static void staging(Bound *b, size_t bytes) {
    if (bytes <= b->dev_bytes) return;
    for (int i = 0; i < 4; ++i) {
        if (b->dev[i]) HIP_OK(hipFree(b->dev[i]));
        HIP_OK(hipMalloc(&b->dev[i], bytes));
    }
    b->dev_bytes = bytes;
}

static void unbind(Bound *b) {
    for (int i = 0; i < 4; ++i)
        if (b->dev[i]) HIP_OK(hipFree(b->dev[i]));
    if (b->plan) af_plan_destroy(b->plan);
    memset(b, 0, sizeof *b);
}

/* host arrays of the binding's element type (Double or Float) from doubles, and back */
static void *to_elem(const double *v, int64_t count, int precision) {
    if (precision == AF_F64) {
        double *o = malloc(sizeof(double) * (size_t)(count ? count : 1));
        memcpy(o, v, sizeof(double) * (size_t)count);
        return o;
    }
    float *o = malloc(sizeof(float) * (size_t)(count ? count : 1));
    for (int64_t i = 0; i < count; ++i) o[i] = (float)v[i];
    return o;
}
static void from_elem(const void *p, int64_t count, int precision, double *v) {
    for (int64_t i = 0; i < count; ++i) v[i] = precision == AF_F64 ? ((const double *)p)[i] : ((const float *)p)[i];
}

/* runPlanned: Planned.transform sign n on `rows` rows of planar complex data (host arrays) */
static af_status run_planned(Bound *b, int64_t n, int64_t rows, const double *re, const double *im, double *ore,
                             double *oim) {
    const size_t es = b->precision == AF_F64 ? 8 : 4, bytes = (size_t)(rows * n) * es;
    void *hre = to_elem(re, rows * n, b->precision), *him = to_elem(im, rows * n, b->precision);
    staging(b, bytes ? bytes : 1);
    HIP_OK(hipMemcpy(b->dev[0], hre, bytes, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(b->dev[1], him, bytes, hipMemcpyHostToDevice));
    const af_status st = af_execute(b->plan, rows, b->dev[0], b->dev[1], b->dev[2], b->dev[3], NULL, NULL);
    if (st == AF_SUCCESS) {
        HIP_OK(hipMemcpy(hre, b->dev[2], bytes, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(him, b->dev[3], bytes, hipMemcpyDeviceToHost));
        from_elem(hre, rows * n, b->precision, ore);
        from_elem(him, rows * n, b->precision, oim);
    }
    free(hre);
    free(him);
    return st;
}

static void test_transform(int precision, int64_t n, int sign, int64_t rows) {
    Bound b;
    char what[128];
    snprintf(what, sizeof what, "transform %s n=%lld sign=%d rows=%lld", precision == AF_F64 ? "Double" : "Float",
             (long long)n, sign, (long long)rows);
    const af_status pc = bind_plan(&b, n, sign, AF_C2C, precision);
    const int64_t cnt = rows * n;
    double *re = malloc(8 * (size_t)cnt), *im = malloc(8 * (size_t)cnt);
    double *ore = malloc(8 * (size_t)cnt), *oim = malloc(8 * (size_t)cnt);
    double *rre = malloc(8 * (size_t)cnt), *rim = malloc(8 * (size_t)cnt);
    for (int64_t i = 0; i < cnt; ++i) {
        re[i] = next_unit();
        im[i] = next_unit();
    }
    /* Float inputs are what the Float array holds */
    if (precision == AF_F32)
        for (int64_t i = 0; i < cnt; ++i) {
            re[i] = (float)re[i];
            im[i] = (float)im[i];
        }
    af_status st = pc;
    if (pc == AF_SUCCESS) st = run_planned(&b, n, rows, re, im, ore, oim);
    /* bins checked: all of them up to 8192, every 64th above (n is a multiple of 64 there) */
    const int64_t kstep = n > 8192 ? 64 : 1;
    if (st == AF_UNSUPPORTED) {
        /* the binding's fallback: the reference function */
        dft(sign, n, rows, re, im, ore, oim, 1);
        printf("%s: AF_UNSUPPORTED, fallback\n", what);
    } else if (st != AF_SUCCESS) {
        printf("%s: status %d (%s) %s\n", what, (int)st, af_status_string(st), af_last_error());
        g_fail = 1;
    }
    dft(sign, n, rows, re, im, rre, rim, kstep);
    const double tol = precision == AF_F64 ? 1e-12 : 1e-6;
    const double err = rel_l2s(ore, oim, rre, rim, cnt, kstep);
    check(st == AF_SUCCESS && err <= tol, what, err, tol);
    unbind(&b);
    free(re), free(im), free(ore), free(oim), free(rre), free(rim);
}

/* the device-resident path: forward then inverse on device pointers, no host copy between */
static void test_device_resident(int precision, int64_t n, int64_t rows) {
    Bound f, i;
    char what[128];
    snprintf(what, sizeof what, "device-resident forward->inverse %s n=%lld", precision == AF_F64 ? "Double" : "Float",
             (long long)n);
    if (bind_plan(&f, n, -1, AF_C2C, precision) != AF_SUCCESS || bind_plan(&i, n, 1, AF_C2C, precision) != AF_SUCCESS) {
        printf("%s: plan creation failed: %s\n", what, af_last_error());
        g_fail = 1;
        return;
    }
    const int64_t cnt = rows * n;
    const size_t es = precision == AF_F64 ? 8 : 4, bytes = (size_t)cnt * es;
    double *re = malloc(8 * (size_t)cnt), *im = malloc(8 * (size_t)cnt);
    double *bre = malloc(8 * (size_t)cnt), *bim = malloc(8 * (size_t)cnt);
    for (int64_t k = 0; k < cnt; ++k) {
        re[k] = next_unit();
        im[k] = next_unit();
        if (precision == AF_F32) {
            re[k] = (float)re[k];
            im[k] = (float)im[k];
        }
    }
    void *hre = to_elem(re, cnt, precision), *him = to_elem(im, cnt, precision);
    staging(&f, bytes);
    HIP_OK(hipMemcpy(f.dev[0], hre, bytes, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(f.dev[1], him, bytes, hipMemcpyHostToDevice));
    af_status st = af_execute(f.plan, rows, f.dev[0], f.dev[1], f.dev[2], f.dev[3], NULL, NULL);
    /* the inverse reads the forward's device output in place */
    if (st == AF_SUCCESS) st = af_execute(i.plan, rows, f.dev[2], f.dev[3], f.dev[2], f.dev[3], NULL, NULL);
    if (st == AF_SUCCESS) {
        HIP_OK(hipMemcpy(hre, f.dev[2], bytes, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(him, f.dev[3], bytes, hipMemcpyDeviceToHost));
        from_elem(hre, cnt, precision, bre);
        from_elem(him, cnt, precision, bim);
        for (int64_t k = 0; k < cnt; ++k) {
            bre[k] /= (double)n;
            bim[k] /= (double)n;
        }
    }
    const double tol = precision == AF_F64 ? 1e-12 : 1e-6;
    const double err = st == AF_SUCCESS ? rel_l2(bre, bim, re, im, cnt) : 1.0;
    check(st == AF_SUCCESS && err <= tol, what, err, tol);
    unbind(&f);
    unbind(&i);
    free(hre), free(him), free(re), free(im), free(bre), free(bim);
}

/* Real.toSpectrum (Planned.transform sign (n/2)) and Real.fromSpectrum (Planned.transform sign (n/2)) */
static void test_real(int precision, int64_t n, int64_t rows) {
    char what[128];
    const int64_t cnt = rows * n;
    const size_t es = precision == AF_F64 ? 8 : 4, bytes = (size_t)cnt * es;
    double *x = malloc(8 * (size_t)cnt), *zero = calloc((size_t)cnt, 8);
    double *sre = malloc(8 * (size_t)cnt), *sim = malloc(8 * (size_t)cnt);
    double *rre = malloc(8 * (size_t)cnt), *rim = malloc(8 * (size_t)cnt), *y = malloc(8 * (size_t)cnt);
    for (int64_t k = 0; k < cnt; ++k) {
        x[k] = next_unit();
        if (precision == AF_F32) x[k] = (float)x[k];
    }
    dft(-1, n, rows, x, zero, rre, rim, 1);
    /* toSpectrum forward */
    Bound t;
    snprintf(what, sizeof what, "toSpectrum %s n=%lld", precision == AF_F64 ? "Double" : "Float", (long long)n);
    af_status st = bind_plan(&t, n, -1, AF_R2C_FULL, precision);
    void *hx = to_elem(x, cnt, precision), *hre = malloc(bytes), *him = malloc(bytes);
    if (st == AF_SUCCESS) {
        staging(&t, bytes);
        HIP_OK(hipMemcpy(t.dev[0], hx, bytes, hipMemcpyHostToDevice));
        st = af_to_spectrum(t.plan, rows, t.dev[0], t.dev[2], t.dev[3], NULL, NULL);
    }
    if (st == AF_SUCCESS) {
        HIP_OK(hipMemcpy(hre, t.dev[2], bytes, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(him, t.dev[3], bytes, hipMemcpyDeviceToHost));
        from_elem(hre, cnt, precision, sre);
        from_elem(him, cnt, precision, sim);
    }
    const double tol = precision == AF_F64 ? 1e-12 : 1e-6;
    double err = st == AF_SUCCESS ? rel_l2(sre, sim, rre, rim, cnt) : 1.0;
    check(st == AF_SUCCESS && err <= tol, what, err, tol);
    unbind(&t);
    /* fromSpectrum inverse of the exact spectrum: n x */
    snprintf(what, sizeof what, "fromSpectrum %s n=%lld", precision == AF_F64 ? "Double" : "Float", (long long)n);
    st = bind_plan(&t, n, 1, AF_C2R_FULL, precision);
    void *sr = to_elem(rre, cnt, precision), *si = to_elem(rim, cnt, precision);
    if (st == AF_SUCCESS) {
        staging(&t, bytes);
        HIP_OK(hipMemcpy(t.dev[0], sr, bytes, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(t.dev[1], si, bytes, hipMemcpyHostToDevice));
        st = af_from_spectrum(t.plan, rows, t.dev[0], t.dev[1], t.dev[2], NULL, NULL);
    }
    if (st == AF_SUCCESS) {
        HIP_OK(hipMemcpy(hre, t.dev[2], bytes, hipMemcpyDeviceToHost));
        from_elem(hre, cnt, precision, y);
        for (int64_t k = 0; k < cnt; ++k) y[k] /= (double)n;
    }
    err = st == AF_SUCCESS ? rel_l2(y, NULL, x, NULL, cnt) : 1.0;
    check(st == AF_SUCCESS && err <= tol, what, err, tol);
    unbind(&t);
    free(x), free(zero), free(sre), free(sim), free(rre), free(rim), free(y);
    free(hx), free(hre), free(him), free(sr), free(si);
}

/* Convolution.Adhoc.karatsuba on Int64 rows: exact, wrap-around */
static void test_karatsuba(int64_t len, int64_t rows) {
    char what[128];
    snprintf(what, sizeof what, "karatsuba Int64 len=%lld rows=%lld", (long long)len, (long long)rows);
    const int64_t ni = rows * len, no = rows * (2 * len - 1);
    int64_t *x = malloc(8 * (size_t)ni), *y = malloc(8 * (size_t)ni);
    int64_t *z = malloc(8 * (size_t)no), *ref = calloc((size_t)no, 8);
    for (int64_t k = 0; k < ni; ++k) {
        x[k] = (int64_t)next_u64();  /* full 64-bit range: products wrap */
        y[k] = (int64_t)next_u64();
    }
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t i = 0; i < len; ++i)
            for (int64_t j = 0; j < len; ++j)
                ref[r * (2 * len - 1) + i + j] =
                    (int64_t)((uint64_t)ref[r * (2 * len - 1) + i + j] +
                              (uint64_t)x[r * len + i] * (uint64_t)y[r * len + j]);
    void *dx, *dy, *dz;
    HIP_OK(hipMalloc(&dx, 8 * (size_t)ni));
    HIP_OK(hipMalloc(&dy, 8 * (size_t)ni));
    HIP_OK(hipMalloc(&dz, 8 * (size_t)no));
    HIP_OK(hipMemcpy(dx, x, 8 * (size_t)ni, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dy, y, 8 * (size_t)ni, hipMemcpyHostToDevice));
    const af_status st = af_karatsuba_i64(rows, len, dx, dy, dz, NULL);
    if (st == AF_SUCCESS) HIP_OK(hipMemcpy(z, dz, 8 * (size_t)no, hipMemcpyDeviceToHost));
    int64_t bad = 0;
    for (int64_t k = 0; k < no; ++k) bad += st != AF_SUCCESS || z[k] != ref[k];
    check(st == AF_SUCCESS && bad == 0, what, (double)bad, 0.0);
    HIP_OK(hipFree(dx));
    HIP_OK(hipFree(dy));
    HIP_OK(hipFree(dz));
    free(x), free(y), free(z), free(ref);
}

int main(void) {
    g_state = 20240601ull;
    printf("library %s\n", af_version());
    for (int p = 0; p < 2; ++p) {
        const int prec = p == 0 ? AF_F64 : AF_F32;
        test_transform(prec, 1000, -1, 3);  /* 2^3 5^3: Stockham */
        test_transform(prec, 4099, -1, 2);  /* prime: Rader */
        test_transform(prec, 4099, 1, 2);
        test_transform(prec, 65536, -1, 2); /* four-step */
        test_transform(prec, 1, -1, 2);
        test_transform(prec, 0, -1, 2);
        test_device_resident(prec, 8192, 4);
        test_real(prec, 2048, 3);
    }
    test_karatsuba(1, 3);
    test_karatsuba(100, 3);
    test_karatsuba(1000, 2);
    /* a request the library refuses (a real transform of odd length) must come back
       AF_INVALID_ARG without a launch, as the reference's precondition */
    {
        Bound b;
        const af_status st = bind_plan(&b, 1001, -1, AF_R2C_FULL, AF_F64);
        check(st == AF_INVALID_ARG, "toSpectrum odd length -> AF_INVALID_ARG", (double)st, 0.0);
        if (st == AF_SUCCESS) unbind(&b);
    }
    if (g_fail) {
        printf("replay FAILED\n");
        return 1;
    }
    printf("replay ok\n");
    return 0;
}
