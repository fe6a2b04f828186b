// Device side of the transform path: complex helpers, register butterflies,
// LDS Stockham stages and the generic tile pass.
//
// Written for gfx950 (CDNA4): 64-wide wavefronts, 160 KiB LDS per CU.  A tile
// of B lines lives in LDS; every stage reads R inputs per butterfly into
// registers, applies the stage twiddles, runs the radix-R DFT and writes the
// R outputs back in Stockham (self-sorting) order, so the result of the last
// stage is in natural order with no bit reversal.
#pragma once
#include <hip/hip_runtime.h>

#include "af_pass.hpp"

namespace af {

template <typename T> struct Cx;
template <> struct Cx<float> { using C = float2; using W = float4; };
template <> struct Cx<double> { using C = double2; using W = double2; };

__device__ __forceinline__ float2 mk(float a, float b) { return make_float2(a, b); }
__device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }
// Loads and stores of element i through descriptor pointers as global (not
// flat) memory instructions.  HIP's vector types move as clang vectors of the
// same shape, which copy across address spaces.
template <typename T> using gmem = __attribute__((address_space(1))) T;
template <typename T> using kmem = __attribute__((address_space(4))) T;  // kernel arguments
template <typename V> struct GVec { typedef V type; };
template <> struct GVec<float2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct GVec<float4> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct GVec<double2> { typedef double type __attribute__((ext_vector_type(2))); };
template <typename V> __device__ __forceinline__ V gld(const void *p, int64_t i) {
    typedef typename GVec<V>::type E;
    const E e = ((const gmem<E> *)p)[i];
    V v;
    __builtin_memcpy(&v, &e, sizeof v);
    return v;
}
template <typename V> __device__ __forceinline__ void gst(void *p, int64_t i, V v) {
    typedef typename GVec<V>::type E;
    E e;
    __builtin_memcpy(&e, &v, sizeof e);
    ((gmem<E> *)p)[i] = e;
}

template <class C> __device__ __forceinline__ C cadd(C a, C b) { return mk(a.x + b.x, a.y + b.y); }
template <class C> __device__ __forceinline__ C csub(C a, C b) { return mk(a.x - b.x, a.y - b.y); }
template <class C> __device__ __forceinline__ C cmul(C a, C b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class C> __device__ __forceinline__ C cconj(C a) { return mk(a.x, -a.y); }
template <class C, class T> __device__ __forceinline__ C cscale(C a, T s) { return mk(a.x * s, a.y * s); }
// a * (s i)
template <class C, class T> __device__ __forceinline__ C mul_si(C a, T s) { return mk(-s * a.y, s * a.x); }

// Twiddle-table entries hold exp(-2 pi i a / N) (the forward sign).  For FP32
// an entry is a compensated pair hi + lo (float4: hi.re, hi.im, lo.re, lo.im)
// generated from FP64; for FP64 it is a double2.
//
// twmul(x, w, s) multiplies x by exp(s 2 pi i a / N): by the table value
// itself for s = -1 and by its conjugate for s = +1.
__device__ __forceinline__ float2 twmul(float2 x, float4 w, float s, bool plain) {
    const float hy = -s * w.y;
    float re = x.x * w.x - x.y * hy;
    float im = x.x * hy + x.y * w.x;
    if (!plain) {
        const float ly = -s * w.w;
        re = fmaf(x.x, w.z, fmaf(-x.y, ly, re));
        im = fmaf(x.x, ly, fmaf(x.y, w.z, im));
    }
    return make_float2(re, im);
}
__device__ __forceinline__ double2 twmul(double2 x, double2 w, double s, bool) {
    const double wy = -s * w.y;
    return make_double2(x.x * w.x - x.y * wy, x.x * wy + x.y * w.x);
}
// multiply by a stored vector entry as it is
template <class C, class W> __device__ __forceinline__ C vmul(C x, W w, bool plain) {
    return twmul(x, w, decltype(x.x)(-1), plain);
}

// Split an FP64 complex value into the table format of precision T.
__device__ __forceinline__ void to_w(double2 v, float4 &w) {
    const float hx = (float)v.x, hy = (float)v.y;
    w = make_float4(hx, hy, (float)(v.x - (double)hx), (float)(v.y - (double)hy));
}
__device__ __forceinline__ void to_w(double2 v, double2 &w) { w = v; }
// plain complex value of the precision (one rounding)
__device__ __forceinline__ void to_w(double2 v, float2 &w) { w = make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ double2 from_w(float4 w) {
    return make_double2((double)w.x + (double)w.z, (double)w.y + (double)w.w);
}
__device__ __forceinline__ double2 from_w(double2 w) { return w; }

// x * w for an FP64 twiddle w, in precision T: for FP32 the product is formed
// in FP64 and rounded once (two conversions, a complex FP64 product and two
// conversions back -- fewer instructions than splitting w into a hi/lo pair for
// a compensated FP32 product, and no less accurate).
__device__ __forceinline__ float2 dmul(float2 x, double2 w, bool plain) {
    if (plain) {
        float4 ww;
        to_w(w, ww);
        return twmul(x, ww, -1.0f, true);
    }
    const double xr = (double)x.x, xi = (double)x.y;
    return make_float2((float)(xr * w.x - xi * w.y), (float)(xr * w.y + xi * w.x));
}
__device__ __forceinline__ double2 dmul(double2 x, double2 w, bool) { return cmul(x, w); }

// ---------------------------------------------------------------------------
// Register butterflies: in-place DFT of R values, exp(s 2 pi i j k / R).
// ---------------------------------------------------------------------------

// x * exp(s 2 pi i k / R) for a compile-time-foldable k in [0, R/2)
template <int R, class C, class T>
__device__ __forceinline__ C mul_wr(C x, int k, T s) {
    using R_ = decltype(x.x);
    constexpr double SQ = 0.707106781186547524401;
    if (k == 0) return x;
    if (4 * k == R) return mul_si(x, s);
    if (8 * k == R) return mk(R_(SQ) * (x.x - s * x.y), R_(SQ) * (x.y + s * x.x));
    if (8 * k == 3 * R) return mk(R_(SQ) * (-x.x - s * x.y), R_(SQ) * (-x.y + s * x.x));
    // remaining cases occur for R = 16: angles k pi / 8 with k odd
    constexpr double C8 = 0.923879532511286756128, S8 = 0.382683432365089771728;
    double c = 0, sn = 0;
    const int k16 = k * (16 / R);
    if (k16 == 1) { c = C8; sn = S8; }
    else if (k16 == 3) { c = S8; sn = C8; }
    else if (k16 == 5) { c = -S8; sn = C8; }
    else { c = -C8; sn = S8; }
    const R_ cc = R_(c), ss = R_(sn) * s;
    return mk(x.x * cc - x.y * ss, x.x * ss + x.y * cc);
}

template <int R, class C, class T>
__device__ __forceinline__ void dft_p2(C *v, T s) {
    if constexpr (R == 1) {
        return;
    } else if constexpr (R == 2) {
        const C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else {
        constexpr int M = R / 2;
        C e[M], o[M];
#pragma unroll
        for (int j = 0; j < M; ++j) { e[j] = v[2 * j]; o[j] = v[2 * j + 1]; }
        dft_p2<M>(e, s);
        dft_p2<M>(o, s);
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const C t = mul_wr<R>(o[k], k, s);
            v[k] = cadd(e[k], t);
            v[k + M] = csub(e[k], t);
        }
    }
}

template <int R> struct OddC;
template <> struct OddC<3> {
    static constexpr double c[3] = {1.0, -0.5, -0.5};
    static constexpr double s[3] = {0.0, 0.866025403784438646764, -0.866025403784438646764};
};
template <> struct OddC<5> {
    static constexpr double c[5] = {1.0, 0.309016994374947424102, -0.809016994374947424102,
                                    -0.809016994374947424102, 0.309016994374947424102};
    static constexpr double s[5] = {0.0, 0.951056516295153572116, 0.587785252292473129169,
                                    -0.587785252292473129169, -0.951056516295153572116};
};
template <> struct OddC<7> {
    static constexpr double c[7] = {1.0, 0.623489801858733530525, -0.222520933956314404289,
                                    -0.900968867902419126236, -0.900968867902419126236,
                                    -0.222520933956314404289, 0.623489801858733530525};
    static constexpr double s[7] = {0.0, 0.781831482468029808708, 0.974927912181823607018,
                                    0.433883739117558120476, -0.433883739117558120476,
                                    -0.974927912181823607018, -0.781831482468029808708};
};
template <> struct OddC<11> {
    static constexpr double c[11] = {1.0, 0.841253532831181168862, 0.415415013001886425529,
                                     -0.142314838273285140444, -0.654860733945285064057,
                                     -0.95949297361449738989, -0.95949297361449738989,
                                     -0.654860733945285064057, -0.142314838273285140444,
                                     0.415415013001886425529, 0.841253532831181168862};
    static constexpr double s[11] = {0.0, 0.540640817455597582108, 0.909631995354518371412,
                                     0.989821441880932732376, 0.755749574354258283774,
                                     0.281732556841429697711, -0.281732556841429697711,
                                     -0.755749574354258283774, -0.989821441880932732376,
                                     -0.909631995354518371412, -0.540640817455597582108};
};
template <> struct OddC<13> {
    static constexpr double c[13] = {1.0, 0.8854560256532098959, 0.568064746731155802512,
                                     0.120536680255323053349, -0.35460488704253562597,
                                     -0.748510748171101098635, -0.970941817426052027157,
                                     -0.970941817426052027157, -0.748510748171101098635,
                                     -0.35460488704253562597, 0.120536680255323053349,
                                     0.568064746731155802512, 0.8854560256532098959};
    static constexpr double s[13] = {0.0, 0.464723172043768545656, 0.82298386589365639458,
                                     0.992708874098053992801, 0.93501624268541482344,
                                     0.663122658240795202377, 0.239315664287557767149,
                                     -0.239315664287557767149, -0.663122658240795202377,
                                     -0.93501624268541482344, -0.992708874098053992801,
                                     -0.82298386589365639458, -0.464723172043768545656};
};

// Odd prime R: pair x_j with x_{R-j} so each output pair (k, R-k) shares the
// cosine and sine sums.
template <int R, class C, class T>
__device__ __forceinline__ void dft_odd(C *v, T s) {
    using R_ = decltype(v[0].x);
    constexpr int H = (R - 1) / 2;
    C a[H], b[H];
    const C x0 = v[0];
    C sum = x0;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        a[j - 1] = cadd(v[j], v[R - j]);
        b[j - 1] = csub(v[j], v[R - j]);
        sum = cadd(sum, a[j - 1]);
    }
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        C t = x0, u = mk(R_(0), R_(0));
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const int m = (j * k) % R;
            const R_ c = R_(OddC<R>::c[m]), sn = R_(OddC<R>::s[m]);
            t = mk(t.x + a[j - 1].x * c, t.y + a[j - 1].y * c);
            u = mk(u.x + b[j - 1].x * sn, u.y + b[j - 1].y * sn);
        }
        v[k] = mk(t.x - s * u.y, t.y + s * u.x);
        v[R - k] = mk(t.x + s * u.y, t.y - s * u.x);
    }
    v[0] = sum;
}

template <int R, class C, class T>
__device__ __forceinline__ void dft(C *v, T s) {
    if constexpr (R == 2 || R == 4 || R == 8 || R == 16) dft_p2<R>(v, s);
    else dft_odd<R>(v, s);
}

// LDS index padding: one spare slot every 2^ps elements spreads strided
// accesses over the banks.
// The pass descriptors always pad with one slot every 16 elements (padshift 4),
// so the shift is a constant here.
__device__ __forceinline__ int pidx(int i, int) { return i + (i >> 4); }

// q = n / d, r = n % d for 0 <= n < 2^22, d >= 1: the float product n * (1/d)
// is within 0.5 of n / d there, so truncation is off by at most one.
__device__ __forceinline__ void divmod_small(int n, int d, float rcp, int &q, int &r) {
    q = (int)((float)n * rcp);
    r = n - q * d;
    if (r < 0) {
        --q;
        r += d;
    } else if (r >= d) {
        ++q;
        r -= d;
    }
}

// A value every lane of the wavefront holds alike (a descriptor field), moved
// to scalar registers: branches on it become scalar branches and the compiler
// no longer treats it as varying per lane (the descriptor is read through a
// generic or LDS pointer, whose loads it must assume differ between lanes).
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni(int64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(uint64_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <typename P> __device__ __forceinline__ P *uni(P *p) { return (P *)(uintptr_t)uni((int64_t)(uintptr_t)p); }
__device__ __forceinline__ double uni(double v) { return __longlong_as_double(uni((int64_t)__double_as_longlong(v))); }

// One Stockham stage of radix R over a tile of B lines of length L.
// Butterfly i of a line reads elements i + j L/R and writes
// (i - k) R + k + j p with k = i mod p, p = product of the previous radices.
// notw: a prime-factor pass whose stages are pairwise coprime (each its own
// group): no stage has twiddles.
template <typename T, int R>
__device__ __forceinline__ void stage(typename Cx<T>::C *lds, const int L, const int ld, const int B,
                                      const int p, const void *__restrict__ tw, const T s,
                                      const int ps, const bool plain, const bool notw) {
    using C = typename Cx<T>::C;
    constexpr int MB = (pass_ept(sizeof(T) == 8) + R - 1) / R;
    const int nbl = L / R;
    const int nbf = B * nbl;
    const int tws = L / (p * R);
    const int tid = threadIdx.x, NT = blockDim.x;
    const float rnbl = 1.0f / (float)nbl, rp = 1.0f / (float)p;
    C v[MB][R];
    int base[MB], ii[MB];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int bf = tid + b * NT;
        base[b] = 0;
        ii[b] = 0;
        if (bf < nbf) {
            int line, i;
            divmod_small(bf, nbl, rnbl, line, i);
            base[b] = line * ld;
            ii[b] = i;
#pragma unroll
            for (int j = 0; j < R; ++j) v[b][j] = lds[pidx(base[b] + i + j * nbl, ps)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int bf = tid + b * NT;
        if (bf < nbf) {
            const int i = ii[b];
            int q, k;
            divmod_small(i, p, rp, q, k);
            if (p > 1 && !notw) {
#pragma unroll
                for (int j = 1; j < R; ++j)
                    v[b][j] = twmul(v[b][j], gld<typename Cx<T>::W>(tw, j * k * tws), s, plain);
            }
            dft<R>(v[b], s);
            const int o = base[b] + (i - k) * R + k;
#pragma unroll
            for (int j = 0; j < R; ++j) lds[pidx(o + j * p, ps)] = v[b][j];
        }
    }
    __syncthreads();
}

template <typename T, typename D>
__device__ __forceinline__ void run_stages(typename Cx<T>::C *lds, const D &d, const T s) {
    const void *tw0 = uni(d.tw);
    const bool plain = (uni(d.flags) & AFP_PLAIN_TW) != 0;
    const bool notw = uni(d.grp_mask) != 0;
    const int L0 = uni(d.L), ld0 = uni(d.ld), B0 = uni(d.B), ps = uni(d.padshift), nst = uni(d.nst);
    for (int st = 0; st < nst; ++st) {
        // the geometry re-read in scalar registers in every stage: values
        // derived from it (per radix) stay inside the stage instead of being
        // hoisted out of the loop for every radix at once
        int L = L0, ld = ld0, B = B0;
        const void *tw = tw0;
        asm volatile("" : "+s"(L), "+s"(ld), "+s"(B), "+s"(tw));
        const int R = uni(d.radix[st]), p = uni(d.pst[st]);
        switch (R) {
        case 2: stage<T, 2>(lds, L, ld, B, p, tw, s, ps, plain, notw); break;
        case 3: stage<T, 3>(lds, L, ld, B, p, tw, s, ps, plain, notw); break;
        case 4: stage<T, 4>(lds, L, ld, B, p, tw, s, ps, plain, notw); break;
        case 5: stage<T, 5>(lds, L, ld, B, p, tw, s, ps, plain, notw); break;
        case 7: stage<T, 7>(lds, L, ld, B, p, tw, s, ps, plain, notw); break;
        case 8: stage<T, 8>(lds, L, ld, B, p, tw, s, ps, plain, notw); break;
        case 11: stage<T, 11>(lds, L, ld, B, p, tw, s, ps, plain, notw); break;
        case 13: stage<T, 13>(lds, L, ld, B, p, tw, s, ps, plain, notw); break;
        default: stage<T, 16>(lds, L, ld, B, p, tw, s, ps, plain, notw); break;
        }
    }
}

// ---------------------------------------------------------------------------
// Element access
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ typename Cx<T>::C ld_c(const void *re, const void *im, int64_t a) {
    if (im) return mk(gld<T>(re, a), gld<T>(im, a));
    return gld<typename Cx<T>::C>(re, a);
}
template <typename T>
__device__ __forceinline__ void st_c(void *re, void *im, int64_t a, typename Cx<T>::C v) {
    if (im) {
        gst<T>(re, a, v.x);
        gst<T>(im, a, v.y);
    } else {
        gst<typename Cx<T>::C>(re, a, v);
    }
}

// The four-step twiddle fields of a descriptor, held in scalar registers.
struct FtTables {
    const void *ft_coarse;
    const void *ft_fine;
    int32_t ftShift, ftSign;
};

// Four-step twiddle W_N^(sign * a), a < N, as the product coarse[a >> shift] *
// fine[a & mask] of two table entries.  FP64: plain double product.  FP32: the
// hi parts' product plus the first-order lo corrections, accurate to about one
// FP32 ulp.
template <typename D>
__device__ __forceinline__ double2 ft_twiddle_t(const D &d, uint64_t a, double) {
    const double2 c = gld<double2>(d.ft_coarse, a >> d.ftShift);
    const double2 f = gld<double2>(d.ft_fine, a & ((uint64_t(1) << d.ftShift) - 1));
    double2 w = cmul(c, f);
    if (d.ftSign > 0) w.y = -w.y;
    return w;
}
// FP64 four-step twiddle from the FP64 tables (both precisions).
template <typename D>
__device__ __forceinline__ double2 ft_twiddle_d(const D &d, uint64_t a) {
    const double2 c = gld<double2>(d.ft_coarse_d, a >> d.ftShift);
    const double2 f = gld<double2>(d.ft_fine_d, a & ((uint64_t(1) << d.ftShift) - 1));
    double2 w = cmul(c, f);
    if (d.ftSign > 0) w.y = -w.y;
    return w;
}
template <typename D>
__device__ __forceinline__ float2 ft_twiddle_t(const D &d, uint64_t a, float) {
    const float4 c = gld<float4>(d.ft_coarse, a >> d.ftShift);
    const float4 f = gld<float4>(d.ft_fine, a & ((uint64_t(1) << d.ftShift) - 1));
    float re = fmaf(c.x, f.z, fmaf(-c.y, f.w, fmaf(c.z, f.x, -c.w * f.y)));
    float im = fmaf(c.x, f.w, fmaf(c.y, f.z, fmaf(c.z, f.y, c.w * f.x)));
    re = fmaf(c.x, f.x, fmaf(-c.y, f.y, re));
    im = fmaf(c.x, f.y, fmaf(c.y, f.x, im));
    return make_float2(re, d.ftSign > 0 ? -im : im);
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b mod 8), and each XCD has its own L2.  Give each XCD a contiguous range of
// tiles instead, so neighbouring tiles -- which share 128-byte lines when a
// tile is narrower than a line -- meet in the same L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t G) {
    const int64_t q = G >> 3, r = G & 7, x = b & 7, k = b >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

struct TileLines {
    int64_t row[64];
    int64_t line[64];
    int valid[64];
};

// Real-signal packing (Real.hs:49-107).
// Untangled full spectrum at index e of the half-length spectrum Z:
// S_e = ev - od, S_{e+L} = ev + od (toSpectrum, Real.hs:53-70).
template <typename T>
__device__ __forceinline__ void untangle(typename Cx<T>::C zk, typename Cx<T>::C zr, typename Cx<T>::W w,
                                         bool plain, typename Cx<T>::C &lo, typename Cx<T>::C &hi) {
    using C = typename Cx<T>::C;
    const C xp = cscale(cadd(zk, zr), T(0.5)), xm = cscale(csub(zk, zr), T(0.5));
    const C ev = mk(xp.x, xm.y);
    const C t = twmul(mk(xm.x, xp.y), w, T(-1), plain);
    const C od = mk(-t.y, t.x);
    lo = csub(ev, od);
    hi = cadd(ev, od);
}

// fromSpectrum's entangling (Real.hs:96-107): z_e = (p0 + p1) + i conj(W_{2L}^e) (p0 - p1)
template <typename T>
__device__ __forceinline__ typename Cx<T>::C entangle(typename Cx<T>::C p0, typename Cx<T>::C p1,
                                                      typename Cx<T>::W w, bool plain) {
    const auto t = twmul(csub(p0, p1), w, T(1), plain);
    return cadd(cadd(p0, p1), mk(-t.y, t.x));
}

// The spectrum product of the real convolution on one pair (e, r = L - e) of
// half-length spectra, untangle -> product -> entangle in closed form.  With
// w = W_{2L}^e, the untangled values are S_e = E + w O, S_{e+L} = E - w O for
// E = (Z_e + conj Z_r) / 2, D = (Z_e - conj Z_r) / 2, O = -i D; putting the
// products of two such spectra through the entangling gives
//   z_e = 2 A_e B_e - c (A_e - conj A_r)(B_e - conj B_r),   c = (1 + w^2) / 2,
//   z_r = 2 A_r B_r - conj of the same correction,
// and with the second spectrum conjugated (the adjoint's correlations)
//   z_e = 2 A_e conj B_e - conj(c) (A_e + conj A_r) conj(B_e - conj B_r),
//   z_r = 2 A_r conj B_r + conj of that correction:
// four complex products and one by c per pair instead of three twiddle
// products and two spectrum products per element.  c = (cos^2, cos sin) of
// w's angle, formed from both parts of the table entry.
__device__ __forceinline__ float2 half_one_plus_sq(float4 w, bool plain) {
    if (plain) return make_float2(w.x * w.x, w.x * w.y);
    return make_float2(fmaf(w.x, w.x, 2.0f * w.x * w.z), fmaf(w.x, w.y, fmaf(w.x, w.w, w.z * w.y)));
}
__device__ __forceinline__ double2 half_one_plus_sq(double2 w, bool) { return make_double2(w.x * w.x, w.x * w.y); }

template <typename T>
__device__ __forceinline__ void rconv_pair(typename Cx<T>::C ae, typename Cx<T>::C ar, typename Cx<T>::C be,
                                           typename Cx<T>::C br, typename Cx<T>::W w, bool plain, bool conj_b,
                                           typename Cx<T>::C &ze, typename Cx<T>::C &zr) {
    using C = typename Cx<T>::C;
    const C c = half_one_plus_sq(w, plain);
    const C db = csub(be, cconj(br));
    if (!conj_b) {
        const C q = cmul(c, cmul(csub(ae, cconj(ar)), db));
        ze = csub(cscale(cmul(ae, be), T(2)), q);
        zr = csub(cscale(cmul(ar, br), T(2)), cconj(q));
    } else {
        const C q = cmul(cconj(c), cmul(cadd(ae, cconj(ar)), cconj(db)));
        ze = csub(cscale(cmul(ae, cconj(be)), T(2)), q);
        zr = cadd(cscale(cmul(ar, cconj(br)), T(2)), cconj(q));
    }
}

// ---------------------------------------------------------------------------
// One tile of a pass: load -> stages -> [spectrum product -> inverse stages]
// -> store.
// ---------------------------------------------------------------------------
// D is AfPass, or AfPass in the kernel-argument (constant) address space.
template <typename T, bool FULL, typename D>
__device__ __forceinline__ void pass_tile(const D &d, const int64_t tile, typename Cx<T>::C *lds, TileLines &tl,
                          typename Cx<T>::C *s_x0, typename Cx<T>::C *s_dc) {
    using C = typename Cx<T>::C;
    using W = typename Cx<T>::W;
    const int tid = threadIdx.x, NT = blockDim.x;
    const int B = uni(d.B), L = uni(d.L), ld = uni(d.ld), ps = uni(d.padshift);
    const int32_t flags = FULL ? uni(d.flags) : (uni(d.flags) & ~kRarePassFlags);
    const bool plain = (flags & AFP_PLAIN_TW) != 0;
    const int64_t nlines_total = d.nrows * d.nlines;

    if (tid < B) {
        const int64_t gl = tile * B + tid;
        const int v = gl < nlines_total;
        tl.valid[tid] = v;
        tl.row[tid] = v ? gl / d.nlines : 0;
        tl.line[tid] = v ? gl - (gl / d.nlines) * d.nlines : 0;
        if (v && (flags & AFP_RADER)) s_x0[tid] = ld_c<T>(d.in_re, d.in_im, tl.row[tid] * d.in_bs);
        if (v && (flags & AFP_X0_CAPTURE) && tl.line[tid] == 0)
            gst<C>(d.x0_out, tl.row[tid], ld_c<T>(d.in_re, d.in_im, tl.row[tid] * d.in_bs));
    }
    __syncthreads();

    // ---- load phase ----
    // A thread holds at most EPT elements of the tile (B L <= EPT * blockDim.x
    // for every launch the planner makes; a remainder loop covers any other):
    // its loads are issued a batch of elements at a time, before the batch's
    // LDS writes, instead of one load latency per element.
    constexpr int EPT = pass_ept(sizeof(T) == 8);
    // elements whose loads are in flight together: as many as fit the 128
    // registers of a 1024-thread workgroup beside the rest of the kernel
    constexpr int kLoadBatch = sizeof(T) == 8 ? 4 : 8, kStoreBatch = sizeof(T) == 8 ? 2 : 4;
    static_assert(EPT % kLoadBatch == 0 && EPT % kStoreBatch == 0, "batches divide EPT");
    const int total = B * L;
    const float rB = 1.0f / (float)B, rL = 1.0f / (float)L;
    const bool in_lf = uni(d.in_lines_fast) != 0;
    const int32_t *const gt_in = uni(d.gt_in), *const in_map = uni(d.in_map);
    const void *const in_re = uni(d.in_re), *const in_im = uni(d.in_im), *const in_mul = uni(d.in_mul);
    const int64_t in_bs = uni(d.in_bs), in_len = uni(d.in_len);
    const uint32_t in_ls = (uint32_t)uni((int32_t)d.in_ls), in_es = (uint32_t)uni((int32_t)d.in_es);
    auto load_elem = [&](const int idx, int &li) -> C {
        int b, e;
        if (in_lf) divmod_small(idx, B, rB, e, b);
        else divmod_small(idx, L, rL, b, e);
        li = pidx(b * ld + (gt_in ? gld<int32_t>(gt_in, e) : e), ps);
        C v = mk(T(0), T(0));
        if (tl.valid[b]) {
            const int64_t r = tl.row[b], l = tl.line[b];
            const int64_t rb = r * in_bs;
            if (flags & AFP_REAL_ENTANGLE) {
                // fromSpectrum (Real.hs:96-107): z = (S_e + S_{e+L}) + i W_{2L}^{+e} (S_e - S_{e+L})
                const C p0 = ld_c<T>(d.in_re, d.in_im, rb + e);
                const C p1 = ld_c<T>(d.in_re, d.in_im, rb + e + L);
                const C t = twmul(csub(p0, p1), gld<W>(d.aux_tw, e), T(1), plain);
                v = cadd(cadd(p0, p1), mk(-t.y, t.x));
            } else if (flags & AFP_UNTANGLE_ADJ) {
                // Adjoint of toSpectrum: x -> F(x) maps reals to the full spectrum, so
                // its adjoint is G -> Re(F^H G) = fromSpectrum(H) with H the Hermitian
                // part (G_k + conj G_{-k}) / 2, which makes F^H H real.  Load H and
                // entangle it exactly as fromSpectrum does.
                const int er = e ? L - e : 0;
                const C g0 = ld_c<T>(d.in_re, d.in_im, rb + e);
                const C g1 = ld_c<T>(d.in_re, d.in_im, rb + e + L);
                const C h0 = ld_c<T>(d.in_re, d.in_im, rb + er + L);
                const C h1 = ld_c<T>(d.in_re, d.in_im, rb + er);
                // Hermitian part of G at e and e+L: (G_e + conj G_{-e}) / 2, where -e of e is
                // (2L - e) mod 2L, i.e. L + er for e in [0, L) (er = (L-e) mod L) unless e = 0.
                C q0, q1;
                if (e == 0) {
                    q0 = mk(g0.x, T(0));            // (G_0 + conj G_0)/2
                    q1 = mk(g1.x, T(0));            // (G_L + conj G_L)/2
                } else {
                    q0 = cscale(cadd(g0, cconj(h0)), T(0.5));
                    q1 = cscale(cadd(g1, cconj(h1)), T(0.5));
                }
                const C t = twmul(csub(q0, q1), gld<W>(d.aux_tw, e), T(1), plain);
                v = cadd(cadd(q0, q1), mk(-t.y, t.x));
            } else {
                // positions inside a row stay below 2^31 (lengths <= 2^30)
                const int64_t pos = (int64_t)((uint32_t)l * in_ls + (uint32_t)e * in_es);
                if (pos < in_len) {
                    const int64_t a = in_map ? (int64_t)gld<int32_t>(in_map, pos) : pos;
                    v = ld_c<T>(in_re, in_im, rb + a);
                    if (in_mul) v = cmul(v, gld<C>(in_mul, pos));
                }
            }
        }
        return v;
    };
#pragma unroll 1
    for (int k0 = 0; k0 < EPT; k0 += kLoadBatch) {
        C vb[kLoadBatch];
        int lb[kLoadBatch];
#pragma unroll
        for (int k = 0; k < kLoadBatch; ++k) {
            lb[k] = -1;
            if (tid + (k0 + k) * NT < total) vb[k] = load_elem(tid + (k0 + k) * NT, lb[k]);
        }
#pragma unroll
        for (int k = 0; k < kLoadBatch; ++k)
            if (lb[k] >= 0) lds[lb[k]] = vb[k];
    }
    {
        for (int idx = tid + EPT * NT; idx < total; idx += NT) {
            int li;
            const C v = load_elem(idx, li);
            lds[li] = v;
        }
    }
    __syncthreads();

    // ---- first stage sequence ----
    const T s1 = T(d.sign);
    run_stages<T>(lds, d, s1);

    // ---- convolution: spectrum product and inverse stages ----
    if (flags & AFP_CONV) {
        const void *const spec = uni(d.spec);
        const int64_t spec_bs = uni(d.spec_bs);
        const uint32_t spec_ls = (uint32_t)uni((int32_t)d.spec_ls);
        auto spec_elem = [&](const int idx) {
            int b, e;
            divmod_small(idx, L, rL, b, e);
            if (!tl.valid[b]) return;
            const int64_t r = tl.row[b], l = tl.line[b];
            const int li = pidx(b * ld + e, ps);
            const int64_t si = r * spec_bs + (int64_t)((uint32_t)l * spec_ls + (uint32_t)e);
            C v = lds[li];
            if (l == 0 && e == 0) {
                if (flags & AFP_RADER) s_dc[b] = v;
                if (flags & AFP_DC_CAPTURE) gst<C>(d.dc_out, r, v);
            }
            if (flags & AFP_SPEC_PLAIN) lds[li] = cmul(v, gld<C>(spec, si));
            else lds[li] = vmul(v, gld<W>(spec, si), plain);
        };
        // unrolled in batches: the spectrum loads (global) do not wait for the
        // LDS writes
#pragma unroll 1
        for (int k0 = 0; k0 < EPT; k0 += kLoadBatch)
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k)
                if (tid + (k0 + k) * NT < total) spec_elem(tid + (k0 + k) * NT);
        for (int idx = tid + EPT * NT; idx < total; idx += NT) spec_elem(idx);
        __syncthreads();
        run_stages<T>(lds, d, -s1);
    }

    // ---- store phase ----
    if (flags & AFP_REAL_UNTANGLE) {
        // toSpectrum (Real.hs:53-70) on the half-length spectrum Z of this line
        for (int idx = tid; idx < total; idx += NT) {
            int b, e;
            divmod_small(idx, L, rL, b, e);
            if (!tl.valid[b]) continue;
            const int64_t ob = tl.row[b] * d.out_bs;
            const C zk = lds[pidx(b * ld + e, ps)];
            const C zr = lds[pidx(b * ld + (e ? L - e : 0), ps)];
            const C xp = cscale(cadd(zk, zr), T(0.5)), xm = cscale(csub(zk, zr), T(0.5));
            const C ev = mk(xp.x, xm.y);
            // odds = i W_{2L}^e (xm.re + i xp.im) with the forward-sign W_{2L}^e
            const C t = twmul(mk(xm.x, xp.y), gld<W>(d.aux_tw, e), T(-1), plain);
            const C od = mk(-t.y, t.x);
            C lo = csub(ev, od), hi = cadd(ev, od);
            if (flags & AFP_SCALE) { lo = cscale(lo, T(d.scale)); hi = cscale(hi, T(d.scale)); }
            st_c<T>(d.out_re, d.out_im, ob + e, lo);
            st_c<T>(d.out_re, d.out_im, ob + e + L, hi);
        }
        __syncthreads();
        return;
    }
    if (flags & AFP_ENTANGLE_ADJ) {
        // Adjoint of fromSpectrum's entangling z_e = S_e (1 + c_e) + S_{e+L} (1 - c_e),
        // c_e = i conj(W_{2L}^e): from the half-length cotangent U (after the
        // opposite-sign sub-transform) dS_e = U_e + conj(c_e) U_e and
        // dS_{e+L} = U_e - conj(c_e) U_e, with conj(c_e) = -i W_{2L}^e.
        for (int idx = tid; idx < total; idx += NT) {
            int b, e;
            divmod_small(idx, L, rL, b, e);
            if (!tl.valid[b]) continue;
            const int64_t ob = tl.row[b] * d.out_bs;
            const C u = lds[pidx(b * ld + e, ps)];
            // c = conj(i W_{2L}^{+e}) * u = -i W_{2L}^{-e} u
            const C t = twmul(u, gld<W>(d.aux_tw, e), T(-1), plain);
            const C c = mk(t.y, -t.x);
            st_c<T>(d.out_re, d.out_im, ob + e, cadd(u, c));
            st_c<T>(d.out_re, d.out_im, ob + e + L, csub(u, c));
        }
        __syncthreads();
        return;
    }
    const bool rader_any = (flags & (AFP_RADER | AFP_RADER_OUT)) != 0;
    const bool out_lf = uni(d.out_lines_fast) != 0;
    const int32_t *const gt_out = uni(d.gt_out), *const out_map = uni(d.out_map);
    const void *const out_mul = uni(d.out_mul);
    void *const out_re = uni(d.out_re), *const out_im = uni(d.out_im);
    const int64_t out_bs = uni(d.out_bs), out_len = uni(d.out_len);
    const uint32_t out_ls = (uint32_t)uni((int32_t)d.out_ls), out_es = (uint32_t)uni((int32_t)d.out_es);
    const T scale = T(uni(d.scale));
    FtTables ft;
    ft.ft_coarse = uni(d.ft_coarse);
    ft.ft_fine = uni(d.ft_fine);
    ft.ftShift = uni(d.ftShift);
    ft.ftSign = uni(d.ftSign);
    // the value and address of output element idx (address -1: not written)
    auto store_elem = [&](const int idx, int64_t &ga) -> C {
        int b, e;
        if (out_lf) divmod_small(idx, B, rB, e, b);
        else divmod_small(idx, L, rL, b, e);
        ga = -1;
        if (!tl.valid[b]) return mk(T(0), T(0));
        const int64_t r = tl.row[b], l = tl.line[b];
        const int64_t o = (int64_t)((uint32_t)l * out_ls + (uint32_t)e * out_es);
        if (o >= out_len) return mk(T(0), T(0));
        C v = lds[pidx(b * ld + (gt_out ? gld<int32_t>(gt_out, e) : e), ps)];
        if (flags & AFP_POSTTW) {
            // the lines of a four-step pass number N / L, so l * e < N: no reduction
            v = cmul(v, ft_twiddle_t(ft, (uint64_t)((uint32_t)l * (uint32_t)e), T(0)));
        }
        if (out_mul) v = cmul(v, gld<C>(out_mul, o));
        if (flags & AFP_SCALE) v = cscale(v, scale);
        if (rader_any) {
            C x0;
            if (flags & AFP_RADER) x0 = s_x0[b];
            else x0 = ld_c<T>(d.x0_re, d.x0_im, r * d.x0_bs);
            v = cadd(v, x0);
        }
        ga = r * out_bs + (out_map ? (int64_t)gld<int32_t>(out_map, o) : o);
        return v;
    };
    // a batch of elements' table loads before the first global store of the
    // batch (the compiler cannot move a load above a store that may alias it)
#pragma unroll 1
    for (int k0 = 0; k0 < EPT; k0 += kStoreBatch) {
        C vb[kStoreBatch];
        int64_t gb[kStoreBatch];
#pragma unroll
        for (int k = 0; k < kStoreBatch; ++k) {
            gb[k] = -1;
            if (tid + (k0 + k) * NT < total) vb[k] = store_elem(tid + (k0 + k) * NT, gb[k]);
        }
#pragma unroll
        for (int k = 0; k < kStoreBatch; ++k)
            if (gb[k] >= 0) st_c<T>(out_re, out_im, gb[k], vb[k]);
    }
    {
        for (int idx = tid + EPT * NT; idx < total; idx += NT) {
            int64_t ga;
            const C v = store_elem(idx, ga);
            if (ga >= 0) st_c<T>(out_re, out_im, ga, v);
        }
    }
    if (rader_any && tid < B && tl.valid[tid] && tl.line[tid] == 0) {
        // X[0] = x0 + sum of the other inputs = x0 + DC of the permuted sequence
        const int64_t r = tl.row[tid];
        C x0, dc;
        if (flags & AFP_RADER) { x0 = s_x0[tid]; dc = s_dc[tid]; }
        else { x0 = ld_c<T>(d.x0_re, d.x0_im, r * d.x0_bs); dc = gld<C>(d.dc_in, r); }
        st_c<T>(d.out_re, d.out_im, r * d.out_bs, cadd(x0, dc));
    }
    __syncthreads();
}

}  // namespace af
