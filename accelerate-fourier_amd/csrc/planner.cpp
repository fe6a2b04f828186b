// Host planner of the MI355X transform path; see planner.hpp.
#include "planner.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "accelerate_fourier.h"

namespace af {

int64_t ceil_pow2(int64_t n) {
    if (n <= 1) return 1;
    int64_t m = 1;
    while (m < n) m <<= 1;
    return m;
}

// ceilingSmooths a b n m: trade one factor a of m for a factor b while the
// product stays >= n (NumberTheory.hs:118-124).
static void ceiling_smooths(int64_t a, int64_t b, int64_t n, int64_t m, std::vector<int64_t> &out) {
    for (;;) {
        out.push_back(m);
        if (m % a != 0) break;
        int64_t k = m * b;
        while (k % a == 0 && k / a >= n) k /= a;
        m = k;
    }
}

int64_t ceil_5smooth(int64_t n) {
    std::vector<int64_t> l3, l5;
    ceiling_smooths(2, 3, n, ceil_pow2(n), l3);
    int64_t best = -1;
    for (int64_t m3 : l3) {
        l5.clear();
        ceiling_smooths(2, 5, n, m3, l5);
        for (int64_t v : l5)
            if (best < 0 || v < best) best = v;
    }
    return best;
}

static bool is_7smooth(int64_t m) {
    for (int p : {2, 3, 5, 7})
        while (m % p == 0) m /= p;
    return m == 1;
}

int64_t ceil_7smooth(int64_t n) {
    int64_t m = n < 1 ? 1 : n;
    while (!is_7smooth(m)) ++m;
    return m;
}

bool is_prime(int64_t n) {
    if (n < 2) return false;
    if (n % 2 == 0) return n == 2;
    for (int64_t d = 3; d * d <= n; d += 2)
        if (n % d == 0) return false;
    return true;
}

int64_t mod_pow(int64_t base, int64_t expo, int64_t m) {
    unsigned __int128 r = 1, b = (uint64_t)(base % m);
    while (expo > 0) {
        if (expo & 1) r = (r * b) % (uint64_t)m;
        b = (b * b) % (uint64_t)m;
        expo >>= 1;
    }
    return (int64_t)r;
}

int64_t multiplicative_generator(int64_t p) {
    const int64_t order = p - 1;
    std::vector<int64_t> facs;
    int64_t m = order;
    for (int64_t q = 2; q * q <= m; ++q)
        if (m % q == 0) {
            facs.push_back(q);
            while (m % q == 0) m /= q;
        }
    if (m > 1) facs.push_back(m);
    for (int64_t g = 1; g < p; ++g) {
        if (mod_pow(g, order, p) != 1) continue;
        bool ok = true;
        for (int64_t q : facs)
            if (mod_pow(g, order / q, p) == 1) { ok = false; break; }
        if (ok) return g;
    }
    return -1;
}

bool stockham_radices(int64_t n, std::vector<int> &radices) {
    radices.clear();
    if (n < 2) return n == 1;
    int e2 = 0;
    while (n % 2 == 0) { n /= 2; ++e2; }
    std::vector<int> odd;
    for (int p : {13, 11, 7, 5, 3})
        while (n % p == 0) { n /= p; odd.push_back(p); }
    if (n != 1) return false;
    for (int i = 0; i < e2 / 4; ++i) radices.push_back(16);
    if (e2 % 4) radices.push_back(1 << (e2 % 4));
    for (int p : odd) radices.push_back(p);
    return true;
}

int elem_bytes(int prec) { return prec == AF_F64 ? 16 : 8; }
int64_t lds_max_len(int prec) { return prec == AF_F64 ? 8192 : 16384; }
// Dynamic LDS budget of one tile of the generic pass: 160 KiB per workgroup
// minus 4 KiB for that kernel's static arrays (its line table and Rader slots,
// 3328 bytes in FP64).
int lds_tile_bytes_cap() { return 156 * 1024; }

static bool three_pass(int64_t N, int64_t lmax, CoreShape &c) {
    int64_t a = 0, b = 0, cc = 0;
    if ((N & (N - 1)) == 0) {
        // 2^k = 2^ka 2^kb 2^kc, balanced, the spare bits on the contiguous rows
        int k = 0;
        while (((int64_t)1 << k) < N) ++k;
        const int kb = k / 3, kc = k / 3, ka = k - kb - kc;
        a = (int64_t)1 << ka;
        b = (int64_t)1 << kb;
        cc = (int64_t)1 << kc;
        if (a > lmax) return false;
    } else {
        // the most balanced smooth factorisation within the LDS limit
        std::vector<int64_t> dv;
        for (int64_t d = 2; d <= lmax && d * d <= N * lmax; ++d)
            if (N % d == 0) dv.push_back(d);
        double best = 1e300;
        for (int64_t x : dv)
            for (int64_t y : dv) {
                if ((N / x) % y) continue;
                const int64_t z = N / x / y;
                if (z < 2 || z > lmax) continue;
                std::vector<int> t;
                if (!stockham_radices(x, t) || !stockham_radices(y, t) || !stockham_radices(z, t)) continue;
                const double lx = std::log((double)x), ly = std::log((double)y), lz = std::log((double)z);
                const double cost = std::max(lx, std::max(ly, lz)) - std::min(lx, std::min(ly, lz));
                if (cost < best - 1e-12) { best = cost; a = z; b = y; cc = x; }
            }
        if (!a) return false;
    }
    c.kind = CORE_THREE;
    c.N1 = a;
    c.N2 = b;
    c.N3 = cc;
    stockham_radices(a, c.r1);
    stockham_radices(b, c.r2);
    stockham_radices(cc, c.r3);
    return true;
}

CoreShape core_shape(int prec, int64_t N, bool allow_three) {
    CoreShape c;
    c.N = N;
    std::vector<int> r;
    if (!stockham_radices(N, r)) return c;
    const int64_t lmax = lds_max_len(prec);
    if (N <= lmax) {
        c.kind = CORE_SINGLE;
        c.r = r;
        return c;
    }
    if (allow_three && (N & (N - 1)) == 0 && N >= ((int64_t)1 << 24) && three_pass(N, lmax, c)) return c;
    // four-step: the most balanced N1 * N2 with both factors within the LDS limit
    int64_t best1 = 0;
    double bestcost = 1e300;
    for (int64_t d = 2; d <= lmax; ++d) {
        if (N % d) continue;
        const int64_t e = N / d;
        if (e > lmax || e < 2) continue;
        std::vector<int> r1, r2;
        if (!stockham_radices(d, r1) || !stockham_radices(e, r2)) continue;
        const double cost = std::fabs(std::log((double)d) - std::log((double)e));
        if (cost < bestcost - 1e-12) { bestcost = cost; best1 = d; }
    }
    if (!best1) {
        if (allow_three) three_pass(N, lmax, c);
        return c;
    }
    c.kind = CORE_FOUR;
    c.N1 = best1;
    c.N2 = N / best1;
    stockham_radices(c.N1, c.r1);
    stockham_radices(c.N2, c.r2);
    return c;
}

bool plan_shape(int prec, int64_t n, int algorithm, PlanShape &out, std::string &why, bool allow_three) {
    out = PlanShape();
    out.n = n;
    if (n == 0) { out.algo = ALG_EMPTY; return true; }
    if (n == 1) { out.algo = ALG_COPY; return true; }
    const bool chirp_only = algorithm == AF_ALG_CHIRP2 || algorithm == AF_ALG_CHIRP235;
    if (!chirp_only) {
        CoreShape c = core_shape(prec, n, allow_three);
        if (c.kind != CORE_NONE) {
            out.algo = ALG_STOCKHAM;
            out.core = c;
            return true;
        }
        if (is_prime(n)) {
            CoreShape cr = core_shape(prec, n - 1);
            if (cr.kind != CORE_NONE) {
                out.algo = ALG_RADER;
                out.core = cr;
                return true;
            }
        }
        if (algorithm == AF_ALG_DECOMPOSE) {
            // transformDecompose never pads: compose sub-plans on the host
            out.algo = ALG_DECOMP;
            if (is_prime(n)) {
                out.dec = DEC_RADER;
                out.da = n - 1;
                return true;
            }
            // the largest prime factor q and its multiplicity: Good-Thomas over
            // (n / q^e, q^e) when that is a coprime split, else Cooley-Tukey (q, n / q)
            int64_t m = n, q = 1;
            for (int64_t f = 2; f * f <= m; ++f)
                while (m % f == 0) {
                    q = f;
                    m /= f;
                }
            if (m > 1) q = std::max(q, m);
            int64_t qe = 1;
            while (n % (qe * q) == 0) qe *= q;
            if (qe < n) {
                out.dec = DEC_GOOD_THOMAS;
                out.da = n / qe;
                out.db = qe;
            } else {
                out.dec = DEC_COOLEY_TUKEY;
                out.da = q;
                out.db = n / q;
            }
            return true;
        }
    }
    int64_t M;
    if (algorithm == AF_ALG_CHIRP2) {
        M = ceil_pow2(2 * n - 1);
    } else if (algorithm == AF_ALG_CHIRP235) {
        M = ceil_5smooth(2 * n - 1);
    } else {
        // the smallest 7-smooth pad, except where a power of two gives a four-step
        // core of power-of-two factors 256 ... 2048: its three convolution passes
        // run on compile-time stage kernels, several times faster per element than
        // the mixed-radix ones, which outweighs the up-to-2x longer pad
        M = ceil_7smooth(2 * n - 1);
        const int64_t M2 = ceil_pow2(2 * n - 1);
        if (M2 >= ((int64_t)1 << 16) && M2 <= ((int64_t)1 << 22)) M = M2;
    }
    CoreShape cb;
    if (algorithm == AF_ALG_AUTO && (M & (M - 1)) == 0 && M >= ((int64_t)1 << 16) && M <= ((int64_t)1 << 22)) {
        // power-of-two convolution core: short columns (N1 = 256 or 512 points,
        // several workgroups per CU) and rows of up to 4096 points, which the
        // convolution's middle pass reads contiguously
        cb.N = M;
        cb.kind = CORE_FOUR;
        cb.N1 = std::max<int64_t>(256, M / 4096);
        cb.N2 = M / cb.N1;
        stockham_radices(cb.N1, cb.r1);
        stockham_radices(cb.N2, cb.r2);
    } else {
        cb = core_shape(prec, M);
    }
    if (cb.kind == CORE_NONE) {
        why = "padded length does not factor within the LDS limit";
        return false;
    }
    out.algo = ALG_BLUESTEIN;
    out.M = M;
    out.core = cb;
    return true;
}

static std::string radix_str(const std::vector<int> &r) {
    std::string s = "[";
    for (size_t i = 0; i < r.size(); ++i) s += (i ? "," : "") + std::to_string(r[i]);
    return s + "]";
}

static std::string core_str(const CoreShape &c) {
    if (c.kind == CORE_SINGLE) return "stockham" + std::to_string(c.N) + radix_str(c.r);
    if (c.kind == CORE_FOUR)
        return "fourstep" + std::to_string(c.N) + "(" + std::to_string(c.N1) + radix_str(c.r1) + "x" +
               std::to_string(c.N2) + radix_str(c.r2) + ")";
    if (c.kind == CORE_THREE)
        return "threepass" + std::to_string(c.N) + "(" + std::to_string(c.N1) + radix_str(c.r1) + "x" +
               std::to_string(c.N2) + radix_str(c.r2) + "x" + std::to_string(c.N3) + radix_str(c.r3) + ")";
    return "none";
}

std::string describe_shape(const PlanShape &s) {
    switch (s.algo) {
    case ALG_EMPTY: return "empty";
    case ALG_COPY: return "copy1";
    case ALG_STOCKHAM: return core_str(s.core);
    case ALG_RADER: return "rader" + std::to_string(s.n) + "(" + core_str(s.core) + ")";
    case ALG_BLUESTEIN: return "bluestein" + std::to_string(s.n) + "(" + core_str(s.core) + ")";
    case ALG_DECOMP: {
        // the sub-plans, described the same way (decompositions of the factors)
        auto sub = [&](int64_t m) {
            PlanShape t;
            std::string w;
            return plan_shape(AF_F64, m, AF_ALG_DECOMPOSE, t, w) ? describe_shape(t) : std::string("?");
        };
        if (s.dec == DEC_RADER) return "rader" + std::to_string(s.n) + "(" + sub(s.da) + ")";
        return std::string(s.dec == DEC_GOOD_THOMAS ? "goodthomas" : "cooleytukey") + std::to_string(s.n) + "(" +
               sub(s.da) + "," + sub(s.db) + ")";
    }
    }
    return "?";
}

}  // namespace af
