// C ABI of the MI355X transform path (include/accelerate_fourier.h).
//
// Plans hold the host planner's decision (planner.hpp) and the device tables
// generated in FP64 on the GPU at creation.  Execution turns a plan plus the
// caller's pointers into one to three AfPass descriptors and launches them on
// the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>
#include <numeric>

#include "accelerate_fourier.h"
#include "kernels.hpp"
#include "planner.hpp"

using namespace af;

namespace {

// text of the last failed HIP call on this thread (af_last_error)
thread_local char g_last_error[256] = "";

af_status hip_status(hipError_t e, const char *call, int line) {
    std::snprintf(g_last_error, sizeof g_last_error, "%s failed at abi.cpp:%d: %s", call, line, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? AF_OUT_OF_MEMORY : AF_HIP_ERROR;
}

#define HIPCHK(x)                                                 \
    do {                                                          \
        hipError_t e__ = (x);                                     \
        if (e__ != hipSuccess) return hip_status(e__, #x, __LINE__); \
    } while (0)
// status of a HIP call, recording the failed call's text
af_status hip_try(hipError_t e, const char *call, int line) { return e == hipSuccess ? AF_SUCCESS : hip_status(e, call, line); }
#define HIPST(x) hip_try((x), #x, __LINE__)
#define AFCHK(x)                                 \
    do {                                         \
        af_status s__ = (x);                     \
        if (s__ != AF_SUCCESS) return s__;       \
    } while (0)

std::once_flag g_init_flag;
af_status g_init_status = AF_SUCCESS;

}  // namespace

// Scratch of af_execute_noalloc: one fixed buffer per GPU, in the code
// object's zero-initialised data (no file space), held until the process
// ends.  Host code never passes `scratch` itself to a kernel: its host-side
// name is only a handle; the buffer's device address comes from
// hipGetSymbolAddress, once per device (noalloc_scratch).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call for that device (device_init) and kept.
af_status noalloc_scratch(int dev, void **ptr) {
    if (dev < 0 || dev >= kMaxDevices) return AF_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return AF_INTERNAL;
        g_noalloc_ptr[dev] = p;
    }
    *ptr = g_noalloc_ptr[dev];
    return AF_SUCCESS;
}

af_status device_init() {
    std::call_once(g_init_flag, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            (void)hipGetLastError();
            g_init_status = AF_NO_DEVICE;
            return;
        }
        g_init_status = AF_SUCCESS;
    });
    if (g_init_status != AF_SUCCESS) return g_init_status;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return AF_UNSUPPORTED;
    // set-up of the allocation-free entry point's buffer on this device
    void *p = nullptr;
    return noalloc_scratch(dev, &p);
}

// Device buffer owned by a plan.
struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct Owner {
    std::vector<std::unique_ptr<DevBuf>> bufs;
    af_status alloc(void **out, size_t bytes) {
        auto b = std::make_unique<DevBuf>();
        if (bytes == 0) bytes = 16;
        HIPCHK(hipMalloc(&b->p, bytes));
        *out = b->p;
        bufs.push_back(std::move(b));
        return AF_SUCCESS;
    }
};

int tw_bytes(int) { return 16; }  // float4 hi/lo or double2

// Device tables of one FFT core.
struct Core {
    CoreShape sh;
    void *tw = nullptr;                 // single: W_N^m, m < N
    void *tw1 = nullptr, *tw2 = nullptr;  // four-step: W_N1, W_N2
    void *ft_coarse = nullptr, *ft_fine = nullptr;  // four-step twiddles W_N, table format
    int ftShift = 0;
    // three passes: W_N3 stage table and the second pass's twiddles W_{N1 N2}
    void *tw3 = nullptr, *ft2_coarse = nullptr, *ft2_fine = nullptr;
    int ft2Shift = 0;
    // FP64 copies (the same tables for FP64 plans)
    void *tw_d = nullptr, *tw1_d = nullptr, *tw2_d = nullptr, *ft_coarse_d = nullptr, *ft_fine_d = nullptr;
    void *tw3_d = nullptr, *ft2_coarse_d = nullptr, *ft2_fine_d = nullptr;
    // single pass in prime-factor order (AfPass::gt_in, gt_out, grp_mask)
    int32_t *gt_in = nullptr, *gt_out = nullptr;
    int32_t gt_mask = 0;
};

// The prime-factor order of a single-pass line of length L whose radices n_1
// ... n_G (stage order) are pairwise coprime: stage g consumes the input digit
// of weight S_g = n_{g+1} ... n_G and produces the output digit of weight
// n_1 ... n_{g-1} (the Stockham stages leave the output in natural order of
// that mixed radix), with no twiddles between the factors.  Input element e
// enters at position sum_g S_g ((e mod n_g) (L/n_g)^-1 mod n_g): the inverse of
// the skew-grid map sum_g (L/n_g) j_g mod L (permuteSkewGrid,
// Planned.hs:597-606, for any number of factors); output element e is read
// from position sum_g (e mod n_g) n_1 ... n_{g-1}, the Chinese-remainder map
// (permuteSkewGridInv, 608-617).  False unless there are at least three
// radices and no two of them share a prime: a repeated prime needs twiddles
// inside its factor, and such lengths keep the Cooley-Tukey order.
// Three radices at least: with two, the maps cost more than the one twiddled
// stage they remove (measured on MI355X: 15 points 85.6 against 80.6 us per
// 32 MiB; from three on prime-factor order is 2-11% faster).
bool prime_factor_radices(const std::vector<int> &radices) {
    if (radices.size() < 3) return false;
    for (size_t a = 0; a < radices.size(); ++a)
        for (size_t b = a + 1; b < radices.size(); ++b)
            if (std::gcd(radices[a], radices[b]) != 1) return false;
    return true;
}

bool prime_factor_maps(int64_t L, const std::vector<int> &radices, std::vector<int32_t> &gin,
                       std::vector<int32_t> &gout, int32_t &mask) {
    if (!prime_factor_radices(radices)) {
        mask = 0;
        return false;
    }
    auto prime_of = [](int r) {
        for (int q : {2, 3, 5, 7, 11, 13})
            if (r % q == 0) return q;
        return r;
    };
    std::vector<int64_t> n;
    mask = 0;
    int last = 0;
    for (size_t s = 0; s < radices.size(); ++s) {
        const int q = prime_of(radices[s]);
        if (s == 0 || q != last) {
            if (s > 0) mask |= 1 << s;
            n.push_back(radices[s]);
        } else {
            n.back() *= radices[s];
        }
        last = q;
    }
    if (n.size() < 2 || n.size() != radices.size()) {
        mask = 0;
        return false;
    }
    const size_t G = n.size();
    std::vector<int64_t> sin(G), sout(G), inv(G);
    int64_t acc = 1;
    for (size_t g = G; g-- > 0;) {
        sin[g] = acc;
        acc *= n[g];
    }
    acc = 1;
    for (size_t g = 0; g < G; ++g) {
        sout[g] = acc;
        acc *= n[g];
        const int64_t c = (L / n[g]) % n[g];
        int64_t x = 1;
        while (x < n[g] && (c * x) % n[g] != 1) ++x;  // n[g] <= 16: a short search
        if (x >= n[g] && n[g] > 1) {
            mask = 0;
            return false;
        }
        inv[g] = x;
    }
    gin.assign(L, 0);
    gout.assign(L, 0);
    for (int64_t e = 0; e < L; ++e) {
        int64_t qi = 0, qo = 0;
        for (size_t g = 0; g < G; ++g) {
            const int64_t r = e % n[g];
            qi += sin[g] * ((r * inv[g]) % n[g]);
            qo += sout[g] * r;
        }
        gin[e] = (int32_t)qi;
        gout[e] = (int32_t)qo;
    }
    return true;
}

// coarse x fine split of the twiddles W_N^a: 4^shift >= N
int ft_shift_for(int64_t N) {
    int shift = 0;
    while (((int64_t)1 << (2 * shift)) < N) ++shift;
    return shift;
}

af_status build_core(int prec, const CoreShape &sh, Core &c, Owner &own, hipStream_t s, bool prime_factor = false) {
    c.sh = sh;
    if (sh.kind == CORE_SINGLE) {
        AFCHK(own.alloc(&c.tw, (size_t)sh.N * tw_bytes(prec)));
        HIPCHK(gen_twiddles(prec, c.tw, sh.N, sh.N, 1, s));
        std::vector<int32_t> gin, gout;
        if (prime_factor && prime_factor_maps(sh.N, sh.r, gin, gout, c.gt_mask)) {
            AFCHK(own.alloc((void **)&c.gt_in, (size_t)sh.N * 4));
            AFCHK(own.alloc((void **)&c.gt_out, (size_t)sh.N * 4));
            HIPCHK(hipMemcpyAsync(c.gt_in, gin.data(), (size_t)sh.N * 4, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(c.gt_out, gout.data(), (size_t)sh.N * 4, hipMemcpyHostToDevice, s));
            HIPCHK(hipStreamSynchronize(s));  // the host maps go out of scope
        }
    } else if (sh.kind == CORE_FOUR || sh.kind == CORE_THREE) {
        AFCHK(own.alloc(&c.tw1, (size_t)sh.N1 * tw_bytes(prec)));
        AFCHK(own.alloc(&c.tw2, (size_t)sh.N2 * tw_bytes(prec)));
        HIPCHK(gen_twiddles(prec, c.tw1, sh.N1, sh.N1, 1, s));
        HIPCHK(gen_twiddles(prec, c.tw2, sh.N2, sh.N2, 1, s));
        c.ftShift = ft_shift_for(sh.N);
        const int64_t T = (int64_t)1 << c.ftShift, nc = (sh.N + T - 1) / T;
        AFCHK(own.alloc(&c.ft_coarse, (size_t)nc * 16));
        AFCHK(own.alloc(&c.ft_fine, (size_t)T * 16));
        HIPCHK(gen_twiddles(prec, c.ft_coarse, sh.N, nc, T, s));
        HIPCHK(gen_twiddles(prec, c.ft_fine, sh.N, T, 1, s));
        if (sh.kind == CORE_THREE) {
            AFCHK(own.alloc(&c.tw3, (size_t)sh.N3 * tw_bytes(prec)));
            HIPCHK(gen_twiddles(prec, c.tw3, sh.N3, sh.N3, 1, s));
            const int64_t N12 = sh.N1 * sh.N2;
            c.ft2Shift = ft_shift_for(N12);
            const int64_t T2 = (int64_t)1 << c.ft2Shift, nc2 = (N12 + T2 - 1) / T2;
            AFCHK(own.alloc(&c.ft2_coarse, (size_t)nc2 * 16));
            AFCHK(own.alloc(&c.ft2_fine, (size_t)T2 * 16));
            HIPCHK(gen_twiddles(prec, c.ft2_coarse, N12, nc2, T2, s));
            HIPCHK(gen_twiddles(prec, c.ft2_fine, N12, T2, 1, s));
        }
    }
    if (prec == PREC_F64) {
        c.tw_d = c.tw;
        c.tw1_d = c.tw1;
        c.tw2_d = c.tw2;
        c.ft_coarse_d = c.ft_coarse;
        c.ft_fine_d = c.ft_fine;
        c.tw3_d = c.tw3;
        c.ft2_coarse_d = c.ft2_coarse;
        c.ft2_fine_d = c.ft2_fine;
        return AF_SUCCESS;
    }
    // FP32 plans keep FP64 copies for the power-of-two kernel's twiddle recurrences
    auto dtab = [&](void **out, int64_t N, int64_t count, int64_t mul) -> af_status {
        AFCHK(own.alloc(out, (size_t)count * 16));
        HIPCHK(gen_twiddles_d(*out, N, count, mul, s));
        return AF_SUCCESS;
    };
    if (sh.kind == CORE_SINGLE) {
        AFCHK(dtab(&c.tw_d, sh.N, sh.N, 1));
    } else if (sh.kind == CORE_FOUR || sh.kind == CORE_THREE) {
        const int64_t T = (int64_t)1 << c.ftShift, nc = (sh.N + T - 1) / T;
        AFCHK(dtab(&c.tw1_d, sh.N1, sh.N1, 1));
        AFCHK(dtab(&c.tw2_d, sh.N2, sh.N2, 1));
        AFCHK(dtab(&c.ft_coarse_d, sh.N, nc, T));
        AFCHK(dtab(&c.ft_fine_d, sh.N, T, 1));
        if (sh.kind == CORE_THREE) {
            const int64_t N12 = sh.N1 * sh.N2, T2 = (int64_t)1 << c.ft2Shift, nc2 = (N12 + T2 - 1) / T2;
            AFCHK(dtab(&c.tw3_d, sh.N3, sh.N3, 1));
            AFCHK(dtab(&c.ft2_coarse_d, N12, nc2, T2));
            AFCHK(dtab(&c.ft2_fine_d, N12, T2, 1));
        }
    }
    return AF_SUCCESS;
}

// ---------------------------------------------------------------------------
// Pass construction
// ---------------------------------------------------------------------------

int pidx_h(int i, int ps) { return ps ? i + (i >> ps) : i; }

int lds_bytes_for(int prec, int L, int B, int ps) {
    const int ld = L + (B > 1 ? 1 : 0);
    return (pidx_h(B * ld - 1, ps) + 1) * elem_bytes(prec);
}

bool tile_fits(int prec, int L, int B) {
    if ((int64_t)B * L > 1024 * pass_ept(prec == PREC_F64)) return false;  // 1024 threads
    return lds_bytes_for(prec, L, B, 4) <= lds_tile_bytes_cap();
}

// Lines per tile for passes over contiguous rows.
int choose_B_rows(int prec, int64_t L, int64_t nlines_total) {
    const int64_t target = prec == AF_F64 ? 2048 : 4096;
    int B = 1;
    while (B < 64 && (2 * B) * L <= target && tile_fits(prec, (int)L, 2 * B) &&
           (nlines_total + 2 * B - 1) / (2 * B) >= 512)
        B *= 2;
    return fast_tile_lines(prec, L, B);
}

// Lines per tile for passes whose loads or stores are strided: runs of 128
// bytes of interleaved elements.  Planar layouts keep the same line count
// (64-byte runs per component array): measured faster on MI355X than doubling
// the lines, for both precisions (64 x 65536 and 2^24 transforms).
int choose_B_cols(int prec, int64_t L) {
    int B = prec == AF_F64 ? 8 : 16;
    while (B > 1 && !tile_fits(prec, (int)L, B)) B /= 2;
    return fast_tile_lines(prec, L, B);
}

struct Endpoint {
    const void *re = nullptr;
    const void *im = nullptr;  // nullptr: interleaved
    int64_t bs = 0;            // row stride in elements
};

// The tile geometry of a pass with B lines per tile: LDS pitch, tiles,
// threads of the generic kernel (ceil(B L / EPT), whole wavefronts, at most
// 1024) and its dynamic LDS.
void set_tile(AfPass &d, int prec, int B) {
    d.B = B;
    d.ld = d.L + (B > 1 ? 1 : 0);
    d.ntiles = (d.nrows * d.nlines + B - 1) / B;
    const int ept = pass_ept(prec == PREC_F64);
    int nt = (int)((((int64_t)B * d.L + ept - 1) / ept + 63) / 64 * 64);
    if (nt < 64) nt = 64;
    if (nt > 1024) nt = 1024;
    d.nthreads = nt;
    d.lds_bytes = lds_bytes_for(prec, d.L, B, d.padshift);
}

AfPass base_pass(int prec, int L, const std::vector<int> &radices, const void *tw, const void *tw_d, int B,
                 int64_t nlines, int64_t nrows, int sign, bool plain) {
    AfPass d;
    std::memset(&d, 0, sizeof d);
    d.L = L;
    d.nst = (int)radices.size();
    int p = 1;
    for (int i = 0; i < d.nst; ++i) {
        d.radix[i] = radices[i];
        d.pst[i] = p;
        p *= radices[i];
    }
    d.padshift = 4;
    d.nlines = nlines;
    d.nrows = nrows;
    set_tile(d, prec, B);
    d.tw = tw;
    d.tw_d = tw_d;
    d.sign = sign;
    d.flags = plain ? AFP_PLAIN_TW : 0;
    d.in_len = INT64_MAX;
    d.out_len = INT64_MAX;
    d.scale = 1.0;
    return d;
}

void set_in(AfPass &d, const Endpoint &e, int64_t es, int64_t ls) {
    d.in_re = e.re;
    d.in_im = e.im;
    d.in_bs = e.bs;
    d.in_es = es;
    d.in_ls = ls;
}
void set_out(AfPass &d, const Endpoint &e, int64_t es, int64_t ls) {
    d.out_re = (void *)e.re;
    d.out_im = (void *)e.im;
    d.out_bs = e.bs;
    d.out_es = es;
    d.out_ls = ls;
}

void set_ft(AfPass &d, const Core &c, int sign) {
    d.flags |= AFP_POSTTW;
    d.ftN = c.sh.N;
    d.ftShift = c.ftShift;
    d.ftSign = sign;
    d.ft_coarse = c.ft_coarse;
    d.ft_fine = c.ft_fine;
    d.ft_coarse_d = c.ft_coarse_d;
    d.ft_fine_d = c.ft_fine_d;
}

// Three passes of N = Na Nb Nc (Na = N1, Nb = N2, Nc = N3), input index
// n = a + Na b + Na Nb c, output index k = kc + Nc kb + Nc Nb ka, from
// W_N^{nk} = W_Nc^{c kc} W_N^{(a + Na b) kc} W_Nb^{b kb} W_{Na Nb}^{a kb} W_Na^{a ka}:
//  1. Nc-point transforms over c (stride Na Nb), lines l = a + Na b, then W_N^{l kc};
//  2. Nb-point transforms over b (stride Na) in each block kc, lines a, then
//     W_{Na Nb}^{a kb}, in place;
//  3. Na-point transforms of the contiguous rows, stored transposed to
//     kc + Nc kb + Nc Nb ka (one launch per transform of the batch).
// Every pass has lines of one factor and tiles of neighbouring lines.
void three_passes(int prec, const Core &c, int64_t nrows, int sign, bool plain, const Endpoint &in,
                  const Endpoint &out, void *scratch, std::vector<AfPass> &ps) {
    const int64_t N = c.sh.N, Na = c.sh.N1, Nb = c.sh.N2, Nc = c.sh.N3, Nab = Na * Nb;
    const size_t eb = (size_t)elem_bytes(prec);
    Endpoint sc;
    sc.re = scratch;
    sc.bs = N;
    AfPass p1 = base_pass(prec, (int)Nc, c.sh.r3, c.tw3, c.tw3_d, choose_B_cols(prec, Nc), Nab, nrows,
                          sign, plain);
    set_in(p1, in, Nab, 1);
    set_out(p1, sc, Nab, 1);
    p1.in_lines_fast = p1.out_lines_fast = 1;
    set_ft(p1, c, sign);
    ps.push_back(p1);
    Endpoint sb;
    sb.re = scratch;
    sb.bs = Nab;
    AfPass p2 = base_pass(prec, (int)Nb, c.sh.r2, c.tw2, c.tw2_d, choose_B_cols(prec, Nb), Na, nrows * Nc, sign, plain);
    set_in(p2, sb, Na, 1);
    set_out(p2, sb, Na, 1);
    p2.in_lines_fast = p2.out_lines_fast = 1;
    p2.flags |= AFP_POSTTW;
    p2.ftN = Nab;
    p2.ftShift = c.ft2Shift;
    p2.ftSign = sign;
    p2.ft_coarse = c.ft2_coarse;
    p2.ft_fine = c.ft2_fine;
    p2.ft_coarse_d = c.ft2_coarse_d;
    p2.ft_fine_d = c.ft2_fine_d;
    ps.push_back(p2);
    for (int64_t r = 0; r < nrows; ++r) {
        AfPass p3 = base_pass(prec, (int)Na, c.sh.r1, c.tw1, c.tw1_d, choose_B_cols(prec, Na), Nc, Nb,
                              sign, plain);
        Endpoint si;
        si.re = (const char *)scratch + (size_t)r * N * eb;
        si.bs = Na;
        set_in(p3, si, 1, Nab);
        Endpoint so = out;
        const size_t rb = out.im ? eb / 2 : eb;
        so.re = (const char *)out.re + (size_t)(r * out.bs) * rb;
        if (out.im) so.im = (const char *)out.im + (size_t)(r * out.bs) * rb;
        so.bs = Nc;
        set_out(p3, so, Nc * Nb, 1);
        p3.out_lines_fast = 1;
        ps.push_back(p3);
    }
}

// Plain transform of `nrows` rows (natural order in and out).  For the
// four-step core, `row_order_out` leaves the spectrum in the row order
// (k1 * N2 + k2 holds X[k1 + N1 k2]) used by the convolution passes.
void c2c_passes(int prec, const Core &c, int64_t nrows, int sign, bool plain, const Endpoint &in,
                const Endpoint &out, void *scratch, bool row_order_out, std::vector<AfPass> &ps) {
    if (c.sh.kind == CORE_SINGLE) {
        // rows are the lines of one long "row": tiles of B rows need no row split
        const int L = (int)c.sh.N;
        AfPass d = base_pass(prec, L, c.sh.r, c.tw, c.tw_d, choose_B_rows(prec, L, nrows), nrows, 1, sign, plain);
        Endpoint i2 = in, o2 = out;
        set_in(d, i2, 1, in.bs);
        set_out(d, o2, 1, out.bs);
        d.in_bs = d.out_bs = 0;
        d.gt_in = c.gt_in;
        d.gt_out = c.gt_out;
        d.grp_mask = c.gt_mask;
        ps.push_back(d);
        return;
    }
    if (c.sh.kind == CORE_THREE) {
        three_passes(prec, c, nrows, sign, plain, in, out, scratch, ps);
        return;
    }
    const int64_t N = c.sh.N, N1 = c.sh.N1, N2 = c.sh.N2;
    Endpoint sc;
    sc.re = scratch;
    sc.bs = N;
    AfPass a = base_pass(prec, (int)N1, c.sh.r1, c.tw1, c.tw1_d, choose_B_cols(prec, N1), N2, nrows, sign,
                         plain);
    set_in(a, in, N2, 1);
    set_out(a, sc, N2, 1);
    a.in_lines_fast = a.out_lines_fast = 1;
    set_ft(a, c, sign);
    ps.push_back(a);
    AfPass b = base_pass(prec, (int)N2, c.sh.r2, c.tw2, c.tw2_d,
                         choose_B_cols(prec, N2), N1, nrows, sign, plain);
    set_in(b, sc, 1, N2);
    if (row_order_out) {
        set_out(b, out, 1, N2);
    } else {
        set_out(b, out, N1, 1);
        b.out_lines_fast = 1;
    }
    ps.push_back(b);
}

// Convolution with a spectrum: F+( spec * F-(pre(x)) ), then post-operations.
// The caller sets the pre/post fields on the first/last returned pass.
// Single core: one pass.  Four-step core: A (columns, forward, twiddle),
// B (rows: forward, spectrum, inverse, conjugate twiddle), C (columns, inverse).
void conv_passes(int prec, const Core &c, int64_t nrows, bool plain, const Endpoint &in, const Endpoint &out,
                 const void *spec, int64_t spec_bs, bool spec_plain, void *scratch, std::vector<AfPass> &ps) {
    const int spf = spec_plain ? AFP_SPEC_PLAIN : 0;
    if (c.sh.kind == CORE_SINGLE) {
        const int L = (int)c.sh.N;
        AfPass d = base_pass(prec, L, c.sh.r, c.tw, c.tw_d, choose_B_rows(prec, L, nrows), 1, nrows, -1, plain);
        set_in(d, in, 1, 0);
        set_out(d, out, 1, 0);
        d.flags |= AFP_CONV | spf;
        d.spec = spec;
        d.spec_bs = spec_bs;
        d.spec_ls = 0;
        ps.push_back(d);
        return;
    }
    const int64_t N = c.sh.N, N1 = c.sh.N1, N2 = c.sh.N2;
    Endpoint sc;
    sc.re = scratch;
    sc.bs = N;
    AfPass a = base_pass(prec, (int)N1, c.sh.r1, c.tw1, c.tw1_d, choose_B_cols(prec, N1), N2, nrows, -1, plain);
    set_in(a, in, N2, 1);
    set_out(a, sc, N2, 1);
    a.in_lines_fast = a.out_lines_fast = 1;
    set_ft(a, c, -1);
    ps.push_back(a);
    // power-of-two rows take the line count of their compile-time kernel
    const int bconv = fast_conv_lines(prec, N2);
    AfPass b = base_pass(prec, (int)N2, c.sh.r2, c.tw2, c.tw2_d, bconv ? bconv : choose_B_rows(prec, N2, nrows * N1), N1,
                         nrows, -1, plain);
    set_in(b, sc, 1, N2);
    set_out(b, sc, 1, N2);
    b.flags |= AFP_CONV | spf;
    b.spec = spec;
    b.spec_bs = spec_bs;
    b.spec_ls = N2;
    set_ft(b, c, +1);
    ps.push_back(b);
    AfPass cc = base_pass(prec, (int)N1, c.sh.r1, c.tw1, c.tw1_d, choose_B_cols(prec, N1), N2, nrows, +1, plain);
    set_in(cc, sc, N2, 1);
    set_out(cc, out, N2, 1);
    cc.in_lines_fast = cc.out_lines_fast = 1;
    ps.push_back(cc);
}

af_status launch_all(int prec, const std::vector<AfPass> &ps, hipStream_t s) {
    for (const AfPass &d : ps) HIPCHK(launch_pass(prec, d, s));
    return AF_SUCCESS;
}

// Stream-ordered scratch for one call.
struct Scratch {
    void *p = nullptr;
    hipStream_t s = nullptr;
    bool owned = false;
    af_status get(size_t bytes, void *user, hipStream_t stream) {
        s = stream;
        if (bytes == 0) return AF_SUCCESS;
        if (user) {
            p = user;
            return AF_SUCCESS;
        }
        HIPCHK(hipMallocAsync(&p, bytes, stream));
        owned = true;
        return AF_SUCCESS;
    }
    ~Scratch() {
        if (owned && p) (void)hipFreeAsync(p, s);
    }
};

int prec_of(const af_plan_desc &d) { return d.precision == AF_F64 ? PREC_F64 : PREC_F32; }

}  // namespace

// ---------------------------------------------------------------------------
// Autotuning records
// ---------------------------------------------------------------------------

// One pass's tuned choice: B lines per tile (0 = the planner's) and the order of
// its radices (0 = as planned, 1 = reversed, 2 = odd radices first).
struct PassTune {
    int B = 0;
    int order = 0;
};

// Radix order of a generic-kernel pass (the Stockham stages accept any order;
// the stage twiddle table W_L^m does not depend on it).
void reorder_radices(AfPass &d, int order) {
    if (order == 0 || d.nst < 2) return;
    // the prime-factor maps hold for the planned order only: a reordered pass
    // runs in Cooley-Tukey order
    d.gt_in = d.gt_out = nullptr;
    d.grp_mask = 0;
    std::vector<int> r(d.radix, d.radix + d.nst);
    if (order == 1) std::reverse(r.begin(), r.end());
    else std::stable_partition(r.begin(), r.end(), [](int x) { return x % 2 == 1; });
    int p = 1;
    for (int i = 0; i < d.nst; ++i) {
        d.radix[i] = r[i];
        d.pst[i] = p;
        p *= r[i];
    }
}

// A pass some kernel can run with its current tile geometry.
bool tile_launchable(int prec, const AfPass &d) {
    if (d.B < 1 || d.B > 64) return false;
    if (fast_ext_eligible(prec, d)) return true;
    if (fast_pass_eligible(prec, d) && fast_tile_lines(prec, d.L, d.B) == d.B) return true;
    return d.nthreads <= 1024 && (int64_t)d.B * d.L <= 1024 * (int64_t)pass_ept(prec == PREC_F64) &&
           d.lds_bytes <= lds_tile_bytes_cap() && (int64_t)d.B * d.L < ((int64_t)1 << 22);
}

std::mutex g_wisdom_mu;
std::map<std::string, std::vector<PassTune>> g_wisdom;

std::string wisdom_key(const af_plan_desc &d, int64_t batch, int planar) {
    return std::string(d.precision == AF_F64 ? "f64" : "f32") + " " + std::to_string(d.algorithm) + " " +
           std::to_string(d.length) + " " + std::to_string(batch) + " " + std::to_string(planar);
}

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------

struct af_plan_s {
    af_plan_desc desc{};
    int prec = PREC_F32;
    bool plain = false;
    PlanShape shape;
    Core core;
    Owner own;
    int32_t *in_map = nullptr, *out_map = nullptr;  // Rader gathers / scatters
    void *chirp = nullptr;                          // Bluestein chirp, plain complex values, length n
    void *spec = nullptr;                           // convolution spectrum, plain complex of the precision
    void *aux_tw = nullptr;                         // real packing twiddles W_n^k, k < n/2
    af_plan dual = nullptr;                         // the opposite sign (adjoint of Rader / Bluestein)
    af_plan sub = nullptr;                          // convolveCyclic of non-smooth lengths, or the half-length
                                                    // transform of unfused real kinds: C2C
    bool real_fused = false;
    // decomposition plans (ALG_DECOMP): Good-Thomas and Cooley-Tukey run
    // dec_a (length a) then dec_b (length b); Rader runs dec_a (p - 1, sign
    // -1) and dec_b (p - 1, sign +1) around the spectrum product.  dec_in /
    // dec_out: the input and output index maps on the device.
    af_plan dec_a = nullptr, dec_b = nullptr;
    int32_t *dec_in = nullptr, *dec_out = nullptr;
    void *dec_spec = nullptr;
    // autotuned tile lines and radix orders of the C2C passes, per (batch, planar)
    std::map<std::pair<int64_t, int>, std::vector<PassTune>> tuned;
    std::mutex tune_mu;
    ~af_plan_s() {
        if (dual) af_plan_destroy(dual);
        if (sub) af_plan_destroy(sub);
        if (dec_a) af_plan_destroy(dec_a);
        if (dec_b) af_plan_destroy(dec_b);
    }
    size_t scratch_per_row() const {
        const size_t e = (size_t)elem_bytes(desc.precision);
        size_t n = 0;
        if (desc.kind == AF_CONV_CYCLIC) {
            if (sub) return e * (size_t)desc.length * 2;
            n = (size_t)core.sh.N * (core.sh.kind == CORE_FOUR ? 2 : 1);
            return e * n;
        }
        if (shape.algo == ALG_DECOMP) return e * 2 * (size_t)shape.n;
        if (core.sh.kind == CORE_FOUR) n = (size_t)core.sh.N + (shape.algo == ALG_RADER ? 2 : 0);
        if (core.sh.kind == CORE_THREE) n = (size_t)core.sh.N;
        return e * n;
    }
};

namespace {

af_status validate_desc(const af_plan_desc *d) {
    if (!d) return AF_INVALID_ARG;
    if (d->length < 0) return AF_INVALID_ARG;
    if (d->sign != 1 && d->sign != -1) return AF_INVALID_ARG;
    if (d->precision != AF_F64 && d->precision != AF_F32) return AF_INVALID_ARG;
    if (d->algorithm < AF_ALG_AUTO || d->algorithm > AF_ALG_CHIRP235) return AF_INVALID_ARG;
    if (d->flags & ~(AF_FLAG_PLAIN_TWIDDLES | AF_FLAG_COOLEY_TUKEY)) return AF_INVALID_ARG;
    switch (d->kind) {
    case AF_C2C: break;
    case AF_R2C_FULL:
    case AF_C2R_FULL:
        if (d->length % 2) return AF_INVALID_ARG;  // Real.hs:45-47, 88-90
        break;
    case AF_CONV_CYCLIC:
        if (d->length < 1) return AF_INVALID_ARG;  // Planned.hs:776-778
        break;
    case AF_REAL_CONV:
        if (d->length % 2 || d->length < 2) return AF_INVALID_ARG;
        break;
    default: return AF_INVALID_ARG;
    }
    return AF_SUCCESS;
}

// Length of the device FFT core a kind needs.
int64_t core_length(const af_plan_desc &d) {
    if (d.kind == AF_R2C_FULL || d.kind == AF_C2R_FULL || d.kind == AF_REAL_CONV) return d.length / 2;
    return d.length;
}

af_status shape_for(const af_plan_desc &d, PlanShape &sh, std::string &why) {
    const int prec = d.precision;
    const int64_t n = core_length(d);
    if (d.length > ((int64_t)1 << 30)) {
        why = "length above 2^30";
        return AF_UNSUPPORTED;
    }
    if (d.kind == AF_C2C || d.kind == AF_CONV_CYCLIC) {
        // three passes only for plain transforms (the convolution passes take one or two)
        if (!plan_shape(prec, n, d.algorithm, sh, why, d.kind == AF_C2C)) return AF_UNSUPPORTED;
        return AF_SUCCESS;
    }
    // real kinds: the half-length transform
    sh = PlanShape();
    if (n == 0) {
        sh.algo = ALG_EMPTY;
        return AF_SUCCESS;
    }
    if (n == 1) {
        sh.n = 1;
        sh.algo = ALG_STOCKHAM;
        sh.core.kind = CORE_SINGLE;
        sh.core.N = 1;
        return AF_SUCCESS;
    }
    if (!plan_shape(prec, n, AF_ALG_AUTO, sh, why)) return AF_UNSUPPORTED;
    return AF_SUCCESS;
}

// Real kinds run as one fused pass when the half length is smooth and fits
// LDS (two half-length buffers for the convolution); otherwise as a
// half-length C2C plan of any kind plus the packing in global memory.
bool real_fusable(const af_plan_desc &d, const PlanShape &sh) {
    if (sh.algo != ALG_STOCKHAM || sh.core.kind != CORE_SINGLE) return false;
    if (d.kind == AF_REAL_CONV)
        return lds_bytes_for(d.precision, (int)sh.core.N, 1, 4) * 2 <= lds_tile_bytes_cap();
    return true;
}

// FP64 forward DFT of `count` interleaved values on the device (plan time).
af_status fft_f64_forward(int64_t N, const void *in, void *out, hipStream_t s) {
    af_plan_desc d{};
    d.length = N;
    d.sign = -1;
    d.kind = AF_C2C;
    d.precision = AF_F64;
    d.algorithm = AF_ALG_AUTO;
    af_plan p = nullptr;
    AFCHK(af_plan_create(&d, &p));
    af_status st = af_execute(p, 1, in, nullptr, out, nullptr, nullptr, s);
    if (st == AF_SUCCESS) st = HIPST(hipStreamSynchronize(s));
    af_plan_destroy(p);
    return st;
}

af_status build_rader(af_plan P, hipStream_t s) {
    const int64_t p = P->shape.n, L = p - 1;
    const int64_t g = multiplicative_generator(p);
    std::vector<int32_t> powg(L), inv_pow(L);
    int64_t x = 1;
    for (int64_t i = 0; i < L; ++i) {
        powg[i] = (int32_t)x;
        x = (int64_t)(((unsigned __int128)x * (uint64_t)g) % (uint64_t)p);
    }
    for (int64_t i = 0; i < L; ++i) inv_pow[i] = powg[(L - i) % L];  // g^-i
    AFCHK(P->own.alloc((void **)&P->in_map, L * 4));
    AFCHK(P->own.alloc((void **)&P->out_map, L * 4));
    HIPCHK(hipMemcpyAsync(P->in_map, inv_pow.data(), L * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(P->out_map, powg.data(), L * 4, hipMemcpyHostToDevice, s));
    // kernel b_i = exp(sign 2 pi i g^i / p), spectrum F-(b) / (p-1)
    void *kb = nullptr, *kf = nullptr;
    HIPCHK(hipMalloc(&kb, L * 16));
    HIPCHK(hipMalloc(&kf, L * 16));
    af_status st = AF_SUCCESS;
    { const af_status h__ = HIPST(gen_rader_kernel_d(kb, P->out_map, p, L, P->desc.sign, s)); if (st == AF_SUCCESS) st = h__; }
    if (st == AF_SUCCESS) st = fft_f64_forward(L, kb, kf, s);
    if (st == AF_SUCCESS) st = P->own.alloc(&P->spec, L * elem_bytes(P->desc.precision));
    if (st == AF_SUCCESS) {
        const bool four = P->core.sh.kind == CORE_FOUR;
        st = HIPST(convert_to_plain(P->prec, P->spec, kf, L, 1.0 / (double)L, four ? P->core.sh.N1 : 0,
                                four ? P->core.sh.N2 : 0, s));
    }
    { const af_status h__ = HIPST(hipStreamSynchronize(s)); if (st == AF_SUCCESS) st = h__; }
    (void)hipFree(kb);
    (void)hipFree(kf);
    return st;
}

af_status build_bluestein(af_plan P, hipStream_t s) {
    const int64_t n = P->shape.n, M = P->shape.M;
    void *cd = nullptr, *hb = nullptr, *hf = nullptr;
    HIPCHK(hipMalloc(&cd, n * 16));
    HIPCHK(hipMalloc(&hb, M * 16));
    HIPCHK(hipMalloc(&hf, M * 16));
    af_status st = AF_SUCCESS;
    // chirp c_k = exp(sign i pi k^2 / n), k < n; kernel h = conj(c) wrapped over M
    { const af_status h__ = HIPST(gen_chirp_d(cd, n, n, P->desc.sign, 0, 0, s)); if (st == AF_SUCCESS) st = h__; }
    // plain complex values of the precision (FP32: each value rounded once from FP64)
    if (st == AF_SUCCESS) st = P->own.alloc(&P->chirp, n * elem_bytes(P->desc.precision));
    if (st == AF_SUCCESS) st = HIPST(convert_to_plain(P->prec, P->chirp, cd, n, 1.0, 0, 0, s));
    if (st == AF_SUCCESS) st = HIPST(gen_chirp_d(hb, n, M, P->desc.sign, M, 1, s));
    if (st == AF_SUCCESS) st = fft_f64_forward(M, hb, hf, s);
    if (st == AF_SUCCESS) st = P->own.alloc(&P->spec, M * elem_bytes(P->desc.precision));
    if (st == AF_SUCCESS) {
        const bool four = P->core.sh.kind == CORE_FOUR;
        st = HIPST(convert_to_plain(P->prec, P->spec, hf, M, 1.0 / (double)M, four ? P->core.sh.N1 : 0,
                                four ? P->core.sh.N2 : 0, s));
    }
    { const af_status h__ = HIPST(hipStreamSynchronize(s)); if (st == AF_SUCCESS) st = h__; }
    (void)hipFree(cd);
    (void)hipFree(hb);
    (void)hipFree(hf);
    return st;
}

std::once_flag g_attr_flag;
void set_kernel_attributes() { init_kernel_attributes(); }

}  // namespace

extern "C" {

const char *af_status_string(af_status s) {
    switch (s) {
    case AF_SUCCESS: return "success";
    case AF_INVALID_ARG: return "invalid argument (a precondition of the reference is violated)";
    case AF_UNSUPPORTED: return "unsupported by this library";
    case AF_NO_DEVICE: return "no gfx950 device";
    case AF_HIP_ERROR: return "HIP runtime error";
    case AF_OUT_OF_MEMORY: return "out of device memory";
    case AF_INTERNAL: return "internal error";
    }
    return "unknown status";
}

const char *af_version(void) { return "accelerate-fourier-amd 0.1 (gfx950)"; }

const char *af_last_error(void) { return g_last_error; }

af_status af_plan_describe(const af_plan_desc *desc, char *buf, size_t buflen) {
    AFCHK(validate_desc(desc));
    if (!buf || buflen == 0) return AF_INVALID_ARG;
    PlanShape sh;
    std::string why;
    af_status st = shape_for(*desc, sh, why);
    std::string s;
    const char *kinds[] = {"c2c", "r2c", "c2r", "convcyclic", "realconv"};
    s = std::string(kinds[desc->kind]) + (desc->precision == AF_F64 ? ".f64:" : ".f32:");
    const bool real_kind = desc->kind == AF_R2C_FULL || desc->kind == AF_C2R_FULL || desc->kind == AF_REAL_CONV;
    if (st != AF_SUCCESS) s += "unsupported(" + why + ")";
    else if (real_kind && sh.algo != ALG_EMPTY && !real_fusable(*desc, sh)) s += "packed(" + describe_shape(sh) + ")";
    else s += describe_shape(sh);
    if (st == AF_SUCCESS && desc->kind == AF_C2C && sh.algo == ALG_STOCKHAM && sh.core.kind == CORE_SINGLE &&
        !(desc->flags & AF_FLAG_COOLEY_TUKEY) && prime_factor_radices(sh.core.r)) {
        const size_t at = s.find("stockham");
        if (at != std::string::npos) s.replace(at, 8, "primefactor");
    }
    std::snprintf(buf, buflen, "%s", s.c_str());
    return st == AF_SUCCESS && s.size() < buflen ? AF_SUCCESS : (st != AF_SUCCESS ? st : AF_INVALID_ARG);
}

static af_status plan_create_impl(const af_plan_desc *desc, af_plan *plan, bool with_dual);

af_status af_plan_create(const af_plan_desc *desc, af_plan *plan) { return plan_create_impl(desc, plan, true); }

// Sub-plans and index maps of a decomposition plan (Planned.hs:255-293 with no
// Bluestein): Good-Thomas maps as af_good_thomas_maps (the reference's
// permuteSkewGrid / permuteSkewGridInv), Rader's g^-q gather and g^r scatter
// and its kernel spectrum F-(b) / (p - 1), b_m = exp(sign 2 pi i g^m / p), from
// an FP64 decomposition of p - 1.
static af_status build_decomp(af_plan P, hipStream_t s) {
    const PlanShape &sh = P->shape;
    auto sub = [&](int64_t len, int sign, af_plan *out) {
        af_plan_desc d = P->desc;
        d.length = len;
        d.sign = sign;
        d.algorithm = AF_ALG_DECOMPOSE;
        return plan_create_impl(&d, out, false);
    };
    const int64_t n = sh.n;
    if (n >= ((int64_t)1 << 31)) return AF_UNSUPPORTED;
    if (sh.dec == DEC_GOOD_THOMAS || sh.dec == DEC_COOLEY_TUKEY) {
        AFCHK(sub(sh.da, P->desc.sign, &P->dec_a));
        AFCHK(sub(sh.db, P->desc.sign, &P->dec_b));
        if (sh.dec == DEC_COOLEY_TUKEY) return AF_SUCCESS;
        std::vector<int64_t> im(n), om(n);
        AFCHK(af_good_thomas_maps(sh.da, sh.db, im.data(), om.data()));
        std::vector<int32_t> im32(im.begin(), im.end()), om32(om.begin(), om.end());
        AFCHK(P->own.alloc((void **)&P->dec_in, n * 4));
        AFCHK(P->own.alloc((void **)&P->dec_out, n * 4));
        HIPCHK(hipMemcpyAsync(P->dec_in, im32.data(), n * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(P->dec_out, om32.data(), n * 4, hipMemcpyHostToDevice, s));
        return HIPST(hipStreamSynchronize(s));
    }
    // Rader: p = n, convolution length L = p - 1
    const int64_t p = n, L = p - 1;
    AFCHK(sub(L, -1, &P->dec_a));
    AFCHK(sub(L, 1, &P->dec_b));
    const int64_t g = multiplicative_generator(p);
    std::vector<int32_t> powg(L), inv_pow(L);
    int64_t x = 1;
    for (int64_t i = 0; i < L; ++i) {
        powg[i] = (int32_t)x;
        x = (int64_t)(((unsigned __int128)x * (uint64_t)g) % (uint64_t)p);
    }
    for (int64_t i = 0; i < L; ++i) inv_pow[i] = powg[(L - i) % L];  // g^-i
    AFCHK(P->own.alloc((void **)&P->dec_in, L * 4));
    AFCHK(P->own.alloc((void **)&P->dec_out, L * 4));
    HIPCHK(hipMemcpyAsync(P->dec_in, inv_pow.data(), L * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(P->dec_out, powg.data(), L * 4, hipMemcpyHostToDevice, s));
    af_plan f64 = nullptr;
    {
        af_plan_desc d = P->desc;
        d.length = L;
        d.sign = -1;
        d.precision = AF_F64;
        d.algorithm = AF_ALG_DECOMPOSE;
        d.flags = 0;
        AFCHK(plan_create_impl(&d, &f64, false));
    }
    void *kb = nullptr, *kf = nullptr;
    af_status st = HIPST(hipMalloc(&kb, L * 16));
    if (st == AF_SUCCESS) st = HIPST(hipMalloc(&kf, L * 16));
    if (st == AF_SUCCESS) st = HIPST(gen_rader_kernel_d(kb, P->dec_out, p, L, P->desc.sign, s));
    if (st == AF_SUCCESS) st = af_execute(f64, 1, kb, nullptr, kf, nullptr, nullptr, s);
    if (st == AF_SUCCESS) st = P->own.alloc(&P->dec_spec, L * elem_bytes(P->desc.precision));
    if (st == AF_SUCCESS) st = HIPST(convert_to_plain(P->prec, P->dec_spec, kf, L, 1.0 / (double)L, 0, 0, s));
    { const af_status h__ = HIPST(hipStreamSynchronize(s)); if (st == AF_SUCCESS) st = h__; }
    if (kb) (void)hipFree(kb);
    if (kf) (void)hipFree(kf);
    af_plan_destroy(f64);
    return st;
}


static af_status plan_create_impl(const af_plan_desc *desc, af_plan *plan, bool with_dual) {
    if (!plan) return AF_INVALID_ARG;
    *plan = nullptr;
    AFCHK(validate_desc(desc));
    PlanShape sh;
    std::string why;
    AFCHK(shape_for(*desc, sh, why));
    AFCHK(device_init());
    std::call_once(g_attr_flag, set_kernel_attributes);
    std::unique_ptr<af_plan_s> P(new af_plan_s);
    P->desc = *desc;
    P->prec = prec_of(*desc);
    P->plain = (desc->flags & AF_FLAG_PLAIN_TWIDDLES) != 0;
    P->shape = sh;
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    af_status st = AF_SUCCESS;
    if (desc->kind == AF_CONV_CYCLIC && sh.algo != ALG_STOCKHAM && sh.algo != ALG_COPY) {
        // non-smooth length: three full transforms and a pointwise product
        af_plan_desc d2 = *desc;
        d2.kind = AF_C2C;
        d2.sign = -1;
        st = af_plan_create(&d2, &P->sub);
        if (st == AF_SUCCESS) {
            CoreShape one;
            one.kind = CORE_SINGLE;
            one.N = 1;
            P->core.sh = one;
        }
    } else if ((desc->kind == AF_R2C_FULL || desc->kind == AF_C2R_FULL || desc->kind == AF_REAL_CONV) &&
               sh.algo != ALG_EMPTY && !real_fusable(*desc, sh)) {
        af_plan_desc d2 = *desc;
        d2.kind = AF_C2C;
        d2.length = sh.n;
        d2.algorithm = AF_ALG_AUTO;
        if (desc->kind == AF_REAL_CONV) d2.sign = -1;
        st = af_plan_create(&d2, &P->sub);
    } else if (sh.algo == ALG_COPY) {
        CoreShape one;
        one.kind = CORE_SINGLE;
        one.N = 1;
        P->core.sh = one;
    } else if (sh.algo == ALG_DECOMP) {
        st = build_decomp(P.get(), s);
    } else if (sh.algo != ALG_EMPTY) {
        // plain transforms of one pass run in prime-factor order unless the
        // caller asks for Cooley-Tukey
        const bool pf = desc->kind == AF_C2C && sh.algo == ALG_STOCKHAM && !(desc->flags & AF_FLAG_COOLEY_TUKEY);
        st = build_core(P->prec, sh.core, P->core, P->own, s, pf);
    }
    if (st == AF_SUCCESS && desc->kind == AF_C2C) {
        if (sh.algo == ALG_RADER) st = build_rader(P.get(), s);
        else if (sh.algo == ALG_BLUESTEIN) st = build_bluestein(P.get(), s);
        if (st == AF_SUCCESS && with_dual &&
            (sh.algo == ALG_RADER || sh.algo == ALG_BLUESTEIN || sh.algo == ALG_DECOMP)) {
            // the adjoint needs the tables of the opposite sign (cacheDuplex, Planned.hs:335-341)
            af_plan_desc d2 = *desc;
            d2.sign = -desc->sign;
            st = plan_create_impl(&d2, &P->dual, false);
        }
    }
    if (desc->kind == AF_R2C_FULL || desc->kind == AF_C2R_FULL || desc->kind == AF_REAL_CONV)
        P->real_fused = sh.algo != ALG_EMPTY && real_fusable(*desc, sh);
    if (st == AF_SUCCESS &&
        (desc->kind == AF_R2C_FULL || desc->kind == AF_C2R_FULL || desc->kind == AF_REAL_CONV) && sh.n >= 1) {
        st = P->own.alloc(&P->aux_tw, (size_t)sh.n * 16);
        if (st == AF_SUCCESS) st = HIPST(gen_twiddles(P->prec, P->aux_tw, 2 * sh.n, sh.n, 1, s));
    }
    { const af_status h__ = HIPST(hipStreamSynchronize(s)); if (st == AF_SUCCESS) st = h__; }
    (void)hipStreamDestroy(s);
    if (st != AF_SUCCESS) return st;
    *plan = P.release();
    return AF_SUCCESS;
}

af_status af_plan_destroy(af_plan plan) {
    if (!plan) return AF_INVALID_ARG;
    delete plan;
    return AF_SUCCESS;
}

af_status af_plan_workspace_size(af_plan plan, int64_t batch, size_t *bytes) {
    if (!plan || !bytes || batch < 0) return AF_INVALID_ARG;
    *bytes = plan->scratch_per_row() * (size_t)batch;
    return AF_SUCCESS;
}

}  // extern "C"

namespace {

Endpoint ep(const void *re, const void *im, int64_t bs) {
    Endpoint e;
    e.re = re;
    e.im = im;
    e.bs = bs;
    return e;
}

// Passes of a complex transform with an explicit sign (the plan's own, or the
// opposite for the adjoint of Stockham plans).  `scratch` holds
// scratch_per_row() bytes per row.
// The tuned choices recorded for (batch, planar) -- the plan's own, else the
// process's wisdom -- applied to the passes where a kernel can run them.
void apply_tuning(af_plan P, int64_t batch, int planar, std::vector<AfPass> &ps) {
    std::vector<PassTune> t;
    {
        std::lock_guard<std::mutex> lk(P->tune_mu);
        auto it = P->tuned.find({batch, planar});
        if (it != P->tuned.end()) t = it->second;
    }
    if (t.empty()) {
        std::lock_guard<std::mutex> lk(g_wisdom_mu);
        auto it = g_wisdom.find(wisdom_key(P->desc, batch, planar));
        if (it == g_wisdom.end()) return;
        t = it->second;
    }
    if (t.size() != ps.size()) return;
    for (size_t i = 0; i < ps.size(); ++i) {
        AfPass d = ps[i];
        reorder_radices(d, t[i].order);
        if (t[i].B) set_tile(d, P->prec, t[i].B);
        if (tile_launchable(P->prec, d)) ps[i] = d;
    }
}

af_status build_c2c_passes(af_plan P, int sign, int64_t batch, const Endpoint &in, const Endpoint &out,
                           void *scratch, std::vector<AfPass> &ps);

// Passes of a complex transform, with the autotuned tile shapes when `tuned`
// (the uneven batch builds untuned passes: its launch groups are fixed when it
// is created).
af_status build_c2c(af_plan P, int sign, int64_t batch, const Endpoint &in, const Endpoint &out, void *scratch,
                    std::vector<AfPass> &ps, bool tuned = true) {
    const size_t first = ps.size();
    AFCHK(build_c2c_passes(P, sign, batch, in, out, scratch, ps));
    if (tuned && ps.size() > first) {
        std::vector<AfPass> mine(ps.begin() + first, ps.end());
        apply_tuning(P, batch, in.im != nullptr, mine);
        std::copy(mine.begin(), mine.end(), ps.begin() + first);
    }
    return AF_SUCCESS;
}

af_status build_c2c_passes(af_plan P, int sign, int64_t batch, const Endpoint &in, const Endpoint &out,
                           void *scratch, std::vector<AfPass> &ps) {
    // decomposition plans are composed on the host, not a sequence of passes
    if (P->shape.algo == ALG_DECOMP) return AF_UNSUPPORTED;
    if (sign != P->desc.sign && (P->shape.algo == ALG_RADER || P->shape.algo == ALG_BLUESTEIN)) {
        // Rader kernels and chirps depend on the sign: the opposite sign is the dual plan's
        if (!P->dual) return AF_INTERNAL;
        P = P->dual;
    }
    const int64_t n = P->shape.n;
    const int prec = P->prec;
    if (n == 0 || batch == 0) return AF_SUCCESS;
    switch (P->shape.algo) {
    case ALG_COPY: {
        AfPass d = base_pass(prec, 1, {}, nullptr, nullptr, 64, 1, batch, sign, P->plain);
        set_in(d, in, 1, 0);
        set_out(d, out, 1, 0);
        ps.push_back(d);
        break;
    }
    case ALG_STOCKHAM:
        c2c_passes(prec, P->core, batch, sign, P->plain, in, out, scratch, false, ps);
        break;
    case ALG_RADER: {
        const int64_t L = n - 1;
        const bool four = P->core.sh.kind == CORE_FOUR;
        const size_t first = ps.size();
        conv_passes(prec, P->core, batch, P->plain, in, out, P->spec, 0, true, scratch, ps);
        AfPass &a = ps[first];
        a.in_map = P->in_map;
        a.in_len = L;
        AfPass &last = ps.back();
        last.out_map = P->out_map;
        last.out_len = L;
        if (!four) {
            last.flags |= AFP_RADER;
        } else {
            // x0 of every row is saved by the first pass, so in-place calls are safe
            void *side = (char *)scratch + (size_t)batch * L * elem_bytes(P->desc.precision);
            void *dc = (char *)side + (size_t)batch * elem_bytes(P->desc.precision);
            a.flags |= AFP_X0_CAPTURE;
            a.x0_out = side;
            ps[first + 1].flags |= AFP_DC_CAPTURE;
            ps[first + 1].dc_out = dc;
            last.flags |= AFP_RADER_OUT;
            last.x0_re = side;
            last.x0_im = nullptr;
            last.x0_bs = 1;
            last.dc_in = dc;
        }
        break;
    }
    case ALG_BLUESTEIN: {
        const size_t first = ps.size();
        conv_passes(prec, P->core, batch, P->plain, in, out, P->spec, 0, true, scratch, ps);
        ps[first].in_len = n;
        ps[first].in_mul = P->chirp;
        ps.back().out_len = n;
        ps.back().out_mul = P->chirp;
        break;
    }
    default: return AF_INTERNAL;
    }
    return AF_SUCCESS;
}

af_status run_c2c(af_plan P, int sign, int64_t batch, const void *in_re, const void *in_im, void *out_re,
                  void *out_im, void *ws, hipStream_t s);

// A decomposition plan (ALG_DECOMP) of its own sign over `batch` rows: the
// steps of Good-Thomas, Cooley-Tukey or Rader (decomp.hip) around the
// sub-plans, intermediates in two interleaved buffers of batch x n.
af_status run_decomp(af_plan P, int64_t batch, const void *in_re, const void *in_im, void *out_re, void *out_im,
                     void *ws, hipStream_t s) {
    if (!P || !P->dec_a || !P->dec_b) return AF_INTERNAL;
    const int64_t n = P->shape.n, a = P->shape.da, b = P->shape.db;
    const size_t eb = (size_t)elem_bytes(P->desc.precision);
    const int prec = P->prec;
    Scratch sc;
    AFCHK(sc.get(2 * (size_t)batch * (size_t)n * eb, ws, s));
    void *t1 = sc.p, *t2 = (char *)sc.p + (size_t)batch * (size_t)n * eb;
    af_plan A = P->dec_a, B = P->dec_b;
    switch (P->shape.dec) {
    case DEC_GOOD_THOMAS:
        // t1[r][j2][j1] = x[r][(a j2 + b j1) mod n]; rows over j1; transpose; rows over j2;
        // out[r][q] = t2[r][(q mod a) b + (q mod b)]
        HIPCHK(launch_dc_gather(prec, in_re, in_im, n, t1, nullptr, n, P->dec_in, batch, s));
        AFCHK(run_c2c(A, A->desc.sign, batch * b, t1, nullptr, t2, nullptr, nullptr, s));
        HIPCHK(launch_transpose(prec, t2, nullptr, t1, nullptr, b, a, batch, s));
        AFCHK(run_c2c(B, B->desc.sign, batch * a, t1, nullptr, t2, nullptr, nullptr, s));
        HIPCHK(launch_dc_gather(prec, t2, nullptr, n, out_re, out_im, n, P->dec_out, batch, s));
        return AF_SUCCESS;
    case DEC_COOLEY_TUKEY:
        // x[r][b j1 + j2] as a x b; transpose; rows over j1; W_n^(j2 k1); transpose;
        // rows over j2; X[k1 + a k2] is the transpose of the a x b result
        HIPCHK(launch_transpose(prec, in_re, in_im, t1, nullptr, a, b, batch, s));
        AFCHK(run_c2c(A, A->desc.sign, batch * b, t1, nullptr, t2, nullptr, nullptr, s));
        HIPCHK(launch_dc_twiddle(prec, t2, a, b, n, P->desc.sign, batch, s));
        HIPCHK(launch_transpose(prec, t2, nullptr, t1, nullptr, b, a, batch, s));
        AFCHK(run_c2c(B, B->desc.sign, batch * a, t1, nullptr, t2, nullptr, nullptr, s));
        HIPCHK(launch_transpose(prec, t2, nullptr, out_re, out_im, a, b, batch, s));
        return AF_SUCCESS;
    case DEC_RADER: {
        // a[q] = x[g^-q]; c = F+(F-(a) * F-(b) / (p - 1)); X[0] = sum x, X[g^r] = x[0] + c[r]
        const int64_t L = n - 1;
        HIPCHK(launch_dc_gather(prec, in_re, in_im, n, t1, nullptr, L, P->dec_in, batch, s));
        AFCHK(run_c2c(A, -1, batch, t1, nullptr, t2, nullptr, nullptr, s));
        HIPCHK(launch_dc_spec_mul(prec, t2, P->dec_spec, L, batch, s));
        AFCHK(run_c2c(B, 1, batch, t2, nullptr, t1, nullptr, nullptr, s));
        HIPCHK(launch_dc_rader_out(prec, in_re, in_im, t1, P->dec_out, n, batch, out_re, out_im, s));
        return AF_SUCCESS;
    }
    }
    return AF_INTERNAL;
}

af_status run_c2c(af_plan P, int sign, int64_t batch, const void *in_re, const void *in_im, void *out_re,
                  void *out_im, void *ws, hipStream_t s) {
    const int64_t n = P->shape.n;
    if (n == 0 || batch == 0) return AF_SUCCESS;
    if (P->shape.algo == ALG_DECOMP)
        return run_decomp(sign == P->desc.sign ? P : P->dual, batch, in_re, in_im, out_re, out_im, ws, s);
    Scratch sc;
    AFCHK(sc.get(P->scratch_per_row() * (size_t)batch, ws, s));
    std::vector<AfPass> ps;
    AFCHK(build_c2c(P, sign, batch, ep(in_re, in_im, n), ep(out_re, out_im, n), sc.p, ps));
    return launch_all(P->prec, ps, s);
}

// Transform along the middle axis of an (outer, n, inner) array with the C2C
// plan P (its own sign).  inner == 1 is the plain batched transform; a plain
// single-pass core runs one strided pass (lines adjacent, elements `inner`
// apart, the geometry of a four-step column pass, in place allowed); any other
// plan transposes to (outer, inner, n), transforms contiguous rows and
// transposes back.
af_status run_axis(af_plan P, int64_t outer, int64_t inner, const void *in_re, const void *in_im, void *out_re,
                   void *out_im, hipStream_t s) {
    const int64_t n = P->shape.n;
    if (n == 0 || outer == 0 || inner == 0) return AF_SUCCESS;
    const int prec = P->prec;
    if (inner == 1 || n == 1)  // contiguous rows (n == 1: rows of one element)
        return run_c2c(P, P->desc.sign, outer * (n == 1 ? inner : 1), in_re, in_im, out_re, out_im, nullptr, s);
    if (P->shape.algo == ALG_STOCKHAM && P->core.sh.kind == CORE_SINGLE && n * inner < ((int64_t)1 << 31)) {
        int B = choose_B_cols(prec, n);
        while (B > 1 && B > inner) B /= 2;
        AfPass d = base_pass(prec, (int)n, P->core.sh.r, P->core.tw, P->core.tw_d, B, inner, outer, P->desc.sign,
                             P->plain);
        set_in(d, ep(in_re, in_im, n * inner), inner, 1);
        set_out(d, ep(out_re, out_im, n * inner), inner, 1);
        d.in_lines_fast = d.out_lines_fast = 1;
        HIPCHK(launch_pass(prec, d, s));
        return AF_SUCCESS;
    }
    const size_t eb = (size_t)elem_bytes(P->desc.precision), tot = (size_t)(outer * n * inner);
    Scratch sc;
    AFCHK(sc.get(2 * tot * eb, nullptr, s));
    void *t1 = sc.p, *t2 = (char *)sc.p + tot * eb;
    HIPCHK(launch_transpose(prec, in_re, in_im, t1, nullptr, n, inner, outer, s));
    AFCHK(run_c2c(P, P->desc.sign, outer * inner, t1, nullptr, t2, nullptr, nullptr, s));
    HIPCHK(launch_transpose(prec, t2, nullptr, out_re, out_im, inner, n, outer, s));
    return AF_SUCCESS;
}

bool bad_ptrs(const void *a, const void *b) { return !a || !b; }

// Unfused real kinds: the half-length C2C plan P->sub plus the packing kernels.
af_status real_general(af_plan P, int op, int64_t batch, const void *a_re, const void *a_im, const void *b,
                       const void *c, void *o1, void *o1_im, void *o2, hipStream_t s) {
    const int prec = P->prec;
    const int64_t n2 = P->shape.n;
    const size_t eb = (size_t)elem_bytes(P->desc.precision);
    const size_t half = (size_t)batch * n2 * eb;
    af_plan sub = P->sub;
    const int pl = P->plain ? 1 : 0;
    Scratch sc;
    switch (op) {
    case 0:  // toSpectrum: transform the real rows viewed as interleaved complex, untangle
        AFCHK(sc.get(half, nullptr, s));
        AFCHK(run_c2c(sub, P->desc.sign, batch, a_re, nullptr, sc.p, nullptr, nullptr, s));
        HIPCHK(launch_pack(prec, PACK_UNTANGLE, sc.p, nullptr, nullptr, o1, o1_im, P->aux_tw, n2, batch, pl, 0, s));
        return AF_SUCCESS;
    case 1:  // fromSpectrum: entangle, transform into the real rows
        AFCHK(sc.get(half, nullptr, s));
        HIPCHK(launch_pack(prec, PACK_ENTANGLE, a_re, a_im, nullptr, sc.p, nullptr, P->aux_tw, n2, batch, pl, 0, s));
        return run_c2c(sub, P->desc.sign, batch, sc.p, nullptr, o1, nullptr, nullptr, s);
    case 2:  // adjoint of toSpectrum: Hermitian part, entangle, inverse-sign transform
        AFCHK(sc.get(half, nullptr, s));
        HIPCHK(launch_pack(prec, PACK_ENTANGLE, a_re, a_im, nullptr, sc.p, nullptr, P->aux_tw, n2, batch, pl, 1, s));
        return run_c2c(sub, +1, batch, sc.p, nullptr, o1, nullptr, nullptr, s);
    case 3:  // adjoint of fromSpectrum: opposite-sign transform, adjoint entangling
        AFCHK(sc.get(half, nullptr, s));
        AFCHK(run_c2c(sub, -P->desc.sign, batch, a_re, nullptr, sc.p, nullptr, nullptr, s));
        HIPCHK(launch_pack(prec, PACK_ENTANGLE_ADJ, sc.p, nullptr, nullptr, o1, o1_im, P->aux_tw, n2, batch, pl, 0, s));
        return AF_SUCCESS;
    case 4: {  // real convolution y = fromSpectrum(S_x * S_k)
        // the product reads partners (k, -k): its output must not alias its inputs
        AFCHK(sc.get(3 * half, nullptr, s));
        void *za = sc.p, *zb = (char *)sc.p + half, *zp = (char *)sc.p + 2 * half;
        AFCHK(run_c2c(sub, -1, batch, a_re, nullptr, za, nullptr, nullptr, s));
        AFCHK(run_c2c(sub, -1, batch, b, nullptr, zb, nullptr, nullptr, s));
        HIPCHK(launch_pack(prec, PACK_PRODUCT, za, nullptr, zb, zp, nullptr, P->aux_tw, n2, batch, pl, 0, s));
        return run_c2c(sub, +1, batch, zp, nullptr, o1, nullptr, nullptr, s);
    }
    default: {  // its adjoint: dk = fromSpectrum(S_w conj S_x), dx = fromSpectrum(S_w conj S_k)
        AFCHK(sc.get(3 * half, nullptr, s));
        void *zw = sc.p, *zx = (char *)sc.p + half, *zp = (char *)sc.p + 2 * half;
        AFCHK(run_c2c(sub, -1, batch, c, nullptr, zw, nullptr, nullptr, s));
        AFCHK(run_c2c(sub, -1, batch, a_re, nullptr, zx, nullptr, nullptr, s));
        HIPCHK(launch_pack(prec, PACK_PRODUCT, zw, nullptr, zx, zp, nullptr, P->aux_tw, n2, batch, pl, 1, s));
        AFCHK(run_c2c(sub, +1, batch, zp, nullptr, o2, nullptr, nullptr, s));
        AFCHK(run_c2c(sub, -1, batch, b, nullptr, zx, nullptr, nullptr, s));
        HIPCHK(launch_pack(prec, PACK_PRODUCT, zw, nullptr, zx, zp, nullptr, P->aux_tw, n2, batch, pl, 1, s));
        return run_c2c(sub, +1, batch, zp, nullptr, o1, nullptr, nullptr, s);
    }
    }
}

}  // namespace

extern "C" {

af_status af_execute(af_plan plan, int64_t batch, const void *in_re, const void *in_im, void *out_re, void *out_im,
                     void *workspace, void *stream) {
    if (!plan || plan->desc.kind != AF_C2C || batch < 0) return AF_INVALID_ARG;
    if (batch > 0 && plan->desc.length > 0 && bad_ptrs(in_re, out_re)) return AF_INVALID_ARG;
    if (batch * std::max<int64_t>(plan->desc.length, 1) > ((int64_t)1 << 40)) return AF_INVALID_ARG;
    return run_c2c(plan, plan->desc.sign, batch, in_re, in_im, out_re, out_im, workspace, (hipStream_t)stream);
}

af_status af_execute_noalloc(af_plan plan, int64_t batch, const void *in_re, const void *in_im, void *out_re,
                             void *out_im, void *stream) {
    if (!plan || plan->desc.kind != AF_C2C || batch < 0) return AF_INVALID_ARG;
    if (batch > 0 && plan->desc.length > 0 && bad_ptrs(in_re, out_re)) return AF_INVALID_ARG;
    if (batch * std::max<int64_t>(plan->desc.length, 1) > ((int64_t)1 << 40)) return AF_INVALID_ARG;
    if (plan->shape.algo == ALG_DECOMP) {
        std::snprintf(g_last_error, sizeof g_last_error,
                      "af_execute_noalloc: decomposition plans run sub-plans that allocate their own scratch");
        return AF_UNSUPPORTED;
    }
    const size_t need = plan->scratch_per_row() * (size_t)batch;
    if (need > kNoAllocScratchBytes) {
        std::snprintf(g_last_error, sizeof g_last_error,
                      "af_execute_noalloc: %zu bytes of scratch needed for %lld rows, the fixed buffer holds %zu",
                      need, (long long)batch, kNoAllocScratchBytes);
        return AF_UNSUPPORTED;
    }
    AFCHK(device_init());
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    void *ws = nullptr;
    AFCHK(noalloc_scratch(dev, &ws));
    // every pass writes the scratch it later reads, so it needs no clearing
    return run_c2c(plan, plan->desc.sign, batch, in_re, in_im, out_re, out_im, need ? ws : nullptr,
                   (hipStream_t)stream);
}

af_status af_plan_num_passes(af_plan plan, int64_t batch, int32_t *npasses) {
    if (!plan || !npasses || batch < 0 || plan->desc.kind != AF_C2C) return AF_INVALID_ARG;
    std::vector<AfPass> ps;
    const int64_t n = plan->desc.length;
    AFCHK(build_c2c(plan, plan->desc.sign, batch, ep((void *)16, nullptr, n), ep((void *)16, nullptr, n),
                    (void *)16, ps));
    *npasses = (int32_t)ps.size();
    return AF_SUCCESS;
}

af_status af_execute_pass(af_plan plan, int32_t index, int64_t batch, const void *in_re, const void *in_im,
                          void *out_re, void *out_im, void *workspace, void *stream) {
    if (!plan || plan->desc.kind != AF_C2C || batch < 0 || index < 0) return AF_INVALID_ARG;
    if (batch > 0 && plan->desc.length > 0 && bad_ptrs(in_re, out_re)) return AF_INVALID_ARG;
    if (plan->scratch_per_row() > 0 && !workspace) return AF_INVALID_ARG;
    const int64_t n = plan->desc.length;
    std::vector<AfPass> ps;
    AFCHK(build_c2c(plan, plan->desc.sign, batch, ep(in_re, in_im, n), ep(out_re, out_im, n), workspace, ps));
    if (index >= (int32_t)ps.size()) return AF_INVALID_ARG;
    HIPCHK(launch_pass(plan->prec, ps[index], (hipStream_t)stream));
    return AF_SUCCESS;
}

// ---- autotuner ----

af_status af_plan_tune(af_plan plan, int64_t batch, int32_t planar, void *stream, char *report, size_t report_len) {
    if (!plan || plan->desc.kind != AF_C2C || batch <= 0 || (planar != 0 && planar != 1)) return AF_INVALID_ARG;
    if (report && report_len) report[0] = 0;
    const int64_t n = plan->desc.length;
    if (n <= 1) return AF_SUCCESS;
    hipStream_t s = (hipStream_t)stream;
    const int prec = plan->prec;
    const size_t eb = (size_t)elem_bytes(plan->desc.precision);
    const size_t bytes = (size_t)batch * (size_t)n * eb;
    // zero-filled buffers: the transform's timing does not depend on the values
    void *in = nullptr, *out = nullptr, *ws = nullptr;
    const size_t wsb = plan->scratch_per_row() * (size_t)batch;
    HIPCHK(hipMalloc(&in, bytes));
    af_status st = HIPST(hipMalloc(&out, bytes));
    if (st == AF_SUCCESS && wsb) st = HIPST(hipMalloc(&ws, wsb));
    if (st == AF_SUCCESS) st = HIPST(hipMemsetAsync(in, 0, bytes, s));
    if (st == AF_SUCCESS) st = HIPST(hipMemsetAsync(out, 0, bytes, s));
    if (st == AF_SUCCESS && wsb) st = HIPST(hipMemsetAsync(ws, 0, wsb, s));
    Endpoint ei, eo;
    if (planar) {
        ei = ep(in, (char *)in + bytes / 2, n);
        eo = ep(out, (char *)out + bytes / 2, n);
    } else {
        ei = ep(in, nullptr, n);
        eo = ep(out, nullptr, n);
    }
    std::vector<AfPass> ps;
    if (st == AF_SUCCESS) st = build_c2c(plan, plan->desc.sign, batch, ei, eo, ws, ps, false);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st == AF_SUCCESS) st = HIPST(hipEventCreate(&e0));
    if (st == AF_SUCCESS) st = HIPST(hipEventCreate(&e1));
    // average microseconds per launch of d: two warm-ups, then the best of five runs of five
    auto time_pass = [&](const AfPass &d, double &us) -> af_status {
        for (int w = 0; w < 2; ++w) HIPCHK(launch_pass(prec, d, s));
        us = 1e30;
        for (int rep = 0; rep < 5; ++rep) {
            HIPCHK(hipEventRecord(e0, s));
            for (int k = 0; k < 5; ++k) HIPCHK(launch_pass(prec, d, s));
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            us = std::min(us, (double)ms * 1e3 / 5.0);
        }
        return AF_SUCCESS;
    };
    std::vector<PassTune> best(ps.size());
    std::string text;
    for (size_t i = 0; i < ps.size() && st == AF_SUCCESS; ++i) {
        const AfPass base = ps[i];
        double t0 = 0;
        st = time_pass(base, t0);
        double tb = t0;
        // generic-kernel passes also try other radix orders
        const bool generic = !fast_ext_eligible(prec, base) &&
                             !(fast_pass_eligible(prec, base) && fast_tile_lines(prec, base.L, base.B) == base.B);
        const int norders = generic && base.nst > 1 ? 3 : 1;
        for (int B = 1; B <= 64 && st == AF_SUCCESS; B *= 2)
            for (int o = 0; o < norders && st == AF_SUCCESS; ++o) {
                if (B == base.B && o == 0) continue;
                AfPass d = base;
                reorder_radices(d, o);
                set_tile(d, prec, B);
                if (!tile_launchable(prec, d)) continue;
                double t = 0;
                st = time_pass(d, t);
                // the incumbent is timed again right after each candidate (its
                // best time only improves), and a candidate must beat it by 3%:
                // run-to-run noise does not flip the choice
                AfPass cur = base;
                if (best[i].B) {
                    reorder_radices(cur, best[i].order);
                    set_tile(cur, prec, best[i].B);
                }
                double tc = 0;
                if (st == AF_SUCCESS) st = time_pass(cur, tc);
                tb = std::min(tb, tc);
                if (best[i].B == 0) t0 = std::min(t0, tc);
                if (st == AF_SUCCESS && t < tb * 0.97) {
                    tb = t;
                    best[i].B = B;
                    best[i].order = o;
                }
            }
        char line[160];
        std::snprintf(line, sizeof line, "pass %zu L=%d: B=%d order=%d %.2f us (planned B=%d %.2f us)\n", i, base.L,
                      best[i].B ? best[i].B : base.B, best[i].order, tb, base.B, t0);
        text += line;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    { const af_status h__ = HIPST(hipStreamSynchronize(s)); if (st == AF_SUCCESS) st = h__; }
    if (ws) (void)hipFree(ws);
    (void)hipFree(out);
    (void)hipFree(in);
    if (st != AF_SUCCESS) return st;
    {
        std::lock_guard<std::mutex> lk(plan->tune_mu);
        plan->tuned[{batch, planar}] = best;
    }
    {
        std::lock_guard<std::mutex> lk(g_wisdom_mu);
        g_wisdom[wisdom_key(plan->desc, batch, planar)] = best;
    }
    if (report && report_len) std::snprintf(report, report_len, "%s", text.c_str());
    return AF_SUCCESS;
}

af_status af_wisdom_export(char *buf, size_t buflen, size_t *needed) {
    std::string text;
    {
        std::lock_guard<std::mutex> lk(g_wisdom_mu);
        for (const auto &kv : g_wisdom) {
            text += kv.first + " :";
            for (const PassTune &t : kv.second) text += " " + std::to_string(t.B) + "/" + std::to_string(t.order);
            text += "\n";
        }
    }
    if (needed) *needed = text.size() + 1;
    if (!buf) return buflen ? AF_INVALID_ARG : AF_SUCCESS;
    if (buflen < text.size() + 1) return AF_INVALID_ARG;
    std::memcpy(buf, text.c_str(), text.size() + 1);
    return AF_SUCCESS;
}

af_status af_wisdom_import(const char *text) {
    if (!text) return AF_INVALID_ARG;
    std::map<std::string, std::vector<PassTune>> add;
    const char *p = text;
    while (*p) {
        const char *e = std::strchr(p, '\n');
        const std::string line(p, e ? (size_t)(e - p) : std::strlen(p));
        p = e ? e + 1 : p + line.size();
        if (line.empty()) continue;
        const size_t c = line.find(" :");
        if (c == std::string::npos) return AF_INVALID_ARG;
        char pr[8];
        long long alg, len, batch, planar;
        if (std::sscanf(line.c_str(), "%7s %lld %lld %lld %lld", pr, &alg, &len, &batch, &planar) != 5)
            return AF_INVALID_ARG;
        if ((std::strcmp(pr, "f32") && std::strcmp(pr, "f64")) || len < 0 || batch <= 0 || planar < 0 || planar > 1)
            return AF_INVALID_ARG;
        std::vector<PassTune> v;
        const char *q = line.c_str() + c + 2;
        int B, o, used;
        while (std::sscanf(q, " %d/%d%n", &B, &o, &used) == 2) {
            if (B < 0 || B > 64 || (B & (B - 1)) || o < 0 || o > 2) return AF_INVALID_ARG;
            v.push_back({B, o});
            q += used;
        }
        add[line.substr(0, c)] = v;
    }
    std::lock_guard<std::mutex> lk(g_wisdom_mu);
    for (auto &kv : add) g_wisdom[kv.first] = kv.second;
    return AF_SUCCESS;
}

af_status af_vjp_execute(af_plan plan, int64_t batch, const void *g_re, const void *g_im, void *out_re, void *out_im,
                         void *workspace, void *stream) {
    if (!plan || plan->desc.kind != AF_C2C || batch < 0) return AF_INVALID_ARG;
    if (batch > 0 && plan->desc.length > 0 && bad_ptrs(g_re, out_re)) return AF_INVALID_ARG;
    // adjoint of the unnormalised DFT with sign s: the DFT with sign -s
    if (plan->dual)
        return run_c2c(plan->dual, plan->dual->desc.sign, batch, g_re, g_im, out_re, out_im, workspace,
                       (hipStream_t)stream);
    return run_c2c(plan, -plan->desc.sign, batch, g_re, g_im, out_re, out_im, workspace, (hipStream_t)stream);
}

af_status af_to_spectrum(af_plan plan, int64_t batch, const void *x, void *out_re, void *out_im, void *workspace,
                         void *stream) {
    (void)workspace;
    if (!plan || plan->desc.kind != AF_R2C_FULL || batch < 0) return AF_INVALID_ARG;
    const int64_t n = plan->desc.length, n2 = n / 2;
    if (n == 0 || batch == 0) return AF_SUCCESS;
    if (bad_ptrs(x, out_re)) return AF_INVALID_ARG;
    if (!plan->real_fused)
        return real_general(plan, 0, batch, x, nullptr, nullptr, nullptr, out_re, out_im, nullptr, (hipStream_t)stream);
    AfPass d = base_pass(plan->prec, (int)n2, plan->core.sh.r, plan->core.tw, plan->core.tw_d, choose_B_rows(plan->prec, n2, batch), 1,
                         batch, plan->desc.sign, plan->plain);
    set_in(d, ep(x, nullptr, n2), 1, 0);
    set_out(d, ep(out_re, out_im, n), 1, 0);
    d.flags |= AFP_REAL_UNTANGLE;
    d.aux_tw = plan->aux_tw;
    HIPCHK(launch_pass(plan->prec, d, (hipStream_t)stream));
    return AF_SUCCESS;
}

af_status af_from_spectrum(af_plan plan, int64_t batch, const void *in_re, const void *in_im, void *y,
                           void *workspace, void *stream) {
    (void)workspace;
    if (!plan || plan->desc.kind != AF_C2R_FULL || batch < 0) return AF_INVALID_ARG;
    const int64_t n = plan->desc.length, n2 = n / 2;
    if (n == 0 || batch == 0) return AF_SUCCESS;
    if (bad_ptrs(in_re, y)) return AF_INVALID_ARG;
    if (!plan->real_fused)
        return real_general(plan, 1, batch, in_re, in_im, nullptr, nullptr, y, nullptr, nullptr, (hipStream_t)stream);
    AfPass d = base_pass(plan->prec, (int)n2, plan->core.sh.r, plan->core.tw, plan->core.tw_d, choose_B_rows(plan->prec, n2, batch), 1,
                         batch, plan->desc.sign, plan->plain);
    set_in(d, ep(in_re, in_im, n), 1, 0);
    set_out(d, ep(y, nullptr, n2), 1, 0);
    d.flags |= AFP_REAL_ENTANGLE;
    d.aux_tw = plan->aux_tw;
    HIPCHK(launch_pass(plan->prec, d, (hipStream_t)stream));
    return AF_SUCCESS;
}

af_status af_vjp_to_spectrum(af_plan plan, int64_t batch, const void *g_re, const void *g_im, void *dx,
                             void *workspace, void *stream) {
    (void)workspace;
    if (!plan || plan->desc.kind != AF_R2C_FULL || batch < 0) return AF_INVALID_ARG;
    if (plan->desc.sign != -1) return AF_UNSUPPORTED;  // the DFT's real restriction only
    const int64_t n = plan->desc.length, n2 = n / 2;
    if (n == 0 || batch == 0) return AF_SUCCESS;
    if (bad_ptrs(g_re, dx)) return AF_INVALID_ARG;
    if (!plan->real_fused)
        return real_general(plan, 2, batch, g_re, g_im, nullptr, nullptr, dx, nullptr, nullptr, (hipStream_t)stream);
    AfPass d = base_pass(plan->prec, (int)n2, plan->core.sh.r, plan->core.tw, plan->core.tw_d, choose_B_rows(plan->prec, n2, batch), 1,
                         batch, +1, plan->plain);
    set_in(d, ep(g_re, g_im, n), 1, 0);
    set_out(d, ep(dx, nullptr, n2), 1, 0);
    d.flags |= AFP_UNTANGLE_ADJ;
    d.aux_tw = plan->aux_tw;
    HIPCHK(launch_pass(plan->prec, d, (hipStream_t)stream));
    return AF_SUCCESS;
}

af_status af_vjp_from_spectrum(af_plan plan, int64_t batch, const void *g, void *out_re, void *out_im,
                               void *workspace, void *stream) {
    (void)workspace;
    if (!plan || plan->desc.kind != AF_C2R_FULL || batch < 0) return AF_INVALID_ARG;
    const int64_t n = plan->desc.length, n2 = n / 2;
    if (n == 0 || batch == 0) return AF_SUCCESS;
    if (bad_ptrs(g, out_re)) return AF_INVALID_ARG;
    if (!plan->real_fused)
        return real_general(plan, 3, batch, g, nullptr, nullptr, nullptr, out_re, out_im, nullptr, (hipStream_t)stream);
    AfPass d = base_pass(plan->prec, (int)n2, plan->core.sh.r, plan->core.tw, plan->core.tw_d, choose_B_rows(plan->prec, n2, batch), 1,
                         batch, -plan->desc.sign, plan->plain);
    set_in(d, ep(g, nullptr, n2), 1, 0);
    set_out(d, ep(out_re, out_im, n), 1, 0);
    d.flags |= AFP_ENTANGLE_ADJ;
    d.aux_tw = plan->aux_tw;
    HIPCHK(launch_pass(plan->prec, d, (hipStream_t)stream));
    return AF_SUCCESS;
}

static AfPass real_conv_geom(af_plan plan, int64_t batch) {
    const int64_t n2 = plan->desc.length / 2;
    AfPass d = base_pass(plan->prec, (int)n2, plan->core.sh.r, plan->core.tw, plan->core.tw_d, 1, 1, batch, -1, plan->plain);
    d.aux_tw = plan->aux_tw;
    d.lds_bytes = 2 * (pidx_h(d.ld, d.padshift) + 1) * elem_bytes(plan->prec);
    return d;
}

af_status af_real_convolve(af_plan plan, int64_t batch, const void *x, const void *k, void *y, void *stream) {
    if (!plan || plan->desc.kind != AF_REAL_CONV || batch < 0) return AF_INVALID_ARG;
    if (batch == 0) return AF_SUCCESS;
    if (!x || !k || !y) return AF_INVALID_ARG;
    if (!plan->real_fused)
        return real_general(plan, 4, batch, x, nullptr, k, nullptr, y, nullptr, nullptr, (hipStream_t)stream);
    HIPCHK(launch_real_conv(plan->prec, real_conv_geom(plan, batch), x, k, y, (hipStream_t)stream));
    return AF_SUCCESS;
}

af_status af_vjp_real_convolve(af_plan plan, int64_t batch, const void *x, const void *k, const void *w, void *dx,
                               void *dk, void *stream) {
    if (!plan || plan->desc.kind != AF_REAL_CONV || batch < 0) return AF_INVALID_ARG;
    if (batch == 0) return AF_SUCCESS;
    if (!x || !k || !w || !dx || !dk) return AF_INVALID_ARG;
    if (!plan->real_fused)
        return real_general(plan, 5, batch, x, nullptr, k, w, dx, nullptr, dk, (hipStream_t)stream);
    HIPCHK(launch_real_conv_vjp(plan->prec, real_conv_geom(plan, batch), x, k, w, dx, dk, (hipStream_t)stream));
    return AF_SUCCESS;
}

// ---- two-for-one real packing (Real.hs:129-236) ----

af_status af_two_to_spectrum(af_plan plan, int64_t batch, const void *x, const void *y, void *a_re, void *a_im,
                             void *b_re, void *b_im, void *stream) {
    if (!plan || plan->desc.kind != AF_C2C || batch < 0) return AF_INVALID_ARG;
    const int64_t n = plan->desc.length;
    if (n == 0 || batch == 0) return AF_SUCCESS;
    if (!x || !y || !a_re || !b_re) return AF_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    // z = x + i y is the planar complex input (re = x, im = y); its spectrum goes
    // to stream-ordered scratch, then untangles into a and b
    Scratch sz;
    AFCHK(sz.get((size_t)batch * (size_t)n * elem_bytes(plan->desc.precision), nullptr, s));
    AFCHK(run_c2c(plan, plan->desc.sign, batch, x, y, sz.p, nullptr, nullptr, s));
    HIPCHK(launch_two_pack(plan->prec, true, sz.p, nullptr, nullptr, nullptr, a_re, a_im, b_re, b_im, n, batch, s));
    return AF_SUCCESS;
}

af_status af_two_from_spectrum(af_plan plan, int64_t batch, const void *a_re, const void *a_im, const void *b_re,
                               const void *b_im, void *x, void *y, void *stream) {
    if (!plan || plan->desc.kind != AF_C2C || batch < 0) return AF_INVALID_ARG;
    const int64_t n = plan->desc.length;
    if (n == 0 || batch == 0) return AF_SUCCESS;
    if (!x || !y || !a_re || !b_re) return AF_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    // entangled spectrum a + i b to scratch; its transform's real and imaginary
    // parts are the two signals (planar output: re -> x, im -> y)
    Scratch sz;
    AFCHK(sz.get((size_t)batch * (size_t)n * elem_bytes(plan->desc.precision), nullptr, s));
    HIPCHK(launch_two_pack(plan->prec, false, a_re, a_im, b_re, b_im, sz.p, nullptr, nullptr, nullptr, n, batch, s));
    return run_c2c(plan, plan->desc.sign, batch, sz.p, nullptr, x, y, nullptr, s);
}

af_status af_untangle_spectra(int32_t precision, int64_t batch, int64_t n, const void *z_re, const void *z_im,
                              void *a_re, void *a_im, void *b_re, void *b_im, void *stream) {
    if ((precision != AF_F64 && precision != AF_F32) || batch < 0 || n < 0) return AF_INVALID_ARG;
    if (n == 0 || batch == 0) return AF_SUCCESS;
    if (!z_re || !a_re || !b_re) return AF_INVALID_ARG;
    // each output reads two input positions: no aliasing of input and outputs
    for (const void *o : {(const void *)a_re, (const void *)a_im, (const void *)b_re, (const void *)b_im})
        if (o && (o == z_re || o == z_im)) return AF_INVALID_ARG;
    AFCHK(device_init());
    HIPCHK(launch_two_pack(precision == AF_F64 ? PREC_F64 : PREC_F32, true, z_re, z_im, nullptr, nullptr, a_re, a_im,
                           b_re, b_im, n, batch, (hipStream_t)stream));
    return AF_SUCCESS;
}

af_status af_entangle_spectra(int32_t precision, int64_t batch, int64_t n, const void *c_re, const void *c_im,
                              const void *d_re, const void *d_im, void *z_re, void *z_im, void *stream) {
    if ((precision != AF_F64 && precision != AF_F32) || batch < 0 || n < 0) return AF_INVALID_ARG;
    if (n == 0 || batch == 0) return AF_SUCCESS;
    if (!c_re || !d_re || !z_re) return AF_INVALID_ARG;
    AFCHK(device_init());
    HIPCHK(launch_two_pack(precision == AF_F64 ? PREC_F64 : PREC_F32, false, c_re, c_im, d_re, d_im, z_re, z_im,
                           nullptr, nullptr, n, batch, (hipStream_t)stream));
    return AF_SUCCESS;
}

// ---- multi-dimensional transforms (Preprocessed.hs:138-161, Adhoc.hs:365-380) ----

static bool plans_c2c(std::initializer_list<af_plan> ps) {
    int prec = -1;
    for (af_plan p : ps) {
        if (!p || p->desc.kind != AF_C2C) return false;
        if (prec >= 0 && p->desc.precision != prec) return false;
        prec = p->desc.precision;
    }
    return true;
}

af_status af_execute_2d(af_plan inner, af_plan outer, int64_t batch, const void *in_re, const void *in_im,
                        void *out_re, void *out_im, void *stream) {
    if (!plans_c2c({inner, outer}) || batch < 0) return AF_INVALID_ARG;
    const int64_t W = inner->desc.length, H = outer->desc.length;
    if (batch == 0 || W == 0 || H == 0) return AF_SUCCESS;
    if (!in_re || !out_re) return AF_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    AFCHK(run_c2c(inner, inner->desc.sign, batch * H, in_re, in_im, out_re, out_im, nullptr, s));
    return run_axis(outer, batch, W, out_re, out_im, out_re, out_im, s);
}

af_status af_execute_3d(af_plan p0, af_plan p1, af_plan p2, int64_t batch, const void *in_re, const void *in_im,
                        void *out_re, void *out_im, void *stream) {
    if (!plans_c2c({p0, p1, p2}) || batch < 0) return AF_INVALID_ARG;
    const int64_t N = p0->desc.length, M = p1->desc.length, K = p2->desc.length;
    if (batch == 0 || N == 0 || M == 0 || K == 0) return AF_SUCCESS;
    if (!in_re || !out_re) return AF_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    AFCHK(run_c2c(p0, p0->desc.sign, batch * K * M, in_re, in_im, out_re, out_im, nullptr, s));
    AFCHK(run_axis(p1, batch * K, N, out_re, out_im, out_re, out_im, s));
    return run_axis(p2, batch, M * N, out_re, out_im, out_re, out_im, s);
}

af_status af_transpose(int32_t precision, int64_t batch, int64_t rows, int64_t cols, const void *in_re,
                       const void *in_im, void *out_re, void *out_im, void *stream) {
    if ((precision != AF_F64 && precision != AF_F32) || batch < 0 || rows < 0 || cols < 0) return AF_INVALID_ARG;
    if (batch == 0 || rows == 0 || cols == 0) return AF_SUCCESS;
    if (!in_re || !out_re || in_re == out_re || (in_im && in_im == out_im)) return AF_INVALID_ARG;
    if ((rows + 31) / 32 > 65535) return AF_UNSUPPORTED;
    AFCHK(device_init());
    HIPCHK(launch_transpose(precision == AF_F64 ? PREC_F64 : PREC_F32, in_re, in_im, out_re, out_im, rows, cols, batch,
                            (hipStream_t)stream));
    return AF_SUCCESS;
}

af_status af_two_to_spectrum_2d(af_plan inner, af_plan outer, int64_t batch, const void *x, const void *y,
                                void *a_re, void *a_im, void *b_re, void *b_im, void *stream) {
    if (!plans_c2c({inner, outer}) || batch < 0) return AF_INVALID_ARG;
    const int64_t W = inner->desc.length, H = outer->desc.length;
    if (batch == 0 || W == 0 || H == 0) return AF_SUCCESS;
    if (!x || !y || !a_re || !b_re) return AF_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    Scratch sz;
    AFCHK(sz.get((size_t)(batch * H * W) * elem_bytes(inner->desc.precision), nullptr, s));
    AFCHK(af_execute_2d(inner, outer, batch, x, y, sz.p, nullptr, stream));
    HIPCHK(launch_two_pack(inner->prec, true, sz.p, nullptr, nullptr, nullptr, a_re, a_im, b_re, b_im, W, batch * H,
                           s, H));
    return AF_SUCCESS;
}

af_status af_two_from_spectrum_2d(af_plan inner, af_plan outer, int64_t batch, const void *a_re, const void *a_im,
                                  const void *b_re, const void *b_im, void *x, void *y, void *stream) {
    if (!plans_c2c({inner, outer}) || batch < 0) return AF_INVALID_ARG;
    const int64_t W = inner->desc.length, H = outer->desc.length;
    if (batch == 0 || W == 0 || H == 0) return AF_SUCCESS;
    if (!x || !y || !a_re || !b_re) return AF_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    Scratch sz;
    AFCHK(sz.get((size_t)(batch * H * W) * elem_bytes(inner->desc.precision), nullptr, s));
    HIPCHK(launch_two_pack(inner->prec, false, a_re, a_im, b_re, b_im, sz.p, nullptr, nullptr, nullptr, W, batch * H,
                           s));
    return af_execute_2d(inner, outer, batch, sz.p, nullptr, x, y, stream);
}

af_status af_untangle_spectra_2d(int32_t precision, int64_t batch, int64_t height, int64_t width, const void *z_re,
                                 const void *z_im, void *a_re, void *a_im, void *b_re, void *b_im, void *stream) {
    if ((precision != AF_F64 && precision != AF_F32) || batch < 0 || height < 0 || width < 0) return AF_INVALID_ARG;
    if (batch == 0 || height == 0 || width == 0) return AF_SUCCESS;
    if (!z_re || !a_re || !b_re) return AF_INVALID_ARG;
    for (const void *o : {(const void *)a_re, (const void *)a_im, (const void *)b_re, (const void *)b_im})
        if (o && (o == z_re || o == z_im)) return AF_INVALID_ARG;
    AFCHK(device_init());
    HIPCHK(launch_two_pack(precision == AF_F64 ? PREC_F64 : PREC_F32, true, z_re, z_im, nullptr, nullptr, a_re, a_im,
                           b_re, b_im, width, batch * height, (hipStream_t)stream, height));
    return AF_SUCCESS;
}

af_status af_convolve_cyclic(af_plan plan, int64_t batch, const void *x_re, const void *x_im, const void *y_re,
                             const void *y_im, void *out_re, void *out_im, void *workspace, void *stream) {
    if (!plan || plan->desc.kind != AF_CONV_CYCLIC || batch < 0) return AF_INVALID_ARG;
    if (batch == 0) return AF_SUCCESS;
    if (!x_re || !y_re || !out_re) return AF_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = plan->desc.length;
    const int prec = plan->prec;
    const size_t eb = (size_t)elem_bytes(plan->desc.precision);
    Scratch sc;
    AFCHK(sc.get(plan->scratch_per_row() * (size_t)batch, workspace, s));
    const Endpoint X = ep(x_re, x_im, n), Y = ep(y_re, y_im, n), O = ep(out_re, out_im, n);
    std::vector<AfPass> ps;
    if (plan->sub) {
        // convolveCyclic (Planned.hs:779-821): F+( F-(x) / n * F-(y) ) with general sub-plans
        void *s1 = sc.p, *s2 = (char *)sc.p + (size_t)batch * n * eb;
        AFCHK(af_execute(plan->sub, batch, x_re, x_im, s1, nullptr, nullptr, s));
        AFCHK(af_execute(plan->sub, batch, y_re, y_im, s2, nullptr, nullptr, s));
        AfPass m = base_pass(prec, 1, {}, nullptr, nullptr, 64, n, batch, -1, plan->plain);
        set_in(m, ep(s1, nullptr, n), 0, 1);
        set_out(m, ep(s1, nullptr, n), 0, 1);
        m.flags |= AFP_CONV | AFP_SPEC_PLAIN | AFP_SCALE;
        m.spec = s2;
        m.spec_bs = n;
        m.spec_ls = 1;
        m.scale = 1.0 / (double)n;
        HIPCHK(launch_pass(prec, m, s));
        return run_c2c(plan->sub, +1, batch, s1, nullptr, out_re, out_im, nullptr, s);
    }
    const bool four = plan->core.sh.kind == CORE_FOUR;
    void *spec = sc.p;
    void *scr = four ? (char *)sc.p + (size_t)batch * n * eb : nullptr;
    // spectrum of y, scaled by 1/n, in the order the convolution passes read it
    c2c_passes(prec, plan->core, batch, -1, plan->plain, Y, ep(spec, nullptr, n), scr, true, ps);
    ps.back().flags |= AFP_SCALE;
    ps.back().scale = 1.0 / (double)n;
    conv_passes(prec, plan->core, batch, plan->plain, X, O, spec, n, true, scr, ps);
    return launch_all(prec, ps, s);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Karatsuba
// ---------------------------------------------------------------------------

namespace {

// P = 2^d s >= Lp with s <= 64: the padded length the LDS kernel works on
int64_t karatsuba_P(int64_t Lp) {
    int d = 0;
    int64_t s = Lp;
    while (s > 64) {
        ++d;
        s = (Lp + ((int64_t)1 << d) - 1) >> d;
    }
    return s << d;
}

// Pinned host staging for the small per-call tables copied to the device (the
// Karatsuba pair table): a pool of hipHostMalloc buffers, each reused once the
// event recorded after its copy reports completion, so a call neither copies
// from pageable memory nor waits on the stream.  The buffers live until the
// process ends.
struct Staging {
    void *host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
};
std::mutex g_stage_mu;
std::vector<Staging *> g_stage_free;

af_status stage_to_device(void *dst, const void *src, size_t bytes, hipStream_t s) {
    Staging *b = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        for (size_t i = 0; i < g_stage_free.size(); ++i)
            if (g_stage_free[i]->cap >= bytes && hipEventQuery(g_stage_free[i]->done) == hipSuccess) {
                b = g_stage_free[i];
                g_stage_free.erase(g_stage_free.begin() + (std::ptrdiff_t)i);
                break;
            }
    }
    if (!b) {
        std::unique_ptr<Staging> nb(new Staging);
        nb->cap = std::max<size_t>(bytes, (size_t)64 << 10);
        HIPCHK(hipHostMalloc(&nb->host, nb->cap, hipHostMallocDefault));
        const hipError_t e = hipEventCreateWithFlags(&nb->done, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipHostFree(nb->host);
            HIPCHK(e);
        }
        b = nb.release();
    }
    std::memcpy(b->host, src, bytes);
    af_status st = HIPST(hipMemcpyAsync(dst, b->host, bytes, hipMemcpyHostToDevice, s));
    if (st == AF_SUCCESS) st = HIPST(hipEventRecord(b->done, s));
    if (st != AF_SUCCESS) (void)hipStreamSynchronize(s);  // the buffer is free again either way
    std::lock_guard<std::mutex> lk(g_stage_mu);
    g_stage_free.push_back(b);
    return st;
}

// Pairs (meta: input offset, length, output offset per pair) whose padded
// length fits the LDS kernel, in one launch, longest first for load balance.
af_status karatsuba_leaves(std::vector<int64_t> meta, const int64_t *x, const int64_t *y, int64_t *z, int cyclic,
                           hipStream_t s) {
    const int64_t npairs = (int64_t)meta.size() / 3;
    if (npairs == 0) return AF_SUCCESS;
    int64_t maxP = 0;
    std::vector<int64_t> order(npairs);
    for (int64_t i = 0; i < npairs; ++i) {
        order[i] = i;
        maxP = std::max(maxP, karatsuba_P(meta[3 * i + 1]));
    }
    if (maxP > kKaratsubaMaxP) return AF_INTERNAL;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return meta[3 * a + 1] > meta[3 * b + 1]; });
    std::vector<int64_t> sorted(3 * npairs);
    for (int64_t i = 0; i < npairs; ++i)
        for (int j = 0; j < 3; ++j) sorted[3 * i + j] = meta[3 * order[i] + j];
    void *dm = nullptr;  // the pair table, then one int32 per pair for the kernels' hand-over
    HIPCHK(hipMallocAsync(&dm, sorted.size() * 8 + (size_t)npairs * 4, s));
    af_status st = stage_to_device(dm, sorted.data(), sorted.size() * 8, s);
    if (st == AF_SUCCESS)
        st = HIPST(launch_karatsuba(x, y, z, (const int64_t *)dm, npairs, (int)maxP, cyclic,
                                    (int32_t *)((char *)dm + sorted.size() * 8), s));
    (void)hipFreeAsync(dm, s);
    return st;
}

// Any lengths.  Pairs longer than the LDS kernel's limit get their top
// Karatsuba levels in global memory: split level by level (k_kar_split) until
// every piece fits, the pieces go through the LDS kernel in one launch, and the
// joins (k_kar_join) run back up, the root joins writing the outputs (wrapped
// when cyclic).
af_status karatsuba_run(const std::vector<int64_t> &meta_in, int64_t npairs, const int64_t *x, const int64_t *y,
                        int64_t *z, int cyclic, hipStream_t s) {
    AFCHK(device_init());
    std::call_once(g_attr_flag, set_kernel_attributes);
    if (npairs == 0) return AF_SUCCESS;
    // a node of the top levels: a root (a caller's pair, offsets into x, y, z)
    // or a piece (offsets into the workspace arrays WX, WY, WZ)
    struct Node {
        int64_t L, xo, zo;
        bool root;
    };
    struct Split {
        Node n;
        int64_t cxo, zco;  // children's operands in WX/WY, children's products in WZ
    };
    std::vector<int64_t> small, pieces;  // leaf meta: roots, workspace pieces
    std::vector<Node> cur;
    for (int64_t p = 0; p < npairs; ++p) {
        const int64_t off = meta_in[3 * p], L = meta_in[3 * p + 1], zo = meta_in[3 * p + 2];
        if (L <= 0) continue;
        if (karatsuba_P(L) <= kKaratsubaMaxP) small.insert(small.end(), {off, L, zo});
        else cur.push_back({L, off, zo, true});
    }
    std::vector<std::vector<Split>> levels;
    int64_t wx = 0, wz = 0;
    while (!cur.empty()) {
        std::vector<Split> lv;
        std::vector<Node> next;
        for (const Node &n : cur) {
            const int64_t h = (n.L + 1) / 2;
            Split sp{n, wx, wz};
            wx += 3 * h;
            wz += 3 * (2 * h - 1);
            for (int c = 0; c < 3; ++c) {
                const Node ch{h, sp.cxo + c * h, sp.zco + c * (2 * h - 1), false};
                if (karatsuba_P(h) <= kKaratsubaMaxP) pieces.insert(pieces.end(), {ch.xo, ch.L, ch.zo});
                else next.push_back(ch);
            }
            lv.push_back(sp);
        }
        levels.push_back(std::move(lv));
        cur.swap(next);
    }
    af_status st = karatsuba_leaves(small, x, y, z, cyclic, s);
    if (st != AF_SUCCESS || levels.empty()) return st;
    int64_t *ws = nullptr;
    HIPCHK(hipMallocAsync((void **)&ws, (size_t)(2 * wx + wz) * sizeof(int64_t), s));
    int64_t *WX = ws, *WY = ws + wx, *WZ = ws + 2 * wx;
    std::vector<KarSplitTask> splits;
    std::vector<KarJoinTask> joins;
    std::vector<size_t> split_at, join_at;  // first task of each level
    for (const auto &lv : levels) {
        split_at.push_back(splits.size());
        join_at.push_back(joins.size());
        for (const Split &sp : lv) {
            const Node &n = sp.n;
            splits.push_back({n.root ? x + n.xo : WX + n.xo, n.root ? y + n.xo : WY + n.xo, n.L, WX + sp.cxo,
                              WY + sp.cxo});
            joins.push_back({WZ + sp.zco, n.L, n.root ? z + n.zo : WZ + n.zo, n.root ? (int64_t)cyclic : 0});
        }
    }
    split_at.push_back(splits.size());
    join_at.push_back(joins.size());
    const size_t sb = splits.size() * sizeof(KarSplitTask), jb = joins.size() * sizeof(KarJoinTask);
    std::vector<unsigned char> tab(sb + jb);
    std::memcpy(tab.data(), splits.data(), sb);
    std::memcpy(tab.data() + sb, joins.data(), jb);
    unsigned char *dt = nullptr;
    st = HIPST(hipMallocAsync((void **)&dt, tab.size(), s));
    if (st == AF_SUCCESS) st = stage_to_device(dt, tab.data(), tab.size(), s);
    const KarSplitTask *dsplit = (const KarSplitTask *)dt;
    const KarJoinTask *djoin = (const KarJoinTask *)(dt + sb);
    for (size_t l = 0; st == AF_SUCCESS && l < levels.size(); ++l)
        st = HIPST(launch_kar_split(dsplit + split_at[l], (int64_t)(split_at[l + 1] - split_at[l]), s));
    if (st == AF_SUCCESS) st = karatsuba_leaves(pieces, WX, WY, WZ, 0, s);
    for (size_t l = levels.size(); st == AF_SUCCESS && l-- > 0;)
        st = HIPST(launch_kar_join(djoin + join_at[l], (int64_t)(join_at[l + 1] - join_at[l]), s));
    if (dt) (void)hipFreeAsync(dt, s);
    (void)hipFreeAsync(ws, s);
    return st;
}

}  // namespace

extern "C" {

af_status af_karatsuba_i64(int64_t batch, int64_t len, const int64_t *x, const int64_t *y, int64_t *z, void *stream) {
    if (batch < 0 || len < 0) return AF_INVALID_ARG;
    if (batch == 0 || len == 0) return AF_SUCCESS;
    if (!x || !y || !z) return AF_INVALID_ARG;
    std::vector<int64_t> meta(3 * batch);
    for (int64_t r = 0; r < batch; ++r) {
        meta[3 * r] = r * len;
        meta[3 * r + 1] = len;
        meta[3 * r + 2] = r * (2 * len - 1);
    }
    return karatsuba_run(meta, batch, x, y, z, 0, (hipStream_t)stream);
}

af_status af_karatsuba_cyclic_i64(int64_t batch, int64_t len, const int64_t *x, const int64_t *y, int64_t *z,
                                  void *stream) {
    if (batch < 0 || len < 0) return AF_INVALID_ARG;
    if (batch == 0 || len == 0) return AF_SUCCESS;
    if (!x || !y || !z) return AF_INVALID_ARG;
    std::vector<int64_t> meta(3 * batch);
    for (int64_t r = 0; r < batch; ++r) {
        meta[3 * r] = r * len;
        meta[3 * r + 1] = len;
        meta[3 * r + 2] = r * len;
    }
    return karatsuba_run(meta, batch, x, y, z, 1, (hipStream_t)stream);
}

af_status af_karatsuba_complex_i64(int64_t batch, int64_t len, int32_t cyclic, const int64_t *xr, const int64_t *xi,
                                   const int64_t *yr, const int64_t *yi, int64_t *zr, int64_t *zi, void *stream) {
    if (batch < 0 || len < 0 || (cyclic != 0 && cyclic != 1)) return AF_INVALID_ARG;
    if (batch == 0 || len == 0) return AF_SUCCESS;
    if (!xr || !xi || !yr || !yi || !zr || !zi) return AF_INVALID_ARG;
    AFCHK(device_init());
    hipStream_t s = (hipStream_t)stream;
    const int64_t nz = cyclic ? len : 2 * len - 1, count = batch * len, zcount = batch * nz;
    // the three operand pairs, then the three products, in one workspace
    int64_t *ws = nullptr;
    HIPCHK(hipMallocAsync((void **)&ws, (size_t)(6 * count + 3 * zcount) * sizeof(int64_t), s));
    int64_t *x3 = ws, *y3 = ws + 3 * count, *p3 = ws + 6 * count;
    std::vector<int64_t> meta(9 * batch);
    for (int64_t q = 0; q < 3 * batch; ++q) {
        meta[3 * q] = q * len;
        meta[3 * q + 1] = len;
        meta[3 * q + 2] = q * nz;
    }
    af_status st = HIPST(launch_gauss_split(xr, xi, yr, yi, x3, y3, count, s));
    if (st == AF_SUCCESS) st = karatsuba_run(meta, 3 * batch, x3, y3, p3, cyclic, s);
    if (st == AF_SUCCESS) st = HIPST(launch_gauss_join(p3, zr, zi, zcount, s));
    (void)hipFreeAsync(ws, s);
    return st;
}

af_status af_karatsuba_ragged_i64(int64_t npairs, const int64_t *offsets, const int64_t *x, const int64_t *y,
                                  int64_t *z, void *stream) {
    if (npairs < 0 || (npairs > 0 && !offsets)) return AF_INVALID_ARG;
    if (npairs == 0) return AF_SUCCESS;
    if (!x || !y || !z) return AF_INVALID_ARG;
    if (offsets[0] != 0) return AF_INVALID_ARG;
    std::vector<int64_t> meta(3 * npairs);
    for (int64_t p = 0; p < npairs; ++p) {
        const int64_t L = offsets[p + 1] - offsets[p];
        if (L < 1) return AF_INVALID_ARG;  // equal, non-empty operand lengths per pair
        meta[3 * p] = offsets[p];
        meta[3 * p + 1] = L;
        meta[3 * p + 2] = 2 * offsets[p] - p;
    }
    return karatsuba_run(meta, npairs, x, y, z, 0, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Host index maps and number theory
// ---------------------------------------------------------------------------

af_status af_rader_permutations(int64_t p, int64_t *perm, int64_t *rev, int64_t *inv) {
    if (!is_prime(p) || p < 3 || !perm || !rev || !inv) return AF_INVALID_ARG;
    const int64_t L = p - 1, g = multiplicative_generator(p);
    int64_t x = 1;
    for (int64_t i = 0; i < L; ++i) {
        perm[i] = x - 1;
        x = (int64_t)(((unsigned __int128)x * (uint64_t)g) % (uint64_t)p);
    }
    for (int64_t i = 0; i < L; ++i) rev[i] = perm[(L - i) % L];
    for (int64_t i = 0; i < L; ++i) inv[perm[i]] = i;
    return AF_SUCCESS;
}

af_status af_good_thomas_maps(int64_t n, int64_t m, int64_t *in_map, int64_t *out_map) {
    if (n < 1 || m < 1 || !in_map || !out_map) return AF_INVALID_ARG;
    int64_t a = n, b = m;
    while (b) { int64_t t = a % b; a = b; b = t; }
    if (a != 1) return AF_INVALID_ARG;
    const int64_t nm = n * m;
    for (int64_t k = 0; k < m; ++k)
        for (int64_t j = 0; j < n; ++j) in_map[k * n + j] = (n * k + m * j) % nm;
    for (int64_t q = 0; q < nm; ++q) out_map[q] = (q % n) * m + (q % m);
    return AF_SUCCESS;
}

int64_t af_multiplicative_generator(int64_t p) { return is_prime(p) ? multiplicative_generator(p) : -1; }
int64_t af_ceiling_5smooth(int64_t n) { return n < 1 ? 1 : ceil_5smooth(n); }
int64_t af_ceiling_power_of_two(int64_t n) { return ceil_pow2(n); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Uneven batches: all rows' passes as grouped launches in one hipGraph
// ---------------------------------------------------------------------------

struct af_batch_s {
    int64_t count = 0;
    int sign = -1;
    int precision = AF_F32;
    int prec = PREC_F32;
    std::vector<int64_t> lengths, offsets;
    std::map<int64_t, af_plan> plans;
    // every row's passes built with placeholder pointers: the launch shapes
    std::vector<std::vector<AfPass>> shapes;
    std::vector<size_t> scratch_need;  // bytes of intermediates per row (256-aligned)
    std::vector<size_t> scratch_off;
    // Rows with intermediates run in chunks of at most chunk_bytes of them
    // (0: one chunk): all passes of one chunk, then the next chunk, every chunk
    // on the same scratch region, so that a chunk's intermediates can stay in
    // the memory-side cache between its passes.
    int64_t chunk_bytes = 0;
    int nchunks = 0;
    // lanes = 2: the rows without intermediates (chunk 0 when there are any)
    // run on a branch of the graph beside the others, so that their small
    // launches overlap the large ones; 1: every launch in one sequence;
    // 3 (unchunked only): as 2, and the rows with intermediates that some pass
    // runs on the generic kernel (mixed radix, Rader) on a third branch beside
    // those all of whose passes run on the power-of-two kernels (Bluestein
    // pads), each set a chunk of its own with its own scratch
    int lanes = 1;
    bool light_chunk = false;
    int mixed_chunk = -1;  // lanes = 3: the chunk on the third branch
    DevBuf scratch_buf, descs_buf, prefix_buf;
    void *scratch = nullptr;
    struct Group {
        int phase = 0, nthreads = 0, lds = 0, chunk = 0;
        bool full = false;  // some descriptor sets a flag of the full pass kernel
        // power-of-two passes of compile-time shape (L, B, conv, sign) run on the
        // extended fast kernel instead of the generic one
        bool ext = false, conv = false;
        int L = 0, B = 0, sign = 0;
        // generic groups: the radix order of its descriptors (reorder_radices),
        // chosen by af_batch_tune; 0 = as planned
        int order = 0;
        std::vector<int64_t> rows;  // transform indices
        int64_t desc_off = 0, prefix_off = 0, total_tiles = 0;
    };
    std::vector<Group> groups;
    AfPass *d_descs = nullptr;
    int64_t *d_prefix = nullptr;
    std::vector<AfPass> h_descs;
    std::vector<int64_t> h_prefix;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *last[4] = {nullptr, nullptr, nullptr, nullptr};
    bool uploaded = false;
    std::string text, summary;
    // Calls share the descriptors and the scratch intermediates: the mutex
    // serialises their host side, and `done` (recorded after each graph launch)
    // orders every launch after the previous one, whatever its stream, and keeps
    // the descriptors from being rewritten while a launch may still read them.
    std::mutex mu;
    hipEvent_t done = nullptr;
    bool recorded = false;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~af_batch_s() {
        if (done) (void)hipEventDestroy(done);
        drop_graph();
        for (auto &kv : plans) af_plan_destroy(kv.second);
    }
};

namespace {

Endpoint offset_ep(const void *re, const void *im, int64_t off, int64_t n, size_t rb) {
    Endpoint e;
    if (im) {
        e.re = (const char *)re + off * rb;
        e.im = (const char *)im + off * rb;
    } else {
        e.re = (const char *)re + off * 2 * rb;
        e.im = nullptr;
    }
    e.bs = n;
    return e;
}

// The passes of row i on the given buffers; its intermediates at
// scratch_base + scratch_off[i].
af_status batch_passes(af_batch B, int64_t i, const void *in_re, const void *in_im, void *out_re, void *out_im,
                       char *scratch_base, std::vector<AfPass> &ps) {
    const int64_t n = B->lengths[i];
    if (n == 0) return AF_SUCCESS;
    af_plan P = B->plans.at(n);
    const size_t rb = B->precision == AF_F64 ? 8 : 4;
    return build_c2c(P, B->sign, 1, offset_ep(in_re, in_im, B->offsets[i], n, rb),
                     offset_ep(out_re, out_im, B->offsets[i], n, rb), scratch_base + B->scratch_off[i], ps, false);
}

// (Re)builds the launch groups, the scratch region and the captured graph for
// a chunk budget.  Rows without intermediates form one chunk, launched first.
// Rows with intermediates are ordered by pass count, intermediate size and
// length (rows of one shape share grouped launches) and cut into chunks of at
// most `chunk_bytes` of intermediates (a row larger than the budget is a chunk
// of its own); every chunk's intermediates start at offset 0 of one region.
// Within a chunk, the k-th passes of its rows are grouped by kernel, as the
// whole batch is when chunk_bytes is 0.
af_status batch_build(af_batch B, int64_t chunk_bytes, int lanes) {
    B->drop_graph();
    B->lanes = lanes;
    B->groups.clear();
    B->h_descs.clear();
    B->h_prefix.clear();
    B->uploaded = false;
    B->chunk_bytes = chunk_bytes;
    const int64_t count = B->count;
    std::vector<int64_t> light, heavy;
    for (int64_t i = 0; i < count; ++i) {
        if (B->shapes[i].empty()) continue;
        (B->scratch_need[i] ? heavy : light).push_back(i);
    }
    std::stable_sort(heavy.begin(), heavy.end(), [&](int64_t a, int64_t b) {
        const size_t pa = B->shapes[a].size(), pb = B->shapes[b].size();
        if (pa != pb) return pa < pb;
        if (B->scratch_need[a] != B->scratch_need[b]) return B->scratch_need[a] < B->scratch_need[b];
        return B->lengths[a] < B->lengths[b];
    });
    std::vector<std::vector<int64_t>> chunks;
    if (!light.empty()) chunks.push_back(light);
    B->light_chunk = !light.empty();
    B->mixed_chunk = -1;
    size_t region = 0, used = 0;
    if (lanes == 3 && chunk_bytes == 0) {
        // two chunks of rows with intermediates that may run at once: disjoint scratch
        std::vector<int64_t> mixed, pow2;
        for (int64_t i : heavy) {
            bool all_fast = true;
            for (const AfPass &d : B->shapes[i]) all_fast = all_fast && fast_ext_eligible(B->prec, d);
            (all_fast ? pow2 : mixed).push_back(i);
        }
        for (const std::vector<int64_t> *set : {&mixed, &pow2}) {
            if (set->empty()) continue;
            if (set == &mixed) B->mixed_chunk = (int)chunks.size();
            chunks.emplace_back();
            for (int64_t i : *set) {
                B->scratch_off[i] = used;
                used += B->scratch_need[i];
                chunks.back().push_back(i);
            }
        }
        region = used;
    } else {
        for (int64_t i : heavy) {
            const size_t need = B->scratch_need[i];
            const bool fresh = chunks.empty() || chunks.back().empty() || B->scratch_need[chunks.back()[0]] == 0 ||
                               (chunk_bytes > 0 && used + need > (size_t)chunk_bytes);
            if (fresh) {
                chunks.emplace_back();
                used = 0;
            }
            B->scratch_off[i] = used;
            used += need;
            region = std::max(region, used);
            chunks.back().push_back(i);
        }
    }
    for (int64_t i : light) B->scratch_off[i] = 0;
    B->nchunks = (int)chunks.size();
    // scratch: one region for every chunk
    if (B->scratch_buf.p) {
        (void)hipFree(B->scratch_buf.p);
        B->scratch_buf.p = nullptr;
    }
    HIPCHK(hipMalloc(&B->scratch_buf.p, std::max<size_t>(region, 256)));
    B->scratch = B->scratch_buf.p;
    const int classes[] = {64, 128, 256, 512, 1024};
    int64_t ndesc = 0, nprefix = 0;
    for (size_t c = 0; c < chunks.size(); ++c) {
        int maxp = 0;
        for (int64_t i : chunks[c]) maxp = std::max(maxp, (int)B->shapes[i].size());
        for (int ph = 0; ph < maxp; ++ph) {
            // groups of this phase: generic thread classes (key 0, class) and
            // extended fast shapes (key 1, L, B, conv, sign)
            std::map<std::array<int, 5>, af_batch_s::Group> gm;
            for (int64_t i : chunks[c]) {
                if ((int)B->shapes[i].size() <= ph) continue;
                const AfPass &d = B->shapes[i][ph];
                std::array<int, 5> key;
                if (fast_ext_eligible(B->prec, d)) {
                    key = {1, d.L, d.B, (d.flags & AFP_CONV) ? 1 : 0, d.sign};
                } else {
                    const int cls = d.nthreads <= 64 ? 0 : d.nthreads <= 128 ? 1 : d.nthreads <= 256 ? 2
                                    : d.nthreads <= 512 ? 3 : 4;
                    key = {0, cls, 0, 0, 0};
                }
                af_batch_s::Group &g = gm[key];
                if (g.rows.empty()) {
                    g.phase = ph;
                    g.chunk = (int)c;
                    if (key[0] == 1) {
                        g.ext = true;
                        g.L = d.L;
                        g.B = d.B;
                        g.conv = key[3] != 0;
                        g.sign = d.sign;
                    } else {
                        g.nthreads = classes[key[1]];
                    }
                }
                g.rows.push_back(i);
                g.lds = std::max(g.lds, d.lds_bytes);
                if (d.flags & kRarePassFlags) g.full = true;
            }
            for (auto &kv : gm) {
                af_batch_s::Group g = kv.second;
                g.desc_off = ndesc;
                g.prefix_off = nprefix;
                int64_t t = 0;
                for (int64_t i : g.rows) {
                    B->h_prefix.push_back(t);
                    t += B->shapes[i][ph].ntiles;
                }
                B->h_prefix.push_back(t);
                g.total_tiles = t;
                ndesc += (int64_t)g.rows.size();
                nprefix += (int64_t)g.rows.size() + 1;
                B->groups.push_back(g);
            }
        }
    }
    B->h_descs.resize(ndesc);
    for (DevBuf *b : {&B->descs_buf, &B->prefix_buf})
        if (b->p) {
            (void)hipFree(b->p);
            b->p = nullptr;
        }
    HIPCHK(hipMalloc(&B->descs_buf.p, std::max<int64_t>(ndesc, 1) * sizeof(AfPass)));
    HIPCHK(hipMalloc(&B->prefix_buf.p, std::max<int64_t>(nprefix, 1) * 8));
    B->d_descs = (AfPass *)B->descs_buf.p;
    B->d_prefix = (int64_t *)B->prefix_buf.p;
    if (nprefix) HIPCHK(hipMemcpy(B->d_prefix, B->h_prefix.data(), nprefix * 8, hipMemcpyHostToDevice));
    // capture the grouped launches once; with two lanes the light chunk's
    // launches go to a second stream forked from and joined back to the first,
    // with three lanes the mixed chunk's to a third (the capture turns the
    // forks and joins into graph dependencies)
    if (!B->groups.empty()) {
        const bool two = lanes >= 2 && B->light_chunk && B->nchunks > 1;
        const bool third = B->mixed_chunk >= 0 && B->nchunks > 1;
        hipStream_t s = nullptr, s2 = nullptr, s3 = nullptr;
        hipEvent_t fork = nullptr, join = nullptr, join3 = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        af_status st = AF_SUCCESS;
        if (two || third) st = HIPST(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        if (two) {
            if (st == AF_SUCCESS) st = HIPST(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
            if (st == AF_SUCCESS) st = HIPST(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        }
        if (third) {
            if (st == AF_SUCCESS) st = HIPST(hipStreamCreateWithFlags(&s3, hipStreamNonBlocking));
            if (st == AF_SUCCESS) st = HIPST(hipEventCreateWithFlags(&join3, hipEventDisableTiming));
        }
        if (st == AF_SUCCESS) st = HIPST(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const bool capturing = st == AF_SUCCESS;
        if ((two || third) && st == AF_SUCCESS) st = HIPST(hipEventRecord(fork, s));
        if (two && st == AF_SUCCESS) st = HIPST(hipStreamWaitEvent(s2, fork, 0));
        if (third && st == AF_SUCCESS) st = HIPST(hipStreamWaitEvent(s3, fork, 0));
        for (const auto &g : B->groups) {
            if (st != AF_SUCCESS) break;
            hipStream_t gs = two && g.chunk == 0 ? s2 : third && g.chunk == B->mixed_chunk ? s3 : s;
            if (g.ext)
                st = HIPST(launch_fast_ext_grouped(B->prec, g.L, g.B, g.conv, g.sign, B->d_descs + g.desc_off,
                                                   B->d_prefix + g.prefix_off, (int)g.rows.size(), g.total_tiles, gs));
            else
                st = HIPST(launch_pass_grouped(B->prec, B->d_descs + g.desc_off, B->d_prefix + g.prefix_off,
                                               (int)g.rows.size(), g.total_tiles, g.nthreads, g.lds, g.full, gs));
        }
        if (two && st == AF_SUCCESS) st = HIPST(hipEventRecord(join, s2));
        if (two && st == AF_SUCCESS) st = HIPST(hipStreamWaitEvent(s, join, 0));
        if (third && st == AF_SUCCESS) st = HIPST(hipEventRecord(join3, s3));
        if (third && st == AF_SUCCESS) st = HIPST(hipStreamWaitEvent(s, join3, 0));
        hipGraph_t graph = nullptr;
        if (capturing) {
            const af_status h__ = HIPST(hipStreamEndCapture(s, &graph));
            if (st == AF_SUCCESS) st = h__;
        }
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (join3) (void)hipEventDestroy(join3);
        if (s2) (void)hipStreamDestroy(s2);
        if (s3) (void)hipStreamDestroy(s3);
        (void)hipStreamDestroy(s);
        if (st != AF_SUCCESS) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        B->graph = graph;
        HIPCHK(hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0));
    }
    // description: algorithm counts, chunks, one line per grouped launch
    B->text = B->summary + " launches=" + std::to_string(B->groups.size()) + " chunks=" +
              std::to_string(B->nchunks) + " lanes=" + std::to_string(B->lanes) +
              " chunk_bytes=" + std::to_string((long long)chunk_bytes) +
              " scratch_bytes=" + std::to_string((unsigned long long)region);
    for (size_t k = 0; k < B->groups.size(); ++k) {
        const auto &g = B->groups[k];
        char line[192];
        if (g.ext)
            std::snprintf(line, sizeof line, "\nlaunch %zu chunk %d phase %d ext L=%d B=%d conv=%d sign=%d rows=%zu first=%lld",
                          k, g.chunk, g.phase, g.L, g.B, g.conv ? 1 : 0, g.sign, g.rows.size(), (long long)g.rows[0]);
        else
            std::snprintf(line, sizeof line, "\nlaunch %zu chunk %d phase %d generic threads=%d full=%d rows=%zu first=%lld",
                          k, g.chunk, g.phase, g.nthreads, g.full ? 1 : 0, g.rows.size(), (long long)g.rows[0]);
        B->text += line;
    }
    return AF_SUCCESS;
}

// The host descriptors of one grouped launch for the given buffers, with the
// group's radix order.
af_status batch_group_descs(af_batch B, const af_batch_s::Group &g, const void *in_re, const void *in_im,
                            void *out_re, void *out_im) {
    std::vector<AfPass> ps;
    int64_t j = g.desc_off;
    for (int64_t i : g.rows) {
        ps.clear();
        AFCHK(batch_passes(B, i, in_re, in_im, out_re, out_im, (char *)B->scratch, ps));
        AfPass d = ps[g.phase];
        if (!g.ext) reorder_radices(d, g.order);  // the stage order changes no tile, thread or LDS count
        // the captured launch fixed this descriptor's tiles and shape
        const int64_t k = j - g.desc_off;
        if (d.ntiles != B->h_prefix[g.prefix_off + k + 1] - B->h_prefix[g.prefix_off + k] ||
            (g.ext && (d.L != g.L || d.B != g.B)) || (!g.ext && (d.nthreads > g.nthreads || d.lds_bytes > g.lds)))
            return AF_INTERNAL;
        B->h_descs[j++] = d;
    }
    return AF_SUCCESS;
}

// The descriptors of every grouped launch for the given buffers, uploaded on s
// (the caller holds B->mu and has waited for the previous launch).
af_status batch_upload(af_batch B, const void *in_re, const void *in_im, void *out_re, void *out_im, hipStream_t s) {
    for (const auto &g : B->groups) AFCHK(batch_group_descs(B, g, in_re, in_im, out_re, out_im));
    if (!B->h_descs.empty())
        HIPCHK(hipMemcpyAsync(B->d_descs, B->h_descs.data(), B->h_descs.size() * sizeof(AfPass), hipMemcpyHostToDevice,
                              s));
    HIPCHK(hipStreamSynchronize(s));
    return AF_SUCCESS;
}

// Default chunk budget: none (one chunk).  On MI355X the 4096-row batch of four
// length classes ran in 10.6 ms unchunked against 11.0 ms with chunks of 128
// or 384 MiB of intermediates (half and one and a half times the 256 MB
// memory-side cache) and 13.3 ms with 32 MiB chunks: the passes did not gain
// from re-reading cache-resident intermediates as much as the extra launches
// and their partly filled last rounds cost.  af_batch_tune measures it again
// for each batch; chunking also bounds the scratch memory (5.7 GB unchunked,
// the chunk budget chunked).  Default lanes: two (the rows without
// intermediates on a branch of their own, whose small launches then overlap
// the large ones): 10.32 against 10.61 ms on that batch, in two runs.
constexpr int64_t kBatchChunkBytes = 0;
constexpr int kBatchLanes = 2;

}  // namespace

extern "C" {

af_status af_batch_create(int64_t count, const int64_t *lengths, int32_t sign, int32_t precision, af_batch *batch) {
    if (!batch) return AF_INVALID_ARG;
    *batch = nullptr;
    if (count < 0 || (count > 0 && !lengths) || (sign != 1 && sign != -1)) return AF_INVALID_ARG;
    if (precision != AF_F64 && precision != AF_F32) return AF_INVALID_ARG;
    for (int64_t i = 0; i < count; ++i)
        if (lengths[i] < 0) return AF_INVALID_ARG;
    AFCHK(device_init());
    std::call_once(g_attr_flag, set_kernel_attributes);
    std::unique_ptr<af_batch_s> B(new af_batch_s);
    B->count = count;
    B->sign = sign;
    B->precision = precision;
    B->prec = precision == AF_F64 ? PREC_F64 : PREC_F32;
    B->lengths.assign(lengths, lengths + count);
    B->offsets.resize(count + 1, 0);
    for (int64_t i = 0; i < count; ++i) B->offsets[i + 1] = B->offsets[i] + lengths[i];
    B->scratch_need.assign(count, 0);
    B->scratch_off.assign(count, 0);
    for (int64_t i = 0; i < count; ++i) {
        const int64_t n = lengths[i];
        if (n == 0) continue;
        if (!B->plans.count(n)) {
            af_plan_desc d{};
            d.length = n;
            d.sign = sign;
            d.kind = AF_C2C;
            d.precision = precision;
            d.algorithm = AF_ALG_AUTO;
            af_plan p = nullptr;
            AFCHK(plan_create_impl(&d, &p, false));
            B->plans[n] = p;
        }
        B->scratch_need[i] = (B->plans[n]->scratch_per_row() + 255) / 256 * 256;
    }
    // pass structure with placeholder pointers
    B->shapes.resize(count);
    const char *fake = (const char *)0x1000;
    for (int64_t i = 0; i < count; ++i)
        AFCHK(batch_passes(B.get(), i, fake, nullptr, (void *)fake, nullptr, (char *)0x100000, B->shapes[i]));
    // algorithm counts for the description
    std::map<std::string, int64_t> counts;
    for (int64_t i = 0; i < count; ++i) {
        if (!lengths[i]) continue;
        const int a = B->plans[lengths[i]]->shape.algo;
        const int k = B->plans[lengths[i]]->core.sh.kind;
        std::string key = a == ALG_RADER ? "rader" : a == ALG_BLUESTEIN ? "bluestein" : a == ALG_COPY ? "copy"
                          : a == ALG_DECOMP ? "decompose" : "stockham";
        key += k == CORE_FOUR ? "/fourstep" : k == CORE_THREE ? "/threepass" : "/single";
        counts[key]++;
    }
    B->summary = "batch" + std::to_string(count);
    for (auto &kv : counts) B->summary += " " + kv.first + "=" + std::to_string(kv.second);
    AFCHK(batch_build(B.get(), kBatchChunkBytes, kBatchLanes));
    *batch = B.release();
    return AF_SUCCESS;
}

af_status af_batch_set_schedule(af_batch B, int64_t chunk_bytes, int32_t lanes) {
    if (!B || chunk_bytes < 0 || lanes < 1 || lanes > 3 || (lanes == 3 && chunk_bytes != 0)) return AF_INVALID_ARG;
    std::lock_guard<std::mutex> lk(B->mu);
    if (B->recorded) HIPCHK(hipEventSynchronize(B->done));
    B->recorded = false;
    return batch_build(B, chunk_bytes, lanes);
}

af_status af_batch_tune(af_batch B, void *stream, char *report, size_t report_len) {
    if (!B) return AF_INVALID_ARG;
    if (report && report_len) report[0] = 0;
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(B->mu);
    if (B->recorded) HIPCHK(hipEventSynchronize(B->done));
    const int64_t total = B->offsets.back();
    if (total == 0) return AF_SUCCESS;
    // zero-filled interleaved buffers: the timing does not depend on the values
    const size_t bytes = (size_t)total * (B->precision == AF_F64 ? 16 : 8);
    DevBuf in, out;
    HIPCHK(hipMalloc(&in.p, bytes));
    HIPCHK(hipMalloc(&out.p, bytes));
    HIPCHK(hipMemsetAsync(in.p, 0, bytes, s));
    HIPCHK(hipMemsetAsync(out.p, 0, bytes, s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int64_t start = B->chunk_bytes;
    const int start_lanes = B->lanes;
    // the batch's current schedule first: it is the incumbent the others must beat
    std::vector<std::pair<int64_t, int>> candidates = {{start, start_lanes}};
    for (const std::pair<int64_t, int> c : {std::pair<int64_t, int>{0, 1},
                                            {0, 2},
                                            {0, 3},
                                            {(int64_t)32 << 20, 2},
                                            {(int64_t)64 << 20, 2},
                                            {(int64_t)128 << 20, 2},
                                            {(int64_t)192 << 20, 2},
                                            {(int64_t)384 << 20, 2}})
        if (c != candidates[0]) candidates.push_back(c);
    std::string text;
    int64_t best = start;
    int best_lanes = start_lanes;
    double tbest = 1e30;
    af_status st = AF_SUCCESS;
    // each candidate: one warm-up, then the best of three launches; the
    // incumbent must be beaten by 3%
    for (const auto &cand : candidates) {
        const int64_t cb = cand.first;
        st = batch_build(B, cb, cand.second);
        if (st == AF_SUCCESS) st = batch_upload(B, in.p, nullptr, out.p, nullptr, s);
        double t = 1e30;
        for (int r = 0; r < 4 && st == AF_SUCCESS; ++r) {
            st = HIPST(hipEventRecord(e0, s));
            if (st == AF_SUCCESS) st = HIPST(hipGraphLaunch(B->exec, s));
            if (st == AF_SUCCESS) st = HIPST(hipEventRecord(e1, s));
            if (st == AF_SUCCESS) st = HIPST(hipEventSynchronize(e1));
            float ms = 0.f;
            if (st == AF_SUCCESS) st = HIPST(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0) t = std::min(t, (double)ms);
        }
        if (st != AF_SUCCESS) break;
        char line[160];
        std::snprintf(line, sizeof line, "chunk_bytes=%lld lanes=%d chunks=%d launches=%zu %.3f ms\n", (long long)cb,
                      cand.second, B->nchunks, B->groups.size(), t);
        text += line;
        if (t < tbest * 0.97) {
            tbest = t;
            best = cb;
            best_lanes = cand.second;
        }
    }
    { const af_status h__ = HIPST(hipStreamSynchronize(s)); if (st == AF_SUCCESS) st = h__; }
    // the chosen budget (or, after a failure, the one the batch had)
    const af_status st2 =
        st == AF_SUCCESS ? batch_build(B, best, best_lanes) : batch_build(B, start, start_lanes);
    B->recorded = false;
    if (st == AF_SUCCESS) st = st2;
    if (st == AF_SUCCESS)
        text += "chosen chunk_bytes=" + std::to_string((long long)best) + " lanes=" + std::to_string(best_lanes) + "\n";
    // then the radix order of every generic grouped launch, each timed alone:
    // as planned, reversed, odd radices first (the incumbent must be beaten by 3%)
    if (st == AF_SUCCESS) st = batch_upload(B, in.p, nullptr, out.p, nullptr, s);
    for (size_t k = 0; k < B->groups.size() && st == AF_SUCCESS; ++k) {
        af_batch_s::Group &g = B->groups[k];
        if (g.ext) continue;
        int maxst = 0;
        for (int64_t j = g.desc_off; j < g.desc_off + (int64_t)g.rows.size(); ++j)
            maxst = std::max(maxst, B->h_descs[j].nst);
        if (maxst < 2) continue;
        double tb = 1e30;
        int bo = 0;
        for (int o = 0; o < 3 && st == AF_SUCCESS; ++o) {
            g.order = o;
            st = batch_group_descs(B, g, in.p, nullptr, out.p, nullptr);
            if (st == AF_SUCCESS)
                st = HIPST(hipMemcpyAsync(B->d_descs + g.desc_off, B->h_descs.data() + g.desc_off,
                                          g.rows.size() * sizeof(AfPass), hipMemcpyHostToDevice, s));
            double t = 1e30;
            for (int r = 0; r < 6 && st == AF_SUCCESS; ++r) {
                st = HIPST(hipEventRecord(e0, s));
                if (st == AF_SUCCESS)
                    st = HIPST(launch_pass_grouped(B->prec, B->d_descs + g.desc_off, B->d_prefix + g.prefix_off,
                                                   (int)g.rows.size(), g.total_tiles, g.nthreads, g.lds, g.full, s));
                if (st == AF_SUCCESS) st = HIPST(hipEventRecord(e1, s));
                if (st == AF_SUCCESS) st = HIPST(hipEventSynchronize(e1));
                float ms = 0.f;
                if (st == AF_SUCCESS) st = HIPST(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) t = std::min(t, (double)ms);
            }
            if (st == AF_SUCCESS && (o == 0 || t < tb * 0.97)) {
                tb = t;
                bo = o;
            }
        }
        g.order = bo;
        char line[128];
        std::snprintf(line, sizeof line, "launch %zu order=%d %.1f us\n", k, bo, tb * 1e3);
        text += line;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    { const af_status h__ = HIPST(hipStreamSynchronize(s)); if (st == AF_SUCCESS) st = h__; }
    B->uploaded = false;  // the next call uploads its buffers' descriptors, with the chosen orders
    if (st == AF_SUCCESS && report && report_len) std::snprintf(report, report_len, "%s", text.c_str());
    return st;
}

af_status af_batch_destroy(af_batch batch) {
    if (!batch) return AF_INVALID_ARG;
    delete batch;
    return AF_SUCCESS;
}

af_status af_batch_describe(af_batch batch, char *buf, size_t buflen) {
    if (!batch || !buf || !buflen) return AF_INVALID_ARG;
    std::snprintf(buf, buflen, "%s", batch->text.c_str());
    return AF_SUCCESS;
}

af_status af_batch_execute(af_batch B, const void *in_re, const void *in_im, void *out_re, void *out_im,
                           void *stream) {
    if (!B) return AF_INVALID_ARG;
    if (B->offsets.back() > 0 && (!in_re || !out_re)) return AF_INVALID_ARG;
    if ((in_im == nullptr) != (out_im == nullptr)) return AF_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    const void *cur[4] = {in_re, in_im, out_re, out_im};
    std::lock_guard<std::mutex> lk(B->mu);
    if (!B->done) HIPCHK(hipEventCreateWithFlags(&B->done, hipEventDisableTiming));
    if (!B->uploaded || std::memcmp(cur, B->last, sizeof cur) != 0) {
        // the previous launch may still read the descriptors about to change
        if (B->recorded) HIPCHK(hipEventSynchronize(B->done));
        AFCHK(batch_upload(B, in_re, in_im, out_re, out_im, s));
        std::memcpy(B->last, cur, sizeof cur);
        B->uploaded = true;
    }
    if (B->groups.empty()) return AF_SUCCESS;
    if (B->recorded) HIPCHK(hipStreamWaitEvent(s, B->done, 0));
    HIPCHK(hipGraphLaunch(B->exec, s));
    HIPCHK(hipEventRecord(B->done, s));
    B->recorded = true;
    return AF_SUCCESS;
}

}  // extern "C"
