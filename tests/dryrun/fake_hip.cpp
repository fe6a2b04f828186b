// Host-only stand-in for the HIP runtime, for checking the library's launches
// on a machine without a GPU.
//
// The library's sources are compiled host-only (hipcc --cuda-host-only) and
// linked against this file instead of the HIP runtime.  Device memory is host
// memory with a registry of live allocations; kernels do not run.  Every
// launch is checked instead: the memory footprint that the kernel's indexing
// implies (from its pass descriptor or scalar arguments) must lie inside live
// allocations, and the launch shape must match what the kernel assumes.
// Graph capture records the launches and graph launches check them again, so
// the descriptors that the batch uploads after capture are checked too.
//
// dry_problem_count() returns the number of problems found, dry_problem(i) their text.
#include <hip/hip_runtime_api.h>

#include <cxxabi.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include <numeric>

#include "af_pass.hpp"

namespace {

// Registration runs from other objects' static constructors, so every global
// here is constructed on first use.
struct Alloc {
    size_t size;
};
struct State {
    std::mutex mu;
    std::map<uintptr_t, Alloc> allocs;
    std::vector<std::string> problems;
    std::map<const void *, std::string> kernels;  // host stub -> demangled device name
    long launches = 0, checked = 0;
    std::string context;
    std::map<std::string, long> per_kernel;  // launches by kernel name
    std::string summary;
};
State &st() {
    static State *s = new State;
    return *s;
}
#define g_mu (st().mu)
#define g_allocs (st().allocs)
#define g_problems (st().problems)
#define g_kernels (st().kernels)
#define g_launches (st().launches)
#define g_checked (st().checked)
#define g_context (st().context)

struct Config {
    dim3 grid, block;
    size_t shmem;
    hipStream_t stream;
};
thread_local std::vector<Config> g_cfg;

struct FakeStream {
    bool capturing = false;
    std::vector<std::function<void()>> captured;
    // stream capture across streams: a stream that waits on an event recorded
    // on a capturing stream joins that capture (its launches go to the same
    // graph) until the capture ends; `pending` marks work not yet joined back
    // to the origin by an event the origin waits on
    struct FakeStream *joined = nullptr;
    bool pending = false;
    std::vector<struct FakeStream *> forks;
};
struct FakeEvent {
    FakeStream *root = nullptr;  // capture the recording stream belonged to
    FakeStream *from = nullptr;
};
FakeStream *capture_root(FakeStream *s) { return !s ? nullptr : s->capturing ? s : s->joined; }
struct FakeGraph {
    std::vector<std::function<void()>> nodes;
};

void problem(const std::string &s) {
    if (g_problems.size() < 2000) g_problems.push_back(g_context.empty() ? s : g_context + ": " + s);
}

// [lo, hi) bytes from p inside one live allocation
bool in_alloc(const void *p, int64_t lo, int64_t hi, const char *what, const std::string &k) {
    if (hi <= lo) return true;
    if (!p) {
        problem(k + ": " + what + " is null but accessed");
        return false;
    }
    const uintptr_t a = (uintptr_t)p;
    auto it = g_allocs.upper_bound(a);
    if (it == g_allocs.begin()) {
        problem(k + ": " + what + " points outside every allocation");
        return false;
    }
    --it;
    const uintptr_t base = it->first, end = base + it->second.size;
    if (a >= end) {
        problem(k + ": " + what + " points outside every allocation");
        return false;
    }
    const int64_t off = (int64_t)(a - base);
    if (off + lo < 0 || off + hi > (int64_t)it->second.size) {
        char buf[256];
        std::snprintf(buf, sizeof buf, ": %s accesses bytes [%lld, %lld) of an allocation of %zu (pointer at +%lld)",
                      what, (long long)(off + lo), (long long)(off + hi), it->second.size, (long long)off);
        problem(k + buf);
        return false;
    }
    return true;
}

int pidx_h(int i, int ps) { return ps ? i + (i >> ps) : i; }

// element ranges of complex data: planar (im != null, T-sized parts) or interleaved
void check_c(const void *re, const void *im, int64_t lo, int64_t hi, int tsz, const char *what, const std::string &k) {
    if (hi <= lo) return;
    if (im) {
        in_alloc(re, lo * tsz, hi * tsz, what, k);
        in_alloc(im, lo * tsz, hi * tsz, what, k);
    } else {
        in_alloc(re, lo * 2 * tsz, hi * 2 * tsz, what, k);
    }
}

// max over positions pos = l*ls + e*es (l < nl, e < L, pos < len) of f(pos), and the max position
template <class F>
int64_t scan_positions(const AfPass &d, int64_t ls, int64_t es, int64_t len, F f) {
    int64_t maxpos = -1;
    if ((double)d.nlines * d.L <= 4e7) {
        for (int64_t l = 0; l < d.nlines; ++l)
            for (int64_t e = 0; e < d.L; ++e) {
                const int64_t pos = l * ls + e * es;
                if (pos < len) {
                    if (pos > maxpos) maxpos = pos;
                    f(pos);
                }
            }
    } else {
        maxpos = std::min(len - 1, (d.nlines - 1) * ls + (int64_t)(d.L - 1) * es);
        for (int64_t pos = 0; pos <= maxpos; ++pos) f(pos);
    }
    return maxpos;
}

// Footprint of one pass descriptor, as pass_tile (generic) or k_fast index it.
void check_pass(const AfPass &d, int tsz, const std::string &k) {
    const int csz = 2 * tsz;
    if (d.L < 1 || d.B < 1 || d.nrows < 0 || d.nlines < 0) {
        problem(k + ": bad geometry");
        return;
    }
    if (d.nrows * d.nlines == 0) return;
    if (d.ntiles * d.B < d.nrows * d.nlines) problem(k + ": tiles do not cover the lines");
    // positions inside a row are computed in 32 bits; four-step indices l * e without reduction
    if ((d.nlines - 1) * d.in_ls + (int64_t)(d.L - 1) * d.in_es >= ((int64_t)1 << 31) ||
        (d.nlines - 1) * d.out_ls + (int64_t)(d.L - 1) * d.out_es >= ((int64_t)1 << 31))
        problem(k + ": row positions above 31 bits");
    if ((d.flags & AFP_POSTTW) && (d.nlines - 1) * (int64_t)(d.L - 1) >= d.ftN)
        problem(k + ": four-step index l * e reaches N");
    if (d.B * (int64_t)d.L >= ((int64_t)1 << 22)) problem(k + ": tile index above the small division range");
    const int64_t R1 = d.nrows - 1;
    // stage tables: tw[j*k*tws] < L entries of 16 bytes
    if (d.nst > 0 || d.L > 1) in_alloc(d.tw, 0, (int64_t)d.L * 16, "tw", k);
    int64_t prod = 1;
    for (int i = 0; i < d.nst; ++i) {
        if (d.pst[i] != prod) problem(k + ": stage products inconsistent");
        prod *= d.radix[i];
    }
    if (prod != d.L) problem(k + ": radices do not multiply to L");
    // ---- load ----
    if (d.flags & (AFP_REAL_ENTANGLE | AFP_UNTANGLE_ADJ)) {
        check_c(d.in_re, d.in_im, 0, R1 * d.in_bs + 2 * (int64_t)d.L, tsz, "in (entangle)", k);
        in_alloc(d.aux_tw, 0, (int64_t)d.L * 16, "aux_tw", k);
    } else {
        int64_t amax = -1, amin = INT64_MAX;
        const int64_t maxpos = scan_positions(d, d.in_ls, d.in_es, d.in_len, [&](int64_t pos) {
            int64_t a = pos;
            if (d.in_map) {
                if (!in_alloc(d.in_map, pos * 4, pos * 4 + 4, "in_map", k)) return;
                a = d.in_map[pos];
            }
            if (a > amax) amax = a;
            if (a < amin) amin = a;
        });
        if (maxpos >= 0) {
            if (d.in_map) in_alloc(d.in_map, 0, (maxpos + 1) * 4, "in_map", k);
            if (amin < 0) problem(k + ": negative input index");
            check_c(d.in_re, d.in_im, std::min<int64_t>(amin, 0), R1 * d.in_bs + amax + 1, tsz, "in", k);
            if (d.in_mul) in_alloc(d.in_mul, 0, (maxpos + 1) * csz, "in_mul", k);  // plain complex values
        }
    }
    if (d.flags & (AFP_RADER | AFP_X0_CAPTURE)) check_c(d.in_re, d.in_im, 0, R1 * d.in_bs + 1, tsz, "in x0", k);
    if (d.flags & AFP_X0_CAPTURE) in_alloc(d.x0_out, 0, d.nrows * csz, "x0_out", k);
    // ---- convolution ----
    if (d.flags & AFP_CONV) {
        const int64_t hi = R1 * d.spec_bs + (d.nlines - 1) * d.spec_ls + d.L;
        in_alloc(d.spec, 0, hi * ((d.flags & AFP_SPEC_PLAIN) ? csz : 16), "spec", k);
        if (d.flags & AFP_DC_CAPTURE) in_alloc(d.dc_out, 0, d.nrows * csz, "dc_out", k);
    }
    // ---- store ----
    if (d.flags & (AFP_REAL_UNTANGLE | AFP_ENTANGLE_ADJ)) {
        check_c(d.out_re, d.out_im, 0, R1 * d.out_bs + 2 * (int64_t)d.L, tsz, "out (untangle)", k);
        in_alloc(d.aux_tw, 0, (int64_t)d.L * 16, "aux_tw", k);
        return;
    }
    int64_t amax = -1, amin = INT64_MAX;
    const int64_t maxo = scan_positions(d, d.out_ls, d.out_es, d.out_len, [&](int64_t o) {
        int64_t a = o;
        if (d.out_map) {
            if (!in_alloc(d.out_map, o * 4, o * 4 + 4, "out_map", k)) return;
            a = d.out_map[o];
        }
        if (a > amax) amax = a;
        if (a < amin) amin = a;
    });
    if (maxo >= 0) {
        if (amin < 0) problem(k + ": negative output index");
        check_c(d.out_re, d.out_im, std::min<int64_t>(amin, 0), R1 * d.out_bs + amax + 1, tsz, "out", k);
        if (d.out_mul) in_alloc(d.out_mul, 0, (maxo + 1) * csz, "out_mul", k);
    }
    if (d.flags & AFP_POSTTW) {
        if (d.ftN < 1) problem(k + ": POSTTW without ftN");
        const int64_t T = (int64_t)1 << d.ftShift;
        in_alloc(d.ft_coarse, 0, (((d.ftN - 1) >> d.ftShift) + 1) * 16, "ft_coarse", k);
        in_alloc(d.ft_fine, 0, std::min(T, d.ftN) * 16, "ft_fine", k);
    }
    if (d.flags & AFP_RADER_OUT) {
        check_c(d.x0_re, d.x0_im, 0, R1 * d.x0_bs + 1, tsz, "x0", k);
        in_alloc(d.dc_in, 0, d.nrows * csz, "dc_in", k);
    }
    if (d.flags & (AFP_RADER | AFP_RADER_OUT)) check_c(d.out_re, d.out_im, 0, R1 * d.out_bs + 1, tsz, "out X0", k);
}

void check_generic_launch(const AfPass &d, int tsz, unsigned block, size_t shmem, const std::string &k0) {
    const int csz = 2 * tsz;
    const std::string k = k0 + " [L=" + std::to_string(d.L) + " B=" + std::to_string(d.B) + " lds=" +
                          std::to_string(shmem) + " threads=" + std::to_string(block) + "]";
    if (block > 1024 || block < 1) problem(k + ": block size out of range");
    if ((int64_t)block * pass_ept(tsz == 8) < (int64_t)d.B * d.L)
        problem(k + ": fewer than B*L/EPT threads (butterflies dropped)");
    if (d.B > 64) problem(k + ": B above the 64 lines of TileLines");
    if (d.ld < d.L + (d.B > 1 ? 1 : 0)) problem(k + ": ld too small");
    const int64_t need = ((int64_t)pidx_h(d.B * d.ld - 1, d.padshift) + 1) * csz;
    if ((int64_t)shmem < need) problem(k + ": dynamic LDS below what the tile indexes");
    const int64_t stat = 64 * (8 + 8 + 4) + 2 * 64 * csz;
    if ((int64_t)shmem + stat > 160 * 1024) problem(k + ": LDS above 160 KiB with the static arrays");
    if (d.gt_in || d.gt_out || d.grp_mask) {
        // prime-factor order: both maps, a group mask within the stages, plain passes only;
        // the maps are permutations of the line (device memory is host memory here)
        if (!d.gt_in || !d.gt_out || !d.grp_mask) problem(k + ": incomplete prime-factor descriptor");
        if (d.grp_mask != (1 << d.nst) - 2) problem(k + ": prime-factor group mask differs from one stage per factor");
        for (int a = 0; a < d.nst; ++a)
            for (int b2 = a + 1; b2 < d.nst; ++b2)
                if (std::gcd(d.radix[a], d.radix[b2]) != 1) problem(k + ": prime-factor radices not pairwise coprime");
        if (d.flags & (AFP_CONV | AFP_POSTTW | kRarePassFlags)) problem(k + ": prime-factor order on a pass with packing, twiddles or a product");
        if (in_alloc(d.gt_in, 0, (int64_t)d.L * 4, "gt_in", k) && in_alloc(d.gt_out, 0, (int64_t)d.L * 4, "gt_out", k)) {
            std::vector<char> seen_in(d.L, 0), seen_out(d.L, 0);
            for (int e = 0; e < d.L; ++e) {
                const int32_t a = d.gt_in[e], b = d.gt_out[e];
                if (a < 0 || a >= d.L || seen_in[a]++) { problem(k + ": gt_in is not a permutation"); break; }
                if (b < 0 || b >= d.L || seen_out[b]++) { problem(k + ": gt_out is not a permutation"); break; }
            }
        }
    }
    check_pass(d, tsz, k);
}

int fast_ld_h(int L, int B) { return B > 1 ? L + (L >> 4) + 1 : L + (L >> 4); }

void check_fast_launch(const AfPass &d, int tsz, int L, int B, int sg, dim3 grid, dim3 block, size_t shmem,
                       const std::string &k) {
    if (d.L != L || d.B != B) problem(k + ": descriptor shape differs from the instantiation");
    if (d.sign != sg) problem(k + ": descriptor sign differs from the instantiation");
    if ((int64_t)grid.x != d.ntiles) problem(k + ": grid differs from ntiles");
    if ((int64_t)block.x != (int64_t)B * L / 16) problem(k + ": block differs from B*L/16");
    const int64_t need = (int64_t)fast_ld_h(L, B) * B * 2 * tsz + (L <= 1024 ? L : 256) * 16;
    if ((int64_t)shmem != need) problem(k + ": LDS size differs from the kernel's layout");
    if (shmem > 160 * 1024) problem(k + ": LDS above 160 KiB");
    if (d.flags & ~(AFP_POSTTW | AFP_SCALE)) problem(k + ": flags the kernel ignores");
    if (d.in_map || d.out_map || d.in_mul || d.out_mul) problem(k + ": maps or vectors the kernel ignores");
    if (d.in_len != INT64_MAX || d.out_len != INT64_MAX) problem(k + ": lengths the kernel ignores");
    if (d.nlines % B) problem(k + ": B does not divide the lines per row");
    const int64_t eb = tsz;
    const int64_t span_i = ((int64_t)(B - 1) * d.in_ls + (int64_t)(L - 1) * d.in_es) * (d.in_im ? eb : 2 * eb) + eb;
    const int64_t span_o = ((int64_t)(B - 1) * d.out_ls + (int64_t)(L - 1) * d.out_es) * (d.out_im ? eb : 2 * eb) + eb;
    if (span_i >= ((int64_t)1 << 32) || span_o >= ((int64_t)1 << 32)) problem(k + ": tile span above 32 bits");
    in_alloc(d.tw_d, 0, (int64_t)L * 16, "tw_d", k);
    if (d.flags & AFP_POSTTW) {
        const int64_t T = (int64_t)1 << d.ftShift;
        in_alloc(d.ft_coarse_d, 0, (((d.ftN - 1) >> d.ftShift) + 1) * 16, "ft_coarse_d", k);
        in_alloc(d.ft_fine_d, 0, std::min(T, d.ftN) * 16, "ft_fine_d", k);
        if (d.ftN & (d.ftN - 1)) problem(k + ": four-step N not a power of two");
    }
    check_pass(d, tsz, k);
}

// k_fast_ext<T, L, B, SG, CONV> and the tiles of one descriptor of k_fast_ext_grouped
void check_ext_desc(const AfPass &d, int tsz, int L, int B, int sg, bool conv, dim3 block, size_t shmem,
                    const std::string &k) {
    if (d.L != L || d.B != B) problem(k + ": descriptor shape differs from the instantiation");
    if (d.sign != sg) problem(k + ": descriptor sign differs from the instantiation");
    if (((d.flags & AFP_CONV) != 0) != conv) problem(k + ": CONV differs from the instantiation");
    if (conv && d.in_lines_fast) problem(k + ": convolution lines must be rows");
    if ((int64_t)block.x != (int64_t)B * L / 16) problem(k + ": block differs from B*L/16");
    const int64_t need = (int64_t)fast_ld_h(L, B) * B * 2 * tsz + (L <= 1024 ? L : 256) * 16;
    if ((int64_t)shmem != need) problem(k + ": LDS size differs from the kernel's layout");
    if (shmem > 160 * 1024) problem(k + ": LDS above 160 KiB");
    if (d.flags & ~(AFP_POSTTW | AFP_SCALE | AFP_CONV | AFP_SPEC_PLAIN)) problem(k + ": flags the kernel ignores");
    if (d.in_map || d.out_map) problem(k + ": maps the kernel ignores");
    if (d.nlines % B) problem(k + ": B does not divide the lines per row");
    const int64_t eb = tsz;
    const int64_t span_i = ((int64_t)(B - 1) * d.in_ls + (int64_t)(L - 1) * d.in_es) * (d.in_im ? eb : 2 * eb) + eb;
    const int64_t span_o = ((int64_t)(B - 1) * d.out_ls + (int64_t)(L - 1) * d.out_es) * (d.out_im ? eb : 2 * eb) + eb;
    if (span_i >= ((int64_t)1 << 32) || span_o >= ((int64_t)1 << 32)) problem(k + ": tile span above 32 bits");
    in_alloc(d.tw_d, 0, (int64_t)L * 16, "tw_d", k);
    in_alloc(d.tw, 0, (int64_t)L * 16, "tw", k);  // the LDS stage table, or W_256 from it
    if (d.flags & AFP_POSTTW) {
        const int64_t T = (int64_t)1 << d.ftShift;
        in_alloc(d.ft_coarse_d, 0, (((d.ftN - 1) >> d.ftShift) + 1) * 16, "ft_coarse_d", k);
        in_alloc(d.ft_fine_d, 0, std::min(T, d.ftN) * 16, "ft_fine_d", k);
        if (d.ftN & (d.ftN - 1)) problem(k + ": four-step N not a power of two");
    }
    check_pass(d, tsz, k);
}

// Optional record of the descriptors of grouped launches (AF_DRYRUN_DUMP=file):
// one line per descriptor "threads lds L B nlines nrows ntiles flags", "--"
// after each launch.
void dump_group(const AfPass *descs, int nd, dim3 block, size_t shmem) {
    const char *f = std::getenv("AF_DRYRUN_DUMP");
    if (!f) return;
    FILE *o = std::fopen(f, "a");
    if (!o) return;
    for (int i = 0; i < nd; ++i)
        std::fprintf(o, "%u %zu %d %d %lld %lld %lld %d\n", block.x, shmem, descs[i].L, descs[i].B,
                     (long long)descs[i].nlines, (long long)descs[i].nrows, (long long)descs[i].ntiles, descs[i].flags);
    std::fprintf(o, "--\n");
    std::fclose(o);
}

int tsize_of(const std::string &n) { return n.find("<double") != std::string::npos ? 8 : 4; }

// template arguments of k_fast<T, L, B, SG>
bool fast_args(const std::string &n, int &L, int &B, int &sg) {
    const size_t a = n.find('<');
    if (a == std::string::npos) return false;
    const char *p = n.c_str() + a + 1;
    while (*p && *p != ',') ++p;
    return std::sscanf(p, ", %d, %d, %d", &L, &B, &sg) == 3;
}

// Checks one launch now; returns a closure that repeats the check (for graphs).
std::function<void()> check_launch(const std::string &name, dim3 grid, dim3 block, size_t shmem, void **args) {
    const int tsz = tsize_of(name);
    if (name.find("af::k_pass<") == 0) {
        const AfPass d = *(const AfPass *)args[0];
        return [=] { check_generic_launch(d, tsz, block.x, shmem, name); };
    }
    if (name.find("af::k_fast<") == 0) {
        const AfPass d = *(const AfPass *)args[0];
        int L = 0, B = 0, sg = 0;
        if (!fast_args(name, L, B, sg)) return [=] { problem(name + ": cannot parse the instantiation"); };
        return [=] { check_fast_launch(d, tsz, L, B, sg, grid, block, shmem, name); };
    }
    if (name.find("af::k_fast_ext<") == 0) {
        const AfPass d = *(const AfPass *)args[0];
        int L = 0, B = 0, sg = 0;
        if (!fast_args(name, L, B, sg)) return [=] { problem(name + ": cannot parse the instantiation"); };
        const bool conv = name.find("true>") != std::string::npos;
        return [=] {
            if ((int64_t)grid.x != d.ntiles) problem(name + ": grid differs from ntiles");
            check_ext_desc(d, tsz, L, B, sg, conv, block, shmem, name);
        };
    }
    if (name.find("af::k_fast_ext_grouped<") == 0) {
        const AfPass *descs = *(const AfPass *const *)args[0];
        const int64_t *prefix = *(const int64_t *const *)args[1];
        const int nd = *(const int *)args[2];
        int L = 0, B = 0, sg = 0;
        if (!fast_args(name, L, B, sg)) return [=] { problem(name + ": cannot parse the instantiation"); };
        const bool conv = name.find("true>") != std::string::npos;
        return [=] {
            if (!in_alloc(descs, 0, (int64_t)nd * sizeof(AfPass), "descs", name)) return;
            if (!in_alloc(prefix, 0, (int64_t)(nd + 1) * 8, "tile_prefix", name)) return;
            if (prefix[0] != 0) problem(name + ": tile prefix does not start at 0");
            if ((int64_t)grid.x != prefix[nd]) problem(name + ": grid differs from the total tiles");
            for (int i = 0; i < nd; ++i) {
                if (prefix[i + 1] - prefix[i] != descs[i].ntiles) problem(name + ": tile prefix differs from ntiles");
                check_ext_desc(descs[i], tsz, L, B, sg, conv, block, shmem, name + " desc " + std::to_string(i));
            }
            dump_group(descs, nd, block, shmem);
        };
    }
    if (name.find("af::k_pass_grouped<") == 0) {
        const AfPass *descs = *(const AfPass *const *)args[0];
        const int64_t *prefix = *(const int64_t *const *)args[1];
        const int nd = *(const int *)args[2];
        return [=] {
            if (!in_alloc(descs, 0, (int64_t)nd * sizeof(AfPass), "descs", name)) return;
            if (!in_alloc(prefix, 0, (int64_t)(nd + 1) * 8, "tile_prefix", name)) return;
            for (int i = 0; i < nd; ++i) {
                if (prefix[i + 1] - prefix[i] != descs[i].ntiles) problem(name + ": tile prefix differs from ntiles");
                check_generic_launch(descs[i], tsz, block.x, shmem, name + " desc " + std::to_string(i));
            }
            dump_group(descs, nd, block, shmem);
        };
    }
    if (name.find("af::k_fast_rconv<") == 0) {
        const AfPass d = *(const AfPass *)args[0];
        const void *in[3] = {*(const void *const *)args[1], *(const void *const *)args[2], *(const void *const *)args[3]};
        void *out[2] = {*(void **)args[4], *(void **)args[5]};
        int L = 0, B = 0;
        const bool vjp = name.find("true>") != std::string::npos;
        const size_t a = name.find('<');
        const char *p = name.c_str() + a + 1;
        while (*p && *p != ',') ++p;
        if (std::sscanf(p, ", %d, %d", &L, &B) != 2) return [=] { problem(name + ": cannot parse the instantiation"); };
        return [=] {
            const int64_t row = 2 * (int64_t)L * tsz;
            if (d.L != L) problem(name + ": descriptor L differs from the instantiation");
            if ((int64_t)grid.x * B != d.nrows) problem(name + ": grid * B differs from the rows");
            if ((int64_t)block.x != (int64_t)B * L / 16) problem(name + ": block differs from B*L/16");
            // FP32: one line buffer (first spectrum in registers); FP64: two
            // + the W_L stage table (L <= 1024) or W_256
            const int64_t need = (tsz == 4 ? 1 : 2) * (int64_t)fast_ld_h(L, B) * B * 2 * tsz + (L <= 1024 ? L : 256) * 16;
            if ((int64_t)shmem != need) problem(name + ": LDS size differs from the kernel's layout");
            for (int i = 0; i < (vjp ? 3 : 2); ++i) in_alloc(in[i], 0, d.nrows * row, "input", name);
            for (int i = 0; i < (vjp ? 2 : 1); ++i) in_alloc(out[i], 0, d.nrows * row, "output", name);
            in_alloc(d.aux_tw, 0, (int64_t)L * 16, "aux_tw", name);
            // W_L (float4 / double2 entries): the LDS stage table, or W_256 from it
            in_alloc(d.tw, 0, (int64_t)L * 16, "tw", name);
            if (L > 1024) in_alloc(d.tw_d, 0, (int64_t)L * 16, "tw_d", name);
        };
    }
    if (name.find("af::k_karatsuba") == 0) {
        // k_karatsuba and k_karatsuba_mfma: the same arguments (the first has one
        // more, skip_mfma) and LDS layouts of their own
        const bool mfma = name.find("af::k_karatsuba_mfma") == 0;
        const void *x = *(const void *const *)args[0];
        const void *y = *(const void *const *)args[1];
        void *z = *(void **)args[2];
        const int64_t *meta = *(const int64_t *const *)args[3];
        const int64_t np = *(const int64_t *)args[4];
        const int cyc = *(const int *)args[5];
        const void *taken = *(const void *const *)args[6];  // one int32 per pair
        return [=] {
            if (!in_alloc(meta, 0, np * 24, "meta", name)) return;
            if (taken && !in_alloc(taken, 0, np * 4, "taken", name)) return;
            if (mfma && !taken) problem(name + ": no hand-over array");
            for (int64_t p = 0; p < np; ++p) {
                const int64_t off = meta[3 * p], L = meta[3 * p + 1], zo = meta[3 * p + 2];
                if (L <= 0) continue;
                in_alloc(x, off * 8, (off + L) * 8, "x", name);
                in_alloc(y, off * 8, (off + L) * 8, "y", name);
                in_alloc(z, zo * 8, (zo + (cyc ? L : 2 * L - 1)) * 8, "z", name);
                // LDS words of the pair: 2P + 2Q + 3^B 2s + 3^(B-1) 4s + 16 * 128
                int d = 0;
                int64_t sl = L;
                while (sl > 64) {
                    ++d;
                    sl = (L + ((int64_t)1 << d) - 1) >> d;
                }
                const int64_t P = sl << d, B = d < 3 ? d : 3;
                int64_t n3 = 1;
                for (int i = 0; i < B; ++i) n3 *= 3;
                const int64_t Qp = ((sl << B) + 15) / 16 * 16;
                const int64_t words = mfma ? 2 * P + 2 * ((Qp + 32) + (Qp / 16 + 32) * 17) + 16 * 16 * 17 + 4 * 16 + 2 * 16
                                           : 2 * P + 2 * (sl << B) + n3 * 2 * sl + (B ? n3 / 3 * 4 * sl : 0) + 16 * 128;
                if (words * 8 > (int64_t)shmem)
                    problem(name + ": a pair needs more LDS than the launch has (L=" + std::to_string(L) + " need " +
                            std::to_string(words * 8) + " have " + std::to_string(shmem) + ")");
                if (block.x != 1024) problem(name + ": block differs from 1024");
            }
        };
    }
    if (name.find("af::k_dc_") == 0) {
        // decomposition plans: gathers through index maps, twiddles, spectrum
        // products and Rader's output, each over `rows` rows
        const int tsz = name.find("<double>") != std::string::npos ? 8 : 4;
        if (name.find("af::k_dc_gather<") == 0) {
            const void *sr = *(const void *const *)args[0], *si = *(const void *const *)args[1];
            const int64_t n = *(const int64_t *)args[2];
            void *dr = *(void **)args[3], *di = *(void **)args[4];
            const int64_t m = *(const int64_t *)args[5];
            const int32_t *map = *(const int32_t *const *)args[6];
            const int64_t rows = *(const int64_t *)args[7];
            return [=] {
                check_c(dr, di, 0, rows * m, tsz, "dst", name);
                check_c(sr, si, 0, rows * n, tsz, "src", name);
                if (!in_alloc(map, 0, m * 4, "map", name)) return;
                for (int64_t i = 0; i < m; ++i)
                    if (map[i] < 0 || map[i] >= n) {
                        problem(name + ": map entry outside the source row");
                        break;
                    }
            };
        }
        if (name.find("af::k_dc_twiddle<") == 0) {
            const void *x = *(const void *const *)args[0];
            const int64_t a = *(const int64_t *)args[1], b = *(const int64_t *)args[2], n = *(const int64_t *)args[3];
            const int64_t rows = *(const int64_t *)args[5];
            return [=] {
                if (a * b != n) problem(name + ": a b differs from n");
                in_alloc(x, 0, rows * n * 2 * tsz, "x", name);
            };
        }
        if (name.find("af::k_dc_spec_mul<") == 0) {
            const void *x = *(const void *const *)args[0], *spec = *(const void *const *)args[1];
            const int64_t L = *(const int64_t *)args[2], rows = *(const int64_t *)args[3];
            return [=] {
                in_alloc(x, 0, rows * L * 2 * tsz, "x", name);
                in_alloc(spec, 0, L * 2 * tsz, "spec", name);
            };
        }
        if (name.find("af::k_dc_rader_out<") == 0) {
            const void *xr = *(const void *const *)args[0], *xi = *(const void *const *)args[1];
            const void *c = *(const void *const *)args[2];
            const int32_t *pw = *(const int32_t *const *)args[3];
            const int64_t p = *(const int64_t *)args[4];
            void *orr = *(void **)args[5], *oi = *(void **)args[6];
            const int64_t rows = grid.x;
            return [=] {
                check_c(xr, xi, 0, rows * p, tsz, "x", name);
                check_c(orr, oi, 0, rows * p, tsz, "out", name);
                in_alloc(c, 0, rows * (p - 1) * 2 * tsz, "c", name);
                if (block.x != 256) problem(name + ": the reduction assumes 256 threads");
                if (!in_alloc(pw, 0, (p - 1) * 4, "pw", name)) return;
                for (int64_t i = 0; i < p - 1; ++i)
                    if (pw[i] < 1 || pw[i] >= p) {
                        problem(name + ": scatter index outside 1 .. p - 1");
                        break;
                    }
            };
        }
        return [=] { problem(name + ": no footprint model"); };
    }
    if (name.find("af::k_kar_split") == 0 || name.find("af::k_kar_join") == 0) {
        // split tasks (x, y, L, cx, cy): x, y hold L words, cx, cy 3 ceil(L/2);
        // join tasks (zc, L, z, cyclic): zc holds 3 (2h - 1), z 2L - 1 (L if cyclic)
        const bool split = name.find("split") != std::string::npos;
        const int64_t *t = *(const int64_t *const *)args[0];
        const int64_t nt = *(const int64_t *)args[1];
        return [=] {
            if (!in_alloc(t, 0, nt * (split ? 40 : 32), "task table", name)) return;
            for (int64_t i = 0; i < nt; ++i) {
                const int64_t *k = t + i * (split ? 5 : 4);
                const int64_t L = k[split ? 2 : 1], h = (L + 1) / 2;
                if (L < 2) problem(name + ": a split needs a length of at least 2");
                if (split) {
                    in_alloc((const void *)k[0], 0, L * 8, "x", name);
                    in_alloc((const void *)k[1], 0, L * 8, "y", name);
                    in_alloc((const void *)k[3], 0, 3 * h * 8, "cx", name);
                    in_alloc((const void *)k[4], 0, 3 * h * 8, "cy", name);
                } else {
                    in_alloc((const void *)k[0], 0, 3 * (2 * h - 1) * 8, "zc", name);
                    in_alloc((const void *)k[2], 0, (k[3] ? L : 2 * L - 1) * 8, "z", name);
                }
            }
            if (block.x != 256) problem(name + ": block differs from 256");
        };
    }
    if (name.find("af::k_gen_twiddles<") == 0) {
        void *out = *(void **)args[0];
        const int64_t count = *(const int64_t *)args[2];
        return [=] { in_alloc(out, 0, count * 16, "out", name); };
    }
    if (name.find("af::k_gen_chirp") == 0) {
        void *out = *(void **)args[0];
        const int64_t count = *(const int64_t *)args[2];
        return [=] { in_alloc(out, 0, count * 16, "out", name); };
    }
    if (name.find("af::k_gen_rader_kernel") == 0) {
        void *out = *(void **)args[0];
        const void *powg = *(const void *const *)args[1];
        const int64_t count = *(const int64_t *)args[3];
        return [=] {
            in_alloc(out, 0, count * 16, "out", name);
            in_alloc(powg, 0, count * 4, "powg", name);
        };
    }
    if (name.find("af::k_two_untangle<") == 0 || name.find("af::k_two_entangle<") == 0) {
        // (p_re, p_im, [q_re, q_im,] o1_re, o1_im, [o2_re, o2_im,] n, rows): every
        // non-null array holds rows * n elements (components or interleaved pairs)
        const bool unt = name.find("untangle") != std::string::npos;
        const void *ptr[6];
        for (int i = 0; i < 6; ++i) ptr[i] = *(const void *const *)args[i];
        const int64_t n = *(const int64_t *)args[6], rows = *(const int64_t *)args[7];
        const int64_t h = unt ? *(const int64_t *)args[8] : 1;
        if (h < 1 || rows % h) return [=] { problem(name + ": rows not a multiple of the matrix height"); };
        return [=] {
            // pairs (re, im): interleaved when im is null
            for (int pr = 0; pr < 3; ++pr) {
                const void *re = ptr[2 * pr], *im = ptr[2 * pr + 1];
                if (!re) {
                    if (im) problem(name + ": imaginary part without a real part");
                    continue;
                }
                const std::string what = unt ? (pr == 0 ? "z" : pr == 1 ? "a" : "b") : (pr == 0 ? "c" : pr == 1 ? "d" : "z");
                check_c(re, im, 0, rows * n, tsz, what.c_str(), name);
            }
            if ((int64_t)grid.x * block.x < n) problem(name + ": grid does not cover the row");
        };
    }
    if (name.find("af::k_transpose<") == 0) {
        const void *ir = *(const void *const *)args[0], *ii = *(const void *const *)args[1];
        const void *orr = *(const void *const *)args[2], *oi = *(const void *const *)args[3];
        const int64_t R = *(const int64_t *)args[4], C = *(const int64_t *)args[5], nb = *(const int64_t *)args[6];
        return [=] {
            check_c(ir, ii, 0, nb * R * C, tsz, "in", name);
            check_c(orr, oi, 0, nb * R * C, tsz, "out", name);
            if ((int64_t)grid.x * 32 < C || (int64_t)grid.y * 32 < R) problem(name + ": grid does not cover the matrix");
            if (block.x != 256) problem(name + ": the kernel assumes 32 x 8 threads");
        };
    }
    if (name.find("af::k_convert<") == 0) {
        void *out = *(void **)args[0];
        const void *in = *(const void *const *)args[1];
        const int64_t count = *(const int64_t *)args[2];
        const int64_t N1 = *(const int64_t *)args[4], N2 = *(const int64_t *)args[5];
        // output element: float2 (plain FP32 spectra) 8 bytes, float4 / double2 16
        const int64_t osz = name.find("HIP_vector_type<float, 2u>") != std::string::npos ? 8 : 16;
        return [=] {
            in_alloc(out, 0, count * osz, "out", name);
            in_alloc(in, 0, count * 16, "in", name);
            if (N1 > 0 && N1 * N2 != count) problem(name + ": N1*N2 differs from count");
        };
    }
    if (name.find("af::k_gauss_split") == 0 || name.find("af::k_gauss_join") == 0) {
        // split (xr, xi, yr, yi, x3, y3, count): four inputs of count words, two
        // outputs of 3 count; join (p3, zr, zi, count): one input of 3 count, two outputs
        const bool split = name.find("split") != std::string::npos;
        const int na = split ? 6 : 3;
        const void *ptr[6];
        for (int i = 0; i < na; ++i) ptr[i] = *(const void *const *)args[i];
        const int64_t count = *(const int64_t *)args[na];
        return [=] {
            for (int i = 0; i < na; ++i) {
                const int64_t words = (split ? i >= 4 : i == 0) ? 3 * count : count;
                in_alloc(ptr[i], 0, words * 8, split ? (i < 4 ? "operand" : "split output") : (i ? "z" : "p3"), name);
            }
            if (block.x != 256) problem(name + ": block differs from 256");
        };
    }
    return [] {};  // kernels without a footprint model (packing)
}

}  // namespace

namespace {
// device variables: their registered size, by host handle
std::map<const void *, size_t> &dev_vars() {
    static auto *m = new std::map<const void *, size_t>;
    return *m;
}
}  // namespace

extern "C" {

// ---- registration ----
void **__hipRegisterFatBinary(const void *) {
    static void *mod = nullptr;
    return &mod;
}
void __hipUnregisterFatBinary(void **) {}
void __hipRegisterFunction(void **, const void *hostFunction, char *deviceFunction, const char *, unsigned int,
                           uint3 *, uint3 *, dim3 *, dim3 *, int *) {
    int rc = 0;
    char *dm = abi::__cxa_demangle(deviceFunction, nullptr, nullptr, &rc);
    std::string n = (rc == 0 && dm) ? dm : deviceFunction;
    if (n.compare(0, 5, "void ") == 0) n = n.substr(5);  // templates demangle with their return type
    g_kernels[hostFunction] = n;
    std::free(dm);
}
void __hipRegisterVar(void **, void *var, char *, const char *, int, size_t size, int, int) {
    dev_vars()[var] = size;
}

hipError_t __hipPushCallConfiguration(dim3 gridDim, dim3 blockDim, size_t sharedMem, hipStream_t stream) {
    g_cfg.push_back({gridDim, blockDim, sharedMem, stream});
    return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3 *gridDim, dim3 *blockDim, size_t *sharedMem, hipStream_t *stream) {
    const Config c = g_cfg.back();
    g_cfg.pop_back();
    *gridDim = c.grid;
    *blockDim = c.block;
    *sharedMem = c.shmem;
    *stream = c.stream;
    return hipSuccess;
}

hipError_t hipLaunchKernel(const void *f, dim3 grid, dim3 block, void **args, size_t shmem, hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_kernels.find(f);
    const std::string name = it == g_kernels.end() ? std::string("unregistered kernel") : it->second;
    ++g_launches;
    ++st().per_kernel[name];
    if (grid.x == 0 || block.x == 0) problem(name + ": empty launch");
    auto fn = check_launch(name, grid, block, shmem, args);
    FakeStream *fs = (FakeStream *)stream;
    if (FakeStream *root = capture_root(fs)) {
        root->captured.push_back(fn);
        if (fs != root) fs->pending = true;
        return hipSuccess;
    }
    ++g_checked;
    fn();
    return hipSuccess;
}

// ---- memory ----
hipError_t hipMalloc(void **p, size_t n) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (n == 0) n = 1;
    void *q = std::malloc(n);
    if (!q) return hipErrorOutOfMemory;
    std::memset(q, 0, n);
    g_allocs[(uintptr_t)q] = {n};
    *p = q;
    return hipSuccess;
}
hipError_t hipMallocAsync(void **p, size_t n, hipStream_t) { return hipMalloc(p, n); }
// a device variable's "device" address: one allocation of its registered size
// per variable (the host handle itself is never a device address)
hipError_t hipGetSymbolAddress(void **p, const void *sym) {
    static std::map<const void *, void *> addr;
    auto it = dev_vars().find(sym);
    if (it == dev_vars().end()) return hipErrorInvalidSymbol;
    void *&a = addr[sym];
    if (!a) {
        const hipError_t e = hipMalloc(&a, it->second);
        if (e != hipSuccess) return e;
    }
    *p = a;
    return hipSuccess;
}
hipError_t hipFree(void *p) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!p) return hipSuccess;
    auto it = g_allocs.find((uintptr_t)p);
    if (it == g_allocs.end()) {
        problem("hipFree of a pointer that is not a live allocation");
        return hipErrorInvalidValue;
    }
    g_allocs.erase(it);
    std::free(p);
    return hipSuccess;
}
hipError_t hipFreeAsync(void *p, hipStream_t) { return hipFree(p); }
hipError_t hipMemsetAsync(void *dst, int v, size_t n, hipStream_t) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!in_alloc(dst, 0, n, "memset dst", "hipMemsetAsync")) return hipErrorInvalidValue;
    std::memset(dst, v, n);
    return hipSuccess;
}
hipError_t hipMemset(void *dst, int v, size_t n) { return hipMemsetAsync(dst, v, n, nullptr); }

// ---- events: every interval reads as 1 ms (the autotuner then keeps the planned shapes) ----
hipError_t hipEventCreate(hipEvent_t *e) {
    *e = (hipEvent_t) new FakeEvent;
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    if (!e) return hipErrorInvalidValue;
    FakeEvent *fe = (FakeEvent *)e;
    fe->root = capture_root((FakeStream *)s);
    fe->from = (FakeStream *)s;
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int) {
    if (!e) return hipErrorInvalidValue;
    FakeEvent *fe = (FakeEvent *)e;
    FakeStream *fs = (FakeStream *)s;
    if (!fe->root || !fs || fe->from == fs) return hipSuccess;
    if (fs == fe->root) {
        // the origin waits on a forked stream: its work so far is joined back
        if (fe->from) fe->from->pending = false;
    } else if (capture_root(fs) != fe->root) {
        fs->joined = fe->root;
        fe->root->forks.push_back(fs);
    }
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
    *ms = 1.0f;
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete (FakeEvent *)e;
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
// pinned host memory: plain host memory here (never a kernel operand)
hipError_t hipHostMalloc(void **p, size_t n, unsigned int) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void *p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind kind) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice) in_alloc(dst, 0, n, "memcpy dst", "hipMemcpy");
    if (kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice) in_alloc(src, 0, n, "memcpy src", "hipMemcpy");
    std::memcpy(dst, src, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind kind, hipStream_t) {
    return hipMemcpy(dst, src, n, kind);
}

// ---- device ----
hipError_t hipGetDeviceCount(int *n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600 *p, int) {
    std::memset(p, 0, sizeof *p);
    std::strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
    std::strcpy(p->name, "dry run");
    p->multiProcessorCount = 256;
    p->sharedMemPerBlock = 160 * 1024;
    p->maxThreadsPerBlock = 1024;
    p->warpSize = 64;
    return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "dry-run HIP error"; }

// ---- streams and graphs ----
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int) {
    *s = (hipStream_t) new FakeStream;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    delete (FakeStream *)s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    if (!s) return hipErrorInvalidValue;
    ((FakeStream *)s)->capturing = true;
    ((FakeStream *)s)->captured.clear();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
    FakeStream *fs = (FakeStream *)s;
    for (FakeStream *f : fs->forks) {
        if (f->pending) problem("stream capture ended with a forked stream's work not joined back");
        f->joined = nullptr;
        f->pending = false;
    }
    fs->forks.clear();
    FakeGraph *fg = new FakeGraph;
    fg->nodes.swap(fs->captured);
    fs->capturing = false;
    *g = (hipGraph_t)fg;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
    FakeGraph *c = new FakeGraph(*(FakeGraph *)g);
    *e = (hipGraphExec_t)c;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &fn : ((FakeGraph *)e)->nodes) {
        ++g_checked;
        fn();
    }
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete (FakeGraph *)g;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete (FakeGraph *)e;
    return hipSuccess;
}

// ---- reporting (read by tests/test_dryrun.py) ----
void dry_context(const char *c) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_context = c ? c : "";
}
long dry_problem_count(void) { return (long)g_problems.size(); }
const char *dry_problem(long i) { return i >= 0 && i < (long)g_problems.size() ? g_problems[i].c_str() : ""; }
long dry_launches(void) { return g_launches; }
long dry_checked(void) { return g_checked; }
long dry_live_allocations(void) { return (long)g_allocs.size(); }
// "name: launches" lines for every kernel launched so far
const char *dry_kernel_summary(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    st().summary.clear();
    for (auto &kv : st().per_kernel) st().summary += kv.first + ": " + std::to_string(kv.second) + "\n";
    return st().summary.c_str();
}

}  // extern "C"
