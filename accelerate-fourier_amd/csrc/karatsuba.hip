// Bit-exact Int64 Karatsuba convolution on gfx950.
//
// The reference (Convolution/Adhoc.hs:36-98) splits each operand in halves
// level after level down to single coefficients (3^d leaf products), then
// recombines with z = zl + u (zm - zl - zr) + u^2 zr.  Over the ring Z/2^64 every
// exact splitting gives the same bits, so here one workgroup handles one pair
// and organises the same algebra around LDS:
//
//  * operands are zero-padded to P = 2^d s (leaf size s <= 64); the d levels
//    are split into T = d - B top levels and B = min(d, 3) bottom levels;
//  * the top levels are walked one subtree at a time: subtree t in {0,1,2}^T
//    takes the sum of the operand blocks of size Q = 2^B s that t selects
//    (t_i = 0: low half, 2: high half, 1: both -- the Karatsuba evaluation);
//  * inside a subtree, one wavefront forms one of its 3^B leaf products (two
//    s-long blocks, schoolbook, each lane owning coefficients lane and
//    lane + s), and the bottom B levels are recombined in LDS, level by level;
//  * the subtree product is added to the LDS result with the shifts and signs
//    of prod_i c_{t_i}(u_i), c_0 = 1 - u, c_1 = u, c_2 = u^2 - u, over the top
//    levels -- one term at a time, so no atomics are needed.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace af {

typedef unsigned long long u64;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Whole-wavefront rotation by one lane (DPP wave_ror:1): lane l receives the
// value of lane l - 1 (mod 64).
__device__ __forceinline__ u64 rot1(u64 v) {
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)v, 0x13C, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(v >> 32), 0x13C, 0xF, 0xF, true);
    return ((u64)(uint32_t)hi << 32) | (uint32_t)lo;
}
// Whole-wavefront shift by one lane (DPP wave_shr:1): lane l receives the
// value of lane l - 1, lane 0 receives 0.
__device__ __forceinline__ u64 shr1(u64 v) {
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)v, 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(v >> 32), 0x138, 0xF, 0xF, true);
    return ((u64)(uint32_t)hi << 32) | (uint32_t)lo;
}
__device__ __forceinline__ u64 readlane64(u64 v, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((u64)hi << 32) | lo;
}

// Base-3 digits of t over `levels` levels (level 1 = most significant digit):
// base has bit (levels - lev) set where the digit is 2, freem where it is 1.
// (Accumulated in locals with selects: conditional updates through the
// references became a scratch array indexed by the digit.)
__device__ __forceinline__ void digits3(int t, int levels, int &base, int &freem) {
    int b = 0, f = 0;
    for (int lev = levels; lev >= 1; --lev) {
        const int dig = t % 3;
        t /= 3;
        const int bit = 1 << (levels - lev);
        b |= dig == 2 ? bit : 0;
        f |= dig == 1 ? bit : 0;
    }
    base = b;
    freem = f;
}

constexpr int kKarThreads = 1024;
constexpr int kKarWaves = kKarThreads / 64;
// operand blocks per pair: P = 2^d s <= 4608 with s <= 64 gives d <= 7, so at
// most 2^(7-3) = 16 blocks of Q = 8 s <= 512 elements
constexpr int kKarMaxBlocks = 16;

// LDS words (u64) of one pair: result 2P, subtree operands 2Q, leaf products
// 3^B 2s, recombination buffer 3^(B-1) 4s, and 128 per wavefront.
__host__ __device__ inline int karatsuba_lds_words(int P, int s, int B) {
    int n3 = 1;
    for (int i = 0; i < B; ++i) n3 *= 3;
    const int Q = s << B;
    return 2 * P + 2 * Q + n3 * 2 * s + (B ? n3 / 3 * 4 * s : 0) + kKarWaves * 128;
}

// Pair geometry: P = 2^d s >= Lp with s <= 64, bottom levels B = min(d, 3),
// top levels T = d - B, subtree operand length Q = 2^B s.
__device__ __forceinline__ void kar_geometry(int64_t Lp, int &d, int &s, int &P, int &B, int &T, int &Q) {
    d = 0;
    int64_t s64 = Lp;
    while (s64 > 64) {
        ++d;
        s64 = (Lp + ((int64_t)1 << d) - 1) >> d;
    }
    s = (int)s64;
    P = s << d;
    B = d < 3 ? d : 3;
    T = d - B;
    Q = s << B;
}

// The pairs the MFMA kernel takes (the other kernel then skips them): every
// subtree product exact in FP64 and convertible back by the 2^52 + 2^51 bias.
// Subtree operands are sums of up to 2^T coefficients, so |X| < 2^(bits(mx) + T),
// |Y| < 2^(bits(my) + T), and every partial sum of the Q-term schoolbook
// products stays below 2^(bits(mx) + bits(my) + 2T + bits(Q)) <= 2^51.
__device__ __forceinline__ bool kar_mfma_takes(u64 mx, u64 my, int T, int Q) {
    const int bits = (64 - __clzll((long long)mx)) + (64 - __clzll((long long)my)) + 2 * T + (32 - __clz(Q));
    return bits <= 51;
}

// An integer-valued double of magnitude below 2^51 as Int64: added to
// 2^52 + 2^51 its value sits in the low mantissa bits, exactly.
__device__ __forceinline__ u64 kar_f64_to_i64(double v) {
    return (u64)__double_as_longlong(v + 6755399441055744.0) - (u64)0x4338000000000000ull;
}

// meta[3p] = input offset, meta[3p+1] = length, meta[3p+2] = output offset;
// taken (may be null): pairs k_karatsuba_mfma has done (taken[p] != 0) are skipped
__global__ void __launch_bounds__(kKarThreads)
k_karatsuba(const u64 *__restrict__ x, const u64 *__restrict__ y, u64 *__restrict__ z,
            const int64_t *__restrict__ meta, int64_t npairs, int cyclic, const int32_t *__restrict__ taken) {
    extern __shared__ __align__(16) u64 sm[];
    const int tid = threadIdx.x, NT = kKarThreads;
    const int lane = tid & 63, wid = tid >> 6;
    for (int64_t pi = blockIdx.x; pi < npairs; pi += gridDim.x) {
        const int64_t off = meta[3 * pi], Lp = meta[3 * pi + 1], zoff = meta[3 * pi + 2];
        if (Lp <= 0 || (taken && taken[pi])) continue;
        int d, s, P, B, T, Q;
        kar_geometry(Lp, d, s, P, B, T, Q);
        int nleaf = 1, nsub = 1;
        for (int i = 0; i < B; ++i) nleaf *= 3;
        for (int i = 0; i < T; ++i) nsub *= 3;
        u64 *acc = sm, *qx = acc + 2 * P, *qy = qx + Q, *R2 = qy + Q, *R1 = R2 + nleaf * 2 * s;
        for (int i = tid; i < 2 * P; i += NT) acc[i] = 0ull;
        // element tid of each of the 2^T operand blocks of size Q, held in
        // registers for the whole pair (Q <= 512 <= NT; 2^T <= kKarMaxBlocks)
        u64 xb[kKarMaxBlocks], yb[kKarMaxBlocks];
        u64 mx = 0, my = 0;  // largest magnitudes of the pair's coefficients
#pragma unroll
        for (int b = 0; b < kKarMaxBlocks; ++b) {
            const int64_t idx = (int64_t)b * Q + tid;
            const bool in = b < (1 << T) && tid < Q && idx < Lp;
            xb[b] = in ? x[off + idx] : 0ull;
            yb[b] = in ? y[off + idx] : 0ull;
            mx = max(mx, (int64_t)xb[b] < 0 ? 0ull - xb[b] : xb[b]);
            my = max(my, (int64_t)yb[b] < 0 ? 0ull - yb[b] : yb[b]);
        }
        // Leaf products in FP64 are exact when every leaf operand and every sum of
        // s products is an integer below 2^53: leaf operands are sums of up to 2^d
        // coefficients, so |X| < 2^(bits(mx) + d), |Y| < 2^(bits(my) + d), and the
        // s-term sums stay below 2^(bits(mx) + bits(my) + 2d + bits(s)).
        u64 *red = R1 + (B ? nleaf / 3 * 4 * s : 0);  // kKarWaves * 128 spare words
        for (int o = 32; o >= 1; o >>= 1) {
            mx = max(mx, (u64)__shfl_xor((unsigned long long)mx, o));
            my = max(my, (u64)__shfl_xor((unsigned long long)my, o));
        }
        if (lane == 0) {
            red[wid] = mx;
            red[kKarWaves + wid] = my;
        }
        __syncthreads();
        // every thread reduces the wavefront maxima (LDS broadcast reads), so the
        // choice below is uniform over the workgroup
        mx = my = 0;
        for (int w = 0; w < kKarWaves; ++w) {
            mx = max(mx, red[w]);
            my = max(my, red[kKarWaves + w]);
        }
        const int bits = (64 - __clzll((long long)mx)) + (64 - __clzll((long long)my)) + 2 * d + (32 - __clz(s));
        const bool exact_f64 = mx != 0 && my != 0 && bits <= 53;
        for (int sub = 0; sub < nsub; ++sub) {
            int tbase, tfree;
            digits3(sub, T, tbase, tfree);
            // subtree operands: the sum of the Q-blocks b with b & ~tfree == tbase
            __syncthreads();
            if (tid < Q) {
                u64 X = 0, Y = 0;
#pragma unroll
                for (int b = 0; b < kKarMaxBlocks; ++b) {
                    const bool sel = (b & ~tfree) == tbase;
                    X += sel ? xb[b] : 0ull;
                    Y += sel ? yb[b] : 0ull;
                }
                qx[tid] = X;
                qy[tid] = Y;
            }
            __syncthreads();
            // leaf products, one wavefront per leaf
            for (int leaf = wid; leaf < nleaf; leaf += kKarWaves) {
                int base, freem;
                digits3(leaf, B, base, freem);
                u64 X = 0, Y = 0;
                if (lane < s) {
                    int m = freem;
                    for (;;) {
                        const int b = base | m;
                        X += qx[b * s + lane];
                        Y += qy[b * s + lane];
                        if (m == 0) break;
                        m = (m - 1) & freem;
                    }
                }
                // schoolbook in registers: x_i is broadcast from lane i and y moves one lane
                // per step.  Rotated, lane l sees y_((l-i) mod 64); shifted (zeros entering at
                // lane 0), it sees y_(l-i) for i <= l and 0 otherwise.  The blocks are zero
                // beyond s, so lane l's z_l is the shifted sum and z_(l+64) is the rotated sum
                // minus the shifted one.
                // (steps run in groups of four up to s rounded up: x_i is zero for
                // i >= s, so the extra steps add nothing, and the groups unroll)
                const int s4 = (s + 3) & ~3;
                u64 lo = 0, hi = 0;
                if (exact_f64) {
                    // FP64 fused multiply-adds: exact, see above
                    const double Xd = (double)(int64_t)X;
                    double yr = (double)(int64_t)Y, ys = yr, dt = 0.0, dl = 0.0;
                    for (int i0 = 0; i0 < s4; i0 += 4) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const double xi =
                                __longlong_as_double((long long)readlane64((u64)__double_as_longlong(Xd), i0 + k));
                            dt = fma(xi, yr, dt);
                            dl = fma(xi, ys, dl);
                            yr = __longlong_as_double((long long)rot1((u64)__double_as_longlong(yr)));
                            ys = __longlong_as_double((long long)shr1((u64)__double_as_longlong(ys)));
                        }
                    }
                    lo = (u64)(int64_t)dl;
                    hi = (u64)(int64_t)(dt - dl);
                } else {
                    u64 yr = Y, ys = Y, t = 0;
                    for (int i0 = 0; i0 < s4; i0 += 4) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const u64 xi = readlane64(X, i0 + k);
                            t += xi * yr;
                            lo += xi * ys;
                            yr = rot1(yr);
                            ys = shr1(ys);
                        }
                    }
                    hi = t - lo;
                }
                u64 *pr = R2 + leaf * 2 * s;  // 2s slots: z_0 .. z_(2s-2) and the zero pad
                if (lane < 2 * s) pr[lane] = lo;
                if (64 + lane < 2 * s) pr[64 + lane] = hi;
            }
            __syncthreads();
            // bottom levels, deepest first: triples (digit 0, 1, 2) of products of
            // 2h slots become one product of 4h slots, z0 + u (z1 - z0 - z2) + u^2 z2
            u64 *src = R2, *dst = R1;
            int h = s, cnt = nleaf;
            for (int lev = B; lev >= 1; --lev) {
                const int ng = cnt / 3, w4 = 4 * h;
                for (int idx = tid; idx < ng * w4; idx += NT) {
                    const int g = idx / w4, k = idx - g * w4;
                    const u64 *z0 = src + 3 * g * 2 * h, *z1 = z0 + 2 * h, *z2 = z1 + 2 * h;
                    u64 v = 0;
                    if (k < 2 * h) v += z0[k];
                    if (k >= h && k < 3 * h) v += z1[k - h] - z0[k - h] - z2[k - h];
                    if (k >= 2 * h) v += z2[k - 2 * h];
                    dst[idx] = v;
                }
                __syncthreads();
                u64 *tmp = src;
                src = dst;
                dst = tmp;
                h *= 2;
                cnt = ng;
            }
            // src: the subtree product (2Q slots); add it for each top-level term
            const int nvar = T - __popc(tfree);
            for (int c = 0; c < (1 << nvar); ++c) {
                int64_t shift = 0;
                bool neg = false;
                int v = 0;
                for (int lev = 1; lev <= T; ++lev) {
                    const int64_t hl = (int64_t)s << (d - lev);
                    const int bit = 1 << (T - lev);
                    const int t = (tbase & bit) ? 2 : ((tfree & bit) ? 1 : 0);
                    if (t == 1) {
                        shift += hl;
                    } else {
                        const int bc = (c >> v) & 1;
                        ++v;
                        if (t == 0) {
                            if (bc) { shift += hl; neg = !neg; }
                        } else {
                            if (bc) shift += 2 * hl;
                            else { shift += hl; neg = !neg; }
                        }
                    }
                }
                for (int k = tid; k < 2 * Q - 1; k += NT) {
                    const u64 p = src[k];
                    acc[shift + k] += neg ? (u64)0 - p : p;
                }
                __syncthreads();
            }
        }
        __syncthreads();
        if (cyclic) {
            for (int64_t k = tid; k < Lp; k += NT)
                z[zoff + k] = acc[k] + ((k + Lp < 2 * Lp - 1) ? acc[k + Lp] : 0ull);
        } else {
            for (int64_t k = tid; k < 2 * Lp - 1; k += NT) z[zoff + k] = acc[k];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Karatsuba with matrix-core subtree products (pairs that kar_mfma_takes accepts)
// ---------------------------------------------------------------------------
//
// The top T levels are walked as in k_karatsuba, but each subtree's product --
// two operands of length Q = 2^B s, integers small enough that FP64 holds every
// partial sum exactly -- is one schoolbook product on the FP64 matrix cores
// instead of B more Karatsuba levels and 3^B leaf products.  With y cut into
// 16-blocks y_b, output block c of z = x y is
//   z_c[r] = sum_delta sum_{q<16} T_delta[r][q] y_{c-delta}[q],
//   T_delta[r][q] = x[16 delta + r - q]  (zero outside the operand),
// so a tile of 16 output blocks c0 ... c0 + 15 is, for each delta, the 16 x 16
// Toeplitz block T_delta times the 16 x 16 matrix of y blocks c0 - delta ...
// c0 + 15 - delta: four v_mfma_f64_16x16x4 per (tile, delta).  Both operands
// come from zero-padded LDS copies (y skewed to 17 words per block, so the
// lanes of one MFMA operand fall in different banks) at fixed offsets.  The
// (tile, delta) units are dealt to the 16 wavefronts in contiguous runs; a
// wavefront stages each finished tile in LDS and adds it, converted to Int64
// (exact), into the LDS result once per top-level term with 64-bit LDS atomics
// on consecutive words.  The atomics commute, so no barrier orders the terms,
// and over the ring of 64-bit words the bits equal every other exact evaluation.
typedef double kar_v4d __attribute__((ext_vector_type(4)));
constexpr int kKarMfmaMaxTerms = 1 << (7 - 3);  // 2^T, T <= 4 for P <= 4608
constexpr int kKarPadX = 16;                      // zero words before x
constexpr int kKarPadB = 16;                      // zero blocks before y
constexpr int kKarStage = 16 * 17;                // one wavefront's tile, skewed

// LDS words of one pair: result 2P, two sets (one subtree computed while the
// next is formed) of x (Qp + 32), skewed y ((Qp/16 + 32) 17) and the term table,
// the wavefronts' tiles and the reduction slots.
__host__ __device__ inline int karatsuba_mfma_lds_words(int P, int Q) {
    const int Qp = (Q + 15) & ~15;
    return 2 * P + 2 * ((Qp + 32) + (Qp / 16 + 32) * 17) + kKarWaves * kKarStage + 4 * kKarMfmaMaxTerms +
           2 * kKarWaves;
}

// taken[p] = 1 for the pairs done here, 0 for those left to k_karatsuba
__global__ void __launch_bounds__(kKarThreads)
k_karatsuba_mfma(const u64 *__restrict__ x, const u64 *__restrict__ y, u64 *__restrict__ z,
                 const int64_t *__restrict__ meta, int64_t npairs, int cyclic, int32_t *__restrict__ taken) {
    extern __shared__ __align__(16) u64 sm[];
    const int tid = threadIdx.x, NT = kKarThreads;
    // the wavefront index as a scalar: the unit runs and their loops are uniform
    const int lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r16 = lane & 15, k4 = lane >> 4;  // MFMA operand row / column and K slot
    for (int64_t pi = blockIdx.x; pi < npairs; pi += gridDim.x) {
        const int64_t off = meta[3 * pi], Lp = meta[3 * pi + 1], zoff = meta[3 * pi + 2];
        if (Lp <= 0) continue;
        int d, s, P, B, T, Q;
        kar_geometry(Lp, d, s, P, B, T, Q);
        const int Qp = (Q + 15) & ~15, nb = Qp / 16;
        const int nxs = Qp + 32, nys = (nb + 32) * 17;
        u64 *acc = sm;
        double *xs = (double *)(acc + 2 * P);  // set k: x at xs + k (nxs + nys), y after it
        double *stage = xs + 2 * (nxs + nys) + wid * kKarStage;
        int64_t *tshift = (int64_t *)(xs + 2 * (nxs + nys) + kKarWaves * kKarStage);  // set k at k 16
        int64_t *tneg = tshift + 2 * kKarMfmaMaxTerms;
        u64 *red = (u64 *)(tneg + 2 * kKarMfmaMaxTerms);
        // element e of each block of one operand, as doubles (exact for the pairs
        // taken; the others leave before using them): threads e < Qp hold x,
        // threads Qp + e hold y
        const bool hold_y = tid >= Qp;
        const int e = hold_y ? tid - Qp : tid;
        const u64 *src = hold_y ? y : x;
        double ob[kKarMaxBlocks];
        u64 mx = 0, my = 0;
#pragma unroll
        for (int b = 0; b < kKarMaxBlocks; ++b) {
            const int64_t idx = (int64_t)b * Q + e;
            const bool in = b < (1 << T) && e < Q && tid < 2 * Qp && idx < Lp;
            const u64 v = in ? src[off + idx] : 0ull;
            const u64 a = (int64_t)v < 0 ? 0ull - v : v;
            if (hold_y) my = max(my, a);
            else mx = max(mx, a);
            ob[b] = (double)(int64_t)v;
        }
        for (int o = 32; o >= 1; o >>= 1) {
            mx = max(mx, (u64)__shfl_xor((unsigned long long)mx, o));
            my = max(my, (u64)__shfl_xor((unsigned long long)my, o));
        }
        if (lane == 0) {
            red[wid] = mx;
            red[kKarWaves + wid] = my;
        }
        __syncthreads();
        mx = my = 0;
        for (int w = 0; w < kKarWaves; ++w) {
            mx = max(mx, red[w]);
            my = max(my, red[kKarWaves + w]);
        }
        const bool take = kar_mfma_takes(mx, my, T, Q);
        if (tid == 0) taken[pi] = take ? 1 : 0;
        if (!take) {
            __syncthreads();  // every thread has read red; k_karatsuba takes this pair
            continue;
        }
        for (int i = tid; i < 2 * P; i += NT) acc[i] = 0ull;
        for (int i = tid; i < 2 * (nxs + nys); i += NT) xs[i] = 0.0;  // the padding stays zero
        __syncthreads();  // zeroed before any operand is formed
        int nsub = 1;
        for (int i = 0; i < T; ++i) nsub *= 3;
        // the (tile, delta) units of one subtree product: tiles of 16 output
        // blocks c0 ..., delta in [c0 - nb + 1, c0 + 15] within [0, nb]
        const int ntile = (2 * nb + 15) / 16;
        int nunit = 0;
        for (int tt = 0; tt < ntile; ++tt) {
            const int c0 = 16 * tt, dlo = max(0, c0 - nb + 1), dhi = min(nb, c0 + 15);
            nunit += dhi >= dlo ? dhi - dlo + 1 : 0;
        }
        const int u_begin = (int)((int64_t)nunit * wid / kKarWaves), u_end = (int)((int64_t)nunit * (wid + 1) / kKarWaves);
        // operands and terms of subtree `sb` into set k
        auto form = [&](int sb, int k) {
            int tbase, tfree;
            digits3(sb, T, tbase, tfree);
            double *xk = xs + k * (nxs + nys), *yk = xk + nxs;
            if (tid < 2 * Qp) {
                // the selected blocks' sum: one FMA per block with a uniform 0/1
                // weight (exact: integers below 2^51)
                double X = 0.0;
#pragma unroll
                for (int b = 0; b < kKarMaxBlocks; ++b) X = fma((b & ~tfree) == tbase ? 1.0 : 0.0, ob[b], X);
                if (hold_y) yk[(kKarPadB + (e >> 4)) * 17 + (e & 15)] = X;
                else xk[kKarPadX + e] = X;
            }
            // the subtree's top-level terms (shift, sign) of prod_i c_{t_i}(u_i)
            const int nvar = T - __popc(tfree);
            if (tid < (1 << nvar)) {
                const int c = tid;
                int64_t shift = 0;
                bool neg = false;
                int v = 0;
                for (int lev = 1; lev <= T; ++lev) {
                    const int64_t hl = (int64_t)s << (d - lev);
                    const int bit = 1 << (T - lev);
                    const int t = (tbase & bit) ? 2 : ((tfree & bit) ? 1 : 0);
                    if (t == 1) {
                        shift += hl;
                    } else {
                        const int bc = (c >> v) & 1;
                        ++v;
                        if (t == 0) {
                            if (bc) { shift += hl; neg = !neg; }
                        } else {
                            if (bc) shift += 2 * hl;
                            else { shift += hl; neg = !neg; }
                        }
                    }
                }
                tshift[k * kKarMfmaMaxTerms + c] = shift;
                tneg[k * kKarMfmaMaxTerms + c] = neg ? 1 : 0;
            }
        };
        form(0, 0);
        __syncthreads();
        // subtree sb from set sb & 1 while the next one is formed into the
        // other set: one barrier per subtree
        for (int sb = 0; sb < nsub; ++sb) {
            const int k = sb & 1;
            int tbase, tfree;
            digits3(sb, T, tbase, tfree);
            const int nvar = T - __popc(tfree);
            const double *xk = xs + k * (nxs + nys), *yk = xk + nxs;
            const int64_t *tsh = tshift + k * kKarMfmaMaxTerms, *tng = tneg + k * kKarMfmaMaxTerms;
            const int nterm = 1 << nvar;
            kar_v4d D = {0.0, 0.0, 0.0, 0.0};
            int cur_tile = -1;
            // a finished tile: D holds output coefficient 16 (c0 + col) + row,
            // row = k4 + 4 v, col = r16; staged, then added with consecutive lanes
            // on consecutive coefficients
            auto flush = [&](int tile) {
#pragma unroll
                for (int v = 0; v < 4; ++v) stage[r16 * 17 + k4 + 4 * v] = D[v];
                wave_sync();
                const int base = 256 * tile;
                u64 val[4], nval[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int e = lane + 64 * q;
                    val[q] = base + e < 2 * Q - 1 ? kar_f64_to_i64(stage[(e >> 4) * 17 + (e & 15)]) : 0ull;
                    nval[q] = (u64)0 - val[q];
                }
                for (int c = 0; c < nterm; ++c) {
                    // the term's shift and sign, uniform: scalar registers
                    const int64_t sh = __builtin_amdgcn_readfirstlane((int)tsh[c]);
                    const bool ng = __builtin_amdgcn_readfirstlane((int)tng[c]) != 0;
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (val[q] != 0ull)
                            atomicAdd((unsigned long long *)&acc[sh + base + lane + 64 * q],
                                      (unsigned long long)(ng ? nval[q] : val[q]));
                }
                wave_sync();
            };
            int u = 0;
            for (int tt = 0; tt < ntile; ++tt) {
                const int c0 = 16 * tt, dlo = max(0, c0 - nb + 1), dhi = min(nb, c0 + 15);
                if (dhi < dlo) continue;
                const int cnt = dhi - dlo + 1;
                const int lo = max(u, u_begin), hi = min(u + cnt, u_end);
                if (lo < hi) {
                    if (cur_tile != tt) {
                        if (cur_tile >= 0) flush(cur_tile);
                        D = (kar_v4d){0.0, 0.0, 0.0, 0.0};
                        cur_tile = tt;
                    }
                    for (int dl = dlo + (lo - u); dl < dlo + (hi - u); ++dl) {
                        // A: T_delta[r16][q] = x[16 delta + r16 - q], q = 4 kap + k4
                        const double *xa = xk + kKarPadX + 16 * dl + r16 - k4 - 12;
                        // B: y block c0 + r16 - delta, element q
                        const double *yv = yk + (c0 + r16 - dl + kKarPadB) * 17 + k4;
                        D = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[12], yv[0], D, 0, 0, 0);
                        D = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[8], yv[4], D, 0, 0, 0);
                        D = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[4], yv[8], D, 0, 0, 0);
                        D = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[0], yv[12], D, 0, 0, 0);
                    }
                }
                u += cnt;
            }
            if (cur_tile >= 0) flush(cur_tile);
            if (sb + 1 < nsub) form(sb + 1, k ^ 1);
            __syncthreads();
        }
        if (cyclic) {
            for (int64_t k = tid; k < Lp; k += NT)
                z[zoff + k] = acc[k] + ((k + Lp < 2 * Lp - 1) ? acc[k + Lp] : 0ull);
        } else {
            for (int64_t k = tid; k < 2 * Lp - 1; k += NT) z[zoff + k] = acc[k];
        }
        __syncthreads();
    }
}

// Convolution.Adhoc.complex (Convolution/Adhoc.hs:114-133) in the Gauss form:
// three real products xr*yr, xi*yi and (xr + xi)*(yr + yi).  The split writes
// the three operand pairs as consecutive blocks of `count` words; the join
// takes the three products (blocks of `count` words) to
// re = p0 - p1, im = p2 - (p0 + p1), all in the ring of 64-bit words.
__global__ void __launch_bounds__(256)
k_gauss_split(const u64 *__restrict__ xr, const u64 *__restrict__ xi, const u64 *__restrict__ yr,
              const u64 *__restrict__ yi, u64 *__restrict__ x3, u64 *__restrict__ y3, int64_t count) {
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += step) {
        const u64 a = xr[i], b = xi[i], c = yr[i], d = yi[i];
        x3[i] = a;
        x3[count + i] = b;
        x3[2 * count + i] = a + b;
        y3[i] = c;
        y3[count + i] = d;
        y3[2 * count + i] = c + d;
    }
}

__global__ void __launch_bounds__(256)
k_gauss_join(const u64 *__restrict__ p3, u64 *__restrict__ zr, u64 *__restrict__ zi, int64_t count) {
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += step) {
        const u64 p0 = p3[i], p1 = p3[count + i], p2 = p3[2 * count + i];
        zr[i] = p0 - p1;
        zi[i] = p2 - (p0 + p1);
    }
}

static unsigned elementwise_grid(int64_t count) {
    const int64_t g = (count + 255) / 256;
    return (unsigned)(g < 8192 ? g : 8192);
}

// Top levels of pairs longer than one workgroup's LDS holds, in global memory.
// A split task takes one pair (x, y) of length L to the three pairs of length
// h = ceil(L/2) of one Karatsuba level (karatsubaReorder, Convolution/Adhoc.hs:66-79):
// low halves, high halves zero-padded to h, and their sums, stored as three
// consecutive h-blocks at cx and cy.  A join task takes the three products of
// such a split (consecutive (2h-1)-blocks at zc) to
// z = zl + u^h (zm - zl - zr) + u^(2h) zr (karatsubaGo, Adhoc.hs:81-98),
// 2L - 1 coefficients, or their cyclic wrap to L (Adhoc.hs:104-111).  All in
// the ring of 64-bit words, so the bits equal any other exact evaluation.
__global__ void __launch_bounds__(256)
k_kar_split(const KarSplitTask *__restrict__ tasks, int64_t ntasks) {
    for (int64_t t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const KarSplitTask k = tasks[t];
        const int64_t h = (k.L + 1) / 2;
        const u64 *px = (const u64 *)k.x, *py = (const u64 *)k.y;
        u64 *cx = (u64 *)k.cx, *cy = (u64 *)k.cy;
        for (int64_t i = threadIdx.x; i < h; i += blockDim.x) {
            const u64 xl = px[i], yl = py[i];
            const u64 xh = h + i < k.L ? px[h + i] : 0ull, yh = h + i < k.L ? py[h + i] : 0ull;
            cx[i] = xl;
            cx[h + i] = xh;
            cx[2 * h + i] = xl + xh;
            cy[i] = yl;
            cy[h + i] = yh;
            cy[2 * h + i] = yl + yh;
        }
    }
}

__device__ __forceinline__ u64 kar_join_at(const u64 *zc, int64_t h, int64_t k) {
    const int64_t w = 2 * h - 1;
    const u64 *zl = zc, *zr = zc + w, *zm = zc + 2 * w;
    u64 v = 0;
    if (k < w) v += zl[k];
    const int64_t m = k - h;
    if (m >= 0 && m < w) v += zm[m] - zl[m] - zr[m];
    const int64_t r = k - 2 * h;
    if (r >= 0 && r < w) v += zr[r];
    return v;
}

__global__ void __launch_bounds__(256)
k_kar_join(const KarJoinTask *__restrict__ tasks, int64_t ntasks) {
    for (int64_t t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const KarJoinTask k = tasks[t];
        const int64_t h = (k.L + 1) / 2, L = k.L;
        const u64 *zc = (const u64 *)k.zc;
        u64 *z = (u64 *)k.z;
        if (k.cyclic) {
            for (int64_t i = threadIdx.x; i < L; i += blockDim.x)
                z[i] = kar_join_at(zc, h, i) + (i + L < 2 * L - 1 ? kar_join_at(zc, h, i + L) : 0ull);
        } else {
            for (int64_t i = threadIdx.x; i < 2 * L - 1; i += blockDim.x) z[i] = kar_join_at(zc, h, i);
        }
    }
}

hipError_t launch_kar_split(const KarSplitTask *dev_tasks, int64_t ntasks, hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    const unsigned g = (unsigned)(ntasks < 65536 ? ntasks : 65536);
    hipLaunchKernelGGL(k_kar_split, dim3(g), dim3(256), 0, s, dev_tasks, ntasks);
    return hipGetLastError();
}

hipError_t launch_kar_join(const KarJoinTask *dev_tasks, int64_t ntasks, hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    const unsigned g = (unsigned)(ntasks < 65536 ? ntasks : 65536);
    hipLaunchKernelGGL(k_kar_join, dim3(g), dim3(256), 0, s, dev_tasks, ntasks);
    return hipGetLastError();
}

hipError_t launch_gauss_split(const int64_t *xr, const int64_t *xi, const int64_t *yr, const int64_t *yi, int64_t *x3,
                              int64_t *y3, int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gauss_split, dim3(elementwise_grid(count)), dim3(256), 0, s, (const u64 *)xr,
                       (const u64 *)xi, (const u64 *)yr, (const u64 *)yi, (u64 *)x3, (u64 *)y3, count);
    return hipGetLastError();
}

hipError_t launch_gauss_join(const int64_t *p3, int64_t *zr, int64_t *zi, int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gauss_join, dim3(elementwise_grid(count)), dim3(256), 0, s, (const u64 *)p3, (u64 *)zr,
                       (u64 *)zi, count);
    return hipGetLastError();
}

void init_karatsuba_attributes() {
    (void)hipFuncSetAttribute((const void *)k_karatsuba, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute((const void *)k_karatsuba_mfma, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

hipError_t launch_karatsuba(const int64_t *x, const int64_t *y, int64_t *z, const int64_t *dev_meta, int64_t npairs,
                            int max_P, int cyclic, int32_t *dev_taken, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    // LDS of the largest pair: P = 2^d s with s <= 64
    int words = 0;
    for (int d = 0; (1 << d) <= max_P; ++d)
        for (int sl = 1; sl <= 64; ++sl)
            if ((sl << d) <= max_P) words = std::max(words, karatsuba_lds_words(sl << d, sl, d < 3 ? d : 3));
    const size_t lds = (size_t)words * sizeof(u64);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const int64_t g = npairs < (int64_t)1 << 30 ? npairs : (int64_t)1 << 30;
    // pairs with small enough values first, on the matrix cores; the other
    // kernel skips those (dev_taken: one int32 per pair)
    int mwords = 0;
    for (int d = 0; (1 << d) <= max_P; ++d)
        for (int sl = 1; sl <= 64; ++sl)
            if ((sl << d) <= max_P) mwords = std::max(mwords, karatsuba_mfma_lds_words(sl << d, sl << (d < 3 ? d : 3)));
    const size_t mlds = (size_t)mwords * sizeof(u64);
    if (mlds > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_karatsuba_mfma, dim3((unsigned)g), dim3(kKarThreads), mlds, s, (const u64 *)x,
                       (const u64 *)y, (u64 *)z, dev_meta, npairs, cyclic, dev_taken);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_karatsuba, dim3((unsigned)g), dim3(kKarThreads), lds, s, (const u64 *)x, (const u64 *)y,
                       (u64 *)z, dev_meta, npairs, cyclic, (const int32_t *)dev_taken);
    return hipGetLastError();
}

}  // namespace af
