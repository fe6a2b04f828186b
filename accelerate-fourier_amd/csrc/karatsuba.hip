// Bit-exact Int64 Karatsuba convolution on gfx950.
//
// The reference (Convolution/Adhoc.hs:36-98) splits each operand in halves
// level after level down to single coefficients (3^d leaf products), then
// recombines with z = zl + u (zm - zl - zr) + u^2 zr.  Over the ring Z/2^64 every
// exact splitting gives the same bits, so here one workgroup handles one pair
// and organises the same algebra around LDS:
//
//  * both operands are zero-padded to P = 2^d s (leaf size s <= 64) in LDS;
//  * leaf t in {0,1,2}^d is the sum of the operand blocks selected by t
//    (t_i = 0: low half, 2: high half, 1: both -- the Karatsuba evaluation);
//  * one wavefront forms one leaf product of two s-long blocks, each lane
//    owning output coefficients lane and lane + s;
//  * the leaf product is added into the LDS result with the signs and shifts of
//    prod_i c_{t_i}(u_i), c_0 = 1 - u, c_1 = u, c_2 = u^2 - u (the Karatsuba
//    interpolation, unrolled over the levels), using 64-bit LDS atomics.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace af {

typedef unsigned long long u64;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// meta[3p] = input offset, meta[3p+1] = length, meta[3p+2] = output offset
__global__ void __launch_bounds__(1024)
k_karatsuba(const u64 *__restrict__ x, const u64 *__restrict__ y, u64 *__restrict__ z,
            const int64_t *__restrict__ meta, int64_t npairs, int cyclic) {
    extern __shared__ __align__(16) u64 sm[];
    const int tid = threadIdx.x, NT = blockDim.x;
    const int lane = tid & 63, wid = tid >> 6, nw = NT >> 6;
    for (int64_t pi = blockIdx.x; pi < npairs; pi += gridDim.x) {
        const int64_t off = meta[3 * pi], Lp = meta[3 * pi + 1], zoff = meta[3 * pi + 2];
        if (Lp <= 0) continue;
        int d = 0;
        int64_t s = Lp;
        while (s > 64) {
            ++d;
            s = (Lp + ((int64_t)1 << d) - 1) >> d;
        }
        const int P = (int)(s << d);
        u64 *xs = sm, *ys = sm + P, *acc = sm + 2 * P, *wb = sm + 4 * P;
        u64 *wx = wb + wid * 128, *wy = wx + 64;
        for (int i = tid; i < P; i += NT) {
            xs[i] = i < Lp ? x[off + i] : 0ull;
            ys[i] = i < Lp ? y[off + i] : 0ull;
        }
        for (int i = tid; i < 2 * P; i += NT) acc[i] = 0ull;
        __syncthreads();

        int nleaves = 1;
        for (int i = 0; i < d; ++i) nleaves *= 3;
        const int si = (int)s;
        for (int leaf = wid; leaf < nleaves; leaf += nw) {
            // digits: level lev (1 = top split) is the lev-th most significant base-3 digit
            // base: levels with digit 2, freem: levels with digit 1 (bit d - lev)
            int base = 0, freem = 0, rem = leaf;
            for (int lev = d; lev >= 1; --lev) {
                const int dig = rem % 3;
                rem /= 3;
                const int bit = 1 << (d - lev);
                if (dig == 2) base |= bit;
                if (dig == 1) freem |= bit;
            }
            // evaluation: sum of the selected blocks
            u64 X = 0, Y = 0;
            if (lane < si) {
                int m = freem;
                for (;;) {
                    const int b = base | m;
                    X += xs[b * si + lane];
                    Y += ys[b * si + lane];
                    if (m == 0) break;
                    m = (m - 1) & freem;
                }
            }
            wx[lane] = X;
            wy[lane] = Y;
            wave_sync();
            // leaf product: lane j owns coefficients j and j + s
            u64 lo = 0, hi = 0;
            if (lane < si) {
                for (int i = 0; i < si; ++i) {
                    const u64 xi = wx[i];
                    const int j = lane - i;
                    if (j >= 0) lo += xi * wy[j];
                    else hi += xi * wy[j + si];
                }
            }
            // interpolation terms
            const int nvar = d - __popc(freem);
            for (int c = 0; c < (1 << nvar); ++c) {
                int64_t shift = 0;
                bool neg = false;
                int v = 0;
                for (int lev = 1; lev <= d; ++lev) {
                    const int64_t h = s << (d - lev);
                    const int bit = 1 << (d - lev);
                    const int t = (base & bit) ? 2 : ((freem & bit) ? 1 : 0);
                    if (t == 1) {
                        shift += h;
                    } else {
                        const int bc = (c >> v) & 1;
                        ++v;
                        if (t == 0) {
                            if (bc) { shift += h; neg = !neg; }
                        } else {
                            if (bc) shift += 2 * h;
                            else { shift += h; neg = !neg; }
                        }
                    }
                }
                if (lane < si) atomicAdd(&acc[shift + lane], neg ? (u64)0 - lo : lo);
                if (lane < si - 1) atomicAdd(&acc[shift + lane + si], neg ? (u64)0 - hi : hi);
            }
            wave_sync();
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

void init_karatsuba_attributes() {
    (void)hipFuncSetAttribute((const void *)k_karatsuba, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

hipError_t launch_karatsuba(const int64_t *x, const int64_t *y, int64_t *z, const int64_t *dev_meta, int64_t npairs,
                            int max_P, int cyclic, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    const int nthreads = 1024;
    const size_t lds = (size_t)(4 * max_P + (nthreads / 64) * 128) * sizeof(u64);
    const int64_t g = npairs < (int64_t)1 << 30 ? npairs : (int64_t)1 << 30;
    hipLaunchKernelGGL(k_karatsuba, dim3((unsigned)g), dim3(nthreads), lds, s, (const u64 *)x, (const u64 *)y,
                       (u64 *)z, dev_meta, npairs, cyclic);
    return hipGetLastError();
}

}  // namespace af
