// Host planner of the MI355X transform path (no HIP calls in here).
//
// The reference plans on the host from a Haskell Int length (Planned.plan,
// Planned.hs:201-293): even lengths by split-radix, odd lengths by Good-Thomas
// over the first coprime Fermat factor pair, primes by Rader and, when Rader
// nests three deep, by Bluestein on ceiling5Smooth(2n-1).  The device plan
// computes the same DFT with the structure that suits gfx950: any length whose
// prime factors are radices of the register butterflies (2, 3, 5, 7, 11, 13)
// runs as a mixed-radix Stockham sequence in LDS (one pass up to the LDS
// limit, two passes of the four-step factorisation N = N1 N2 beyond it, three
// passes N = N1 N2 N3 for the largest plain transforms);
// primes p with smooth p-1 run Rader's convolution of length p-1; everything
// else runs Bluestein with a smooth padded length.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace af {

// NumberTheory.hs:94-101
int64_t ceil_pow2(int64_t n);
// NumberTheory.hs:103-124
int64_t ceil_5smooth(int64_t n);
// smallest m >= n whose prime factors are all in {2,3,5,7} (Bluestein pad)
int64_t ceil_7smooth(int64_t n);
bool is_prime(int64_t n);
// NumberTheory.hs:32-47: the smallest primitive root modulo a prime
int64_t multiplicative_generator(int64_t p);
int64_t mod_pow(int64_t base, int64_t expo, int64_t m);

// Radices of the register butterflies, or false when n has another prime factor.
bool stockham_radices(int64_t n, std::vector<int> &radices);

// Largest single-pass length and per-tile element budget for a precision.
int64_t lds_max_len(int prec);
int lds_tile_bytes_cap();
int elem_bytes(int prec);

enum CoreKind { CORE_NONE = 0, CORE_SINGLE = 1, CORE_FOUR = 2, CORE_THREE = 3 };
struct CoreShape {
    int kind = CORE_NONE;
    int64_t N = 0;
    std::vector<int> r;        // single pass radices
    int64_t N1 = 0, N2 = 0;    // four-step factors; three-pass: N = N1 N2 N3 with
    std::vector<int> r1, r2;   // input index a + N1 b + N1 N2 c (a fastest)
    int64_t N3 = 0;
    std::vector<int> r3;
};
// Shape of an FFT of length N by passes (single, four-step or, with
// allow_three, three passes); CORE_NONE if N is not smooth or does not factor
// within the LDS limit.  Three passes are used for powers of two from 2^24 on
// (every pass then has short lines and wide tiles) and for smooth lengths that
// two passes cannot hold; only plain transforms take them.
CoreShape core_shape(int prec, int64_t N, bool allow_three = false);

enum PlanAlgo {
    ALG_EMPTY = 0,     // n = 0: nothing to do
    ALG_COPY = 1,      // n = 1: identity
    ALG_STOCKHAM = 2,  // smooth n
    ALG_RADER = 3,     // prime n, Rader convolution of length n-1
    ALG_BLUESTEIN = 4, // chirp-z with padded length M
    ALG_DECOMP = 5     // transformDecompose of a length with a prime factor no device plan serves
};
// The split of a decomposition plan (Planned.hs:255-293 without Bluestein),
// composed on the host from sub-plans: Good-Thomas for coprime a b (the
// reference's skew-grid maps, no twiddles), Cooley-Tukey for a prime power
// n = a b, Rader for a prime whose p - 1 = a needs a decomposition itself.
enum DecompKind { DEC_GOOD_THOMAS = 1, DEC_COOLEY_TUKEY = 2, DEC_RADER = 3 };
struct PlanShape {
    int algo = ALG_EMPTY;
    int64_t n = 0;
    int64_t M = 0;     // Bluestein padded length
    CoreShape core;    // the FFT that runs on the device
    int dec = 0;       // ALG_DECOMP: DecompKind, with its factors a (and b)
    int64_t da = 0, db = 0;
};
// algorithm: af_algorithm (auto, decompose, chirp2, chirp235)
bool plan_shape(int prec, int64_t n, int algorithm, PlanShape &out, std::string &why, bool allow_three = false);
std::string describe_shape(const PlanShape &s);

}  // namespace af
