// The running sum s + t_0 + t_1 + ... + t_{n-1}, rounded after every add in this order (a Fortran
// loop `s = s + t(i)`), computed by a whole workgroup instead of one lane, bit for bit.
//
// Let s be a normal double in the binade [2^e, 2^(e+1)) and u = 2^(e-52) its spacing, so s = M u
// with M an integer in [2^52, 2^53).  For t >= 0 the exact s + t is at least 2^e, and while it stays
// below 2^(e+1) - u/2 the nearest double is the nearest multiple of u: fl(s + t) = (M + k) u with k
// the nearest integer to t / u (exact: a power-of-two scaling), unless t / u is exactly half-way
// between two integers (then the tie goes to the even M + k, which depends on M).  So as long as
// every partial P_i = M + k_0 + ... + k_i stays at or below 2^53 - 1 and no t_i is a tie, the
// running sums are P_i u exactly: an integer prefix sum, which a workgroup forms in parallel.
//
// One pass from position pos counts every remaining term in s's binade; a wave scan and the waves'
// totals give every thread the prefix before and after its terms.  If the whole rest stays in the
// binade with no tie, the sum is (M + total) u.  Otherwise the one thread whose prefixes straddle
// the limit 2^53 - 1 - M finds the first term that leaves the binade; that term, or the first tie
// if it comes earlier, is added with a real floating-point add to the valid partial before it, and
// the next pass starts after it in the binade the sum has reached.  A running sum that is zero,
// subnormal or not finite, and terms that are negative or NaN, fall back to the plain sequential
// adds (one lane), so every input gives the sequential result.
//
// Counts are held as doubles: integers, exact while a prefix stays below 2^53, and a prefix past
// the limit stays past it when rounded, so every comparison with the limit is exact.
// This header holds the arithmetic shared by the device pass and its host check.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

namespace moed {

constexpr uint64_t kSumLim = (1ull << 53) - 1;  // largest valid partial M + prefix

__host__ __device__ inline bool sum_normal(double s) { return s >= DBL_MIN && s <= DBL_MAX; }

// t >= 0 in units of 2^(e - 52), rounded to the nearest integer and capped at 2^53 (a term that
// large leaves the binade by itself); tie: t is exactly half-way between two multiples
__host__ __device__ inline double sum_count(double t, int e, bool& tie) {
    double q = ldexp(t, 52 - e);
    q = q < 0x1p53 ? q : 0x1p53;
    q = q > 0.0 ? q : 0.0;
    tie = q < 0x1p52 && q - floor(q) == 0.5;
    return rint(q);
}

}  // namespace moed
