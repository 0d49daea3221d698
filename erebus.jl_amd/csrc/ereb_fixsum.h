// ereb_fixsum.h -- order-independent sums of Float64 contributions, for the
// cotangent grids of the pullbacks (a marker's cotangent spreads to the grid
// nodes of its stencils, from many markers at once).
//
// Every contribution v is rounded once to a multiple of 2^-S and added as a
// two-word integer: u = v * 2^(S-40) is split into hi = floor(u) and
// lo = rint((u - hi) * 2^40), and the words are summed with 64-bit integer
// atomics.  Integer addition is associative, so the sum is the same whatever
// order the atomics land in: the grids are bit-identical from run to run.
// The scale comes from a first pass over the same contributions that takes
// the largest |v| (atomicMax on the bit pattern, which orders non-negative
// doubles): with |v| * 2^(S-40) < 2^36, a node takes up to 2^27 hi words
// and 2^24 lo words without overflow, and each contribution is rounded to
// 2^-76 of the largest one.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

namespace ereb {

constexpr double FIX_LO = 0x1p40;   // weight of the hi word in lo units
constexpr double FIX_HI_MAX = 0x1p36;

// One grid's accumulator: words w[2k] (hi, signed) and w[2k+1] (lo, >= 0)
// of node k, and the scale up = 2^(S-40).  With w == nullptr the grid is not
// wanted and contributions are dropped.  The scale comes either from the
// host (up, after a first pass recorded the largest |v| in maxbits), or, when
// bound != nullptr, from a bound on |v| that a kernel wrote there before
// (fix_up_of_bound, read by every kernel that adds or reads the grid).
struct FixGrid {
    unsigned long long* w;
    unsigned long long* maxbits;  // first pass: largest |v| as bits
    double up;
    const unsigned long long* bound = nullptr;  // |v| <= this (as bits), or null
};

// the two words of v at scale up; false when v is outside the scale's range
// (or not finite)
__device__ __forceinline__ bool fix_split(double v, double up, unsigned long long& H,
                                          unsigned long long& L) {
    const double u = v * up;
    const double hi = floor(u);
    if (!(fabs(hi) < FIX_HI_MAX)) return false;
    H = (unsigned long long)(long long)hi;
    L = (unsigned long long)rint((u - hi) * FIX_LO);  // u - hi in [0, 1), exact
    return true;
}

// PASS 0: record |v| in maxbits.  PASS 1: add v to node k.  `flag` is set
// when a contribution lies outside the range of the scale (or is not
// finite), which the host reports as an error.
template <int PASS>
__device__ __forceinline__ void fix_add(const FixGrid& F, int64_t k, double v, int* flag) {
    if (!F.w || v == 0.0) return;
    if (PASS == 0) {
        atomicMax(F.maxbits, (unsigned long long)__double_as_longlong(fabs(v)));
        return;
    }
    unsigned long long H, L;
    if (!fix_split(v, F.up, H, L)) {
        atomicOr(flag, 1);
        return;
    }
    atomicAdd(&F.w[2 * k], H);
    atomicAdd(&F.w[2 * k + 1], L);
}

// the scale of a pass whose largest |contribution| is vmax: vmax * up < 2^36
inline double fix_scale_of(double vmax) {
    return vmax > 0.0 ? std::ldexp(1.0, std::min(35 - std::ilogb(vmax), 1000)) : 1.0;
}

// Bounds for the single-pass form: a contribution is a marker cotangent
// times factors of magnitude at most 1 times a bilinear weight; FIX_WEIGHT_MAX
// allows weights up to 256 (a position up to 15 cells outside its clamped
// cell).  The scale then rounds contributions to 2^-68 of the largest
// marker cotangent instead of 2^-76 of the largest contribution.
constexpr double FIX_WEIGHT_MAX = 256.0;

__device__ __forceinline__ double fix_up_of_bound(const unsigned long long* bound) {
    const double v = __longlong_as_double((long long)*bound) * FIX_WEIGHT_MAX;
    if (!(v > 0.0) || !(v < 1e300)) return 1.0;  // zero, or not finite: flagged on use
    return ldexp(1.0, min(35 - ilogb(v), 1000));
}

// the scale of F inside a kernel
__device__ __forceinline__ double fix_up(const FixGrid& F) {
    return F.bound ? fix_up_of_bound(F.bound) : F.up;
}


// value of node k: carry the lo words' whole units into hi, then one rounding
__device__ __forceinline__ double fix_value(const unsigned long long* w, int64_t k, double up) {
    const unsigned long long lo = w[2 * k + 1];
    const long long hi = (long long)w[2 * k] + (long long)(lo >> 40);
    const double f = (double)(lo & ((1ull << 40) - 1)) * (1.0 / FIX_LO);
    return ((double)hi + f) / up;
}

}  // namespace ereb
