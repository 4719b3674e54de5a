// Reciprocals 1/n of the small integers the harmonic recurrences divide by (n <= RECIP_MAX):
// each entry is the correctly rounded FP64 value, the same number the division 1.0 / n gives,
// so a lookup (one cached load) replaces a runtime FP64 division (a dozen instructions) without
// changing any result.  One copy per translation unit (device code is not linked across units).
#pragma once
#include "fmm_common.hpp"

namespace fmm {

constexpr int RECIP_MAX = MAX_ORDER * MAX_ORDER;   // l^2 - m^2 and 2i for degrees <= MAX_ORDER

struct RecipTable {
    double v[RECIP_MAX + 1];
    constexpr RecipTable() : v() {
        for (int i = 1; i <= RECIP_MAX; i++) v[i] = 1.0 / (double)i;
    }
};
static __constant__ RecipTable RECIP = RecipTable();

__device__ __forceinline__ double recip(int n) { return RECIP.v[n]; }

// the same reciprocals rounded to FP32, for the FP32 recurrences
struct RecipTableF {
    float v[RECIP_MAX + 1];
    constexpr RecipTableF() : v() {
        for (int i = 1; i <= RECIP_MAX; i++) v[i] = (float)(1.0 / (double)i);
    }
};
static __constant__ RecipTableF RECIPF = RecipTableF();

__device__ __forceinline__ float recipf(int n) { return RECIPF.v[n]; }

}  // namespace fmm
