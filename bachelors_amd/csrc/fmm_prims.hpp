#pragma once
#include "fmm_common.hpp"

namespace fmm {
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, DevBuf& tmp, hipStream_t st);
// stable LSD radix sort on the low end_bit bits of the keys
void sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const int* vin, int* vout, size_t n, int end_bit,
                    DevBuf& tmp, hipStream_t st);
// the same sort by the onesweep algorithm at every size (the octree point sorts)
void sort_points_u32(const uint32_t* kin, uint32_t* kout, const int* vin, int* vout, size_t n, int end_bit,
                     DevBuf& tmp, hipStream_t st);
void sort_pairs_u64(const uint64_t* kin, uint64_t* kout, const int* vin, int* vout, size_t n, int end_bit,
                    DevBuf& tmp, hipStream_t st);
}  // namespace fmm
