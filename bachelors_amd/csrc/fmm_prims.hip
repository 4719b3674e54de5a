// Device-wide scan and radix-sort primitives (rocPRIM), kept in one translation unit
// so the templates are instantiated once.  Used for the octree partition, the
// interaction-list compaction and the class sort of far pairs.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "fmm_prims.hpp"

namespace fmm {

// The octree point sorts take rocPRIM's onesweep path (a histogram pass plus one launch per
// 8-bit digit): the default switches to a block-sort + merge path, about 20 launches, for up
// to 2^20 items, which is slower for 32-bit keys at that size.  The pair sorts keep the default.
typedef rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config, rocprim::default_config, 0>
    OnesweepSort;

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, DevBuf& tmp, hipStream_t st) {
    if (n == 0) return;
    size_t bytes = 0;
    FMM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), st));
    void* t = tmp.reserve(bytes);
    FMM_HIP(rocprim::exclusive_scan(t, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), st));
}

void sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const int* vin, int* vout, size_t n, int end_bit,
                    DevBuf& tmp, hipStream_t st) {
    if (n == 0) return;
    size_t bytes = 0;
    FMM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, st));
    void* t = tmp.reserve(bytes);
    FMM_HIP(rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, st));
}

void sort_points_u32(const uint32_t* kin, uint32_t* kout, const int* vin, int* vout, size_t n, int end_bit,
                     DevBuf& tmp, hipStream_t st) {
    if (n == 0) return;
    size_t bytes = 0;
    FMM_HIP(rocprim::radix_sort_pairs<OnesweepSort>(nullptr, bytes, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, st));
    void* t = tmp.reserve(bytes);
    FMM_HIP(rocprim::radix_sort_pairs<OnesweepSort>(t, bytes, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, st));
}

void sort_pairs_u64(const uint64_t* kin, uint64_t* kout, const int* vin, int* vout, size_t n, int end_bit,
                    DevBuf& tmp, hipStream_t st) {
    if (n == 0) return;
    size_t bytes = 0;
    FMM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, st));
    void* t = tmp.reserve(bytes);
    FMM_HIP(rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, st));
}

}  // namespace fmm
