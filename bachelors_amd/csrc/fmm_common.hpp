// Shared definitions of the MI355X Biot-Savart FMM library (host + device).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace fmm {

// constants of the reference (FMM/FMMGPU/math.hpp:12-16)
constexpr double REF_PI = 3.14159265359;
constexpr double REF_MU0 = 1.2566370614e-6;
constexpr int MAX_LEVELS = 22;     // depth cap 21 (levels 0..21); the reference has none
constexpr int MAX_ORDER = 20;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FMM_HIP(x)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            throw ::fmm::Error(3, std::string(#x " failed: ") + hipGetErrorString(e_) + " at " + \
                                      __FILE__ + ":" + std::to_string(__LINE__));                  \
    } while (0)

inline int nlen(int p) { return (p + 1) * (p + 1); }
inline int pad_to(int v, int m) { return (v + m - 1) / m * m; }
inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Grow-only device buffer: reused across runs so steady-state calls do not allocate.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            if (ptr) (void)hipFree(ptr);
            ptr = o.ptr; bytes = o.bytes;
            o.ptr = nullptr; o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { if (ptr) (void)hipFree(ptr); }
    void* reserve(size_t b) {
        if (b <= bytes && ptr) return ptr;
        if (ptr) {
            // growth is rare (first use of a size); make sure no queued kernel still reads it
            FMM_HIP(hipDeviceSynchronize());
            FMM_HIP(hipFree(ptr));
        }
        ptr = nullptr;
        size_t nb = b < 256 ? 256 : b;
        FMM_HIP(hipMalloc(&ptr, nb));
        bytes = nb;
        return ptr;
    }
    template <class T> T* as() const { return (T*)ptr; }
    template <class T> T* get(size_t n) { return (T*)reserve(n * sizeof(T)); }
};

// Grow-only pinned host array (the host mirrors of device arrays): device-to-host copies into
// it are plain DMA, without the staging of pageable memory.  resize() keeps no contents.
template <class T> struct PinnedVec {
    T* p = nullptr;
    size_t n = 0, cap = 0;
    PinnedVec() = default;
    PinnedVec(const PinnedVec&) = delete;
    PinnedVec& operator=(const PinnedVec&) = delete;
    ~PinnedVec() { if (p) (void)hipHostFree(p); }
    void resize(size_t m) {
        if (m > cap) {
            if (p) FMM_HIP(hipHostFree(p));
            p = nullptr;
            FMM_HIP(hipHostMalloc((void**)&p, std::max<size_t>(m, 1) * sizeof(T), hipHostMallocDefault));
            cap = m;
        }
        n = m;
    }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
    T* data() { return p; }
    const T* data() const { return p; }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
};

struct ScaleTable {
    double s[MAX_LEVELS];
};

// ---- per-tree device description -------------------------------------------------
// Node arrays use canonical BFS numbering: level-k nodes are contiguous, and the 8
// children of a subdivided node are consecutive ids in reference child order.
// grid size for n work items in workgroups of b
inline int blocks(long long n, int b = 256) { return (int)((n + b - 1) / b); }

// compute units of the current device (queried once per process)
inline int num_cus() {
    static const int n = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
        return v;
    }();
    return n;
}

struct TreeDev {
    int num_nodes = 0;
    int num_slots = 0;                 // points kept by the tree (leaf DFS order)
    int num_levels = 0;                // deepest level + 1
    int level_begin[MAX_LEVELS + 2] = {0};
    double half[MAX_LEVELS][3];        // box half sizes per level (uniform within a level)
    double radius[MAX_LEVELS];         // |half| per level (box.cpp:41-46)
    double scale[MAX_LEVELS];          // expansion scale sigma per level
    DevBuf center;                     // double3 per node
    DevBuf parent, first_child, arrivals, npts, begin, end, level;  // int per node
    DevBuf slots;                      // int per slot: input point index
    PinnedVec<int> h_first_child, h_npts;         // host mirrors the evaluation reads
    PinnedVec<int> h_begin, h_end, h_level, h_parent;   // fetched on demand (fetch_tree_details)
    PinnedVec<double> h_center;
    bool details_on_host = false;
    int dropped = 0;                   // input points outside the root
    bool sorted_build = false;         // built by the one-sort path (else level by level)
};

}  // namespace fmm
