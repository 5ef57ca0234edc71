// Device helpers shared by the kernel sources.
#pragma once
#include <hip/hip_runtime.h>

namespace hipbeagle {

constexpr int WG = 256;

// Pointers that arrive inside job structs are generic; casting them to the
// global address space turns flat accesses (which wait on both the vector
// and the scalar counters) into global loads and stores.
#define GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const GLOBAL T* gin(const void* p) {
    return (const GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ GLOBAL T* gout(void* p) {
    return (GLOBAL T*)p;
}

// Job and item of this workgroup under a level work list (TileList in
// kernels.h): the largest j with first[j] <= blockIdx.x, found by a binary
// search over wave-uniform (scalar) loads; without a list the grid is
// (items, jobs).
template <typename List>
__device__ __forceinline__ void tile_item(const List& w, int& job, int& item) {
    if (!w.first) {
        job = blockIdx.y;
        item = blockIdx.x;
        return;
    }
    const int b = blockIdx.x;
    int lo = 0, hi = w.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (w.first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    job = lo;
    item = b - w.first[lo];
}

// x = f 2^e with f in [0.5, 1)
__device__ __forceinline__ int exponent_of(double x) {
    int e;
    frexp(x, &e);
    return e;
}
__device__ __forceinline__ int exponent_of(float x) {
    int e;
    frexpf(x, &e);
    return e;
}
__device__ __forceinline__ double pow2(double x, int e) { return ldexp(x, e); }
__device__ __forceinline__ float pow2(float x, int e) { return ldexpf(x, e); }

}  // namespace hipbeagle
