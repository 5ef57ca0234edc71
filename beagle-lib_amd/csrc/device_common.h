// Device helpers shared by the gfx950 kernel files (kernels.hip, nodes.hip):
// 4-real loads and stores, workgroup and wave sums, buffer resources with
// empty ranges, the 4 x 5 matrix product with the pad column.
#pragma once
#include "kernels.h"
#include "device_util.h"

namespace hipbeagle {

template <typename R>
__device__ __forceinline__ void load4(const R* p, R (&v)[4]);
template <>
__device__ __forceinline__ void load4<double>(const double* p, double (&v)[4]) {
    const double2* q = reinterpret_cast<const double2*>(p);
    double2 a = q[0], b = q[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
template <>
__device__ __forceinline__ void load4<float>(const float* p, float (&v)[4]) {
    float4 a = *reinterpret_cast<const float4*>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void load4(const GLOBAL double* p, double (&v)[4]) {
    const GLOBAL dvec2* q = (const GLOBAL dvec2*)p;
    const dvec2 a = q[0], b = q[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void load4(const GLOBAL float* p, float (&v)[4]) {
    const fvec4 a = *(const GLOBAL fvec4*)p;
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void store4(GLOBAL double* p, const double (&v)[4]) {
    GLOBAL dvec2* q = (GLOBAL dvec2*)p;
    q[0] = dvec2{v[0], v[1]};
    q[1] = dvec2{v[2], v[3]};
}
__device__ __forceinline__ void store4(GLOBAL float* p, const float (&v)[4]) {
    *(GLOBAL fvec4*)p = fvec4{v[0], v[1], v[2], v[3]};
}
// The 64 lanes' 4-state FP64 vectors of 64 consecutive patterns (lane l
// holds pattern l; t points at pattern 0) stored as two fully contiguous
// 1 KB writes instead of two writes of 16-byte halves 32 bytes apart:
// write h covers patterns [32 h, 32 h + 32), lane l writing half (l & 1)
// of pattern 32 h + l / 2, fetched from that pattern's lane.  All 64 lanes
// must be active.
__device__ __forceinline__ void store4_tile(GLOBAL double* t, const double (&v)[4], int lane) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int src = 32 * h + (lane >> 1);
        const double a = __shfl(v[0], src, 64), b = __shfl(v[1], src, 64);
        const double c = __shfl(v[2], src, 64), d = __shfl(v[3], src, 64);
        *(GLOBAL dvec2*)(t + 128 * h + 2 * lane) = (lane & 1) ? dvec2{c, d} : dvec2{a, b};
    }
}

template <typename R>
__device__ __forceinline__ void store4(R* p, const R (&v)[4]);
template <>
__device__ __forceinline__ void store4<double>(double* p, const double (&v)[4]) {
    double2* q = reinterpret_cast<double2*>(p);
    q[0] = make_double2(v[0], v[1]);
    q[1] = make_double2(v[2], v[3]);
}
template <>
__device__ __forceinline__ void store4<float>(float* p, const float (&v)[4]) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}

// Sum over a 256-lane workgroup; the result is valid in thread 0.  The tree
// is fixed for a given launch shape, so sums are reproducible run to run.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nw; ++w) r += sh[w];
    }
    __syncthreads();
    return r;
}

constexpr int TILE = 64;

// One 16-byte vector of reals.
template <typename R>
struct Vec16;
template <>
struct Vec16<float> {
    typedef fvec4 T;
    static constexpr int N = 4;
};
template <>
struct Vec16<double> {
    typedef dvec2 T;
    static constexpr int N = 2;
};

// Buffer resource over `base` when `on`, else an empty range over `valid`:
// an access to an empty range reads 0 and touches no memory, so a load a
// step does not need costs one instruction, no branch and no traffic.  The
// base and range are wave-uniform (job fields); readfirstlane states it,
// so the resource stays in scalar registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t range_if(const void* base, bool on,
                                                           const void* valid) {
    const unsigned long long a = reinterpret_cast<unsigned long long>(on ? base : valid);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    const int n = __builtin_amdgcn_readfirstlane(on ? -1 : 0);
    void* b = reinterpret_cast<void*>(((unsigned long long)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(b, (short)0, n, 0x00020000);
}

// Buffer resource over `bytes` bytes at `base` when `on`, else an empty
// range over `valid`: stores to an empty range (or past `bytes`) are
// dropped, so a store a step does not need costs one instruction and no
// branch.  Base and size are wave-uniform.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t range_n(const void* base, bool on,
                                                          unsigned bytes, const void* valid) {
    const unsigned long long a = reinterpret_cast<unsigned long long>(on ? base : valid);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    const int n = __builtin_amdgcn_readfirstlane(on ? (int)bytes : 0);
    void* b = reinterpret_cast<void*>(((unsigned long long)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(b, (short)0, n, 0x00020000);
}

// Buffer resource of n bytes at `base` (n from a job record, so already
// wave-uniform; -1 the whole buffer, 0 an empty range: no access reaches
// memory, so a null base is never dereferenced).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t range_w(const void* base, int n) {
    // readfirstlane states the uniformity (a no-op on scalar registers), so
    // no access needs a loop over distinct resources
    const unsigned long long a = reinterpret_cast<unsigned long long>(base);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    void* b = reinterpret_cast<void*>(((unsigned long long)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(b, (short)0, __builtin_amdgcn_readfirstlane(n),
                                             0x00020000);
}

// four reals at byte offset `at` of a buffer resource
template <typename R>
__device__ __forceinline__ void load4_range(__amdgpu_buffer_rsrc_t r, unsigned at, R (&v)[4]) {
    if (sizeof(R) == 4) {
        const auto q = __builtin_amdgcn_raw_buffer_load_b128(r, at, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = (R)__uint_as_float(q[i]);
    } else {
        const auto q0 = __builtin_amdgcn_raw_buffer_load_b128(r, at, 0, 0);
        const auto q1 = __builtin_amdgcn_raw_buffer_load_b128(r, at + 16, 0, 0);
        v[0] = (R)__hiloint2double((int)q0[1], (int)q0[0]);
        v[1] = (R)__hiloint2double((int)q0[3], (int)q0[2]);
        v[2] = (R)__hiloint2double((int)q1[1], (int)q1[0]);
        v[3] = (R)__hiloint2double((int)q1[3], (int)q1[2]);
    }
}

// t_i = sum_j M[i][j] x_j with x_4 selecting the pad column
template <typename R>
__device__ __forceinline__ void mat4(const R* m, const R (&x)[5], R (&t)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
        t[i] = m[i * 5 + 0] * x[0] + m[i * 5 + 1] * x[1] + m[i * 5 + 2] * x[2] +
               m[i * 5 + 3] * x[3] + m[i * 5 + 4] * x[4];
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// v moved across lanes by a DPP control; lanes without a source read 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

// v moved across lanes by a DPP control into the rows ROWS selects; the
// other rows read an unspecified value (no instruction clears them).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move_rows(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave with DPP steps (quad swaps, row shifts,
// row broadcasts); the total ends in lane 63 and is returned to all lanes.
// The broadcasts leave rows 0 and 2 unspecified: lane 63 reads row 2's
// sum before the first broadcast and row 1's (complete) after it.
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_move<0xB1, 0xf>(v);    // quad_perm [1,0,3,2]
    v += dpp_move<0x4E, 0xf>(v);    // quad_perm [2,3,0,1]: every lane holds its quad's sum
    v += dpp_move<0x114, 0xf>(v);   // row_shr:4
    v += dpp_move<0x118, 0xf>(v);   // row_shr:8: lanes 12-15 of each row hold the row sum
    v += dpp_move_rows<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_move_rows<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3: lane 63 holds the total
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

}  // namespace hipbeagle
