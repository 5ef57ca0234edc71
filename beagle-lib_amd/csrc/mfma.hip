// Matrix-core (MFMA) kernels for state counts above 4 (amino acids S = 20,
// codons S = 61).
//
// The post-order update of one category is two small GEMMs and a Hadamard
// product: dest = (P1 X1) * (P2 X2) with X of shape S x patterns
// (libhmsbeagle/GPU/kernels/kernelsX.cu:441-600 computes the same sums one
// lane per output element).  Here a wavefront owns one category and a tile
// of 32 patterns, and runs v_mfma_f32_32x32x2f32 with
//   A = P (32 output states x 2 inner states per instruction, from the
//       repacked matrix, one float4 per lane per four instructions),
//   B = X (2 inner states x 32 patterns, one dword per lane read straight
//       from the pattern-minor tiles of the partials buffer),
// so lane l ends up holding pattern l % 32 and, in element r of the
// accumulator of M block mb, output state 32 mb + 8 (r / 4) + 4 (l / 32) +
// r % 4.  Inner states are assigned to instructions as k = 2 step + l / 32;
// the repacked matrix follows the same assignment, padding with zeros
// (states >= S), and a partials read past the last state falls outside the
// tile's buffer resource and reads 0.  A tip-state child instead reads its
// column P[., s] from the matrix repacked as rows by state (row S is the pad
// column, 1 for P(t)).
//
// Rescaling is fused: the per-pattern maximum over states comes from the
// lane's registers and its partner lane (l ^ 32), the maximum over
// categories from one LDS exchange between the workgroup's waves.
#include "kernels.h"
#include "device_util.h"

#include <cmath>

namespace hipbeagle {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// x = hi + lo with hi = bf16(x) (round to nearest even) and lo = bf16(x - hi):
// 16 significant bits of x in two BF16 numbers
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// ---------------------------------------------------------------------------
// Operand access through buffer resources.  Base and size are wave-uniform
// (job fields, the category, the tile), so a resource sits in scalar
// registers and every load is one instruction with a per-lane offset fixed
// for the kernel and a compile-time immediate: no 64-bit address arithmetic
// per load.  A load past a resource's size reads 0 and a store past it is
// dropped, which gives the padded states (rows >= S) for free.

typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc rsrc(const void* base, int bytes) {
    const unsigned long long a = reinterpret_cast<unsigned long long>(base);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    void* b = reinterpret_cast<void*>(((unsigned long long)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(b, (short)0, __builtin_amdgcn_readfirstlane(bytes),
                                             0x00020000);
}

// One category's S x PTILE tile of a partials buffer as two resources of up
// to 32 rows (states), so that every row offset within a resource fits the
// 12-bit immediate of the instruction: rows >= S fall outside both.
struct Tile {
    Rsrc r[2];
};
__device__ __forceinline__ Tile tile_of(const void* buf, long tbase, int S) {
    const float* b = (const float*)buf + tbase;
    return Tile{{rsrc(b, min(S, 32) * PTILE * 4), rsrc(b + 32 * PTILE, max(S - 32, 0) * PTILE * 4)}};
}
// element (row, lane's pattern) of a tile; `vo` holds the lane's byte offset
// (pattern n, plus the lane's row offset), `row` is compile-time
__device__ __forceinline__ float tile_load(const Tile& t, int row, unsigned vo) {
    return __uint_as_float(
        __builtin_amdgcn_raw_buffer_load_b32(t.r[row >> 5], vo + (row & 31) * PTILE * 4, 0, 0));
}
__device__ __forceinline__ void tile_store(const Tile& t, int row, unsigned vo, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), t.r[row >> 5],
                                          vo + (row & 31) * PTILE * 4, 0, 0);
}
// four floats at byte offset vo + off of a packed-matrix resource (`off`
// compile-time, placed in the scalar offset)
__device__ __forceinline__ floatx4 pack_load4(Rsrc r, unsigned vo, int off) {
    return __builtin_bit_cast(floatx4, __builtin_amdgcn_raw_buffer_load_b128(r, vo, off, 0));
}

// ---------------------------------------------------------------------------
// matrix repacking: one thread per output float

template <typename R>
__global__ __launch_bounds__(WG) void k_pack_mats(const PackJob* __restrict__ jobs, int njobs,
                                                  int S, int C, const R* __restrict__ mats,
                                                  float* __restrict__ pack) {
    const int per = mfma_pack_reals(S);
    const long t = (long)blockIdx.x * WG + threadIdx.x;
    if (t >= (long)njobs * C * per) return;
    const int j = (int)(t / ((long)C * per));
    const int rem = (int)(t - (long)j * C * per);
    const int c = rem / per, e = rem - c * per;
    const PackJob J = jobs[j];
    const R* M = mats + J.src + (long)c * S * (S + 1);
    const int MB = mfma_mblocks(S), KG = mfma_kgroups(S);
    float v = 0.f;
    // element (i, k) of the matrix as used, or of its transpose as stored
    const bool tin = (J.kind & PACK_TRANSPOSED_INPUT) != 0;
    const int kind = J.kind & ~PACK_TRANSPOSED_INPUT;
    auto at = [&](int i, int k) { return (float)(tin ? M[k * (S + 1) + i] : M[i * (S + 1) + k]); };
    if (kind == PACK_OPERAND) {
        // [mb][k group g][lane][q]: row 32 mb + lane % 32, column 8 g + 2 q + lane / 32
        if (e < MB * KG * 256) {
            const int mb = e / (KG * 256), r = e - mb * KG * 256;
            const int g = r / 256, lane = (r & 255) >> 2, q = r & 3;
            const int i = mb * 32 + (lane & 31), k = 8 * g + 2 * q + (lane >> 5);
            if (i < S && k < S) v = at(i, k);
        }
    } else if (kind == PACK_PRE_SELF) {
        // M^T u with u in the accumulator layout: [mo][group g][lane][e];
        // step s = 4 g + e holds accumulator element r = s % 16 of block
        // s / 16, i.e. inner state 32 (s / 16) + 8 (r / 4) + 4 (lane / 32) + r % 4;
        // output row j = 32 mo + lane % 32; value M[k][j]
        const int NG = MB * 4;
        if (e < MB * NG * 256) {
            const int mo = e / (NG * 256), r0 = e - mo * NG * 256;
            const int g = r0 / 256, lane = (r0 & 255) >> 2, q = r0 & 3;
            const int s4 = 4 * g + q, mb = s4 >> 4, r = s4 & 15;
            const int k = mb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int j = mo * 32 + (lane & 31);
            if (j < S && k < S) v = at(k, j);
        }
    } else if (kind == PACK_OPERAND_BF16) {
        // [mb][k step t][part: hi, lo][lane][8 BF16]: row 32 mb + lane % 32,
        // column 16 t + 8 (lane / 32) + e; one float holds elements 2e' and 2e'+1
        const int KT = mfma_ksteps16(S);
        if (e < MB * KT * 512) {
            unsigned short half[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int el = 2 * e + u;
                const int ee = el & 7, lane = (el >> 3) & 63, part = (el >> 9) & 1;
                const int blk = el >> 10, t = blk % KT, mb = blk / KT;
                const int i = mb * 32 + (lane & 31), k = 16 * t + 8 * (lane >> 5) + ee;
                const float x = (i < S && k < S) ? at(i, k) : 0.f;
                __bf16 hi, lo;
                split_bf16(x, hi, lo);
                half[u] = __builtin_bit_cast(unsigned short, part ? lo : hi);
            }
            v = __builtin_bit_cast(float, (unsigned)half[0] | ((unsigned)half[1] << 16));
        }
    } else if (kind == PACK_TIP_PRODUCT) {
        return;   // k_pack_tip_products
    } else if (kind == PACK_TIP_COLUMN) {
        // tip branch of the fused gradient, [state s (0..S)][row j]:
        // P[j][s mod S]
        const int W = MB * 32;
        if (e < (S + 1) * W) {
            const int s = e / W, j = e - s * W;
            if (j < S) v = at(j, s % S);
        }
    } else if (kind == PACK_PRE_SELF_BF16) {
        // M^T u with u in the accumulator layout on v_mfma_f32_32x32x16_bf16:
        // [mo][instruction s][part: hi, lo][lane][8 BF16]; slot e of lane
        // half h in instruction s is accumulator element 8 (s % 2) + e of
        // block s / 2, i.e. inner state 32 (s / 2) + 16 (s % 2) + 8 (e / 4) +
        // 4 h + e % 4, which the lane already holds; output row j = 32 mo +
        // lane % 32; value M[k][j]
        const int NS = 2 * MB;
        if (e < MB * NS * 512) {
            unsigned short half[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int el = 2 * e + u;
                const int ee = el & 7, lane = (el >> 3) & 63, part = (el >> 9) & 1;
                const int blk = el >> 10, s = blk % NS, mo = blk / NS;
                const int k = 32 * (s >> 1) + 16 * (s & 1) + 8 * (ee >> 2) + 4 * (lane >> 5) + (ee & 3);
                const int j = mo * 32 + (lane & 31);
                const float x = (j < S && k < S) ? at(k, j) : 0.f;
                __bf16 hi, lo;
                split_bf16(x, hi, lo);
                half[u] = __builtin_bit_cast(unsigned short, part ? lo : hi);
            }
            v = __builtin_bit_cast(float, (unsigned)half[0] | ((unsigned)half[1] << 16));
        }
    } else {
        // [state s (S: the pad column)][row i]; a transposed input has no pad
        // row: a missing state reads 0 there (the derivative matrices' pad)
        const int W = MB * 32;
        if (e < (S + 1) * W) {
            const int s = e / W, i = e - s * W;
            if (i < S) v = tin ? (s < S ? (float)M[s * (S + 1) + i] : 0.f)
                               : (float)M[i * (S + 1) + s];
        }
    }
    pack[J.dst + (long)c * per + e] = v;
}

// PACK_TIP_PRODUCT blocks, one workgroup per (job, category): P and D as
// used staged in LDS, then [state s (0..S)][row j] = (P D)[j][s] = sum_i
// P[j][i] D[i][s] summed in FP64 (s = S: the derivative matrix's pad column,
// 0 for a transposed input as in the rows-by-state pack).
template <typename R>
__global__ __launch_bounds__(WG) void k_pack_tip_products(const PackJob* __restrict__ jobs,
                                                          int S, int C, const R* __restrict__ mats,
                                                          float* __restrict__ pack) {
    __shared__ float sm[64 * 64], sd[64 * 65];
    const PackJob J = jobs[blockIdx.x];
    const int c = blockIdx.y;
    const bool tin = (J.kind & PACK_TRANSPOSED_INPUT) != 0;
    const R* M = mats + J.src + (long)c * S * (S + 1);
    const R* D = mats + J.src2 + (long)c * S * (S + 1);
    for (int e = threadIdx.x; e < S * S; e += WG) {
        const int j = e / S, i = e - j * S;
        sm[j * 64 + i] = (float)(tin ? M[i * (S + 1) + j] : M[j * (S + 1) + i]);
    }
    for (int e = threadIdx.x; e < S * (S + 1); e += WG) {
        const int i = e / (S + 1), s = e - i * (S + 1);
        sd[i * 65 + s] = (float)(tin ? (s < S ? D[s * (S + 1) + i] : (R)0) : D[i * (S + 1) + s]);
    }
    __syncthreads();
    const int W = mfma_mblocks(S) * 32, per = mfma_pack_reals(S);
    float* out = pack + J.dst + (long)c * per;
    for (int e = threadIdx.x; e < (S + 1) * W; e += WG) {
        const int s = e / W, j = e - s * W;
        double acc = 0.0;
        if (j < S)
            for (int i = 0; i < S; ++i) acc += (double)sm[j * 64 + i] * (double)sd[i * 65 + s];
        out[e] = (float)acc;
    }
}

template <typename R>
void launch_pack_tip_products(const PackJob* jobs, int njobs, int S, int C, const void* mats,
                              float* pack, hipStream_t st) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_pack_tip_products<R>, dim3(njobs, C), dim3(WG), 0, st, jobs, S, C,
                       reinterpret_cast<const R*>(mats), pack);
}

template <typename R>
void launch_pack_matrices(const PackJob* jobs, int njobs, int S, int C, const void* mats,
                          float* pack, hipStream_t st) {
    const long n = (long)njobs * C * mfma_pack_reals(S);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pack_mats<R>, dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, st, jobs,
                       njobs, S, C, reinterpret_cast<const R*>(mats), pack);
}

// ---------------------------------------------------------------------------
// post-order

// Operands of one pipeline stage, k group g of one partials child: the
// matrix (A; lane offset va = 16 l) and the partials (B; lane offset
// vx = 4 l, which is row h and pattern n of the tile).
template <int MB>
__device__ __forceinline__ void load_a(Rsrc a, unsigned va, int g, int KG, floatx4 (&am)[MB]) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) am[mb] = pack_load4(a, va, (mb * KG + g) * 1024);
}
__device__ __forceinline__ void load_x(const Tile& x, unsigned vx, int g, float (&b)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) b[q] = tile_load(x, 8 * g + 2 * q, vx);
}

template <int MB>
__device__ __forceinline__ void mfma_stage(floatx16 (&y)[MB], const floatx4 (&am)[MB],
                                           const float (&b)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
            y[mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(am[mb][q], b[q], y[mb], 0, 0, 0);
}

// Split-BF16 stages: k step t (16 inner states) of one partials child.
// A (v_mfma_f32_32x32x16_bf16): lane l holds P[32 mb + l % 32][16 t + 8 (l / 32) + e]
// as packed hi and lo parts; B: X[16 t + 8 (l / 32) + e][pattern l % 32].
// lane offsets: va = 16 l, vx = 1024 h + 4 n (row 8 h, pattern n)
template <int MB>
__device__ __forceinline__ void load_a16(Rsrc a, unsigned va, int t, int KT, u32x4 (&ah)[MB],
                                         u32x4 (&al)[MB]) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        ah[mb] = __builtin_amdgcn_raw_buffer_load_b128(a, va, ((mb * KT + t) * 2 + 0) * 1024, 0);
        al[mb] = __builtin_amdgcn_raw_buffer_load_b128(a, va, ((mb * KT + t) * 2 + 1) * 1024, 0);
    }
}
__device__ __forceinline__ void load_x16(const Tile& x, unsigned vx, int t, float (&b)[8]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) b[e] = tile_load(x, 16 * t + e, vx);
}
// NT = 3: Pl Xh + Ph Xl + Ph Xh (small terms first); NT = 4 adds Pl Xl.
template <int MB, int NT>
__device__ __forceinline__ void mfma16_stage(floatx16 (&y)[MB], const u32x4 (&ah)[MB],
                                             const u32x4 (&al)[MB], const float (&b)[8]) {
    bf16x8 xh, xl;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        __bf16 hi, lo;
        split_bf16(b[e], hi, lo);
        xh[e] = hi;
        xl[e] = lo;
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const bf16x8 ph = __builtin_bit_cast(bf16x8, ah[mb]);
        const bf16x8 pl = __builtin_bit_cast(bf16x8, al[mb]);
        if (NT == 4) y[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pl, xl, y[mb], 0, 0, 0);
        y[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pl, xh, y[mb], 0, 0, 0);
        y[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ph, xl, y[mb], 0, 0, 0);
        y[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ph, xh, y[mb], 0, 0, 0);
    }
}

// Split-BF16 stream of NA partials children, KT stages each, operands PF
// stages ahead (as stream_mfma).
template <int MB, int KT, int NA, int NT>
__device__ __forceinline__ void stream_bf16(const Tile& x0, Rsrc a0, const Tile& x1, Rsrc a1,
                                            unsigned vx, unsigned va, floatx16 (&y)[2][MB]) {
    constexpr int PF = 2, T = NA * KT;
#pragma unroll
    for (int k = 0; k < NA; ++k)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) y[k][mb] = floatx16{0};
    u32x4 ah[PF][MB], al[PF][MB];
    float b[PF][8];
#pragma unroll
    for (int j = 0; j < PF && j < T; ++j) {
        load_a16<MB>(j < KT ? a0 : a1, va, j % KT, KT, ah[j], al[j]);
        load_x16(j < KT ? x0 : x1, vx, j % KT, b[j]);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < T; ++j) {
        mfma16_stage<MB, NT>(y[j / KT], ah[j % PF], al[j % PF], b[j % PF]);
        if (j + PF < T) {
            const int t = (j + PF) % KT;
            load_a16<MB>(j + PF < KT ? a0 : a1, va, t, KT, ah[j % PF], al[j % PF]);
            load_x16(j + PF < KT ? x0 : x1, vx, t, b[j % PF]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Column P[., s] of a tip-state child in the accumulator layout.
template <int MB>
__device__ __forceinline__ void state_rows(const float* pk, const int* states, int p, int h,
                                           floatx16 (&y)[MB]) {
    const int s = gin<int>(states)[p];
    const GLOBAL float* row = gin<float>(pk) + s * (MB * 32) + 4 * h;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const floatx4 v = *(const GLOBAL floatx4*)(row + mb * 32 + 8 * q);
            y[mb][4 * q + 0] = v.x;
            y[mb][4 * q + 1] = v.y;
            y[mb][4 * q + 2] = v.z;
            y[mb][4 * q + 3] = v.w;
        }
}

// The MFMA stream of NA partials children (KG stages each, child k into
// y[k]) with matrix operands loaded PF stages ahead.  XU: the partials of
// every stage are loaded up front (the whole tile in flight, one exposed
// latency per tile); otherwise they follow the matrix operands PF stages
// ahead.  Straight-line code, so the wait before each stage covers only
// what that stage reads.
template <int MB, int KG, int NA, bool XU>
__device__ __forceinline__ void stream_mfma(const Tile& x0, Rsrc a0, const Tile& x1, Rsrc a1,
                                            unsigned vx, unsigned va, floatx16 (&y)[2][MB]) {
    constexpr int PF = 3, T = NA * KG;
    constexpr int XB = XU ? T : PF;     // partials buffers
#pragma unroll
    for (int k = 0; k < NA; ++k)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) y[k][mb] = floatx16{0};
    floatx4 am[PF][MB];
    float b[XB][4];
#pragma unroll
    for (int j = 0; j < PF && j < T; ++j) {
        load_a<MB>(j < KG ? a0 : a1, va, j % KG, KG, am[j]);
        if (!XU) load_x(j < KG ? x0 : x1, vx, j % KG, b[j]);
    }
    if (XU) {
#pragma unroll
        for (int j = 0; j < T; ++j) load_x(j < KG ? x0 : x1, vx, j % KG, b[j % XB]);
    }
    // keep the stage order: the scheduler would otherwise regroup the loads
    // and wait for all of them before the first instruction
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < T; ++j) {
        mfma_stage<MB>(y[j / KG], am[j % PF], b[j % XB]);
        if (j + PF < T) {
            const int g = (j + PF) % KG;
            load_a<MB>(j + PF < KG ? a0 : a1, va, g, KG, am[j % PF]);
            if (!XU) load_x(j + PF < KG ? x0 : x1, vx, g, b[j % XB]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// One post-order operation per blockIdx.y (a dependency level), one tile of
// 32 patterns per blockIdx.x, one wave per category.  The partials children
// (0, 1 or 2 of them) form one stream of KG-group stages whose operands are
// loaded PF stages ahead; tip-state children are gathered afterwards.
// Scaling semantics as the generic path (BeagleCPUImpl.hpp:1555-1577,
// 4008-4133).
// Scaling and store of one operation's output tile (post- or pre-order).
struct ScaleOut {
    int mode, logScalers;
    const double* scaleRead;
    double* scaleWrite;
    double* cumulative;
    const double* accA;
    const double* accB;
    const short* e1;
    const short* e2;
    short* ed;
    void* dest;
};

// Semantics of the generic path (BeagleCPUImpl.hpp:1555-1577, 4008-4133):
// SCALE_READ divides by the stored raw factor; SCALE_WRITE (manual) divides
// by the per-pattern maximum over states and categories (0 -> 1), stores it
// (or its log) plus the children's factors, and adds log(max) to the
// cumulative buffer; AUTO multiplies by 2^-e (max = f 2^e) and stores the
// exponent sum.  The maximum over categories goes through LDS (`red`).
template <int MB, bool AUTO>
__device__ __forceinline__ void rescale_store(floatx16 (&v)[MB], const ScaleOut& J, long tbase,
                                              int S, int C, int c, int n, int h, int p, bool in,
                                              float (*red)[32]) {
    if (J.mode == SCALE_READ) {
        const float f = (float)gin<double>(J.scaleRead)[p];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) v[mb][r] = v[mb][r] / f;
    }
    if (AUTO || J.mode == SCALE_WRITE) {
        float mx = 0.f;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) mx = fmaxf(mx, v[mb][r]);
        mx = fmaxf(mx, __shfl_xor(mx, 32));
        if (h == 0) red[c][n] = mx;
        __syncthreads();
        for (int cc = 0; cc < C; ++cc) mx = fmaxf(mx, red[cc][n]);
        if (AUTO) {
            const int e = mx > 0.f ? exponent_of(mx) : 0;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int r = 0; r < 16; ++r) v[mb][r] = pow2(v[mb][r], -e);
            if (c == 0 && h == 0 && in) {
                int t = e;
                if (J.e1) t += gin<short>(J.e1)[p];
                if (J.e2) t += gin<short>(J.e2)[p];
                gout<short>(J.ed)[p] = (short)t;
            }
        } else {
            if (mx == 0.f) mx = 1.f;
            const float inv = 1.f / mx;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) v[mb] *= inv;
            if (c == 0 && h == 0 && in) {
                const double lm = log((double)mx);
                double w = J.logScalers ? lm : (double)mx;
                if (J.cumulative) gout<double>(J.cumulative)[p] += lm;
                if (J.accA) w += gin<double>(J.accA)[p];
                if (J.accB) w += gin<double>(J.accB)[p];
                gout<double>(J.scaleWrite)[p] = w;
            }
        }
    }
    // rows >= S fall outside the tile's resources; a lane past the
    // operation's patterns stores out of range
    const Tile out = tile_of(J.dest, tbase, S);
    const unsigned vo = in ? 4 * n + 512 * h : 0x40000000u;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) tile_store(out, mb * 32 + 8 * (r >> 2) + (r & 3), vo, v[mb][r]);
}

// MODE 0: FP32 matrix cores, partials pipelined; 1: FP32, whole partials
// tile up front; 3, 4: split BF16 with 3 or 4 products.  KX: k groups of 8
// (FP32) or k steps of 16 (BF16).
template <int MB, int KX, bool AUTO, int MODE>
__global__ __launch_bounds__(512) void k_postm(const PostJob* __restrict__ jobs, Dims d,
                                               const float* __restrict__ pack, TileList w) {
    __shared__ float red[8][32];
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = threadIdx.x & 63, n = l & 31, h = l >> 5;
    int job, item;
    tile_item(w, job, item);
    const PostJob& J = jobs[job];
    const int tp0 = ((w.first ? J.p0 / PTILE : 0) + item) * PTILE;
    if (tp0 >= J.p1 || tp0 + PTILE <= J.p0) return;    // uniform over the workgroup
    const int p = tp0 + n;
    const bool in = p >= J.p0 && p < J.p1;
    const int S = d.S;
    const long per = mfma_pack_reals(S);
    const long tbase = pbase(d, c, tp0);
    // the partials children in list order: stream 0 is child 1 unless that
    // child is tip states, stream 1 is child 2 when both are partials
    const Tile x0 = tile_of(J.s1 ? J.x2 : J.x1, tbase, S), x2 = tile_of(J.x2, tbase, S);
    const Rsrc a0 = rsrc(pack + (J.s1 ? J.pk2 : J.pk1) + c * per, -1);
    const Rsrc a2 = rsrc(pack + J.pk2 + c * per, -1);
    floatx16 y[2][MB];
    const int na = (J.s1 ? 0 : 1) + (J.s2 ? 0 : 1);
    if (MODE <= 1) {
        if (na == 2) stream_mfma<MB, KX, 2, MODE == 1>(x0, a0, x2, a2, 4 * l, 16 * l, y);
        else if (na == 1) stream_mfma<MB, KX, 1, MODE == 1>(x0, a0, x2, a2, 4 * l, 16 * l, y);
    } else {
        const unsigned vx = 1024 * h + 4 * n;
        if (na == 2) stream_bf16<MB, KX, 2, MODE>(x0, a0, x2, a2, vx, 16 * l, y);
        else if (na == 1) stream_bf16<MB, KX, 1, MODE>(x0, a0, x2, a2, vx, 16 * l, y);
    }
    // tip-state children take the accumulator slots the stream left free
    if (J.s2) state_rows<MB>(pack + J.pk2 + c * per, J.s2, p, h, y[1]);
    if (J.s1) {
        if (J.s2) state_rows<MB>(pack + J.pk1 + c * per, J.s1, p, h, y[0]);
        else state_rows<MB>(pack + J.pk1 + c * per, J.s1, p, h, y[1]);
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) y[0][mb] *= y[1][mb];
    const ScaleOut so{J.mode, J.logScalers, J.scaleRead, J.scaleWrite, J.cumulative, J.accA,
                      J.accB, J.e1, J.e2, J.ed, J.dest};
    rescale_store<MB, AUTO>(y[0], so, tbase, S, d.C, c, n, h, p, in, red);
}


void launch_post_mfma(const PostJob* djobs, int njobs, const Dims& d, const float* pack,
                      int variant, hipStream_t st, TileList w, int items) {
    if (njobs <= 0 || (w.first && items <= 0)) return;
    const dim3 grid = w.first ? dim3(items) : dim3(d.Ppad / PTILE, njobs), block(64 * d.C);
    const int KG = mfma_kgroups(d.S);
#define POSTM_V(M, K, A)                                                                       \
    if (variant == 1)                                                                          \
        hipLaunchKernelGGL((k_postm<M, K, A, 1>), grid, block, 0, st, djobs, d, pack, w);     \
    else                                                                                       \
        hipLaunchKernelGGL((k_postm<M, K, A, 0>), grid, block, 0, st, djobs, d, pack, w);
#define POSTM(M, K)                                                                            \
    case K:                                                                                    \
        if (d.autoScale) { POSTM_V(M, K, true) } else { POSTM_V(M, K, false) }                 \
        break;
    switch (KG) {
        POSTM(1, 1) POSTM(1, 2) POSTM(1, 3) POSTM(1, 4)
        POSTM(2, 5) POSTM(2, 6) POSTM(2, 7) POSTM(2, 8)
        default: break;
    }
#undef POSTM
#undef POSTM_V
}

void launch_post_bf16(const PostJob* djobs, int njobs, const Dims& d, const float* pack,
                      int terms, hipStream_t st, TileList w, int items) {
    if (njobs <= 0 || (w.first && items <= 0)) return;
    const dim3 grid = w.first ? dim3(items) : dim3(d.Ppad / PTILE, njobs), block(64 * d.C);
    const int KT = mfma_ksteps16(d.S);
#define POSTB_V(M, K, A)                                                                       \
    if (terms == 4)                                                                            \
        hipLaunchKernelGGL((k_postm<M, K, A, 4>), grid, block, 0, st, djobs, d, pack, w);     \
    else                                                                                       \
        hipLaunchKernelGGL((k_postm<M, K, A, 3>), grid, block, 0, st, djobs, d, pack, w);
#define POSTB(M, K)                                                                            \
    case K:                                                                                    \
        if (d.autoScale) { POSTB_V(M, K, true) } else { POSTB_V(M, K, false) }                 \
        break;
    switch (KT) {
        POSTB(1, 1) POSTB(1, 2) POSTB(2, 3) POSTB(2, 4)
        default: break;
    }
#undef POSTB
#undef POSTB_V
}

// ---------------------------------------------------------------------------
// Pre-order (S > 4, FP32): dest = P_self^T (pre_parent * (P_sib x_sib))
// (BeagleCPUImpl.hpp:4530-4662; the reference GPU computes the same sums in
// libhmsbeagle/GPU/kernels/kernelsXDerivatives.cu:1-82).  The first product
// runs as a post-order child stream; its result, multiplied by the parent's
// pre-order partials, stays in the accumulator registers and is the B
// operand of the second product: inner state k of instruction s = 4 g + e is
// the accumulator element the lane already holds (PACK_PRE_SELF orders
// P_self^T the same way), so no data moves between lanes.

// z = M^T u, u in the accumulator layout; A operands loaded PF groups ahead.
template <int MB>
__device__ __forceinline__ void gemm_from_acc(Rsrc a, unsigned va, const floatx16 (&u)[MB],
                                              floatx16 (&z)[MB]) {
    constexpr int NG = MB * 4, PF = 3;
#pragma unroll
    for (int mo = 0; mo < MB; ++mo) z[mo] = floatx16{0};
    floatx4 am[PF][MB];
#pragma unroll
    for (int j = 0; j < PF && j < NG; ++j)
#pragma unroll
        for (int mo = 0; mo < MB; ++mo) am[j][mo] = pack_load4(a, va, (mo * NG + j) * 1024);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int g = 0; g < NG; ++g) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int s4 = 4 * g + e;
#pragma unroll
            for (int mo = 0; mo < MB; ++mo)
                z[mo] = __builtin_amdgcn_mfma_f32_32x32x2f32(am[g % PF][mo][e], u[s4 >> 4][s4 & 15],
                                                            z[mo], 0, 0, 0);
        }
        if (g + PF < NG) {
#pragma unroll
            for (int mo = 0; mo < MB; ++mo)
                am[g % PF][mo] = pack_load4(a, va, (mo * NG + g + PF) * 1024);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// z = M^T u on the BF16 matrix cores (both operands split hi/lo, NT
// products), u in the accumulator layout: instruction s takes the lane's
// own elements 8 (s % 2) ... 8 (s % 2) + 7 of block s / 2 as its B operand
// (PACK_PRE_SELF_BF16 orders M^T the same way); A operands PF instructions
// ahead.
template <int MB, int NT>
__device__ __forceinline__ void gemm_from_acc16(Rsrc a, unsigned va, const floatx16 (&u)[MB],
                                                floatx16 (&z)[MB]) {
    constexpr int NS = 2 * MB, PF = 2;
#pragma unroll
    for (int mo = 0; mo < MB; ++mo) z[mo] = floatx16{0};
    u32x4 ah[PF][MB], al[PF][MB];
#pragma unroll
    for (int j = 0; j < PF && j < NS; ++j) load_a16<MB>(a, va, j, NS, ah[j], al[j]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        float b[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) b[e] = u[s >> 1][8 * (s & 1) + e];
        mfma16_stage<MB, NT>(z, ah[s % PF], al[s % PF], b);
        if (s + PF < NS) load_a16<MB>(a, va, s + PF, NS, ah[s % PF], al[s % PF]);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// values of a partials tile in the accumulator layout (rows >= S read 0);
// lane offset vo = 4 n + 512 h (row 4 h, pattern n)
template <int MB>
__device__ __forceinline__ void load_acc(const Tile& t, unsigned vo, floatx16 (&v)[MB]) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) v[mb][r] = tile_load(t, mb * 32 + 8 * (r >> 2) + (r & 3), vo);
}

__device__ __forceinline__ double half_wave_sum(double v) {
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Derivative tile sums of one branch from its per-category dot products
// s3 = (pre . x, pre . (D1 x), pre . (D2 x)) of the lane's pattern (lane pair
// h = 0, 1 already added): categories combined in FP64 through LDS, r_k =
// N_k / D, tile sums of w r1, w r1^2 and w r2 into rows row * 3 + k of
// `partial` (ceil(P/32) tiles), as the edge-derivative kernel writes them.
__device__ __forceinline__ void branch_sums(const float (&s3)[3], const double* cw,
                                            const double* pw, double* partial, int row, int tile,
                                            const Dims& d, int c, int n, int h, int l, int p,
                                            double (*dred)[3][32]) {
    const double wc = cw[c];
    if (h == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) dred[c][k][n] = wc * (double)s3[k];
    }
    __syncthreads();
    const int ntp = (d.P + PTILE - 1) / PTILE;
    if (c == 0) {
        double D = 0.0, N1 = 0.0, N2 = 0.0;
        for (int cc = 0; cc < d.C; ++cc) {
            D += dred[cc][0][n];
            N1 += dred[cc][1][n];
            N2 += dred[cc][2][n];
        }
        const bool use = p < d.P && h == 0;
        const double r1 = use ? N1 / D : 0.0, r2 = use ? N2 / D : 0.0;
        const double wp = use ? pw[p] : 0.0;
        const double s1 = half_wave_sum(wp * r1), sq = half_wave_sum(wp * r1 * r1);
        const double s2 = half_wave_sum(wp * r2);
        if (l == 0 && tile < ntp) {
            GLOBAL double* pt = gout<double>(partial);
            pt[((long)row * 3 + 0) * ntp + tile] = s1;
            pt[((long)row * 3 + 1) * ntp + tile] = sq;
            pt[((long)row * 3 + 2) * ntp + tile] = s2;
        }
    }
}

// MODE 0: FP32 matrix cores (KX = k groups of 8); 3, 4: split BF16 with 3
// or 4 products (KX = k steps of 16).  DV: the fused gradient's launch,
// every job also forms its branch's derivative sums (pkT >= 0: a tip
// branch, from u; else from pre(v) in registers and v's post-order
// partials, with D1 and D2 packed as P_self is).
template <int MB, int KX, bool AUTO, int MODE, bool DV>
__global__ __launch_bounds__(512, 3) void k_prem(const PreJob* __restrict__ jobs, Dims d,
                                              const float* __restrict__ pack, TileList w,
                                              const double* __restrict__ cw,
                                              const double* __restrict__ pw, double* partial) {
    __shared__ float red[8][32];
    __shared__ double dred[8][3][32];
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = threadIdx.x & 63, n = l & 31, h = l >> 5;
    int job, item;
    tile_item(w, job, item);
    const PreJob& J = jobs[job];
    const int tp0 = ((w.first ? J.p0 / PTILE : 0) + item) * PTILE;
    if (tp0 >= J.p1 || tp0 + PTILE <= J.p0) return;    // uniform over the workgroup
    const int p = tp0 + n;
    const bool in = p >= J.p0 && p < J.p1;
    const int S = d.S;
    const long per = mfma_pack_reals(S);
    const long tbase = pbase(d, c, tp0);
    floatx16 par[MB];
    load_acc<MB>(tile_of(J.parent, tbase, S), 4 * n + 512 * h, par);
    floatx16 y[2][MB];
    const float* psib = pack + J.pkSib + c * per;
    if (J.ss) {
        state_rows<MB>(psib, J.ss, p, h, y[0]);
    } else {
        const Tile xs = tile_of(J.xs, tbase, S);
        const Rsrc as = rsrc(psib, -1);
        if (MODE == 0) stream_mfma<MB, KX, 1, false>(xs, as, xs, as, 4 * l, 16 * l, y);
        else stream_bf16<MB, KX, 1, MODE>(xs, as, xs, as, 1024 * h + 4 * n, 16 * l, y);
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) y[0][mb] *= par[mb];
    if (DV && J.pkT >= 0) {
        // tip branch of the fused gradient, from u = y[0] without forming
        // pre(t) = P_t^T u: with the tip's state s, pre(t)_{s mod S} =
        // u . P_t[., s mod S] and pre(t) . (D_k x) = u . (P_t D_k)[., s]
        // (BeagleCPUImpl.hpp:2086-2119 forms the same ratios from pre(t))
        float s3[3] = {0.f, 0.f, 0.f};
        const int pks[3] = {J.pkT, J.pkM1, J.pkM2};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            floatx16 t[MB];
            state_rows<MB>(pack + pks[k] + c * per, J.sself, p, h, t);
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int r = 0; r < 16; ++r) s3[k] += y[0][mb][r] * t[mb][r];
            s3[k] += __shfl_xor(s3[k], 32);
        }
        branch_sums(s3, cw, pw, partial, J.oidx, tp0 / PTILE, d, c, n, h, l, p, dred);
    }
    if (!J.dest) return;   // no pre-order partials kept (uniform)
    const Rsrc aself = rsrc(pack + J.pkSelf + c * per, -1);
    if (MODE == 0) gemm_from_acc<MB>(aself, 16 * l, y[0], y[1]);
    else gemm_from_acc16<MB, MODE>(aself, 16 * l, y[0], y[1]);
    if (DV && J.pkT < 0) {
        // internal branch: pre(v) . (D_k x) = x . (D_k^T pre(v)) with
        // pre(v) = y[1] (before its rescaling, which the ratios do not see)
        // and x = v's post-order partials in the accumulator layout
        floatx16 x[MB], t[MB];
        load_acc<MB>(tile_of(J.xself, tbase, S), 4 * n + 512 * h, x);
        float s3[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) s3[0] += y[1][mb][r] * x[mb][r];
#pragma unroll
        for (int k = 1; k < 3; ++k) {
            const Rsrc ad = rsrc(pack + (k == 1 ? J.pkM1 : J.pkM2) + c * per, -1);
            if (MODE == 0) gemm_from_acc<MB>(ad, 16 * l, y[1], t);
            else gemm_from_acc16<MB, MODE>(ad, 16 * l, y[1], t);
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int r = 0; r < 16; ++r) s3[k] += t[mb][r] * x[mb][r];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) s3[k] += __shfl_xor(s3[k], 32);
        branch_sums(s3, cw, pw, partial, J.oidx, tp0 / PTILE, d, c, n, h, l, p, dred);
    }
    const ScaleOut so{J.mode, J.logScalers, nullptr, J.scaleWrite, nullptr, J.accA,
                      J.accB, J.e1, J.e2, J.ed, J.dest};
    rescale_store<MB, AUTO>(y[1], so, tbase, S, d.C, c, n, h, p, in, red);
}

void launch_pre_mfma(const PreJob* djobs, int njobs, const Dims& d, const float* pack,
                     int terms, hipStream_t st, TileList w, int items, const double* cw,
                     const double* pw, double* partial) {
    if (njobs <= 0 || (w.first && items <= 0)) return;
    const dim3 grid = w.first ? dim3(items) : dim3(d.Ppad / PTILE, njobs), block(64 * d.C);
#define PREM_V(M, K, A, MODE)                                                                  \
    if (partial)                                                                               \
        hipLaunchKernelGGL((k_prem<M, K, A, MODE, true>), grid, block, 0, st, djobs, d, pack, w, \
                           cw, pw, partial);                                                   \
    else                                                                                       \
        hipLaunchKernelGGL((k_prem<M, K, A, MODE, false>), grid, block, 0, st, djobs, d, pack, \
                           w, cw, pw, partial);
#define PREM(M, K, MODE)                                                                       \
    case K:                                                                                    \
        if (d.autoScale) { PREM_V(M, K, true, MODE) } else { PREM_V(M, K, false, MODE) }       \
        break;
    if (terms == 0) {
        switch (mfma_kgroups(d.S)) {
            PREM(1, 1, 0) PREM(1, 2, 0) PREM(1, 3, 0) PREM(1, 4, 0)
            PREM(2, 5, 0) PREM(2, 6, 0) PREM(2, 7, 0) PREM(2, 8, 0)
            default: break;
        }
    } else if (terms == 4) {
        switch (mfma_ksteps16(d.S)) {
            PREM(1, 1, 4) PREM(1, 2, 4) PREM(2, 3, 4) PREM(2, 4, 4)
            default: break;
        }
    } else {
        switch (mfma_ksteps16(d.S)) {
            PREM(1, 1, 3) PREM(1, 2, 3) PREM(2, 3, 3) PREM(2, 4, 3)
            default: break;
        }
    }
#undef PREM
#undef PREM_V
}

// ---------------------------------------------------------------------------
// Edge derivatives (S > 4, FP32), BeagleCPUImpl.hpp:2086-2119 semantics: per
// node and pattern N_k = sum_c w_c pre . (D_k x), D = sum_c w_c pre . x (a tip:
// (D_k x)_i = D_k[i][state], D = sum_c w_c pre[state % S]); r_k = N_k / D.
// D_k x runs as a child stream; the dot products with pre use the
// accumulator layout, per category in FP32, combined over categories in
// FP64 through LDS.  Tile sums of w r1, w r1^2 (and w r2) per node.


// MODE and KX as k_prem.
template <int MB, int KX, int ND, int MODE>
__global__ __launch_bounds__(512) void k_derivm(const DerivJob* __restrict__ jobs, Dims d,
                                                const float* __restrict__ pack,
                                                const double* __restrict__ cw,
                                                const double* __restrict__ pw, double* sites,
                                                double* __restrict__ partial) {
    __shared__ double red[8][3][32];
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = threadIdx.x & 63, n = l & 31, h = l >> 5;
    const DerivJob& J = jobs[blockIdx.y];
    const int tile = blockIdx.x, ntiles = gridDim.x;
    const int p = tile * PTILE + n;
    const bool in = p < d.P;
    const int S = d.S;
    const long per = mfma_pack_reals(S);
    const long tbase = pbase(d, c, tile * PTILE);
    const float* p1 = pack + J.pk1 + c * per;
    const float* p2 = pack + (ND == 2 ? J.pk2 : J.pk1) + c * per;
    floatx16 y[2][MB], z[MB];
    int st = 0;
    if (J.states) {
        st = gin<int>(J.states)[p];
        state_rows<MB>(p1, J.states, p, h, y[0]);
        if (ND == 2) state_rows<MB>(p2, J.states, p, h, y[1]);
    } else {
        const Tile x = tile_of(J.post, tbase, S);
        if (MODE == 0)
            stream_mfma<MB, KX, ND, false>(x, rsrc(p1, -1), x, rsrc(p2, -1), 4 * l, 16 * l, y);
        else
            stream_bf16<MB, KX, ND, MODE>(x, rsrc(p1, -1), x, rsrc(p2, -1), 1024 * h + 4 * n,
                                          16 * l, y);
    }
    load_acc<MB>(tile_of(J.pre, tbase, S), 4 * n + 512 * h, z);
    float den = 0.f, n1 = 0.f, n2 = 0.f;
    if (J.states) {
        const int sd = st % S;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = mb * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
                den += i == sd ? z[mb][r] : 0.f;
                n1 += z[mb][r] * y[0][mb][r];
                if (ND == 2) n2 += z[mb][r] * y[1][mb][r];
            }
    } else {
        floatx16 x[MB];
        load_acc<MB>(tile_of(J.post, tbase, S), 4 * n + 512 * h, x);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                den += z[mb][r] * x[mb][r];
                n1 += z[mb][r] * y[0][mb][r];
                if (ND == 2) n2 += z[mb][r] * y[1][mb][r];
            }
    }
    den += __shfl_xor(den, 32);
    n1 += __shfl_xor(n1, 32);
    if (ND == 2) n2 += __shfl_xor(n2, 32);
    const double wc = cw[c];
    if (h == 0) {
        red[c][0][n] = wc * (double)den;
        red[c][1][n] = wc * (double)n1;
        red[c][2][n] = ND == 2 ? wc * (double)n2 : 0.0;
    }
    __syncthreads();
    if (c != 0) return;
    double D = 0.0, N1 = 0.0, N2 = 0.0;
    for (int cc = 0; cc < d.C; ++cc) {
        D += red[cc][0][n];
        N1 += red[cc][1][n];
        N2 += red[cc][2][n];
    }
    const bool use = in && h == 0;
    const double r1 = use ? N1 / D : 0.0, r2 = use ? N2 / D : 0.0, w = use ? pw[p] : 0.0;
    if (ND == 1 && sites && use) sites[(long)J.row * d.P + p] = r1;
    constexpr int NR = ND == 2 ? 3 : 2;
    const double s1 = half_wave_sum(w * r1), sq = half_wave_sum(w * r1 * r1);
    const double s2 = ND == 2 ? half_wave_sum(w * r2) : 0.0;
    if (l == 0) {
        GLOBAL double* pt = gout<double>(partial);
        const long row = (long)J.row * NR;
        pt[(row + 0) * ntiles + tile] = s1;
        pt[(row + 1) * ntiles + tile] = sq;
        if (ND == 2) pt[(row + 2) * ntiles + tile] = s2;
    }
}

void launch_deriv_mfma(const DerivJob* djobs, int njobs, int nd, const Dims& d,
                       const float* pack, const double* cw, const double* pw, double* sites,
                       double* partial, int terms, hipStream_t st) {
    if (njobs <= 0) return;
    const dim3 grid((d.P + PTILE - 1) / PTILE, njobs), block(64 * d.C);
#define DERIVM_V(M, K, N, MODE)                                                                \
    hipLaunchKernelGGL((k_derivm<M, K, N, MODE>), grid, block, 0, st, djobs, d, pack, cw, pw,  \
                       sites, partial);
#define DERIVM(M, K, MODE)                                                                     \
    case K:                                                                                    \
        if (nd == 2) { DERIVM_V(M, K, 2, MODE) } else { DERIVM_V(M, K, 1, MODE) }              \
        break;
    if (terms == 0) {
        switch (mfma_kgroups(d.S)) {
            DERIVM(1, 1, 0) DERIVM(1, 2, 0) DERIVM(1, 3, 0) DERIVM(1, 4, 0)
            DERIVM(2, 5, 0) DERIVM(2, 6, 0) DERIVM(2, 7, 0) DERIVM(2, 8, 0)
            default: break;
        }
    } else if (terms == 4) {
        switch (mfma_ksteps16(d.S)) {
            DERIVM(1, 1, 4) DERIVM(1, 2, 4) DERIVM(2, 3, 4) DERIVM(2, 4, 4)
            default: break;
        }
    } else {
        switch (mfma_ksteps16(d.S)) {
            DERIVM(1, 1, 3) DERIVM(1, 2, 3) DERIVM(2, 3, 3) DERIVM(2, 4, 3)
            default: break;
        }
    }
#undef DERIVM
#undef DERIVM_V
}

// ---------------------------------------------------------------------------
// Rate-matrix gradient (S = 5...64, FP32 partials) on the matrix cores: the
// X_c = sum_p (w_c w_p / L_p) top_p x_p^T of one branch (a pre-order
// operation with its own post-order operand attached) over a chunk of
// 32-pattern tiles, one wave per category (BeagleCPUImpl.hpp:2122-2237
// holds the cross products these sums make exact).  Per tile:
//   y   = P_sib x_sib, Px = P x           (one stream of up to two children,
//                                          or a tip's column)
//   top = pre(parent) * y, L_c = sum_i top_i (P x)_i   (accumulator layout;
//                                          the lane pair h = 0, 1 adds)
//   L   = sum_c w_c L_c                   (LDS, double-buffered by tile)
//   X_c += (coef top) x^T                 (v_mfma_f32_32x32x2f32, k = two
//                                          patterns per instruction)
// For the last product the scaled top vectors and the own operand x (its
// partials, or the one-hot of a tip state; a missing state's zero row, as
// its pad column does not depend on Q) are staged in the wave's LDS block
// as [state][pattern]: lane l then reads A[i][k] = T[32 ib + l % 32][k0 + l / 32]
// and B[k][j] = X[32 jb + l % 32][k0 + l / 32].  X_c is summed in FP32 over
// the chunk (every term is non-negative) and written as FP64 to
// out[o][c][chunk][S^2]; the chunks are added in FP64 by k_qgx_eigen.  At
// most 4 categories; two workgroups per CU (256 registers per wave with no
// spills, and the LDS of two workgroups).
constexpr int QGM_LD = PTILE + 1;   // row stride of the staged tiles (floats)

__host__ __device__ constexpr size_t qgm_lds_bytes(int MB, int C) {
    return ((size_t)C * 2 * MB * 32 * QGM_LD + 2 * 4 * 32) * sizeof(float);
}

template <int MB, int KG>
__global__ __launch_bounds__(256, 2) void k_qgm(const PreJob* __restrict__ jobs, Dims d,
                                             const float* __restrict__ pack,
                                             const double* __restrict__ cw,
                                             const double* __restrict__ pw, int tilesPerChunk,
                                             double* __restrict__ out) {
    extern __shared__ float qsm[];
    constexpr int RW = MB * 32;   // staged rows
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = threadIdx.x & 63, n = l & 31, h = l >> 5;
    const int S = d.S, C = d.C;
    const int o = blockIdx.y;
    const PreJob& J = jobs[o];
    float* T = qsm + (size_t)c * 2 * RW * QGM_LD;
    float* X = T + RW * QGM_LD;
    float* red = qsm + (size_t)C * 2 * RW * QGM_LD;   // [2 tiles][4 categories][32]
    const long per = mfma_pack_reals(S);
    const float* psib = pack + J.pkSib + c * per;
    const float* pself = pack + J.pkSelf + c * per;
    const double wc = cw[c];
    floatx16 acc[MB][MB];
#pragma unroll
    for (int ib = 0; ib < MB; ++ib)
#pragma unroll
        for (int jb = 0; jb < MB; ++jb) acc[ib][jb] = floatx16{0};
    const int ntiles = d.Ppad / PTILE;
    const int t0 = blockIdx.x * tilesPerChunk, t1 = min(ntiles, t0 + tilesPerChunk);
    for (int tile = t0; tile < t1; ++tile) {
        const int tp0 = tile * PTILE, p = tp0 + n;
        const long tbase = pbase(d, c, tp0);
        floatx16 ys[MB], yo[MB];
        {
            floatx16 y[2][MB];
            const Tile xs = tile_of(J.xs, tbase, S), xo = tile_of(J.xself, tbase, S);
            const Rsrc as = rsrc(psib, -1), ao = rsrc(pself, -1);
            if (!J.ss && !J.sself) {
                stream_mfma<MB, KG, 2, false>(xs, as, xo, ao, 4 * l, 16 * l, y);
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) {
                    ys[mb] = y[0][mb];
                    yo[mb] = y[1][mb];
                }
            } else if (!J.ss) {
                stream_mfma<MB, KG, 1, false>(xs, as, xs, as, 4 * l, 16 * l, y);
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) ys[mb] = y[0][mb];
                state_rows<MB>(pself, J.sself, p, h, yo);
            } else if (!J.sself) {
                stream_mfma<MB, KG, 1, false>(xo, ao, xo, ao, 4 * l, 16 * l, y);
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) yo[mb] = y[0][mb];
                state_rows<MB>(psib, J.ss, p, h, ys);
            } else {
                state_rows<MB>(psib, J.ss, p, h, ys);
                state_rows<MB>(pself, J.sself, p, h, yo);
            }
        }
        floatx16 par[MB];
        load_acc<MB>(tile_of(J.parent, tbase, S), 4 * n + 512 * h, par);
        float lc = 0.f;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            ys[mb] *= par[mb];   // top
#pragma unroll
            for (int r = 0; r < 16; ++r) lc += ys[mb][r] * yo[mb][r];
        }
        lc += __shfl_xor(lc, 32);
        float* rb = red + (tile & 1) * 4 * 32;
        if (h == 0) rb[c * 32 + n] = lc;
        __syncthreads();
        double L = 0.0;
        for (int cc = 0; cc < C; ++cc) L += cw[cc] * (double)rb[cc * 32 + n];
        const float coef = (p < d.P && L != 0.0) ? (float)(pw[p] * wc / L) : 0.f;
        // stage coef top and x as [state][pattern]; the wave's previous
        // tile's reads of its block ended before this tile's barrier
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                T[(mb * 32 + 8 * (r >> 2) + 4 * h + (r & 3)) * QGM_LD + n] = coef * ys[mb][r];
        if (J.sself) {
            const int s = gin<int>(J.sself)[p];
#pragma unroll
            for (int u = 0; u < RW / 2; ++u) {
                const int j = 2 * u + h;
                X[j * QGM_LD + n] = s == j ? 1.f : 0.f;
            }
        } else {
            const Tile xo = tile_of(J.xself, tbase, S);
#pragma unroll
            for (int u = 0; u < RW / 2; ++u) X[(2 * u + h) * QGM_LD + n] = tile_load(xo, 2 * u, 4 * l);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
#pragma unroll
        for (int kk = 0; kk < PTILE / 2; ++kk) {
            float a[MB], b[MB];
#pragma unroll
            for (int ib = 0; ib < MB; ++ib) {
                a[ib] = T[(ib * 32 + n) * QGM_LD + 2 * kk + h];
                b[ib] = X[(ib * 32 + n) * QGM_LD + 2 * kk + h];
            }
#pragma unroll
            for (int ib = 0; ib < MB; ++ib)
#pragma unroll
                for (int jb = 0; jb < MB; ++jb)
                    acc[ib][jb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[ib], b[jb], acc[ib][jb],
                                                                       0, 0, 0);
        }
        // the block's reads done before the next tile writes it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    }
    double* dst = out + (((long)o * C + c) * gridDim.x + blockIdx.x) * S * S;
#pragma unroll
    for (int ib = 0; ib < MB; ++ib)
#pragma unroll
        for (int jb = 0; jb < MB; ++jb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = ib * 32 + 8 * (r >> 2) + 4 * h + (r & 3), j = jb * 32 + n;
                if (i < S && j < S) dst[i * S + j] = (double)acc[ib][jb][r];
            }
}

size_t qgrad_mfma_lds(const Dims& d) { return qgm_lds_bytes(mfma_mblocks(d.S), d.C); }

void launch_qgrad_mfma(const PreJob* djobs, int nops, const Dims& d, const float* pack,
                       const double* cw, const double* pw, const double* eig,
                       const double* tau, double* xs, int chunks, double* z, hipStream_t st) {
    if (nops <= 0) return;
    const int ntiles = d.Ppad / PTILE;
    const int per = (ntiles + chunks - 1) / chunks;
    const dim3 grid(chunks, nops), block(64 * d.C);
    const size_t lds = qgrad_mfma_lds(d);
    const int KG = mfma_kgroups(d.S);
#define QGM(M, K)                                                                              \
    case K:                                                                                    \
        hipLaunchKernelGGL((k_qgm<M, K>), grid, block, lds, st, djobs, d, pack, cw, pw, per, xs); \
        break;
    switch (KG) {
        QGM(1, 1) QGM(1, 2) QGM(1, 3) QGM(1, 4)
        QGM(2, 5) QGM(2, 6) QGM(2, 7) QGM(2, 8)
        default: break;
    }
#undef QGM
    launch_qgx_eigen(xs, chunks, nops, d.S, d.C, eig, tau, z, st);
}

// ---------------------------------------------------------------------------
// FP64: v_mfma_f64_16x16x4f64.  Lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; element r of its accumulator is D[4 r + l / 16][l % 16]
// (rows interleave across the four lane groups).  A wavefront owns one
// category and 16 patterns; M blocks of 16 output states; a stage is 8
// inner states, k = 8 g + 4 q + l / 16 for instruction q of group g.

typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef double doublex2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(WG) void k_pack_mats64(const PackJob* __restrict__ jobs, int njobs,
                                                    int S, int C,
                                                    const double* __restrict__ mats,
                                                    double* __restrict__ pack) {
    const int per = mfma_pack_doubles(S);
    const long t = (long)blockIdx.x * WG + threadIdx.x;
    if (t >= (long)njobs * C * per) return;
    const int j = (int)(t / ((long)C * per));
    const int rem = (int)(t - (long)j * C * per);
    const int c = rem / per, e = rem - c * per;
    const PackJob J = jobs[j];
    const double* M = mats + J.src + (long)c * S * (S + 1);
    const int MB = mfma_mblocks16(S), KG = mfma_kgroups(S);
    double v = 0.0;
    const bool tin = (J.kind & PACK_TRANSPOSED_INPUT) != 0;
    const int kind = J.kind & ~PACK_TRANSPOSED_INPUT;
    auto at = [&](int i, int k) { return tin ? M[k * (S + 1) + i] : M[i * (S + 1) + k]; };
    if (kind == PACK_OPERAND) {
        // [mb][group g][lane][q]: row 16 mb + lane % 16, column 8 g + 4 q + lane / 16
        if (e < MB * KG * 128) {
            const int mb = e / (KG * 128), r = e - mb * KG * 128;
            const int g = r / 128, lane = (r & 127) >> 1, q = r & 1;
            const int i = mb * 16 + (lane & 15), k = 8 * g + 4 * q + (lane >> 4);
            if (i < S && k < S) v = at(i, k);
        }
    } else if (kind == PACK_PRE_SELF) {
        // M^T u with u in the accumulator layout: [mo][group g][lane][q];
        // step s = 2 g + q holds element r = s % 4 of block s / 4, i.e.
        // inner state 16 (s / 4) + 4 r + lane / 16; output row j = 16 mo + lane % 16
        const int NG = MB * 2;
        if (e < MB * NG * 128) {
            const int mo = e / (NG * 128), r0 = e - mo * NG * 128;
            const int g = r0 / 128, lane = (r0 & 127) >> 1, q = r0 & 1;
            const int s4 = 2 * g + q, mb = s4 >> 2, r = s4 & 3;
            const int k = mb * 16 + 4 * r + (lane >> 4), j = mo * 16 + (lane & 15);
            if (j < S && k < S) v = at(k, j);
        }
    } else {
        // [state s (S: the pad column)][mb][lane group kq][r]: row 16 mb + 4 r + kq
        const int W = MB * 16;
        if (e < (S + 1) * W) {
            const int s = e / W, w = e - s * W;
            const int mb = w >> 4, kq = (w & 15) >> 2, r = w & 3;
            const int i = mb * 16 + 4 * r + kq;
            if (i < S) v = tin ? (s < S ? M[s * (S + 1) + i] : 0.0) : M[i * (S + 1) + s];
        }
    }
    pack[J.dst / 2 + (long)c * per + e] = v;
}

void launch_pack_matrices64(const PackJob* jobs, int njobs, int S, int C, const void* mats,
                            float* pack, hipStream_t st) {
    const long n = (long)njobs * C * mfma_pack_doubles(S);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pack_mats64, dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, st, jobs,
                       njobs, S, C, reinterpret_cast<const double*>(mats),
                       reinterpret_cast<double*>(pack));
}

template <int MB>
__device__ __forceinline__ void load_a64(const GLOBAL doublex2* a, int g, int KG,
                                         doublex2 (&am)[MB]) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) am[mb] = a[(mb * KG + g) * 64];
}
__device__ __forceinline__ void load_x64(const GLOBAL double* xb, int g, int kq, int S,
                                         double (&b)[2]) {
#pragma unroll
    for (int q = 0; q < 2; ++q) b[q] = xb[min(8 * g + 4 * q + kq, S - 1) * PTILE];
}
template <int MB>
__device__ __forceinline__ void mfma64_stage(doublex4 (&y)[MB], const doublex2 (&am)[MB],
                                             const double (&b)[2]) {
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
            y[mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[mb][q], b[q], y[mb], 0, 0, 0);
}

// FP64 stream of NA partials children (KG stages each, child k into y[k]),
// operands loaded PF stages ahead; straight-line code.
template <int MB, int KG, int NA>
__device__ __forceinline__ void stream_mfma64(const GLOBAL double* x0, const GLOBAL doublex2* a0,
                                              const GLOBAL double* x1, const GLOBAL doublex2* a1,
                                              int kq, int S, doublex4 (&y)[2][MB]) {
    constexpr int PF = 3, T = NA * KG;
#pragma unroll
    for (int k = 0; k < NA; ++k)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) y[k][mb] = doublex4{0.0, 0.0, 0.0, 0.0};
    doublex2 am[PF][MB];
    double b[PF][2];
#pragma unroll
    for (int j = 0; j < PF && j < T; ++j) {
        load_a64<MB>(j < KG ? a0 : a1, j % KG, KG, am[j]);
        load_x64(j < KG ? x0 : x1, j % KG, kq, S, b[j]);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < T; ++j) {
        mfma64_stage<MB>(y[j / KG], am[j % PF], b[j % PF]);
        if (j + PF < T) {
            const int g = (j + PF) % KG;
            load_a64<MB>(j + PF < KG ? a0 : a1, g, KG, am[j % PF]);
            load_x64(j + PF < KG ? x0 : x1, g, kq, S, b[j % PF]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Column P[., s] of a tip-state child: the lane's four rows 16 mb + 4 r + kq
// are contiguous in the packed rows.
template <int MB>
__device__ __forceinline__ void state_rows64(const double* pk, const int* states, int p, int kq,
                                             doublex4 (&y)[MB]) {
    const int s = gin<int>(states)[p];
    const GLOBAL double* row = gin<double>(pk) + s * (MB * 16) + 4 * kq;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const doublex2 u = *(const GLOBAL doublex2*)(row + mb * 16);
        const doublex2 w = *(const GLOBAL doublex2*)(row + mb * 16 + 2);
        y[mb] = doublex4{u.x, u.y, w.x, w.y};
    }
}

// Scaling and store of an FP64 output tile (accumulator layout of
// v_mfma_f64_16x16x4f64); semantics as rescale_store.
template <int MB, bool AUTO>
__device__ __forceinline__ void rescale_store64(doublex4 (&v)[MB], const ScaleOut& J, long tbase,
                                                int S, int C, int c, int n, int kq, int p,
                                                bool in, double (*red)[16]) {
    if (J.mode == SCALE_READ) {
        const double f = gin<double>(J.scaleRead)[p];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) v[mb][r] = v[mb][r] / f;
    }
    if (AUTO || J.mode == SCALE_WRITE) {
        double mx = 0.0;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) mx = fmax(mx, v[mb][r]);
        mx = fmax(mx, __shfl_xor(mx, 16));
        mx = fmax(mx, __shfl_xor(mx, 32));
        if (kq == 0) red[c][n] = mx;
        __syncthreads();
        for (int cc = 0; cc < C; ++cc) mx = fmax(mx, red[cc][n]);
        if (AUTO) {
            const int e = mx > 0.0 ? exponent_of(mx) : 0;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int r = 0; r < 4; ++r) v[mb][r] = pow2(v[mb][r], -e);
            if (c == 0 && kq == 0 && in) {
                int t = e;
                if (J.e1) t += gin<short>(J.e1)[p];
                if (J.e2) t += gin<short>(J.e2)[p];
                gout<short>(J.ed)[p] = (short)t;
            }
        } else {
            if (mx == 0.0) mx = 1.0;
            const double inv = 1.0 / mx;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) v[mb] *= inv;
            if (c == 0 && kq == 0 && in) {
                const double lm = log(mx);
                double w = J.logScalers ? lm : mx;
                if (J.cumulative) gout<double>(J.cumulative)[p] += lm;
                if (J.accA) w += gin<double>(J.accA)[p];
                if (J.accB) w += gin<double>(J.accB)[p];
                gout<double>(J.scaleWrite)[p] = w;
            }
        }
    }
    GLOBAL double* out = gout<double>(J.dest) + tbase + n;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = mb * 16 + 4 * r + kq;
            if (i < S && in) out[i * PTILE] = v[mb][r];
        }
}

// FP64 post-order: one operation per blockIdx.y, 16 patterns per
// blockIdx.x, one wave per category; scaling as k_postm.
template <int MB, int KG, bool AUTO>
__global__ __launch_bounds__(512) void k_postd(const PostJob* __restrict__ jobs, Dims d,
                                               const float* __restrict__ pack) {
    __shared__ double red[8][16];
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = threadIdx.x & 63, n = l & 15, kq = l >> 4;
    const PostJob& J = jobs[blockIdx.y];
    const int tp0 = blockIdx.x * 16;
    if (tp0 >= J.p1 || tp0 + 16 <= J.p0) return;    // uniform over the workgroup
    const int p = tp0 + n;
    const bool in = p >= J.p0 && p < J.p1;
    const int S = d.S;
    const long per = mfma_pack_doubles(S);
    const long tbase = pbase(d, c, tp0);
    const GLOBAL double* x1 = gin<double>(J.x1) + tbase + n;
    const GLOBAL double* x2 = gin<double>(J.x2) + tbase + n;
    const double* p1 = reinterpret_cast<const double*>(pack + J.pk1) + c * per;
    const double* p2 = reinterpret_cast<const double*>(pack + J.pk2) + c * per;
    const GLOBAL doublex2* a1 = (const GLOBAL doublex2*)gin<double>(p1) + l;
    const GLOBAL doublex2* a2 = (const GLOBAL doublex2*)gin<double>(p2) + l;
    doublex4 y[2][MB];
    const int na = (J.s1 ? 0 : 1) + (J.s2 ? 0 : 1);
    if (na == 2) stream_mfma64<MB, KG, 2>(x1, a1, x2, a2, kq, S, y);
    else if (na == 1)
        stream_mfma64<MB, KG, 1>(J.s1 ? x2 : x1, J.s1 ? a2 : a1, x2, a2, kq, S, y);
    if (J.s2) state_rows64<MB>(p2, J.s2, p, kq, y[1]);
    if (J.s1) {
        if (J.s2) state_rows64<MB>(p1, J.s1, p, kq, y[0]);
        else state_rows64<MB>(p1, J.s1, p, kq, y[1]);
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) y[0][mb] *= y[1][mb];
    const ScaleOut so{J.mode, J.logScalers, J.scaleRead, J.scaleWrite, J.cumulative, J.accA,
                      J.accB, J.e1, J.e2, J.ed, J.dest};
    rescale_store64<MB, AUTO>(y[0], so, tbase, S, d.C, c, n, kq, p, in, red);
}

void launch_post_mfma64(const PostJob* djobs, int njobs, const Dims& d, const float* pack,
                        hipStream_t st) {
    if (njobs <= 0) return;
    const dim3 grid(d.Ppad / 16, njobs), block(64 * d.C);
    const int KG = mfma_kgroups(d.S);
#define POSTD(M, K)                                                                            \
    case K:                                                                                    \
        if (d.autoScale)                                                                       \
            hipLaunchKernelGGL((k_postd<M, K, true>), grid, block, 0, st, djobs, d, pack);    \
        else                                                                                   \
            hipLaunchKernelGGL((k_postd<M, K, false>), grid, block, 0, st, djobs, d, pack);   \
        break;
    switch (KG) {
        POSTD(1, 1) POSTD(1, 2) POSTD(2, 3) POSTD(2, 4)
        POSTD(3, 5) POSTD(3, 6) POSTD(4, 7) POSTD(4, 8)
        default: break;
    }
#undef POSTD
}

// FP64 pre-order (as k_prem): z = M^T u with u in the accumulator layout;
// instruction s = 2 g + q uses inner states 16 (s / 4) + 4 (s % 4) + l / 16,
// the rows the lane's element s % 4 of block s / 4 already holds.
template <int MB>
__device__ __forceinline__ void gemm_from_acc64(const double* pk, int l, const doublex4 (&u)[MB],
                                                doublex4 (&z)[MB]) {
    constexpr int NG = MB * 2, PF = 3;
    const GLOBAL doublex2* a = (const GLOBAL doublex2*)gin<double>(pk) + l;
#pragma unroll
    for (int mo = 0; mo < MB; ++mo) z[mo] = doublex4{0.0, 0.0, 0.0, 0.0};
    doublex2 am[PF][MB];
#pragma unroll
    for (int j = 0; j < PF && j < NG; ++j)
#pragma unroll
        for (int mo = 0; mo < MB; ++mo) am[j][mo] = a[(mo * NG + j) * 64];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int g = 0; g < NG; ++g) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int s4 = 2 * g + q;
#pragma unroll
            for (int mo = 0; mo < MB; ++mo)
                z[mo] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[g % PF][mo][q], u[s4 >> 2][s4 & 3],
                                                            z[mo], 0, 0, 0);
        }
        if (g + PF < NG) {
#pragma unroll
            for (int mo = 0; mo < MB; ++mo) am[g % PF][mo] = a[(mo * NG + g + PF) * 64];
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

template <int MB>
__device__ __forceinline__ void load_acc64(const void* x, long tbase, int n, int kq, int S,
                                           doublex4 (&v)[MB]) {
    const GLOBAL double* xb = gin<double>(x) + tbase + n;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = mb * 16 + 4 * r + kq;
            v[mb][r] = i < S ? xb[i * PTILE] : 0.0;
        }
}

template <int MB, int KG, bool AUTO>
__global__ __launch_bounds__(512) void k_prem64(const PreJob* __restrict__ jobs, Dims d,
                                                const float* __restrict__ pack) {
    __shared__ double red[8][16];
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = threadIdx.x & 63, n = l & 15, kq = l >> 4;
    const PreJob& J = jobs[blockIdx.y];
    const int tp0 = blockIdx.x * 16;
    if (tp0 >= J.p1 || tp0 + 16 <= J.p0) return;    // uniform over the workgroup
    const int p = tp0 + n;
    const bool in = p >= J.p0 && p < J.p1;
    const int S = d.S;
    const long per = mfma_pack_doubles(S);
    const long tbase = pbase(d, c, tp0);
    doublex4 par[MB];
    load_acc64<MB>(J.parent, tbase, n, kq, S, par);
    doublex4 y[2][MB];
    const double* psib = reinterpret_cast<const double*>(pack + J.pkSib) + c * per;
    if (J.ss) {
        state_rows64<MB>(psib, J.ss, p, kq, y[0]);
    } else {
        const GLOBAL double* xs = gin<double>(J.xs) + tbase + n;
        const GLOBAL doublex2* as = (const GLOBAL doublex2*)gin<double>(psib) + l;
        stream_mfma64<MB, KG, 1>(xs, as, xs, as, kq, S, y);
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) y[0][mb] *= par[mb];
    gemm_from_acc64<MB>(reinterpret_cast<const double*>(pack + J.pkSelf) + c * per, l, y[0],
                        y[1]);
    const ScaleOut so{J.mode, J.logScalers, nullptr, J.scaleWrite, nullptr, J.accA,
                      J.accB, J.e1, J.e2, J.ed, J.dest};
    rescale_store64<MB, AUTO>(y[1], so, tbase, S, d.C, c, n, kq, p, in, red);
}

void launch_pre_mfma64(const PreJob* djobs, int njobs, const Dims& d, const float* pack,
                       hipStream_t st) {
    if (njobs <= 0) return;
    const dim3 grid(d.Ppad / 16, njobs), block(64 * d.C);
    const int KG = mfma_kgroups(d.S);
#define PREM64(M, K)                                                                           \
    case K:                                                                                    \
        if (d.autoScale)                                                                       \
            hipLaunchKernelGGL((k_prem64<M, K, true>), grid, block, 0, st, djobs, d, pack);   \
        else                                                                                   \
            hipLaunchKernelGGL((k_prem64<M, K, false>), grid, block, 0, st, djobs, d, pack);  \
        break;
    switch (KG) {
        PREM64(1, 1) PREM64(1, 2) PREM64(2, 3) PREM64(2, 4)
        PREM64(3, 5) PREM64(3, 6) PREM64(4, 7) PREM64(4, 8)
        default: break;
    }
#undef PREM64
}

// FP64 edge derivatives (as k_derivm; 16-pattern tiles, FP64 throughout)
__device__ __forceinline__ double quarter_wave_sum(double v) {
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <int MB, int KG, int ND>
__global__ __launch_bounds__(512) void k_derivm64(const DerivJob* __restrict__ jobs, Dims d,
                                                  const float* __restrict__ pack,
                                                  const double* __restrict__ cw,
                                                  const double* __restrict__ pw, double* sites,
                                                  double* __restrict__ partial) {
    __shared__ double red[8][3][16];
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = threadIdx.x & 63, n = l & 15, kq = l >> 4;
    const DerivJob& J = jobs[blockIdx.y];
    const int tile = blockIdx.x, ntiles = gridDim.x;
    const int p = tile * 16 + n;
    const bool in = p < d.P;
    const int S = d.S;
    const long per = mfma_pack_doubles(S);
    const long tbase = pbase(d, c, tile * 16);
    const double* p1 = reinterpret_cast<const double*>(pack + J.pk1) + c * per;
    const double* p2 = reinterpret_cast<const double*>(pack + (ND == 2 ? J.pk2 : J.pk1)) + c * per;
    doublex4 y[2][MB], z[MB];
    int st = 0;
    if (J.states) {
        st = gin<int>(J.states)[p];
        state_rows64<MB>(p1, J.states, p, kq, y[0]);
        if (ND == 2) state_rows64<MB>(p2, J.states, p, kq, y[1]);
    } else {
        const GLOBAL double* x = gin<double>(J.post) + tbase + n;
        stream_mfma64<MB, KG, ND>(x, (const GLOBAL doublex2*)gin<double>(p1) + l, x,
                                  (const GLOBAL doublex2*)gin<double>(p2) + l, kq, S, y);
    }
    load_acc64<MB>(J.pre, tbase, n, kq, S, z);
    double den = 0.0, n1 = 0.0, n2 = 0.0;
    if (J.states) {
        const int sd = st % S;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = mb * 16 + 4 * r + kq;
                den += i == sd ? z[mb][r] : 0.0;
                n1 += z[mb][r] * y[0][mb][r];
                if (ND == 2) n2 += z[mb][r] * y[1][mb][r];
            }
    } else {
        doublex4 x[MB];
        load_acc64<MB>(J.post, tbase, n, kq, S, x);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                den += z[mb][r] * x[mb][r];
                n1 += z[mb][r] * y[0][mb][r];
                if (ND == 2) n2 += z[mb][r] * y[1][mb][r];
            }
    }
#pragma unroll
    for (int o = 16; o <= 32; o <<= 1) {
        den += __shfl_xor(den, o);
        n1 += __shfl_xor(n1, o);
        if (ND == 2) n2 += __shfl_xor(n2, o);
    }
    const double wc = cw[c];
    if (kq == 0) {
        red[c][0][n] = wc * den;
        red[c][1][n] = wc * n1;
        red[c][2][n] = ND == 2 ? wc * n2 : 0.0;
    }
    __syncthreads();
    if (c != 0) return;
    double D = 0.0, N1 = 0.0, N2 = 0.0;
    for (int cc = 0; cc < d.C; ++cc) {
        D += red[cc][0][n];
        N1 += red[cc][1][n];
        N2 += red[cc][2][n];
    }
    const bool use = in && kq == 0;
    const double r1 = use ? N1 / D : 0.0, r2 = use ? N2 / D : 0.0, w = use ? pw[p] : 0.0;
    if (ND == 1 && sites && use) sites[(long)J.row * d.P + p] = r1;
    constexpr int NR = ND == 2 ? 3 : 2;
    const double s1 = quarter_wave_sum(w * r1), sq = quarter_wave_sum(w * r1 * r1);
    const double s2 = ND == 2 ? quarter_wave_sum(w * r2) : 0.0;
    if (l == 0) {
        GLOBAL double* pt = gout<double>(partial);
        const long row = (long)J.row * NR;
        pt[(row + 0) * ntiles + tile] = s1;
        pt[(row + 1) * ntiles + tile] = sq;
        if (ND == 2) pt[(row + 2) * ntiles + tile] = s2;
    }
}

void launch_deriv_mfma64(const DerivJob* djobs, int njobs, int nd, const Dims& d,
                         const float* pack, const double* cw, const double* pw, double* sites,
                         double* partial, hipStream_t st) {
    if (njobs <= 0) return;
    const dim3 grid((d.P + 15) / 16, njobs), block(64 * d.C);
    const int KG = mfma_kgroups(d.S);
#define DERIVM64(M, K)                                                                         \
    case K:                                                                                    \
        if (nd == 2)                                                                           \
            hipLaunchKernelGGL((k_derivm64<M, K, 2>), grid, block, 0, st, djobs, d, pack, cw,  \
                               pw, sites, partial);                                            \
        else                                                                                   \
            hipLaunchKernelGGL((k_derivm64<M, K, 1>), grid, block, 0, st, djobs, d, pack, cw,  \
                               pw, sites, partial);                                            \
        break;
    switch (KG) {
        DERIVM64(1, 1) DERIVM64(1, 2) DERIVM64(2, 3) DERIVM64(2, 4)
        DERIVM64(3, 5) DERIVM64(3, 6) DERIVM64(4, 7) DERIVM64(4, 8)
        default: break;
    }
#undef DERIVM64
}

template void launch_pack_tip_products<float>(const PackJob*, int, int, int, const void*, float*,
                                              hipStream_t);
template void launch_pack_matrices<float>(const PackJob*, int, int, int, const void*, float*,
                                          hipStream_t);

}  // namespace hipbeagle
