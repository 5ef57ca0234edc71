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
// (states >= S), so a partials read past the last state is clamped to the
// last state and multiplied by zero.  A tip-state child instead reads its
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
    if (J.kind == PACK_OPERAND) {
        // [mb][k group g][lane][q]: row 32 mb + lane % 32, column 8 g + 2 q + lane / 32
        if (e < MB * KG * 256) {
            const int mb = e / (KG * 256), r = e - mb * KG * 256;
            const int g = r / 256, lane = (r & 255) >> 2, q = r & 3;
            const int i = mb * 32 + (lane & 31), k = 8 * g + 2 * q + (lane >> 5);
            if (i < S && k < S) v = (float)M[i * (S + 1) + k];
        }
    } else if (J.kind == PACK_OPERAND_BF16) {
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
                const float x = (i < S && k < S) ? (float)M[i * (S + 1) + k] : 0.f;
                __bf16 hi, lo;
                split_bf16(x, hi, lo);
                half[u] = __builtin_bit_cast(unsigned short, part ? lo : hi);
            }
            v = __builtin_bit_cast(float, (unsigned)half[0] | ((unsigned)half[1] << 16));
        }
    } else {
        // [state s (S: the pad column)][row i]
        const int W = MB * 32;
        if (e < (S + 1) * W) {
            const int s = e / W, i = e - s * W;
            if (i < S) v = (float)M[i * (S + 1) + s];
        }
    }
    pack[J.dst + (long)c * per + e] = v;
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
// matrix (A) and the partials (B).
template <int MB>
__device__ __forceinline__ void load_a(const GLOBAL floatx4* a, int g, int KG,
                                       floatx4 (&am)[MB]) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) am[mb] = a[(mb * KG + g) * 64];
}
__device__ __forceinline__ void load_x(const GLOBAL float* xb, int g, int h, int S,
                                       float (&b)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) b[q] = xb[min(8 * g + 2 * q + h, S - 1) * PTILE];
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
template <int MB>
__device__ __forceinline__ void load_a16(const GLOBAL u32x4* a, int t, int KT, u32x4 (&ah)[MB],
                                         u32x4 (&al)[MB]) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        ah[mb] = a[((mb * KT + t) * 2 + 0) * 64];
        al[mb] = a[((mb * KT + t) * 2 + 1) * 64];
    }
}
__device__ __forceinline__ void load_x16(const GLOBAL float* xb, int t, int h, int S,
                                         float (&b)[8]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) b[e] = xb[min(16 * t + 8 * h + e, S - 1) * PTILE];
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
__device__ __forceinline__ void stream_bf16(const GLOBAL float* x0, const GLOBAL u32x4* a0,
                                            const GLOBAL float* x1, const GLOBAL u32x4* a1,
                                            int h, int S, floatx16 (&y)[2][MB]) {
    constexpr int PF = 2, T = NA * KT;
#pragma unroll
    for (int k = 0; k < NA; ++k)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) y[k][mb] = floatx16{0};
    u32x4 ah[PF][MB], al[PF][MB];
    float b[PF][8];
#pragma unroll
    for (int j = 0; j < PF && j < T; ++j) {
        load_a16<MB>(j < KT ? a0 : a1, j % KT, KT, ah[j], al[j]);
        load_x16(j < KT ? x0 : x1, j % KT, h, S, b[j]);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < T; ++j) {
        mfma16_stage<MB, NT>(y[j / KT], ah[j % PF], al[j % PF], b[j % PF]);
        if (j + PF < T) {
            const int t = (j + PF) % KT;
            load_a16<MB>(j + PF < KT ? a0 : a1, t, KT, ah[j % PF], al[j % PF]);
            load_x16(j + PF < KT ? x0 : x1, t, h, S, b[j % PF]);
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
__device__ __forceinline__ void stream_mfma(const GLOBAL float* x0, const GLOBAL floatx4* a0,
                                            const GLOBAL float* x1, const GLOBAL floatx4* a1,
                                            int h, int S, floatx16 (&y)[2][MB]) {
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
        load_a<MB>(j < KG ? a0 : a1, j % KG, KG, am[j]);
        if (!XU) load_x(j < KG ? x0 : x1, j % KG, h, S, b[j]);
    }
    if (XU) {
#pragma unroll
        for (int j = 0; j < T; ++j) load_x(j < KG ? x0 : x1, j % KG, h, S, b[j % XB]);
    }
    // keep the stage order: the scheduler would otherwise regroup the loads
    // and wait for all of them before the first instruction
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < T; ++j) {
        mfma_stage<MB>(y[j / KG], am[j % PF], b[j % XB]);
        if (j + PF < T) {
            const int g = (j + PF) % KG;
            load_a<MB>(j + PF < KG ? a0 : a1, g, KG, am[j % PF]);
            if (!XU) load_x(j + PF < KG ? x0 : x1, g, h, S, b[j % XB]);
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
// MODE 0: FP32 matrix cores, partials pipelined; 1: FP32, whole partials
// tile up front; 3, 4: split BF16 with 3 or 4 products.  KX: k groups of 8
// (FP32) or k steps of 16 (BF16).
template <int MB, int KX, bool AUTO, int MODE>
__global__ __launch_bounds__(512) void k_postm(const PostJob* __restrict__ jobs, Dims d,
                                               const float* __restrict__ pack) {
    __shared__ float red[8][32];
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int l = threadIdx.x & 63, n = l & 31, h = l >> 5;
    const PostJob& J = jobs[blockIdx.y];
    const int tp0 = blockIdx.x * PTILE;
    if (tp0 >= J.p1 || tp0 + PTILE <= J.p0) return;    // uniform over the workgroup
    const int p = tp0 + n;
    const bool in = p >= J.p0 && p < J.p1;
    const int S = d.S;
    const long per = mfma_pack_reals(S);
    const long tbase = pbase(d, c, tp0);
    // the partials children in list order: stream 0 is child 1 unless that
    // child is tip states, stream 1 is child 2 when both are partials
    const GLOBAL float* x1 = gin<float>(J.x1) + tbase + n;
    const GLOBAL float* x2 = gin<float>(J.x2) + tbase + n;
    const GLOBAL float* p1 = gin<float>(pack + J.pk1 + c * per);
    const GLOBAL float* p2 = gin<float>(pack + J.pk2 + c * per);
    const GLOBAL float* xs0 = J.s1 ? x2 : x1;
    const GLOBAL float* ps0 = J.s1 ? p2 : p1;
    floatx16 y[2][MB];
    const int na = (J.s1 ? 0 : 1) + (J.s2 ? 0 : 1);
    if (MODE <= 1) {
        const GLOBAL floatx4* a0 = (const GLOBAL floatx4*)ps0 + l;
        const GLOBAL floatx4* a2 = (const GLOBAL floatx4*)p2 + l;
        if (na == 2) stream_mfma<MB, KX, 2, MODE == 1>(xs0, a0, x2, a2, h, S, y);
        else if (na == 1) stream_mfma<MB, KX, 1, MODE == 1>(xs0, a0, x2, a2, h, S, y);
    } else {
        const GLOBAL u32x4* a0 = (const GLOBAL u32x4*)ps0 + l;
        const GLOBAL u32x4* a2 = (const GLOBAL u32x4*)p2 + l;
        if (na == 2) stream_bf16<MB, KX, 2, MODE>(xs0, a0, x2, a2, h, S, y);
        else if (na == 1) stream_bf16<MB, KX, 1, MODE>(xs0, a0, x2, a2, h, S, y);
    }
    // tip-state children take the accumulator slots the stream left free
    if (J.s2) state_rows<MB>(pack + J.pk2 + c * per, J.s2, p, h, y[1]);
    if (J.s1) {
        if (J.s2) state_rows<MB>(pack + J.pk1 + c * per, J.s1, p, h, y[0]);
        else state_rows<MB>(pack + J.pk1 + c * per, J.s1, p, h, y[1]);
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) y[0][mb] *= y[1][mb];
    floatx16 (&v)[MB] = y[0];
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
        for (int cc = 0; cc < d.C; ++cc) mx = fmaxf(mx, red[cc][n]);
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
    GLOBAL float* out = gout<float>(J.dest) + tbase + n;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = mb * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
            if (i < S && in) out[i * PTILE] = v[mb][r];
        }
}

void launch_post_mfma(const PostJob* djobs, int njobs, const Dims& d, const float* pack,
                      int variant, hipStream_t st) {
    if (njobs <= 0) return;
    const dim3 grid(d.Ppad / PTILE, njobs), block(64 * d.C);
    const int KG = mfma_kgroups(d.S);
#define POSTM_V(M, K, A)                                                                       \
    if (variant == 1)                                                                          \
        hipLaunchKernelGGL((k_postm<M, K, A, 1>), grid, block, 0, st, djobs, d, pack);        \
    else                                                                                       \
        hipLaunchKernelGGL((k_postm<M, K, A, 0>), grid, block, 0, st, djobs, d, pack);
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
                      int terms, hipStream_t st) {
    if (njobs <= 0) return;
    const dim3 grid(d.Ppad / PTILE, njobs), block(64 * d.C);
    const int KT = mfma_ksteps16(d.S);
#define POSTB_V(M, K, A)                                                                       \
    if (terms == 4)                                                                            \
        hipLaunchKernelGGL((k_postm<M, K, A, 4>), grid, block, 0, st, djobs, d, pack);        \
    else                                                                                       \
        hipLaunchKernelGGL((k_postm<M, K, A, 3>), grid, block, 0, st, djobs, d, pack);
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
    if (J.kind == PACK_OPERAND) {
        // [mb][group g][lane][q]: row 16 mb + lane % 16, column 8 g + 4 q + lane / 16
        if (e < MB * KG * 128) {
            const int mb = e / (KG * 128), r = e - mb * KG * 128;
            const int g = r / 128, lane = (r & 127) >> 1, q = r & 1;
            const int i = mb * 16 + (lane & 15), k = 8 * g + 4 * q + (lane >> 4);
            if (i < S && k < S) v = M[i * (S + 1) + k];
        }
    } else {
        // [state s (S: the pad column)][mb][lane group kq][r]: row 16 mb + 4 r + kq
        const int W = MB * 16;
        if (e < (S + 1) * W) {
            const int s = e / W, w = e - s * W;
            const int mb = w >> 4, kq = (w & 15) >> 2, r = w & 3;
            const int i = mb * 16 + 4 * r + kq;
            if (i < S) v = M[i * (S + 1) + s];
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
    doublex4 (&v)[MB] = y[0];
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
        for (int cc = 0; cc < d.C; ++cc) mx = fmax(mx, red[cc][n]);
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

template void launch_pack_matrices<float>(const PackJob*, int, int, int, const void*, float*,
                                          hipStream_t);

}  // namespace hipbeagle
