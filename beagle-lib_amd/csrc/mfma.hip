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

// y = P x of one child over the tile (layout above).  pk: the child's packed
// matrix for this category; tbase: element offset of the tile in a
// partials buffer; p: the lane's pattern.
template <int MB>
__device__ __forceinline__ void child_mfma(const void* x, const int* states, const float* pk,
                                           long tbase, int S, int KG, int l, int p,
                                           floatx16 (&y)[MB]) {
    const int n = l & 31, h = l >> 5;
    if (states) {
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
        return;
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) y[mb] = floatx16{0};
    const GLOBAL float* xb = gin<float>(x) + tbase + n;
    const GLOBAL floatx4* a = (const GLOBAL floatx4*)gin<float>(pk) + l;
    for (int g = 0; g < KG; ++g) {
        floatx4 am[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) am[mb] = a[(mb * KG + g) * 64];
        float b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = min(8 * g + 2 * q + h, S - 1);
            b[q] = xb[k * PTILE];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
                y[mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(am[mb][q], b[q], y[mb], 0, 0, 0);
    }
}

// One post-order operation per blockIdx.y (a dependency level), one tile of
// 32 patterns per blockIdx.x, one wave per category.  Scaling semantics as
// the generic path (BeagleCPUImpl.hpp:1555-1577, 4008-4133).
template <int MB, bool AUTO>
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
    const int S = d.S, KG = (S + 7) / 8;
    const long per = mfma_pack_reals(S);
    const long tbase = pbase(d, c, tp0);
    floatx16 y[MB], z[MB];
    child_mfma<MB>(J.x1, J.s1, pack + J.pk1 + c * per, tbase, S, KG, l, p, y);
    child_mfma<MB>(J.x2, J.s2, pack + J.pk2 + c * per, tbase, S, KG, l, p, z);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) y[mb] *= z[mb];
    if (J.mode == SCALE_READ) {
        const float f = (float)gin<double>(J.scaleRead)[p];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) y[mb][r] = y[mb][r] / f;
    }
    if (AUTO || J.mode == SCALE_WRITE) {
        float mx = 0.f;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) mx = fmaxf(mx, y[mb][r]);
        mx = fmaxf(mx, __shfl_xor(mx, 32));
        if (h == 0) red[c][n] = mx;
        __syncthreads();
        for (int cc = 0; cc < d.C; ++cc) mx = fmaxf(mx, red[cc][n]);
        if (AUTO) {
            const int e = mx > 0.f ? exponent_of(mx) : 0;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int r = 0; r < 16; ++r) y[mb][r] = pow2(y[mb][r], -e);
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
            for (int mb = 0; mb < MB; ++mb) y[mb] *= inv;
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
            if (i < S && in) out[i * PTILE] = y[mb][r];
        }
}

void launch_post_mfma(const PostJob* djobs, int njobs, const Dims& d, const float* pack,
                      hipStream_t st) {
    if (njobs <= 0) return;
    const dim3 grid(d.Ppad / PTILE, njobs), block(64 * d.C);
    const int MB = mfma_mblocks(d.S);
#define POSTM(M, A) hipLaunchKernelGGL((k_postm<M, A>), grid, block, 0, st, djobs, d, pack)
    if (MB == 1) {
        if (d.autoScale) POSTM(1, true);
        else POSTM(1, false);
    } else {
        if (d.autoScale) POSTM(2, true);
        else POSTM(2, false);
    }
#undef POSTM
}

template void launch_pack_matrices<float>(const PackJob*, int, int, int, const void*, float*,
                                          hipStream_t);

}  // namespace hipbeagle
