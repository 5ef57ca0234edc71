// gfx950 kernels over node jobs (4 states): the node-fused pre-order +
// branch-derivative traversal (k_pren4 in FP64 and FP32 with any category
// count up to 8, k_pren4f on the FP32 matrix cores for 1, 2 and 4
// categories), the rate-matrix gradient (k_qgrad4, k_qgrad_finish), the
// root-frequency gradient and the gather of each job's matrices.  Layouts
// as in kernels.hip: partials [category][pattern][state], one 64-pattern
// tile per workgroup, one wave per category.
#include "kernels.h"
#include "device_common.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <type_traits>

namespace hipbeagle {

// ---------------------------------------------------------------------------
// Node-fused pre-order + derivatives (4 states).  One job per internal node
// v with children a, b: pre(v), x_a and x_b are loaded once (pre(v) comes
// from registers when the previous job produced it) and both children's
// pre-order partials and branch derivative sums are formed.  Pipelining as
// in k_post4: fixed loads into alternating register sets, unused slots read
// the dummy tile, stores of lanes past P or of outputs that are not kept go
// to the dummy tile.  With automatic scaling the exponents are loaded with
// the operands, one job ahead, and the pattern weight is loaded once.

// A node job's six matrices for one category (P_a, P_b, D1_a, D2_a, D1_b,
// D2_b; 20 reals each, contiguous after launch_gather_node_matrices) are
// loaded one job ahead as one 16-byte vector per lane, staged in the
// wave's own LDS block and read back as broadcasts, which keeps them out
// of scalar registers.  Every lane stores its vector (the block holds 64 of
// them; lanes past the matrices repeat the last one) so no branch
// separates the lanes.
constexpr int NODE_MATS = kNodeJobMats;
constexpr int NODE_F32_BLOCK = 128;       // floats of a wave's block in the FP32 node kernels
// ... and with the rate-matrix gradient: the matrices, then the factors of
// the job's two branches for the wave's category ([2][16] floats)
constexpr int NODE_F32_QBLOCK = NODE_F32_BLOCK + 32;
// A category's block of gathered job data for the FP32 node kernels: the
// six matrices column-major, padding, the factors -- the layout of the
// wave's LDS block, so it is staged as it is loaded (40 lanes x 16 bytes).
constexpr int NODE_JM = kNodeJobF32;
static_assert(NODE_JM == NODE_F32_QBLOCK, "FP32 job block = LDS block");
constexpr int NODE_MAT_BLOCK = 64 * 16;   // bytes of one wave's block

// 32-bit addressing: a wave-uniform base pointer plus a per-lane byte
// offset below 4 GiB, so the base stays in scalar registers and an address
// costs one VGPR (the host uses these kernels only for buffers < 4 GiB).
template <typename T>
__device__ __forceinline__ const GLOBAL T* at(const void* base, unsigned boff) {
    return (const GLOBAL T*)((const GLOBAL char*)base + boff);
}
template <typename T>
__device__ __forceinline__ GLOBAL T* at_w(void* base, unsigned boff) {
    return (GLOBAL T*)((GLOBAL char*)base + boff);
}

template <typename R>
struct NodeOps4 {
    R u[4];                    // parent pre-order partials (kPar 0)
    R xa[4], xb[4];            // children's post-order partials
    int sa, sb;                // children's tip states
    int eu, ea, eb;            // exponents (automatic scaling)
    typename Vec16<R>::T mv;   // this lane's share of the job's matrices (and factors)
};

// Byte offsets: `off` of this lane's 4-state vector in a partials buffer,
// `doff` of its vector in the dummy tile; tip states at 4p (dummy 4 lane),
// exponents at 2p (dummy 2 lane).
template <typename R, bool AUTO>
__device__ __forceinline__ void node_issue(const NodeJob& J, const void* dx, const void* ds,
                                           const void* dE, const R* __restrict__ jm, int c,
                                           unsigned off, unsigned doff, int p, int lane,
                                           NodeOps4<R>& o) {
    // (scratch-tile reads here, not empty buffer ranges: those take more
    // registers and this kernel is at its budget)
    const bool lp = J.kPar == 0, pa = J.k[0] == 0, pb = J.k[1] == 0;
    load4(at<R>(lp ? J.parent : dx, lp ? off : doff), o.u);
    load4(at<R>(pa ? J.x[0] : dx, pa ? off : doff), o.xa);
    load4(at<R>(pb ? J.x[1] : dx, pb ? off : doff), o.xb);
    o.sa = *at<int>(pa ? ds : J.x[0], pa ? 4u * lane : 4u * p);
    o.sb = *at<int>(pb ? ds : J.x[1], pb ? 4u * lane : 4u * p);
    {
        constexpr int V = Vec16<R>::N;
        const int l = min(lane, NODE_MATS / V - 1);
        o.mv = *at<typename Vec16<R>::T>(jm, (unsigned)((J.mbase + c * NODE_MATS + l * V) *
                                                        (int)sizeof(R)));
    }
    if (AUTO) {
        const bool a = lp && J.eParent != nullptr, b = J.ex[0] != nullptr,
                   e = J.ex[1] != nullptr;
        o.eu = *at<short>(a ? (const void*)J.eParent : dE, a ? 2u * p : 2u * lane);
        o.ea = *at<short>(b ? (const void*)J.ex[0] : dE, b ? 2u * p : 2u * lane);
        o.eb = *at<short>(e ? (const void*)J.ex[1] : dE, e ? 2u * p : 2u * lane);
    }
}

// A child's operand as a 5-vector: its partials, or for a tip the one-hot
// state whose fifth entry selects the matrices' pad column (missing state).
template <typename R>
__device__ __forceinline__ void child5(int kind, int st, const R (&x)[4], R (&o)[5]) {
    const bool isS = kind == 1;
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = isS ? (st == j ? (R)1 : (R)0) : x[j];
    o[4] = (isS && st >= 4) ? (R)1 : (R)0;
}

template <typename R>
__device__ __forceinline__ R dot4(const R (&a)[4], const R (&b)[4]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

// Derivative numerators and denominator of one branch for one category,
// in the partials' precision (categories are combined in FP64):
// den = pre . x (a tip: pre[state & 3], as the reference computes it),
// n1 = pre . (D1 x), n2 = pre . (D2 x).  One matrix is read at a time.
template <typename R>
__device__ __forceinline__ void branch_sums(int kind, int st, const R (&pre)[4], const R (&x)[5],
                                            const R* d1, const R* d2, R& den, R& n1, R& n2) {
    if (kind == 1) {
        // one-hot weights, not an index: an indexed pick is lowered to a
        // private-memory load
        const int s = st & 3;
        den = pre[0] * (R)(s == 0) + pre[1] * (R)(s == 1) + pre[2] * (R)(s == 2) +
              pre[3] * (R)(s == 3);
    } else {
        den = pre[0] * x[0] + pre[1] * x[1] + pre[2] * x[2] + pre[3] * x[3];
    }
    R t[4];
    __builtin_amdgcn_sched_barrier(0);
    mat4(d1, x, t);
    n1 = dot4(pre, t);
    __builtin_amdgcn_sched_barrier(0);
    mat4(d2, x, t);
    n2 = dot4(pre, t);
    __builtin_amdgcn_sched_barrier(0);
}

// t_j = sum_i M[i][j] u_i: the transpose product of the pre-order update
template <typename R>
__device__ __forceinline__ void mat4t(const R* m, const R (&u)[4], R (&t)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
        t[j] = m[0 * 5 + j] * u[0] + m[1 * 5 + j] * u[1] + m[2 * 5 + j] * u[2] +
               m[3 * 5 + j] * u[3];
}

template <typename R, bool AUTO>
__device__ __forceinline__ void node_compute(const NodeJob& J, const NodeOps4<R>& o, R (&pa)[4],
                                             R (&pb)[4], int& pea, int& peb, R* smat,
                                             double* sdv, int& k, int c, int lane, int p, int P,
                                             int C, unsigned off, unsigned doff, R* dummyW,
                                             void* dummyEw, const double* __restrict__ cw,
                                             double w, double* partial, int tile, int ntiles) {
    const bool in = p < P;
    R u[4], xa[5], xb[5], t[4], ua[4], ub[4], va[4], vb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = J.kPar == 2 ? pa[i] : (J.kPar == 3 ? pb[i] : o.u[i]);
    const int eu = J.kPar == 2 ? pea : (J.kPar == 3 ? peb : o.eu);
    child5(J.k[0], o.sa, o.xa, xa);
    child5(J.k[1], o.sb, o.xb, xb);
    // this wave's LDS block for the job's matrices (two alternating blocks)
    R* lm = smat + (c * 2 + (k & 1)) * (NODE_MAT_BLOCK / (int)sizeof(R));
    *(typename Vec16<R>::T*)(lm + lane * Vec16<R>::N) = o.mv;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    // scheduling barriers keep one matrix in registers at a time
    __builtin_amdgcn_sched_barrier(0);
    mat4(lm, xa, t);                     // P_a x_a
#pragma unroll
    for (int i = 0; i < 4; ++i) ub[i] = u[i] * t[i];
    __builtin_amdgcn_sched_barrier(0);
    mat4(lm + 20, xb, t);                // P_b x_b
#pragma unroll
    for (int i = 0; i < 4; ++i) ua[i] = u[i] * t[i];
    __builtin_amdgcn_sched_barrier(0);
    mat4t(lm, ua, va);                   // pre(a) = P_a^T (pre(v) * P_b x_b)
    __builtin_amdgcn_sched_barrier(0);
    mat4t(lm + 20, ub, vb);              // pre(b) = P_b^T (pre(v) * P_a x_a)
    // derivative sums of both branches, formed before rescaling (the
    // ratios do not depend on it)
    R dA, n1A, n2A, dB, n1B, n2B;
    branch_sums(J.k[0], o.sa, va, xa, lm + 40, lm + 60, dA, n1A, n2A);
    branch_sums(J.k[1], o.sb, vb, xb, lm + 80, lm + 100, dB, n1B, n2B);
    const double wc = cw[c];
    // LDS exchange [2][C][8][TILE]: max a, max b, den/n1/n2 of a, of b
    double* buf = sdv + (k & 1) * C * 8 * TILE;
    ++k;
    double* mine = buf + c * 8 * TILE + lane;
    mine[0 * TILE] = in ? (double)fmax(fmax(va[0], va[1]), fmax(va[2], va[3])) : 0.0;
    mine[1 * TILE] = in ? (double)fmax(fmax(vb[0], vb[1]), fmax(vb[2], vb[3])) : 0.0;
    mine[2 * TILE] = wc * (double)dA;
    mine[3 * TILE] = wc * (double)n1A;
    mine[4 * TILE] = wc * (double)n2A;
    mine[5 * TILE] = wc * (double)dB;
    mine[6 * TILE] = wc * (double)n1B;
    mine[7 * TILE] = wc * (double)n2B;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    const bool keepA = in && J.dest[0] != nullptr;
    const bool keepB = in && J.dest[1] != nullptr;
    int outA = 0, outB = 0;
    if (AUTO) {
        double ma = buf[lane], mb = buf[TILE + lane];
        for (int cc = 1; cc < C; ++cc) {
            ma = fmax(ma, buf[cc * 8 * TILE + lane]);
            mb = fmax(mb, buf[cc * 8 * TILE + TILE + lane]);
        }
        const int ea = ma > 0.0 ? exponent_of((R)ma) : 0;
        const int eb = mb > 0.0 ? exponent_of((R)mb) : 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            va[i] = pow2(va[i], -ea);
            vb[i] = pow2(vb[i], -eb);
        }
        // pre(a) carries the exponents of pre(v) and of x_b; pre(b) those of
        // pre(v) and x_a
        outA = ea + eu + (J.ex[1] ? o.eb : 0);
        outB = eb + eu + (J.ex[0] ? o.ea : 0);
        *at_w<short>(keepA ? (void*)J.ed[0] : dummyEw, keepA ? 2u * p : 2u * lane) = (short)outA;
        *at_w<short>(keepB ? (void*)J.ed[1] : dummyEw, keepB ? 2u * p : 2u * lane) = (short)outB;
    }
    // per-tile sums of child `ch` by wave ch % C: w r1, w r1^2, w r2 with
    // r = N / D summed over categories
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
        if (c != ch % C) continue;
        double D = 0, N1 = 0, N2 = 0;
        for (int cc = 0; cc < C; ++cc) {
            const double* q = buf + cc * 8 * TILE + (2 + 3 * ch) * TILE + lane;
            D += q[0];
            N1 += q[TILE];
            N2 += q[2 * TILE];
        }
        const double r1 = in ? N1 / D : 0.0, r2 = in ? N2 / D : 0.0;
        const double s1 = wave_sum_dpp(w * r1), sq = wave_sum_dpp(w * r1 * r1),
                     s2 = wave_sum_dpp(w * r2);
        if (lane == 0) {
            GLOBAL double* pt = gout<double>(partial);
            const long row = (long)J.oidx[ch] * 3;
            pt[(row + 0) * ntiles + tile] = s1;
            pt[(row + 1) * ntiles + tile] = sq;
            pt[(row + 2) * ntiles + tile] = s2;
        }
    }
    store4(at_w<R>(keepA ? J.dest[0] : (void*)dummyW, keepA ? off : doff), va);
    store4(at_w<R>(keepB ? J.dest[1] : (void*)dummyW, keepB ? off : doff), vb);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        pa[i] = va[i];
        pb[i] = vb[i];
    }
    pea = outA;
    peb = outB;
}

#define NODE_KERNEL_SETUP                                                                     \
    extern __shared__ double sdv[]; /* [2][C][8][TILE] exchange, then [C][2] matrix blocks */ \
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);                          \
    const int lane = threadIdx.x & 63;                                                        \
    const int tile = blockIdx.x;                                                              \
    const int p = tile * TILE + lane; /* < Ppad: pattern weights are zero-padded */           \
    R* smat = reinterpret_cast<R*>(sdv + 2 * C * 8 * TILE);                                   \
    const unsigned off = (unsigned)((c * cs + (long)p * 4) * (long)sizeof(R));                \
    const unsigned doff = (unsigned)((c * TILE + lane) * 4 * (int)sizeof(R));                 \
    const int ntiles = (P + TILE - 1) / TILE;                                                 \
    const double w = pw[p];                                                                   \
    void* dummyEw = const_cast<char*>(reinterpret_cast<const char*>(dummyE)) + 128;           \
    int k = 0, pea = 0, peb = 0;                                                              \
    R pa[4] = {0, 0, 0, 0}, pb[4] = {0, 0, 0, 0};

template <typename R, bool AUTO>
__global__ __launch_bounds__(512) void k_pren4(const NodeJob* __restrict__ jobs, int njobs,
                                               int chunk, int P, int C, long cs,
                                               const R* __restrict__ jm, const R* dummyX,
                                               const int* dummyS, R* dummyW,
                                               const short* dummyE, const double* __restrict__ cw,
                                               const double* __restrict__ pw, double* partial) {
    NODE_KERNEL_SETUP
    // workgroup (tile, y) walks jobs [y chunk, (y + 1) chunk)
    jobs += blockIdx.y * chunk;
    njobs = min(chunk, njobs - (int)blockIdx.y * chunk);
    NodeOps4<R> S0, S1;
    NodeJob J0 = jobs[0], J1;
    node_issue<R, AUTO>(J0, dummyX, dummyS, dummyE, jm, c, off, doff, p, lane, S0);
    for (int o = 0;; o += 2) {
        J1 = jobs[min(o + 1, njobs - 1)];
        node_issue<R, AUTO>(J1, dummyX, dummyS, dummyE, jm, c, off, doff, p, lane, S1);
        node_compute<R, AUTO>(J0, S0, pa, pb, pea, peb, smat, sdv, k, c, lane, p, P, C, off, doff,
                              dummyW, dummyEw, cw, w, partial, tile, ntiles);
        if (o + 1 >= njobs) break;
        J0 = jobs[min(o + 2, njobs - 1)];
        node_issue<R, AUTO>(J0, dummyX, dummyS, dummyE, jm, c, off, doff, p, lane, S0);
        node_compute<R, AUTO>(J1, S1, pa, pb, pea, peb, smat, sdv, k, c, lane, p, P, C, off, doff,
                              dummyW, dummyEw, cw, w, partial, tile, ntiles);
        if (o + 2 >= njobs) break;
    }
}

// One dependency level of node jobs: one job per workgroup, grid (tiles,
// jobs); every parent is read from memory.
template <typename R, bool AUTO>
__global__ __launch_bounds__(512) void k_pren4_level(const NodeJob* __restrict__ jobs, int P,
                                                     int C, long cs, const R* __restrict__ jm,
                                                     const R* dummyX, const int* dummyS,
                                                     R* dummyW, const short* dummyE,
                                                     const double* __restrict__ cw,
                                                     const double* __restrict__ pw,
                                                     double* partial) {
    NODE_KERNEL_SETUP
    NodeOps4<R> S0;
    const NodeJob J = jobs[blockIdx.y];
    node_issue<R, AUTO>(J, dummyX, dummyS, dummyE, jm, c, off, doff, p, lane, S0);
    node_compute<R, AUTO>(J, S0, pa, pb, pea, peb, smat, sdv, k, c, lane, p, P, C, off, doff,
                          dummyW, dummyEw, cw, w, partial, tile, ntiles);
}
#undef NODE_KERNEL_SETUP

// ---------------------------------------------------------------------------
// FP32 node jobs with the category count C as a template parameter (C = 1,
// 2, 4).  Same jobs and results as node_compute up to FP32 rounding; the
// work per job is cut down to the 4x4 products:
//  - a tip child's products P x and D x are its state's matrix column,
//    read from the wave's LDS block with the state as the lane's index (the
//    pad column for a missing state), not 20 products with a one-hot vector;
//  - categories exchange FP32 values, [2][8 quantities][C][TILE] floats;
//  - the derivative ratios are formed and summed over the tile's patterns
//    in FP32 (the tile sums are added in FP64 across tiles), by DPP steps;
//  - the two waves that form the sums alternate from job to job, so every
//    SIMD does the same work.

template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_movf(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}

// Three sums over the 64 lanes of a wave at once, each in the order of a
// pairwise tree (lane pairs, quads, row halves, rows, then (R0 + R1) +
// (R2 + R3) over the four 16-lane rows): the first step leaves a's pair sums
// in the even lanes and b's in the odd ones of one register, c's in the
// even lanes of a second; the quad and row steps keep the lane parity; the
// two registers' row totals are then added across rows by one permlane16
// and one permlane32 swap.  a ends in lane 14, b in lane 15, c in lane 30.
__device__ __forceinline__ void wave_sum3_f32(float a, float b, float c, float& sa, float& sb,
                                              float& sc) {
    const bool odd = threadIdx.x & 1;
    float x = (odd ? b : a) + dpp_movf<0xB1, 0xf>(odd ? a : b);
    float y = c + dpp_movf<0xB1, 0xf>(c);
    x += dpp_movf<0x4E, 0xf>(x);
    y += dpp_movf<0x4E, 0xf>(y);
    x += dpp_movf<0x114, 0xf>(x);
    y += dpp_movf<0x114, 0xf>(y);
    x += dpp_movf<0x118, 0xf>(x);
    y += dpp_movf<0x118, 0xf>(y);
    // rows: [x0, y0, x2, y2] and [x1, y1, x3, y3]
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false,
                                                    false);
    const float r = __uint_as_float(p[0]) + __uint_as_float(p[1]);
    const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(r), __float_as_uint(r), false,
                                                    false);
    const float t = __uint_as_float(q[0]) + __uint_as_float(q[1]);
    sa = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t), 14));
    sb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t), 15));
    sc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t), 30));
}

// Largest of four non-negative floats: their bit patterns order as the
// values do (+0 lowest), so integer maxima (no NaN quieting of the inputs).
__device__ __forceinline__ float max4_nonneg(const float (&v)[4]) {
    return __int_as_float(max(max(__float_as_int(v[0]), __float_as_int(v[1])),
                              max(__float_as_int(v[2]), __float_as_int(v[3]))));
}

// The FP32 node kernels read each matrix of a job column by column from the
// wave's LDS block: [5 columns][4 rows] (column 4 is the pad column), so a
// column is one aligned 16-byte read.
__device__ __forceinline__ fvec4 column(const float* m, int j) {
    return *(const fvec4*)(m + 4 * j);
}

// 4x4 products on the matrix cores, lane = pattern: v_mfma_f32_4x4x1f32
// treats the wave as 16 blocks of 4 lanes; in instruction k lane l supplies
// A[l % 4][k] and B[k][l % 4] of block l / 4, and element r of lane l's
// accumulator is D[r][l % 4] of that block.  With B[k][.] = x[k] of the
// lane's own pattern and A[r][k] = M[r][k] (the same matrix in every block),
// lane l ends with (M x)[0..3] for its pattern after four instructions.
// r = lane % 4.

// t = M x: A from row r of M, M[r][k] = column k's entry r
__device__ __forceinline__ void mfma_rows(const float* m, int r, const float (&x)[4],
                                          float (&t)[4]) {
    fvec4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m[4 * k + r], x[k], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = acc[i];
}

// t = M^T u: A[r][k] = M[k][r], column r of M
__device__ __forceinline__ void mfma_cols(const float* m, int r, const float (&u)[4],
                                          float (&t)[4]) {
    const fvec4 col = column(m, r);
    fvec4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = __builtin_amdgcn_mfma_f32_4x4x1f32(col[k], u[k], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = acc[i];
}

// ---------------------------------------------------------------------------
// FP32 rate-matrix gradient (4 states) on the matrix cores: the sums of
// qw_compute (H += F o ((V^T u)(V^-1 x)^T) w_c w_p / L_p per branch) for
// lanes that hold one pattern of one category each, the four lanes of every
// 4-lane block the same category.  In v_mfma_f32_4x4x1f32 three chains of
// instructions give per block, with lane 4b + s holding component s of the
// block's four patterns i (the products transpose):
//   alpha_i = V^T (coef_i top_i)   A = the lane's own vector, B = V[k][s]
//   beta_i  = V^-1 x_i             A = the lane's own vector, B = V^-1[s][k]
//   M^T     = sum_i beta_i alpha_i^T   (A = beta_i, B = alpha_i; lane 4b + s
//             ends with row s of M = sum_i alpha_i beta_i^T)
// and the lane adds F[s][.] o M[s][.] to its four FP64 accumulators (F of
// its category and branch).  V and V^-1 enter rounded to FP32: on the
// 512-taxon, 50,000-pattern workload dlogL/dQ stays within 2.5e-8 of the
// FP64 path (relative to its largest entry), as with V split into FP32
// hi + lo parts (2.1e-8), at 12 instructions per branch instead of 20.
// Lanes with the same s are summed at the end.

struct QgVec {
    float vh[4];   // V[k][s], s = lane % 4
    float wh[4];   // V^-1[s][k]
};

// f: the lane's four factors F[s][.] (FP64, or FP32 where a job's sums are
// formed in FP32 and added in FP64 afterwards); H accumulates in T.
template <typename FT, typename T>
__device__ __forceinline__ void qgrad_block(const float (&ta)[4], const float (&x)[4],
                                            const QgVec& v, const FT* f, T (&Hr)[4]) {
    fvec4 al = {0.0f, 0.0f, 0.0f, 0.0f}, be = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        al = __builtin_amdgcn_mfma_f32_4x4x1f32(ta[k], v.vh[k], al, 0, 0, 0);
        be = __builtin_amdgcn_mfma_f32_4x4x1f32(x[k], v.wh[k], be, 0, 0, 0);
    }
    fvec4 m = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 4; ++i) m = __builtin_amdgcn_mfma_f32_4x4x1f32(be[i], al[i], m, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) Hr[i] += (T)f[i] * (T)m[i];
}

// The same with V^T and V^-1 (FP32, [2][4][4]: row s of each) in LDS, read
// where the products need them rather than held in registers.
template <typename FT, typename T>
__device__ __forceinline__ void qgrad_block(const float (&ta)[4], const float (&x)[4],
                                            const float* sv, int s, const FT* f, T (&Hr)[4]) {
    QgVec v;
    const fvec4 a = *(const fvec4*)(sv + 4 * s), b = *(const fvec4*)(sv + 16 + 4 * s);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v.vh[k] = a[k];
        v.wh[k] = b[k];
    }
    qgrad_block(ta, x, v, f, Hr);
}

// V^T and V^-1 rows (FP32) for LDS: thread t < 32 loads entry t
// (qgrad_vector_load) and stores it (qgrad_vector_store) once the job's
// operand loads are issued, so the wait for it does not hold them back.
__device__ __forceinline__ double qgrad_vector_load(const double* __restrict__ eig) {
    const int t = min((int)threadIdx.x, 31);
    return eig[t < 16 ? (t & 3) * 4 + (t >> 2) : t];
}

// ... and after them the rows of a tip's one-hot vectors: e_0 .. e_3 and a
// zero row (a missing state or a lane past the patterns), sv[32 + 4 code]
__device__ __forceinline__ void qgrad_vector_store(double e, float* sv) {
    const int t = threadIdx.x;
    if (t < 32) sv[t] = (float)e;
    else if (t < 52) sv[t] = (t - 32) % 5 == 0 && t < 48 ? 1.0f : 0.0f;
}

__device__ __forceinline__ void qgrad_vectors(const double* __restrict__ eig, int s, QgVec& v) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v.vh[k] = (float)eig[k * 4 + s];
        v.wh[k] = (float)eig[16 + s * 4 + k];
    }
}

// t = M x for a child: its partials (matrix cores), or for a tip the
// column of its state (s in [0, 4], a per-lane read; column 4 is the pad
// column of a missing state).
__device__ __forceinline__ void child_product(bool tip, int s, int r, const float (&x)[4],
                                              const float* m, float (&t)[4]) {
    if (tip) {
        const fvec4 c = column(m, s);
#pragma unroll
        for (int i = 0; i < 4; ++i) t[i] = c[i];
    } else {
        mfma_rows(m, r, x, t);
    }
}

// Derivative numerators and denominator of one branch for one category:
// den = pre . x (a tip: pre[state & 3], as the reference computes it),
// n1 = pre . (D1 x), n2 = pre . (D2 x).
__device__ __forceinline__ void branch_sums_f32(bool tip, int s, int r, const float (&pre)[4],
                                                const float (&x)[4], const float* d1,
                                                const float* d2, float& den, float& n1,
                                                float& n2) {
    float t[4];
    if (tip) {
        // bit tests, not an index: an indexed pick is lowered to a
        // private-memory load
        const float lo = (s & 1) ? pre[1] : pre[0], hi = (s & 1) ? pre[3] : pre[2];
        den = (s & 2) ? hi : lo;
    } else {
        den = dot4(pre, x);
    }
    child_product(tip, s, r, x, d1, t);
    n1 = dot4(pre, t);
    child_product(tip, s, r, x, d2, t);
    n2 = dot4(pre, t);
}

// A register with no particular value and no instruction to set it, for
// operand slots a job does not load (its uses are selected or branched
// away by the same job fields that skip the load).
template <typename T>
__device__ __forceinline__ T unset() {
    T x;
    asm volatile("" : "=v"(x));
    return x;
}

// Operand loads of an FP32 node job without the dummy tile.
template <bool AUTO, bool QG = false>
__device__ __forceinline__ void node_issue_direct(const NodeJob& J, const float* __restrict__ jm,
                                                  int c, unsigned off, int p, int lane,
                                                  NodeOps4<float>& o, bool valid = true) {
    // every operand slot is loaded, from an empty range (reads 0, no
    // traffic) where the job has no such operand (or, !valid, for no job):
    // no branches, so no value merges that would make the compiler wait
    // for earlier loads before issuing the later ones
    load4_range<float>(range_w(J.parent, valid ? J.rU : 0), off, o.u);
    load4_range<float>(range_w(J.x[0], valid ? J.rX[0] : 0), off, o.xa);
    load4_range<float>(range_w(J.x[1], valid ? J.rX[1] : 0), off, o.xb);
    o.sa = (int)__builtin_amdgcn_raw_buffer_load_b32(range_w(J.x[0], valid ? J.rS[0] : 0),
                                                     4u * p, 0, 0);
    o.sb = (int)__builtin_amdgcn_raw_buffer_load_b32(range_w(J.x[1], valid ? J.rS[1] : 0),
                                                     4u * p, 0, 0);
    {
        // the category's block: matrices (lanes 0-29), with the rate-matrix
        // gradient padding and factors too (lanes 30-39); later lanes repeat
        // the last vector
        const int l = min(lane, (QG ? NODE_JM : NODE_MATS) / 4 - 1);
        const auto q = __builtin_amdgcn_raw_buffer_load_b128(
            range_if(jm, valid, jm),
            (unsigned)((J.mbase + c * NODE_JM + l * 4) * (int)sizeof(float)), 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) o.mv[i] = __uint_as_float(q[i]);
    }
    if (AUTO) {
        o.eu = (short)__builtin_amdgcn_raw_buffer_load_b16(
            range_w(J.eParent, valid ? J.rEu : 0), 2u * p, 0, 0);
        o.ea = (short)__builtin_amdgcn_raw_buffer_load_b16(
            range_w(J.ex[0], valid ? J.rEx[0] : 0), 2u * p, 0, 0);
        o.eb = (short)__builtin_amdgcn_raw_buffer_load_b16(
            range_w(J.ex[1], valid ? J.rEx[1] : 0), 2u * p, 0, 0);
    }
}

template <int C, bool AUTO, bool QG = false, int NPAR = 2>
__device__ __forceinline__ void node_compute_f32(const NodeJob& J, const NodeOps4<float>& o,
                                                 float (&pa)[4], float (&pb)[4], int& pea,
                                                 int& peb, float* smat, float* sx, int& k, int c,
                                                 int lane, int p, int P, unsigned off,
                                                 const float (&wc)[C], float w, double* partial,
                                                 int tile, int ntiles, const float* sv,
                                                 float (&Hj)[4]) {
    constexpr int NQ = QG ? 9 : 8;   // quantities exchanged per pattern and category
    constexpr int BLK = QG ? NODE_F32_QBLOCK : NODE_F32_BLOCK;
    const bool in = p < P;
    float u[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = J.kPar == 2 ? pa[i] : (J.kPar == 3 ? pb[i] : o.u[i]);
    const int eu = J.kPar == 2 ? pea : (J.kPar == 3 ? peb : o.eu);
    const bool tipA = J.k[0] == 1, tipB = J.k[1] == 1;
    // tip states are stored in [0, 4]; the clamp keeps the column index in
    // the matrix whatever a lane without a tip read
    const int sa = min((unsigned)o.sa, 4u), sb = min((unsigned)o.sb, 4u);
    // NPAR 2: alternate buffers from job to job (one barrier per job);
    // NPAR 1: one job per workgroup
    const int par = NPAR == 2 ? (k & 1) : 0;
    float* lm = smat + (c * NPAR + par) * BLK;
    float* lf = lm + NODE_F32_BLOCK;   // QG: [2][16] factors (FP32)
    *(fvec4*)(lm + min(lane, BLK / 4 - 1) * 4) = o.mv;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    const int r = lane & 3;
    float t[4], ua[4], ub[4], va[4], vb[4];
    child_product(tipA, sa, r, o.xa, lm, t);       // P_a x_a
#pragma unroll
    for (int i = 0; i < 4; ++i) ub[i] = u[i] * t[i];
    child_product(tipB, sb, r, o.xb, lm + 20, t);  // P_b x_b
#pragma unroll
    for (int i = 0; i < 4; ++i) ua[i] = u[i] * t[i];
    // QG: this category's likelihood of the pattern, sum pre(v) P_a x_a P_b x_b
    const float Lc = QG ? ub[0] * t[0] + ub[1] * t[1] + ub[2] * t[2] + ub[3] * t[3] : 0.0f;
    mfma_cols(lm, r, ua, va);                      // pre(a) = P_a^T (pre(v) * P_b x_b)
    mfma_cols(lm + 20, r, ub, vb);                 // pre(b) = P_b^T (pre(v) * P_a x_a)
    float dA, n1A, n2A, dB, n1B, n2B;
    branch_sums_f32(tipA, sa, r, va, o.xa, lm + 40, lm + 60, dA, n1A, n2A);
    branch_sums_f32(tipB, sb, r, vb, o.xb, lm + 80, lm + 100, dB, n1B, n2B);
    // exchange: max a, max b, den/n1/n2 of a, of b (QG: and Lc)
    float* buf = sx + par * NQ * C * TILE;
    ++k;
    float* mine = buf + c * TILE + lane;
    mine[0 * C * TILE] = in ? max4_nonneg(va) : 0.0f;
    mine[1 * C * TILE] = in ? max4_nonneg(vb) : 0.0f;
    mine[2 * C * TILE] = dA;
    mine[3 * C * TILE] = n1A;
    mine[4 * C * TILE] = n2A;
    mine[5 * C * TILE] = dB;
    mine[6 * C * TILE] = n1B;
    mine[7 * C * TILE] = n2B;
    if (QG) mine[8 * C * TILE] = Lc;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    int outA = 0, outB = 0;
    if (AUTO) {
        int ima = __float_as_int(buf[lane]), imb = __float_as_int(buf[C * TILE + lane]);
#pragma unroll
        for (int cc = 1; cc < C; ++cc) {
            ima = max(ima, __float_as_int(buf[cc * TILE + lane]));
            imb = max(imb, __float_as_int(buf[(C + cc) * TILE + lane]));
        }
        const float ma = __int_as_float(ima), mb = __int_as_float(imb);
        const int ea = ma > 0.0f ? exponent_of(ma) : 0;
        const int eb = mb > 0.0f ? exponent_of(mb) : 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            va[i] = pow2(va[i], -ea);
            vb[i] = pow2(vb[i], -eb);
        }
        outA = ea + eu + (J.ex[1] ? o.eb : 0);
        outB = eb + eu + (J.ex[0] ? o.ea : 0);
        // wave 0 stores the exponents; patterns past P fall outside the range
        __builtin_amdgcn_raw_buffer_store_b16(
            (unsigned short)outA,
            range_w(J.ed[0], c == 0 ? J.rEd[0] : 0),
            2u * p, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b16(
            (unsigned short)outB,
            range_w(J.ed[1], c == 0 ? J.rEd[1] : 0),
            2u * p, 0, 0);
    }
    // tile sums of child ch (w r1, w r1^2, w r2, r = N / D over categories)
    // by wave (ch + 2 par) % C
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
        const bool mine_ch = c == (ch + 2 * par) % C;
        float s1 = 0.0f, sq = 0.0f, s2 = 0.0f;
        if (mine_ch) {
            const float* q = buf + (2 + 3 * ch) * C * TILE + lane;
            float D = wc[0] * q[0], N1 = wc[0] * q[C * TILE], N2 = wc[0] * q[2 * C * TILE];
#pragma unroll
            for (int cc = 1; cc < C; ++cc) {
                D += wc[cc] * q[cc * TILE];
                N1 += wc[cc] * q[(C + cc) * TILE];
                N2 += wc[cc] * q[(2 * C + cc) * TILE];
            }
            const float inv = __builtin_amdgcn_rcpf(D);   // 1 ulp
            const float r1 = in ? N1 * inv : 0.0f, r2 = in ? N2 * inv : 0.0f;
            const float a = w * r1;
            wave_sum3_f32(a, a * r1, w * r2, s1, sq, s2);
        }
        // every lane stores the same wave sums (one write per address);
        // the other waves' stores fall in an empty range
        const auto rp = range_w(partial, mine_ch ? -1 : 0);
        const unsigned row = (unsigned)J.oidx[ch] * 3u;
        const double sv3[3] = {(double)s1, (double)sq, (double)s2};
#pragma unroll
        for (int k3 = 0; k3 < 3; ++k3) {
            typedef unsigned u2 __attribute__((ext_vector_type(2)));
            const u2 d = {(unsigned)__double2loint(sv3[k3]), (unsigned)__double2hiint(sv3[k3])};
            __builtin_amdgcn_raw_buffer_store_b64(
                d, rp, ((row + k3) * (unsigned)ntiles + (unsigned)tile) * 8u, 0, 0);
        }
    }
    {
        // pre-order partials of the children that keep them; patterns past
        // P fall outside the category's range
        typedef unsigned u4 __attribute__((ext_vector_type(4)));
        const unsigned cat = off - 16u * (unsigned)p;   // the category's first byte
        const u4 qa = {__float_as_uint(va[0]), __float_as_uint(va[1]), __float_as_uint(va[2]),
                       __float_as_uint(va[3])};
        const u4 qb = {__float_as_uint(vb[0]), __float_as_uint(vb[1]), __float_as_uint(vb[2]),
                       __float_as_uint(vb[3])};
        __builtin_amdgcn_raw_buffer_store_b128(
            qa, range_w((const char*)J.dest[0] + cat, J.rD[0]), 16u * p, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b128(
            qb, range_w((const char*)J.dest[1] + cat, J.rD[1]), 16u * p, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        pa[i] = va[i];
        pb[i] = vb[i];
    }
    pea = outA;
    peb = outB;
    if (QG) {
        // rate-matrix gradient of both branches (after the stores, which
        // then drain meanwhile): top_a = pre(v) * P_b x_b
        // (ua), top_b = pre(v) * P_a x_a (ub); L_p = sum_c w_c Lc, the same
        // order of additions in every wave; scale factors cancel in
        // top x^T / L_p.  The job's sums are formed in FP32 (Hj) and added
        // in FP64 by the caller.
#pragma unroll
        for (int i = 0; i < 4; ++i) Hj[i] = 0.0f;
        float L = 0.0f, wcc = wc[0];
#pragma unroll
        for (int cc = 0; cc < C; ++cc) {
            L += wc[cc] * buf[(8 * C + cc) * TILE + lane];
            if (c == cc) wcc = wc[cc];
        }
        // coef is zero past the patterns (their partials are zero or
        // finite, never written by the passes): ta is then zero there
        const float coef = in ? (w * wcc) * __builtin_amdgcn_rcpf(L) : 0.0f;
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
            const bool tip = ch == 0 ? tipA : tipB;
            const int st = ch == 0 ? sa : sb;
            float ta[4], x[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) ta[i] = coef * (ch == 0 ? ua[i] : ub[i]);
            if (tip) {
                // a tip's vector is one-hot; a missing state (4) reads the
                // pad column, which does not depend on Q: the zero row;
                // lanes past the patterns add nothing
                const fvec4 e = *(const fvec4*)(sv + 32 + 4 * (in && st < 4 ? st : 4));
#pragma unroll
                for (int i = 0; i < 4; ++i) x[i] = e[i];
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) x[i] = in ? (ch == 0 ? o.xa[i] : o.xb[i]) : 0.0f;
            }
            qgrad_block(ta, x, sv, r, lf + ch * 16 + r * 4, Hj);
        }
    }
}

#define NODE_F32_SETUP(NQ, NPAR)                                                              \
    extern __shared__ double sdv[]; /* [NPAR][NQ][C][TILE] floats, [C][NPAR] matrix blocks */ \
    static_assert(NODE_MATS <= NODE_F32_BLOCK, "matrix block");                               \
    float* sx = reinterpret_cast<float*>(sdv);                                                \
    float* smat = sx + (NPAR) * (NQ) * C * TILE;                                              \
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);                          \
    const int lane = threadIdx.x & 63;                                                        \
    const int tile = blockIdx.x;                                                              \
    const int p = tile * TILE + lane; /* < Ppad: pattern weights are zero-padded */           \
    const unsigned off = (unsigned)((c * cs + (long)p * 4) * (long)sizeof(float));            \
    const int ntiles = (P + TILE - 1) / TILE;                                                 \
    const float w = (float)pw[p];                                                             \
    float wc[C];                                                                              \
    _Pragma("unroll") for (int cc = 0; cc < C; ++cc) wc[cc] = __int_as_float(                 \
        __builtin_amdgcn_readfirstlane(__float_as_int((float)cw[cc])));                       \
    int k = 0, pea = 0, peb = 0;                                                              \
    float pa[4] = {0, 0, 0, 0}, pb[4] = {0, 0, 0, 0};

// v summed over the 16 lanes with the same lane % 4: two row rotations
// (DPP), then the row pairs and the wave halves swapped (permlane16/32
// swaps); lanes 0-3 hold the sums for s = 0-3.
__device__ __forceinline__ double swap_sum(double v, bool half32) {
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    const auto a = half32 ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                          : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = half32 ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                          : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __hiloint2double((int)b[0], (int)a[0]) + __hiloint2double((int)b[1], (int)a[1]);
}

__device__ __forceinline__ double sum_by_s(double v) {
    v += dpp_move<0x124, 0xf>(v);   // row_ror:4
    v += dpp_move<0x128, 0xf>(v);   // row_ror:8: the row's four lanes of this s
    v = swap_sum(v, false);         // rows 0 + 1, 2 + 3
    return swap_sum(v, true);       // halves
}

// With the rate-matrix gradient (QG): the workgroup's H (the lanes' shares
// summed over its jobs, then over the lanes with the same lane % 4 and over
// the category waves in order) is written to its own slot of hacc
// ([slot][16], slot = base + chunk * tiles + tile); every slot of a pass is
// written once, by one workgroup, so the pass needs no zeroed accumulator
// and its sums do not depend on the timing of the workgroups.
// The same for one job's FP32 sums (one-job workgroups): the row sums of
// each s by two DPP rotations, the rows' and waves' sums (16 per wave)
// through LDS, added in FP64 in a fixed order by wave 0.
template <int C>
__device__ __forceinline__ void qgrad_store_f32(float (&Hj)[4], int c, int lane, float* red,
                                                double* __restrict__ hacc, long slot) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        Hj[i] += dpp_movf<0x124, 0xf>(Hj[i]);   // row_ror:4
        Hj[i] += dpp_movf<0x128, 0xf>(Hj[i]);   // row_ror:8: the row's sum for this s
    }
    if ((lane & 15) < 4) {
#pragma unroll
        for (int i = 0; i < 4; ++i) red[(c * 4 + (lane >> 4)) * 16 + (lane & 3) * 4 + i] = Hj[i];
    }
    __syncthreads();
    if (c == 0 && lane < 16) {
        double h = 0.0;
#pragma unroll
        for (int q = 0; q < 4 * C; ++q) h += (double)red[q * 16 + lane];
        gout<double>(hacc)[slot * 16 + lane] = h;
    }
}

template <int C>
__device__ __forceinline__ void qgrad_store(double (&Hr)[4], int c, int lane, double* red,
                                            double* __restrict__ hacc, long slot) {
#pragma unroll
    for (int i = 0; i < 4; ++i) Hr[i] = sum_by_s(Hr[i]);
    if (lane < 4) {
#pragma unroll
        for (int i = 0; i < 4; ++i) red[c * 16 + lane * 4 + i] = Hr[i];
    }
    __syncthreads();
    if (c == 0 && lane < 16) {
        double h = red[lane];
#pragma unroll
        for (int cc = 1; cc < C; ++cc) h += red[cc * 16 + lane];
        gout<double>(hacc)[slot * 16 + lane] = h;
    }
}

// LDS of the FP32 node kernels: the exchange and the waves' matrix (and
// factor) blocks, NPAR of each (2 for a list of jobs, 1 for one job per
// workgroup), and with QG the workgroup's H by category and V^T | V^-1
template <int C, bool QG, int NPAR>
constexpr size_t node_f32_lds() {
    return (size_t)(NPAR * (QG ? 9 : 8) * C * TILE +
                    C * NPAR * (QG ? NODE_F32_QBLOCK : NODE_F32_BLOCK)) * sizeof(float) +
           (QG ? 64 * C * sizeof(float) + 52 * sizeof(float) : 0);
}

// Each job's operands are loaded when the job starts (node_issue_direct):
// at most 64 VGPRs, so eight waves per SIMD and a 100,000-pattern grid
// resident at once, which covers the loads' latency better than loading
// the next job ahead in more registers does.
template <int C, bool AUTO, bool QG>
__global__ __launch_bounds__(64 * C, QG ? 4 : 7) void k_pren4f(
    const NodeJob* __restrict__ jobs, int njobs, int chunk, int P, long cs,
    const float* __restrict__ jm, const double* __restrict__ cw, const double* __restrict__ pw,
    double* partial, const double* __restrict__ eig, double* __restrict__ hacc, long base) {
    NODE_F32_SETUP(QG ? 9 : 8, 2)
    double* red = reinterpret_cast<double*>(smat + C * 2 * NODE_F32_QBLOCK);
    float* sv = reinterpret_cast<float*>(red + 16 * C);
    double Hr[4] = {0.0, 0.0, 0.0, 0.0};
    const double ev = QG ? qgrad_vector_load(eig) : 0.0;
    jobs += blockIdx.y * chunk;
    njobs = min(chunk, njobs - (int)blockIdx.y * chunk);
    for (int o = 0; o < njobs; ++o) {
        const NodeJob J = jobs[o];
        NodeOps4<float> S0;
        node_issue_direct<AUTO, QG>(J, jm, c, off, p, lane, S0);
        if (QG && o == 0) qgrad_vector_store(ev, sv);   // read after the job's barrier
        float Hj[4];
        node_compute_f32<C, AUTO, QG>(J, S0, pa, pb, pea, peb, smat, sx, k, c, lane, p, P, off,
                                      wc, w, partial, tile, ntiles, sv, Hj);
        if (QG) {
#pragma unroll
            for (int i = 0; i < 4; ++i) Hr[i] += (double)Hj[i];
        }
    }
    if (QG) qgrad_store<C>(Hr, c, lane, red, hacc, base + (long)blockIdx.y * gridDim.x + blockIdx.x);
}

// A batched level walked in chunks of independent jobs, software-
// pipelined: while job o computes, the operands of job o + 1 are in flight
// (two operand register sets, the loop unrolled by two so no loaded
// register is copied).  Every job issues the same memory operations -- the
// loads and stores a job does not need go to empty buffer ranges, and the
// one past the chunk's end loads nothing -- so the compiler's waits count
// exactly the next job's loads and never the stores.
template <int C, bool AUTO, bool QG>
__global__ __launch_bounds__(64 * C, QG ? 3 : 4) void k_pren4f_pipe(
    const NodeJob* __restrict__ jobs, int njobs, int chunk, int P, long cs,
    const float* __restrict__ jm, const double* __restrict__ cw, const double* __restrict__ pw,
    double* partial, const double* __restrict__ eig, double* __restrict__ hacc, long base) {
    NODE_F32_SETUP(QG ? 9 : 8, 2)
    double* red = reinterpret_cast<double*>(smat + C * 2 * NODE_F32_QBLOCK);
    float* sv = reinterpret_cast<float*>(red + 16 * C);
    double Hr[4] = {0.0, 0.0, 0.0, 0.0};
    const double ev = QG ? qgrad_vector_load(eig) : 0.0;
    jobs += blockIdx.y * chunk;
    const int n = min(chunk, njobs - (int)blockIdx.y * chunk);
    NodeOps4<float> A, B;
    NodeJob JA = jobs[0], JB;
    node_issue_direct<AUTO, QG>(JA, jm, c, off, p, lane, A);
    if (QG) qgrad_vector_store(ev, sv);   // read after the first job's barrier
    for (int o = 0;; o += 2) {
        JB = jobs[min(o + 1, n - 1)];
        node_issue_direct<AUTO, QG>(JB, jm, c, off, p, lane, B, o + 1 < n);
        float Hj[4];
        node_compute_f32<C, AUTO, QG>(JA, A, pa, pb, pea, peb, smat, sx, k, c, lane, p, P, off,
                                      wc, w, partial, tile, ntiles, sv, Hj);
        if (QG) {
#pragma unroll
            for (int i = 0; i < 4; ++i) Hr[i] += (double)Hj[i];
        }
        if (o + 1 >= n) break;
        JA = jobs[min(o + 2, n - 1)];
        node_issue_direct<AUTO, QG>(JA, jm, c, off, p, lane, A, o + 2 < n);
        node_compute_f32<C, AUTO, QG>(JB, B, pa, pb, pea, peb, smat, sx, k, c, lane, p, P, off,
                                      wc, w, partial, tile, ntiles, sv, Hj);
        if (QG) {
#pragma unroll
            for (int i = 0; i < 4; ++i) Hr[i] += (double)Hj[i];
        }
        if (o + 2 >= n) break;
    }
    if (QG) qgrad_store<C>(Hr, c, lane, red, hacc, base + (long)blockIdx.y * gridDim.x + blockIdx.x);
}

template <int C, bool AUTO, bool QG>
__global__ __launch_bounds__(64 * C, QG ? 7 : 8) void k_pren4f_level(
    const NodeJob* __restrict__ jobs, int P, long cs, const float* __restrict__ jm,
    const double* __restrict__ cw, const double* __restrict__ pw, double* partial,
    const double* __restrict__ eig, double* __restrict__ hacc, long base) {
    NODE_F32_SETUP(QG ? 9 : 8, 1)
    NodeOps4<float> S0;
    float* red = smat + C * NODE_F32_QBLOCK;   // [C][4 rows][16]
    float* sv = red + 64 * C;
    float Hj[4];
    const double ev = QG ? qgrad_vector_load(eig) : 0.0;
    const NodeJob J = jobs[blockIdx.y];
    node_issue_direct<AUTO, QG>(J, jm, c, off, p, lane, S0);
    if (QG) qgrad_vector_store(ev, sv);   // read after the job's barrier
    node_compute_f32<C, AUTO, QG, 1>(J, S0, pa, pb, pea, peb, smat, sx, k, c, lane, p, P, off,
                                     wc, w, partial, tile, ntiles, sv, Hj);
    if (QG)
        qgrad_store_f32<C>(Hj, c, lane, red, hacc, base + (long)blockIdx.y * gridDim.x + blockIdx.x);
}
#undef NODE_F32_SETUP

// sum over the four categories of a pattern held by lanes j, j + 16, j + 32,
// j + 48 (the same order of additions in every lane)
__device__ __forceinline__ float cat_sum(float v) {
    v += __shfl_xor(v, 16, 64);
    return v + __shfl_xor(v, 32, 64);
}

template <int C, bool QG>
static void launch_pre_nodes_f32(const NodeJob* djobs, int njobs, bool batched, int chunk,
                                 const Dims& d, const float* m, const double* cw,
                                 const double* pw, double* partial, hipStream_t st,
                                 NodeQGrad* q) {
    const int tiles = (d.P + TILE - 1) / TILE;
    const dim3 block(TILE * C);
    const double* eig = QG ? q->eig : nullptr;
    double* hacc = QG ? q->hacc : nullptr;
    const long base = QG ? q->used : 0;
    int rows;
    if (batched && chunk <= 1) {
        rows = njobs;
        const size_t lds = node_f32_lds<C, QG, 1>();
        if (d.autoScale) hipLaunchKernelGGL((k_pren4f_level<C, true, QG>), dim3(tiles, njobs), block, lds, st, djobs, d.P, d.catStride, m, cw, pw, partial, eig, hacc, base);
        else hipLaunchKernelGGL((k_pren4f_level<C, false, QG>), dim3(tiles, njobs), block, lds, st, djobs, d.P, d.catStride, m, cw, pw, partial, eig, hacc, base);
    } else if (batched) {
        // chunks of independent jobs: the pipelined walk
        chunk = std::min(chunk, njobs);
        rows = (njobs + chunk - 1) / chunk;
        const dim3 grid(tiles, rows);
        const size_t lds = node_f32_lds<C, QG, 2>();
        if (d.autoScale) hipLaunchKernelGGL((k_pren4f_pipe<C, true, QG>), grid, block, lds, st, djobs, njobs, chunk, d.P, d.catStride, m, cw, pw, partial, eig, hacc, base);
        else hipLaunchKernelGGL((k_pren4f_pipe<C, false, QG>), grid, block, lds, st, djobs, njobs, chunk, d.P, d.catStride, m, cw, pw, partial, eig, hacc, base);
    } else {
        // a fused list: one chunk, parents forwarded from registers
        chunk = njobs;
        rows = 1;
        const dim3 grid(tiles, rows);
        const size_t lds = node_f32_lds<C, QG, 2>();
        if (d.autoScale) hipLaunchKernelGGL((k_pren4f<C, true, QG>), grid, block, lds, st, djobs, njobs, chunk, d.P, d.catStride, m, cw, pw, partial, eig, hacc, base);
        else hipLaunchKernelGGL((k_pren4f<C, false, QG>), grid, block, lds, st, djobs, njobs, chunk, d.P, d.catStride, m, cw, pw, partial, eig, hacc, base);
    }
    if (QG) q->used += (long)rows * tiles;
}

// ---------------------------------------------------------------------------
// Rate-matrix gradient (4 states).  For one eigen-system Q = V diag(lam) V^-1
// and tau = r_c t, the Frechet derivative of P = exp(tau Q) is
//   dP = V (F o (V^-1 dQ V)) V^-1,  F_kl = (e^{tau lam_k} - e^{tau lam_l}) /
//        (lam_k - lam_l)  (tau e^{tau lam_k} when lam_k = lam_l),
// so dlogL/dQ = V^-T H V^T with
//   H = sum over branches, categories c and patterns p of
//       w_c (w_p / L_p) F o ((V^T u)(V^-1 x)^T),
// u the branch's top vector pre(v) * (P_sib x_sib) and x its post-order
// operand.  A node job v (children a, b) gives both branches: u_a = pre(v)
// * P_b x_b, u_b = pre(v) * P_a x_a, and L_p = sum_c w_c sum_i pre(v)_i
// (P_a x_a)_i (P_b x_b)_i.  Each lane accumulates H (FP64) for its pattern
// and category over every job; scale factors cancel in u x^T / L_p.  The
// jobs are independent (all pre-order partials are computed), so a
// workgroup walks the list with the next job's loads in flight.

template <typename R>
struct QOps4 {
    R u[4], xa[4], xb[4];
    int sa, sb;
    typename Vec16<R>::T mv;   // this lane's share of the job's matrices
    dvec2 fv;                  // this lane's share of the job's F values
};

template <typename R>
__device__ __forceinline__ void q_issue(const NodeJob& J, int jidx, const void* dx, const void* ds,
                                        const R* __restrict__ jm, const double* __restrict__ jf,
                                        int c, int C, unsigned off, unsigned doff, int p, int lane,
                                        QOps4<R>& o) {
    const bool pa = J.k[0] == 0, pb = J.k[1] == 0;
    load4(at<R>(J.parent, off), o.u);
    load4_range<R>(range_if(J.x[0], pa, jm), off, o.xa);
    load4_range<R>(range_if(J.x[1], pb, jm), off, o.xb);
    o.sa = (int)__builtin_amdgcn_raw_buffer_load_b32(range_if(J.x[0], !pa, jm), 4u * p, 0, 0);
    o.sb = (int)__builtin_amdgcn_raw_buffer_load_b32(range_if(J.x[1], !pb, jm), 4u * p, 0, 0);
    constexpr int V = Vec16<R>::N;
    const int l = min(lane, NODE_MATS / V - 1);
    o.mv = *at<typename Vec16<R>::T>(jm, (unsigned)((J.mbase + c * NODE_MATS + l * V) *
                                                    (int)sizeof(R)));
    // lanes 0-7: child a's 16 factors of category c, lanes 8-15: child b's
    const int fl = min(lane, 15);
    o.fv = *at<dvec2>(jf, (unsigned)((J.mF[fl >> 3] + c * 16 + 2 * (fl & 7)) * 8));
}

template <typename R>
__device__ __forceinline__ void q_compute(const NodeJob& J, const QOps4<R>& o, double (&H)[16],
                                          R* smat, double* sfb, double* sx,
                                          const double* sV, int& k, int c, int lane, int p,
                                          int P, int C, const double* __restrict__ cw, double w) {
    const bool in = p < P;
    R xa[5], xb[5], ya[4], yb[4];
    child5(J.k[0], o.sa, o.xa, xa);
    child5(J.k[1], o.sb, o.xb, xb);
    R* lm = smat + (c * 2 + (k & 1)) * (NODE_MAT_BLOCK / (int)sizeof(R));
    double* lf = sfb + (c * 2 + (k & 1)) * 32;
    *(typename Vec16<R>::T*)(lm + lane * Vec16<R>::N) = o.mv;
    *(dvec2*)(lf + 2 * min(lane, 15)) = o.fv;   // lanes past 15 repeat lane 15's pair
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    mat4(lm, xa, ya);        // P_a x_a
    mat4(lm + 20, xb, yb);   // P_b x_b
    double uab[4], dc = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uab[i] = (double)o.u[i] * (double)ya[i] * (double)yb[i];
        dc += uab[i];
    }
    // L_p: category sums exchanged through LDS
    double* buf = sx + (k & 1) * C * TILE;
    ++k;
    buf[c * TILE + lane] = cw[c] * dc;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    double D = 0.0;
    for (int cc = 0; cc < C; ++cc) D += buf[cc * TILE + lane];
    const double coef = in ? cw[c] * w / D : 0.0;
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
        const R(&x)[5] = ch == 0 ? xa : xb;
        const R(&y)[4] = ch == 0 ? yb : ya;   // the sibling's P x
        double top[4], xv[4], al[4], be[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            top[i] = (double)o.u[i] * (double)y[i];
            // a missing tip state reads the pad column, which does not
            // depend on Q: its one-hot part is zero, so it adds nothing
            xv[i] = (double)x[i];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            al[q] = coef * (sV[0 * 4 + q] * top[0] + sV[1 * 4 + q] * top[1] +
                            sV[2 * 4 + q] * top[2] + sV[3 * 4 + q] * top[3]);
            be[q] = sV[16 + q * 4 + 0] * xv[0] + sV[16 + q * 4 + 1] * xv[1] +
                    sV[16 + q * 4 + 2] * xv[2] + sV[16 + q * 4 + 3] * xv[3];
        }
        const double* F = lf + ch * 16;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) H[q * 4 + r] += F[q * 4 + r] * al[q] * be[r];
    }
}

template <typename R>
__global__ __launch_bounds__(512) void k_qgrad4(const NodeJob* __restrict__ jobs, int njobs, int P,
                                                int C, long cs, const R* __restrict__ jm,
                                                const double* __restrict__ jf,
                                                const double* __restrict__ eig,
                                                const R* dummyX, const int* dummyS,
                                                const double* __restrict__ cw,
                                                const double* __restrict__ pw, double* partial) {
    // LDS: [2][C][TILE] exchange, V | V^-1 (32), [C][2][32] F blocks, then
    // [C][2] matrix blocks
    extern __shared__ double sx[];
    double* sV = sx + 2 * C * TILE;
    double* sfb = sV + 32;
    R* smat = reinterpret_cast<R*>(sfb + C * 2 * 32);
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x;
    const int p = tile * TILE + lane;
    if (threadIdx.x < 32) sV[threadIdx.x] = eig[threadIdx.x];
    __syncthreads();
    const unsigned off = (unsigned)((c * cs + (long)p * 4) * (long)sizeof(R));
    const unsigned doff = (unsigned)((c * TILE + lane) * 4 * (int)sizeof(R));
    const double w = pw[p];
    double H[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) H[q] = 0.0;
    int k = 0;
    // this workgroup's share of the (independent) jobs: chunk blockIdx.y
    const int chunk = (njobs + gridDim.y - 1) / gridDim.y;
    jobs += blockIdx.y * chunk;
    njobs = min(chunk, njobs - (int)blockIdx.y * chunk);
    QOps4<R> S0, S1;
    // a chunk past the end of the list (at most the last) only writes zeros
    NodeJob J0, J1;
    if (njobs > 0) {
        J0 = jobs[0];
        q_issue<R>(J0, 0, dummyX, dummyS, jm, jf, c, C, off, doff, p, lane, S0);
    }
    for (int o = 0; njobs > 0; o += 2) {
        const int o1 = min(o + 1, njobs - 1);
        J1 = jobs[o1];
        q_issue<R>(J1, o1, dummyX, dummyS, jm, jf, c, C, off, doff, p, lane, S1);
        q_compute<R>(J0, S0, H, smat, sfb, sx, sV, k, c, lane, p, P, C, cw, w);
        if (o + 1 >= njobs) break;
        const int o2 = min(o + 2, njobs - 1);
        J0 = jobs[o2];
        q_issue<R>(J0, o2, dummyX, dummyS, jm, jf, c, C, off, doff, p, lane, S0);
        q_compute<R>(J1, S1, H, smat, sfb, sx, sV, k, c, lane, p, P, C, cw, w);
        if (o + 2 >= njobs) break;
    }
    // per-(tile, chunk, category) sums of H: rows kl of [16][tiles * chunks * C]
    const long row = (long)gridDim.x * gridDim.y * C;
    const long col = ((long)blockIdx.y * gridDim.x + tile) * C + c;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const double s = wave_sum_dpp(H[q]);
        if (lane == 0) partial[q * row + col] = s;
    }
}

// The same sums with one wave per 16 patterns holding all four categories
// (lane = 16 c + j, pattern 16 w + j of wave w): the per-pattern likelihood
// L_p = sum_c w_c sum_i u_i (P_a x_a)_i (P_b x_b)_i is two lane exchanges
// (lanes 16 and 32 apart) instead of an LDS exchange and a workgroup
// barrier per job, so waves run independently; V and V^-1 are wave-uniform
// (scalar registers).  Each wave walks one chunk of the (independent) jobs
// with the next job's loads in flight: its pattern operands into the other
// of two register sets (the loop is unrolled by two), and its matrices P_a,
// P_b (all categories, [2][C][4][5]) and Frechet factors ([2][C][16]),
// loaded by the wave's lanes together, staged through a wave-private LDS
// slot when the job starts (LDS reads wait on their own counter, so the
// loads in flight are not waited for).  H is summed over the 64 lanes at
// the end.
template <typename R>
struct QWOps {
    R u[4], xa[4], xb[4];
    int sa, sb;
    typename Vec16<R>::T m0, m1;   // this lane's share of the job's matrices
    dvec2 f;                       // this lane's share of its Frechet factors
};

template <typename R>
__device__ __forceinline__ void qw_issue(const NodeJob& J, unsigned off, int p, int lane,
                                         const R* __restrict__ mats,
                                         const double* __restrict__ F, QWOps<R>& o) {
    constexpr int V = Vec16<R>::N, NV = 2 * 80 / V;   // vectors of the two matrices
    load4(at<R>(J.parent, off), o.u);
    if (J.k[0] == 1) o.sa = *at<int>(J.x[0], 4u * p);
    else load4(at<R>(J.x[0], off), o.xa);
    if (J.k[1] == 1) o.sb = *at<int>(J.x[1], 4u * p);
    else load4(at<R>(J.x[1], off), o.xb);
    // vectors l and (FP64: 80 vectors) l + 64 of [P_a | P_b] (each 80 reals,
    // all categories)
    const int l0 = min(lane, NV - 1);
    const int h0 = l0 < NV / 2 ? 0 : 1;
    o.m0 = *(const GLOBAL typename Vec16<R>::T*)(gin<R>(mats) + J.moP[h0] + (l0 - h0 * NV / 2) * V);
    if (NV > 64) {
        const int l1 = min(lane + 64, NV - 1);
        const int h1 = l1 < NV / 2 ? 0 : 1;
        o.m1 = *(const GLOBAL typename Vec16<R>::T*)(gin<R>(mats) + J.moP[h1] +
                                                      (l1 - h1 * NV / 2) * V);
    }
    // factors: lanes 0-31 child a's [C][16], lanes 32-63 child b's
    o.f = *(const GLOBAL dvec2*)(gin<double>(F) + J.mF[lane >> 5] + 2 * (lane & 31));
}

template <typename R>
__device__ __forceinline__ void qw_stage(const QWOps<R>& o, int lane, R* sm, double* sf) {
    constexpr int V = Vec16<R>::N, NV = 2 * 80 / V;
    *(typename Vec16<R>::T*)(sm + min(lane, NV - 1) * V) = o.m0;
    if (NV > 64) *(typename Vec16<R>::T*)(sm + min(lane + 64, NV - 1) * V) = o.m1;
    *(dvec2*)(sf + 2 * lane) = o.f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

template <typename R>
__device__ __forceinline__ void qw_compute(const NodeJob& J, const QWOps<R>& o, const R* sm,
                                           const double* sf, int c, bool in, double w,
                                           double wc, const double (&V)[16],
                                           const double (&Vi)[16], double (&H)[16]) {
    const bool tipA = J.k[0] == 1, tipB = J.k[1] == 1;
    const R* Pa = sm + c * 20;
    const R* Pb = sm + 80 + c * 20;
    // a tip reads the column of its state (column 4: the pad column of a
    // missing state); its own vector is one-hot, and a missing state's pad
    // column does not depend on Q, so it contributes nothing
    const int ca = min((unsigned)o.sa, 4u), cb = min((unsigned)o.sb, 4u);
    R ya[4], yb[4];   // P x, in the partials' precision as the products are formed
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        ya[i] = tipA ? Pa[i * 5 + ca]
                     : Pa[i * 5] * o.xa[0] + Pa[i * 5 + 1] * o.xa[1] + Pa[i * 5 + 2] * o.xa[2] +
                           Pa[i * 5 + 3] * o.xa[3];
        yb[i] = tipB ? Pb[i * 5 + cb]
                     : Pb[i * 5] * o.xb[0] + Pb[i * 5 + 1] * o.xb[1] + Pb[i * 5 + 2] * o.xb[2] +
                           Pb[i * 5 + 3] * o.xb[3];
    }
    double L = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) L += (double)o.u[i] * (double)ya[i] * (double)yb[i];
    L *= wc;
    L += __shfl_xor(L, 16, 64);
    L += __shfl_xor(L, 32, 64);
    const double coef = in ? w / L : 0.0;
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
        const bool tip = ch == 0 ? tipA : tipB;
        const int st = ch == 0 ? ca : cb;
        double al[4], be[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double a = 0.0, b = 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double top = (double)o.u[i] * (double)(ch == 0 ? yb[i] : ya[i]);
                a += V[i * 4 + q] * top;
                const double x = (double)(ch == 0 ? o.xa[i] : o.xb[i]);
                b += Vi[q * 4 + i] * (tip ? (st == i ? 1.0 : 0.0) : x);
            }
            al[q] = coef * a;
            be[q] = b;
        }
        const double* Fc = sf + ch * 64 + c * 16;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) H[q * 4 + r] += Fc[q * 4 + r] * (al[q] * be[r]);
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_qgrad4w(const NodeJob* __restrict__ jobs, int njobs,
                                                 int P, long cs, const R* __restrict__ mats,
                                                 const double* __restrict__ F,
                                                 const double* __restrict__ eig,
                                                 const double* __restrict__ cw,
                                                 const double* __restrict__ pw,
                                                 double* __restrict__ partial) {
    // LDS per wave: the job's matrices [2][4][20] reals, factors [2][4][16]
    __shared__ double sfb[4][128];
    __shared__ R smb[4][160];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane >> 4, j = lane & 15;
    const int wave = blockIdx.x * 4 + wv;
    const int nwaves = gridDim.x * 4;
    const int p = wave * 16 + j;            // < Ppad (a multiple of 64)
    const bool in = p < P;
    const double wc = cw[c];
    const double w = pw[p] * wc;            // pattern weights are zero past P
    const unsigned off = (unsigned)((c * cs + (long)p * 4) * (long)sizeof(R));
    R* sm = smb[wv];
    double* sf = sfb[wv];
    double V[16], Vi[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        V[q] = eig[q];
        Vi[q] = eig[16 + q];
    }
    double H[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) H[q] = 0.0;
    const int chunk = (njobs + gridDim.y - 1) / gridDim.y;
    const int o0 = blockIdx.y * chunk, o1 = min(njobs, o0 + chunk);
    QWOps<R> A, B;
    if (o0 < o1) qw_issue<R>(jobs[o0], off, p, lane, mats, F, A);
    for (int o = o0; o < o1; o += 2) {
        const NodeJob& Ja = jobs[o];
        qw_stage<R>(A, lane, sm, sf);
        if (o + 1 < o1) qw_issue<R>(jobs[o + 1], off, p, lane, mats, F, B);
        qw_compute<R>(Ja, A, sm, sf, c, in, w, wc, V, Vi, H);
        if (o + 1 >= o1) break;
        const NodeJob& Jb = jobs[o + 1];
        qw_stage<R>(B, lane, sm, sf);
        if (o + 2 < o1) qw_issue<R>(jobs[o + 2], off, p, lane, mats, F, A);
        qw_compute<R>(Jb, B, sm, sf, c, in, w, wc, V, Vi, H);
    }
    const long cols = (long)nwaves * gridDim.y;
    const long col = (long)blockIdx.y * nwaves + wave;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const double sum = wave_sum_dpp(H[q]);
        if (lane == 0) partial[q * cols + col] = sum;
    }
}

// t = M x with M row-major [4][5] (the model's layout), lane = pattern
__device__ __forceinline__ void mfma_rows5(const float* m, int r, const float (&x)[4],
                                           float (&t)[4]) {
    fvec4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = __builtin_amdgcn_mfma_f32_4x4x1f32(m[r * 5 + k], x[k], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = acc[i];
}

__device__ __forceinline__ void qm_compute(const NodeJob& J, const QWOps<float>& o,
                                           const float* sm, const double* sf, int c, int lane,
                                           bool in, float w, float wc, const QgVec& v,
                                           double (&Hr)[4]) {
    const bool tipA = J.k[0] == 1, tipB = J.k[1] == 1;
    const float* Pa = sm + c * 20;
    const float* Pb = sm + 80 + c * 20;
    const int r = lane & 3;
    const int ca = min((unsigned)o.sa, 4u), cb = min((unsigned)o.sb, 4u);
    float ya[4], yb[4];
    if (tipA) {
#pragma unroll
        for (int i = 0; i < 4; ++i) ya[i] = Pa[i * 5 + ca];
    } else {
        mfma_rows5(Pa, r, o.xa, ya);
    }
    if (tipB) {
#pragma unroll
        for (int i = 0; i < 4; ++i) yb[i] = Pb[i * 5 + cb];
    } else {
        mfma_rows5(Pb, r, o.xb, yb);
    }
    float L = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) L += o.u[i] * ya[i] * yb[i];
    L = cat_sum(wc * L);
    const float coef = in ? (w * wc) / L : 0.0f;
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
        const bool tip = ch == 0 ? tipA : tipB;
        const int st = ch == 0 ? ca : cb;
        float ta[4], x[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            // a tip's vector is one-hot; a missing state (4) reads the pad
            // column, which does not depend on Q: zero; lanes past the
            // patterns add nothing
            ta[i] = in ? coef * (o.u[i] * (ch == 0 ? yb[i] : ya[i])) : 0.0f;
            const float xv = ch == 0 ? o.xa[i] : o.xb[i];
            x[i] = !in ? 0.0f : tip ? (st == i ? 1.0f : 0.0f) : xv;
        }
        qgrad_block(ta, x, v, sf + ch * 64 + c * 16 + r * 4, Hr);
    }
}

// grid (tiles of 64 patterns, chunks of the job list), 4 waves of 16
// patterns x 4 categories; partial rows [16][cols], column (chunk, wave)
__global__ __launch_bounds__(256, 4) void k_qgrad4m(const NodeJob* __restrict__ jobs, int njobs,
                                                 int P, long cs, const float* __restrict__ mats,
                                                 const double* __restrict__ F,
                                                 const double* __restrict__ eig,
                                                 const double* __restrict__ cw,
                                                 const double* __restrict__ pw,
                                                 double* __restrict__ partial) {
    __shared__ double sfb[4][128];
    __shared__ float smb[4][160];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane >> 4, j = lane & 15;
    const int wave = blockIdx.x * 4 + wv;
    const int nwaves = gridDim.x * 4;
    const int p = wave * 16 + j;            // < Ppad (a multiple of 64)
    const bool in = p < P;
    const float wc = (float)cw[c];
    const float w = (float)pw[p];           // pattern weights are zero past P
    const unsigned off = (unsigned)((c * cs + (long)p * 4) * (long)sizeof(float));
    float* sm = smb[wv];
    double* sf = sfb[wv];
    QgVec v;
    qgrad_vectors(eig, lane & 3, v);
    double Hr[4] = {0.0, 0.0, 0.0, 0.0};
    const int chunk = (njobs + gridDim.y - 1) / gridDim.y;
    const int o0 = blockIdx.y * chunk, o1 = min(njobs, o0 + chunk);
    QWOps<float> A, B;
    if (o0 < o1) qw_issue<float>(jobs[o0], off, p, lane, mats, F, A);
    for (int o = o0; o < o1; o += 2) {
        const NodeJob& Ja = jobs[o];
        qw_stage<float>(A, lane, sm, sf);
        if (o + 1 < o1) qw_issue<float>(jobs[o + 1], off, p, lane, mats, F, B);
        qm_compute(Ja, A, sm, sf, c, lane, in, w, wc, v, Hr);
        if (o + 1 >= o1) break;
        const NodeJob& Jb = jobs[o + 1];
        qw_stage<float>(B, lane, sm, sf);
        if (o + 2 < o1) qw_issue<float>(jobs[o + 2], off, p, lane, mats, F, A);
        qm_compute(Jb, B, sm, sf, c, lane, in, w, wc, v, Hr);
    }
    // lanes with the same lane % 4 hold the same entries of H
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int m = 4; m < 64; m <<= 1) Hr[i] += __shfl_xor(Hr[i], m, 64);
    }
    const long cols = (long)nwaves * gridDim.y;
    const long col = (long)blockIdx.y * nwaves + wave;
    if (lane < 4) {
#pragma unroll
        for (int i = 0; i < 4; ++i) partial[(lane * 4 + i) * cols + col] = Hr[i];
    }
}

// Chunks of the job list the Q-gradient kernels spread over the grid's
// second dimension: about 16 waves per SIMD (1,024 SIMDs), at least 16 jobs
// per chunk.
int qgrad_chunks(const Dims& d, int njobs) {
    const int tiles = (d.P + TILE - 1) / TILE;
    const int waves = d.C == 4 ? 4 * tiles : tiles * d.C;
    const int want = (16384 + waves - 1) / waves;
    return std::max(1, std::min(want, (njobs + 15) / 16));
}

int qgrad_partial_cols(const Dims& d, int njobs) {
    const int tiles = (d.P + TILE - 1) / TILE;
    const int chunks = qgrad_chunks(d, njobs);
    return d.C == 4 ? 4 * tiles * chunks : tiles * chunks * d.C;
}

template <typename R>
void launch_qgrad(const NodeJob* djobs, int njobs, const Dims& d, const void* jobMats,
                  const void* mats, const double* jobF, const double* eig, void* dummy,
                  const double* cw, const double* pw, double* partial, hipStream_t st) {
    if (njobs <= 0) return;
    if (d.C == 4) {
        const int tiles = (d.P + TILE - 1) / TILE;
        if (std::is_same<R, float>::value) {
            hipLaunchKernelGGL(k_qgrad4m, dim3(tiles, qgrad_chunks(d, njobs)), dim3(256), 0, st,
                               djobs, njobs, d.P, d.catStride,
                               reinterpret_cast<const float*>(mats), jobF, eig, cw, pw, partial);
            return;
        }
        hipLaunchKernelGGL((k_qgrad4w<R>), dim3(tiles, qgrad_chunks(d, njobs)), dim3(256), 0, st,
                           djobs, njobs, d.P, d.catStride, reinterpret_cast<const R*>(mats), jobF,
                           eig, cw, pw, partial);
        return;
    }
    const R* dummyX = reinterpret_cast<const R*>(dummy);
    R* dummyW = const_cast<R*>(dummyX) + 8 * TILE * 4;
    const double* dummyD = reinterpret_cast<const double*>(dummyW + 8 * TILE * 4);
    const int* dummyS = reinterpret_cast<const int*>(dummyD + TILE);
    const int tiles = (d.P + TILE - 1) / TILE;
    const size_t lds = (size_t)(2 * d.C * TILE + 32 + d.C * 2 * 32) * sizeof(double) +
                       (size_t)d.C * 2 * NODE_MAT_BLOCK;
    const int chunks = qgrad_chunks(d, njobs);
    hipLaunchKernelGGL((k_qgrad4<R>), dim3(tiles, chunks), dim3(TILE * d.C), lds, st, djobs, njobs, d.P,
                       d.C, d.catStride, reinterpret_cast<const R*>(jobMats), jobF, eig, dummyX,
                       dummyS, cw, pw, partial);
}

// Sum of the node pass's H slots ([n][16] doubles) in two fixed-order
// steps: blocks of 16 slots x 16 entries, each thread adding at most
// kSlotSumLoads slots with its loads issued together; then one block per
// entry adds the block sums.
constexpr int kSlotSumLoads = 8;

__global__ __launch_bounds__(256) void k_slot_sums16(const double* __restrict__ in, long n,
                                                     double* __restrict__ part) {
    __shared__ double sh[16][16];
    const int j = threadIdx.x & 15, r = threadIdx.x >> 4;
    double x[kSlotSumLoads];
#pragma unroll
    for (int q = 0; q < kSlotSumLoads; ++q) {
        const long i = ((long)blockIdx.x * kSlotSumLoads + q) * 16 + r;
        x[q] = i < n ? in[i * 16 + j] : 0.0;
    }
    double v = x[0];
#pragma unroll
    for (int q = 1; q < kSlotSumLoads; ++q) v += x[q];
    sh[r][j] = v;
    __syncthreads();
    if (threadIdx.x < 16) {
        double s = sh[0][j];
        for (int q = 1; q < 16; ++q) s += sh[q][j];
        part[blockIdx.x * 16 + j] = s;
    }
}

__global__ __launch_bounds__(256) void k_slot_column_sums(const double* __restrict__ part,
                                                          int nb, double* __restrict__ out) {
    __shared__ double sh[4];
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) v += part[(long)b * 16 + blockIdx.x];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

void launch_qgrad_slot_sums(const double* slots, long n, double* part, double* out,
                            hipStream_t st) {
    const int nb = (int)std::max<long>(1, (n + 16 * kSlotSumLoads - 1) / (16 * kSlotSumLoads));
    hipLaunchKernelGGL(k_slot_sums16, dim3(nb), dim3(256), 0, st, slots, n, part);
    hipLaunchKernelGGL(k_slot_column_sums, dim3(16), dim3(256), 0, st, part, nb, out);
}

// dlogL/dQ = V^-T H V^T (one wave), the root-frequency sums copied after it.
__global__ __launch_bounds__(64) void k_qgrad_finish(const double* __restrict__ sums,
                                                     const double* __restrict__ eig, int nf,
                                                     double* __restrict__ out) {
    const int t = threadIdx.x;
    if (t < 16) {
        const int a = t >> 2, b = t & 3;
        const double* V = eig;
        const double* Vi = eig + 16;
        double g = 0.0;
        for (int k = 0; k < 4; ++k)
            for (int l = 0; l < 4; ++l) g += Vi[k * 4 + a] * sums[k * 4 + l] * V[b * 4 + l];
        out[t] = g;
    } else if (t < 16 + nf) {
        out[t] = sums[t];
    }
}

void launch_qgrad_finish(const double* sums, const double* eig, int nf, double* out,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_qgrad_finish, dim3(1), dim3(64), 0, st, sums, eig, nf, out);
}

// dlogL/dpi_i through the root distribution: sum_p w_p (sum_c w_c R_cpi) /
// (sum_i pi_i sum_c w_c R_cpi), one thread per pattern, block sums per i
// (rows of [S][blocks]).  Scale factors of the root partials cancel.
template <typename R>
__global__ __launch_bounds__(WG) void k_root_freq_grad(const R* __restrict__ root,
                                                        const double* __restrict__ cw,
                                                        const double* __restrict__ freqs,
                                                        const double* __restrict__ pw,
                                                        double* __restrict__ partial, Dims d) {
    __shared__ double sh[WG / 64];
    const int p = blockIdx.x * WG + threadIdx.x;
    const int nb = gridDim.x;
    double L = 0.0;
    if (p < d.P) {
        for (int i = 0; i < d.S; ++i) {
            double g = 0.0;
            for (int c = 0; c < d.C; ++c) g += cw[c] * (double)root[pbase(d, c, p) + (long)i * d.sStride];
            L += freqs[i] * g;
        }
    }
    for (int i = 0; i < d.S; ++i) {
        double v = 0.0;
        if (p < d.P) {
            double g = 0.0;
            for (int c = 0; c < d.C; ++c) g += cw[c] * (double)root[pbase(d, c, p) + (long)i * d.sStride];
            v = pw[p] * g / L;
        }
        const double s = block_sum(v, sh);
        if (threadIdx.x == 0) partial[(long)i * nb + blockIdx.x] = s;
    }
}

template <typename R>
void launch_root_freq_grad(const void* root, const double* cw, const double* freqs,
                           const double* pw, double* partial, const Dims& d, hipStream_t st) {
    const int nb = (d.P + WG - 1) / WG;
    hipLaunchKernelGGL(k_root_freq_grad<R>, dim3(nb), dim3(WG), 0, st,
                       reinterpret_cast<const R*>(root), cw, freqs, pw, partial, d);
}

// The same sums for many states: one wave per 64 patterns and group of 16
// states, grid (ceil(P / 64), ceil(S / 16)); every lane forms its pattern's
// likelihood, then each of the group's states is summed over the wave by
// DPP steps (no workgroup barrier per state).  Rows i of [S][ceil(P / 64)].
template <typename R>
__global__ __launch_bounds__(64) void k_root_freq_grad_wide(const R* __restrict__ root,
                                                            const double* __restrict__ cw,
                                                            const double* __restrict__ freqs,
                                                            const double* __restrict__ pw,
                                                            double* __restrict__ partial, Dims d) {
    const int p = blockIdx.x * 64 + threadIdx.x;
    const bool in = p < d.P;
    const int q = in ? p : 0;   // lanes past the patterns read pattern 0 and add 0
    double L = 0.0;
    for (int i = 0; i < d.S; ++i) {
        double g = 0.0;
        for (int c = 0; c < d.C; ++c) g += cw[c] * (double)root[pbase(d, c, q) + (long)i * d.sStride];
        L += freqs[i] * g;
    }
    const int i1 = min(d.S, (int)(blockIdx.y + 1) * 16);
    for (int i = blockIdx.y * 16; i < i1; ++i) {
        double g = 0.0;
        for (int c = 0; c < d.C; ++c) g += cw[c] * (double)root[pbase(d, c, q) + (long)i * d.sStride];
        const double s = wave_sum_dpp(in ? pw[q] * g / L : 0.0);
        if (threadIdx.x == 0) partial[(long)i * gridDim.x + blockIdx.x] = s;
    }
}

int root_freq_wide_blocks(const Dims& d) { return (d.P + 63) / 64; }

template <typename R>
void launch_root_freq_grad_wide(const void* root, const double* cw, const double* freqs,
                                const double* pw, double* partial, const Dims& d,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_root_freq_grad_wide<R>, dim3(root_freq_wide_blocks(d), (d.S + 15) / 16),
                       dim3(64), 0, st, reinterpret_cast<const R*>(root), cw, freqs, pw, partial,
                       d);
}

template <typename R>
__global__ __launch_bounds__(WG) void k_gather_node_mats(const NodeJob* __restrict__ jobs,
                                                         int njobs, int C,
                                                         const R* __restrict__ mats,
                                                         R* __restrict__ out, bool colMajor,
                                                         const double* __restrict__ F) {
    const int stride = colMajor ? NODE_JM : NODE_MATS;   // reals of a category's block
    const long t = (long)blockIdx.x * WG + threadIdx.x;
    const long per = (long)C * stride;
    if (t >= per * njobs) return;
    const int j = (int)(t / per);
    const int rem = (int)(t - (long)j * per);
    const int c = rem / stride, r = rem - c * stride;
    const NodeJob& J = jobs[j];
    R v = (R)0;
    int d = r;
    if (r < NODE_MATS) {
        const int m = r / 20, e = r - m * 20;
        // row-major [4][5] as stored, or column-major [5][4]
        d = colMajor ? m * 20 + (e % 5) * 4 + e / 5 : r;
        const int mo = m == 0 ? J.moP[0] : m == 1 ? J.moP[1] : m == 2 ? J.moD1[0]
                     : m == 3 ? J.moD2[0] : m == 4 ? J.moD1[1] : J.moD2[1];
        v = mats[(long)mo + c * 20 + e];
    } else if (r >= NODE_F32_BLOCK && F) {
        // the branches' Frechet factors of category c, [child][16]
        const int f = r - NODE_F32_BLOCK;
        v = (R)F[(long)J.mF[f >> 4] + c * 16 + (f & 15)];
    }
    out[(long)J.mbase + c * stride + d] = v;
}

template <typename R>
void launch_gather_node_matrices(const NodeJob* djobs, int njobs, const Dims& d,
                                 const void* mats, void* jobMats, hipStream_t st, bool colMajor,
                                 const double* F) {
    const long n = (long)njobs * d.C * (colMajor ? NODE_JM : NODE_MATS);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_node_mats<R>, dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, st,
                       djobs, njobs, d.C, reinterpret_cast<const R*>(mats),
                       reinterpret_cast<R*>(jobMats), colMajor, F);
}

template <typename R>
void launch_pre_nodes(const NodeJob* djobs, int njobs, bool batched, int chunk, const Dims& d,
                      const void* jobMats, void* dummy, const double* cw, const double* pw,
                      double* partial, hipStream_t st, NodeQGrad* q) {
    if (njobs <= 0) return;
    const R* dummyX = reinterpret_cast<const R*>(dummy);
    R* dummyW = const_cast<R*>(dummyX) + 8 * TILE * 4;
    const double* dummyD = reinterpret_cast<const double*>(dummyW + 8 * TILE * 4);
    const int* dummyS = reinterpret_cast<const int*>(dummyD + TILE);
    const short* dummyE = reinterpret_cast<const short*>(dummyS + TILE);
    const R* m = reinterpret_cast<const R*>(jobMats);
    if (std::is_same<R, float>::value && node_f32_supported(d.C)) {
        const float* mf = reinterpret_cast<const float*>(m);
        if (q) {
            // a pass never needs more slots than jobs x tiles
            if (q->used + (long)njobs * ((d.P + TILE - 1) / TILE) > q->capacity)
                throw std::runtime_error("rate-matrix gradient slots exhausted");
            if (d.C == 4) launch_pre_nodes_f32<4, true>(djobs, njobs, batched, chunk, d, mf, cw, pw, partial, st, q);
            else if (d.C == 2) launch_pre_nodes_f32<2, true>(djobs, njobs, batched, chunk, d, mf, cw, pw, partial, st, q);
            else launch_pre_nodes_f32<1, true>(djobs, njobs, batched, chunk, d, mf, cw, pw, partial, st, q);
        } else {
            if (d.C == 4) launch_pre_nodes_f32<4, false>(djobs, njobs, batched, chunk, d, mf, cw, pw, partial, st, q);
            else if (d.C == 2) launch_pre_nodes_f32<2, false>(djobs, njobs, batched, chunk, d, mf, cw, pw, partial, st, q);
            else launch_pre_nodes_f32<1, false>(djobs, njobs, batched, chunk, d, mf, cw, pw, partial, st, q);
        }
        return;
    }
    if (q) throw std::runtime_error("the rate-matrix gradient pass needs FP32 with 1, 2 or 4 categories");
    const int tiles = (d.P + TILE - 1) / TILE;
    const size_t lds = (size_t)2 * d.C * 8 * TILE * sizeof(double) +
                       (size_t)d.C * 2 * NODE_MAT_BLOCK;
    dim3 block(TILE * d.C);
    if (batched && chunk <= 1) {
        if (d.autoScale) hipLaunchKernelGGL((k_pren4_level<R, true>), dim3(tiles, njobs), block, lds, st, djobs, d.P, d.C, d.catStride, m, dummyX, dummyS, dummyW, dummyE, cw, pw, partial);
        else hipLaunchKernelGGL((k_pren4_level<R, false>), dim3(tiles, njobs), block, lds, st, djobs, d.P, d.C, d.catStride, m, dummyX, dummyS, dummyW, dummyE, cw, pw, partial);
    } else {
        // a fused list (chunk = the whole list) or a batched level walked in
        // chunks of independent jobs
        if (!batched || chunk > njobs) chunk = njobs;
        const dim3 grid(tiles, (njobs + chunk - 1) / chunk);
        if (d.autoScale) hipLaunchKernelGGL((k_pren4<R, true>), grid, block, lds, st, djobs, njobs, chunk, d.P, d.C, d.catStride, m, dummyX, dummyS, dummyW, dummyE, cw, pw, partial);
        else hipLaunchKernelGGL((k_pren4<R, false>), grid, block, lds, st, djobs, njobs, chunk, d.P, d.C, d.catStride, m, dummyX, dummyS, dummyW, dummyE, cw, pw, partial);
    }
}


#define INSTANTIATE(R)                                                                          \
    template void launch_gather_node_matrices<R>(const NodeJob*, int, const Dims&, const void*, \
                                                 void*, hipStream_t, bool, const double*);                     \
    template void launch_qgrad<R>(const NodeJob*, int, const Dims&, const void*, const void*,   \
                                  const double*, const double*, void*, const double*,           \
                                  const double*, double*, hipStream_t);                         \
    template void launch_root_freq_grad<R>(const void*, const double*, const double*,          \
                                           const double*, double*, const Dims&, hipStream_t);     \
    template void launch_root_freq_grad_wide<R>(const void*, const double*, const double*,     \
                                                const double*, double*, const Dims&, hipStream_t);   \
    template void launch_pre_nodes<R>(const NodeJob*, int, bool, int, const Dims&, const void*, \
                                      void*, const double*, const double*, double*,             \
                                      hipStream_t, NodeQGrad*);

INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace hipbeagle
