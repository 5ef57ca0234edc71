// gfx950 kernels of the MI355X BEAGLE plugin.
//
// Partials live in HBM as [category][pattern (padded to Ppad)][state] for
// S <= 4, so consecutive lanes touch consecutive patterns and every 4-state
// load/store is a full 16-byte vector access; for S > 4 as pattern-minor
// tiles [category][pattern tile][state][32 patterns] (kernels.h, pbase).  Transition matrices
// keep the reference CPU layout [category][row i][S+1 columns] with the pad
// column that a missing tip state reads (1 for P(t), 0 for derivatives).
// Scale buffers and every reduction are FP64 whatever the partials
// precision, except the FP32 node kernels' per-tile derivative sums (FP32
// within a 64-pattern tile, FP64 across tiles).
//
// The 4-state traversals are fused: one launch walks the whole operation
// list, each lane owning one pattern (all categories) from the first
// operation to the last.  Patterns are independent, so per-lane program
// order is the only dependency a traversal has; children written earlier
// in the same launch are re-read by the lane that wrote them, from L2.
// Tree shape (a 127-deep caterpillar or a balanced tree) does not change
// the launch count.
#include "kernels.h"
#include "device_util.h"

#include <cmath>
#include <cstdlib>
#include <type_traits>

namespace hipbeagle {

// ---------------------------------------------------------------------------
// small helpers

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

// out[j] = sum over n of in[j*stride + n]-style partials: one workgroup per
// output, strided lane sums then a fixed tree.
__global__ __launch_bounds__(WG) void k_reduce_rows(const double* __restrict__ in, int n,
                                                    int rows, double* __restrict__ out,
                                                    int accumulate) {
    __shared__ double sh[WG / 64];
    const int row = blockIdx.x;
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += WG) v += in[(long)row * n + i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) {
        if (accumulate) out[row] += v;
        else out[row] = v;
    }
}

// ---------------------------------------------------------------------------
// transition matrices: P = clamp(sum_k C_ijk e^(lambda_k r t)), dP, d2P
// (EigenDecompositionCube.hpp:105-322 semantics; computed in FP64, stored in
// the instance precision).

// state counts whose V and V^-1 fit the 64 KB of LDS a workgroup may use
constexpr int kMatLdsMaxS = 62;

// Complex eigen-system (EigenDecompositionSquare.hpp:108-185): the rows of
// V^-1 are scaled by e^(a d) for a real eigenvalue a, or taken in conjugate
// pairs (a +- ib, b the imaginary part of the pair's first row) and rotated,
//   row_i' = e^(a d) (cos(b d) row_i + sin(b d) row_i+1),
//   row_i+1' = e^(a d) (cos(b d) row_i+1 - sin(b d) row_i),
// then P = V T, clamped at 0, pad column 1.  Like the reference, no
// derivative matrices.  S <= kMatLdsMaxS (V and T staged in LDS).
template <typename R>
__device__ void matrices_complex(const MatrixJob& J, int S, double r, int c, double* sh) {
    double* ex = sh;
    double* ec = sh + S;
    double* es = sh + 2 * S;
    double* V = sh + 3 * S;
    double* T = V + S * S;
    int* kind = reinterpret_cast<int*>(T + S * S);   // 0 real, 1 first of a pair, 2 second
    const double d = J.t * r;
    for (int k = threadIdx.x; k < S; k += WG) {
        const double e = exp(J.evals[k] * d), b = J.evals[S + k] * d;
        ex[k] = e;
        ec[k] = e * cos(b);
        es[k] = e * sin(b);
    }
    for (int t = threadIdx.x; t < S * S; t += WG) {
        V[t] = J.vv[t];
        T[t] = J.vv[S * S + t];
    }
    if (threadIdx.x == 0) {
        // pairs are found in row order, as the reference walks them (a
        // nonzero imaginary part on the last row has no partner: real)
        for (int i = 0; i < S;) {
            if (J.evals[S + i] != 0.0 && i + 1 < S) {
                kind[i] = 1;
                kind[i + 1] = 2;
                i += 2;
            } else {
                kind[i++] = 0;
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < S * S; t += WG) {
        const int i = t / S, j = t - i * S;
        if (kind[i] == 0) {
            T[t] *= ex[i];
        } else if (kind[i] == 1) {   // this thread owns (i, j) and (i + 1, j)
            const double a = T[t], b = T[t + S];
            T[t] = ec[i] * a + es[i] * b;
            T[t + S] = ec[i] * b - es[i] * a;
        }
    }
    __syncthreads();
    const int W = S + 1;
    R* Pm = reinterpret_cast<R*>(J.P) + (long)c * S * W;
    for (int t = threadIdx.x; t < S * S; t += WG) {
        const int i = t / S, j = t - i * S;
        double s0 = 0.0;
        for (int k = 0; k < S; ++k) s0 += V[i * S + k] * T[k * S + j];
        Pm[i * W + j] = (R)(s0 > 0.0 ? s0 : 0.0);
    }
    for (int i = threadIdx.x; i < S; i += WG) Pm[i * W + S] = (R)1.0;
}

template <typename R>
__global__ __launch_bounds__(WG) void k_matrices(const MatrixJob* __restrict__ jobs, int S, int C) {
    extern __shared__ double sh[];
    double* ex = sh;
    double* e1 = sh + S;
    double* e2 = sh + 2 * S;
    const MatrixJob J = jobs[blockIdx.x];
    if (J.cat >= 0 && blockIdx.y != 0) return;   // one category per job
    const int c = J.cat >= 0 ? J.cat : (int)blockIdx.y;
    const double r = J.cat >= 0 ? 1.0 : J.rates[c];
    if (J.cplx) {
        matrices_complex<R>(J, S, r, c, sh);
        return;
    }
    const bool deriv = J.D1 != nullptr;
    for (int k = threadIdx.x; k < S; k += WG) {
        const double lam = J.evals[k];
        const double sl = lam * r;
        // same association as the reference: lambda*(t*r) without
        // derivatives, (lambda*r)*t with them
        const double e = deriv ? exp(sl * J.t) : exp(lam * (J.t * r));
        ex[k] = e;
        e1[k] = sl * e;
        e2[k] = sl * e1[k];
    }
    if (J.F && threadIdx.x < 16) {
        // Frechet factors of exp(tau Q), tau = r t, for the rate-matrix
        // gradient: F_ab = (e^{tau l_a} - e^{tau l_b}) / (l_a - l_b), as
        // e^{tau l_b} expm1(tau d) / d (no cancellation near d = 0), and
        // tau e^{tau l} at d = 0
        const int a = threadIdx.x >> 2, b = threadIdx.x & 3;
        const double tau = r * J.t, la = J.evals[a], lb = J.evals[b], dl = la - lb;
        const double eb = exp(tau * lb);
        J.F[c * 16 + threadIdx.x] = dl == 0.0 ? tau * eb : eb * expm1(tau * dl) / dl;
    }
    __syncthreads();
    const int W = S + 1;
    R* Pm = reinterpret_cast<R*>(J.P) + (long)c * S * W;
    R* D1 = deriv ? reinterpret_cast<R*>(J.D1) + (long)c * S * W : nullptr;
    R* D2 = J.D2 ? reinterpret_cast<R*>(J.D2) + (long)c * S * W : nullptr;
    if (S <= kMatLdsMaxS) {
        // V and V^-1 staged in LDS; C_ijk = V_ik Vinv_kj formed on the fly
        // (the same product the cube holds, so results are identical)
        double* V = sh + 3 * S;
        double* Vi = V + S * S;
        for (int t = threadIdx.x; t < S * S; t += WG) {
            V[t] = J.vv[t];
            Vi[t] = J.vv[S * S + t];
        }
        __syncthreads();
        // four columns per thread: four independent sums share V_ik
        const int JG = (S + 3) / 4;
        for (int w = threadIdx.x; w < S * JG; w += WG) {
            const int i = w / JG, j0 = (w - i * JG) * 4;
            double s0[4] = {0.0, 0.0, 0.0, 0.0}, s1[4] = {0.0, 0.0, 0.0, 0.0},
                   s2[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
            for (int k = 0; k < S; ++k) {
                const double v = V[i * S + k];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double cv = j0 + u < S ? v * Vi[k * S + j0 + u] : 0.0;
                    s0[u] += cv * ex[k];
                    if (deriv) {
                        s1[u] += cv * e1[k];
                        s2[u] += cv * e2[k];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + u;
                if (j >= S) break;
                Pm[i * W + j] = (R)(s0[u] > 0.0 ? s0[u] : 0.0);
                if (D1) D1[i * W + j] = (R)s1[u];
                if (D2) D2[i * W + j] = (R)s2[u];
            }
        }
    } else {
        for (int idx = threadIdx.x; idx < S * S; idx += WG) {
            const int i = idx / S, j = idx - i * S;
            const double* cm = J.cube + (long)idx * S;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (int k = 0; k < S; ++k) {
                const double cv = cm[k];
                s0 += cv * ex[k];
                s1 += cv * e1[k];
                s2 += cv * e2[k];
            }
            Pm[i * W + j] = (R)(s0 > 0.0 ? s0 : 0.0);
            if (D1) D1[i * W + j] = (R)s1;
            if (D2) D2[i * W + j] = (R)s2;
        }
    }
    for (int i = threadIdx.x; i < S; i += WG) {
        Pm[i * W + S] = (R)1.0;
        if (D1) D1[i * W + S] = (R)0.0;
        if (D2) D2[i * W + S] = (R)0.0;
    }
}

template <typename R>
void launch_matrices(const MatrixJob* jobs, int njobs, int S, int C, hipStream_t st) {
    if (njobs <= 0) return;
    dim3 grid(njobs, C);
    const size_t lds = (3 * S + (S <= kMatLdsMaxS ? 2 * S * S : 0)) * sizeof(double) +
                       (S <= kMatLdsMaxS ? S * sizeof(int) : 0);   // complex: row kinds
    hipLaunchKernelGGL(k_matrices<R>, grid, dim3(WG), lds, st, jobs, S, C);
}

// out[c] = A[c] B[c] over S states (in the instance precision, as the
// reference sums in its REALTYPE), pad column 1 (BeagleCPUImpl.hpp:1096-1162).
template <typename R>
__global__ __launch_bounds__(WG) void k_convolve(const ConvolveJob* __restrict__ jobs, int S) {
    const ConvolveJob J = jobs[blockIdx.x];
    const int c = blockIdx.y, W = S + 1;
    const R* A = reinterpret_cast<const R*>(J.A) + (long)c * S * W;
    const R* B = reinterpret_cast<const R*>(J.B) + (long)c * S * W;
    R* O = reinterpret_cast<R*>(J.out) + (long)c * S * W;
    for (int idx = threadIdx.x; idx < S * S; idx += WG) {
        const int i = idx / S, j = idx - i * S;
        R sum = 0;
        for (int k = 0; k < S; ++k) sum += A[i * W + k] * B[k * W + j];
        O[i * W + j] = sum;
    }
    for (int i = threadIdx.x; i < S; i += WG) O[i * W + S] = (R)1.0;
}

template <typename R>
void launch_convolve(const ConvolveJob* jobs, int njobs, int S, int C, hipStream_t st) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_convolve<R>, dim3(njobs, C), dim3(WG), 0, st, jobs, S);
}

// ---------------------------------------------------------------------------
// 4-state fused traversals

// Workgroup layout of the 4-state traversals: C waves over one tile of 64
// patterns; wave w owns category w, lane l owns pattern tile*64+l.  A
// wave's matrices are wave-uniform (scalar loads), its partials loads and
// stores are 64 consecutive 16/32-byte vectors.  The per-pattern maximum
// over categories that rescaling needs is exchanged through LDS; only
// rescaled operations pay that barrier.
constexpr int TILE = 64;

// Cross-category per-pattern maximum; `k` counts rescaled operations so
// consecutive exchanges alternate LDS buffers and one barrier suffices.
template <typename R>
__device__ __forceinline__ R tile_max(R m, R (*smax)[8][TILE], int& k, int c, int lane, int C) {
    R (*buf)[TILE] = smax[k & 1];
    ++k;
    buf[c][lane] = m;
    // LDS-only release/acquire around the barrier: outstanding global stores
    // need not complete for this exchange
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    R mx = buf[0][lane];
    for (int cc = 1; cc < C; ++cc) mx = fmax(mx, buf[cc][lane]);
    return mx;
}


// Rescale one lane's four values given the tile maximum.  Manual modes
// (reference semantics): divide by max (0 -> 1), wave 0 stores max or
// log(max), adds log(max) to the cumulative buffer and, under
// scaling-always, the children's node factors.  Automatic mode: multiply
// by 2^-e where max = f 2^e, f in [0.5,1) (exact), and wave 0 stores the
// node's exponent e + e(child1) + e(child2).
template <typename R, bool AUTO>
__device__ __forceinline__ void rescale4(R (&v)[4], R mx, int c, int p, double* scaleWrite,
                                         int logScalers, double* cumulative, const double* accA,
                                         const double* accB, const short* e1, const short* e2,
                                         short* ed) {
    if (AUTO) {
        const int e = mx > (R)0 ? exponent_of(mx) : 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = pow2(v[i], -e);
        if (c == 0) {
            int t = e;
            if (e1) t += gin<short>(e1)[p];
            if (e2) t += gin<short>(e2)[p];
            gout<short>(ed)[p] = (short)t;
        }
    } else {
        if (mx == (R)0) mx = (R)1;
        const R inv = (R)1 / mx;
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] *= inv;
        if (c == 0) {
            const double lm = log((double)mx);
            double w = logScalers ? lm : (double)mx;
            if (cumulative) gout<double>(cumulative)[p] += lm;
            if (accA) w += gin<double>(accA)[p];
            if (accB) w += gin<double>(accB)[p];
            gout<double>(scaleWrite)[p] = w;
        }
    }
}

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

// A 4-state operand as held in registers between its load and its use:
// partials values, or a tip state (expanded at use to a one-hot vector whose
// fifth entry selects the matrix pad column, so a missing state reads the
// pad exactly as the reference does).
template <typename R>
struct Operand4 {
    R x[4];
    int st;
    int e;        // pipelined post-order: the operand's scale exponent
    double sr;    // pipelined post-order (first operand): the read-scale factor
    typename Vec16<R>::T mv;   // pipelined post-order (first operand): this
                               // lane's 16 bytes of the step's two matrices
};

template <typename R>
__device__ __forceinline__ void fetch4(const void* xp, const int* s, long off, int p,
                                       Operand4<R>& o) {
    if (s) o.st = gin<int>(s)[p];
    else load4(gin<R>(xp) + off, o.x);
}

// t_i = sum_j M[i][j] x_j over the 4 states (plus the pad column for a
// missing tip state); M is one category's 4 x 5 block.
template <typename R>
__device__ __forceinline__ void apply4(const R* m, bool isState, const Operand4<R>& o,
                                       R (&t)[4]) {
    if (isState) {
        const int st = o.st;
        const R x0 = st == 0, x1 = st == 1, x2 = st == 2, x3 = st == 3, x4 = st >= 4;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            t[i] = m[i * 5 + 0] * x0 + m[i * 5 + 1] * x1 + m[i * 5 + 2] * x2 +
                   m[i * 5 + 3] * x3 + m[i * 5 + 4] * x4;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            t[i] = m[i * 5 + 0] * o.x[0] + m[i * 5 + 1] * o.x[1] + m[i * 5 + 2] * o.x[2] +
                   m[i * 5 + 3] * o.x[3];
    }
}

// Post-order traversal, software-pipelined.  While operation o computes,
// the loads of operation o+1's operands are in flight.  To let the
// compiler wait only for those loads (gfx9 counts stores in the same
// vmcnt as loads, so a conservative vmcnt(0) would also wait for the
// previous operation's stores), every step issues the same loads into the
// same registers: unused operand slots read a one-tile dummy buffer that
// stays in cache, and the loop is unrolled by two with alternating register
// sets so no loaded register is ever copied.  Operand kinds are decided on
// the host: partials (0), tip states (1), or the previous operation's
// output forwarded from registers (2).
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

// One operand's loads, issued whatever its kind so that every step of a
// pipelined kernel issues the same loads: partials (kind 0) at element
// offset `off`, tip states (kind 1) at 4p bytes; the one the kind does not
// read comes from an empty range (0).  Partials buffers are under 4 GiB
// (byte offsets in 32 bits).
template <typename R>
__device__ __forceinline__ void issue4(const void* x, const int* s, int kind, const void* valid,
                                       long off, int p, Operand4<R>& o) {
    load4_range<R>(range_if(x, kind == 0, valid), (unsigned)(off * (long)sizeof(R)), o.x);
    o.st = (int)__builtin_amdgcn_raw_buffer_load_b32(range_if(s, kind == 1, valid),
                                                     4u * (unsigned)p, 0, 0);
}

// Operand value by kind, formed with selects (no branches, so no register
// copies of in-flight loads).
template <typename R>
__device__ __forceinline__ void operand4(int kind, const Operand4<R>& o, const R (&prev)[4],
                                         R (&x)[5]) {
    const int st = o.st;
    const bool isS = kind == 1, isF = kind == 2;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const R oh = st == j ? (R)1 : (R)0;
        x[j] = isS ? oh : (isF ? prev[j] : o.x[j]);
    }
    x[4] = (isS && st >= 4) ? (R)1 : (R)0;
}

// t_i = sum_j M[i][j] x_j with x_4 selecting the pad column
template <typename R>
__device__ __forceinline__ void mat4(const R* m, const R (&x)[5], R (&t)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
        t[i] = m[i * 5 + 0] * x[0] + m[i * 5 + 1] * x[1] + m[i * 5 + 2] * x[2] +
               m[i * 5 + 3] * x[3] + m[i * 5 + 4] * x[4];
}

// One pipelined step's loads: both operands and, with automatic scaling,
// their exponents (a partials operand's exponent array, else the dummy),
// without it the read-scale factor (the dummy unless SCALE_READ); issued a
// step ahead like the operands themselves.
// The step's two 4 x 5 matrices for category c (20 reals each) arrive as
// one 16-byte vector per lane: lanes [0, n) the first matrix, [n, 2n) the
// second, the rest repeat the last vector (n = 20 reals / 16 bytes).
template <typename R>
__device__ __forceinline__ void post_issue_mats(const PostJob& J, const R* __restrict__ mats,
                                                int c, int lane, Operand4<R>& A) {
    constexpr int V = Vec16<R>::N, n = 20 / V;
    const int l = min(lane, 2 * n - 1);
    const int mo = l < n ? J.mo1 : J.mo2;
    A.mv = *reinterpret_cast<const GLOBAL typename Vec16<R>::T*>(
        gin<R>(mats) + mo + c * 20 + (l < n ? l : l - n) * V);
}

template <typename R, bool AUTO>
__device__ __forceinline__ void post_issue(const PostJob& J, const void* dx, const int* ds,
                                           const double* dD, const short* dE, long off,
                                           long doff, int p, int lane, Operand4<R>& A,
                                           Operand4<R>& B) {
    issue4(J.x1, J.s1, J.k1, dx, off, p, A);
    issue4(J.x2, J.s2, J.k2, dx, off, p, B);
    if (AUTO) {
        const bool l1 = J.k1 == 0 && J.e1 != nullptr, l2 = J.k2 == 0 && J.e2 != nullptr;
        A.e = (short)__builtin_amdgcn_raw_buffer_load_b16(range_if(J.e1, l1, dx), 2u * p, 0, 0);
        B.e = (short)__builtin_amdgcn_raw_buffer_load_b16(range_if(J.e2, l2, dx), 2u * p, 0, 0);
    } else {
        const bool r = J.mode == SCALE_READ;
        const auto q = __builtin_amdgcn_raw_buffer_load_b64(range_if(J.scaleRead, r, dx), 8u * p,
                                                            0, 0);
        A.sr = __hiloint2double((int)q[1], (int)q[0]);
    }
}

// t = M x for one operand of a pipelined post-order step, M the 4 x 5
// block in LDS: a tip's state column (column 4, the pad column, for a
// missing state), read with the state as the lane's index; otherwise M
// times the loaded or forwarded partials, the pad column left out (it
// multiplies 0).  The same values as the one-hot product.
template <typename R>
__device__ __forceinline__ void step_product(int kind, const Operand4<R>& o, const R (&prev)[4],
                                             const R* m, R (&t)[4]) {
    if (kind == 1) {
        const int s = min((unsigned)o.st, 4u);
#pragma unroll
        for (int i = 0; i < 4; ++i) t[i] = m[i * 5 + s];
    } else {
        R x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = kind == 2 ? prev[j] : o.x[j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            t[i] = m[i * 5 + 0] * x[0] + m[i * 5 + 1] * x[1] + m[i * 5 + 2] * x[2] +
                   m[i * 5 + 3] * x[3];
    }
}

// Automatic (power-of-two) rescale inside the pipelined pre-order kernel:
// every wave of the tile loads both exponents and stores the result's (the
// same value from each wave), so the step's memory operations do not
// depend on the wave or on which operands are tips.
template <typename R>
__device__ __forceinline__ void auto_rescale4(R (&v)[4], R mx, const short* e1, const short* e2,
                                              short* ed, const short* dummyE, int p, int lane,
                                              bool in) {
    const int e = mx > (R)0 ? exponent_of(mx) : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = pow2(v[i], -e);
    const int x1 = e1 ? gin<short>(e1)[p] : gin<short>(dummyE)[lane];
    const int x2 = e2 ? gin<short>(e2)[p] : gin<short>(dummyE)[lane];
    short* dst = in ? ed + p : const_cast<short*>(dummyE) + 64 + lane;
    *gout<short>(dst) = (short)(e + (e1 ? x1 : 0) + (e2 ? x2 : 0));
}

// Automatic (power-of-two) rescale inside the pipelined post-order kernel:
// every wave of the tile stores the node's exponent (the same value from
// each wave).  A child's exponent is the prefetched one, the previous
// step's output exponent (forwarded operand) or 0 (tip).
template <typename R>
__device__ __forceinline__ void post_auto_rescale4(R (&v)[4], R mx, const PostJob& J,
                                              const Operand4<R>& A, const Operand4<R>& B,
                                              int& prevE, const short* dummyE, int p, int lane,
                                              bool in) {
    const int e = mx > (R)0 ? exponent_of(mx) : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = pow2(v[i], -e);
    const int x1 = J.k1 == 2 ? prevE : (J.e1 ? A.e : 0);
    const int x2 = J.k2 == 2 ? prevE : (J.e2 ? B.e : 0);
    const int out = e + x1 + x2;
    short* dst = in ? J.ed + p : const_cast<short*>(dummyE) + 64 + lane;
    *gout<short>(dst) = (short)out;
    prevE = out;
}

template <typename R, bool AUTO>
__device__ __forceinline__ void post_compute(const PostJob& J, const Operand4<R>& A,
                                             const Operand4<R>& B, R (&prev)[4], int& prevE,
                                             R* smat, R (*smax)[8][TILE],
                                             int& k, int c, int lane, int p, int P, int C,
                                             long off, long doff, R* dummyW,
                                             const short* dummyE) {
    const bool in = p < P && p >= J.p0 && p < J.p1;
    R a[4], b[4], v[4];
    // the wave's own block of the matrices: written by the lanes that
    // carry them, read back as broadcasts (a tip's state column per lane)
    {
        constexpr int V = Vec16<R>::N;
        R* lm = smat + c * 40;
        *reinterpret_cast<typename Vec16<R>::T*>(lm + min(lane, 40 / V - 1) * V) = A.mv;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        step_product(J.k1, A, prev, lm, a);
        step_product(J.k2, B, prev, lm + 20, b);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = a[i] * b[i];
    if (!AUTO && J.mode == SCALE_READ) {
        const R sr = (R)A.sr;
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = v[i] / sr;
    }
    if (AUTO) {
        const R m = fmax(fmax(v[0], v[1]), fmax(v[2], v[3]));
        const R mx = tile_max<R>(in ? m : (R)0, smax, k, c, lane, C);
        post_auto_rescale4(v, mx, J, A, B, prevE, dummyE, p, lane, in);
    } else if (J.mode == SCALE_WRITE) {
        const R m = fmax(fmax(v[0], v[1]), fmax(v[2], v[3]));
        const R mx = tile_max<R>(in ? m : (R)0, smax, k, c, lane, C);
        if (in)
            rescale4<R, false>(v, mx, c, p, J.scaleWrite, J.logScalers, J.cumulative, J.accA,
                               J.accB, J.e1, J.e2, J.ed);
    }
    // FP64 tiles wholly inside the operation's pattern range: contiguous
    // stores; otherwise lanes outside the range store into a scratch tile
    const int t0 = p - lane;
    if (sizeof(R) == 8 && t0 >= J.p0 && t0 + TILE <= min(J.p1, P)) {
        store4_tile(gout<double>(reinterpret_cast<double*>(J.dest) + (off - 4 * lane)),
                    reinterpret_cast<const double(&)[4]>(v), lane);
    } else {
        R* dst = in ? reinterpret_cast<R*>(J.dest) + off : dummyW + doff;
        store4(gout<R>(dst), v);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) prev[i] = v[i];
}

template <typename R, bool AUTO>
__global__ __launch_bounds__(512) void k_post4(const PostJob* __restrict__ jobs, int njobs,
                                               int chunk, int P, int C, long cs,
                                               const R* __restrict__ mats, const R* dummyX,
                                               const int* dummyS, R* dummyW,
                                               const double* dummyD, const short* dummyE) {
    __shared__ R smax[2][8][TILE];
    __shared__ __attribute__((aligned(16))) R smat[8 * 40];   // [category][2 matrices][20]
    // workgroup (tile, y) walks operations [y chunk, (y + 1) chunk)
    jobs += blockIdx.y * chunk;
    njobs = min(chunk, njobs - (int)blockIdx.y * chunk);
    // wave index = category; readfirstlane makes it an SGPR so matrix
    // addresses are provably uniform and become scalar loads
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * TILE + lane;
    const long off = c * cs + (long)p * 4;
    const long doff = (long)(c * TILE + lane) * 4;
    int k = 0, prevE = 0;
    R prev[4] = {0, 0, 0, 0};
    Operand4<R> A0, B0, A1, B1;
    PostJob J0 = jobs[0], J1;
    post_issue<R, AUTO>(J0, dummyX, dummyS, dummyD, dummyE, off, doff, p, lane, A0, B0);
    post_issue_mats(J0, mats, c, lane, A0);
    for (int o = 0;; o += 2) {
        // operands of the next step; past the end the last operation's are
        // reloaded (in bounds, unused) so every step issues the same loads
        J1 = jobs[min(o + 1, njobs - 1)];
        post_issue<R, AUTO>(J1, dummyX, dummyS, dummyD, dummyE, off, doff, p, lane, A1, B1);
        post_issue_mats(J1, mats, c, lane, A1);
        post_compute<R, AUTO>(J0, A0, B0, prev, prevE, smat, smax, k, c, lane, p, P, C, off, doff,
                              dummyW, dummyE);
        if (o + 1 >= njobs) break;
        J0 = jobs[min(o + 2, njobs - 1)];
        post_issue<R, AUTO>(J0, dummyX, dummyS, dummyD, dummyE, off, doff, p, lane, A0, B0);
        post_issue_mats(J0, mats, c, lane, A0);
        post_compute<R, AUTO>(J1, A1, B1, prev, prevE, smat, smax, k, c, lane, p, P, C, off, doff,
                              dummyW, dummyE);
        if (o + 2 >= njobs) break;
    }
}

// Level-batched post-order: every operation of one dependency level in a
// single launch, grid = (pattern tiles, operations); each workgroup runs one
// operation on one tile.  Used for the wide levels of bushy trees, where it
// exposes ops x tiles independent workgroups instead of one serial chain
// per tile.
template <typename R, bool AUTO>
__global__ __launch_bounds__(512) void k_post4_level(const PostJob* __restrict__ jobs, int P,
                                                     int C, long cs,
                                                     const R* __restrict__ mats) {
    __shared__ R smax[2][8][TILE];
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * TILE + lane;
    const long off = c * cs + (long)p * 4;
    int k = 0;
    const PostJob J = jobs[blockIdx.y];
    const bool in = p < P && p >= J.p0 && p < J.p1;
    Operand4<R> A, B;
    fetch4(J.x1, J.s1, off, p, A);
    fetch4(J.x2, J.s2, off, p, B);
    R v[4];
    {
        R a[4], b[4];
        apply4(mats + J.mo1 + c * 20, J.s1 != nullptr, A, a);
        apply4(mats + J.mo2 + c * 20, J.s2 != nullptr, B, b);
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = a[i] * b[i];
        if (J.mode == SCALE_READ) {
            const R sr = (R)gin<double>(J.scaleRead)[p];
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] = v[i] / sr;
        }
    }
    if (AUTO || J.mode == SCALE_WRITE) {
        const R m = fmax(fmax(v[0], v[1]), fmax(v[2], v[3]));
        const R mx = tile_max<R>(in ? m : (R)0, smax, k, c, lane, C);
        if (in)
            rescale4<R, AUTO>(v, mx, c, p, J.scaleWrite, J.logScalers, J.cumulative, J.accA,
                              J.accB, J.e1, J.e2, J.ed);
    }
    if (in) store4(gout<R>(J.dest) + off, v);
}

// Pre-order traversal, pipelined like the post-order one (same fixed
// per-step loads into alternating register sets).  Operand kinds, decided
// on the host: parent 0 = load, 2 = the previous operation's output,
// 3 = the previous operation's parent (the two children of a node are
// consecutive operations); sibling and (with DERIV) the node's own
// post-order operand 0 = partials, 1 = tip states.
//
// With DERIV each operation also yields its branch's derivative sums:
// with pre = the new pre-order partials and post = the node's post-order
// partials, per pattern
//   den = sum_c w_c pre.post,  n1 = sum_c w_c pre.(D1 post),  n2 likewise
// (a tip uses column `state` of D and pre[state & 3] for den, as the
// reference does), then sum_p w_p n1/den, sum_p w_p (n1/den)^2 and
// sum_p w_p n2/den are accumulated per 64-pattern tile.  This is
// updatePrePartials followed by two calculateEdgeDerivatives calls, with
// the pre- and post-order partials read once.
template <typename R>
struct PreOps4 {
    Operand4<R> U, B, Y;   // parent, sibling, own post-order
};

template <typename R, bool DERIV>
__device__ __forceinline__ void pre_issue(const PreJob& J, const void* dx, const int* ds, long off,
                                          long doff, int p, int lane, PreOps4<R>& o) {
    issue4(J.parent, nullptr, J.kPar == 0 ? 0 : 2, dx, off, p, o.U);
    issue4(J.xs, J.ss, J.kSib, dx, off, p, o.B);
    if (DERIV) issue4(J.xself, J.sself, J.kSelf, dx, off, p, o.Y);
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

template <typename R, bool AUTO, bool DERIV>
__device__ __forceinline__ void pre_compute(const PreJob& J, const PreOps4<R>& o, R (&prev)[4],
                                            R (&uprev)[4], const R* __restrict__ mats,
                                            double* sdv, int& k, int c, int lane,
                                            int p, int P, int C, long off, long doff, R* dummyW,
                                            const short* dummyE, const double* __restrict__ cw,
                                            const double* __restrict__ pw, double* partial,
                                            int tile, int ntiles) {
    const bool in = p < P && p >= J.p0 && p < J.p1;
    R u[4], xb[5], b[4], v[4];
    // parent operand by kind (selects, no copies of in-flight registers)
#pragma unroll
    for (int i = 0; i < 4; ++i)
        u[i] = J.kPar == 2 ? prev[i] : (J.kPar == 3 ? uprev[i] : o.U.x[i]);
    operand4(J.kSib, o.B, prev, xb);
    mat4(mats + J.moSib + c * 20, xb, b);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uprev[i] = u[i];
        u[i] *= b[i];
    }
    const R* ms = mats + J.moSelf + c * 20;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        v[j] = ms[0 * 5 + j] * u[0] + ms[1 * 5 + j] * u[1] + ms[2 * 5 + j] * u[2] +
               ms[3 * 5 + j] * u[3];
    const bool rescale = AUTO || J.mode == SCALE_WRITE;
    // The per-pattern maximum (for rescaling) and the derivative sums over
    // categories are exchanged across the tile's category waves in one LDS
    // round; derivative ratios do not depend on the rescaling, so they are
    // formed from the values before it.
    double den = 0, n1 = 0, n2 = 0;
    if (DERIV) {
        R y[5], t1[4], t2[4];
        operand4(J.kSelf, o.Y, prev, y);
        mat4(mats + J.moD1 + c * 20, y, t1);
        mat4(mats + J.moD2 + c * 20, y, t2);
        if (J.kSelf == 1) {
            const int s = o.Y.st & 3;
            den = (double)(s == 0 ? v[0] : s == 1 ? v[1] : s == 2 ? v[2] : v[3]);
        } else {
            den = (double)v[0] * (double)y[0] + (double)v[1] * (double)y[1] +
                  (double)v[2] * (double)y[2] + (double)v[3] * (double)y[3];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            n1 += (double)v[i] * (double)t1[i];
            n2 += (double)v[i] * (double)t2[i];
        }
        const double wc = cw[c];
        den *= wc;
        n1 *= wc;
        n2 *= wc;
    }
    // A job without a destination (extension: derivative sums only) keeps
    // nothing; its stores go to the dummy tile.
    const bool keep = in && J.dest != nullptr;
    if (DERIV || rescale) {
        // LDS exchange buffer [2][C][4][TILE], alternating between jobs
        double* buf = sdv + (k & 1) * C * 4 * TILE;
        ++k;
        const R m = fmax(fmax(v[0], v[1]), fmax(v[2], v[3]));
        buf[(c * 4 + 0) * TILE + lane] = in ? (double)m : 0.0;
        if (DERIV) {
            buf[(c * 4 + 1) * TILE + lane] = den;
            buf[(c * 4 + 2) * TILE + lane] = n1;
            buf[(c * 4 + 3) * TILE + lane] = n2;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        if (rescale) {
            double mx = buf[lane];
            for (int cc = 1; cc < C; ++cc) mx = fmax(mx, buf[(cc * 4) * TILE + lane]);
            if (AUTO) {
                auto_rescale4(v, (R)mx, J.e1, J.e2, J.ed, dummyE, p, lane, keep);
            } else if (in) {
                rescale4<R, false>(v, (R)mx, c, p, J.scaleWrite, J.logScalers, nullptr, J.accA,
                                   J.accB, J.e1, J.e2, J.ed);
            }
        }
        if (DERIV && c == 0) {
            double D = 0, N1 = 0, N2 = 0;
            for (int cc = 0; cc < C; ++cc) {
                D += buf[(cc * 4 + 1) * TILE + lane];
                N1 += buf[(cc * 4 + 2) * TILE + lane];
                N2 += buf[(cc * 4 + 3) * TILE + lane];
            }
            const double w = in ? pw[p] : 0.0;
            const double r1 = in ? N1 / D : 0.0, r2 = in ? N2 / D : 0.0;
            const double a = wave_sum(w * r1), bsq = wave_sum(w * r1 * r1), e = wave_sum(w * r2);
            if (lane == 0) {
                GLOBAL double* pt = gout<double>(partial);
                pt[((long)J.oidx * 3 + 0) * ntiles + tile] = a;
                pt[((long)J.oidx * 3 + 1) * ntiles + tile] = bsq;
                pt[((long)J.oidx * 3 + 2) * ntiles + tile] = e;
            }
        }
    }
    R* dst = keep ? reinterpret_cast<R*>(J.dest) + off : dummyW + doff;
    store4(gout<R>(dst), v);
#pragma unroll
    for (int i = 0; i < 4; ++i) prev[i] = v[i];
}

template <typename R, bool AUTO, bool DERIV>
__global__ __launch_bounds__(512) void k_pre4(const PreJob* __restrict__ jobs, int njobs, int P,
                                              int C, long cs, const R* __restrict__ mats,
                                              const R* dummyX, const int* dummyS, R* dummyW,
                                              const short* dummyE, const double* __restrict__ cw,
                                              const double* __restrict__ pw, double* partial) {
    extern __shared__ double sdv[];   // [2][C][4][TILE]
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x;
    const int p = tile * TILE + lane;
    const long off = c * cs + (long)p * 4;
    const long doff = (long)(c * TILE + lane) * 4;
    const int ntiles = (P + TILE - 1) / TILE;
    int k = 0;
    R prev[4] = {0, 0, 0, 0}, uprev[4] = {0, 0, 0, 0};
    PreOps4<R> S0, S1;
    PreJob J0 = jobs[0], J1;
    pre_issue<R, DERIV>(J0, dummyX, dummyS, off, doff, p, lane, S0);
    for (int o = 0;; o += 2) {
        J1 = jobs[min(o + 1, njobs - 1)];
        pre_issue<R, DERIV>(J1, dummyX, dummyS, off, doff, p, lane, S1);
        pre_compute<R, AUTO, DERIV>(J0, S0, prev, uprev, mats, sdv, k, c, lane, p, P, C, off, doff,
                                    dummyW, dummyE, cw, pw, partial, tile, ntiles);
        if (o + 1 >= njobs) break;
        J0 = jobs[min(o + 2, njobs - 1)];
        pre_issue<R, DERIV>(J0, dummyX, dummyS, off, doff, p, lane, S0);
        pre_compute<R, AUTO, DERIV>(J1, S1, prev, uprev, mats, sdv, k, c, lane, p, P, C, off, doff,
                                    dummyW, dummyE, cw, pw, partial, tile, ntiles);
        if (o + 2 >= njobs) break;
    }
}

// Level-batched pre-order: one operation per workgroup, grid = (tiles, ops
// of one dependency level).
template <typename R, bool AUTO, bool DERIV>
__global__ __launch_bounds__(512) void k_pre4_level(const PreJob* __restrict__ jobs, int P, int C,
                                                    long cs, const R* __restrict__ mats,
                                                    const R* dummyX, const int* dummyS,
                                                    R* dummyW, const short* dummyE,
                                                    const double* __restrict__ cw,
                                                    const double* __restrict__ pw,
                                                    double* partial) {
    extern __shared__ double sdv[];   // [2][C][4][TILE]
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x;
    const int p = tile * TILE + lane;
    const long off = c * cs + (long)p * 4;
    const long doff = (long)(c * TILE + lane) * 4;
    const int ntiles = (P + TILE - 1) / TILE;
    int k = 0;
    R prev[4] = {0, 0, 0, 0}, uprev[4] = {0, 0, 0, 0};
    PreOps4<R> S0;
    const PreJob J = jobs[blockIdx.y];
    pre_issue<R, DERIV>(J, dummyX, dummyS, off, doff, p, lane, S0);
    pre_compute<R, AUTO, DERIV>(J, S0, prev, uprev, mats, sdv, k, c, lane, p, P, C, off, doff,
                                dummyW, dummyE, cw, pw, partial, tile, ntiles);
}

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
constexpr int NODE_MATS = 6 * 20;
constexpr int NODE_F32_BLOCK = 128;       // floats of a wave's block in the FP32 node kernels
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
    typename Vec16<R>::T mv;   // this lane's share of the job's matrices
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

template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_movf_rows(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, ROWS, 0xf, false));
}

// Sum over the 64 lanes, the steps of wave_sum_dpp in FP32.
__device__ __forceinline__ float wave_sum_f32(float v) {
    v += dpp_movf<0xB1, 0xf>(v);
    v += dpp_movf<0x4E, 0xf>(v);
    v += dpp_movf<0x114, 0xf>(v);
    v += dpp_movf<0x118, 0xf>(v);
    v += dpp_movf_rows<0x142, 0xa>(v);
    v += dpp_movf_rows<0x143, 0xc>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
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

// Operand loads of an FP32 node job without the dummy tile: uniform
// branches skip what the job does not read (a forwarded parent, the
// partials of a tip child, the states of a partials child).  Used where no
// later job's loads are in flight, so the branches cost no extra waits.
// A register with no particular value and no instruction to set it, for
// operand slots a job does not load (its uses are selected or branched
// away by the same job fields that skip the load).
template <typename T>
__device__ __forceinline__ T unset() {
    T x;
    asm volatile("" : "=v"(x));
    return x;
}

template <bool AUTO>
__device__ __forceinline__ void node_issue_direct(const NodeJob& J, const float* __restrict__ jm,
                                                  int c, unsigned off, int p, int lane,
                                                  NodeOps4<float>& o) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        o.u[i] = unset<float>();
        o.xa[i] = unset<float>();
        o.xb[i] = unset<float>();
    }
    o.sa = unset<int>();
    o.sb = unset<int>();
    o.eu = unset<int>();
    o.ea = unset<int>();
    o.eb = unset<int>();
    if (J.kPar == 0) load4(at<float>(J.parent, off), o.u);
    if (J.k[0] == 0) load4(at<float>(J.x[0], off), o.xa);
    else o.sa = *at<int>(J.x[0], 4u * p);
    if (J.k[1] == 0) load4(at<float>(J.x[1], off), o.xb);
    else o.sb = *at<int>(J.x[1], 4u * p);
    {
        const int l = min(lane, NODE_MATS / 4 - 1);
        o.mv = *at<fvec4>(jm, (unsigned)((J.mbase + c * NODE_MATS + l * 4) * (int)sizeof(float)));
    }
    if (AUTO) {
        if (J.kPar == 0 && J.eParent != nullptr) o.eu = *at<short>(J.eParent, 2u * p);
        if (J.ex[0] != nullptr) o.ea = *at<short>(J.ex[0], 2u * p);
        if (J.ex[1] != nullptr) o.eb = *at<short>(J.ex[1], 2u * p);
    }
}

template <int C, bool AUTO>
__device__ __forceinline__ void node_compute_f32(const NodeJob& J, const NodeOps4<float>& o,
                                                 float (&pa)[4], float (&pb)[4], int& pea,
                                                 int& peb, float* smat, float* sx, int& k, int c,
                                                 int lane, int p, int P, unsigned off,
                                                 const float (&wc)[C], float w, double* partial,
                                                 int tile, int ntiles) {
    const bool in = p < P;
    float u[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = J.kPar == 2 ? pa[i] : (J.kPar == 3 ? pb[i] : o.u[i]);
    const int eu = J.kPar == 2 ? pea : (J.kPar == 3 ? peb : o.eu);
    const bool tipA = J.k[0] == 1, tipB = J.k[1] == 1;
    // tip states are stored in [0, 4]; the clamp keeps the column index in
    // the matrix whatever a lane without a tip read
    const int sa = min((unsigned)o.sa, 4u), sb = min((unsigned)o.sb, 4u);
    const int par = k & 1;
    float* lm = smat + (c * 2 + par) * NODE_F32_BLOCK;
    *(fvec4*)(lm + min(lane, NODE_MATS / 4 - 1) * 4) = o.mv;
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
    mfma_cols(lm, r, ua, va);                      // pre(a) = P_a^T (pre(v) * P_b x_b)
    mfma_cols(lm + 20, r, ub, vb);                 // pre(b) = P_b^T (pre(v) * P_a x_a)
    float dA, n1A, n2A, dB, n1B, n2B;
    branch_sums_f32(tipA, sa, r, va, o.xa, lm + 40, lm + 60, dA, n1A, n2A);
    branch_sums_f32(tipB, sb, r, vb, o.xb, lm + 80, lm + 100, dB, n1B, n2B);
    // exchange: max a, max b, den/n1/n2 of a, of b
    float* buf = sx + par * 8 * C * TILE;
    ++k;
    float* mine = buf + c * TILE + lane;
    mine[0 * C * TILE] = in ? fmaxf(fmaxf(va[0], va[1]), fmaxf(va[2], va[3])) : 0.0f;
    mine[1 * C * TILE] = in ? fmaxf(fmaxf(vb[0], vb[1]), fmaxf(vb[2], vb[3])) : 0.0f;
    mine[2 * C * TILE] = dA;
    mine[3 * C * TILE] = n1A;
    mine[4 * C * TILE] = n2A;
    mine[5 * C * TILE] = dB;
    mine[6 * C * TILE] = n1B;
    mine[7 * C * TILE] = n2B;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    const bool keepA = in && J.dest[0] != nullptr;
    const bool keepB = in && J.dest[1] != nullptr;
    int outA = 0, outB = 0;
    if (AUTO) {
        float ma = buf[lane], mb = buf[C * TILE + lane];
#pragma unroll
        for (int cc = 1; cc < C; ++cc) {
            ma = fmaxf(ma, buf[cc * TILE + lane]);
            mb = fmaxf(mb, buf[(C + cc) * TILE + lane]);
        }
        const int ea = ma > 0.0f ? exponent_of(ma) : 0;
        const int eb = mb > 0.0f ? exponent_of(mb) : 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            va[i] = pow2(va[i], -ea);
            vb[i] = pow2(vb[i], -eb);
        }
        outA = ea + eu + (J.ex[1] ? o.eb : 0);
        outB = eb + eu + (J.ex[0] ? o.ea : 0);
        if (keepA) *at_w<short>(J.ed[0], 2u * p) = (short)outA;
        if (keepB) *at_w<short>(J.ed[1], 2u * p) = (short)outB;
    }
    // tile sums of child ch (w r1, w r1^2, w r2, r = N / D over categories)
    // by wave (ch + 2 par) % C
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
        if (c != (ch + 2 * par) % C) continue;
        const float* q = buf + (2 + 3 * ch) * C * TILE + lane;
        float D = wc[0] * q[0], N1 = wc[0] * q[C * TILE], N2 = wc[0] * q[2 * C * TILE];
#pragma unroll
        for (int cc = 1; cc < C; ++cc) {
            D += wc[cc] * q[cc * TILE];
            N1 += wc[cc] * q[(C + cc) * TILE];
            N2 += wc[cc] * q[(2 * C + cc) * TILE];
        }
        const float inv = 1.0f / D;
        const float r1 = in ? N1 * inv : 0.0f, r2 = in ? N2 * inv : 0.0f;
        const float a = w * r1;
        const float s1 = wave_sum_f32(a), sq = wave_sum_f32(a * r1), s2 = wave_sum_f32(w * r2);
        if (lane == 0) {
            GLOBAL double* pt = gout<double>(partial);
            const long row = (long)J.oidx[ch] * 3;
            pt[(row + 0) * ntiles + tile] = (double)s1;
            pt[(row + 1) * ntiles + tile] = (double)sq;
            pt[(row + 2) * ntiles + tile] = (double)s2;
        }
    }
    if (keepA) store4(at_w<float>(J.dest[0], off), va);
    if (keepB) store4(at_w<float>(J.dest[1], off), vb);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        pa[i] = va[i];
        pb[i] = vb[i];
    }
    pea = outA;
    peb = outB;
}

#define NODE_F32_SETUP                                                                        \
    extern __shared__ double sdv[]; /* [2][8][C][TILE] floats, then [C][2] matrix blocks */   \
    static_assert(NODE_MATS <= NODE_F32_BLOCK, "matrix block");                               \
    float* sx = reinterpret_cast<float*>(sdv);                                                \
    float* smat = sx + 2 * 8 * C * TILE;                                                      \
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

// Each job's operands are loaded when the job starts (node_issue_direct):
// at most 64 VGPRs, so eight waves per SIMD and a 100,000-pattern grid
// resident at once, which covers the loads' latency better than loading
// the next job ahead in more registers does.
template <int C, bool AUTO>
__global__ __launch_bounds__(64 * C, 7) void k_pren4f(const NodeJob* __restrict__ jobs, int njobs,
                                                      int chunk, int P, long cs,
                                                      const float* __restrict__ jm,
                                                      const double* __restrict__ cw,
                                                      const double* __restrict__ pw,
                                                      double* partial) {
    NODE_F32_SETUP
    jobs += blockIdx.y * chunk;
    njobs = min(chunk, njobs - (int)blockIdx.y * chunk);
    for (int o = 0; o < njobs; ++o) {
        const NodeJob J = jobs[o];
        NodeOps4<float> S0;
        node_issue_direct<AUTO>(J, jm, c, off, p, lane, S0);
        node_compute_f32<C, AUTO>(J, S0, pa, pb, pea, peb, smat, sx, k, c, lane, p, P, off, wc, w,
                                  partial, tile, ntiles);
    }
}

template <int C, bool AUTO>
__global__ __launch_bounds__(64 * C) void k_pren4f_level(const NodeJob* __restrict__ jobs,
                                                         int P, long cs,
                                                         const float* __restrict__ jm,
                                                         const double* __restrict__ cw,
                                                         const double* __restrict__ pw,
                                                         double* partial) {
    NODE_F32_SETUP
    NodeOps4<float> S0;
    const NodeJob J = jobs[blockIdx.y];
    node_issue_direct<AUTO>(J, jm, c, off, p, lane, S0);
    node_compute_f32<C, AUTO>(J, S0, pa, pb, pea, peb, smat, sx, k, c, lane, p, P, off, wc, w,
                              partial, tile, ntiles);
}
#undef NODE_F32_SETUP

template <int C>
static void launch_pre_nodes_f32(const NodeJob* djobs, int njobs, bool batched, int chunk,
                                 const Dims& d, const float* m, const double* cw,
                                 const double* pw, double* partial, hipStream_t st) {
    const int tiles = (d.P + TILE - 1) / TILE;
    const size_t lds = (size_t)(2 * 8 * C * TILE + C * 2 * NODE_F32_BLOCK) * sizeof(float);
    const dim3 block(TILE * C);
    if (batched && chunk <= 1) {
        if (d.autoScale) hipLaunchKernelGGL((k_pren4f_level<C, true>), dim3(tiles, njobs), block, lds, st, djobs, d.P, d.catStride, m, cw, pw, partial);
        else hipLaunchKernelGGL((k_pren4f_level<C, false>), dim3(tiles, njobs), block, lds, st, djobs, d.P, d.catStride, m, cw, pw, partial);
    } else {
        if (!batched || chunk > njobs) chunk = njobs;
        const dim3 grid(tiles, (njobs + chunk - 1) / chunk);
        if (d.autoScale) hipLaunchKernelGGL((k_pren4f<C, true>), grid, block, lds, st, djobs, njobs, chunk, d.P, d.catStride, m, cw, pw, partial);
        else hipLaunchKernelGGL((k_pren4f<C, false>), grid, block, lds, st, djobs, njobs, chunk, d.P, d.catStride, m, cw, pw, partial);
    }
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
    QOps4<R> S0, S1;
    NodeJob J0 = jobs[0], J1;
    q_issue<R>(J0, 0, dummyX, dummyS, jm, jf, c, C, off, doff, p, lane, S0);
    for (int o = 0;; o += 2) {
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
    // per-tile, per-category sums of H: rows kl of [16][tiles * C]
    const int ntiles = gridDim.x;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const double s = wave_sum_dpp(H[q]);
        if (lane == 0) partial[(long)q * ntiles * C + (long)tile * C + c] = s;
    }
}

template <typename R>
void launch_qgrad(const NodeJob* djobs, int njobs, const Dims& d, const void* jobMats,
                  const double* jobF, const double* eig, void* dummy, const double* cw,
                  const double* pw, double* partial, hipStream_t st) {
    if (njobs <= 0) return;
    const R* dummyX = reinterpret_cast<const R*>(dummy);
    R* dummyW = const_cast<R*>(dummyX) + 8 * TILE * 4;
    const double* dummyD = reinterpret_cast<const double*>(dummyW + 8 * TILE * 4);
    const int* dummyS = reinterpret_cast<const int*>(dummyD + TILE);
    const int tiles = (d.P + TILE - 1) / TILE;
    const size_t lds = (size_t)(2 * d.C * TILE + 32 + d.C * 2 * 32) * sizeof(double) +
                       (size_t)d.C * 2 * NODE_MAT_BLOCK;
    hipLaunchKernelGGL((k_qgrad4<R>), dim3(tiles), dim3(TILE * d.C), lds, st, djobs, njobs, d.P,
                       d.C, d.catStride, reinterpret_cast<const R*>(jobMats), jobF, eig, dummyX,
                       dummyS, cw, pw, partial);
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

// ---------------------------------------------------------------------------
// dlogL/dQ for any state count up to 64, one pre-order operation (branch)
// at a time.  With top = pre(parent) * (P_sib x_sib), x the branch's own
// post-order operand and P its matrix:
//   coef_p = w_p / sum_c w_c sum_i top_i (P x)_i                   (k_qgx_coef)
//   X_c    = sum_p coef_p w_c top_c(p) x_c(p)^T                      (k_qgx_outer)
//   H     += (V^T X_c V^-T) o F(r_c t),                              (k_qgx_eigen)
//   F_ab   = (e^(tau l_a) - e^(tau l_b)) / (l_a - l_b)
// the sums k_qgrad4 forms for 4 states.  The likelihood in coef is formed
// from the same operands as top and x, so scale factors and exponents
// cancel.  Patterns go in tiles of QGX_T, staged in LDS as [S][QGX_LD]
// doubles (row stride QGX_LD: column reads of consecutive states do not
// share a bank); matrices are read from memory (cached: every pattern of a
// workgroup reads the same rows).

constexpr int QGX_T = 32;
constexpr int QGX_LD = QGX_T + 1;
static_assert(WG == 256, "the Q-gradient register blocks map 256 threads to 16 x 16");
constexpr int QGX_NE = 64 * 64 / WG;   // S x S entries per thread, S <= 64 (k_qgx_eigen)

// Tiles are staged in the partials' own precision T = R (the products are
// formed in FP64 from them).  With FP32 partials k_qgx_outer, which walks
// many tiles of one category, also stages that category's sibling matrix
// in LDS (qgx_mstage); k_qgx_coef takes one tile per workgroup, where
// staging two matrices per category costs more than reading their rows.
// LDS, [doubles][T tiles][matrix][ints]: k_qgx_coef: [16][QGX_T] partial
// sums and L; k_qgx_outer: the tile's coef, the staged matrix.  At most
// 55,296 bytes (S = 64).
template <typename R>
constexpr bool qgx_mstage() { return sizeof(R) == 4; }

template <typename R>
static size_t qgx_lds_bytes(int S, bool coef) {
    const size_t d = coef ? 16 * QGX_T + QGX_T : QGX_T;
    const size_t mats = (!coef && qgx_mstage<R>()) ? (size_t)S * (S + 1) * sizeof(R) : 0;
    return d * sizeof(double) + (size_t)3 * S * QGX_LD * sizeof(R) + mats +
           2 * QGX_T * sizeof(int);
}

// Stage category c of the tile at p0: A = pre(parent), B = the sibling's
// post-order partials (not for a tip sibling), X = own operand (one-hot of
// a tip state < S; zero for a missing state, whose pad column does not
// depend on Q).  Tip states into sts (sibling) and sto (own).
template <typename R>
__device__ void qgx_stage(const PreJob& J, const Dims& d, int c, int p0, R* A, R* B, R* X,
                          int* sts, int* sto) {
    const R* par = reinterpret_cast<const R*>(J.parent);
    const R* xs = reinterpret_cast<const R*>(J.xs);
    const R* xo = reinterpret_cast<const R*>(J.xself);
    for (int t = threadIdx.x; t < d.S * QGX_T; t += WG) {
        const int i = t / QGX_T, q = t % QGX_T, p = p0 + q;
        const long e = pbase(d, c, p) + (long)i * d.sStride;
        A[i * QGX_LD + q] = par[e];
        if (xs) B[i * QGX_LD + q] = xs[e];
        X[i * QGX_LD + q] = xo ? xo[e] : (J.sself[p] == i ? (R)1 : (R)0);
    }
    if (threadIdx.x < QGX_T) {
        const int p = p0 + threadIdx.x;
        sts[threadIdx.x] = J.ss ? J.ss[p] : 0;
        sto[threadIdx.x] = J.sself ? J.sself[p] : 0;
    }
}

// Category c's matrix rows ([S][S + 1] reals) into LDS when staged.
template <typename R>
__device__ void qgx_stage_matrix(const R* M, int S, int c, R* dst) {
    const R* src = M + (long)c * S * (S + 1);
    for (int t = threadIdx.x; t < S * (S + 1); t += WG) dst[t] = src[t];
}

// Register blocks: the 256 threads of a workgroup as a 16 x 16 grid (bi,
// bj); products cover states 4 bi + u (u < 4) at patterns 2 bj + w (w < 2),
// the outer product entries (4 bi + u, 4 bj + v).  Rows and columns past S
// are clamped to S - 1 (their results are discarded), so every read stays
// inside the staged tiles.

// (M v)_i at pattern q for the thread's block: M the category's rows
// ([S][S + 1] reals, pad column S; in LDS or memory), v a staged tile, or
// for a tip the rows' entries at the pattern's state (the pad for a
// missing one).
template <typename R, typename T>
__device__ __forceinline__ void qgx_mv_block(const R* M, int S, const T* v, bool tip,
                                             const int* st, int i0, int q0,
                                             double (&out)[4][2]) {
    const R* r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) r[u] = M + (long)min(i0 + u, S - 1) * (S + 1);
    if (tip) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int w = 0; w < 2; ++w) out[u][w] = (double)r[u][st[q0 + w]];
        return;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) out[u][0] = out[u][1] = 0.0;
    for (int k = 0; k < S; ++k) {
        const double v0 = (double)v[k * QGX_LD + q0], v1 = (double)v[k * QGX_LD + q0 + 1];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double m = (double)r[u][k];
            out[u][0] += m * v0;
            out[u][1] += m * v1;
        }
    }
}

// grid (Ppad / QGX_T, 1, operations): coef[o][p]
template <typename R>
__global__ __launch_bounds__(WG) void k_qgx_coef(const PreJob* __restrict__ jobs, Dims d,
                                                 const double* __restrict__ cw,
                                                 const double* __restrict__ pw,
                                                 double* __restrict__ coef) {
    extern __shared__ double qs[];
    const int S = d.S;
    double* red = qs;                                 // [16][QGX_T]
    double* L = red + 16 * QGX_T;                     // [QGX_T]
    R* A = reinterpret_cast<R*>(L + QGX_T);
    R* B = A + S * QGX_LD;
    R* X = B + S * QGX_LD;
    int* sts = reinterpret_cast<int*>(X + S * QGX_LD);
    int* sto = sts + QGX_T;
    const int o = blockIdx.z;
    const PreJob J = jobs[o];
    const int p0 = blockIdx.x * QGX_T;
    const int bi = threadIdx.x / 16, i0 = 4 * bi, q0 = 2 * (threadIdx.x % 16);
    if (threadIdx.x < QGX_T) L[threadIdx.x] = 0.0;
    for (int c = 0; c < d.C; ++c) {
        __syncthreads();
        qgx_stage<R>(J, d, c, p0, A, B, X, sts, sto);
        __syncthreads();
        const R* Ms = reinterpret_cast<const R*>(J.mSib) + (long)c * S * (S + 1);
        const R* Mo = reinterpret_cast<const R*>(J.mSelf) + (long)c * S * (S + 1);
        double tb[4][2], yb[4][2];
        qgx_mv_block(Ms, S, B, J.ss != nullptr, sts, i0, q0, tb);   // P_sib x_sib
        qgx_mv_block(Mo, S, X, J.sself != nullptr, sto, i0, q0, yb);  // P x
        double part[2] = {0.0, 0.0};
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u < S)
#pragma unroll
                for (int w = 0; w < 2; ++w)
                    part[w] += (double)A[(i0 + u) * QGX_LD + q0 + w] * tb[u][w] * yb[u][w];
        red[bi * QGX_T + q0] = part[0];
        red[bi * QGX_T + q0 + 1] = part[1];
        __syncthreads();
        if (threadIdx.x < QGX_T) {
            double t = 0.0;
            for (int k = 0; k < 16; ++k) t += red[k * QGX_T + threadIdx.x];
            L[threadIdx.x] += cw[c] * t;
        }
    }
    __syncthreads();
    if (threadIdx.x < QGX_T) {
        const int p = p0 + threadIdx.x;
        const double l = L[threadIdx.x];
        coef[(long)o * d.Ppad + p] = (p < d.P && l != 0.0) ? pw[p] / l : 0.0;
    }
}

// grid (chunks, C, operations): this chunk's X_c, out[o][c][chunk][S * S].
// The coef-scaled top vectors are kept in FP64 registers and written over
// the parent's tile in T (FP32 partials: the top values at FP32 precision,
// as the likelihood path holds them).
template <typename R>
__global__ __launch_bounds__(WG) void k_qgx_outer(const PreJob* __restrict__ jobs, Dims d,
                                                  const double* __restrict__ cw,
                                                  const double* __restrict__ coef,
                                                  int tilesPerChunk, double* __restrict__ out) {
    extern __shared__ double qs[];
    constexpr bool MS = qgx_mstage<R>();
    const int S = d.S;
    double* cf = qs;
    R* A = reinterpret_cast<R*>(cf + QGX_T);
    R* B = A + S * QGX_LD;
    R* X = B + S * QGX_LD;
    R* lMs = X + S * QGX_LD;
    int* sts = reinterpret_cast<int*>(lMs + (MS ? S * (S + 1) : 0));
    int* sto = sts + QGX_T;
    const int o = blockIdx.z, c = blockIdx.y;
    const PreJob J = jobs[o];
    const R* gMs = reinterpret_cast<const R*>(J.mSib);
    if (MS) qgx_stage_matrix(gMs, S, c, lMs);   // visible after the first tile's barrier
    const R* Ms = MS ? lMs : gMs + (long)c * S * (S + 1);
    const double wc = cw[c];
    const int bi = threadIdx.x / 16, bj = threadIdx.x % 16;
    const int i0 = 4 * bi, q0 = 2 * bj, j0 = 4 * bj;
    const bool live = i0 < S && j0 < S;   // outer-product block inside S x S
    int ri[4], rj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        ri[u] = min(i0 + u, S - 1) * QGX_LD;
        rj[u] = min(j0 + u, S - 1) * QGX_LD;
    }
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    const int ntiles = d.Ppad / QGX_T;
    const int t0 = blockIdx.x * tilesPerChunk, t1 = min(ntiles, t0 + tilesPerChunk);
    for (int tile = t0; tile < t1; ++tile) {
        const int p0 = tile * QGX_T;
        __syncthreads();
        qgx_stage<R>(J, d, c, p0, A, B, X, sts, sto);
        if (threadIdx.x < QGX_T) cf[threadIdx.x] = wc * coef[(long)o * d.Ppad + p0 + threadIdx.x];
        __syncthreads();
        double tb[4][2];
        qgx_mv_block(Ms, S, B, J.ss != nullptr, sts, i0, q0, tb);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int w = 0; w < 2; ++w) tb[u][w] *= cf[q0 + w] * (double)A[ri[u] + q0 + w];
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u < S) {
                A[(i0 + u) * QGX_LD + q0] = (R)tb[u][0];
                A[(i0 + u) * QGX_LD + q0 + 1] = (R)tb[u][1];
            }
        __syncthreads();
        if (live)
            for (int k = 0; k < QGX_T; ++k) {
                double a[4], x[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    a[u] = (double)A[ri[u] + k];
                    x[u] = (double)X[rj[u] + k];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * x[v];
            }
    }
    double* dst = out + (((long)o * d.C + c) * gridDim.x + blockIdx.x) * S * S;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (i0 + u < S && j0 + v < S) dst[(i0 + u) * S + j0 + v] = acc[u][v];
}

// grid (C, operations): sums the chunks of X_c, forms (V^T X_c V^-T) o F
// into z[(a S + b)][o C + c].  eig = V | V^-1 (row-major) | eigenvalues;
// tau[o C + c] = r_c t of the operation's branch.
__global__ __launch_bounds__(WG) void k_qgx_eigen(const double* __restrict__ xs, int chunks,
                                                  int nops, int S, int C,
                                                  const double* __restrict__ eig,
                                                  const double* __restrict__ tau,
                                                  double* __restrict__ z) {
    extern __shared__ double qs[];
    double* Xs = qs;   // S x S
    const int c = blockIdx.x, o = blockIdx.y;
    const double* V = eig;
    const double* Vi = eig + S * S;
    const double* lam = eig + 2 * S * S;
    const double* src = xs + ((long)o * C + c) * chunks * S * S;
    for (int e = threadIdx.x; e < S * S; e += WG) {
        double v = 0.0;
        for (int k = 0; k < chunks; ++k) v += src[(long)k * S * S + e];
        Xs[e] = v;
    }
    __syncthreads();
    double t[QGX_NE];
#pragma unroll
    for (int m = 0; m < QGX_NE; ++m) {   // T = V^T X: T[a][j] = sum_i V[i][a] X[i][j]
        const int e = threadIdx.x + m * WG;
        t[m] = 0.0;
        if (e < S * S) {
            const int a = e / S, j = e % S;
            for (int i = 0; i < S; ++i) t[m] += V[i * S + a] * Xs[i * S + j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < QGX_NE; ++m) {
        const int e = threadIdx.x + m * WG;
        if (e < S * S) Xs[e] = t[m];
    }
    __syncthreads();
    const double tv = tau[o * C + c];
    const long cols = (long)nops * C;
#pragma unroll
    for (int m = 0; m < QGX_NE; ++m) {   // Y = T V^-T: Y[a][b] = sum_j T[a][j] Vi[b][j]
        const int e = threadIdx.x + m * WG;
        if (e < S * S) {
            const int a = e / S, b = e % S;
            double y = 0.0;
            for (int j = 0; j < S; ++j) y += Xs[a * S + j] * Vi[b * S + j];
            const double dl = lam[a] - lam[b];
            // tau e^(tau l) at dl = 0, expm1 near it (no cancellation)
            const double f = dl == 0.0 ? tv * exp(tv * lam[b]) : exp(tv * lam[b]) * expm1(tv * dl) / dl;
            z[(long)e * cols + o * C + c] = y * f;
        }
    }
}

template <typename R>
void launch_qgrad_generic(const PreJob* djobs, int nops, const Dims& d, const double* cw,
                          const double* pw, const double* eig, const double* tau, double* coef,
                          double* xs, int chunks, double* z, hipStream_t st) {
    if (nops <= 0) return;
    const int ntiles = d.Ppad / QGX_T;
    hipLaunchKernelGGL(k_qgx_coef<R>, dim3(ntiles, 1, nops), dim3(WG), qgx_lds_bytes<R>(d.S, true), st,
                       djobs, d, cw, pw, coef);
    const int per = (ntiles + chunks - 1) / chunks;
    hipLaunchKernelGGL(k_qgx_outer<R>, dim3(chunks, d.C, nops), dim3(WG), qgx_lds_bytes<R>(d.S, false), st,
                       djobs, d, cw, coef, per, xs);
    hipLaunchKernelGGL(k_qgx_eigen, dim3(d.C, nops), dim3(WG), (size_t)d.S * d.S * sizeof(double),
                       st, xs, chunks, nops, d.S, d.C, eig, tau, z);
}

// ---------------------------------------------------------------------------
// generic state count: one launch per operation, one lane per output
// element (category, pattern, state); a second pass rescales.

template <typename R>
__device__ __forceinline__ R child_term(const R* x, const int* s, const R* m, int S, long xoff,
                                        int ss, int p, int c, int i) {
    const R* row = m + ((long)c * S + i) * (S + 1);
    if (s) return row[s[p]];
    const R* xv = x + xoff;
    R acc = 0;
    for (int j = 0; j < S; ++j) acc += row[j] * xv[(long)j * ss];
    return acc;
}

template <typename R>
__global__ __launch_bounds__(WG) void k_postx(const PostJob* __restrict__ jobs, int o, Dims d) {
    const PostJob J = jobs[o];
    const long n = (long)d.C * (J.p1 - J.p0) * d.S;
    const long t = (long)blockIdx.x * WG + threadIdx.x;
    if (t >= n) return;
    const int i = t % d.S;
    const long cp = t / d.S;
    const int np = J.p1 - J.p0;
    const int p = J.p0 + (int)(cp % np);
    const int c = (int)(cp / np);
    const long off = pbase(d, c, p);
    const R a = child_term(reinterpret_cast<const R*>(J.x1), J.s1,
                           reinterpret_cast<const R*>(J.m1), d.S, off, d.sStride, p, c, i);
    const R b = child_term(reinterpret_cast<const R*>(J.x2), J.s2,
                           reinterpret_cast<const R*>(J.m2), d.S, off, d.sStride, p, c, i);
    R v = a * b;
    if (J.mode == SCALE_READ) v = v / (R)J.scaleRead[p];
    reinterpret_cast<R*>(J.dest)[off + (long)i * d.sStride] = v;
}

// pre-order, generic S, phase 1: u[c,p,i] = pre_parent[c,p,i] * term_i
template <typename R>
__global__ __launch_bounds__(WG) void k_prex_u(const PreJob* __restrict__ jobs, int o, Dims d,
                                               R* __restrict__ u) {
    const PreJob J = jobs[o];
    const int np = J.p1 - J.p0;
    const long n = (long)d.C * np * d.S;
    const long t = (long)blockIdx.x * WG + threadIdx.x;
    if (t >= n) return;
    const int i = t % d.S;
    const long cp = t / d.S;
    const int p = J.p0 + (int)(cp % np);
    const int c = (int)(cp / np);
    const long off = pbase(d, c, p);
    const R b = child_term(reinterpret_cast<const R*>(J.xs), J.ss,
                           reinterpret_cast<const R*>(J.mSib), d.S, off, d.sStride, p, c, i);
    const long e = off + (long)i * d.sStride;
    u[e] = reinterpret_cast<const R*>(J.parent)[e] * b;
}

// phase 2: dest[c,p,j] = sum_i M[i][j] u[c,p,i]  (M read transposed when the
// client passed P^T, the reference GPU convention for S > 4)
template <typename R>
__global__ __launch_bounds__(WG) void k_prex_out(const PreJob* __restrict__ jobs, int o, Dims d,
                                                 const R* __restrict__ u, int transposedSelf) {
    const PreJob J = jobs[o];
    const int np = J.p1 - J.p0;
    const long n = (long)d.C * np * d.S;
    const long t = (long)blockIdx.x * WG + threadIdx.x;
    if (t >= n) return;
    const int j = t % d.S;
    const long cp = t / d.S;
    const int p = J.p0 + (int)(cp % np);
    const int c = (int)(cp / np);
    const long off = pbase(d, c, p);
    const int ss = d.sStride;
    const R* M = reinterpret_cast<const R*>(J.mSelf) + (long)c * d.S * (d.S + 1);
    const int W = d.S + 1;
    R acc = 0;
    if (transposedSelf)
        for (int i = 0; i < d.S; ++i) acc += M[j * W + i] * u[off + (long)i * ss];
    else
        for (int i = 0; i < d.S; ++i) acc += M[i * W + j] * u[off + (long)i * ss];
    reinterpret_cast<R*>(J.dest)[off + (long)j * ss] = acc;
}

// rescale pass for the generic path: one lane per pattern.
template <typename R>
__global__ __launch_bounds__(WG) void k_rescalex(R* __restrict__ dest, int p0, int p1,
                                                 double* scaleWrite, int logScalers,
                                                 double* cumulative, const double* accA,
                                                 const double* accB, const short* e1,
                                                 const short* e2, short* ed, Dims d) {
    const int p = p0 + blockIdx.x * WG + threadIdx.x;
    if (p >= p1) return;
    R mx = 0;
    for (int c = 0; c < d.C; ++c)
        for (int i = 0; i < d.S; ++i) mx = fmax(mx, dest[pbase(d, c, p) + (long)i * d.sStride]);
    if (d.autoScale) {
        const int e = mx > (R)0 ? exponent_of(mx) : 0;
        for (int c = 0; c < d.C; ++c)
            for (int i = 0; i < d.S; ++i) {
                R& x = dest[pbase(d, c, p) + (long)i * d.sStride];
                x = pow2(x, -e);
            }
        int t = e;
        if (e1) t += e1[p];
        if (e2) t += e2[p];
        ed[p] = (short)t;
        return;
    }
    if (mx == (R)0) mx = (R)1;
    const R inv = (R)1 / mx;
    for (int c = 0; c < d.C; ++c)
        for (int i = 0; i < d.S; ++i) dest[pbase(d, c, p) + (long)i * d.sStride] *= inv;
    const double lm = log((double)mx);
    double w = logScalers ? lm : (double)mx;
    if (cumulative) cumulative[p] += lm;
    if (accA) w += accA[p];
    if (accB) w += accB[p];
    scaleWrite[p] = w;
}

// Traversal launchers: the fused 4-state kernels, or the generic-S path
// whose grids come from the host copy of each job.
template <typename R>
void launch_post_levels(const PostJob* djobs, const int* levelStart, int nlevels, const Dims& d,
                        const void* mats, hipStream_t st) {
    for (int l = 0; l < nlevels; ++l) {
        const int n = levelStart[l + 1] - levelStart[l];
        dim3 grid((d.P + TILE - 1) / TILE, n);
        if (d.autoScale)
            hipLaunchKernelGGL((k_post4_level<R, true>), grid, dim3(TILE * d.C), 0, st,
                               djobs + levelStart[l], d.P, d.C, d.catStride,
                               reinterpret_cast<const R*>(mats));
        else
            hipLaunchKernelGGL((k_post4_level<R, false>), grid, dim3(TILE * d.C), 0, st,
                               djobs + levelStart[l], d.P, d.C, d.catStride,
                               reinterpret_cast<const R*>(mats));
    }
}

template <typename R>
void launch_post(const PostJob* djobs, const PostJob* hjobs, int njobs, int chunk,
                 const Dims& d, const void* mats, void* dummy, hipStream_t st) {
    // dummy: scratch laid out as [read tile 8x64x4 R][write tile 8x64x4 R]
    //        [64 doubles][64 ints][128 shorts]
    const R* dummyX = reinterpret_cast<const R*>(dummy);
    R* dummyW = const_cast<R*>(dummyX) + 8 * TILE * 4;
    const double* dummyD = reinterpret_cast<const double*>(dummyW + 8 * TILE * 4);
    const int* dummyS = reinterpret_cast<const int*>(dummyD + TILE);
    const short* dummyE = reinterpret_cast<const short*>(dummyS + TILE);
    if (njobs <= 0) return;
    if (d.S == 4 && d.C <= 8) {
        if (chunk <= 0) chunk = njobs;
        dim3 grid((d.P + TILE - 1) / TILE, (njobs + chunk - 1) / chunk);
        if (d.autoScale)
            hipLaunchKernelGGL((k_post4<R, true>), grid, dim3(TILE * d.C), 0, st, djobs, njobs,
                               chunk, d.P, d.C, d.catStride, reinterpret_cast<const R*>(mats),
                               dummyX, dummyS, dummyW, dummyD, dummyE);
        else
            hipLaunchKernelGGL((k_post4<R, false>), grid, dim3(TILE * d.C), 0, st, djobs, njobs,
                               chunk, d.P, d.C, d.catStride, reinterpret_cast<const R*>(mats),
                               dummyX, dummyS, dummyW, dummyD, dummyE);
        return;
    }
    for (int o = 0; o < njobs; ++o) {
        const PostJob& J = hjobs[o];
        const long n = (long)d.C * (J.p1 - J.p0) * d.S;
        if (n <= 0) continue;
        hipLaunchKernelGGL(k_postx<R>, dim3((n + WG - 1) / WG), dim3(WG), 0, st, djobs, o, d);
        if (J.mode == SCALE_WRITE || d.autoScale)
            hipLaunchKernelGGL(k_rescalex<R>, dim3((J.p1 - J.p0 + WG - 1) / WG), dim3(WG), 0, st,
                               reinterpret_cast<R*>(J.dest), J.p0, J.p1, J.scaleWrite,
                               J.logScalers, J.cumulative, J.accA, J.accB, J.e1, J.e2, J.ed, d);
    }
}

// LDS exchange of the 4-state pre-order kernels: [2][C][4][TILE] doubles
// (used by jobs that rescale or form derivative sums).
static size_t pre_lds_bytes(int C) { return (size_t)2 * C * 4 * TILE * sizeof(double); }

template <typename R>
void launch_pre_levels(const PreJob* djobs, const int* levelStart, int nlevels, const Dims& d,
                       const void* mats, void* dummy, const double* cw, const double* pw,
                       double* partial, hipStream_t st) {
    const R* dummyX = reinterpret_cast<const R*>(dummy);
    R* dummyW = const_cast<R*>(dummyX) + 8 * TILE * 4;
    const double* dummyD = reinterpret_cast<const double*>(dummyW + 8 * TILE * 4);
    const int* dummyS = reinterpret_cast<const int*>(dummyD + TILE);
    const short* dummyE = reinterpret_cast<const short*>(dummyS + TILE);
    const R* m = reinterpret_cast<const R*>(mats);
    for (int l = 0; l < nlevels; ++l) {
        const int n = levelStart[l + 1] - levelStart[l];
        dim3 grid((d.P + TILE - 1) / TILE, n);
        dim3 block(TILE * d.C);
        const size_t lds = pre_lds_bytes(d.C);
        const PreJob* j = djobs + levelStart[l];
#define PREL(A, D)                                                                                \
    hipLaunchKernelGGL((k_pre4_level<R, A, D>), grid, block, lds, st, j, d.P, d.C, d.catStride, m, \
                       dummyX, dummyS, dummyW, dummyE, cw, pw, partial)
        if (partial) {
            if (d.autoScale) PREL(true, true);
            else PREL(false, true);
        } else {
            if (d.autoScale) PREL(true, false);
            else PREL(false, false);
        }
#undef PREL
    }
}

template <typename R>
__global__ __launch_bounds__(WG) void k_gather_node_mats(const NodeJob* __restrict__ jobs,
                                                         int njobs, int C,
                                                         const R* __restrict__ mats,
                                                         R* __restrict__ out, bool colMajor) {
    const long t = (long)blockIdx.x * WG + threadIdx.x;
    const long per = (long)C * NODE_MATS;
    if (t >= per * njobs) return;
    const int j = (int)(t / per);
    const int rem = (int)(t - (long)j * per);
    const int c = rem / NODE_MATS, r = rem - c * NODE_MATS;
    const int m = r / 20, e = r - m * 20;
    // row-major [4][5] as stored, or column-major [5][4]
    const int d = colMajor ? m * 20 + (e % 5) * 4 + e / 5 : r;
    const NodeJob& J = jobs[j];
    const int mo = m == 0 ? J.moP[0] : m == 1 ? J.moP[1] : m == 2 ? J.moD1[0]
                 : m == 3 ? J.moD2[0] : m == 4 ? J.moD1[1] : J.moD2[1];
    out[(long)J.mbase + c * NODE_MATS + d] = mats[(long)mo + c * 20 + e];
}

template <typename R>
void launch_gather_node_matrices(const NodeJob* djobs, int njobs, const Dims& d,
                                 const void* mats, void* jobMats, hipStream_t st, bool colMajor) {
    const long n = (long)njobs * d.C * NODE_MATS;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_node_mats<R>, dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, st,
                       djobs, njobs, d.C, reinterpret_cast<const R*>(mats),
                       reinterpret_cast<R*>(jobMats), colMajor);
}

template <typename R>
void launch_pre_nodes(const NodeJob* djobs, int njobs, bool batched, int chunk, const Dims& d,
                      const void* jobMats, void* dummy, const double* cw, const double* pw,
                      double* partial, hipStream_t st) {
    if (njobs <= 0) return;
    const R* dummyX = reinterpret_cast<const R*>(dummy);
    R* dummyW = const_cast<R*>(dummyX) + 8 * TILE * 4;
    const double* dummyD = reinterpret_cast<const double*>(dummyW + 8 * TILE * 4);
    const int* dummyS = reinterpret_cast<const int*>(dummyD + TILE);
    const short* dummyE = reinterpret_cast<const short*>(dummyS + TILE);
    const R* m = reinterpret_cast<const R*>(jobMats);
    if (std::is_same<R, float>::value && node_f32_supported(d.C)) {
        const float* mf = reinterpret_cast<const float*>(m);
        if (d.C == 4) launch_pre_nodes_f32<4>(djobs, njobs, batched, chunk, d, mf, cw, pw, partial, st);
        else if (d.C == 2) launch_pre_nodes_f32<2>(djobs, njobs, batched, chunk, d, mf, cw, pw, partial, st);
        else launch_pre_nodes_f32<1>(djobs, njobs, batched, chunk, d, mf, cw, pw, partial, st);
        return;
    }
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

template <typename R>
void launch_pre(const PreJob* djobs, const PreJob* hjobs, int njobs, const Dims& d,
                int transposedSelf, void* scratch, const void* mats, void* dummy,
                const double* cw, const double* pw, double* partial, hipStream_t st) {
    if (njobs <= 0) return;
    if (d.S == 4 && d.C <= 8 && !transposedSelf) {
        const R* dummyX = reinterpret_cast<const R*>(dummy);
        R* dummyW = const_cast<R*>(dummyX) + 8 * TILE * 4;
        const double* dummyD = reinterpret_cast<const double*>(dummyW + 8 * TILE * 4);
        const int* dummyS = reinterpret_cast<const int*>(dummyD + TILE);
        const short* dummyE = reinterpret_cast<const short*>(dummyS + TILE);
        const R* m = reinterpret_cast<const R*>(mats);
        dim3 grid((d.P + TILE - 1) / TILE);
        dim3 block(TILE * d.C);
        const size_t lds = pre_lds_bytes(d.C);
#define PRE4(A, D)                                                                                \
    hipLaunchKernelGGL((k_pre4<R, A, D>), grid, block, lds, st, djobs, njobs, d.P, d.C,           \
                       d.catStride, m, dummyX, dummyS, dummyW, dummyE, cw, pw, partial)
        if (partial) {
            if (d.autoScale) PRE4(true, true);
            else PRE4(false, true);
        } else {
            if (d.autoScale) PRE4(true, false);
            else PRE4(false, false);
        }
#undef PRE4
        return;
    }
    for (int o = 0; o < njobs; ++o) {
        const PreJob& J = hjobs[o];
        const long n = (long)d.C * (J.p1 - J.p0) * d.S;
        if (n <= 0) continue;
        dim3 grid((n + WG - 1) / WG);
        hipLaunchKernelGGL(k_prex_u<R>, grid, dim3(WG), 0, st, djobs, o, d,
                           reinterpret_cast<R*>(scratch));
        hipLaunchKernelGGL(k_prex_out<R>, grid, dim3(WG), 0, st, djobs, o, d,
                           reinterpret_cast<const R*>(scratch), transposedSelf);
        if (J.mode == SCALE_WRITE || d.autoScale)
            hipLaunchKernelGGL(k_rescalex<R>, dim3((J.p1 - J.p0 + WG - 1) / WG), dim3(WG), 0, st,
                               reinterpret_cast<R*>(J.dest), J.p0, J.p1, J.scaleWrite,
                               J.logScalers, nullptr, J.accA, J.accB, J.e1, J.e2, J.ed, d);
    }
}

// ---------------------------------------------------------------------------
// root log-likelihood: site[p] = log(sum_i pi_i sum_c w_c root[c,p,i]) +
// cumulative[p]; per-workgroup sums of weight*site, then one fixed-order
// reduction (BeagleCPUImpl.hpp:2630-2691 semantics, FP64 sum).

template <typename R>
__global__ __launch_bounds__(WG) void k_root(const R* __restrict__ root,
                                             const double* __restrict__ cw,
                                             const double* __restrict__ freqs,
                                             const double* __restrict__ cum,
                                             const double* __restrict__ pw,
                                             const short* __restrict__ ex,
                                             double* __restrict__ site, double* __restrict__ bsum,
                                             int p0, int p1, Dims d) {
    __shared__ double sh[WG / 64];
    const int p = p0 + blockIdx.x * WG + threadIdx.x;
    double contrib = 0.0;
    if (p < p1) {
        double sum = 0.0;
        if (d.S == 4) {
            double t[4] = {0, 0, 0, 0};
            for (int c = 0; c < d.C; ++c) {
                R v[4];
                load4(root + c * d.catStride + (long)p * 4, v);
                for (int i = 0; i < 4; ++i) t[i] += (double)v[i] * cw[c];
            }
            for (int i = 0; i < 4; ++i) sum += freqs[i] * t[i];
        } else {
            for (int i = 0; i < d.S; ++i) {
                double t = 0.0;
                for (int c = 0; c < d.C; ++c)
                    t += (double)root[pbase(d, c, p) + (long)i * d.sStride] * cw[c];
                sum += freqs[i] * t;
            }
        }
        double s = log(sum);
        if (cum) s += cum[p];
        if (ex) s += 0.69314718055994530942 * (double)ex[p];
        site[p] = s;
        contrib = s * pw[p];
    }
    const double b = block_sum(contrib, sh);
    if (threadIdx.x == 0) bsum[blockIdx.x] = b;
}

// Several roots combined into one likelihood per pattern
// (calcRootLogLikelihoodsMulti, BeagleCPUImpl.hpp:2474-2586): subset s
// contributes sum_i pi_s[i] sum_c w_s[c] root_s[c,p,i]; where scale factors
// apply, each is multiplied by exp(scale_s[p] - max_s scale_s[p]) and the
// maximum is added back after the log.  A subset's scale is its cumulative
// factors (logs) or its root's automatic-scaling exponents.
template <typename R>
__global__ __launch_bounds__(WG) void k_root_multi(const RootSubset* __restrict__ subs, int count,
                                                   const double* __restrict__ pw,
                                                   double* __restrict__ site,
                                                   double* __restrict__ bsum, Dims d) {
    __shared__ double sh[WG / 64];
    const int p = blockIdx.x * WG + threadIdx.x;
    double contrib = 0.0;
    if (p < d.P) {
        const double ln2 = 0.69314718055994530942;
        bool scaled = false;
        double mx = 0.0;
        for (int s = 0; s < count; ++s) {
            const RootSubset& U = subs[s];
            if (!U.cum && !U.ex) continue;
            const double sc = (U.cum ? U.cum[p] : 0.0) + (U.ex ? ln2 * (double)U.ex[p] : 0.0);
            mx = scaled ? fmax(mx, sc) : sc;
            scaled = true;
        }
        double total = 0.0;
        for (int s = 0; s < count; ++s) {
            const RootSubset& U = subs[s];
            const R* root = reinterpret_cast<const R*>(U.root);
            double sum = 0.0;
            for (int i = 0; i < d.S; ++i) {
                double t = 0.0;
                for (int c = 0; c < d.C; ++c)
                    t += (double)root[pbase(d, c, p) + (long)i * d.sStride] * U.cw[c];
                sum += U.freqs[i] * t;
            }
            if (scaled) {
                const double sc = (U.cum ? U.cum[p] : 0.0) + (U.ex ? ln2 * (double)U.ex[p] : 0.0);
                sum *= exp(sc - mx);
            }
            total += sum;
        }
        const double v = log(total) + mx;
        site[p] = v;
        contrib = v * pw[p];
    }
    const double b = block_sum(contrib, sh);
    if (threadIdx.x == 0) bsum[blockIdx.x] = b;
}

template <typename R>
void launch_root_multi(const RootSubset* subs, int count, const double* pw, double* site,
                       double* bsum, const Dims& d, hipStream_t st) {
    const int nb = (d.P + WG - 1) / WG;
    hipLaunchKernelGGL(k_root_multi<R>, dim3(nb), dim3(WG), 0, st, subs, count, pw, site, bsum, d);
}

// Per-category root log-likelihoods, no category weights
// (calcRootLogLikelihoodsPerCategory, BeagleCPUImpl.hpp:2589-2625):
// out[c P + p] = log(sum_i root[c,p,i] pi_i) + the pattern's scale.
template <typename R>
__global__ __launch_bounds__(WG) void k_root_per_category(const R* __restrict__ root,
                                                          const double* __restrict__ freqs,
                                                          const double* __restrict__ cum,
                                                          const short* __restrict__ ex,
                                                          double* __restrict__ out, Dims d) {
    const int p = blockIdx.x * WG + threadIdx.x, c = blockIdx.y;
    if (p >= d.P) return;
    double sum = 0.0;
    for (int i = 0; i < d.S; ++i) sum += (double)root[pbase(d, c, p) + (long)i * d.sStride] * freqs[i];
    double v = log(sum);
    if (cum) v += cum[p];
    if (ex) v += 0.69314718055994530942 * (double)ex[p];
    out[(long)c * d.P + p] = v;
}

template <typename R>
void launch_root_per_category(const void* root, const double* freqs, const double* cum,
                              const short* ex, double* out, const Dims& d, hipStream_t st) {
    hipLaunchKernelGGL(k_root_per_category<R>, dim3((d.P + WG - 1) / WG, d.C), dim3(WG), 0, st,
                       reinterpret_cast<const R*>(root), freqs, cum, ex, out, d);
}

// S > 4 (pattern-minor tiles): a workgroup takes 32 patterns, its eight
// wavefront halves take every eighth state, so each load instruction reads
// whole 128-byte rows; the eight partial sums are added in a fixed order.
template <typename R>
__global__ __launch_bounds__(WG) void k_root_tiled(const R* __restrict__ root,
                                                   const double* __restrict__ cw,
                                                   const double* __restrict__ freqs,
                                                   const double* __restrict__ cum,
                                                   const double* __restrict__ pw,
                                                   const short* __restrict__ ex,
                                                   double* __restrict__ site,
                                                   double* __restrict__ bsum, int p0, int p1,
                                                   Dims d) {
    __shared__ double part[WG / 32][32];
    __shared__ double sh[WG / 64];
    const int n = threadIdx.x & 31, slice = threadIdx.x >> 5;
    const int p = p0 + blockIdx.x * 32 + n;
    double acc = 0.0;
    if (p < p1) {
        for (int i = slice; i < d.S; i += WG / 32) {
            double t = 0.0;
            for (int c = 0; c < d.C; ++c)
                t += (double)root[pbase(d, c, p) + (long)i * PTILE] * cw[c];
            acc += freqs[i] * t;
        }
    }
    part[slice][n] = acc;
    __syncthreads();
    double contrib = 0.0;
    if (slice == 0 && p < p1) {
        double sum = 0.0;
        for (int k = 0; k < WG / 32; ++k) sum += part[k][n];
        double v = log(sum);
        if (cum) v += cum[p];
        if (ex) v += 0.69314718055994530942 * (double)ex[p];
        site[p] = v;
        contrib = v * pw[p];
    }
    const double b = block_sum(contrib, sh);
    if (threadIdx.x == 0) bsum[blockIdx.x] = b;
}

template <typename R>
static int root_blocks_and_launch(const void* root, const short* rootExp, const double* cw,
                                  const double* freqs, const double* cum, const double* pw,
                                  double* site, double* bsum, int p0, int p1, const Dims& d,
                                  hipStream_t st) {
    if (d.sStride == PTILE) {
        const int nb = (p1 - p0 + 31) / 32;
        if (nb > 0)
            hipLaunchKernelGGL(k_root_tiled<R>, dim3(nb), dim3(WG), 0, st,
                               reinterpret_cast<const R*>(root), cw, freqs, cum, pw, rootExp, site,
                               bsum, p0, p1, d);
        return nb;
    }
    const int nb = (p1 - p0 + WG - 1) / WG;
    if (nb > 0)
        hipLaunchKernelGGL(k_root<R>, dim3(nb), dim3(WG), 0, st, reinterpret_cast<const R*>(root),
                           cw, freqs, cum, pw, rootExp, site, bsum, p0, p1, d);
    return nb;
}

int root_block_count(const Dims& d, int p0, int p1) {
    if (p1 <= p0) return 0;
    return d.sStride == PTILE ? (p1 - p0 + 31) / 32 : (p1 - p0 + WG - 1) / WG;
}

template <typename R>
void launch_root(const void* root, const short* rootExp, const double* catWeights,
                 const double* freqs, const double* cumulative, const double* patternWeights,
                 double* siteLogL, double* blockSums, double* result, const Dims& d,
                 hipStream_t st) {
    const int nb = root_blocks_and_launch<R>(root, rootExp, catWeights, freqs, cumulative,
                                             patternWeights, siteLogL, blockSums, 0, d.P, d, st);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(WG), 0, st, blockSums, nb, 1, result, 0);
}

template <typename R>
void launch_root_range(const void* root, const short* rootExp, const double* catWeights,
                       const double* freqs, const double* cumulative,
                       const double* patternWeights, double* siteLogL, double* blockSums,
                       double* result, int p0, int p1, const Dims& d, hipStream_t st) {
    if (p1 <= p0) {
        launch_fill(result, 0.0, 1, st);
        return;
    }
    const int nb = root_blocks_and_launch<R>(root, rootExp, catWeights, freqs, cumulative,
                                             patternWeights, siteLogL, blockSums, p0, p1, d, st);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(WG), 0, st, blockSums, nb, 1, result, 0);
}

// ---------------------------------------------------------------------------
// edge derivatives: per node and pattern
//   num = sum_c w_c sum_i pre_i sum_j D_ij post_j,  den = sum_c w_c pre.post
// A state child uses D[i][state] and pre[state % S] (the reference's
// indexing, BeagleCPUImpl.hpp:2086-2119).  Outputs ratio per site and
// workgroup partial sums of w*ratio and w*ratio^2.

template <typename R>
__global__ __launch_bounds__(WG) void k_edge_deriv(const EdgeJob* __restrict__ jobs,
                                                   const double* __restrict__ cw,
                                                   const double* __restrict__ pw,
                                                   double* __restrict__ sites,
                                                   double* __restrict__ bsum, int transposedD,
                                                   Dims d) {
    __shared__ double sh[WG / 64];
    const int n = blockIdx.y;
    const EdgeJob J = jobs[n];
    const int p = blockIdx.x * WG + threadIdx.x;
    const int S = d.S, W = S + 1;
    double r1 = 0.0, r2 = 0.0;
    if (p < d.P) {
        const R* pre = reinterpret_cast<const R*>(J.pre);
        const R* post = reinterpret_cast<const R*>(J.post);
        const R* D = reinterpret_cast<const R*>(J.D);
        double num = 0.0, den = 0.0;
        for (int c = 0; c < d.C; ++c) {
            const long off = pbase(d, c, p);
            const int ss = d.sStride;
            const R* Dc = D + (long)c * S * W;
            double nc = 0.0, dc = 0.0;
            if (J.states) {
                const int s = J.states[p];
                for (int i = 0; i < S; ++i) {
                    double dij;
                    if (transposedD) dij = s < S ? (double)Dc[s * W + i] : 0.0;
                    else dij = (double)Dc[i * W + s];
                    nc += dij * (double)pre[off + (long)i * ss];
                }
                dc = (double)pre[off + (long)(s % S) * ss];
            } else {
                for (int i = 0; i < S; ++i) {
                    double t = 0.0;
                    if (transposedD)
                        for (int j = 0; j < S; ++j)
                            t += (double)Dc[j * W + i] * (double)post[off + (long)j * ss];
                    else
                        for (int j = 0; j < S; ++j)
                            t += (double)Dc[i * W + j] * (double)post[off + (long)j * ss];
                    const double pi = (double)pre[off + (long)i * ss];
                    nc += t * pi;
                    dc += (double)post[off + (long)i * ss] * pi;
                }
            }
            num += cw[c] * nc;
            den += cw[c] * dc;
        }
        const double r = num / den;
        if (sites) sites[(long)n * d.P + p] = r;
        r1 = r * pw[p];
        r2 = r * r * pw[p];
    }
    const int nb = gridDim.x;
    const double a = block_sum(r1, sh);
    const double b = block_sum(r2, sh);
    if (threadIdx.x == 0) {
        bsum[(long)(2 * n) * nb + blockIdx.x] = a;
        bsum[(long)(2 * n + 1) * nb + blockIdx.x] = b;
    }
}

// 4-state edge derivatives: one lane per pattern over all categories, 16-
// or 32-byte vector loads of the pre- and post-order partials, the job's
// derivative matrix through scalar loads.  Same sums and per-site ratios as
// k_edge_deriv (a tip's denominator is pre[state % 4], as in the
// reference).  Lanes past the last pattern read the last pattern and
// contribute nothing.
template <typename R>
__global__ __launch_bounds__(WG) void k_edge_deriv4(const EdgeJob* __restrict__ jobs,
                                                    const double* __restrict__ cw,
                                                    const double* __restrict__ pw,
                                                    double* __restrict__ sites,
                                                    double* __restrict__ bsum, Dims d) {
    __shared__ double sh[WG / 64];
    const int n = blockIdx.y;
    const EdgeJob J = jobs[n];
    const int p = blockIdx.x * WG + threadIdx.x;
    const int pc = min(p, d.P - 1);
    const R* __restrict__ D = reinterpret_cast<const R*>(J.D);
    double num = 0.0, den = 0.0;
    if (J.states) {
        const int s = gin<int>(J.states)[pc];
        for (int c = 0; c < d.C; ++c) {
            R pre[4];
            load4(gin<R>(J.pre) + c * d.catStride + (long)pc * 4, pre);
            const R* Dc = D + c * 20;
            double nc = 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const R* row = Dc + i * 5;
                const R dij = s == 0 ? row[0] : s == 1 ? row[1] : s == 2 ? row[2]
                                                : s == 3 ? row[3] : row[4];
                nc += (double)dij * (double)pre[i];
            }
            const int sm = s & 3;
            const R dc = sm == 0 ? pre[0] : sm == 1 ? pre[1] : sm == 2 ? pre[2] : pre[3];
            num += cw[c] * nc;
            den += cw[c] * (double)dc;
        }
    } else {
        for (int c = 0; c < d.C; ++c) {
            R pre[4], post[4];
            const long off = c * d.catStride + (long)pc * 4;
            load4(gin<R>(J.pre) + off, pre);
            load4(gin<R>(J.post) + off, post);
            const R* Dc = D + c * 20;
            double nc = 0.0, dc = 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const R* row = Dc + i * 5;
                const double t = (double)row[0] * (double)post[0] + (double)row[1] * (double)post[1] +
                                 (double)row[2] * (double)post[2] + (double)row[3] * (double)post[3];
                nc += t * (double)pre[i];
                dc += (double)post[i] * (double)pre[i];
            }
            num += cw[c] * nc;
            den += cw[c] * dc;
        }
    }
    double r1 = 0.0, r2 = 0.0;
    if (p < d.P) {
        const double r = num / den;
        if (sites) sites[(long)n * d.P + p] = r;
        r1 = r * pw[p];
        r2 = r * r * pw[p];
    }
    const int nb = gridDim.x;
    const double a = block_sum(r1, sh);
    const double b = block_sum(r2, sh);
    if (threadIdx.x == 0) {
        bsum[(long)(2 * n) * nb + blockIdx.x] = a;
        bsum[(long)(2 * n + 1) * nb + blockIdx.x] = b;
    }
}

__global__ __launch_bounds__(WG) void k_split_pairs(const double* __restrict__ rows, int njobs,
                                                    double* sums, double* sumSq) {
    const int n = blockIdx.x * WG + threadIdx.x;
    if (n >= njobs) return;
    if (sums) sums[n] = rows[2 * n];
    if (sumSq) sumSq[n] = rows[2 * n + 1];
}

template <typename R>
void launch_edge_derivatives(const EdgeJob* jobs, int njobs, const double* catWeights,
                             const double* patternWeights, double* sites, double* blockSums,
                             double* sums, double* sumSq, int transposedD, const Dims& d,
                             hipStream_t st) {
    if (njobs <= 0) return;
    const int nb = (d.P + WG - 1) / WG;
    if (d.S == 4 && d.Sp == 4 && !transposedD)
        hipLaunchKernelGGL(k_edge_deriv4<R>, dim3(nb, njobs), dim3(WG), 0, st, jobs, catWeights,
                           patternWeights, sites, blockSums, d);
    else
        hipLaunchKernelGGL(k_edge_deriv<R>, dim3(nb, njobs), dim3(WG), 0, st, jobs, catWeights,
                           patternWeights, sites, blockSums, transposedD, d);
    // rows (2n, 2n+1) hold the per-workgroup sums of node n
    double* rowsOut = blockSums + (long)2 * njobs * nb;
    hipLaunchKernelGGL(k_reduce_rows, dim3(2 * njobs), dim3(WG), 0, st, blockSums, nb, 2 * njobs,
                       rowsOut, 0);
    hipLaunchKernelGGL(k_split_pairs, dim3((njobs + WG - 1) / WG), dim3(WG), 0, st, rowsOut, njobs,
                       sums, sumSq);
}

// ---------------------------------------------------------------------------
// cross products: out[k*S+j] += sum_n sum_p (w_p/den_p) sum_c w_c r_c t_n
// pre[c,p,k] post[c,p,j]; a state child is the one-hot vector of its state
// (all ones when missing), which reproduces the reference's special cases
// (BeagleCPUImpl.hpp:2122-2237, BeagleCPU4StateImpl.hpp:1088-1254).  Every
// workgroup writes an S x S partial; a fixed-order column reduction adds
// them into out.

// 4 states: one lane per pattern walks a subset of the nodes (node group
// blockIdx.y) keeping the 16 sums in registers.
constexpr int XP4_GROUPS = 8;

template <typename R>
__global__ __launch_bounds__(WG) void k_cross4(const EdgeJob* __restrict__ jobs, int njobs,
                                               const double* __restrict__ cw,
                                               const double* __restrict__ rates,
                                               const double* __restrict__ pw,
                                               double* __restrict__ partial, Dims d) {
    __shared__ double sh[WG / 64];
    const int p = blockIdx.x * WG + threadIdx.x;
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0;
    if (p < d.P) {
        for (int n = blockIdx.y; n < njobs; n += gridDim.y) {
            const EdgeJob J = jobs[n];
            const R* pre = reinterpret_cast<const R*>(J.pre);
            const R* post = reinterpret_cast<const R*>(J.post);
            double tmp[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) tmp[e] = 0.0;
            double den = 0.0;
            for (int c = 0; c < d.C; ++c) {
                const long off = c * d.catStride + (long)p * 4;
                R a[4], b[4];
                load4(pre + off, a);
                if (J.states) {
                    const int s = J.states[p];
#pragma unroll
                    for (int j = 0; j < 4; ++j) b[j] = (s >= 4 || s == j) ? (R)1 : (R)0;
                } else {
                    load4(post + off, b);
                }
                double dd = 0.0;
#pragma unroll
                for (int i = 0; i < 4; ++i) dd += (double)a[i] * (double)b[i];
                den += dd * cw[c];
                const double ws = cw[c] * (rates[c] * J.scale);
#pragma unroll
                for (int k = 0; k < 4; ++k)
#pragma unroll
                    for (int j = 0; j < 4; ++j) tmp[k * 4 + j] += (double)a[k] * ((double)b[j] * ws);
            }
            const double f = pw[p] / den;
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[e] += tmp[e] * f;
        }
    }
    const long row = (long)blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const double v = block_sum(acc[e], sh);
        if (threadIdx.x == 0) partial[row * 16 + e] = v;
    }
}

// any S <= 64: workgroups walk (node, 16-pattern chunk) pairs, staging the
// chunk through LDS; each lane owns up to 16 of the S x S sums.
constexpr int XP_CHUNK = 16;
constexpr int XP_EPT = 16;
constexpr int XP_BLOCKS = 512;

template <typename R>
__global__ __launch_bounds__(WG) void k_cross(const EdgeJob* __restrict__ jobs, int njobs,
                                              const double* __restrict__ cw,
                                              const double* __restrict__ rates,
                                              const double* __restrict__ pw,
                                              double* __restrict__ partial, Dims d) {
    extern __shared__ double sh[];
    const int S = d.S, C = d.C, S2 = S * S;
    double* spre = sh;                              // [XP_CHUNK][C][S]
    double* spost = spre + XP_CHUNK * C * S;        // [XP_CHUNK][C][S]
    double* sa = spost + XP_CHUNK * C * S;          // [XP_CHUNK][C]
    double acc[XP_EPT];
#pragma unroll
    for (int t = 0; t < XP_EPT; ++t) acc[t] = 0.0;
    const int nchunks = (d.P + XP_CHUNK - 1) / XP_CHUNK;
    const long npairs = (long)njobs * nchunks;
    for (long pair = blockIdx.x; pair < npairs; pair += gridDim.x) {
        const int n = (int)(pair / nchunks);
        const int p0 = (int)(pair % nchunks) * XP_CHUNK;
        const EdgeJob J = jobs[n];
        const R* pre = reinterpret_cast<const R*>(J.pre);
        const R* post = reinterpret_cast<const R*>(J.post);
        for (int t = threadIdx.x; t < XP_CHUNK * C * S; t += WG) {
            const int i = t % S;
            const int c = (t / S) % C;
            const int q = t / (S * C);
            const int p = p0 + q;
            double a = 0.0, b = 0.0;
            if (p < d.P) {
                const long off = pbase(d, c, p) + (long)i * d.sStride;
                a = (double)pre[off];
                if (J.states) {
                    const int s = J.states[p];
                    b = (s >= S || s == i) ? 1.0 : 0.0;
                } else {
                    b = (double)post[off];
                }
            }
            spre[t] = a;
            spost[t] = b;
        }
        __syncthreads();
        if (threadIdx.x < XP_CHUNK) {
            const int q = threadIdx.x;
            const int p = p0 + q;
            double den = 0.0;
            for (int c = 0; c < C; ++c) {
                double dd = 0.0;
                for (int i = 0; i < S; ++i)
                    dd += spre[(q * C + c) * S + i] * spost[(q * C + c) * S + i];
                den += dd * cw[c];
            }
            const double f = p < d.P ? pw[p] / den : 0.0;
            for (int c = 0; c < C; ++c)
                sa[q * C + c] = p < d.P ? f * (cw[c] * (rates[c] * J.scale)) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < XP_EPT; ++t) {
            const int e = threadIdx.x + t * WG;
            if (e < S2) {
                const int k = e / S, j = e - k * S;
                double v = 0.0;
                for (int q = 0; q < XP_CHUNK; ++q)
                    for (int c = 0; c < C; ++c) {
                        const int o = (q * C + c) * S;
                        v += sa[q * C + c] * spre[o + k] * spost[o + j];
                    }
                acc[t] += v;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < XP_EPT; ++t) {
        const int e = threadIdx.x + t * WG;
        if (e < S2) partial[(long)blockIdx.x * S2 + e] = acc[t];
    }
}

// out[e] += sum over rows of partial[row][e]
__global__ __launch_bounds__(WG) void k_reduce_cols(const double* __restrict__ partial, int rows,
                                                    int cols, double* out) {
    const int e = blockIdx.x * WG + threadIdx.x;
    if (e >= cols) return;
    double acc = 0.0;
    for (int r = 0; r < rows; ++r) acc += partial[(long)r * cols + e];
    out[e] += acc;
}

void launch_reduce_rows(const double* in, int n, int rows, double* out, hipStream_t st) {
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_reduce_rows, dim3(rows), dim3(WG), 0, st, in, n, rows, out, 0);
}

int cross_products_rows(const Dims& d) {
    if (d.S == 4) return XP4_GROUPS * ((d.P + WG - 1) / WG);
    return XP_BLOCKS;
}

template <typename R>
void launch_cross_products(const EdgeJob* jobs, int njobs, const double* catWeights,
                           const double* catRates, const double* patternWeights,
                           double* blockPartials, double* out, const Dims& d, hipStream_t st) {
    if (njobs <= 0) return;
    const int cols = d.S * d.S;
    const int rows = cross_products_rows(d);
    if (d.S == 4) {
        const int nb = (d.P + WG - 1) / WG;
        hipLaunchKernelGGL(k_cross4<R>, dim3(nb, XP4_GROUPS), dim3(WG), 0, st, jobs, njobs,
                           catWeights, catRates, patternWeights, blockPartials, d);
    } else {
        const size_t shm = (size_t)(2 * XP_CHUNK * d.C * d.S + XP_CHUNK * d.C) * sizeof(double);
        hipLaunchKernelGGL(k_cross<R>, dim3(XP_BLOCKS), dim3(WG), shm, st, jobs, njobs, catWeights,
                           catRates, patternWeights, blockPartials, d);
    }
    hipLaunchKernelGGL(k_reduce_cols, dim3((cols + WG - 1) / WG), dim3(WG), 0, st, blockPartials,
                       rows, cols, out);
}

// ---------------------------------------------------------------------------
// edge log-likelihood with optional first/second derivatives along the edge
// (BeagleCPUImpl.hpp:3209-3322, 3868-3994).

template <typename R>
__global__ __launch_bounds__(WG) void k_edge_ll(const R* __restrict__ par, const R* child,
                                                const int* states, const R* Pm, const R* D1,
                                                const R* D2, const double* __restrict__ cw,
                                                const double* __restrict__ f,
                                                const double* cumA, const double* cumB,
                                                const short* exA, const short* exB,
                                                const double* __restrict__ pw, double* siteL,
                                                double* siteD1, double* siteD2,
                                                double* __restrict__ bsum, int p0, int p1,
                                                Dims d) {
    __shared__ double sh[WG / 64];
    const int p = p0 + blockIdx.x * WG + threadIdx.x;
    const int S = d.S, W = S + 1;
    double cl = 0.0, c1 = 0.0, c2 = 0.0;
    if (p < p1) {
        double acc[3] = {0.0, 0.0, 0.0};
        const R* Ms[3] = {Pm, D1, D2};
        for (int m = 0; m < 3; ++m) {
            if (!Ms[m]) continue;
            for (int c = 0; c < d.C; ++c) {
                const long off = pbase(d, c, p);
                const int ss = d.sStride;
                const R* Mc = Ms[m] + (long)c * S * W;
                double s = 0.0;
                for (int i = 0; i < S; ++i) {
                    double t;
                    if (states) {
                        t = (double)Mc[i * W + states[p]];
                    } else {
                        t = 0.0;
                        for (int j = 0; j < S; ++j)
                            t += (double)Mc[i * W + j] * (double)child[off + (long)j * ss];
                    }
                    s += f[i] * (double)par[off + (long)i * ss] * t;
                }
                acc[m] += cw[c] * s;
            }
        }
        double L = log(acc[0]);
        if (cumA) L += cumA[p];
        if (cumB) L += cumB[p];
        if (exA) L += 0.69314718055994530942 * (double)exA[p];
        if (exB) L += 0.69314718055994530942 * (double)exB[p];
        siteL[p] = L;
        cl = L * pw[p];
        if (D1) {
            const double g = acc[1] / acc[0];
            siteD1[p] = g;
            c1 = g * pw[p];
            if (D2) {
                const double h = acc[2] / acc[0] - g * g;
                siteD2[p] = h;
                c2 = h * pw[p];
            }
        }
    }
    const int nb = gridDim.x;
    const double a = block_sum(cl, sh);
    const double b = block_sum(c1, sh);
    const double e = block_sum(c2, sh);
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = a;
        bsum[nb + blockIdx.x] = b;
        bsum[2 * nb + blockIdx.x] = e;
    }
}

template <typename R>
void launch_edge_loglik(const void* parent, const void* child, const int* childStates,
                        const void* P, const void* D1, const void* D2,
                        const double* catWeights, const double* freqs,
                        const double* cumA, const double* cumB, const short* exA,
                        const short* exB, const double* patternWeights, double* siteL,
                        double* siteD1, double* siteD2, double* blockSums, double* result,
                        int p0, int p1, const Dims& d, hipStream_t st) {
    if (p1 <= p0) {
        launch_fill(result, 0.0, 3, st);
        return;
    }
    const int nb = (p1 - p0 + WG - 1) / WG;
    hipLaunchKernelGGL(k_edge_ll<R>, dim3(nb), dim3(WG), 0, st, reinterpret_cast<const R*>(parent),
                       reinterpret_cast<const R*>(child), childStates,
                       reinterpret_cast<const R*>(P), reinterpret_cast<const R*>(D1),
                       reinterpret_cast<const R*>(D2), catWeights, freqs, cumA, cumB, exA, exB,
                       patternWeights, siteL, siteD1, siteD2, blockSums, p0, p1, d);
    hipLaunchKernelGGL(k_reduce_rows, dim3(3), dim3(WG), 0, st, blockSums, nb, 3, result, 0);
}

// ---------------------------------------------------------------------------
// scale-buffer bookkeeping (BeagleCPUImpl.hpp:2763-2908 semantics)

__global__ __launch_bounds__(WG) void k_accumulate(double* cum, const double* const* bufs,
                                                   int nbufs, double sign, int logScalers, int p0,
                                                   int p1) {
    const int p = p0 + blockIdx.x * WG + threadIdx.x;
    if (p >= p1) return;
    double v = cum[p];
    for (int b = 0; b < nbufs; ++b) {
        const double s = bufs[b][p];
        v += sign * (logScalers ? s : log(s));
    }
    cum[p] = v;
}

void launch_accumulate(double* cum, const double* const* bufs, int nbufs, double sign,
                       int logScalers, int p0, int p1, hipStream_t st) {
    if (p1 <= p0) return;
    hipLaunchKernelGGL(k_accumulate, dim3((p1 - p0 + WG - 1) / WG), dim3(WG), 0, st, cum, bufs,
                       nbufs, sign, logScalers, p0, p1);
}

__global__ __launch_bounds__(WG) void k_fill(double* dst, double v, long n) {
    const long i = (long)blockIdx.x * WG + threadIdx.x;
    if (i < n) dst[i] = v;
}

__global__ __launch_bounds__(WG) void k_fill_short(short* dst, short v, long n) {
    const long i = (long)blockIdx.x * WG + threadIdx.x;
    if (i < n) dst[i] = v;
}

void launch_fill_short(short* dst, short v, long n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_short, dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, st, dst, v,
                       n);
}

void launch_fill(double* dst, double v, long n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3((n + WG - 1) / WG), dim3(WG), 0, st, dst, v, n);
}

// dst[c][p][i] = v[i] for every category and pattern (root pre-order
// partials = state frequencies).
template <typename R>
__global__ __launch_bounds__(WG) void k_tile_vector(R* dst, const double* __restrict__ v, Dims d) {
    const long t = (long)blockIdx.x * WG + threadIdx.x;
    const long n = (long)d.C * d.Ppad * d.S;
    if (t >= n) return;
    const int i = t % d.S;
    const long cp = t / d.S;
    const long p = cp % d.Ppad, c = cp / d.Ppad;
    dst[pbase(d, (int)c, (int)p) + (long)i * d.sStride] = (R)v[i];
}

template <typename R>
void launch_tile_vector(void* dst, const double* v, const Dims& d, hipStream_t st) {
    const long n = (long)d.C * d.Ppad * d.S;
    hipLaunchKernelGGL(k_tile_vector<R>, dim3((n + WG - 1) / WG), dim3(WG), 0, st,
                       reinterpret_cast<R*>(dst), v, d);
}

template <typename R>
__global__ __launch_bounds__(WG) void k_pad_set(R* dst, long n, R v) {
    const long i = (long)blockIdx.x * WG + threadIdx.x;
    if (i < n) dst[i] = v;
}

template <typename R>
void launch_pad_set(void* dst, long n, R v, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pad_set<R>, dim3((n + WG - 1) / WG), dim3(WG), 0, st,
                       reinterpret_cast<R*>(dst), n, v);
}

// Per category: dst[j][i] = src[i][j] for the S x S block; each row keeps
// its own pad entry.  Staged through LDS so src == dst is allowed.
template <typename R>
__global__ __launch_bounds__(WG) void k_transpose(const void* const* src, void* const* dst, int S) {
    extern __shared__ double sh[];
    const int m = blockIdx.x, c = blockIdx.y;
    const int W = S + 1;
    const R* a = reinterpret_cast<const R*>(src[m]) + (long)c * S * W;
    R* b = reinterpret_cast<R*>(dst[m]) + (long)c * S * W;
    for (int t = threadIdx.x; t < S * W; t += WG) sh[t] = (double)a[t];
    __syncthreads();
    for (int t = threadIdx.x; t < S * W; t += WG) {
        const int i = t / W, j = t - i * W;
        b[t] = (R)(j == S ? sh[i * W + S] : sh[j * W + i]);
    }
}

template <typename R>
void launch_transpose(const void* const* src, void* const* dst, int count, int S, int C,
                      hipStream_t st) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_transpose<R>, dim3(count, C), dim3(WG), (size_t)S * (S + 1) * sizeof(double),
                       st, src, dst, S);
}

// ---------------------------------------------------------------------------
// explicit instantiations

#define INSTANTIATE(R)                                                                          \
    template void launch_convolve<R>(const ConvolveJob*, int, int, int, hipStream_t);           \
    template void launch_matrices<R>(const MatrixJob*, int, int, int, hipStream_t);             \
    template void launch_post<R>(const PostJob*, const PostJob*, int, int, const Dims&,         \
                                 const void*,                                                    \
                                 void*, hipStream_t);                                            \
    template void launch_post_levels<R>(const PostJob*, const int*, int, const Dims&,            \
                                        const void*, hipStream_t);                               \
    template void launch_pre<R>(const PreJob*, const PreJob*, int, const Dims&, int, void*,      \
                                const void*, void*, const double*, const double*, double*,       \
                                hipStream_t);                                                    \
    template void launch_gather_node_matrices<R>(const NodeJob*, int, const Dims&, const void*, \
                                                 void*, hipStream_t, bool);                  \
    template void launch_qgrad<R>(const NodeJob*, int, const Dims&, const void*, const double*,  \
                                  const double*, void*, const double*, const double*, double*,  \
                                  hipStream_t);                                                 \
    template void launch_root_freq_grad<R>(const void*, const double*, const double*,          \
                                           const double*, double*, const Dims&, hipStream_t);   \
    template void launch_root_multi<R>(const RootSubset*, int, const double*, double*, double*,  \
                                       const Dims&, hipStream_t);                               \
    template void launch_root_per_category<R>(const void*, const double*, const double*,        \
                                              const short*, double*, const Dims&, hipStream_t);  \
    template void launch_qgrad_generic<R>(const PreJob*, int, const Dims&, const double*,       \
                                          const double*, const double*, const double*, double*,  \
                                          double*, int, double*, hipStream_t);                  \
    template void launch_pre_nodes<R>(const NodeJob*, int, bool, int, const Dims&, const void*, \
                                      void*, const double*, const double*, double*,         \
                                      hipStream_t);                                         \
    template void launch_pre_levels<R>(const PreJob*, const int*, int, const Dims&, const void*, \
                                       void*, const double*, const double*, double*,             \
                                       hipStream_t);                                             \
    template void launch_root<R>(const void*, const short*, const double*, const double*,        \
                                 const double*, const double*, double*, double*, double*,        \
                                 const Dims&, hipStream_t);                                      \
    template void launch_root_range<R>(const void*, const short*, const double*, const double*,  \
                                       const double*, const double*, double*, double*, double*,  \
                                       int, int, const Dims&, hipStream_t);                      \
    template void launch_edge_derivatives<R>(const EdgeJob*, int, const double*, const double*,  \
                                             double*, double*, double*, double*, int,            \
                                             const Dims&, hipStream_t);                          \
    template void launch_cross_products<R>(const EdgeJob*, int, const double*, const double*,    \
                                           const double*, double*, double*, const Dims&,         \
                                           hipStream_t);                                         \
    template void launch_edge_loglik<R>(const void*, const void*, const int*, const void*,       \
                                        const void*, const void*, const double*, const double*,  \
                                        const double*, const double*, const short*, const short*, \
                                        const double*, double*, double*, double*, double*,       \
                                        double*, int, int, const Dims&, hipStream_t);            \
    template void launch_transpose<R>(const void* const*, void* const*, int, int, int,           \
                                      hipStream_t);                                              \
    template void launch_pad_set<R>(void*, long, R, hipStream_t);                               \
    template void launch_tile_vector<R>(void*, const double*, const Dims&, hipStream_t);

INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace hipbeagle
