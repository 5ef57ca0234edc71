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
#include "device_common.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace hipbeagle {

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
    MatrixJob J = jobs[blockIdx.x];
    if (J.tHost) J.t = *J.tHost;   // a graph replay: the length set since capture
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

__global__ __launch_bounds__(WG) void k_copy_to_host(double* __restrict__ dst,
                                                     const double* __restrict__ src, long n) {
    for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG)
        dst[i] = src[i];
}

void launch_copy_to_host(double* dstHost, const double* src, long n, hipStream_t st) {
    if (n <= 0) return;
    const int nb = (int)std::min<long>((n + WG - 1) / WG, 64);
    hipLaunchKernelGGL(k_copy_to_host, dim3(nb), dim3(WG), 0, st, dstHost, src, n);
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
                                                     const R* __restrict__ mats, TileList w) {
    __shared__ R smax[2][8][TILE];
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    int job, item;
    tile_item(w, job, item);
    const PostJob J = jobs[job];
    const int p = ((w.first ? J.p0 / TILE : 0) + item) * TILE + lane;
    const long off = c * cs + (long)p * 4;
    int k = 0;
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

// The same with T pattern tiles per workgroup: grid = (ceil(tiles / T),
// operations).  Each wave (one category) issues the loads of its T tiles
// before any arithmetic, so the latency every workgroup pays before its
// first useful byte (the job record, the matrices, the first operands) is
// paid once per T tiles, and the T tiles' category maxima share one LDS
// exchange and one barrier.  Tiles past the last are clamped for their
// loads and store nothing.
template <typename R, bool AUTO, int T>
__global__ __launch_bounds__(512) void k_post4_level_mt(const PostJob* __restrict__ jobs, int P,
                                                        int C, long cs,
                                                        const R* __restrict__ mats) {
    __shared__ R smax[T][8][TILE];
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int ntiles = (P + TILE - 1) / TILE;
    const PostJob J = jobs[blockIdx.y];
    Operand4<R> A[T], B[T];
    int p[T];
    bool in[T];
#pragma unroll
    for (int i = 0; i < T; ++i) {
        const int t = (int)blockIdx.x * T + i;
        p[i] = min(t, ntiles - 1) * TILE + lane;
        in[i] = t < ntiles && p[i] < P && p[i] >= J.p0 && p[i] < J.p1;
        const long off = c * cs + (long)p[i] * 4;
        fetch4(J.x1, J.s1, off, p[i], A[i]);
        fetch4(J.x2, J.s2, off, p[i], B[i]);
    }
    const R* m1 = mats + J.mo1 + c * 20;
    const R* m2 = mats + J.mo2 + c * 20;
    R v[T][4];
#pragma unroll
    for (int i = 0; i < T; ++i) {
        R a[4], b[4];
        apply4(m1, J.s1 != nullptr, A[i], a);
        apply4(m2, J.s2 != nullptr, B[i], b);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[i][q] = a[q] * b[q];
        if (J.mode == SCALE_READ) {
            const R sr = (R)gin<double>(J.scaleRead)[p[i]];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[i][q] = v[i][q] / sr;
        }
    }
    if (AUTO || J.mode == SCALE_WRITE) {
#pragma unroll
        for (int i = 0; i < T; ++i) {
            const R m = fmax(fmax(v[i][0], v[i][1]), fmax(v[i][2], v[i][3]));
            smax[i][c][lane] = in[i] ? m : (R)0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
#pragma unroll
        for (int i = 0; i < T; ++i) {
            R mx = smax[i][0][lane];
            for (int cc = 1; cc < C; ++cc) mx = fmax(mx, smax[i][cc][lane]);
            if (in[i])
                rescale4<R, AUTO>(v[i], mx, c, p[i], J.scaleWrite, J.logScalers, J.cumulative,
                                  J.accA, J.accB, J.e1, J.e2, J.ed);
        }
    }
#pragma unroll
    for (int i = 0; i < T; ++i)
        if (in[i]) store4(gout<R>(J.dest) + c * cs + (long)p[i] * 4, v[i]);
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
                                                    double* partial, TileList w) {
    extern __shared__ double sdv[];   // [2][C][4][TILE]
    const int c = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    int job, item;
    tile_item(w, job, item);
    const PreJob J = jobs[job];
    const int tile = (w.first ? J.p0 / TILE : 0) + item;
    const int p = tile * TILE + lane;
    const long off = c * cs + (long)p * 4;
    const long doff = (long)(c * TILE + lane) * 4;
    const int ntiles = (P + TILE - 1) / TILE;
    int k = 0;
    R prev[4] = {0, 0, 0, 0}, uprev[4] = {0, 0, 0, 0};
    PreOps4<R> S0;
    pre_issue<R, DERIV>(J, dummyX, dummyS, off, doff, p, lane, S0);
    pre_compute<R, AUTO, DERIV>(J, S0, prev, uprev, mats, sdv, k, c, lane, p, P, C, off, doff,
                                dummyW, dummyE, cw, pw, partial, tile, ntiles);
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

void launch_qgx_eigen(const double* xs, int chunks, int nops, int S, int C, const double* eig,
                      const double* tau, double* z, hipStream_t st) {
    hipLaunchKernelGGL(k_qgx_eigen, dim3(C, nops), dim3(WG), (size_t)S * S * sizeof(double), st, xs,
                       chunks, nops, S, C, eig, tau, z);
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
    launch_qgx_eigen(xs, chunks, nops, d.S, d.C, eig, tau, z, st);
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

// Waves a level launch should leave per tile-count step (two rounds of
// eight waves on each of the 1,024 SIMDs).
constexpr long kPostLevelWaves = 16384;

// Traversal launchers: the fused 4-state kernels, or the generic-S path
// whose grids come from the host copy of each job.
template <typename R>
void launch_post_levels(const PostJob* djobs, const PostJob* hjobs, const int* levelStart,
                        int nlevels, const Dims& d, const void* mats, hipStream_t st,
                        TileList w, int items) {
    const int tiles = (d.P + TILE - 1) / TILE;
    const R* m = reinterpret_cast<const R*>(mats);
    if (w.first) {
        // one level over pattern ranges: one tile per workgroup, listed
        if (items <= 0) return;
        const PostJob* lj = djobs + levelStart[0];
        if (d.autoScale) hipLaunchKernelGGL((k_post4_level<R, true>), dim3(items), dim3(TILE * d.C), 0, st, lj, d.P, d.C, d.catStride, m, w);
        else hipLaunchKernelGGL((k_post4_level<R, false>), dim3(items), dim3(TILE * d.C), 0, st, lj, d.P, d.C, d.catStride, m, w);
        return;
    }
    for (int l = 0; l < nlevels; ++l) {
        const int n = levelStart[l + 1] - levelStart[l];
        const PostJob* lj = djobs + levelStart[l];
        // FP32 tiles per workgroup: 4 where most operations join two tips
        // (each tile moves few bytes), else 2, as long as about two rounds
        // of waves stay on the chip (kPostLevelWaves); 1 for narrow levels
        // and FP64 (two tiles' FP64 operands take 115 registers).
        int tipTip = 0;
        for (int o = levelStart[l]; o < levelStart[l + 1]; ++o)
            tipTip += hjobs[o].s1 != nullptr && hjobs[o].s2 != nullptr;
        const long waves = (long)n * tiles * d.C;
        int T = sizeof(R) == 4 && waves >= 2 * kPostLevelWaves ? 2 : 1;
        if (T == 2 && 2 * tipTip >= n && waves >= 4 * kPostLevelWaves) T = 4;
        const dim3 block(TILE * d.C);
        if constexpr (sizeof(R) == 4) {
            if (T > 1) {
                const dim3 grid((tiles + T - 1) / T, n);
                if (T == 4) {
                    if (d.autoScale) hipLaunchKernelGGL((k_post4_level_mt<R, true, 4>), grid, block, 0, st, lj, d.P, d.C, d.catStride, m);
                    else hipLaunchKernelGGL((k_post4_level_mt<R, false, 4>), grid, block, 0, st, lj, d.P, d.C, d.catStride, m);
                } else {
                    if (d.autoScale) hipLaunchKernelGGL((k_post4_level_mt<R, true, 2>), grid, block, 0, st, lj, d.P, d.C, d.catStride, m);
                    else hipLaunchKernelGGL((k_post4_level_mt<R, false, 2>), grid, block, 0, st, lj, d.P, d.C, d.catStride, m);
                }
                continue;
            }
        }
        {
            const dim3 grid(tiles, n);
            if (d.autoScale) hipLaunchKernelGGL((k_post4_level<R, true>), grid, block, 0, st, lj, d.P, d.C, d.catStride, m, TileList{nullptr, 0});
            else hipLaunchKernelGGL((k_post4_level<R, false>), grid, block, 0, st, lj, d.P, d.C, d.catStride, m, TileList{nullptr, 0});
        }
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
                       double* partial, hipStream_t st, TileList w, int items) {
    if (w.first && (partial || nlevels != 1)) w.first = nullptr;   // sums need every tile's slot
    if (w.first && items <= 0) return;
    const R* dummyX = reinterpret_cast<const R*>(dummy);
    R* dummyW = const_cast<R*>(dummyX) + 8 * TILE * 4;
    const double* dummyD = reinterpret_cast<const double*>(dummyW + 8 * TILE * 4);
    const int* dummyS = reinterpret_cast<const int*>(dummyD + TILE);
    const short* dummyE = reinterpret_cast<const short*>(dummyS + TILE);
    const R* m = reinterpret_cast<const R*>(mats);
    for (int l = 0; l < nlevels; ++l) {
        const int n = levelStart[l + 1] - levelStart[l];
        const dim3 grid = w.first ? dim3(items) : dim3((d.P + TILE - 1) / TILE, n);
        dim3 block(TILE * d.C);
        const size_t lds = pre_lds_bytes(d.C);
        const PreJob* j = djobs + levelStart[l];
#define PREL(A, D)                                                                                \
    hipLaunchKernelGGL((k_pre4_level<R, A, D>), grid, block, lds, st, j, d.P, d.C, d.catStride, m, \
                       dummyX, dummyS, dummyW, dummyE, cw, pw, partial, w)
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
                                             int p0, int p1, Dims d,
                                             double* __restrict__ fpart = nullptr) {
    __shared__ double sh[WG / 64];
    const int p = p0 + blockIdx.x * WG + threadIdx.x;
    double contrib = 0.0;
    double fg[4] = {0.0, 0.0, 0.0, 0.0};   // fpart (4 states): w_p t_i / L_p
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
            if (fpart)
                for (int i = 0; i < 4; ++i) fg[i] = pw[p] * t[i] / sum;
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
    if (fpart) {
        // dlogL/dpi_i through the root distribution, per workgroup: rows
        // [4][gridDim.x] (scale factors cancel in t_i / L_p)
        for (int i = 0; i < 4; ++i) {
            const double f = block_sum(fg[i], sh);
            if (threadIdx.x == 0) fpart[(long)i * gridDim.x + blockIdx.x] = f;
        }
    }
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

// The root's workgroup sums for a later reduction: rows[0][nb] of
// w_p site_p and, 4 states, rows[1..4][nb] of the root-frequency gradient.
// Returns nb.
template <typename R>
int launch_root_rows(const void* root, const short* rootExp, const double* catWeights,
                     const double* freqs, const double* cumulative, const double* patternWeights,
                     double* siteLogL, double* rows, const Dims& d, hipStream_t st) {
    const int nb = (d.P + WG - 1) / WG;
    hipLaunchKernelGGL(k_root<R>, dim3(nb), dim3(WG), 0, st, reinterpret_cast<const R*>(root),
                       catWeights, freqs, cumulative, patternWeights, rootExp, siteLogL, rows, 0,
                       d.P, d, rows + nb);
    return nb;
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
__global__ __launch_bounds__(WG) void k_tile_vector(R* dst, const double* __restrict__ v, Dims d,
                                                    short* __restrict__ ex) {
    const long t = (long)blockIdx.x * WG + threadIdx.x;
    const long n = (long)d.C * d.Ppad * d.S;
    if (ex && t < d.Ppad) ex[t] = 0;   // the buffer's scaling exponents
    if (t >= n) return;
    const int i = t % d.S;
    const long cp = t / d.S;
    const long p = cp % d.Ppad, c = cp / d.Ppad;
    dst[pbase(d, (int)c, (int)p) + (long)i * d.sStride] = (R)v[i];
}

template <typename R>
void launch_tile_vector(void* dst, const double* v, const Dims& d, hipStream_t st, short* ex) {
    const long n = (long)d.C * d.Ppad * d.S;
    hipLaunchKernelGGL(k_tile_vector<R>, dim3((n + WG - 1) / WG), dim3(WG), 0, st,
                       reinterpret_cast<R*>(dst), v, d, ex);
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
    template void launch_post_levels<R>(const PostJob*, const PostJob*, const int*, int, const Dims&,            \
                                        const void*, hipStream_t, TileList, int);                \
    template void launch_pre<R>(const PreJob*, const PreJob*, int, const Dims&, int, void*,      \
                                const void*, void*, const double*, const double*, double*,       \
                                hipStream_t);                                                    \
    template void launch_root_multi<R>(const RootSubset*, int, const double*, double*, double*,  \
                                       const Dims&, hipStream_t);                               \
    template void launch_root_per_category<R>(const void*, const double*, const double*,        \
                                              const short*, double*, const Dims&, hipStream_t);  \
    template void launch_qgrad_generic<R>(const PreJob*, int, const Dims&, const double*,       \
                                          const double*, const double*, const double*, double*,  \
                                          double*, int, double*, hipStream_t);                  \
    template void launch_pre_levels<R>(const PreJob*, const int*, int, const Dims&, const void*, \
                                       void*, const double*, const double*, double*,             \
                                       hipStream_t, TileList, int);                              \
    template int launch_root_rows<R>(const void*, const short*, const double*, const double*,    \
                                     const double*, const double*, double*, double*,            \
                                     const Dims&, hipStream_t);                                 \
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
    template void launch_tile_vector<R>(void*, const double*, const Dims&, hipStream_t, short*);

INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace hipbeagle
