// Device-side job descriptors shared by the host instance and the gfx950
// kernels, and the launch entry points.  Host code resolves every buffer
// index of an API call into device pointers once, so kernels never chase
// index tables: each job below is read with scalar (wave-uniform) loads.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hipbeagle {

// Rescale modes of one traversal operation (reference semantics,
// libhmsbeagle/CPU/BeagleCPUImpl.hpp:1555-1577).
enum ScaleMode : int {
    SCALE_NONE = 0,     // plain product
    SCALE_WRITE = 1,    // compute, then divide by the per-pattern maximum
    SCALE_READ = 2,     // divide by an existing (raw) factor
};

// One post-order operation.  A child is either partials (x != null) or
// compact tip states (s != null).  Matrices are [C][S][S+1] with the pad
// column read by a missing state.
struct PostJob {
    const void* x1; const int* s1; const void* m1;
    const void* x2; const int* s2; const void* m2;
    void* dest;
    double* scaleWrite;        // SCALE_WRITE target
    const double* scaleRead;   // SCALE_READ source
    double* cumulative;        // += log(max) when rescaling (may be null)
    const double* accA;        // scaling-always: child node scale buffers
    const double* accB;        //   added into scaleWrite (may be null)
    const short* e1;           // automatic scaling: children's and the
    const short* e2;           //   destination's per-pattern power-of-two
    short* ed;                 //   exponents (null for tips)
    int mode;
    int logScalers;            // store log(max) rather than max
    int p0, p1;                // pattern range (a partition, or [0, P))
    int mo1, mo2;              // m1, m2 as element offsets into the matrix slab
    int k1, k2;                // operand kinds: 0 partials, 1 tip states, 2 the
                               //   previous operation's output (same pattern range)
    int pk1, pk2;              // MFMA path: m1, m2 repacked, offsets in the pack slab
};

// One pre-order operation: dest = P_self^T (pre_parent * (P_sib x_sib)).
struct PreJob {
    const void* parent; const void* mSelf;
    const void* xs; const int* ss; const void* mSib;
    void* dest;
    double* scaleWrite;
    const double* accA;        // scaling-always accumulation (may be null)
    const double* accB;
    const short* e1;           // automatic scaling exponents: parent, sibling,
    const short* e2;           //   destination
    short* ed;
    int mode;
    int logScalers;
    int p0, p1;
    int moSelf, moSib;         // mSelf, mSib as element offsets into the matrix slab
    int kPar, kSib;            // operand kinds (see k_pre4)
    // fused derivatives: the node's own post-order operand and the two
    // derivative matrices (element offsets)
    const void* xself; const int* sself;
    int kSelf, moD1, moD2;
    int oidx;                  // row of this operation in the derivative output
    int pkSelf, pkSib;         // MFMA path (S > 4): packed P_self^T and P_sib
    // MFMA path, tip branch of the fused gradient: PACK_TIP_COLUMN of P_self
    // and PACK_TIP_PRODUCT of P_self with both derivative matrices (-1: none)
    int pkT, pkM1, pkM2;
};

// The two pre-order operations of one node's children a (slot 0) and b
// (slot 1), fused with both branches' derivative sums:
//   pre(a) = P_a^T (pre(v) * (P_b x_b)),  pre(b) = P_b^T (pre(v) * (P_a x_a))
// where x_a, x_b are the children's post-order partials (or tip states).
// Each of pre(v), x_a and x_b is read once for both operations.
struct NodeJob {
    const void* parent;        // pre-order partials of the node (kPar 0)
    const void* x[2];          // children's post-order partials, or tip states (k == 1)
    void* dest[2];             // children's pre-order partials; null: not kept
    const short* eParent;      // automatic scaling exponents: the node's pre-order
    const short* ex[2];        //   partials, the children's post-order partials
    short* ed[2];              //   (null for tips), the children's pre-order partials
    int k[2];                  // 0 partials, 1 tip states
    int kPar;                  // 0 load; 2, 3: previous job's dest[0], dest[1] (registers)
    int moP[2], moD1[2], moD2[2];  // matrix element offsets in the matrix slab
    int oidx[2];               // operation index of each child in the derivative output
    int mbase;                 // this job's block in the gathered matrices (reals)
    int mF[2];                 // the children's branch factors in the factor slab (doubles)
    // buffer-range sizes of the FP32 node kernels' accesses, set on the host
    // (-1: the whole buffer, 0: an empty range, so the access is skipped):
    // loads of the parent, each child's partials and states, the exponents;
    // stores of each child's pre-order partials (P patterns, bytes) and
    // exponents (bytes)
    int rU, rX[2], rS[2], rEu, rEx[2];
    int rD[2], rEd[2];
};

// One transition-matrix update: P, dP/dt, d2P/dt2 for all categories.
struct MatrixJob {
    void* P; void* D1; void* D2;      // D1/D2 may be null
    const double* cube;               // S^3, C_ijk = V_ik Vinv_kj
    const double* vv;                 // [V | Vinv], each S x S row-major
    const double* evals;              // S real parts, then S imaginary parts
    const double* rates;              // C category rates
    double t;
    int cat;                          // >= 0: only this category, rate 1 (model categories)
    int cplx;                         // complex eigen-system (conjugate pairs as 2 x 2 blocks)
    double* F;                        // 4 real states: the branch's Frechet factors
                                      //   [C][4][4] for the rate-matrix gradient, or null
    const double* tHost;              // non-null in a captured graph: t read from this
                                      //   (host) address at every replay
};

// One root of a several-root likelihood (calcRootLogLikelihoodsMulti).
struct RootSubset {
    const void* root;
    const double* cw;      // category weights
    const double* freqs;   // state frequencies
    const double* cum;     // cumulative scale factors (logs) or null
    const short* ex;       // automatic-scaling exponents or null
};

// One matrix product per category: out = A B, pad column 1.
struct ConvolveJob {
    const void* A; const void* B; void* out;
};

// The rate-matrix gradient formed inside the FP32 node pass (4 states,
// C = 1, 2, 4): Frechet factors (NodeJob::mF offsets into F), V | V^-1
// (FP64, row-major), and H slots hacc [capacity][16]: each launch writes
// slots [used, used + rows * tiles) once (rows = its grid's job rows) and
// advances used; a pass of n jobs needs at most n * tiles slots.
struct NodeQGrad {
    const double* F;
    const double* eig;
    double* hacc;
    long capacity;
    long used;
};
// Sum of n H slots ([n][16]) into out[16], through part: [ceil(n / 128)][16]
// block sums.
void launch_qgrad_slot_sums(const double* slots, long n, double* part, double* out,
                            hipStream_t st);

// One node of an edge-derivative or cross-product request.
struct EdgeJob {
    const void* pre;       // pre-order partials of the node
    const void* post;      // post-order partials (null if tip states)
    const int* states;     // tip states (null if partials)
    const void* D;         // derivative matrix [C][S][S+1]
    double scale;          // cross products: edge length
};

// Partials layout.  S <= 4: one row of Sp states per pattern,
// [C][Ppad][Sp].  S > 4: pattern-minor tiles of PTILE patterns,
// [C][Ppad/PTILE][S][PTILE], so that a wavefront reading one state of
// consecutive patterns touches whole 128-byte lines (the MFMA kernels hold
// one pattern per lane).  Element (c, p, i) sits at
// pbase(d, c, p) + i * d.sStride.
constexpr int PTILE = 32;

struct Dims {
    int S;        // states
    int P;        // patterns
    int C;        // categories
    int Ppad;     // padded pattern count (multiple of 64)
    int Sp;       // states per pattern in storage (= S)
    long catStride;  // Ppad * Sp
    int autoScale;   // automatic power-of-two rescaling of every operation
    int sStride;     // distance between states of one pattern: 1 or PTILE
};

__host__ __device__ inline long pbase(const Dims& d, int c, int p) {
    return c * d.catStride + (d.sStride == 1 ? (long)p * d.Sp
                                             : (long)(p / PTILE) * PTILE * d.Sp + p % PTILE);
}

// Work list of one level launch whose operations cover pattern ranges
// (operations by partition): workgroup b runs item b - first[j] of job j,
// first[j] <= b < first[j + 1] counting the items (tiles, or groups of tiles)
// of the jobs before j, from the tile that holds the job's first pattern.
// first == nullptr: the grid is (tiles, jobs) over the whole pattern range.
// (BeagleGPUImpl.hpp:2466-2572 lists (operation, pattern block) entries for
// one grid in the same way.)
struct TileList {
    const int* first;
    int n;
};

// One child matrix of a post-order operation repacked for the MFMA kernels
// (mfma.hip): src is the matrix's element offset in the matrix slab, dst
// the block's offset in the pack slab (C * mfma_pack_reals(S) floats).
struct PackJob {
    int src, dst;
    int kind;      // PACK_OPERAND: partials child, PACK_ROWS: tip-state child
    int src2;      // PACK_TIP_PRODUCT: the derivative matrix's element offset
};
enum PackKind : int {
    PACK_OPERAND = 0,       // A operand of M x (partials): rows i, inner states k
    PACK_ROWS = 1,          // M[., s] by state s (tip-state operand)
    PACK_OPERAND_BF16 = 2,  // PACK_OPERAND as hi/lo BF16 parts
    PACK_PRE_SELF = 3,      // A operand of M^T u with u held in accumulator registers
    PACK_PRE_SELF_BF16 = 4, // PACK_PRE_SELF for the BF16 matrix cores, as hi/lo parts
    PACK_TIP_COLUMN = 5,    // P[., s mod S] by state s = 0..S (tip-branch denominators)
    PACK_TIP_PRODUCT = 6,   // (P D)[., s] by state s = 0..S (tip-branch numerators)
    // added to a kind: the stored matrix is the transpose (the reference GPU
    // convention for pre-order self and derivative matrices, S > 4)
    PACK_TRANSPOSED_INPUT = 8,
};

// MFMA path shapes: M blocks of 32 output states, K in groups of 8 (four
// steps of the 32x32x2 FP32 instruction, k = 2 step + half-wave).
__host__ __device__ inline int mfma_mblocks(int S) { return (S + 31) / 32; }
__host__ __device__ inline int mfma_kgroups(int S) { return (S + 7) / 8; }
// split-BF16 path: K steps of 16 (v_mfma_f32_32x32x16_bf16)
__host__ __device__ inline int mfma_ksteps16(int S) { return (S + 15) / 16; }
// floats per category of one packed child matrix
__host__ __device__ inline int mfma_pack_reals(int S) {
    const int a = mfma_mblocks(S) * mfma_kgroups(S) * 256;   // FP32 operand blocks
    const int q = mfma_mblocks(S) * mfma_mblocks(S) * 1024;   // pre-order self blocks
    const int b = mfma_mblocks(S) * mfma_ksteps16(S) * 512;  // hi/lo BF16 operand blocks
    const int t = (S + 1) * mfma_mblocks(S) * 32;            // rows by state
    int m = a > t ? a : t;
    m = m > b ? m : b;
    m = m > q ? m : q;
    return (m + 63) / 64 * 64;
}
// FP64 path: M blocks of 16 output states (v_mfma_f64_16x16x4f64), the same k
// groups of 8; doubles per category of one packed child matrix
__host__ __device__ inline int mfma_mblocks16(int S) { return (S + 15) / 16; }
__host__ __device__ inline int mfma_pack_doubles(int S) {
    const int a = mfma_mblocks16(S) * mfma_kgroups(S) * 128;
    const int t = (S + 1) * mfma_mblocks16(S) * 16;
    const int q = mfma_mblocks16(S) * mfma_mblocks16(S) * 256;   // pre-order self blocks
    int m = a > t ? a : t;
    m = m > q ? m : q;
    return (m + 31) / 32 * 32;
}
// S range and category count the MFMA post-order kernel covers
__host__ __device__ inline bool mfma_post_supported(int S, int C) { return S > 4 && S <= 64 && C <= 8; }

// ---- launchers (mfma.hip) ----
template <typename R>
void launch_pack_matrices(const PackJob* jobs, int njobs, int S, int C, const void* mats,
                          float* pack, hipStream_t st);
// The PACK_TIP_PRODUCT jobs (which launch_pack_matrices skips), S <= 64.
template <typename R>
void launch_pack_tip_products(const PackJob* jobs, int njobs, int S, int C, const void* mats,
                              float* pack, hipStream_t st);
// One launch per dependency level: grid (pattern tiles, njobs).  variant 0:
// partials loaded a few k groups ahead; 1: the whole partials tile up front.
// w.first: a work list of `items` 32-pattern tiles (TileList).
void launch_post_mfma(const PostJob* djobs, int njobs, const Dims& d, const float* pack,
                      int variant, hipStream_t st, TileList w = {nullptr, 0}, int items = 0);
// One node of an MFMA edge-derivative request (S > 4, FP32): its pre- and
// post-order partials (or tip states) and one or two packed derivative
// matrices; `row` is its index in the request (output rows row * NR + k).
struct DerivJob {
    const void* pre;
    const void* post;
    const int* states;
    int pk1, pk2;
    int row;
};

// Pre-order (S > 4, FP32 partials) per dependency level, grid (Ppad/32,
// njobs): the jobs' pkSelf (PACK_PRE_SELF) and pkSib (PACK_OPERAND /
// PACK_ROWS) offsets in the pack slab.  terms = 3 or 4: split-BF16 matrix
// cores (PACK_PRE_SELF_BF16, PACK_OPERAND_BF16), 0: FP32 matrix cores.
// Jobs with pkT >= 0 (sself: the tip's states) also write their branch's
// three derivative tile sums to partial (rows oidx * 3 + k, ceil(P/32)
// tiles), weighted by cw (categories) and pw (patterns).
void launch_pre_mfma(const PreJob* djobs, int njobs, const Dims& d, const float* pack,
                     int terms, hipStream_t st, TileList w = {nullptr, 0}, int items = 0,
                     const double* cw = nullptr, const double* pw = nullptr,
                     double* partial = nullptr);
// Edge derivatives (S > 4, FP32 partials): nd = 1 or 2 derivative matrices;
// per job and 32-pattern tile, partial sums rows [row * NR + k][tile] (NR =
// 2: sum w r1, sum w r1^2; NR = 3: also sum w r2) with r = N / D over
// categories; sites (nd = 1, may be null): r1 per job and pattern.  terms
// as launch_pre_mfma (matrices PACK_OPERAND_BF16 when 3 or 4).
void launch_deriv_mfma(const DerivJob* djobs, int njobs, int nd, const Dims& d,
                       const float* pack, const double* cw, const double* pw, double* sites,
                       double* partial, int terms, hipStream_t st);

// FP64 partials on v_mfma_f64_16x16x4f64: the matrices packed with
// launch_pack_matrices64 (pack offsets in floats, even), grid (Ppad/16, njobs).
void launch_pack_matrices64(const PackJob* jobs, int njobs, int S, int C, const void* mats,
                            float* pack, hipStream_t st);
void launch_post_mfma64(const PostJob* djobs, int njobs, const Dims& d, const float* pack,
                        hipStream_t st);
// FP64 pre-order and edge derivatives (as launch_pre_mfma / launch_deriv_mfma;
// grids over 16-pattern tiles, partial-sum rows [row * NR + k][P/16 tiles]).
void launch_pre_mfma64(const PreJob* djobs, int njobs, const Dims& d, const float* pack,
                       hipStream_t st);
void launch_deriv_mfma64(const DerivJob* djobs, int njobs, int nd, const Dims& d,
                         const float* pack, const double* cw, const double* pw, double* sites,
                         double* partial, hipStream_t st);
// The same on the BF16 matrix cores with hi/lo split operands (terms: 3 or 4
// products per split pair); the matrices packed as PACK_OPERAND_BF16.
void launch_post_bf16(const PostJob* djobs, int njobs, const Dims& d, const float* pack,
                      int terms, hipStream_t st, TileList w = {nullptr, 0}, int items = 0);

// ---- launchers (kernels.hip) ----
// n doubles from device memory to (device-visible) pinned host memory.
void launch_copy_to_host(double* dstHost, const double* src, long n, hipStream_t st);
template <typename R>
void launch_matrices(const MatrixJob* jobs, int njobs, int S, int C, hipStream_t st);
template <typename R>
void launch_convolve(const ConvolveJob* jobs, int njobs, int S, int C, hipStream_t st);
// Traversals take the device copy of the jobs (read by the kernels) and the
// host copy (which decides launch shapes on the generic-S path).
template <typename R>
// dummy: the instance's scratch tile (dummy_bytes()), read and written by
// operand slots and lanes that have no real data.
void launch_post(const PostJob* djobs, const PostJob* hjobs, int njobs, int chunk, const Dims& d,
                 const void* mats, void* dummy, hipStream_t st);
inline size_t dummy_bytes() { return 2 * 8 * 64 * 4 * sizeof(double) + 64 * (8 + 4) + 128 * 2; }
// 4-state post-order by dependency level: jobs are ordered by level and
// level l occupies [levelStart[l], levelStart[l+1]); hjobs is the host copy
// of djobs (the launch shape of each level depends on its operands).
template <typename R>
// w.first (one level only): a work list of `items` 64-pattern tiles.
void launch_post_levels(const PostJob* djobs, const PostJob* hjobs, const int* levelStart,
                        int nlevels, const Dims& d, const void* mats, hipStream_t st,
                        TileList w = {nullptr, 0}, int items = 0);
// partial != null: fused derivative sums, 3 x ntiles doubles per job
// (ntiles = ceil(P/64)), reduced afterwards with launch_reduce_rows.
template <typename R>
void launch_pre(const PreJob* djobs, const PreJob* hjobs, int njobs, const Dims& d,
                int transposedSelf, void* scratch, const void* mats, void* dummy,
                const double* cw, const double* pw, double* partial, hipStream_t st);
void launch_reduce_rows(const double* in, int n, int rows, double* out, hipStream_t st);
template <typename R>
// w.first (one level, partial == null only): a work list of `items` tiles.
void launch_pre_levels(const PreJob* djobs, const int* levelStart, int nlevels, const Dims& d,
                       const void* mats, void* dummy, const double* cw, const double* pw,
                       double* partial, hipStream_t st, TileList w = {nullptr, 0},
                       int items = 0);
// Node-fused pre-order + derivative traversal (4 states): batched == true
// runs the njobs jobs (one dependency level) in one grid over (tiles, jobs),
// otherwise one launch walks the list in order.
template <typename R>
void launch_pre_nodes(const NodeJob* djobs, int njobs, bool batched, int chunk, const Dims& d,
                      const void* jobMats, void* dummy, const double* cw, const double* pw,
                      double* partial, hipStream_t st, NodeQGrad* q = nullptr);
// Rate-matrix gradient over node jobs (4 states): per-tile, per-category
// sums of H (16 rows of tiles * C); jobF holds F per (job, category, child)
// as [job][category][2][16], eig = V | V^-1 row-major.
// Chunks of the job list the 4-state Q-gradient kernels spread over their
// grid's second dimension, and the columns of their [16][cols] partial sums.
int qgrad_chunks(const Dims& d, int njobs);
int qgrad_partial_cols(const Dims& d, int njobs);
// jobMats: the gathered job matrices (C != 4); mats: the matrix slab (C = 4);
// factors: the factor slab (NodeJob::mF offsets); eig: [V | V^-1] (FP64)
template <typename R>
void launch_qgrad(const NodeJob* djobs, int njobs, const Dims& d, const void* jobMats,
                  const void* mats, const double* factors, const double* eig, void* dummy,
                  const double* cw, const double* pw, double* partial, hipStream_t st);
// out[0..15] = V^-T H V^T from the summed H (4 states); out[16..16+nf) =
// sums[16..16+nf) (the root-frequency gradient, passed through)
void launch_qgrad_finish(const double* sums, const double* eig, int nf, double* out,
                         hipStream_t st);
// Rate-matrix gradient for any S <= 64 over pre-order jobs (one branch
// each, own operands attached): coef [nops][Ppad], xs [nops][C][chunks][S*S]
// work; z receives (V^T X V^-T) o F as rows of [S*S][nops * C], summed
// with launch_reduce_rows.  eig = V | V^-1 | eigenvalues, tau [nops][C].
template <typename R>
void launch_qgrad_generic(const PreJob* djobs, int nops, const Dims& d, const double* cw,
                          const double* pw, const double* eig, const double* tau, double* coef,
                          double* xs, int chunks, double* z, hipStream_t st);
// The chunks of X summed and transformed: z = (V^T X V^-T) o F (k_qgx_eigen).
void launch_qgx_eigen(const double* xs, int chunks, int nops, int S, int C, const double* eig,
                      const double* tau, double* z, hipStream_t st);
// The same gradient for S = 5...64 with FP32 partials on the matrix cores:
// the jobs carry packed P_sib (pkSib) and P (pkSelf), each PACK_OPERAND or,
// for a tip-state operand, PACK_ROWS; X chunks as in launch_qgrad_generic
// (no coef work), then launch_qgx_eigen.  qgrad_mfma_lds: the LDS a
// workgroup needs (C waves of one branch).
size_t qgrad_mfma_lds(const Dims& d);
void launch_qgrad_mfma(const PreJob* djobs, int nops, const Dims& d, const float* pack,
                       const double* cw, const double* pw, const double* eig,
                       const double* tau, double* xs, int chunks, double* z, hipStream_t st);
// Several roots, one likelihood per pattern: site[p], block sums of
// w_p site[p] in bsum (ceil(P / WG) of them).
template <typename R>
void launch_root_multi(const RootSubset* subs, int count, const double* pw, double* site,
                       double* bsum, const Dims& d, hipStream_t st);
// Per-category root log-likelihoods out[c * P + p] (no category weights).
template <typename R>
void launch_root_per_category(const void* root, const double* freqs, const double* cum,
                              const short* ex, double* out, const Dims& d, hipStream_t st);
// Root-frequency gradient block sums: rows i of [S][ceil(P / 256)].
template <typename R>
void launch_root_freq_grad(const void* root, const double* cw, const double* freqs,
                           const double* pw, double* partial, const Dims& d, hipStream_t st);
// The same for many states, rows i of [S][root_freq_wide_blocks(d)].
int root_freq_wide_blocks(const Dims& d);
template <typename R>
void launch_root_freq_grad_wide(const void* root, const double* cw, const double* freqs,
                                const double* pw, double* partial, const Dims& d, hipStream_t st);
// Copies each node job's six matrices (P_a, P_b, D1_a, D2_a, D1_b, D2_b) into
// jobMats as [job][category][6][20] reals, so a job's matrices for one
// category are contiguous; job j's block starts at djobs[j].mbase.  Each
// matrix keeps its [4 rows][5 columns] order, or with colMajor becomes
// [5 columns][4 rows] (the FP32 node kernels, node_f32_supported), and
// then a category's block is kNodeJobF32 reals: the six matrices, 8 of
// padding and the job's two branches' Frechet factors [2][16] (FP32, from
// F at NodeJob::mF; zero without F).
constexpr int kNodeJobMats = 6 * 20;
constexpr int kNodeJobF32 = 160;
template <typename R>
void launch_gather_node_matrices(const NodeJob* djobs, int njobs, const Dims& d,
                                 const void* mats, void* jobMats, hipStream_t st,
                                 bool colMajor = false, const double* F = nullptr);
// 4-state FP32 node jobs with these category counts run on the kernels that
// read column-major job matrices
inline bool node_f32_supported(int C) { return C == 1 || C == 2 || C == 4; }
template <typename R>
void launch_root_range(const void* root, const short* rootExp, const double* catWeights,
                       const double* freqs, const double* cumulative,
                       const double* patternWeights, double* siteLogL, double* blockSums,
                       double* result, int p0, int p1, const Dims& d, hipStream_t st);
// k_root's workgroup sums without the reduction (4 states, sStride 1):
// rows [1 + 4][nb], w_p log L_p then the root-frequency gradient; returns nb.
template <typename R>
int launch_root_rows(const void* root, const short* rootExp, const double* catWeights,
                     const double* freqs, const double* cumulative, const double* patternWeights,
                     double* siteLogL, double* rows, const Dims& d, hipStream_t st);
template <typename R>
void launch_root(const void* root, const short* rootExp, const double* catWeights,
                 const double* freqs, const double* cumulative, const double* patternWeights,
                 double* siteLogL, double* blockSums, double* result, const Dims& d,
                 hipStream_t st);
template <typename R>
void launch_edge_derivatives(const EdgeJob* jobs, int njobs, const double* catWeights,
                             const double* patternWeights, double* sites, double* blockSums,
                             double* sums, double* sumSq, int transposedD, const Dims& d,
                             hipStream_t st);
template <typename R>
void launch_cross_products(const EdgeJob* jobs, int njobs, const double* catWeights,
                           const double* catRates, const double* patternWeights,
                           double* blockPartials, double* out, const Dims& d, hipStream_t st);
template <typename R>
void launch_edge_loglik(const void* parent, const void* child, const int* childStates,
                        const void* P, const void* D1, const void* D2,
                        const double* catWeights, const double* freqs,
                        const double* cumA, const double* cumB, const short* exA,
                        const short* exB, const double* patternWeights, double* siteL,
                        double* siteD1, double* siteD2, double* blockSums, double* result,
                        int p0, int p1, const Dims& d, hipStream_t st);

// Rows of S*S partial sums the cross-product launch writes into its workspace.
int cross_products_rows(const Dims& d);
// Workgroup sums launch_root / launch_root_range write for patterns [p0, p1).
int root_block_count(const Dims& d, int p0, int p1);

// Fixed device scratch of the allocation-free root log-likelihood
// (noalloc.hip): 256 KiB per GPU, the last double the result slot, the rest
// workgroup sums (root_block_count must not exceed kNoAllocScratchDoubles - 1:
// at most 1,048,544 patterns for S > 4, 8,388,352 for S <= 4).
constexpr int kNoAllocScratchDoubles = 32768;
// Device address of this device's scratch array (looked up once per device,
// with that device current); null if the lookup failed.
double* noalloc_scratch(int device);

void launch_accumulate(double* cum, const double* const* bufs, int nbufs, double sign,
                       int logScalers, int p0, int p1, hipStream_t st);
void launch_fill(double* dst, double v, long n, hipStream_t st);
void launch_fill_short(short* dst, short v, long n, hipStream_t st);
template <typename R>
void launch_transpose(const void* const* src, void* const* dst, int count, int S, int C,
                      hipStream_t st);
template <typename R>
void launch_pad_set(void* dst, long n, R v, hipStream_t st);
template <typename R>
void launch_tile_vector(void* dst, const double* v, const Dims& d, hipStream_t st,
                        short* ex = nullptr);

}  // namespace hipbeagle
