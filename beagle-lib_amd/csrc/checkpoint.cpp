// Gradient of one tree under a partials budget (beagleExtCheckpointedGradient).
//
// The client describes the tree by its post-order operations over node ids
// (tips: their tip buffers; internal nodes: ids that are not buffers) and
// hands over a pool of partials buffers.  The call keeps the post-order
// partials of checkpoint nodes only (height a multiple of the interval, and
// the root) and recomputes the others from the checkpoints below them
// during the pre-order pass, just before the node jobs that read them.
// Pre-order partials live in pool buffers from the job that writes them to
// the job that reads them; tips' are not stored.  Every step runs through
// the instance's ordinary traversal paths (updatePartials, the root
// log-likelihood, the fused pre-order + derivative pass), whose results do
// not depend on how the operations are batched, so the gradient is the one
// a full evaluation gives.
#include "beagle.h"
#include "instance.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <functional>

namespace hipbeagle {

namespace {

struct TreeOp {
    int dest, c1, m1, c2, m2;
};

// One step of the schedule, on physical buffers.
struct Step {
    enum Kind { POST, ROOT, PRE } kind;
    std::vector<int> rows;                 // POST, PRE: 7-int operations
    std::vector<int> post, d1, d2, out;    // PRE: own post-order buffer, derivative matrices,
                                           //      output node id per operation
    int root = -1, rootPre = -1;           // ROOT
};

struct Schedule {
    std::vector<Step> steps;
    int peak = 0;          // most pool buffers in use at once
    int checkpoints = 0;   // internal nodes whose post-order partials are kept
    int recomputed = 0;    // post-order operations run again in the pre-order pass
    int outSize = 0;       // largest node id + 1
    int maxBatch = 0;      // most operations in one PRE step
};

class Pool {
public:
    Pool(const int* b, int n) : free_(b, b + n) { std::reverse(free_.begin(), free_.end()); }
    bool take(int& b) {
        if (free_.empty()) return false;
        b = free_.back();
        free_.pop_back();
        peak_ = std::max(peak_, ++used_);
        return true;
    }
    void give(int b) {
        free_.push_back(b);
        --used_;
    }
    int peak() const { return peak_; }

private:
    std::vector<int> free_;
    int used_ = 0, peak_ = 0;
};

// The schedule for checkpoint interval k; false when the tree is malformed
// (rc = OUT_OF_RANGE) or the pool runs out (rc = GENERAL).
// tipPre: tips' pre-order partials are stored too (paths other than the
// fused 4-state one need a destination for every operation).
bool makeSchedule(const std::vector<TreeOp>& ops, int tipCount, const int* d1, const int* d2,
                  const int* pool, int poolCount, int k, bool tipPre, Schedule& S, int& rc) {
    rc = BEAGLE_ERROR_OUT_OF_RANGE;
    S = Schedule{};
    if (ops.empty() || k < 1) return false;
    int maxId = tipCount - 1;
    for (const TreeOp& o : ops) maxId = std::max({maxId, o.dest, o.c1, o.c2});
    const int n = maxId + 1;
    S.outSize = n;
    std::vector<int> opOf(n, -1), height(n, 0), parent(n, -1), above(n, -1);
    for (int i = 0; i < (int)ops.size(); ++i) {
        const TreeOp& o = ops[i];
        if (o.dest < tipCount || opOf[o.dest] >= 0) return false;
        for (int c : {o.c1, o.c2}) {
            if (c < 0 || (c >= tipCount && opOf[c] < 0) || parent[c] >= 0 || c == o.dest)
                return false;
        }
        opOf[o.dest] = i;
        height[o.dest] = 1 + std::max(height[o.c1], height[o.c2]);
        parent[o.c1] = parent[o.c2] = o.dest;
        above[o.c1] = o.m1;
        above[o.c2] = o.m2;
    }
    const int root = ops.back().dest;
    for (const TreeOp& o : ops)
        if (o.dest != root && parent[o.dest] < 0) return false;   // not a single tree
    auto internal = [&](int v) { return v >= tipCount; };
    std::vector<char> ckpt(n, 0);
    for (const TreeOp& o : ops) {
        ckpt[o.dest] = o.dest == root || height[o.dest] % k == 0;
        S.checkpoints += ckpt[o.dest];
    }
    rc = BEAGLE_ERROR_GENERAL;
    Pool P(pool, poolCount);
    std::vector<int> loc(n, -1);
    for (int t = 0; t < std::min(tipCount, n); ++t) loc[t] = t;
    auto row = [](std::vector<int>& r, int d, int c1, int m1, int c2, int m2) {
        r.insert(r.end(), {d, BEAGLE_OP_NONE, BEAGLE_OP_NONE, c1, m1, c2, m2});
    };

    // post-order pass: every node computed, non-checkpoint nodes released
    // once their parent has read them
    Step post{Step::POST};
    for (const TreeOp& o : ops) {
        int b;
        if (!P.take(b)) return false;
        row(post.rows, b, loc[o.c1], o.m1, loc[o.c2], o.m2);
        loc[o.dest] = b;
        for (int c : {o.c1, o.c2})
            if (internal(c) && !ckpt[c]) {
                P.give(loc[c]);
                loc[c] = -1;
            }
    }
    S.steps.push_back(std::move(post));
    Step rs{Step::ROOT};
    rs.root = loc[root];
    if (!P.take(rs.rootPre)) return false;
    std::vector<int> pre(n, -1);   // pre-order partials buffer of each internal node
    pre[root] = rs.rootPre;
    S.steps.push_back(std::move(rs));

    // pre-order pass, node jobs in depth-first order
    Step batch{Step::PRE};
    std::vector<int> release;   // buffers freed once the current batch is enqueued
    auto flush = [&]() {
        if (!batch.rows.empty()) {
            S.maxBatch = std::max(S.maxBatch, (int)batch.out.size());
            S.steps.push_back(std::move(batch));
            batch = Step{Step::PRE};
        }
        for (int b : release) P.give(b);
        release.clear();
    };
    Step again{Step::POST};
    std::function<bool(int)> materialize = [&](int v) -> bool {
        if (loc[v] >= 0) return true;
        const TreeOp& o = ops[opOf[v]];
        if (!materialize(o.c1) || !materialize(o.c2)) return false;
        int b;
        if (!P.take(b)) return false;
        row(again.rows, b, loc[o.c1], o.m1, loc[o.c2], o.m2);
        loc[v] = b;
        ++S.recomputed;
        return true;
    };
    std::vector<int> stack{root};
    while (!stack.empty()) {
        const int v = stack.back();
        stack.pop_back();
        const TreeOp& o = ops[opOf[v]];
        if (loc[o.c1] < 0 || loc[o.c2] < 0) {
            flush();
            if (!materialize(o.c1) || !materialize(o.c2)) return false;
            S.steps.push_back(std::move(again));
            again = Step{Step::POST};
        }
        for (int c : {o.c1, o.c2}) {
            if ((internal(c) || tipPre) && !P.take(pre[c])) return false;
            if (!internal(c) && tipPre) release.push_back(pre[c]);
        }
        const int sib[2][2] = {{o.c1, o.c2}, {o.c2, o.c1}};
        for (const auto& s : sib) {
            const int c = s[0], z = s[1];
            row(batch.rows, internal(c) || tipPre ? pre[c] : BEAGLE_OP_NONE, pre[v], above[c],
                loc[z], above[z]);
            batch.post.push_back(loc[c]);
            batch.d1.push_back(d1 ? d1[c] : 0);
            batch.d2.push_back(d2 ? d2[c] : 0);
            batch.out.push_back(c);
        }
        // after this job: v's pre-order partials and its children's
        // post-order partials are no longer read
        release.push_back(pre[v]);
        for (int c : {o.c1, o.c2})
            if (internal(c)) {
                release.push_back(loc[c]);
                loc[c] = -1;
            }
        for (int c : {o.c2, o.c1})
            if (internal(c)) stack.push_back(c);
    }
    flush();
    S.peak = P.peak();
    rc = BEAGLE_SUCCESS;
    return true;
}

// The schedule for `interval`, or for interval 0 the smallest interval
// whose schedule fits the pool; *used is the interval taken.
int chooseSchedule(const std::vector<TreeOp>& tree, int tipCount, const int* d1, const int* d2,
                   const int* pool, int poolCount, int interval, bool tipPre, Schedule& S,
                   int* used) {
    int rc = BEAGLE_ERROR_GENERAL;
    if (interval > 0) {
        if (!makeSchedule(tree, tipCount, d1, d2, pool, poolCount, interval, tipPre, S, rc))
            return rc;
        *used = interval;
        return BEAGLE_SUCCESS;
    }
    for (int k = 1; k <= (int)tree.size() + 1; ++k) {
        if (makeSchedule(tree, tipCount, d1, d2, pool, poolCount, k, tipPre, S, rc)) {
            *used = k;
            return BEAGLE_SUCCESS;
        }
        if (rc != BEAGLE_ERROR_GENERAL) return rc;
    }
    return BEAGLE_ERROR_GENERAL;
}

// Post-order operations (7 ints each) as tree operations; scale indices
// must be BEAGLE_OP_NONE.
int readTree(const int* ops, int count, std::vector<TreeOp>& tree) {
    if (count <= 0 || !ops) return BEAGLE_ERROR_OUT_OF_RANGE;
    tree.resize(count);
    for (int i = 0; i < count; ++i) {
        const int* q = ops + (size_t)i * BEAGLE_OP_COUNT;
        if (q[1] != BEAGLE_OP_NONE || q[2] != BEAGLE_OP_NONE) return BEAGLE_ERROR_NO_IMPLEMENTATION;
        tree[i] = TreeOp{q[0], q[3], q[4], q[5], q[6]};
    }
    return BEAGLE_SUCCESS;
}

}  // namespace

int planCheckpoints(int tipCount, const int* ops, int count, int bufferCount, int interval,
                    int tipPreorder, int* outInterval, int* outPeak, int* outCheckpoints,
                    int* outRecomputed) {
    if (tipCount < 1 || bufferCount <= 0 || interval < 0) return BEAGLE_ERROR_OUT_OF_RANGE;
    std::vector<TreeOp> tree;
    int rc = readTree(ops, count, tree);
    if (rc != BEAGLE_SUCCESS) return rc;
    std::vector<int> pool(bufferCount);
    for (int i = 0; i < bufferCount; ++i) pool[i] = i;
    Schedule S;
    int used = 0;
    rc = chooseSchedule(tree, tipCount, nullptr, nullptr, pool.data(), bufferCount, interval,
                        tipPreorder != 0, S, &used);
    if (rc != BEAGLE_SUCCESS) return rc;
    if (outInterval) *outInterval = used;
    if (outPeak) *outPeak = S.peak;
    if (outCheckpoints) *outCheckpoints = S.checkpoints;
    if (outRecomputed) *outRecomputed = S.recomputed;
    return BEAGLE_SUCCESS;
}

int Instance::checkpointedGradient(const int* ops, int count, const int* d1Idx, const int* d2Idx,
                                   int wIdx, int fIdx, const int* pool, int poolCount,
                                   int interval, int* outInterval, int* outPeak,
                                   double* outLogL, double* outD1, double* outD1Sq,
                                   double* outD2) {
    if (capturing_) return BEAGLE_ERROR_GENERAL;
    if (count <= 0 || !ops || !d1Idx || !d2Idx || !pool || poolCount <= 0 || interval < 0)
        return BEAGLE_ERROR_OUT_OF_RANGE;
    // one pattern range, no scale buffers, matrices as the post-order uses
    // them (not the transposed pre-order convention)
    if (partitionCount_ > 1 || (flags_ & BEAGLE_FLAG_SCALING_ALWAYS) || preTransposed_)
        return BEAGLE_ERROR_NO_IMPLEMENTATION;
    if (!validEigen(wIdx) || hCatWeights_[wIdx].empty() || !validEigen(fIdx))
        return BEAGLE_ERROR_OUT_OF_RANGE;
    std::vector<TreeOp> tree;
    int rc = readTree(ops, count, tree);
    if (rc != BEAGLE_SUCCESS) return rc;
    for (const TreeOp& o : tree) {
        if (!validMatrix(o.m1) || !validMatrix(o.m2)) return BEAGLE_ERROR_OUT_OF_RANGE;
        if (o.c1 < tipCount_ && !validBuffer(o.c1)) return BEAGLE_ERROR_OUT_OF_RANGE;
        if (o.c2 < tipCount_ && !validBuffer(o.c2)) return BEAGLE_ERROR_OUT_OF_RANGE;
    }
    for (int i = 0; i < poolCount; ++i)
        if (!validBuffer(pool[i]) || pool[i] < tipCount_) return BEAGLE_ERROR_OUT_OF_RANGE;
    {
        std::vector<int> sorted(pool, pool + poolCount);   // each buffer once
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return BEAGLE_ERROR_OUT_OF_RANGE;
    }
    // the schedule: the given interval, or the smallest that fits the pool
    const bool fused = dims_.S == 4 && dims_.C <= 8;
    Schedule S;
    rc = chooseSchedule(tree, tipCount_, d1Idx, d2Idx, pool, poolCount, interval, !fused, S,
                        &interval);
    if (rc != BEAGLE_SUCCESS) return rc;
    for (int i = 0; i < count; ++i)
        for (int c : {tree[i].c1, tree[i].c2})
            if (c < S.outSize && (!validMatrix(d1Idx[c]) || !validMatrix(d2Idx[c])))
                return BEAGLE_ERROR_OUT_OF_RANGE;
    if (outInterval) *outInterval = interval;
    if (outPeak) *outPeak = S.peak;
    if (!outLogL && !outD1 && !outD1Sq && !outD2) return BEAGLE_SUCCESS;   // plan only

    // run it: one pinned staging area for the log-likelihood and every
    // batch's sums (the fused path), one wait at the end
    size_t total = 0;
    for (const Step& s : S.steps) total += s.out.size();
    reserveStage(1 + 3 * total);
    std::vector<std::pair<size_t, const Step*>> staged;
    std::vector<std::array<double, 4>> direct;   // other paths: node, sums
    double logL = 0.0;
    size_t at = 1;
    for (const Step& s : S.steps) {
        const int m = (int)s.out.size();
        if (s.kind == Step::POST) {
            rc = updatePartials(s.rows.data(), (int)s.rows.size() / BEAGLE_OP_COUNT,
                                BEAGLE_OP_NONE, false);
        } else if (s.kind == Step::ROOT) {
            rc = enqueueRoot(s.root, wIdx, fIdx, BEAGLE_OP_NONE);
            if (rc == BEAGLE_SUCCESS) {
                stageOut(dResult_, 1, 0);
                if (!fused) {
                    // the other paths deliver each batch's sums through the
                    // staging area: read the log-likelihood now
                    sync();
                    logL = hStage_[0];
                }
                rc = setRootPrePartials(&s.rootPre, &fIdx, 1);
            }
        } else if (fused) {
            const double* sums = nullptr;
            rc = enqueuePreDerivatives(s.rows.data(), m, s.post.data(), s.d1.data(), s.d2.data(),
                                       wIdx, true, &sums);
            if (rc == BEAGLE_SUCCESS) {
                stageOut(sums, 3 * (size_t)m, at);
                staged.push_back({at, &s});
                at += 3 * (size_t)m;
            }
        } else {
            std::vector<double> a(m), b(m), c(m);
            rc = updatePrePartialsWithDerivatives(s.rows.data(), m, s.post.data(), s.d1.data(),
                                                  s.d2.data(), wIdx, a.data(), b.data(), c.data());
            for (int j = 0; j < m && rc == BEAGLE_SUCCESS; ++j)
                direct.push_back({(double)s.out[j], a[j], b[j], c[j]});
        }
        if (rc != BEAGLE_SUCCESS) return rc;
    }
    sync();
    lastLogL_ = fused ? hStage_[0] : logL;
    if (outLogL) *outLogL = lastLogL_;
    for (const auto& st : staged) {
        const double* h = hStage_ + st.first;
        const Step& s = *st.second;
        for (size_t j = 0; j < s.out.size(); ++j) {
            const int v = s.out[j];
            if (outD1) outD1[v] = h[3 * j];
            if (outD1Sq) outD1Sq[v] = h[3 * j + 1];
            if (outD2) outD2[v] = h[3 * j + 2];
        }
    }
    for (const auto& d : direct) {
        const int v = (int)d[0];
        if (outD1) outD1[v] = d[1];
        if (outD1Sq) outD1Sq[v] = d[2];
        if (outD2) outD2[v] = d[3];
    }
    return lastLogL_ != lastLogL_ ? BEAGLE_ERROR_FLOATING_POINT : BEAGLE_SUCCESS;
}

}  // namespace hipbeagle
