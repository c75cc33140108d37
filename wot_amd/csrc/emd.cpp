// Exact earth mover's distance between two weighted point sets: the
// transportation problem min sum_ij F_ij C_ij, F 1 = a, F^T 1 = b, F >= 0,
// which the reference solves with POT's emd2 (wot/ot/util.py:202-237,
// numItermax = 1e7).  Primal network simplex on the complete bipartite graph
// (sources i, sinks j, arc i -> j with cost C_ij, uncapacitated) plus one
// artificial root joined to every node by a big-cost arc (initial feasible
// spanning tree).
//
// Tree representation: parent / pred arc / direction of the pred arc per
// node, children as doubly linked sibling lists, depth for the join-node
// search, node potentials pi with reduced cost C_ij + pi_i - pi_j, and the
// flow of every tree arc stored on its lower node (non-tree arcs carry 0).
// Pricing: block search over the real arcs (block ~ sqrt(n m)), entering arc
// = most negative reduced cost in the first block that has one (rows walked
// without divisions, four independent argmin lanes).  Leaving arc:
// the last blocking arc of the cycle in its orientation (strongly feasible
// trees, no cycling).  A pivot re-hangs the subtree cut off by the leaving arc
// below the entering arc, shifting its potentials by the entering arc's
// reduced cost: O(cycle + subtree) per pivot.  Host code: exact pivoting is
// sequential; the device computes the cost matrix.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "../../include/wotk.h"

namespace wotk {
int fail(int code, const char *fmt, ...);

namespace {

constexpr int kUp = 1;     // pred arc goes node -> parent
constexpr int kDown = -1;  // pred arc goes parent -> node

struct Simplex {
    int64_t n, m, N, root;
    const double *C;
    int64_t ldc;
    double art_cost, eps;
    // tree
    std::vector<int64_t> parent, pred, first_child, next_sib, prev_sib, depth;
    std::vector<int> dir;
    std::vector<double> pi, flow;
    std::vector<int64_t> stack;
    int64_t next_arc = 0, block;

    // arc e < n m: i = e / m -> sink n + e % m; e >= n m: artificial arc of node e - n m
    int64_t src(int64_t e) const { return e < n * m ? e / m : (is_source(e - n * m) ? e - n * m : root); }
    int64_t tgt(int64_t e) const { return e < n * m ? n + e % m : (is_source(e - n * m) ? root : e - n * m); }
    bool is_source(int64_t v) const { return v < n; }
    double cost(int64_t e) const { return e < n * m ? C[(e / m) * ldc + (e % m)] : art_cost; }
    double reduced(int64_t e) const { return cost(e) + pi[src(e)] - pi[tgt(e)]; }

    void link(int64_t v, int64_t p) {
        parent[v] = p;
        prev_sib[v] = -1;
        next_sib[v] = first_child[p];
        if (first_child[p] >= 0) prev_sib[first_child[p]] = v;
        first_child[p] = v;
    }
    void cut(int64_t v) {
        const int64_t p = parent[v];
        if (prev_sib[v] >= 0) next_sib[prev_sib[v]] = next_sib[v];
        else first_child[p] = next_sib[v];
        if (next_sib[v] >= 0) prev_sib[next_sib[v]] = prev_sib[v];
        parent[v] = -1;
        prev_sib[v] = next_sib[v] = -1;
    }

    void init(const double *a, const double *b) {
        N = n + m + 1;
        root = n + m;
        parent.assign(N, -1);
        pred.assign(N, -1);
        first_child.assign(N, -1);
        next_sib.assign(N, -1);
        prev_sib.assign(N, -1);
        depth.assign(N, 1);
        dir.assign(N, kUp);
        pi.assign(N, 0.0);
        flow.assign(N, 0.0);
        double cmax = 0.0;
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j < m; ++j) cmax = std::max(cmax, std::fabs(C[i * ldc + j]));
        art_cost = (cmax + 1.0) * (double)(n + m);
        eps = 64.0 * std::numeric_limits<double>::epsilon() * std::max(cmax, 1.0);
        depth[root] = 0;
        for (int64_t v = n + m - 1; v >= 0; --v) {
            pred[v] = n * m + v;
            if (is_source(v)) {  // artificial arc v -> root carries the supply
                dir[v] = kUp;
                flow[v] = a[v];
                pi[v] = -art_cost;
            } else {             // artificial arc root -> v carries the demand
                dir[v] = kDown;
                flow[v] = b[v - n];
                pi[v] = art_cost;
            }
            link(v, root);
        }
        block = std::max<int64_t>(16, (int64_t)std::sqrt((double)(n * m)));
    }

    // block search: the most negative reduced cost in the first block that has
    // one; arcs are visited row by row from next_arc on (cyclically), a block
    // being `block` consecutive arcs
    int64_t entering() {
        const int64_t E = n * m;
        int64_t i = next_arc / m, j = next_arc % m;
        int64_t seen = 0, in_block = 0, best = -1;
        double best_rc = -eps;
        const double *pj = pi.data() + n;
        while (seen < E) {
            const double *Crow = C + i * ldc;
            const double pii = pi[i];
            int64_t stop = std::min<int64_t>(m, j + (block - in_block));
            stop = std::min<int64_t>(stop, j + (E - seen));
            // four independent argmin lanes (first index wins ties, as a sequential scan)
            double lrc[4] = {best_rc, best_rc, best_rc, best_rc};
            int64_t lj[4] = {-1, -1, -1, -1};
            int64_t jj = j;
            for (; jj + 4 <= stop; jj += 4)
                for (int l = 0; l < 4; ++l) {
                    const double rc = Crow[jj + l] + pii - pj[jj + l];
                    if (rc < lrc[l]) {
                        lrc[l] = rc;
                        lj[l] = jj + l;
                    }
                }
            for (; jj < stop; ++jj) {
                const double rc = Crow[jj] + pii - pj[jj];
                if (rc < lrc[0]) {
                    lrc[0] = rc;
                    lj[0] = jj;
                }
            }
            int64_t bj = -1;
            double brc = best_rc;
            for (int l = 0; l < 4; ++l)
                if (lj[l] >= 0 && (lrc[l] < brc || (lrc[l] == brc && bj >= 0 && lj[l] < bj))) {
                    brc = lrc[l];
                    bj = lj[l];
                }
            if (bj >= 0) {
                best_rc = brc;
                best = i * m + bj;
            }
            const int64_t cnt = stop - j;
            seen += cnt;
            in_block += cnt;
            j = stop;
            if (j == m) {
                j = 0;
                if (++i == n) i = 0;
            }
            if (in_block >= block) {
                if (best >= 0) {
                    next_arc = i * m + j;
                    return best;
                }
                in_block = 0;
            }
        }
        if (best >= 0) {
            next_arc = i * m + j;
            return best;
        }
        return -1;
    }

    int64_t join(int64_t u, int64_t v) const {
        while (u != v) {
            if (depth[u] >= depth[v]) u = parent[u];
            else v = parent[v];
        }
        return u;
    }

    bool pivot(int64_t e) {
        const int64_t s = src(e), t = tgt(e);
        const int64_t w = join(s, t);
        // leaving arc: along s -> join the flow of UP arcs decreases, along t -> join that of DOWN arcs
        double delta = std::numeric_limits<double>::infinity();
        int64_t u_out = -1;
        int side = 0;
        for (int64_t u = s; u != w; u = parent[u])
            if (dir[u] == kUp) {
                const double d = std::max(flow[u], 0.0);
                if (d < delta) {
                    delta = d;
                    u_out = u;
                    side = 1;
                }
            }
        for (int64_t u = t; u != w; u = parent[u])
            if (dir[u] == kDown) {
                const double d = std::max(flow[u], 0.0);
                if (d <= delta) {
                    delta = d;
                    u_out = u;
                    side = 2;
                }
            }
        if (side == 0) return false;  // unbounded: cannot happen with the artificial tree
        const double rc = reduced(e);
        // push delta around the cycle
        if (delta > 0.0) {
            for (int64_t u = s; u != w; u = parent[u]) flow[u] -= dir[u] * delta;
            for (int64_t u = t; u != w; u = parent[u]) flow[u] += dir[u] * delta;
        }
        // re-hang: the path u_in .. u_out reverses, u_in hangs below v_in by the entering arc
        const int64_t u_in = side == 1 ? s : t, v_in = side == 1 ? t : s;
        int64_t carry_pred = e, carry_parent = v_in;
        double carry_flow = delta;
        int64_t u = u_in;
        while (true) {
            const int64_t old_parent = parent[u], old_pred = pred[u];
            const double old_flow = flow[u];
            cut(u);
            link(u, carry_parent);
            pred[u] = carry_pred;
            dir[u] = (src(carry_pred) == u) ? kUp : kDown;
            flow[u] = carry_flow;
            if (u == u_out) break;
            carry_parent = u;
            carry_pred = old_pred;
            carry_flow = old_flow;
            u = old_parent;
        }
        // potentials and depths of the re-hung subtree (rooted at u_in)
        const double shift = side == 1 ? -rc : rc;
        stack.clear();
        stack.push_back(u_in);
        while (!stack.empty()) {
            const int64_t v = stack.back();
            stack.pop_back();
            pi[v] += shift;
            depth[v] = depth[parent[v]] + 1;
            for (int64_t c = first_child[v]; c >= 0; c = next_sib[c]) stack.push_back(c);
        }
        return true;
    }
};

}  // namespace
}  // namespace wotk

extern "C" int wotk_emd(const double *a, int64_t n, const double *b, int64_t m, const double *C, int64_t ldc,
                        int64_t max_iter, double *cost_out, double *flow_out, int64_t *iterations_out,
                        int32_t *status_out) {
    if (!a || !b || !C || n <= 0 || m <= 0 || ldc < m || !cost_out)
        return wotk::fail(WOTK_EINVAL, "emd: bad arguments (n %lld, m %lld, ld %lld)", (long long)n, (long long)m,
                          (long long)ldc);
    double sa = 0.0, sb = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (!(a[i] >= 0.0)) return wotk::fail(WOTK_EINVAL, "emd: negative or NaN source weight at %lld", (long long)i);
        sa += a[i];
    }
    for (int64_t j = 0; j < m; ++j) {
        if (!(b[j] >= 0.0)) return wotk::fail(WOTK_EINVAL, "emd: negative or NaN target weight at %lld", (long long)j);
        sb += b[j];
    }
    if (std::fabs(sa - sb) > 1e-9 * std::max(sa, sb))
        return wotk::fail(WOTK_EINVAL, "emd: weights sum to %.17g and %.17g", sa, sb);
    wotk::Simplex S;
    S.n = n;
    S.m = m;
    S.C = C;
    S.ldc = ldc;
    S.init(a, b);
    int64_t it = 0;
    int32_t status = 0;
    while (true) {
        if (max_iter > 0 && it >= max_iter) {
            status = 1;
            break;
        }
        const int64_t e = S.entering();
        if (e < 0) break;
        if (!S.pivot(e)) return wotk::fail(WOTK_EINVAL, "emd: unbounded pivot (internal error)");
        ++it;
    }
    double total = 0.0;
    if (flow_out)
        for (int64_t i = 0; i < n * m; ++i) flow_out[i] = 0.0;
    for (int64_t v = 0; v < n + m; ++v) {
        const int64_t e = S.pred[v];
        if (e < n * m) {
            const double f = S.flow[v];
            total += f * C[(e / m) * ldc + (e % m)];
            if (flow_out) flow_out[(e / m) * m + (e % m)] = f;
        }
    }
    *cost_out = total;
    if (iterations_out) *iterations_out = it;
    if (status_out) *status_out = status;
    return WOTK_OK;
}
