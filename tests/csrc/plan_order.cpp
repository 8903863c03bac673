// C wrapper around the host planner (elfmm_tree.cpp, elfmm_plan.cpp) that
// checks the launch orders the planner gives its task lists, without a GPU:
// translation tiles interleaved over the 8 XCDs, near-field and P2M boxes
// heaviest first. Each check compares against a recomputation from the
// tree, not against the planner's own bookkeeping.
#include <algorithm>
#include <map>
#include <set>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "../../elfmmlib3d_amd/csrc/elfmm_internal.h"

using namespace elfmm;

namespace {

// The tiles of one launch cover, for every operator (a_off), its pairs
// [first, first + count) once, each pair group of ML_P pairs with every
// ML_M-row block of the operator's h rows once. Empty tiles (npair 0) are
// padding. Returns false on any gap, overlap or stray tile.
bool tiles_cover(const std::vector<M2LTile> &tiles, int npairs, const std::vector<int64_t> *pair_op)
{
    std::map<int64_t, std::vector<const M2LTile *>> by_op;
    for (auto &t : tiles)
        if (t.npair > 0) by_op[t.a_off].push_back(&t);
    std::vector<int> seen(npairs, 0);
    for (auto &kv : by_op) {
        std::map<int, std::set<int>> rows;   // pair0 -> row0 set
        std::map<int, int> np;
        int h = kv.second[0]->h;
        for (auto *t : kv.second) {
            if (t->h != h || t->npair > ML_P || t->row0 % ML_M || t->row0 >= h) return false;
            if (!rows[t->pair0].insert(t->row0).second) return false;
            if (np.count(t->pair0) && np[t->pair0] != t->npair) return false;
            np[t->pair0] = t->npair;
        }
        int nrb = (h + ML_M - 1) / ML_M;
        for (auto &r : rows) {
            if ((int)r.second.size() != nrb) return false;
            for (int q = r.first; q < r.first + np[r.first]; q++) {
                if (q < 0 || q >= npairs) return false;
                if (pair_op && (*pair_op)[q] != kv.first) return false;
                seen[q]++;
            }
        }
    }
    for (int v : seen)
        if (v != 1) return false;
    return true;
}

double xcd_spread(const std::vector<M2LTile> &tiles)
{
    double w[8] = {0}, tot = 0;
    for (size_t i = 0; i < tiles.size(); i++)
        if (tiles[i].npair > 0) {
            w[i % 8] += tiles[i].h + 40.0;
            tot += tiles[i].h + 40.0;
        }
    return tot > 0 ? *std::max_element(w, w + 8) / (tot / 8) : 1.0;
}

}  // namespace

// out[0] M2L tiles cover every (pair group, row block) once, each pair
//        under the operator of its offset
// out[1] largest XCD share of M2L work / mean share
// out[2] M2M and L2L tiles of every level cover their pairs
// out[3] near-field boxes: the leaves with points and sources nearby,
//        each with its self + list-1 sources, heaviest first
// out[4] P2M boxes: sourced leaves of level >= 2, most sources first
// out[5] number of M2L tiles (padding excluded)
extern "C" int plan_order_check(const double *src, int n, const double *tgt, int nt, int iprec, double *out)
{
    try {
        int itable[343];
        list2_orders(iprec_eps(iprec), itable);
        Operators ops;
        int p = expansion_order(iprec_eps(iprec));
        build_operators(p, itable, ops);
        Tree T;
        if (build_tree(src, n, tgt, nt, iprec_nbox(iprec, n + nt), 0, T)) return -2;
        Plan P;
        build_plan(T, n, &ops, true, false, 4, P, false);   // four Laplace right-hand sides per apply

        // M2L: operator of each pair from its offset
        std::vector<int64_t> op(P.m2l.src.size());
        for (size_t q = 0; q < op.size(); q++) {
            int j = P.m2l.src[q], b = P.m2l.dst[q];
            int o = (T.ijk[3 * j] - T.ijk[3 * b] + 3) * 49 + (T.ijk[3 * j + 1] - T.ijk[3 * b + 1] + 3) * 7 +
                    (T.ijk[3 * j + 2] - T.ijk[3 * b + 2] + 3);
            op[q] = ops.m2l_off[o];
        }
        out[0] = tiles_cover(P.m2l_tiles, (int)op.size(), &op) ? 1.0 : 0.0;
        out[1] = xcd_spread(P.m2l_tiles);
        bool tp = true;
        for (auto *L : {&P.m2m_l, &P.l2l_l})
            for (auto &g : *L) tp = tp && tiles_cover(g.tiles, (int)g.src.size(), nullptr);
        out[2] = tp ? 1.0 : 0.0;

        // near field: recompute each leaf's source list and work
        std::map<int, std::multiset<int>> want;
        for (int b = 0; b < T.nboxes; b++) {
            if (T.nkids[b] || T.src_count[b] + T.tgt_count[b] == 0) continue;
            std::multiset<int> s;
            for (int q = T.lstart[1][b]; q < T.lstart[1][b + 1]; q++)
                if (T.src_count[T.litems[1][q]]) s.insert(T.litems[1][q]);
            if (T.src_count[b]) s.insert(b);
            if (!s.empty()) want[b] = s;
        }
        bool nf = P.nf_boxes.size() == want.size() && P.nf_start.size() == P.nf_boxes.size() + 1;
        double prev = 1e300;
        for (size_t q = 0; nf && q < P.nf_boxes.size(); q++) {
            int b = P.nf_boxes[q];
            std::multiset<int> got(P.nf_items.begin() + P.nf_start[q], P.nf_items.begin() + P.nf_start[q + 1]);
            double ns = 0;
            for (int j : got) ns += T.src_count[j];
            double w = ns * (T.src_count[b] + T.tgt_count[b]);
            nf = want.count(b) && want[b] == got && w <= prev;
            prev = w;
        }
        out[3] = nf ? 1.0 : 0.0;

        std::vector<int> pm;
        for (int b = 0; b < T.nboxes; b++)
            if (!T.nkids[b] && T.level[b] >= 2 && T.src_count[b] > 0) pm.push_back(b);
        bool p2m = std::is_permutation(pm.begin(), pm.end(), P.p2m_boxes.begin(), P.p2m_boxes.end());
        for (size_t q = 1; p2m && q < P.p2m_boxes.size(); q++)
            p2m = T.src_count[P.p2m_boxes[q]] <= T.src_count[P.p2m_boxes[q - 1]];
        out[4] = p2m ? 1.0 : 0.0;
        out[5] = (double)std::count_if(P.m2l_tiles.begin(), P.m2l_tiles.end(),
                                       [](const M2LTile &t) { return t.npair > 0; });
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}
