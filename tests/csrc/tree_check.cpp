// C wrapper around the library's host octree (elfmm_tree.cpp) so the tests
// can compare it with the reference's d3tstrcr without a GPU.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../elfmmlib3d_amd/csrc/elfmm_internal.h"

static elfmm::Tree g_tree;

extern "C" int tree_build(const double *src, int n, const double *targ, int nt, int nbox, int imindlin)
{
    g_tree = elfmm::Tree();
    if (elfmm::build_tree(src, n, targ, nt, nbox, imindlin, g_tree)) return -1;
    return g_tree.nboxes;
}

// sizes: [0] nboxes, [1] nlev, [2..6] total items of lists 1..5
extern "C" void tree_sizes(int *out)
{
    out[0] = g_tree.nboxes;
    out[1] = g_tree.nlev;
    for (int t = 1; t <= 5; t++) out[1 + t] = (int)g_tree.litems[t].size();
}

// boxes[nb][7]: level, i, j, k (1-based), parent (0-based, -1 root),
// nkids, src_count; centers[nb][3]; iz[n]; iztarg[nt]; lstart/litems per list
extern "C" void tree_get(int *boxes, double *centers, int *iz, int *iztarg, int *lstart, int *litems)
{
    const elfmm::Tree &T = g_tree;
    for (int b = 0; b < T.nboxes; b++) {
        int *o = boxes + 7 * b;
        o[0] = T.level[b];
        o[1] = T.ijk[3 * b];
        o[2] = T.ijk[3 * b + 1];
        o[3] = T.ijk[3 * b + 2];
        o[4] = T.parent[b];
        o[5] = T.nkids[b];
        o[6] = T.src_count[b];
    }
    std::memcpy(centers, T.centers.data(), 3 * sizeof(double) * T.nboxes);
    if (!T.iz.empty()) std::memcpy(iz, T.iz.data(), sizeof(int) * T.iz.size());
    if (!T.iztarg.empty()) std::memcpy(iztarg, T.iztarg.data(), sizeof(int) * T.iztarg.size());
    for (int t = 1; t <= 5; t++) {
        std::memcpy(lstart, T.lstart[t].data(), sizeof(int) * (T.nboxes + 1));
        lstart += T.nboxes + 1;
        if (!T.litems[t].empty()) std::memcpy(litems, T.litems[t].data(), sizeof(int) * T.litems[t].size());
        litems += T.litems[t].size();
    }
}

// boxes[nb][18]: level, i, j, k, parent, nkids, src_start, src_count,
// tgt_start, tgt_count, kids[8] (the layout of elfmm_gfx950_octree_get)
extern "C" void tree_boxes_full(int *boxes)
{
    const elfmm::Tree &T = g_tree;
    for (int b = 0; b < T.nboxes; b++) {
        int *o = boxes + 18 * b;
        o[0] = T.level[b];
        for (int d = 0; d < 3; d++) o[1 + d] = T.ijk[3 * b + d];
        o[4] = T.parent[b];
        o[5] = T.nkids[b];
        o[6] = T.src_start[b];
        o[7] = T.src_count[b];
        o[8] = T.tgt_start[b];
        o[9] = T.tgt_count[b];
        for (int q = 0; q < 8; q++) o[10 + q] = T.kids[8 * b + q];
    }
}

// Near-field store size of the iter engine in the reference's word layout
// (its lused for icomp 4): plan the iter near lists on the tree just built.
extern "C" double plan_iter_ref_words(int n, int ifsl, int ifdl, int ifpt, int ifst, int ifptt, int ifstt)
{
    const elfmm::Tree &T = g_tree;
    elfmm::Plan P;
    bool src_eval = ifpt || ifst;
    elfmm::build_plan(T, n, nullptr, src_eval, false, 4, P, true);
    double words = 0;
    int nl = ifsl + ifdl;
    for (size_t q = 0; q < P.nf_boxes.size(); q++) {
        int b = P.nf_boxes[q];
        double ntri = 0;
        for (int k = P.nf_start[q]; k < P.nf_start[q + 1]; k++) ntri += T.src_count[P.nf_items[k]];
        words += nl * ntri *
                 ((src_eval ? T.src_count[b] : 0) * 3.0 * (3 * ifpt + 9 * ifst) +
                  T.tgt_count[b] * 3.0 * (3 * ifptt + 9 * ifstt));
    }
    return words;
}

// Forest of the trees of nt point sets (sources only, packed one after
// another, counts cnt[k]): compares the forest's plan with the trees' own
// plans. out[0] = forest boxes, out[1] = sum of tree boxes, out[2..3] =
// near pairs (forest, sum), out[4..5] = M2L pairs, out[6..7] = P2M boxes,
// out[8..9] = P2L items, out[10..11] = M2T items, out[12] = 1 when every
// forest box's lists map back to its tree's lists.
extern "C" int forest_check(const double *src, const int *cnt, int ntr, int nbox, int p, double *out)
{
    std::vector<elfmm::Tree> trees(ntr);
    std::vector<const elfmm::Tree *> tp;
    std::vector<int> so, to;
    int off = 0;
    for (int k = 0; k < ntr; k++) {
        if (elfmm::build_tree(src + 3 * (size_t)off, cnt[k], nullptr, 0, nbox, 0, trees[k])) return -1;
        tp.push_back(&trees[k]);
        so.push_back(off);
        to.push_back(0);
        off += cnt[k];
    }
    elfmm::Tree F;
    elfmm::build_forest(tp, so, to, F);
    elfmm::Operators ops;
    int itab[343];
    elfmm::list2_orders(1e-3, itab);
    elfmm::build_operators(p, itab, ops);
    elfmm::Plan PF;
    elfmm::build_plan(F, off, &ops, true, false, 4, PF, false);
    for (int i = 0; i < 13; i++) out[i] = 0;
    out[0] = F.nboxes;
    out[2] = PF.stats[6];
    out[4] = PF.stats[3];
    out[6] = (double)PF.p2m_boxes.size();
    out[8] = (double)PF.p2l_items.size();
    out[10] = (double)PF.m2t_items.size();
    for (int k = 0; k < ntr; k++) {
        elfmm::Plan P;
        elfmm::build_plan(trees[k], cnt[k], &ops, true, false, 4, P, false);
        out[1] += trees[k].nboxes;
        out[3] += P.stats[6];
        out[5] += P.stats[3];
        out[7] += (double)P.p2m_boxes.size();
        out[9] += (double)P.p2l_items.size();
        out[11] += (double)P.m2t_items.size();
    }
    // every forest box: same level, ijk, counts and lists (mapped) as its tree's box
    bool ok = true;
    std::vector<int> base(ntr + 1, 0);
    for (int q = 0; q < F.nboxes && ok; q++) {
        int s0 = F.src_start[q];
        int k = (int)(std::upper_bound(so.begin(), so.end(), s0) - so.begin()) - 1;
        const elfmm::Tree &T = trees[k];
        // the tree's box with this level and ijk
        int b = -1;
        for (int c = T.laddr[2 * F.level[q]]; c < T.laddr[2 * F.level[q]] + T.laddr[2 * F.level[q] + 1]; c++)
            if (T.ijk[3 * c] == F.ijk[3 * q] && T.ijk[3 * c + 1] == F.ijk[3 * q + 1] &&
                T.ijk[3 * c + 2] == F.ijk[3 * q + 2] && T.src_start[c] + so[k] == s0)
                b = c;
        if (b < 0 || T.src_count[b] != F.src_count[q] || T.nkids[b] != F.nkids[q]) ok = false;
        for (int t = 1; t <= 5 && ok; t++) {
            int nf = F.lstart[t][q + 1] - F.lstart[t][q], nt_ = T.lstart[t][b + 1] - T.lstart[t][b];
            if (nf != nt_) ok = false;
            for (int e = 0; e < nf && ok; e++) {
                int fb = F.litems[t][F.lstart[t][q] + e], tb = T.litems[t][T.lstart[t][b] + e];
                if (F.src_start[fb] != T.src_start[tb] + so[k] || F.level[fb] != T.level[tb]) ok = false;
            }
        }
    }
    out[12] = ok ? 1 : 0;
    return 0;
}
