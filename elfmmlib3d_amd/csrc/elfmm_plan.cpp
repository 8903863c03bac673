// Host planner, part 2: the task lists of one apply, built from the octree
// on the host worker pool.
//
// Which boxes take part in each step follows the reference's
// lfmm3dtrilhesstargmain (src/lfmm3dtrilhess.f:517-1215) and lfmm3d_list2
// (src/lfmm3drouts.f:297-496): multipoles of sourced leaves at level >= 2,
// M2M/L2L between sourced boxes, M2L over list 2 from level 2 down (target
// boxes without targets pruned when no source outputs are requested), P2L
// over list 4, M2T over list 3, L2T at leaves of level >= 2, near field over
// self + list 1. Each translation step is grouped by operator so that one
// dense matrix serves many (source, destination) pairs.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "elfmm_internal.h"

namespace elfmm {
namespace {

void add_pairs(GemmLaunch &g, int64_t a_off, int h, double phi, const int *src, const int *dst, const double *scale,
               int np, int refl = 0)
{
    int base = (int)g.src.size();
    g.src.insert(g.src.end(), src, src + np);
    g.dst.insert(g.dst.end(), dst, dst + np);
    g.scale.insert(g.scale.end(), scale, scale + np);
    for (int p0 = 0; p0 < np; p0 += ML_P)
        for (int r0 = 0; r0 < h; r0 += ML_M)
            g.tiles.push_back({a_off, h, r0, base + p0, std::min(ML_P, np - p0), phi, refl});
}

// Launch order of translation tiles: workgroup b of a launch runs on XCD
// b % 8 and takes tiles[b]. Tiles of one operator are cut into runs, the
// runs go heaviest first to the least loaded XCD (work of a tile: its K
// extent h plus the output adds, worth about 40 coefficients), and each
// XCD takes its runs highest order first, so every XCD ends at about the
// same time and an operator's tiles stay together in one XCD's L2.
// Slots past an XCD's last tile hold empty tiles (npair = 0).
void xcd_order(std::vector<M2LTile> &tiles)
{
    constexpr int NX = 8;
    if (tiles.size() < 2 * NX) return;
    auto work = [](const M2LTile &t) { return t.h + 40.0; };
    double total = 0;
    for (auto &t : tiles) total += work(t);
    double cap = total / (NX * 16);
    struct Run {
        size_t b, e;
        double w;
        int h;
    };
    std::vector<Run> runs;
    for (size_t i = 0; i < tiles.size(); i++) {
        if (runs.empty() || tiles[i].a_off != tiles[runs.back().b].a_off || runs.back().w >= cap)
            runs.push_back({i, i, 0.0, tiles[i].h});
        runs.back().e = i + 1;
        runs.back().w += work(tiles[i]);
    }
    std::vector<int> by_w(runs.size());
    for (size_t r = 0; r < runs.size(); r++) by_w[r] = (int)r;
    std::stable_sort(by_w.begin(), by_w.end(), [&](int a, int b) { return runs[a].w > runs[b].w; });
    std::vector<int> bin[NX];
    double load[NX] = {0};
    for (int r : by_w) {
        int x = (int)(std::min_element(load, load + NX) - load);
        load[x] += runs[r].w;
        bin[x].push_back(r);
    }
    size_t nmax = 0;
    std::vector<M2LTile> seq[NX];
    for (int x = 0; x < NX; x++) {
        std::stable_sort(bin[x].begin(), bin[x].end(), [&](int a, int b) {
            return runs[a].h != runs[b].h ? runs[a].h > runs[b].h : a < b;
        });
        for (int r : bin[x]) seq[x].insert(seq[x].end(), tiles.begin() + runs[r].b, tiles.begin() + runs[r].e);
        nmax = std::max(nmax, seq[x].size());
    }
    M2LTile empty{0, 0, 0, 0, 0, 0.0};
    tiles.assign(nmax * NX, empty);
    for (int x = 0; x < NX; x++)
        for (size_t i = 0; i < seq[x].size(); i++) tiles[i * NX + x] = seq[x][i];
}

// octant of child c in parent b: bit d set when c is the upper half in d
inline int octant(const Tree &T, int c, int b)
{
    int o = 0;
    for (int d = 0; d < 3; d++)
        if (T.ijk[3 * c + d] != 2 * (T.ijk[3 * b + d] - 1) + 1) o |= 1 << d;
    return o;
}

inline int offset_index(const Tree &T, int j, int b)
{
    return (T.ijk[3 * j] - T.ijk[3 * b] + 3) * 49 + (T.ijk[3 * j + 1] - T.ijk[3 * b + 1] + 3) * 7 +
           (T.ijk[3 * j + 2] - T.ijk[3 * b + 2] + 3);
}

// M2M (up) or L2L (down) launch for one level: pairs grouped by octant,
// each octant's rotated operator (two h x h matrices, h = nc / 2).
GemmLaunch tree_pass(const Tree &T, int lev, int nc, const double *oct_phi, bool up)
{
    int h = nc / 2;
    GemmLaunch g;
    std::vector<int> s[8], d[8];
    for (int b = T.laddr[2 * lev]; b < T.laddr[2 * lev] + T.laddr[2 * lev + 1]; b++) {
        if (!T.nkids[b] || (up && T.src_count[b] == 0)) continue;
        for (int q = 0; q < T.nkids[b]; q++) {
            int c = T.kids[8 * b + q];
            if (up && T.src_count[c] == 0) continue;
            int o = octant(T, c, b);
            s[o].push_back(up ? c : b);
            d[o].push_back(up ? b : c);
        }
    }
    for (int o = 0; o < 8; o++)
        if (!s[o].empty()) {
            std::vector<double> one(s[o].size(), 1.0);
            add_pairs(g, (int64_t)o * 2 * h * h, h, oct_phi[o], s[o].data(), d[o].data(), one.data(),
                      (int)s[o].size());
        }
    xcd_order(g.tiles);
    return g;
}

constexpr int CHUNK = 64;

// Split each slot's task range [off[s], off[s+1]) into evaluation wave tasks
// of at most EW_PTS points: (slot, first task, count) triples.
template <class I>
void wave_tasks(const std::vector<I> &off, std::vector<int> &wt)
{
    wt.clear();
    for (size_t s = 0; s + 1 < off.size(); s++)
        for (I t = off[s]; t < off[s + 1]; t += EW_PTS) {
            wt.push_back((int)s);
            wt.push_back((int)t);
            wt.push_back((int)std::min<I>(EW_PTS, off[s + 1] - t));
        }
}

}  // namespace

void build_plan_near(const Tree &T, bool src_eval, bool direct_only, bool iter, Plan &P)
{
    int nb = T.nboxes;
    auto npts = [&](int b) { return (src_eval ? T.src_count[b] : 0) + T.tgt_count[b]; };
    P.nf_boxes.clear();
    P.nf_start.clear();
    P.nf_items.clear();
    // leaves: near field over self + list 1. With iter, lists 3 and 4 join
    // the near field; the reference evaluates a box's list 4 at every point
    // below it, so a leaf also takes the list 4 of each of its ancestors and
    // every point is handled once, by its leaf.
    P.nf_start.push_back(0);
    for (int b = 0; b < nb; b++) {
        if (T.nkids[b] || !npts(b)) continue;
        size_t before = P.nf_items.size();
        for (int q = T.lstart[1][b]; q < T.lstart[1][b + 1]; q++) {
            int j = T.litems[1][q];
            if (T.src_count[j]) P.nf_items.push_back(j);
        }
        if (T.src_count[b]) P.nf_items.push_back(b);
        if (iter) {
            for (int q = T.lstart[3][b]; q < T.lstart[3][b + 1]; q++) {
                int j = T.litems[3][q];
                if (T.src_count[j]) P.nf_items.push_back(j);
            }
            for (int a = b; a >= 0; a = T.parent[a])
                for (int q = T.lstart[4][a]; q < T.lstart[4][a + 1]; q++) {
                    int j = T.litems[4][q];
                    if (T.src_count[j]) P.nf_items.push_back(j);
                }
        }
        if (P.nf_items.size() == before) continue;
        P.nf_boxes.push_back(b);
        P.nf_start.push_back((int)P.nf_items.size());
    }
    {
        // one workgroup per box, heaviest (points x triangles) first, so
        // the long boxes do not start last
        size_t nq = P.nf_boxes.size();
        std::vector<double> w(nq);
        std::vector<int> ord(nq);
        for (size_t q = 0; q < nq; q++) {
            double ns = 0;
            for (int k = P.nf_start[q]; k < P.nf_start[q + 1]; k++) ns += T.src_count[P.nf_items[k]];
            w[q] = ns * npts(P.nf_boxes[q]);
            ord[q] = (int)q;
        }
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return w[a] > w[b]; });
        std::vector<int> bx(nq), st(1, 0), it;
        it.reserve(P.nf_items.size());
        for (size_t q = 0; q < nq; q++) {
            int o = ord[q];
            bx[q] = P.nf_boxes[o];
            it.insert(it.end(), P.nf_items.begin() + P.nf_start[o], P.nf_items.begin() + P.nf_start[o + 1]);
            st.push_back((int)it.size());
        }
        P.nf_boxes.swap(bx);
        P.nf_start.swap(st);
        P.nf_items.swap(it);
    }
    if (direct_only) P.nf_boxes.clear();
    double npair_near = 0;
    for (size_t q = 0; q < P.nf_boxes.size(); q++) {
        double ns = 0;
        for (int k = P.nf_start[q]; k < P.nf_start[q + 1]; k++) ns += T.src_count[P.nf_items[k]];
        npair_near += npts(P.nf_boxes[q]) * ns;
    }
    P.stats[6] = npair_near;
}

void plan_reset(Plan &P)
{
    GemmLaunch keep;
    keep.src.swap(P.m2l.src);
    keep.dst.swap(P.m2l.dst);
    keep.scale.swap(P.m2l.scale);
    P = Plan();
    P.m2l.src.swap(keep.src);
    P.m2l.dst.swap(keep.dst);
    P.m2l.scale.swap(keep.scale);
    for (double &v : P.stats) v = 0;
}

void build_plan(const Tree &T, int n, const Operators *ops, bool src_eval, bool direct_only, int nrhs,
                Plan &P, bool iter)
{
    plan_reset(P);
    build_plan_near(T, src_eval, direct_only, iter, P);
    build_plan_far(T, n, ops, src_eval, nrhs, P, iter);
}

void build_plan_far(const Tree &T, int n, const Operators *ops, bool src_eval, int nrhs, Plan &P, bool iter)
{
    int nb = T.nboxes;
    int p = ops ? ops->p : 0;
    int nc = ncoef(p);
    int nchunk = (nb + CHUNK - 1) / CHUNK;
    if (!ops) {
        P.m2l.src.clear();
        P.m2l.dst.clear();
        P.m2l.scale.clear();
    }
    P.sides.resize(nb);
    for (int b = 0; b < nb; b++) P.sides[b] = T.box_side.empty() ? T.size / (double)(1 << T.level[b]) : T.box_side[b];
    auto npts = [&](int b) { return (src_eval ? T.src_count[b] : 0) + T.tgt_count[b]; };

    if (ops) {
        for (int b = 0; b < nb; b++)
            if (!T.nkids[b] && T.level[b] >= 2 && T.src_count[b] > 0) P.p2m_boxes.push_back(b);
        // one workgroup per box, most sources first
        std::stable_sort(P.p2m_boxes.begin(), P.p2m_boxes.end(),
                         [&](int a, int c) { return T.src_count[a] > T.src_count[c]; });
        // M2M children at lev+1 -> parents at lev, for lev = nlev-1 .. 2;
        // L2L parents at lev -> children, for lev = 2 .. nlev-1
        for (int lev = T.nlev - 1; lev >= 2; lev--) {
            GemmLaunch g = tree_pass(T, lev, nc, ops->oct_phi, true);
            if (!g.tiles.empty()) P.m2m_l.push_back(std::move(g));
        }
        for (int lev = 2; lev <= T.nlev - 1; lev++) {
            GemmLaunch g = tree_pass(T, lev, nc, ops->oct_phi, false);
            if (!g.tiles.empty()) P.l2l_l.push_back(std::move(g));
        }

        // M2L over list 2, grouped by offset: per-chunk counts by offset,
        // then every chunk fills its slice of each offset's run, so the
        // pairs of one offset stay in increasing target-box order.
        std::vector<int> cnt(343 * (size_t)nchunk, 0);
        auto m2l_target = [&](int b) { return T.level[b] >= 2 && (src_eval || T.tgt_count[b] > 0); };
        parallel_chunks(nchunk, [&](int ch, int) {
            int *c = &cnt[343 * (size_t)ch];
            for (int b = ch * CHUNK; b < std::min(nb, (ch + 1) * CHUNK); b++) {
                if (!m2l_target(b)) continue;
                for (int q = T.lstart[2][b]; q < T.lstart[2][b + 1]; q++) {
                    int j = T.litems[2][q];
                    if (T.src_count[j]) c[offset_index(T, j, b)]++;
                }
            }
        });
        std::vector<int> ostart(344, 0);
        for (int o = 0; o < 343; o++) {
            int run = ostart[o];
            for (int ch = 0; ch < nchunk; ch++) {
                int v = cnt[343 * (size_t)ch + o];
                cnt[343 * (size_t)ch + o] = run;
                run += v;
            }
            ostart[o + 1] = run;
        }
        int npair = ostart[343];
        // the pair arrays keep the previous plan's memory (plan_reset): the
        // fill below writes every entry, so no allocation or clearing here
        std::vector<int> &ps = P.m2l.src, &pd = P.m2l.dst;
        std::vector<double> &psc = P.m2l.scale;
        ps.resize(npair);
        pd.resize(npair);
        psc.resize(npair);
        parallel_chunks(nchunk, [&](int ch, int) {
            int *pos = &cnt[343 * (size_t)ch];
            for (int b = ch * CHUNK; b < std::min(nb, (ch + 1) * CHUNK); b++) {
                if (!m2l_target(b)) continue;
                double isb = 1.0 / P.sides[b];
                for (int q = T.lstart[2][b]; q < T.lstart[2][b + 1]; q++) {
                    int j = T.litems[2][q];
                    if (!T.src_count[j]) continue;
                    int k = pos[offset_index(T, j, b)]++;
                    ps[k] = j;
                    pd[k] = b;
                    psc[k] = isb;
                }
            }
        });
        // work counts: dense operator flops as executed and the reference's
        // point-and-shoot count per pair and right-hand side, two rotations
        // R(p) = 4 sum (2n+1)^2 and a z-shift Z(p) = 6 sum_j sum_k (p+1-|k|)
        std::vector<double> alg_m2l(p + 1, 0.0);
        for (int pt = 0; pt <= p; pt++) {
            double R = 0, Z = 0;
            for (int a = 0; a <= pt; a++) R += 4.0 * (2 * a + 1) * (2 * a + 1);
            for (int jj = 0; jj <= pt; jj++)
                for (int kk = -jj; kk <= jj; kk++) Z += 6.0 * (pt + 1 - std::abs(kk));
            alg_m2l[pt] = 2 * R + Z;
        }
        double dense = 0, alg = 0;
        // the pairs are already in offset order; tiles per offset cover
        // ML_P pairs by ML_M coefficients of that offset's rotated operator.
        // Offsets sharing an operator lay out their tiles together (tile
        // slot order: offsets by operator), so that xcd_order keeps them in
        // one run.
        std::vector<int> oord(343);
        for (int o = 0; o < 343; o++) oord[o] = o;
        std::stable_sort(oord.begin(), oord.end(), [&](int a, int b) { return ops->m2l_off[a] < ops->m2l_off[b]; });
        std::vector<int64_t> tstart(343, 0);
        int64_t ntile = 0;
        for (int q = 0; q < 343; q++) {
            int o = oord[q];
            int m = ostart[o + 1] - ostart[o];
            int h = ncoef(ops->m2l_pt[o]) / 2;
            tstart[o] = ntile;
            ntile += m ? (int64_t)((m + ML_P - 1) / ML_P) * ((h + ML_M - 1) / ML_M) : 0;
            if (!m) continue;
            // two real h x h products per pair and right-hand side
            dense += m * 2.0 * 2.0 * h * (double)h * nrhs;
            alg += m * alg_m2l[ops->m2l_pt[o]] * nrhs;
        }
        P.m2l_tiles.resize(ntile);
        parallel_chunks(343, [&](int o, int) {
            int m = ostart[o + 1] - ostart[o];
            if (!m) return;
            int h = ncoef(ops->m2l_pt[o]) / 2;
            M2LTile *t = &P.m2l_tiles[tstart[o]];
            for (int p0 = 0; p0 < m; p0 += ML_P)
                for (int r0 = 0; r0 < h; r0 += ML_M)
                    *t++ = M2LTile{ops->m2l_off[o], h, r0, ostart[o] + p0, std::min(ML_P, m - p0), ops->m2l_phi[o],
                                   ops->m2l_refl[o]};
        });
        xcd_order(P.m2l_tiles);
        P.stats[3] = npair;
        P.stats[4] = dense;
        P.stats[5] = alg;

        // P2L: local of box j from the sources of i with j in list3(i)
        // (src/lfmm3dtrilhess.f:760-829); owner j, sources list4(j).
        // M2T: multipole of i at the points of j with j in list4(i)
        // (src/lfmm3dtrilhess.f:864-981); owner j, sources list3(j).
        P.p2l_start.push_back(0);
        P.m2t_start.push_back(0);
        for (int j = 0; j < nb && !iter; j++) {
            int c = 0;
            for (int q = T.lstart[4][j]; q < T.lstart[4][j + 1]; q++) {
                int i = T.litems[4][q];
                if (!T.src_count[i]) continue;
                P.p2l_items.push_back(i);
                c++;
            }
            if (c) {
                P.p2l_boxes.push_back(j);
                P.p2l_start.push_back((int)P.p2l_items.size());
            }
            if (!npts(j)) continue;
            c = 0;
            for (int q = T.lstart[3][j]; q < T.lstart[3][j + 1]; q++) {
                int i = T.litems[3][q];
                if (!T.src_count[i]) continue;
                P.m2t_items.push_back(i);
                c++;
            }
            if (c) {
                P.m2t_boxes.push_back(j);
                P.m2t_start.push_back((int)P.m2t_items.size());
            }
        }
        // P2L boxes in decreasing order of source count: one workgroup per
        // box, so the longest ones start first and the launch ends evenly
        int n4 = (int)P.p2l_boxes.size();
        std::vector<int64_t> work(n4, 0);
        for (int q = 0; q < n4; q++)
            for (int k = P.p2l_start[q]; k < P.p2l_start[q + 1]; k++) work[q] += T.src_count[P.p2l_items[k]];
        std::vector<int> ord(n4);
        for (int q = 0; q < n4; q++) ord[q] = q;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return work[a] > work[b]; });
        std::vector<int> bx(n4), st(1, 0), it;
        it.reserve(P.p2l_items.size());
        for (int q = 0; q < n4; q++) {
            int o = ord[q];
            bx[q] = P.p2l_boxes[o];
            it.insert(it.end(), P.p2l_items.begin() + P.p2l_start[o], P.p2l_items.begin() + P.p2l_start[o + 1]);
            st.push_back((int)it.size());
        }
        P.p2l_boxes.swap(bx);
        P.p2l_start.swap(st);
        P.p2l_items.swap(it);
    }

    // L2T: the leaves with points at level >= 2
    for (int b = 0; b < nb; b++)
        if (!T.nkids[b] && npts(b) && ops && T.level[b] >= 2) P.l2t_boxes.push_back(b);

    // evaluation tasks: L2T one per point of each leaf; M2T one per (point,
    // list-3 box), grouped by the box whose field operand is read
    auto fill_points = [&](int b, int *tp) {
        int k = 0;
        if (src_eval)
            for (int i = T.src_start[b]; i < T.src_start[b] + T.src_count[b]; i++, k++) tp[k] = i;
        for (int i = T.tgt_start[b]; i < T.tgt_start[b] + T.tgt_count[b]; i++, k++) tp[k] = n + i;
    };
    {
        int nl = (int)P.l2t_boxes.size();
        std::vector<int> off(nl + 1, 0);
        for (int q = 0; q < nl; q++) off[q + 1] = off[q] + npts(P.l2t_boxes[q]);
        P.l2t_tpt.resize(off[nl]);
        parallel_chunks((nl + CHUNK - 1) / CHUNK, [&](int ch, int) {
            for (int q = ch * CHUNK; q < std::min(nl, (ch + 1) * CHUNK); q++)
                fill_points(P.l2t_boxes[q], &P.l2t_tpt[off[q]]);
        });
        wave_tasks(off, P.l2t_wt);
    }
    if (!P.m2t_items.empty()) {
        // slots in order of first use; the owners of each slot in order
        std::vector<int> slot_of(nb, -1), nown;
        for (size_t q = 0; q < P.m2t_boxes.size(); q++)
            for (int k = P.m2t_start[q]; k < P.m2t_start[q + 1]; k++) {
                int i = P.m2t_items[k];
                if (slot_of[i] < 0) {
                    slot_of[i] = (int)P.m2t_slot_box.size();
                    P.m2t_slot_box.push_back(i);
                    nown.push_back(0);
                }
                nown[slot_of[i]]++;
            }
        int ns = (int)P.m2t_slot_box.size();
        std::vector<int> ostart(ns + 1, 0);
        for (int sl = 0; sl < ns; sl++) ostart[sl + 1] = ostart[sl] + nown[sl];
        std::vector<int> owners(ostart[ns]), fill(ostart.begin(), ostart.end() - 1);
        for (size_t q = 0; q < P.m2t_boxes.size(); q++)
            for (int k = P.m2t_start[q]; k < P.m2t_start[q + 1]; k++)
                owners[fill[slot_of[P.m2t_items[k]]]++] = P.m2t_boxes[q];
        std::vector<int64_t> toff(ns + 1, 0);
        for (int sl = 0; sl < ns; sl++) {
            int64_t t = 0;
            for (int k = ostart[sl]; k < ostart[sl + 1]; k++) t += npts(owners[k]);
            toff[sl + 1] = toff[sl] + t;
        }
        P.m2t_tpt.resize(toff[ns]);
        parallel_chunks((ns + CHUNK - 1) / CHUNK, [&](int ch, int) {
            for (int sl = ch * CHUNK; sl < std::min(ns, (ch + 1) * CHUNK); sl++) {
                int64_t at = toff[sl];
                for (int k = ostart[sl]; k < ostart[sl + 1]; k++) {
                    fill_points(owners[k], &P.m2t_tpt[at]);
                    at += npts(owners[k]);
                }
            }
        });
        wave_tasks(toff, P.m2t_wt);
        // workgroups of up to 4 consecutive wave tasks of one slot
        int nw = (int)P.m2t_wt.size() / 3;
        for (int w = 0; w < nw; w++)
            if (w == 0 || P.m2t_wt[3 * w] != P.m2t_wt[3 * (w - 1)] || w - P.m2t_bt.back() == 4) P.m2t_bt.push_back(w);
        P.m2t_bt.push_back(nw);
    }

    // remaining work counts (elfmm_gfx950_last_stats)
    P.stats[0] = nb;
    P.stats[1] = T.nlev;
    P.stats[2] = p;
    P.stats[7] = (double)std::count_if(P.m2l_tiles.begin(), P.m2l_tiles.end(), [](const M2LTile &t) { return t.npair > 0; });
    double m2m_pairs = 0, l2l_pairs = 0;
    for (auto &g : P.m2m_l) m2m_pairs += g.src.size();
    for (auto &g : P.l2l_l) l2l_pairs += g.src.size();
    P.stats[8] = m2m_pairs;
    P.stats[9] = l2l_pairs;
    P.stats[10] = (double)P.p2m_boxes.size();
    P.stats[11] = (double)P.p2l_items.size();
    P.stats[12] = (double)P.m2t_items.size();
    P.stats[13] = (double)P.l2t_boxes.size();
}

void build_mindlin_plan(const Tree &T, const Operators &ops, MindlinPlan &MP)
{
    MP = MindlinPlan();
    int nb = T.nboxes;
    MP.pw_slot.assign(nb, -1);
    for (int b = 0; b < nb; b++) {
        if (T.level[b] < 2 || !T.src_count[b]) continue;
        MP.src_boxes.push_back(b);
        MP.pw_slot[b] = MP.npw++;
        if (T.nkids[b] && T.level[b] < 64) MP.b_scale_boxes[T.level[b]].push_back(b);
    }
    std::vector<int> cnt(343, 0);
    std::vector<std::pair<int, int>> near_pairs;   // (offset, pair index)
    std::vector<int> ns, nd;
    MP.far_start.push_back(0);
    MP.near_start.push_back(0);
    for (int b = 0; b < nb; b++) {
        if (T.level[b] < 2 || !T.tgt_count[b]) continue;
        int nfar = 0, nnear = 0;
        for (int q = T.lstart[2][b]; q < T.lstart[2][b + 1]; q++) {
            int j = T.litems[2][q];
            if (!T.src_count[j]) continue;
            // z offset of target and source centres in box sides (src/lfmm3droutsmindlin.f:329-333)
            double side = T.size / (double)(1 << T.level[j]);
            double zdis = (T.centers[3 * b + 2] - T.centers[3 * j + 2]) / side;
            if (std::fabs(zdis) < 1.5) {
                MP.near_items.push_back(j);
                nnear++;
                int o = offset_index(T, j, b);
                near_pairs.push_back({o, (int)ns.size()});
                ns.push_back(j);
                nd.push_back(b);
            } else {
                MP.far_items.push_back(j);
                nfar++;
            }
        }
        if (nfar) {
            MP.far_boxes.push_back(b);
            MP.far_start.push_back((int)MP.far_items.size());
        }
        if (nnear) {
            MP.near_boxes.push_back(b);
            MP.near_start.push_back((int)MP.near_items.size());
        }
    }
    // near pairs grouped by offset, tiles of the full-order operators
    std::stable_sort(near_pairs.begin(), near_pairs.end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
    GemmLaunch &g = MP.m2l;
    for (size_t k = 0; k < near_pairs.size();) {
        size_t e = k;
        while (e < near_pairs.size() && near_pairs[e].first == near_pairs[k].first) e++;
        int o = near_pairs[k].first;
        std::vector<int> s(e - k), d(e - k);
        std::vector<double> sc(e - k);
        for (size_t i = k; i < e; i++) {
            s[i - k] = ns[near_pairs[i].second];
            d[i - k] = nd[near_pairs[i].second];
            sc[i - k] = (double)(1 << T.level[d[i - k]]) / T.size;
        }
        int h = ncoef(ops.m2l_pt[o]) / 2;
        add_pairs(g, ops.m2l_off[o], h, ops.m2l_phi[o], s.data(), d.data(), sc.data(), (int)(e - k), ops.m2l_refl[o]);
        k = e;
    }
    MP.m2l_tiles = g.tiles;
}

void build_forest(const std::vector<const Tree *> &trees, const std::vector<int> &soff, const std::vector<int> &toff,
                  Tree &F)
{
    F = Tree();
    int nt = (int)trees.size();
    int L = 0;
    for (const Tree *t : trees) L = std::max(L, t->nlev);
    // new id of box b of tree k: level-major, trees in order within a level
    std::vector<std::vector<int>> nid(nt);
    std::vector<std::pair<int, int>> order;                 // (tree, box) by new id
    F.laddr.clear();
    for (int lev = 0; lev <= L; lev++) {
        int first = (int)order.size();
        for (int k = 0; k < nt; k++) {
            const Tree &T = *trees[k];
            if (nid[k].empty()) nid[k].assign(T.nboxes, -1);
            if (lev > T.nlev) continue;
            for (int b = T.laddr[2 * lev]; b < T.laddr[2 * lev] + T.laddr[2 * lev + 1]; b++) {
                nid[k][b] = (int)order.size();
                order.push_back({k, b});
            }
        }
        F.laddr.push_back(first);
        F.laddr.push_back((int)order.size() - first);
    }
    int nb = (int)order.size();
    F.nboxes = nb;
    F.nlev = L;
    F.level.resize(nb);
    F.ijk.resize(3 * (size_t)nb);
    F.parent.resize(nb);
    F.kids.assign(8 * (size_t)nb, -1);
    F.nkids.resize(nb);
    F.src_start.resize(nb);
    F.src_count.resize(nb);
    F.tgt_start.resize(nb);
    F.tgt_count.resize(nb);
    F.centers.resize(3 * (size_t)nb);
    F.box_side.resize(nb);
    for (int t = 1; t <= 5; t++) F.lstart[t].assign(nb + 1, 0);
    for (int q = 0; q < nb; q++) {
        int k = order[q].first, b = order[q].second;
        const Tree &T = *trees[k];
        F.level[q] = T.level[b];
        for (int d = 0; d < 3; d++) {
            F.ijk[3 * q + d] = T.ijk[3 * b + d];
            F.centers[3 * q + d] = T.centers[3 * b + d];
        }
        F.parent[q] = T.parent[b] < 0 ? -1 : nid[k][T.parent[b]];
        F.nkids[q] = T.nkids[b];
        for (int c = 0; c < T.nkids[b]; c++) F.kids[8 * q + c] = nid[k][T.kids[8 * b + c]];
        F.src_start[q] = soff[k] + T.src_start[b];
        F.src_count[q] = T.src_count[b];
        F.tgt_start[q] = toff[k] + T.tgt_start[b];
        F.tgt_count[q] = T.tgt_count[b];
        F.box_side[q] = T.size / (double)(1 << T.level[b]);
        for (int t = 1; t <= 5; t++) F.lstart[t][q + 1] = F.lstart[t][q] + T.lstart[t][b + 1] - T.lstart[t][b];
    }
    for (int t = 1; t <= 5; t++) {
        F.litems[t].resize(F.lstart[t][nb]);
        parallel_chunks((nb + 255) / 256, [&](int ch, int) {
            for (int q = ch * 256; q < std::min(nb, (ch + 1) * 256); q++) {
                int k = order[q].first, b = order[q].second;
                const Tree &T = *trees[k];
                int o = F.lstart[t][q];
                for (int e = T.lstart[t][b]; e < T.lstart[t][b + 1]; e++) F.litems[t][o++] = nid[k][T.litems[t][e]];
            }
        });
    }
    int ns = 0, ng = 0;
    for (int k = 0; k < nt; k++) {
        ns = std::max(ns, soff[k] + (int)trees[k]->iz.size());
        ng = std::max(ng, toff[k] + (int)trees[k]->iztarg.size());
    }
    F.iz.resize(ns);
    F.iztarg.resize(ng);
    for (int k = 0; k < nt; k++) {
        const Tree &T = *trees[k];
        for (size_t i = 0; i < T.iz.size(); i++) F.iz[soff[k] + i] = soff[k] + T.iz[i];
        for (size_t i = 0; i < T.iztarg.size(); i++) F.iztarg[toff[k] + i] = toff[k] + T.iztarg[i];
    }
}

}  // namespace elfmm
