// Host planner, part 1: FMM parameters and the adaptive octree with its
// interaction lists.
//
// The tree must reproduce the reference's boxes, point order and list order
// exactly (src/d3tstrcr.f:31-2350), because the reference FMM's truncation
// choices are tied to the box structure. The rules kept alike, each cited
// where it is implemented: the root cube, the level-by-level split test,
// the floating centre formula, the z-then-y-then-x stable partition with
// ties going low, the compacted child slots, the touch test with a relative
// tolerance, the order in which list entries are generated, and the
// half-space root rule.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "elfmm_internal.h"

namespace elfmm {

double iprec_eps(int iprec)
{
    switch (iprec) {
    case -2: return .5e-0;
    case -1: return .5e-1;
    case 0: return .5e-2;
    case 1: return .5e-3;
    case 2: return .5e-6;
    case 3: return .5e-9;
    case 4: return .5e-12;
    case 5: return .5e-15;
    default: return 0.0;
    }
}

int iprec_nbox(int iprec, int npts)
{
    // src/lfmm3dtrilhess.f:226-235 (integer division as in Fortran)
    switch (iprec) {
    case -2: return 8 / 3;
    case -1: return 15 / 3;
    case 0: return 30 / 3;
    case 1: return 60 / 3;
    case 2: return 120 / 3;
    case 3: return 240 / 3;
    case 4: return 480 / 3;
    case 5: return 700 / 3;
    default: return npts;
    }
}

int iprec_nbox_part(int iprec, int npts)
{
    // src/lfmm3dparthess.f:272-280
    switch (iprec) {
    case -2: return 40;
    case -1: return 50;
    case 0: return 80;
    case 1: return 160;
    case 2: return 400;
    case 3: return 800;
    case 4: return 1200;
    case 5: return 1400;
    default: return npts;
    }
}

int iprec_nbox_iter(int iprec, int npts)
{
    // src/lfmm3dtrilhess_iter_1.f:246-255 (bscale = 2)
    switch (iprec) {
    case -2: return 16;
    case -1: return 30;
    case 0: return 60;
    case 1: return 120;
    case 2: return 240;
    case 3: return 480;
    case 4: return 960;
    case 5: return 1400;
    default: return npts;
    }
}

// Decay of (sqrt(3)/2)^n / z1^(n+1) below eps (src/l3dterms.f:42-85).
static int terms_for(double eps, double z1)
{
    double z2 = std::sqrt(3.0) / 2.0;
    double jfun = z2, hfun = 1.0 / (z1 * z1);
    for (int j = 2; j <= 1000; j++) {
        hfun = hfun / z1;
        jfun = jfun * z2;
        if (jfun * hfun < eps) return j;
    }
    return 1000;
}

int expansion_order(double eps) { return terms_for(eps, 1.5); }

void list2_orders(double eps, int *itable)
{
    // separation of the nearest points of boxes at offset (ii,jj,kk),
    // src/l3dterms.f:166-205, then the symmetric fill of :208-253
    int nt[4][4][4];
    for (int ii = 2; ii <= 3; ii++)
        for (int jj = 0; jj <= 3; jj++)
            for (int kk = 0; kk <= 3; kk++) {
                double dx = ii, dy = jj, dz = kk;
                if (dx > 0) dx -= .5;
                if (dy > 0) dy -= .5;
                if (dz > 0) dz -= .5;
                nt[ii][jj][kk] = terms_for(eps, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
    for (int i = -3; i <= 3; i++)
        for (int j = -3; j <= 3; j++)
            for (int k = -3; k <= 3; k++) {
                int v = 0, ai = std::abs(i), aj = std::abs(j), ak = std::abs(k);
                if (ai > 1)
                    v = nt[ai][aj][ak];
                else if (aj > 1)
                    v = nt[aj][ai][ak];
                if (ai <= 1 && aj <= 1 && ak > 1) v = ai >= aj ? nt[ak][ai][aj] : nt[ak][aj][ai];
                itable[(i + 3) * 49 + (j + 3) * 7 + (k + 3)] = v;
            }
}

namespace {

struct Pt {
    double x[3];
    int id;
};

// Eight octants of one box (src/d3tstrcr.f:1712-1897): the reference splits
// by z, then y, then x with stable two-way partitions, a coordinate equal
// to the centre going low. That order is a stable sort by the octant key
// o = 4*(z>c) + 2*(y>c) + (x>c), done here as a counting sort over the
// box's contiguous point records.
inline int octant_key(const double *x, const double *c)
{
    return (x[2] > c[2] ? 4 : 0) | (x[1] > c[1] ? 2 : 0) | (x[0] > c[0] ? 1 : 0);
}

void split_octants(Pt *pts, int n, const double *c, int cnt[8], std::vector<Pt> &tmp)
{
    for (int o = 0; o < 8; o++) cnt[o] = 0;
    if (n == 0) return;
    tmp.resize(n);
    std::vector<unsigned char> key(n);
    for (int i = 0; i < n; i++) {
        int o = octant_key(pts[i].x, c);
        key[i] = (unsigned char)o;
        cnt[o]++;
    }
    int pos[8];
    pos[0] = 0;
    for (int o = 1; o < 8; o++) pos[o] = pos[o - 1] + cnt[o - 1];
    for (int i = 0; i < n; i++) tmp[pos[key[i]]++] = pts[i];
    std::copy(tmp.begin(), tmp.end(), pts);
}

// The same stable split of one large box by all pool threads: per-chunk
// counts, offsets by (octant, chunk), scatter, copy back.
void split_octants_par(Pt *pts, int n, const double *c, int cnt[8], std::vector<Pt> &tmp,
                       std::vector<unsigned char> &key)
{
    const int CH = 8192;
    int nch = (n + CH - 1) / CH;
    tmp.resize(n);
    key.resize(n);
    std::vector<int> cc(8 * (size_t)nch, 0);
    parallel_chunks(nch, [&](int ch, int) {
        int a = ch * CH, e = std::min(n, a + CH);
        int *k8 = &cc[8 * (size_t)ch];
        for (int i = a; i < e; i++) {
            int o = octant_key(pts[i].x, c);
            key[i] = (unsigned char)o;
            k8[o]++;
        }
    });
    for (int o = 0; o < 8; o++) {
        cnt[o] = 0;
        for (int ch = 0; ch < nch; ch++) cnt[o] += cc[8 * (size_t)ch + o];
    }
    int base = 0;
    for (int o = 0; o < 8; o++) {
        int run = base;
        for (int ch = 0; ch < nch; ch++) {
            int v = cc[8 * (size_t)ch + o];
            cc[8 * (size_t)ch + o] = run;
            run += v;
        }
        base += cnt[o];
    }
    parallel_chunks(nch, [&](int ch, int) {
        int a = ch * CH, e = std::min(n, a + CH);
        int *pos = &cc[8 * (size_t)ch];
        for (int i = a; i < e; i++) tmp[pos[key[i]]++] = pts[i];
    });
    parallel_chunks(nch, [&](int ch, int) {
        int a = ch * CH, e = std::min(n, a + CH);
        std::copy(tmp.begin() + a, tmp.begin() + e, pts + a);
    });
}

// Two boxes touch or overlap (src/d3tstrcr.f:1082-1179). The reference
// compares floating corners with a tolerance of min(side)/1e4; corners of
// the same cube grid differ only by rounding, so the test is exactly the
// closed-interval test on integer box coordinates at the finer level.
inline bool touching(const int *ijk_a, int la, const int *ijk_b, int lb)
{
    int L = la > lb ? la : lb;
    int sa = L - la, sb = L - lb;
    for (int d = 0; d < 3; d++) {
        long lo_a = (long)(ijk_a[d] - 1) << sa, hi_a = (long)ijk_a[d] << sa;
        long lo_b = (long)(ijk_b[d] - 1) << sb, hi_b = (long)ijk_b[d] << sb;
        if (lo_a > hi_b || lo_b > hi_a) return false;
    }
    return true;
}

// Chunked list output: per-box counts plus the items of each chunk of
// consecutive boxes in box order, assembled into CSR at the end.
struct ChunkLists {
    std::vector<int> cnt;                      // per box
    std::vector<std::vector<int>> items;       // per chunk
};

constexpr int BOX_CHUNK = 64;
constexpr int MAXCOLL = 26;

}  // namespace

int build_tree(const double *src, int n, const double *targ, int nt, int nbox, int imindlin,
               Tree &T)
{
    if (n < 1) return 128;
    // root cube over sources and targets (src/d3tstrcr.f:1364-1402)
    double mn[3] = {src[0], src[1], src[2]}, mx[3] = {src[0], src[1], src[2]};
    auto grow = [&](const double *p) {
        for (int d = 0; d < 3; d++) {
            if (p[d] < mn[d]) mn[d] = p[d];
            if (p[d] > mx[d]) mx[d] = p[d];
        }
    };
    for (int i = 0; i < n; i++) grow(src + 3 * i);
    for (int i = 0; i < nt; i++) grow(targ + 3 * i);
    double size = mx[0] - mn[0];
    if (mx[1] - mn[1] > size) size = mx[1] - mn[1];
    if (mx[2] - mn[2] > size) size = mx[2] - mn[2];
    for (int d = 0; d < 3; d++) T.center[d] = (mn[d] + mx[d]) / 2;
    if (imindlin) {
        // root centred on z=0 and large enough to hold the mirror images
        T.center[2] = 0;
        double zm = std::max(std::fabs(mn[2]), std::fabs(mx[2]));
        if (size < 2 * zm) size = 2 * zm;
    }
    T.size = size;
    int nth = host_threads();

    // point records in tree order; permutations are read back at the end
    std::vector<Pt> ps(n), pt(nt);
    parallel_chunks((n + 16383) / 16384, [&](int ch, int) {
        for (int i = ch * 16384; i < std::min(n, (ch + 1) * 16384); i++)
            ps[i] = Pt{{src[3 * i], src[3 * i + 1], src[3 * i + 2]}, i};
    });
    parallel_chunks((nt + 16383) / 16384, [&](int ch, int) {
        for (int i = ch * 16384; i < std::min(nt, (ch + 1) * 16384); i++)
            pt[i] = Pt{{targ[3 * i], targ[3 * i + 1], targ[3 * i + 2]}, i};
    });
    T.level = {0};
    T.ijk = {1, 1, 1};
    T.parent = {-1};
    T.kids.assign(8, -1);
    T.nkids = {0};
    T.src_start = {0};
    T.src_count = {n};
    T.tgt_start = {0};
    T.tgt_count = {nt};
    T.laddr = {0, 1};
    T.nlev = 0;

    // level-by-level subdivision (src/d3tstrcr.f:1497-1545)
    std::vector<std::vector<Pt>> tmps(nth);
    std::vector<unsigned char> keybuf;
    for (int lev = 0; lev < 100; lev++) {
        int first = T.laddr[2 * lev], count = T.laddr[2 * lev + 1];
        double side = size / (double)(1 << lev), side2 = side / 2;
        std::vector<int> sp;
        for (int b = first; b < first + count; b++)
            if (T.src_count[b] > nbox || T.tgt_count[b] > nbox) sp.push_back(b);
        if (sp.empty()) break;
        std::vector<int> cs(8 * sp.size()), ct(8 * sp.size());
        auto centre = [&](int b, double *c) {
            // box centre of d3tcentf (src/d3tstrcr.f:1682-1705)
            for (int d = 0; d < 3; d++)
                c[d] = T.center[d] - size / 2 + (T.ijk[3 * b + d] - 1) * side + side2;
        };
        if ((int)sp.size() >= 2 * nth) {
            parallel_chunks((int)sp.size(), [&](int q, int tid) {
                int b = sp[q];
                double c[3];
                centre(b, c);
                split_octants(ps.data() + T.src_start[b], T.src_count[b], c, &cs[8 * q], tmps[tid]);
                split_octants(pt.data() + T.tgt_start[b], T.tgt_count[b], c, &ct[8 * q], tmps[tid]);
            });
        } else {
            for (size_t q = 0; q < sp.size(); q++) {
                int b = sp[q];
                double c[3];
                centre(b, c);
                auto split = [&](std::vector<Pt> &P, int s0, int cnum, int *out) {
                    if (cnum >= 32768 && nth > 1)
                        split_octants_par(P.data() + s0, cnum, c, out, tmps[0], keybuf);
                    else
                        split_octants(P.data() + s0, cnum, c, out, tmps[0]);
                };
                split(ps, T.src_start[b], T.src_count[b], &cs[8 * q]);
                split(pt, T.tgt_start[b], T.tgt_count[b], &ct[8 * q]);
            }
        }
        int nnew = 0;
        for (size_t q = 0; q < sp.size(); q++) {
            int b = sp[q];
            int s0 = T.src_start[b], t0 = T.tgt_start[b];
            for (int o = 0; o < 8; o++) {
                int sa = s0, ta = t0, ns_ = cs[8 * q + o], nt_ = ct[8 * q + o];
                s0 += ns_;
                t0 += nt_;
                if (ns_ == 0 && nt_ == 0) continue;   // empty octants make no box
                int kid = (int)T.level.size();
                T.level.push_back(lev + 1);
                T.ijk.push_back(2 * (T.ijk[3 * b] - 1) + 1 + (o & 1));
                T.ijk.push_back(2 * (T.ijk[3 * b + 1] - 1) + 1 + ((o >> 1) & 1));
                T.ijk.push_back(2 * (T.ijk[3 * b + 2] - 1) + 1 + ((o >> 2) & 1));
                T.parent.push_back(b);
                for (int k = 0; k < 8; k++) T.kids.push_back(-1);
                T.nkids.push_back(0);
                T.src_start.push_back(sa);
                T.src_count.push_back(ns_);
                T.tgt_start.push_back(ta);
                T.tgt_count.push_back(nt_);
                T.kids[8 * b + T.nkids[b]] = kid;   // compacted child slots
                T.nkids[b]++;
                nnew++;
            }
        }
        T.laddr.push_back(first + count);
        T.laddr.push_back(nnew);
        T.nlev = lev + 1;
    }
    T.iz.resize(n);
    T.iztarg.resize(nt);
    for (int i = 0; i < n; i++) T.iz[i] = ps[i].id;
    for (int i = 0; i < nt; i++) T.iztarg[i] = pt[i].id;
    int nb = (int)T.level.size();
    T.nboxes = nb;

    // exact centres of d3tcentc (src/d3tstrcr.f:1596-1676)
    T.centers.resize(3 * nb);
    for (int b = 0; b < nb; b++) {
        double side = size / (double)(1 << T.level[b]), side2 = side / 2;
        for (int d = 0; d < 3; d++)
            T.centers[3 * b + d] = T.center[d] - size / 2 + (T.ijk[3 * b + d] - 1) * side + side2;
    }
    auto touch = [&](int a, int b) { return touching(&T.ijk[3 * a], T.level[a], &T.ijk[3 * b], T.level[b]); };
    int nchunk = (nb + BOX_CHUNK - 1) / BOX_CHUNK;

    // lists 5 (colleagues) and 2 (well separated children of the parent's
    // colleagues), src/d3tstrcr.f:760-828; level by level, since a box reads
    // its parent's colleagues
    std::vector<int> l5(MAXCOLL * (size_t)nb), n5(nb, 0);
    ChunkLists L2{std::vector<int>(nb, 0), std::vector<std::vector<int>>(nchunk)};
    for (int lev = 1; lev <= T.nlev; lev++) {
        int f = T.laddr[2 * lev], e = f + T.laddr[2 * lev + 1];
        int c0 = f / BOX_CHUNK, c1 = (e - 1) / BOX_CHUNK;
        // chunks straddling a level boundary are split between two levels;
        // each level writes its own boxes' part in box order
        std::vector<std::vector<int>> part(c1 - c0 + 1);
        parallel_chunks(c1 - c0 + 1, [&](int k, int) {
            int a = std::max(f, (c0 + k) * BOX_CHUNK), z = std::min(e, (c0 + k + 1) * BOX_CHUNK);
            std::vector<int> &out = part[k];
            for (int b = a; b < z; b++) {
                int dad = T.parent[b];
                int ncoll = 1 + n5[dad];
                int before = (int)out.size();
                const int *ib = &T.ijk[3 * b];
                for (int ci = 0; ci < ncoll; ci++) {
                    int coll = ci == 0 ? dad : l5[MAXCOLL * (size_t)dad + ci - 1];
                    for (int q = 0; q < T.nkids[coll]; q++) {
                        int kid = T.kids[8 * coll + q];
                        if (kid == b) continue;
                        // same level: touching means every coordinate
                        // differs by at most one
                        const int *ik = &T.ijk[3 * kid];
                        if (std::abs(ik[0] - ib[0]) <= 1 && std::abs(ik[1] - ib[1]) <= 1 &&
                            std::abs(ik[2] - ib[2]) <= 1)
                            l5[MAXCOLL * (size_t)b + n5[b]++] = kid;
                        else
                            out.push_back(kid);
                    }
                }
                L2.cnt[b] = (int)out.size() - before;
            }
        });
        for (int k = 0; k <= c1 - c0; k++) {
            auto &dst = L2.items[c0 + k];
            dst.insert(dst.end(), part[k].begin(), part[k].end());
        }
    }

    // lists 1 and 3 for childless boxes: depth-first descent into each
    // colleague, children taken from the last compacted slot backwards
    // (src/d3tstrcr.f:919-1076). A coarse leaf found touching a finer one
    // is also appended to the finer leaf's list 1 (src/d3tstrcr.f:1021):
    // those appends happen before the finer leaf's own entries (it has the
    // larger index) and in increasing order of the coarse leaf, so they are
    // collected as (fine, coarse) pairs and placed first.
    ChunkLists L1{std::vector<int>(nb, 0), std::vector<std::vector<int>>(nchunk)};
    ChunkLists L3{std::vector<int>(nb, 0), std::vector<std::vector<int>>(nchunk)};
    std::vector<std::vector<int>> app(nchunk);    // (fine, coarse) pairs per chunk
    struct Frame {
        int box, left;
    };
    parallel_chunks(nchunk, [&](int ch, int) {
        std::vector<Frame> stack;
        stack.reserve(128);
        std::vector<int> &o1 = L1.items[ch], &o3 = L3.items[ch], &ap = app[ch];
        for (int b = ch * BOX_CHUNK; b < std::min(nb, (ch + 1) * BOX_CHUNK); b++) {
            if (T.nkids[b] > 0 || T.level[b] == 0) continue;
            int b1 = (int)o1.size(), b3 = (int)o3.size();
            for (int k = 0; k < n5[b]; k++) {
                int c0 = l5[MAXCOLL * (size_t)b + k];
                stack.clear();
                stack.push_back({c0, T.nkids[c0]});
                while (!stack.empty()) {
                    Frame &fr = stack.back();
                    int j = fr.box;
                    if (!touch(b, j)) {
                        o3.push_back(j);
                        stack.pop_back();
                        stack.back().left--;
                        continue;
                    }
                    if (T.nkids[j] == 0) {
                        o1.push_back(j);
                        if (T.level[j] != T.level[b]) {
                            ap.push_back(j);
                            ap.push_back(b);
                        }
                        stack.pop_back();
                        if (!stack.empty()) stack.back().left--;
                        continue;
                    }
                    if (fr.left == 0) {
                        stack.pop_back();
                        if (!stack.empty()) stack.back().left--;
                        continue;
                    }
                    int kid = T.kids[8 * j + fr.left - 1];
                    stack.push_back({kid, T.nkids[kid]});
                }
            }
            L1.cnt[b] = (int)o1.size() - b1;
            L3.cnt[b] = (int)o3.size() - b3;
        }
    });

    auto to_csr = [&](ChunkLists &L, int t) {
        T.lstart[t].assign(nb + 1, 0);
        for (int b = 0; b < nb; b++) T.lstart[t][b + 1] = T.lstart[t][b] + L.cnt[b];
        T.litems[t].resize(T.lstart[t][nb]);
        parallel_chunks(nchunk, [&](int ch, int) {
            int b0 = ch * BOX_CHUNK;
            if (b0 >= nb) return;
            std::copy(L.items[ch].begin(), L.items[ch].end(), T.litems[t].begin() + T.lstart[t][b0]);
        });
    };
    to_csr(L2, 2);
    to_csr(L3, 3);
    // list 5 from the flat colleague table
    T.lstart[5].assign(nb + 1, 0);
    for (int b = 0; b < nb; b++) T.lstart[5][b + 1] = T.lstart[5][b] + n5[b];
    T.litems[5].resize(T.lstart[5][nb]);
    for (int b = 0; b < nb; b++)
        std::copy(&l5[MAXCOLL * (size_t)b], &l5[MAXCOLL * (size_t)b] + n5[b], T.litems[5].begin() + T.lstart[5][b]);
    // list 1: appended coarse leaves first, then the box's own entries
    {
        std::vector<int> acnt(nb, 0);
        for (auto &ap : app)
            for (size_t k = 0; k < ap.size(); k += 2) acnt[ap[k]]++;
        T.lstart[1].assign(nb + 1, 0);
        for (int b = 0; b < nb; b++) T.lstart[1][b + 1] = T.lstart[1][b] + acnt[b] + L1.cnt[b];
        T.litems[1].resize(T.lstart[1][nb]);
        std::vector<int> fill(nb);
        for (int b = 0; b < nb; b++) fill[b] = T.lstart[1][b];
        for (auto &ap : app)
            for (size_t k = 0; k < ap.size(); k += 2) T.litems[1][fill[ap[k]]++] = ap[k + 1];
        parallel_chunks(nchunk, [&](int ch, int) {
            size_t pos = 0;
            for (int b = ch * BOX_CHUNK; b < std::min(nb, (ch + 1) * BOX_CHUNK); b++) {
                std::copy(L1.items[ch].begin() + pos, L1.items[ch].begin() + pos + L1.cnt[b],
                          T.litems[1].begin() + fill[b]);
                pos += L1.cnt[b];
            }
        });
    }
    // list 4 is the dual of list 3, in increasing box order
    {
        T.lstart[4].assign(nb + 1, 0);
        for (int j : T.litems[3]) T.lstart[4][j + 1]++;
        for (int b = 0; b < nb; b++) T.lstart[4][b + 1] += T.lstart[4][b];
        T.litems[4].resize(T.lstart[4][nb]);
        std::vector<int> fill(T.lstart[4].begin(), T.lstart[4].end() - 1);
        for (int b = 0; b < nb; b++)
            for (int q = T.lstart[3][b]; q < T.lstart[3][b + 1]; q++) T.litems[4][fill[T.litems[3][q]]++] = b;
    }
    return 0;
}

}  // namespace elfmm
