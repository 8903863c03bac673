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
// o = 4*(z>c) + 2*(y>c) + (x>c), done here as one counting pass over the
// box's contiguous point records. Returns the per-octant counts.
void split_octants(Pt *pts, int n, const double *c, int cnt[8], std::vector<Pt> &tmp)
{
    for (int o = 0; o < 8; o++) cnt[o] = 0;
    if (n == 0) return;
    tmp.resize(n);
    std::vector<unsigned char> key(n);
    for (int i = 0; i < n; i++) {
        const double *x = pts[i].x;
        int o = (x[2] > c[2] ? 4 : 0) | (x[1] > c[1] ? 2 : 0) | (x[0] > c[0] ? 1 : 0);
        key[i] = (unsigned char)o;
        cnt[o]++;
    }
    int pos[8];
    pos[0] = 0;
    for (int o = 1; o < 8; o++) pos[o] = pos[o - 1] + cnt[o - 1];
    for (int i = 0; i < n; i++) tmp[pos[key[i]]++] = pts[i];
    std::copy(tmp.begin(), tmp.end(), pts);
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

    // point records in tree order; permutations are read back at the end
    std::vector<Pt> ps(n), pt(nt);
    for (int i = 0; i < n; i++) ps[i] = Pt{{src[3 * i], src[3 * i + 1], src[3 * i + 2]}, i};
    for (int i = 0; i < nt; i++) pt[i] = Pt{{targ[3 * i], targ[3 * i + 1], targ[3 * i + 2]}, i};
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

    std::vector<Pt> tmp;
    for (int lev = 0; lev < 100; lev++) {
        int first = T.laddr[2 * lev], count = T.laddr[2 * lev + 1];
        int nnew = 0;
        double side = size / (double)(1 << lev), side2 = side / 2;
        for (int b = first; b < first + count; b++) {
            if (T.src_count[b] <= nbox && T.tgt_count[b] <= nbox) continue;
            // box centre of d3tcentf (src/d3tstrcr.f:1682-1705)
            double c[3];
            for (int d = 0; d < 3; d++)
                c[d] = T.center[d] - size / 2 + (T.ijk[3 * b + d] - 1) * side + side2;
            int cs[8], ct[8];
            split_octants(ps.data() + T.src_start[b], T.src_count[b], c, cs, tmp);
            split_octants(pt.data() + T.tgt_start[b], T.tgt_count[b], c, ct, tmp);
            int s0 = T.src_start[b], t0 = T.tgt_start[b];
            for (int o = 0; o < 8; o++) {
                int sa = s0, ta = t0;
                s0 += cs[o];
                t0 += ct[o];
                if (cs[o] == 0 && ct[o] == 0) continue;   // empty octants make no box
                int kid = (int)T.level.size();
                T.level.push_back(lev + 1);
                T.ijk.push_back(2 * (T.ijk[3 * b] - 1) + 1 + (o & 1));
                T.ijk.push_back(2 * (T.ijk[3 * b + 1] - 1) + 1 + ((o >> 1) & 1));
                T.ijk.push_back(2 * (T.ijk[3 * b + 2] - 1) + 1 + ((o >> 2) & 1));
                T.parent.push_back(b);
                for (int q = 0; q < 8; q++) T.kids.push_back(-1);
                T.nkids.push_back(0);
                T.src_start.push_back(sa);
                T.src_count.push_back(cs[o]);
                T.tgt_start.push_back(ta);
                T.tgt_count.push_back(ct[o]);
                T.kids[8 * b + T.nkids[b]] = kid;   // compacted child slots
                T.nkids[b]++;
                nnew++;
            }
        }
        if (nnew == 0) break;
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

    // lists 5 (colleagues) and 2 (well separated children of the parent's
    // colleagues), src/d3tstrcr.f:760-828
    std::vector<std::vector<int>> L[6];
    for (int t = 1; t <= 5; t++) L[t].assign(nb, {});
    for (int b = 0; b < nb; b++) {
        L[5][b].reserve(26);
        L[2][b].reserve(64);
    }
    for (int b = 1; b < nb; b++) {
        int dad = T.parent[b];
        int ncoll = 1 + (int)L[5][dad].size();
        for (int ci = 0; ci < ncoll; ci++) {
            int coll = ci == 0 ? dad : L[5][dad][ci - 1];
            for (int q = 0; q < T.nkids[coll]; q++) {
                int kid = T.kids[8 * coll + q];
                if (kid == b) continue;
                if (touch(kid, b))
                    L[5][b].push_back(kid);
                else
                    L[2][b].push_back(kid);
            }
        }
    }
    // lists 1 and 3 for childless boxes: depth-first descent into each
    // colleague, children taken from the last compacted slot backwards
    // (src/d3tstrcr.f:919-1076)
    struct Frame {
        int box, left;
    };
    std::vector<Frame> stack;
    stack.reserve(128);
    for (int b = 0; b < nb; b++) {
        if (T.nkids[b] > 0 || T.level[b] == 0) continue;
        for (int c0 : L[5][b]) {
            stack.clear();
            stack.push_back({c0, T.nkids[c0]});
            while (!stack.empty()) {
                Frame &f = stack.back();
                int j = f.box;
                if (!touch(b, j)) {
                    L[3][b].push_back(j);
                    stack.pop_back();
                    stack.back().left--;
                    continue;
                }
                if (T.nkids[j] == 0) {
                    L[1][b].push_back(j);
                    if (T.level[j] != T.level[b]) L[1][j].push_back(b);
                    stack.pop_back();
                    if (!stack.empty()) stack.back().left--;
                    continue;
                }
                if (f.left == 0) {
                    stack.pop_back();
                    if (!stack.empty()) stack.back().left--;
                    continue;
                }
                int kid = T.kids[8 * j + f.left - 1];
                stack.push_back({kid, T.nkids[kid]});
            }
        }
    }
    // list 4 is the dual of list 3
    for (int b = 0; b < nb; b++)
        for (int j : L[3][b]) L[4][j].push_back(b);

    for (int t = 1; t <= 5; t++) {
        T.lstart[t].assign(nb + 1, 0);
        size_t tot = 0;
        for (int b = 0; b < nb; b++) tot += L[t][b].size();
        T.litems[t].resize(tot);
        size_t pos = 0;
        for (int b = 0; b < nb; b++) {
            T.lstart[t][b] = (int)pos;
            for (int j : L[t][b]) T.litems[t][pos++] = j;
        }
        T.lstart[t][nb] = (int)pos;
    }
    return 0;
}

}  // namespace elfmm
