// Device octree: the adaptive octree and interaction lists of the
// reference's d3tstrcr (src/d3tstrcr.f:31-2350), built on the GPU with the
// same boxes, point order and list order as the host planner's build_tree
// (elfmm_tree.cpp), which follows the reference rule by rule.
//
// Every point gets the octant path of its cell down to OCT_D levels: at
// each level the octant of the box that holds it, by the reference's
// floating centre formula (src/d3tstrcr.f:1682-1705) and its "coordinate >
// centre" test with ties low (src/d3tstrcr.f:1786-1829, :1882), three bits
// per level with z most significant (child o = 4z + 2y + x, :1257-1258).
// The cell of a point at a level does not depend on which boxes split, so
// one stable radix sort of all points (sources and targets together) by
// path puts every box's points together and the boxes of a level in the
// reference's order (parents in order, children in octant order). Levels
// are then opened one at a time (src/d3tstrcr.f:1497-1545): a box of the
// level splits when it holds more than nbox sources or targets, runs of
// equal path prefixes among the points of split boxes are the next
// level's boxes (empty octants make no box), and children fill their
// parent's slots in octant order. Within a leaf the reference keeps the
// input order (its partitions are stable), so the final permutations are
// stable sorts of the points, in input order, by leaf.
//
// Lists (src/d3tstrcr.f:702-1179): colleagues and list 2 level by level
// from the parent's colleagues' children, in slot order; lists 1 and 3 by
// the reference's depth-first descent into each colleague (children from
// the last slot backwards), a coarse leaf found touching a finer one also
// going to the front of the finer leaf's list 1 in increasing order; list
// 4 the dual of list 3 in increasing box order. Those orders come from
// sorted (box, box) pairs.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <vector>

#include "elfmm_internal.h"
#include "elfmm_octree.h"

namespace elfmm {
namespace {

constexpr int OCT_D = 21;           // levels of the path key (3 bits each, 63 bits)
constexpr int OCT_COLL = 26;        // colleagues per box
constexpr int OCT_L2 = 216;         // list-2 entries per box: children of 27 boxes
constexpr int OCT_T = 256;

#define OCTOK(expr)                              \
    do {                                         \
        hipError_t e_ = (expr);                  \
        if (e_ != hipSuccess) return e_;         \
    } while (0)

inline unsigned blocks(size_t n) { return (unsigned)((n + OCT_T - 1) / OCT_T); }

// ---- root cube (src/d3tstrcr.f:1364-1402; half space :1380-1383)

__global__ void k_oct_bbox(const double *__restrict__ src, int n, const double *__restrict__ tgt, int nt,
                           double *__restrict__ part)
{
    __shared__ double s[6][OCT_T];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (long i = (long)blockIdx.x * OCT_T + threadIdx.x; i < (long)n + nt; i += (long)gridDim.x * OCT_T) {
        const double *p = i < n ? src + 3 * i : tgt + 3 * (i - n);
        for (int d = 0; d < 3; d++) {
            v[d] = fmin(v[d], p[d]);
            v[3 + d] = fmax(v[3 + d], p[d]);
        }
    }
    for (int k = 0; k < 6; k++) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = OCT_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; k++)
                s[k][threadIdx.x] = k < 3 ? fmin(s[k][threadIdx.x], s[k][threadIdx.x + w])
                                          : fmax(s[k][threadIdx.x], s[k][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) part[6 * blockIdx.x + k] = s[k][0];
}

// root[0..2] centre, root[3] size, as build_tree computes them
__global__ void k_oct_root(const double *__restrict__ part, int nblk, int imindlin, double *__restrict__ root)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < nblk; b++)
        for (int d = 0; d < 3; d++) {
            mn[d] = fmin(mn[d], part[6 * b + d]);
            mx[d] = fmax(mx[d], part[6 * b + 3 + d]);
        }
    double size = __dsub_rn(mx[0], mn[0]);
    if (__dsub_rn(mx[1], mn[1]) > size) size = __dsub_rn(mx[1], mn[1]);
    if (__dsub_rn(mx[2], mn[2]) > size) size = __dsub_rn(mx[2], mn[2]);
    for (int d = 0; d < 3; d++) root[d] = __dadd_rn(mn[d], mx[d]) * 0.5;
    if (imindlin) {
        root[2] = 0.0;
        double zm = fmax(fabs(mn[2]), fabs(mx[2]));
        if (size < 2.0 * zm) size = 2.0 * zm;
    }
    root[3] = size;
}

// ---- octant paths

__global__ void k_oct_keys(const double *__restrict__ src, int n, const double *__restrict__ tgt, int nt,
                           const double *__restrict__ root, uint64_t *__restrict__ key, int *__restrict__ idx)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    if (i >= n + nt) return;
    const double *p = i < n ? src + 3 * (size_t)i : tgt + 3 * (size_t)(i - n);
    const double size = root[3];
    double lo[3];
    for (int d = 0; d < 3; d++) lo[d] = __dsub_rn(root[d], size * 0.5);
    int ijk[3] = {1, 1, 1};
    uint64_t k = 0;
    for (int l = 0; l < OCT_D; l++) {
        // centre of the level-l box holding the point: (centre - size/2) +
        // (i - 1) * side + side/2, each operation rounded on its own
        double side = ldexp(size, -l), side2 = ldexp(size, -l - 1);
        int o = 0;
        for (int d = 0; d < 3; d++) {
            double c = __dadd_rn(__dadd_rn(lo[d], __dmul_rn((double)(ijk[d] - 1), side)), side2);
            int bit = p[d] > c ? 1 : 0;
            o |= bit << d;
            ijk[d] = 2 * (ijk[d] - 1) + 1 + bit;
        }
        k = (k << 3) | (uint64_t)o;
    }
    key[i] = k;
    idx[i] = i;
}

__global__ void k_oct_srctag(const int *__restrict__ P, int N, int n, int *__restrict__ tag)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    if (i < N) tag[i] = P[i] < n ? 1 : 0;
}

// ---- boxes

struct BoxArrays {
    int *level, *ijk, *parent, *kids, *nkids, *ss, *sc, *ts, *tc, *split, *firstkid;
};

__device__ void oct_new_box(const BoxArrays &B, int b, int lev, int i0, int i1, int i2, int par, int ss, int ts)
{
    B.level[b] = lev;
    B.ijk[3 * b] = i0;
    B.ijk[3 * b + 1] = i1;
    B.ijk[3 * b + 2] = i2;
    B.parent[b] = par;
    for (int q = 0; q < 8; q++) B.kids[8 * b + q] = -1;
    B.nkids[b] = 0;
    B.ss[b] = ss;
    B.ts[b] = ts;
    B.sc[b] = 0;
    B.tc[b] = 0;
    B.split[b] = 0;
    B.firstkid[b] = INT_MAX;
}

__global__ void k_oct_root_box(BoxArrays B, int n, int nt, int nbox, int N, int *__restrict__ active,
                               int *__restrict__ cur)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    int sp = (n > nbox || nt > nbox) ? 1 : 0;
    if (i == 0) {
        oct_new_box(B, 0, 0, 1, 1, 1, -1, 0, 0);
        B.sc[0] = n;
        B.tc[0] = nt;
        B.split[0] = sp;
    }
    if (i < N) {
        active[i] = sp;
        cur[i] = 0;
    }
}

// first point of each run of equal level-l prefixes among the points of
// split boxes
__global__ void k_oct_heads(const uint64_t *__restrict__ K, const int *__restrict__ active, int N, int sh,
                            int *__restrict__ head)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    if (i >= N) return;
    int h = 0;
    if (active[i]) h = (i == 0 || !active[i - 1] || (K[i] >> sh) != (K[i - 1] >> sh)) ? 1 : 0;
    head[i] = h;
}

__global__ void k_oct_boxes(BoxArrays B, const uint64_t *__restrict__ K, const int *__restrict__ active,
                            const int *__restrict__ head, const int *__restrict__ incl,
                            const int *__restrict__ sexcl, int N, int lev, int sh, int base,
                            int *__restrict__ cur, int *__restrict__ dcount)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    if (i >= N) return;
    if (i == N - 1) dcount[0] = incl[N - 1];
    if (!active[i]) return;
    int b = base + incl[i] - 1;
    if (head[i]) {
        uint64_t pr = K[i] >> sh;
        int c[3] = {0, 0, 0};
        for (int t = 0; t < lev; t++) {
            int o = (int)((pr >> (3 * (lev - 1 - t))) & 7);
            for (int d = 0; d < 3; d++) c[d] = 2 * c[d] + ((o >> d) & 1);
        }
        oct_new_box(B, b, lev, c[0] + 1, c[1] + 1, c[2] + 1, cur[i], sexcl[i], i - sexcl[i]);
    }
    cur[i] = b;
}

// counts of a new box from its last point: the points of a box are one
// run of the sorted order, starting at ss + ts (sources and targets before
// it), so the counts follow from the running source count
__global__ void k_oct_counts(BoxArrays B, const uint64_t *__restrict__ K, const int *__restrict__ active,
                             const int *__restrict__ cur, const int *__restrict__ tag,
                             const int *__restrict__ sexcl, int N, int sh)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    if (i >= N || !active[i]) return;
    bool tail = i == N - 1 || !active[i + 1] || (K[i] >> sh) != (K[i + 1] >> sh);
    if (!tail) return;
    int b = cur[i];
    int s_upto = sexcl[i] + tag[i];                  // sources up to and including point i
    B.sc[b] = s_upto - B.ss[b];
    B.tc[b] = (i + 1 - s_upto) - B.ts[b];
}

__global__ void k_oct_split(BoxArrays B, const int *__restrict__ head, const int *__restrict__ cur, int N,
                            int nbox, int *__restrict__ dcount)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    if (i >= N || !head[i]) return;
    int b = cur[i];
    int sp = (B.sc[b] > nbox || B.tc[b] > nbox) ? 1 : 0;
    B.split[b] = sp;
    if (sp) atomicAdd(&dcount[1], 1);
    atomicMin(&B.firstkid[B.parent[b]], b);
}

__global__ void k_oct_kids(BoxArrays B, const int *__restrict__ head, const int *__restrict__ cur,
                           int *__restrict__ active, int N)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    if (i >= N || !active[i]) return;
    int b = cur[i];
    if (head[i]) {
        int p = B.parent[b];
        B.kids[8 * p + (b - B.firstkid[p])] = b;   // children of a box are consecutive, in octant order
        atomicAdd(&B.nkids[p], 1);
    }
    active[i] = B.split[b];
}

// position of every point's leaf in tree order (the index of the leaf's
// first point among all sorted points), in input order (sources, then
// targets). Box numbers are level-major, so they do not order leaves of
// different levels.
__global__ void k_oct_leaf(BoxArrays B, const int *__restrict__ P, const int *__restrict__ cur, int N, int n,
                           unsigned *__restrict__ sleaf, unsigned *__restrict__ tleaf)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    if (i >= N) return;
    int o = P[i], b = cur[i];
    unsigned pos = (unsigned)(B.ss[b] + B.ts[b]);
    if (o < n)
        sleaf[o] = pos;
    else
        tleaf[o - n] = pos;
}

__global__ void k_oct_iota(int *__restrict__ a, int n)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    if (i < n) a[i] = i;
}

// ---- lists

// colleagues (list 5) and list 2 of the boxes of one level, one wave per
// box: the candidates (children of the parent and of the parent's
// colleagues, in colleague then slot order) are taken 64 at a time and
// compacted in order with ballots
__global__ __launch_bounds__(OCT_T) void k_oct_coll(BoxArrays B, int first, int count, int *__restrict__ n5,
                                                    int *__restrict__ l5, int *__restrict__ n2, int *__restrict__ l2)
{
    int t = (blockIdx.x * OCT_T + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (t >= count) return;                          // whole waves leave together
    int b = first + t;
    int dad = B.parent[b];
    int ncoll = 1 + n5[dad];
    const int i0 = B.ijk[3 * b], i1 = B.ijk[3 * b + 1], i2 = B.ijk[3 * b + 2];
    const uint64_t below = (1ull << lane) - 1;
    int c5 = 0, c2 = 0;
    for (int f0 = 0; f0 < 8 * ncoll; f0 += 64) {
        int f = f0 + lane, ci = f >> 3, q = f & 7;
        bool valid = false, near = false;
        int kid = -1;
        if (ci < ncoll) {
            int coll = ci == 0 ? dad : l5[OCT_COLL * (size_t)dad + ci - 1];
            if (q < B.nkids[coll]) {
                kid = B.kids[8 * coll + q];
                if (kid != b) {
                    valid = true;
                    const int *ik = &B.ijk[3 * kid];
                    near = abs(ik[0] - i0) <= 1 && abs(ik[1] - i1) <= 1 && abs(ik[2] - i2) <= 1;
                }
            }
        }
        uint64_t m5 = __ballot(valid && near), m2 = __ballot(valid && !near);
        if (valid && near) l5[OCT_COLL * (size_t)b + c5 + __popcll(m5 & below)] = kid;
        if (valid && !near) l2[OCT_L2 * (size_t)b + c2 + __popcll(m2 & below)] = kid;
        c5 += __popcll(m5);
        c2 += __popcll(m2);
    }
    if (lane == 0) {
        n5[b] = c5;
        n2[b] = c2;
    }
}

__device__ __forceinline__ bool oct_touch(const BoxArrays &B, int a, int b)
{
    int la = B.level[a], lb = B.level[b];
    int L = la > lb ? la : lb;
    int sa = L - la, sb = L - lb;
    for (int d = 0; d < 3; d++) {
        long lo_a = (long)(B.ijk[3 * a + d] - 1) << sa, hi_a = (long)B.ijk[3 * a + d] << sa;
        long lo_b = (long)(B.ijk[3 * b + d] - 1) << sb, hi_b = (long)B.ijk[3 * b + d] << sb;
        if (lo_a > hi_b || lo_b > hi_a) return false;
    }
    return true;
}

// lists 1 and 3 of the leaves by the reference's descent, one thread per
// (box, colleague) entry q of the colleague CSR (owner[q] = box): fill = 0
// counts (c1, c3, ca per entry), fill = 1 writes own list-1 entries (t1 at
// o1), list 3 (l3 at o3, with (j, b) pairs p4 for list 4) and the (fine,
// coarse) pairs pa at oa. Entries of one box are consecutive, in colleague
// order, so the exclusive sums over entries give each box's lists in the
// reference's order.
__global__ __launch_bounds__(OCT_T) void k_oct_desc(BoxArrays B, int nq, const int *__restrict__ owner, const int *__restrict__ coll,
                           int fill, int *__restrict__ c1, int *__restrict__ c3, int *__restrict__ ca,
                           const int *__restrict__ o1, const int *__restrict__ o3, const int *__restrict__ oa,
                           int *__restrict__ t1, int *__restrict__ l3, uint64_t *__restrict__ p4,
                           uint64_t *__restrict__ pa)
{
    int q = blockIdx.x * OCT_T + threadIdx.x;
    if (q >= nq) return;
    int b = owner[q];
    int k1 = 0, k3 = 0, ka = 0;
    if (B.nkids[b] == 0 && B.level[b] > 0) {
        // the descent of the reference without a stack: children are
        // consecutive boxes, so the next one to visit (the previous slot)
        // and the way back up follow from the parent links
        const int c0 = coll[q];
        int j = c0;
        bool down = true;                            // j is to be examined (else: j is finished)
        while (true) {
            if (down) {
                if (!oct_touch(B, b, j)) {
                    if (fill) {
                        l3[o3[q] + k3] = j;
                        p4[o3[q] + k3] = ((uint64_t)j << 32) | (uint32_t)b;
                    }
                    k3++;
                    down = false;
                } else if (B.nkids[j] == 0) {
                    if (fill) t1[o1[q] + k1] = j;
                    k1++;
                    if (B.level[j] != B.level[b]) {
                        if (fill) pa[oa[q] + ka] = ((uint64_t)j << 32) | (uint32_t)b;
                        ka++;
                    }
                    down = false;
                } else {
                    j = B.kids[8 * j + B.nkids[j] - 1];   // last child first
                }
                continue;
            }
            if (j == c0) break;
            int par = B.parent[j];
            int slot = j - B.kids[8 * par];
            if (slot > 0) {
                j = B.kids[8 * par + slot - 1];
                down = true;
            } else {
                j = par;                                  // all children of par done
            }
        }
    }
    if (!fill) {
        c1[q] = k1;
        c3[q] = k3;
        ca[q] = ka;
    }
}

// per-box starts from per-entry starts: box b's entries begin at st5[b]
__global__ void k_oct_boxstart(int nb, const int *__restrict__ st5, const int *__restrict__ oq,
                               int *__restrict__ ob)
{
    int b = blockIdx.x * OCT_T + threadIdx.x;
    if (b <= nb) ob[b] = oq[st5[b]];
}

// count of sorted pairs per first box
__global__ void k_oct_paircount(const uint64_t *__restrict__ p, int np, int *__restrict__ cnt)
{
    int k = blockIdx.x * OCT_T + threadIdx.x;
    if (k < np) atomicAdd(&cnt[(int)(p[k] >> 32)], 1);
}

// list 1: appended coarse leaves (sorted pairs) first, then own entries
__global__ void k_oct_list1_app(const uint64_t *__restrict__ pa, int np, const int *__restrict__ start1,
                                const int *__restrict__ aexcl, int *__restrict__ items1)
{
    int k = blockIdx.x * OCT_T + threadIdx.x;
    if (k >= np) return;
    int j = (int)(pa[k] >> 32);
    items1[start1[j] + (k - aexcl[j])] = (int)(pa[k] & 0xffffffffu);
}

__global__ void k_oct_list1_own(int nb, const int *__restrict__ o1, const int *__restrict__ acnt,
                                const int *__restrict__ t1, const int *__restrict__ start1, int *__restrict__ items1)
{
    int b = blockIdx.x * OCT_T + threadIdx.x;
    if (b >= nb) return;
    for (int t = 0; t < o1[b + 1] - o1[b]; t++) items1[start1[b] + acnt[b] + t] = t1[o1[b] + t];
}

__global__ void k_oct_low32(const uint64_t *__restrict__ p, int np, int *__restrict__ out)
{
    int k = blockIdx.x * OCT_T + threadIdx.x;
    if (k < np) out[k] = (int)(p[k] & 0xffffffffu);
}

// rows of a fixed-stride table into CSR (owner, if given: the row of each
// entry)
__global__ void k_oct_rows(int nb, int stride, const int *__restrict__ cnt, const int *__restrict__ tab,
                           const int *__restrict__ start, int *__restrict__ items, int *__restrict__ owner)
{
    int b = blockIdx.x * OCT_T + threadIdx.x;
    if (b >= nb) return;
    for (int t = 0; t < cnt[b]; t++) {
        items[start[b] + t] = tab[(size_t)stride * b + t];
        if (owner) owner[start[b] + t] = b;
    }
}

// list-1 length: appended entries plus own entries
__global__ void k_oct_cnt1(int nb, const int *__restrict__ acnt, const int *__restrict__ o1, int *__restrict__ c)
{
    int i = blockIdx.x * OCT_T + threadIdx.x;
    if (i < nb) c[i] = acnt[i] + o1[i + 1] - o1[i];
    if (i == nb) c[i] = 0;
}

int bits_for(long v)
{
    int b = 1;
    while (b < 63 && (1L << b) <= v) b++;
    return b;
}

}  // namespace

struct OctreeWork::Impl {
    struct Buf {
        void *p = nullptr;
        size_t bytes = 0;
        ~Buf()
        {
            if (p) (void)hipFree(p);
        }
        hipError_t ensure(size_t b)
        {
            if (b <= bytes) return hipSuccess;
            if (p) (void)hipFree(p);
            p = nullptr;
            bytes = 0;
            size_t nb = std::max<size_t>(b + b / 4, 256);
            hipError_t e = hipMalloc(&p, nb);
            if (e == hipSuccess) bytes = nb;
            return e;
        }
        // grow keeping the first `keep` bytes
        hipError_t grow(size_t b, size_t keep, hipStream_t s)
        {
            if (b <= bytes) return hipSuccess;
            void *q = nullptr;
            size_t nb = std::max<size_t>(2 * b, 256);
            hipError_t e = hipMalloc(&q, nb);
            if (e != hipSuccess) return e;
            if (p && keep) {
                e = hipMemcpyAsync(q, p, keep, hipMemcpyDeviceToDevice, s);
                if (e != hipSuccess) {
                    (void)hipFree(q);
                    return e;
                }
                e = hipStreamSynchronize(s);
                if (e != hipSuccess) {
                    (void)hipFree(q);
                    return e;
                }
            }
            if (p) (void)hipFree(p);
            p = q;
            bytes = nb;
            return hipSuccess;
        }
        template <class T>
        T *get()
        {
            return (T *)p;
        }
    };
    Buf part, root, key, key2, idx, idx2, tag, sexcl, head, incl, active, cur, dcount, tmp;
    Buf box[11];                                    // level, ijk, parent, kids, nkids, ss, sc, ts, tc, split, firstkid
    Buf sleaf, tleaf, iota, skey, iz, izt;
    Buf n5, l5, n2, l2, owner, c1, c3, ca, o1q, o3q, oaq, o1, t1, p4, p4s, pa, pas, acnt, aexcl, cnt4, cnt1;
    Buf st[6], it[6];
    size_t cap = 0;                                 // boxes the box arrays hold
};

OctreeWork::OctreeWork() : impl(new Impl()) {}
OctreeWork::~OctreeWork() { delete impl; }

hipError_t build_tree_dev(const double *d_src, int n, const double *d_tgt, int nt, int nbox, int imindlin,
                          Tree &T, OctreeWork &work, hipStream_t s, bool &too_deep)
{
    too_deep = false;
    if (n < 1) return hipErrorInvalidValue;
    OctreeWork::Impl &W = *work.impl;
    const int N = n + nt;
    // cub scratch sized for the largest call below
    auto cub_sort64 = [&](const uint64_t *ki, uint64_t *ko, const int *vi, int *vo, int num, int endbit) -> hipError_t {
        size_t tb = 0;
        OCTOK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, ki, ko, vi, vo, num, 0, endbit, s));
        OCTOK(W.tmp.ensure(tb));
        return hipcub::DeviceRadixSort::SortPairs(W.tmp.p, tb, ki, ko, vi, vo, num, 0, endbit, s);
    };
    auto cub_sortkeys64 = [&](const uint64_t *ki, uint64_t *ko, int num, int endbit) -> hipError_t {
        size_t tb = 0;
        OCTOK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, ki, ko, num, 0, endbit, s));
        OCTOK(W.tmp.ensure(tb));
        return hipcub::DeviceRadixSort::SortKeys(W.tmp.p, tb, ki, ko, num, 0, endbit, s);
    };
    auto cub_sort32 = [&](const unsigned *ki, unsigned *ko, const int *vi, int *vo, int num, int endbit) -> hipError_t {
        size_t tb = 0;
        OCTOK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, ki, ko, vi, vo, num, 0, endbit, s));
        OCTOK(W.tmp.ensure(tb));
        return hipcub::DeviceRadixSort::SortPairs(W.tmp.p, tb, ki, ko, vi, vo, num, 0, endbit, s);
    };
    auto cub_incl = [&](const int *in, int *out, int num) -> hipError_t {
        size_t tb = 0;
        OCTOK(hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, out, num, s));
        OCTOK(W.tmp.ensure(tb));
        return hipcub::DeviceScan::InclusiveSum(W.tmp.p, tb, in, out, num, s);
    };
    auto cub_excl = [&](const int *in, int *out, int num) -> hipError_t {
        size_t tb = 0;
        OCTOK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, num, s));
        OCTOK(W.tmp.ensure(tb));
        return hipcub::DeviceScan::ExclusiveSum(W.tmp.p, tb, in, out, num, s);
    };

    // root cube
    const int nblk = 120;
    OCTOK(W.part.ensure(6 * nblk * 8));
    OCTOK(W.root.ensure(4 * 8));
    hipLaunchKernelGGL(k_oct_bbox, dim3(nblk), dim3(OCT_T), 0, s, d_src, n, d_tgt, nt, W.part.get<double>());
    hipLaunchKernelGGL(k_oct_root, dim3(1), dim3(1), 0, s, W.part.get<double>(), nblk, imindlin,
                       W.root.get<double>());
    OCTOK(hipGetLastError());

    // paths, sorted
    OCTOK(W.key.ensure((size_t)N * 8));
    OCTOK(W.key2.ensure((size_t)N * 8));
    OCTOK(W.idx.ensure((size_t)N * 4));
    OCTOK(W.idx2.ensure((size_t)N * 4));
    hipLaunchKernelGGL(k_oct_keys, dim3(blocks(N)), dim3(OCT_T), 0, s, d_src, n, d_tgt, nt, W.root.get<double>(),
                       W.key.get<uint64_t>(), W.idx.get<int>());
    OCTOK(hipGetLastError());
    OCTOK(cub_sort64(W.key.get<uint64_t>(), W.key2.get<uint64_t>(), W.idx.get<int>(), W.idx2.get<int>(), N, 3 * OCT_D));
    const uint64_t *K = W.key2.get<uint64_t>();
    const int *P = W.idx2.get<int>();
    for (OctreeWork::Impl::Buf *b : {&W.tag, &W.sexcl, &W.head, &W.incl, &W.active, &W.cur}) OCTOK(b->ensure((size_t)N * 4));
    OCTOK(W.dcount.ensure(2 * 4));
    hipLaunchKernelGGL(k_oct_srctag, dim3(blocks(N)), dim3(OCT_T), 0, s, P, N, n, W.tag.get<int>());
    OCTOK(hipGetLastError());
    OCTOK(cub_excl(W.tag.get<int>(), W.sexcl.get<int>(), N));

    // boxes, level by level
    const size_t ints_per_box[11] = {1, 3, 1, 8, 1, 1, 1, 1, 1, 1, 1};
    auto ensure_boxes = [&](size_t need, size_t keep) -> hipError_t {
        if (need <= W.cap) return hipSuccess;
        for (int a = 0; a < 11; a++) OCTOK(W.box[a].grow(need * ints_per_box[a] * 4, keep * ints_per_box[a] * 4, s));
        W.cap = W.box[0].bytes / 4;
        for (int a = 0; a < 11; a++) W.cap = std::min(W.cap, W.box[a].bytes / (4 * ints_per_box[a]));
        return hipSuccess;
    };
    auto arrays = [&]() {
        return BoxArrays{W.box[0].get<int>(), W.box[1].get<int>(), W.box[2].get<int>(), W.box[3].get<int>(),
                         W.box[4].get<int>(), W.box[5].get<int>(), W.box[6].get<int>(), W.box[7].get<int>(),
                         W.box[8].get<int>(), W.box[9].get<int>(), W.box[10].get<int>()};
    };
    OCTOK(ensure_boxes((size_t)N + 1, 0));
    hipLaunchKernelGGL(k_oct_root_box, dim3(blocks(N)), dim3(OCT_T), 0, s, arrays(), n, nt, nbox, N,
                       W.active.get<int>(), W.cur.get<int>());
    OCTOK(hipGetLastError());
    std::vector<int> laddr = {0, 1};
    int nb = 1, nlev = 0;
    int nsplit = (n > nbox || nt > nbox) ? 1 : 0;
    int hc[2];
    for (int lev = 1; nsplit > 0; lev++) {
        if (lev > OCT_D) {
            too_deep = true;
            return hipSuccess;
        }
        OCTOK(ensure_boxes((size_t)nb + N, (size_t)nb));
        BoxArrays B = arrays();
        int sh = 3 * (OCT_D - lev);
        hipLaunchKernelGGL(k_oct_heads, dim3(blocks(N)), dim3(OCT_T), 0, s, K, W.active.get<int>(), N, sh,
                           W.head.get<int>());
        OCTOK(hipGetLastError());
        OCTOK(cub_incl(W.head.get<int>(), W.incl.get<int>(), N));
        OCTOK(hipMemsetAsync(W.dcount.p, 0, 8, s));
        hipLaunchKernelGGL(k_oct_boxes, dim3(blocks(N)), dim3(OCT_T), 0, s, B, K, W.active.get<int>(),
                           W.head.get<int>(), W.incl.get<int>(), W.sexcl.get<int>(), N, lev, sh, nb, W.cur.get<int>(),
                           W.dcount.get<int>());
        hipLaunchKernelGGL(k_oct_counts, dim3(blocks(N)), dim3(OCT_T), 0, s, B, K, W.active.get<int>(),
                           W.cur.get<int>(), W.tag.get<int>(), W.sexcl.get<int>(), N, sh);
        hipLaunchKernelGGL(k_oct_split, dim3(blocks(N)), dim3(OCT_T), 0, s, B, W.head.get<int>(), W.cur.get<int>(),
                           N, nbox, W.dcount.get<int>());
        hipLaunchKernelGGL(k_oct_kids, dim3(blocks(N)), dim3(OCT_T), 0, s, B, W.head.get<int>(), W.cur.get<int>(),
                           W.active.get<int>(), N);
        OCTOK(hipGetLastError());
        OCTOK(hipMemcpyAsync(hc, W.dcount.p, 8, hipMemcpyDeviceToHost, s));
        OCTOK(hipStreamSynchronize(s));
        laddr.push_back(nb);
        laddr.push_back(hc[0]);
        nb += hc[0];
        nlev = lev;
        nsplit = hc[1];
    }
    BoxArrays B = arrays();

    // point permutations: input order, stably by leaf
    const int lbits = bits_for(nb), nbits = bits_for(N);
    for (OctreeWork::Impl::Buf *b : {&W.sleaf, &W.tleaf, &W.iota, &W.skey, &W.iz, &W.izt}) OCTOK(b->ensure((size_t)std::max(n, nt) * 4 + 4));
    hipLaunchKernelGGL(k_oct_leaf, dim3(blocks(N)), dim3(OCT_T), 0, s, B, P, W.cur.get<int>(), N, n,
                       W.sleaf.get<unsigned>(), W.tleaf.get<unsigned>());
    hipLaunchKernelGGL(k_oct_iota, dim3(blocks(std::max(n, nt))), dim3(OCT_T), 0, s, W.iota.get<int>(), std::max(n, nt));
    OCTOK(hipGetLastError());
    OCTOK(cub_sort32(W.sleaf.get<unsigned>(), W.skey.get<unsigned>(), W.iota.get<int>(), W.iz.get<int>(), n, nbits));
    if (nt)
        OCTOK(cub_sort32(W.tleaf.get<unsigned>(), W.skey.get<unsigned>(), W.iota.get<int>(), W.izt.get<int>(), nt,
                         nbits));

    // colleagues and list 2, level by level
    OCTOK(W.n5.ensure((size_t)nb * 4 + 4));
    OCTOK(W.l5.ensure((size_t)nb * OCT_COLL * 4));
    OCTOK(W.n2.ensure((size_t)nb * 4 + 4));
    OCTOK(W.l2.ensure((size_t)nb * OCT_L2 * 4));
    OCTOK(hipMemsetAsync(W.n5.p, 0, 4, s));
    OCTOK(hipMemsetAsync(W.n2.p, 0, 4, s));
    for (int lev = 1; lev <= nlev; lev++) {
        int f = laddr[2 * lev], c = laddr[2 * lev + 1];
        hipLaunchKernelGGL(k_oct_coll, dim3(blocks((size_t)c * 64)), dim3(OCT_T), 0, s, B, f, c, W.n5.get<int>(),
                           W.l5.get<int>(), W.n2.get<int>(), W.l2.get<int>());
    }
    OCTOK(hipGetLastError());

    // lists 5 and 2 from their tables (counts with a trailing zero); each
    // list-5 entry also names its box, for the descent below
    for (int t = 1; t <= 5; t++) OCTOK(W.st[t].ensure((size_t)(nb + 1) * 4));
    OCTOK(hipMemsetAsync(W.n5.get<int>() + nb, 0, 4, s));
    OCTOK(hipMemsetAsync(W.n2.get<int>() + nb, 0, 4, s));
    OCTOK(cub_excl(W.n5.get<int>(), W.st[5].get<int>(), nb + 1));
    OCTOK(cub_excl(W.n2.get<int>(), W.st[2].get<int>(), nb + 1));
    int t5 = 0, t2 = 0;
    OCTOK(hipMemcpyAsync(&t5, W.st[5].get<int>() + nb, 4, hipMemcpyDeviceToHost, s));
    OCTOK(hipMemcpyAsync(&t2, W.st[2].get<int>() + nb, 4, hipMemcpyDeviceToHost, s));
    OCTOK(hipStreamSynchronize(s));
    OCTOK(W.it[5].ensure((size_t)t5 * 4 + 4));
    OCTOK(W.it[2].ensure((size_t)t2 * 4 + 4));
    OCTOK(W.owner.ensure((size_t)t5 * 4 + 4));
    hipLaunchKernelGGL(k_oct_rows, dim3(blocks(nb)), dim3(OCT_T), 0, s, nb, OCT_COLL, W.n5.get<int>(), W.l5.get<int>(),
                       W.st[5].get<int>(), W.it[5].get<int>(), W.owner.get<int>());
    hipLaunchKernelGGL(k_oct_rows, dim3(blocks(nb)), dim3(OCT_T), 0, s, nb, OCT_L2, W.n2.get<int>(), W.l2.get<int>(),
                       W.st[2].get<int>(), W.it[2].get<int>(), nullptr);
    OCTOK(hipGetLastError());

    // lists 1 and 3 (and the pairs of the appends and of list 4), per
    // list-5 entry; counts with a trailing zero
    for (OctreeWork::Impl::Buf *b : {&W.c1, &W.c3, &W.ca, &W.o1q, &W.o3q, &W.oaq})
        OCTOK(b->ensure((size_t)(t5 + 1) * 4));
    for (OctreeWork::Impl::Buf *b : {&W.o1, &W.acnt, &W.aexcl, &W.cnt4, &W.cnt1}) OCTOK(b->ensure((size_t)(nb + 1) * 4));
    if (t5)
        hipLaunchKernelGGL(k_oct_desc, dim3(blocks(t5)), dim3(OCT_T), 0, s, B, t5, W.owner.get<int>(),
                           W.it[5].get<int>(), 0, W.c1.get<int>(), W.c3.get<int>(), W.ca.get<int>(), nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr);
    OCTOK(hipGetLastError());
    for (OctreeWork::Impl::Buf *b : {&W.c1, &W.c3, &W.ca}) OCTOK(hipMemsetAsync(b->get<int>() + t5, 0, 4, s));
    OCTOK(cub_excl(W.c1.get<int>(), W.o1q.get<int>(), t5 + 1));
    OCTOK(cub_excl(W.c3.get<int>(), W.o3q.get<int>(), t5 + 1));
    OCTOK(cub_excl(W.ca.get<int>(), W.oaq.get<int>(), t5 + 1));
    int tot[3];
    OCTOK(hipMemcpyAsync(&tot[0], W.o1q.get<int>() + t5, 4, hipMemcpyDeviceToHost, s));
    OCTOK(hipMemcpyAsync(&tot[1], W.o3q.get<int>() + t5, 4, hipMemcpyDeviceToHost, s));
    OCTOK(hipMemcpyAsync(&tot[2], W.oaq.get<int>() + t5, 4, hipMemcpyDeviceToHost, s));
    OCTOK(hipStreamSynchronize(s));
    const int n1own = tot[0], n3 = tot[1], nap = tot[2];
    OCTOK(W.t1.ensure((size_t)n1own * 4 + 4));
    for (OctreeWork::Impl::Buf *b : {&W.p4, &W.p4s}) OCTOK(b->ensure((size_t)n3 * 8 + 8));
    for (OctreeWork::Impl::Buf *b : {&W.pa, &W.pas}) OCTOK(b->ensure((size_t)nap * 8 + 8));
    OCTOK(W.it[3].ensure((size_t)n3 * 4 + 4));
    OCTOK(W.it[4].ensure((size_t)n3 * 4 + 4));
    if (t5)
        hipLaunchKernelGGL(k_oct_desc, dim3(blocks(t5)), dim3(OCT_T), 0, s, B, t5, W.owner.get<int>(),
                           W.it[5].get<int>(), 1, nullptr, nullptr, nullptr, W.o1q.get<int>(), W.o3q.get<int>(),
                           W.oaq.get<int>(), W.t1.get<int>(), W.it[3].get<int>(), W.p4.get<uint64_t>(),
                           W.pa.get<uint64_t>());
    // per-box starts of the own list-1 entries and of list 3
    hipLaunchKernelGGL(k_oct_boxstart, dim3(blocks(nb + 1)), dim3(OCT_T), 0, s, nb, W.st[5].get<int>(),
                       W.o1q.get<int>(), W.o1.get<int>());
    hipLaunchKernelGGL(k_oct_boxstart, dim3(blocks(nb + 1)), dim3(OCT_T), 0, s, nb, W.st[5].get<int>(),
                       W.o3q.get<int>(), W.st[3].get<int>());
    OCTOK(hipGetLastError());
    const int pbits = 32 + lbits;
    // list 4: (j, b) pairs sorted, counted per j
    if (n3) OCTOK(cub_sortkeys64(W.p4.get<uint64_t>(), W.p4s.get<uint64_t>(), n3, pbits));
    OCTOK(hipMemsetAsync(W.cnt4.p, 0, (size_t)(nb + 1) * 4, s));
    if (n3) {
        hipLaunchKernelGGL(k_oct_paircount, dim3(blocks(n3)), dim3(OCT_T), 0, s, W.p4s.get<uint64_t>(), n3,
                           W.cnt4.get<int>());
        hipLaunchKernelGGL(k_oct_low32, dim3(blocks(n3)), dim3(OCT_T), 0, s, W.p4s.get<uint64_t>(), n3,
                           W.it[4].get<int>());
    }
    OCTOK(hipGetLastError());
    OCTOK(cub_excl(W.cnt4.get<int>(), W.st[4].get<int>(), nb + 1));
    // list 1: appended coarse leaves first, then the leaf's own entries
    if (nap) OCTOK(cub_sortkeys64(W.pa.get<uint64_t>(), W.pas.get<uint64_t>(), nap, pbits));
    OCTOK(hipMemsetAsync(W.acnt.p, 0, (size_t)(nb + 1) * 4, s));
    if (nap)
        hipLaunchKernelGGL(k_oct_paircount, dim3(blocks(nap)), dim3(OCT_T), 0, s, W.pas.get<uint64_t>(), nap,
                           W.acnt.get<int>());
    OCTOK(hipGetLastError());
    OCTOK(cub_excl(W.acnt.get<int>(), W.aexcl.get<int>(), nb + 1));
    hipLaunchKernelGGL(k_oct_cnt1, dim3(blocks(nb + 1)), dim3(OCT_T), 0, s, nb, W.acnt.get<int>(), W.o1.get<int>(),
                       W.cnt1.get<int>());
    OCTOK(hipGetLastError());
    OCTOK(cub_excl(W.cnt1.get<int>(), W.st[1].get<int>(), nb + 1));
    OCTOK(W.it[1].ensure((size_t)(n1own + nap) * 4 + 4));
    if (nap)
        hipLaunchKernelGGL(k_oct_list1_app, dim3(blocks(nap)), dim3(OCT_T), 0, s, W.pas.get<uint64_t>(), nap,
                           W.st[1].get<int>(), W.aexcl.get<int>(), W.it[1].get<int>());
    hipLaunchKernelGGL(k_oct_list1_own, dim3(blocks(nb)), dim3(OCT_T), 0, s, nb, W.o1.get<int>(), W.acnt.get<int>(),
                       W.t1.get<int>(), W.st[1].get<int>(), W.it[1].get<int>());
    OCTOK(hipGetLastError());

    // to the host
    T = Tree();
    T.nboxes = nb;
    T.nlev = nlev;
    T.laddr = laddr;
    double root[4];
    OCTOK(hipMemcpyAsync(root, W.root.p, 32, hipMemcpyDeviceToHost, s));
    T.level.resize(nb);
    T.ijk.resize(3 * (size_t)nb);
    T.parent.resize(nb);
    T.kids.resize(8 * (size_t)nb);
    T.nkids.resize(nb);
    T.src_start.resize(nb);
    T.src_count.resize(nb);
    T.tgt_start.resize(nb);
    T.tgt_count.resize(nb);
    std::vector<int> *dst[9] = {&T.level, &T.ijk, &T.parent, &T.kids, &T.nkids, &T.src_start, &T.src_count,
                                &T.tgt_start, &T.tgt_count};
    for (int a = 0; a < 9; a++)
        OCTOK(hipMemcpyAsync(dst[a]->data(), W.box[a].p, dst[a]->size() * 4, hipMemcpyDeviceToHost, s));
    T.iz.resize(n);
    T.iztarg.resize(nt);
    OCTOK(hipMemcpyAsync(T.iz.data(), W.iz.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    if (nt) OCTOK(hipMemcpyAsync(T.iztarg.data(), W.izt.p, (size_t)nt * 4, hipMemcpyDeviceToHost, s));
    const int tot_items[6] = {0, n1own + nap, t2, n3, n3, t5};
    for (int t = 1; t <= 5; t++) {
        T.lstart[t].resize(nb + 1);
        T.litems[t].resize(tot_items[t]);
        OCTOK(hipMemcpyAsync(T.lstart[t].data(), W.st[t].p, (size_t)(nb + 1) * 4, hipMemcpyDeviceToHost, s));
        if (tot_items[t])
            OCTOK(hipMemcpyAsync(T.litems[t].data(), W.it[t].p, (size_t)tot_items[t] * 4, hipMemcpyDeviceToHost, s));
    }
    OCTOK(hipStreamSynchronize(s));
    for (int d = 0; d < 3; d++) T.center[d] = root[d];
    T.size = root[3];
    // box centres of d3tcentc (src/d3tstrcr.f:1596-1676), as build_tree
    T.centers.resize(3 * (size_t)nb);
    for (int b = 0; b < nb; b++) {
        double side = T.size / (double)(1 << T.level[b]), side2 = side / 2;
        for (int d = 0; d < 3; d++)
            T.centers[3 * b + d] = T.center[d] - T.size / 2 + (T.ijk[3 * b + d] - 1) * side + side2;
    }
    return hipSuccess;
}

}  // namespace elfmm
