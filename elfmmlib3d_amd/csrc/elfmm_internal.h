// Internal declarations shared by the host planner and the gfx950 kernels.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

namespace elfmm {

// Host worker pool (elfmm_pool.cpp): fn(chunk, thread) is called once for
// every chunk in [0, nchunks), chunks taken dynamically by up to
// host_threads() threads; returns when all chunks are done.
int host_threads();
void parallel_chunks(int nchunks, const std::function<void(int, int)> &fn);

// FMM precision flag -> tolerance and leaf size, as the reference's
// lfmm3dtrilhesstarg (src/lfmm3dtrilhess.f:216-235): nbox = (120*2^k)/3.
double iprec_eps(int iprec);
int iprec_nbox(int iprec, int npts);
// expansion order of l3dterms (src/l3dterms.f:42-85)
int expansion_order(double eps);
// list-2 truncation table of l3dterms_list2 (src/l3dterms.f:144-256),
// itable[(i+3)*49+(j+3)*7+(k+3)] for offsets i,j,k in [-3,3]
void list2_orders(double eps, int *itable);

// Adaptive octree and interaction lists with the box numbering, point order
// and list order of d3tstrcr (src/d3tstrcr.f:31). Box ids are 0-based here;
// point permutations are 0-based (original index of the k-th sorted point).
struct Tree {
    int nboxes = 0, nlev = 0;
    double center[3] = {0, 0, 0}, size = 0;
    // per box
    std::vector<int> level, ijk, parent, kids;     // ijk: 3/box (1-based), kids: 8/box (-1 = none)
    std::vector<int> nkids, src_start, src_count, tgt_start, tgt_count;
    std::vector<double> centers;                   // 3/box
    std::vector<int> laddr;                        // 2/level: first box, count
    std::vector<int> iz, iztarg;                   // sorted -> original
    // lists 1..5 in CSR, 0-based box ids; index [t] for t in 1..5
    std::vector<int> lstart[6], litems[6];
};

// imindlin selects the half-space root rule of src/d3tstrcr_mindlin.f:1380.
int build_tree(const double *src, int n, const double *targ, int nt, int nbox,
               int imindlin, Tree &T);

// Packed expansion layout: for degree l and order 0<=m<=l the complex
// coefficient (re, im) sits at 2*(l*(l+1)/2 + m); ncoef(p) doubles total.
inline int ncoef(int p) { return (p + 1) * (p + 2); }

// Dense translation operators on packed, box-size-scaled expansions.
struct Operators {
    int p = 0;
    int nc = 0;                          // ncoef(p)
    std::vector<double> m2m;             // 8 octants, nc x nc each (row-major)
    std::vector<double> l2l;             // 8 octants
    // M2L per offset (i,j,k) in [-3,3]^3 with max|.|>=2: truncation order
    // pt and matrix of ncoef(pt)^2 at offset m2l_off[o].
    int m2l_pt[343];
    int64_t m2l_off[343];
    std::vector<double> m2l;
};
void build_operators(int p, const int *itable, Operators &ops);

}  // namespace elfmm
