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
// Leaf size of the point-source Laplace FMM (lfmm3dparthesstarg).
int iprec_nbox_part(int iprec, int npts);
// leaf size of the iter engine (src/lfmm3dtrilhess_iter_1.f:246-255): twice
// the table above
int iprec_nbox_iter(int iprec, int npts);
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
    // side of each box when the boxes are not all of one root cube (a
    // forest of octrees); empty: size / 2^level
    std::vector<double> box_side;
};

// Forest of octrees: the trees of several point sets as one tree-like
// structure for one plan. Boxes are renumbered level-major (every tree's
// level-0 box, then every tree's level-1 boxes, ...), points keep their
// trees' order with the sources (targets) of tree k starting at soff[k]
// (toff[k]), and iz / iztarg map to the packed input order.
void build_forest(const std::vector<const Tree *> &trees, const std::vector<int> &soff,
                  const std::vector<int> &toff, Tree &F);

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
    // M2M / L2L per octant, rotated into the xz-plane (as M2L): two real
    // (nc/2)^2 matrices each, real-part then imaginary-part, and the azimuth
    std::vector<double> m2m;
    std::vector<double> l2l;
    double oct_phi[8];
    // M2L per offset (i,j,k) in [-3,3]^3 with max|.|>=2: truncation order
    // pt, azimuth phi of the translation, and at m2l_off[o] the two real
    // (ncoef(pt)/2)^2 matrices of the translation rotated into the xz-plane
    // (real parts, then imaginary parts; see build_operators).
    int m2l_pt[343];
    int64_t m2l_off[343];
    // 1 when offset o uses the stored operator of its mirror image in z:
    // R(i, j, -k) = S R(i, j, k) S with S = diag((-1)^(n+m))
    int m2l_refl[343];
    double m2l_phi[343];
    std::vector<double> m2l;
};
void build_operators(int p, const int *itable, Operators &ops);

// Gather table of the field operand of L2T (irr = false, harmonics to
// degree p) and M2T (irr = true, degree p + 2). Entry (pos*2 + part)*9 + q
// gives the real (part 0) or minus the imaginary (part 1) part of field q
// (0 value, 1-3 d/dx d/dy d/dz, 4 xx, 5 yy, 6 xy, 7 xz, 8 yz) of one
// packed real expansion of order p, at the harmonic with stream position
// pos (m outer, l inner; padded to an even count with zero entries), times
// the weight of the real field sum (1 for m = 0, else 2) and
// 1/sqrt((l+m)!(l-m)!): the sum of c[t] * E[idx[t]], t < 3 (E as doubles,
// re/im interleaved). Every such entry has at most three terms.
struct alignas(16) FieldTerm {   // 48 bytes; the device reads idx and c[0..1] as 16-byte vectors
    int idx[4];
    double c[4];
};
void build_field_table(int p, bool irr, std::vector<FieldTerm> &tab);

// One tile of a rotated translation (M2L, M2M, L2L): ML_M coefficients
// (real and imaginary parts) of one operator for up to ML_P (source,
// destination) pairs.
struct M2LTile {
    int64_t a_off;   // offset of the real-part matrix (h x h, row-major); imaginary part follows
    int h;           // coefficients = ncoef(pt) / 2
    int row0;        // first output coefficient of this tile
    int pair0;       // first pair of this tile
    int npair;       // <= ML_P
    double phi;      // azimuth of the translation
    int refl = 0;    // 1: the stored operator is that of the z-mirrored offset (signs (-1)^(n+m) on both sides)
};
constexpr int ML_M = 64, ML_P = 32;

// One launch of rotated translations (M2M or L2L of one level): pairs and
// their tiles, run by the same kernel as M2L.
struct GemmLaunch {
    std::vector<M2LTile> tiles;
    std::vector<int> src, dst;
    std::vector<double> scale;
};
// offsets of one GemmLaunch's arrays in a call's device plan buffer (kept
// by the call, not in the plan, which iter calls of one geometry share)
struct LaunchOff {
    size_t tiles = 0, src = 0, dst = 0, sc = 0;
};

// Task lists of one apply, all box ids 0-based in tree order; points are
// numbered sources first (0..n-1), then targets (n..n+nt-1).
struct Plan {
    std::vector<double> sides;                       // per box
    std::vector<int> p2m_boxes;                      // leaves forming a multipole
    std::vector<GemmLaunch> m2m_l, l2l_l;            // one launch per level
    GemmLaunch m2l;                                  // all list-2 pairs, grouped by offset (src/dst/scale)
    std::vector<M2LTile> m2l_tiles;                  // rotated-M2L tiles over those pairs
    std::vector<int> p2l_boxes, p2l_start, p2l_items;   // box <- sources of its list-4 boxes
    std::vector<int> m2t_boxes, m2t_start, m2t_items;   // box <- multipoles of its list-3 boxes
    std::vector<int> l2t_boxes, nf_boxes, nf_start, nf_items;
    std::vector<int> l2t_tpt;                        // points of each l2t slot, slot after slot
    std::vector<int> m2t_slot_box, m2t_tpt;          // points of each m2t slot, slot after slot
    // evaluation wave tasks: (slot, first task, task count <= EW_PTS) triples
    // over the slot-contiguous (point, slot) tasks above
    std::vector<int> l2t_wt, m2t_wt;
    // M2T workgroups: first wave task of each (all of one slot, at most 4),
    // then the total
    std::vector<int> m2t_bt;
    double stats[16];                                // as elfmm_gfx950_last_stats
};

constexpr int EW_PTS = 32;                           // points per evaluation wave task

// Interaction plan of one apply (the work of the reference's
// lfmm3dtrilhesstargmain steps 1-7 and of the near-field pass, as task
// lists). ops == nullptr plans the direct sum only; nrhs is the number of
// right-hand sides the work counts assume. iter selects the iter engine's
// split: no P2L/M2T, and lists 3 and 4 join self + list 1 in the near
// field (src/elfmm3dtria_iter_1.f:1061-1110).
void build_plan(const Tree &T, int n, const Operators *ops, bool src_eval, bool direct_only, int nrhs,
                Plan &P, bool iter = false);
// The two halves of build_plan, for a caller that starts the near field
// before the far-field lists exist: build_plan_near fills the near-field
// lists (nf_*, stats[6]) of P, build_plan_far everything else; neither
// clears what the other fills (P starts from plan_reset).
void build_plan_near(const Tree &T, bool src_eval, bool direct_only, bool iter, Plan &P);
// Empties P for a new plan (zero stats); the M2L pair arrays keep their
// memory and length, which the next build_plan_far resizes and overwrites
// (no fresh allocation or clearing of its largest arrays per apply).
void plan_reset(Plan &P);
void build_plan_far(const Tree &T, int n, const Operators *ops, bool src_eval, int nrhs, Plan &P, bool iter);

// Task lists of the Mindlin harmonic far field (lfmm3d_list2_mindlin,
// src/lfmm3droutsmindlin.f:28-529) on its own octree, beside the Laplace
// plan of the same tree (P2M boxes, M2M/L2L launches, L2T tasks).
struct MindlinPlan {
    std::vector<int> src_boxes;                // boxes of level >= 2 with sources
    std::vector<int> b_scale_boxes[64];        // per level: boxes with sources and kids (B rescaled after M2M)
    // per target box of level >= 2 with targets: far list-2 sources (|dz| >=
    // 1.5 sides: plane waves), near ones (proxy-charge local); CSR
    std::vector<int> far_boxes, far_start, far_items;
    std::vector<int> near_boxes, near_start, near_items;
    GemmLaunch m2l;                            // near pairs, grouped by offset (full order)
    std::vector<M2LTile> m2l_tiles;
    std::vector<int> pw_slot;                  // per box: plane-wave slot (-1: none)
    int npw = 0;
};
// ops: translation operators of full order p for the near offsets
// (|dz| <= 1) of list 2.
void build_mindlin_plan(const Tree &T, const Operators &ops, MindlinPlan &MP);

}  // namespace elfmm
