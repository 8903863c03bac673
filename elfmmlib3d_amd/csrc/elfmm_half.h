// Host launchers of the half-space (Mindlin) kernels (elfmm_half.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace elfmm {

// Image near field over the near boxes of an iter plan (see k_hnear).
struct HalfNearArgs {
    const int64_t *tb;          // first integral of each near box (nq + 1)
    int nq;
    const int *boxes, *lstart, *litems, *src_start, *src_count, *tgt_start, *ntri_q;
    const int64_t *sbase;       // store (or tmp) base of each near box
    const double *pts, *tri_s, *nrm_s, *ssl_s, *sdl_s;
    int n, ifsl, ifdl, ifpt, ifst, units, W;
    double lam, mu;
};

// bytes of the work-queue scratch (counter + per-wave stacks) for ntask items
size_t adap_stack_bytes(int64_t ntask);
hipError_t launch_hnear(const HalfNearArgs &a, int64_t ntask, void *scratch, double *out, hipStream_t st);
hipError_t launch_hnear_sum(int n, const int *boxes, const int *tgt_start, const int64_t *sbase, const int *ntri_q,
                            const int *pstart, int nq, int npts, int nl, const double *tmp, double *nf,
                            hipStream_t st);
// image half of the direct sum at npt points (caller's order), added to
// pt_out (3 per point) and st_out (3x3 per point)
size_t hdirect_scratch_bytes(int npt, int n);
hipError_t launch_hdirect(int npt, const double *pts, int n, const double *tri, const double *nrm, const double *ssl,
                          const double *sdl, int ifsl, int ifdl, int ifpt, int ifst, double lam, double mu,
                          void *scratch, double *pt_out, double *st_out, hipStream_t st);
hipError_t launch_mirror(int n, const double *in, double *out, hipStream_t st);
// image part of the point-source half-space direct sum at the ne sources (if
// ne > 0: ne == n) and nt targets, caller's order, added to the outputs
hipError_t launch_pimage(int ne, int n, int nt, const double *src, const double *tgt, const double *ssl,
                         const double *sdl, const double *sdv, int ifsl, int ifdl, double lam, double mu, int ifpt,
                         double *pt, int ifst, double *st, int ifptt, double *ptt, int ifstt, double *stt,
                         hipStream_t s);

// ---- Mindlin harmonic far field (elfmm_mindlin.hip) ----
struct MindArgs {
    int p, ifsl, ifdl;
    double lam, mu;
    double tsize;                    // root box side
    const int *level, *src_start, *src_count;
    const double *centers;
    const double *tri, *nrm, *ssl, *sdl;   // triangles of the Mindlin tree's sources (mirrored), tree order
    int nc;                          // doubles per expansion (ncoef(p))
    double *mp, *lo;                 // [box][4][nc]
    int nqtri;                       // order of the P2M triangle rule: 2, 4 or 6 (3, 6 or 12 nodes)
};

// Plane-wave discretisation on the device (per physical sample j: lambda,
// cos u, sin u, unused; per lambda: first Fourier mode / sample, counts).
struct MindPw {
    int nlams, nexptot, nexptotp;
    const double *tab;        // per physical sample j: lam, cos u, sin u, 1/np
    const int *jnl;           // lambda index of physical sample j
    const int *foff, *poff;   // per lambda: first Fourier mode / physical sample
    const int *nfour, *nphys;
    const double *rlams, *whts;
};

struct MindDev {
    MindArgs args;
    MindPw pw;
    const int *slot;              // per box: plane-wave / proxy slot (-1: none)
    const double *ycst;           // ylgndr(p, 1/sqrt(3))
    int nq, nquad;                // proxy charges per ring; sphere quadrature size
    double *pwo, *skel;           // outgoing plane waves, proxy charges
    const double *gl, *ytab, *eph, *wl;
    double *gmom;                 // P2M moments, mind_p2m_scratch_bytes
};
// scratch of launch_mind_p2m for nbx boxes
size_t mind_p2m_scratch_bytes(int p, int nbx);
hipError_t launch_mind_p2m(const MindDev &d, const int *boxes, int nbx, hipStream_t st);
hipError_t launch_mind_bscale(const MindDev &d, const int *boxes, int nbx, hipStream_t st);
hipError_t launch_mind_bmod(const MindDev &d, const int *boxes, int nbx, hipStream_t st);
hipError_t launch_mind_proxy(const MindDev &d, const int *tb, const int *ts, const int *ti, int nbx, hipStream_t st);
hipError_t launch_mind_far(const MindDev &d, const int *tb, const int *ts, const int *ti, int nbx, hipStream_t st);
hipError_t launch_mind_assemble(int n, int nt, const int *izt, const double *pts, const double *lap, double mu,
                                int ifpt, double *ptt, int ifst, double *stt, hipStream_t st);
size_t mind_proxy_lds(int p, int nquad, int nq);
size_t mind_far_lds(int nexptotp, int nexptot);
void mind_set_lds_limits(size_t proxy, size_t far);

}  // namespace elfmm
