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
// Mindlin harmonic-image far field (lfmm3dmindlintriatarg on the mirrored
// triangles, src/elhfmm3dtria_iter.f:795-831), added to the target outputs
int mindlin_far_dev(int iprec, double rlam, double rmu, int n, const double *d_tri, const double *d_nrm,
                    const double *d_src, int ifsl, const double *d_ssl, int ifdl, const double *d_sdl, int nt,
                    const double *d_tgt, int ifptt, double *d_ptt, int ifstt, double *d_stt, hipStream_t stream);
hipError_t launch_mirror(int n, const double *in, double *out, hipStream_t st);

}  // namespace elfmm
