// Host build of the library's near-field pair code (elfmm_math.h, the code
// the near-field kernels run) for CPU tests.
#include "../../elfmmlib3d_amd/csrc/elfmm_math.h"

extern "C" void nf_host_pair(const double *tri, const double *ssl, const double *sdl, const double *P,
                             int iself, double rlam, double rmu, int ifsl, int ifdl, double *out9)
{
    double rec[elfmm::NF_REC];
    elfmm::nf_make_record(tri, ssl, sdl, rec);
    for (int i = 0; i < 9; i++) out9[i] = 0.0;
    elfmm::nf_pair(rec, P, iself, rlam, rmu, (rlam + rmu) / (rlam + 2 * rmu), ifsl, ifdl, out9);
}

extern "C" void nf_host_basis(const double *tri, const double *P, int iself, double rlam, double rmu,
                              int ifsl, int ifdl, double *out54)
{
    double rec[elfmm::NF_REC];
    elfmm::nf_make_record(tri, nullptr, nullptr, rec);
    for (int i = 0; i < 54; i++) out54[i] = 0.0;
    elfmm::nf_pair_basis(rec, P, iself, rlam, rmu, (rlam + rmu) / (rlam + 2 * rmu), ifsl, ifdl,
                         [&](int layer, int k, const double *v) {
                             for (int i = 0; i < 9; i++) out54[(layer * 3 + k) * 9 + i] = v[i];
                         });
}
