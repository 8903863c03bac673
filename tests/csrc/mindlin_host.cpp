// Host build of the Mindlin far-field operations (elfmm_mindlin.h, the code
// the Mindlin far-field kernels run) for CPU tests against the reference.
#include <vector>

#include "../../elfmmlib3d_amd/csrc/elfmm_mindlin.h"

using namespace elfmm;

extern "C" int mh_total_phys(int iprec) { return pw_total_phys(pw_table(iprec)); }

extern "C" void mh_pw_from_multipole(int iprec, int p, const double *mt, double *phys)
{
    pw_from_multipole(pw_table(iprec), p, mt, phys);
}

extern "C" void mh_pw_scale_b(int iprec, double lam, double mu, double rscale, double *phys)
{
    pw_scale_b(pw_table(iprec), lam, mu, rscale, phys);
}

extern "C" void mh_pw_shift_add(int iprec, const double *src, double xd, double yd, double zd, double *dst)
{
    pw_shift_add(pw_table(iprec), src, xd, yd, zd, dst);
}

extern "C" void mh_pw_to_local_add(int iprec, int p, double side, const double *phys, double *lo)
{
    std::vector<double> four(4096);
    pw_to_local_add(pw_table(iprec), p, side, phys, four.data(), lo);
}

extern "C" void mh_skeleton(int p, const double *mt, const double *c, double rscale, double radius, int nq, double *q,
                            double *xs)
{
    std::vector<double> y((p + 1) * (p + 1));
    skeleton_from_b(p, mt, c, rscale, radius, nq, q, xs, y.data());
}

extern "C" void mh_b_remainder(int p, const double *ms, double *mt) { b_remainder_dz2(p, ms, mt); }

extern "C" void mh_gauss_legendre(int n, double *x, double *w) { gauss_legendre(n, x, w); }

extern "C" void mh_proxy_local_add(int p, int nx, const double *q, const double *xs, const double *ct, double radius,
                                   double rscale, double rscale_src, double lam, double mu, double *lo)
{
    int nquad = 2 * p;
    std::vector<double> x(nquad), w(nquad), work(nquad * nquad + (p + 1) * (p + 1));
    gauss_legendre(nquad, x.data(), w.data());
    proxy_local_add(p, nx, q, xs, ct, radius, rscale, rscale_src, lam, mu, nquad, x.data(), w.data(), work.data(), lo);
}
