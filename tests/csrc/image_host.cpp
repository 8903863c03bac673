// Host build of the library's half-space image kernels and adaptive
// quadrature (elfmm_image.h, the code the image near-field kernels run) for
// CPU tests against the reference's routines.
#include "../../elfmmlib3d_amd/csrc/elfmm_image.h"

static const double kRule6[12][3] = {ELFMM_RULE6};

extern "C" void image_host_point(int layer, double lam, double mu, const double *x, const double *dens,
                                 const double *nv, const double *y, int ifpt, int ifst, double *f12)
{
    if (layer == 0) elfmm::image_sl(lam, mu, x, dens, y, ifpt, ifst, f12);
    else elfmm::image_dl(lam, mu, x, dens, nv, y, ifpt, ifst, f12);
}

extern "C" int image_host_adap(int layer, double lam, double mu, const double *tri, const double *dens,
                               const double *nv, const double *y, int ifpt, int ifst, double *rint12, int *nsub)
{
    elfmm::AdapTask T;
    for (int k = 0; k < 9; k++) T.tri[k] = tri[k];
    for (int k = 0; k < 3; k++) {
        T.dens[k] = dens[k];
        T.nv[k] = nv[k];
        T.y[k] = y[k];
    }
    T.lam = lam;
    T.mu = mu;
    T.layer = layer;
    T.ifpt = ifpt;
    T.ifst = ifst;
    return elfmm::adap_integrate_host(T, kRule6, rint12, nsub);
}
