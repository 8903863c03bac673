// Host planner, part 2: dense translation operators.
//
// Expansions use Racah-normalised solid harmonics with the Condon-Shortley
// phase,
//   R_l^m(x) = sqrt((l-m)!/(l+m)!) r^l P_l^m(cos t) e^{i m phi},
//   I_l^m(x) = sqrt((l-m)!/(l+m)!) P_l^m(cos t) e^{i m phi} / r^(l+1),
// so that 1/|x-y| = sum conj(R_l^m(y)) I_l^m(x) for |y|<|x|. Multipole and
// local coefficients of a box of side s are stored scaled,
//   Mt_l^m = M_l^m / s^l,  Lt_l^m = L_l^m * s^l,
// which makes every operator below independent of the level (M2L carries a
// runtime factor 1/s). Real fields need only m >= 0 (c^{-m} =
// (-1)^m conj(c^m)); the packed real layout is documented in
// elfmm_internal.h.
//
// The operators are the exact translations of the truncated expansions the
// reference applies (M2M/L2L at order p, src/l3dmpmpfinal4.f:182,
// src/l3dloclocfinal4.f:155; M2L truncated to the list-2 order table,
// src/lfmm3drouts.f:404-419 with src/l3dmplocfinal4.f:373), built here from
// the addition theorems
//   R_l^m(r+a) = sum sqrt(C(l+m,j+k) C(l-m,j-k)) R_j^k(r) R_{l-j}^{m-k}(a)
//   I_l^m(r+a) = sum (-1)^(j+k) sqrt(C(l+j-m+k,j+k) C(l+j+m-k,j-k))
//                    R_j^k(r) I_{l+j}^{m-k}(a)        (|r| < |a|).
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "elfmm_internal.h"

namespace elfmm {

using cd = std::complex<double>;

namespace {

struct Harm {
    int p;
    std::vector<cd> v;   // (l, m>=0) at l*(l+1)/2+m
    cd get(int l, int m) const
    {
        if (m >= 0) return v[l * (l + 1) / 2 + m];
        cd c = std::conj(v[l * (l + 1) / 2 - m]);
        return (m & 1) ? -c : c;
    }
};

Harm regular(int p, double x, double y, double z)
{
    Harm h{p, std::vector<cd>((p + 1) * (p + 2) / 2)};
    double r2 = x * x + y * y + z * z;
    cd diag = 1.0;
    for (int m = 0; m <= p; m++) {
        if (m > 0) diag = -cd(x, y) * std::sqrt((2.0 * m - 1) / (2.0 * m)) * diag;
        cd a = diag, b = 0.0;
        h.v[m * (m + 1) / 2 + m] = a;
        if (m + 1 <= p) {
            b = std::sqrt(2.0 * m + 1) * z * a;
            h.v[(m + 1) * (m + 2) / 2 + m] = b;
        }
        for (int l = m + 2; l <= p; l++) {
            cd c = ((2.0 * l - 1) * z * b - std::sqrt((l - 1.0 + m) * (l - 1.0 - m)) * r2 * a) /
                   std::sqrt((l + m) * (double)(l - m));
            h.v[l * (l + 1) / 2 + m] = c;
            a = b;
            b = c;
        }
    }
    return h;
}

Harm irregular(int p, double x, double y, double z)
{
    Harm h{p, std::vector<cd>((p + 1) * (p + 2) / 2)};
    double r2 = x * x + y * y + z * z;
    cd diag = 1.0 / std::sqrt(r2);
    for (int m = 0; m <= p; m++) {
        if (m > 0) diag = -cd(x, y) / r2 * std::sqrt((2.0 * m - 1) / (2.0 * m)) * diag;
        cd a = diag, b = 0.0;
        h.v[m * (m + 1) / 2 + m] = a;
        if (m + 1 <= p) {
            b = std::sqrt(2.0 * m + 1) * z / r2 * a;
            h.v[(m + 1) * (m + 2) / 2 + m] = b;
        }
        for (int l = m + 2; l <= p; l++) {
            cd c = ((2.0 * l - 1) * z * b - std::sqrt((l - 1.0 + m) * (l - 1.0 - m)) * a) /
                   (std::sqrt((l + m) * (double)(l - m)) * r2);
            h.v[l * (l + 1) / 2 + m] = c;
            a = b;
            b = c;
        }
    }
    return h;
}

double binom(int n, int k)
{
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

// Add the complex map out(jo,ko) += K * in(l,m) (m of either sign) to the
// real packed matrix A (nrow x ncol).
void add_entry(std::vector<double> &A, size_t base, int ncol, int jo, int ko, int l, int m, cd K)
{
    int am = m < 0 ? -m : m;
    int ro = 2 * (jo * (jo + 1) / 2 + ko);
    int ci = 2 * (l * (l + 1) / 2 + am);
    double *Re = &A[base + (size_t)ro * ncol];
    double *Im = &A[base + (size_t)(ro + 1) * ncol];
    if (m >= 0) {
        // K * (a + i b)
        Re[ci] += K.real();
        Re[ci + 1] += -K.imag();
        Im[ci] += K.imag();
        Im[ci + 1] += K.real();
    } else {
        // K * (-1)^am * conj(a + i b)
        double s = (am & 1) ? -1.0 : 1.0;
        Re[ci] += s * K.real();
        Re[ci + 1] += s * K.imag();
        Im[ci] += s * K.imag();
        Im[ci + 1] += -s * K.real();
    }
}

}  // namespace

void build_operators(int p, const int *itable, Operators &ops)
{
    ops.p = p;
    int nc = ncoef(p);
    ops.nc = nc;
    ops.m2m.assign((size_t)8 * nc * nc, 0.0);
    ops.l2l.assign((size_t)8 * nc * nc, 0.0);
    std::vector<double> pw(p + 1);
    for (int l = 0; l <= p; l++) pw[l] = std::ldexp(1.0, -l);
    for (int o = 0; o < 8; o++) {
        // child centre minus parent centre, in units of the parent side
        double t[3] = {(o & 1) ? 0.25 : -0.25, (o & 2) ? 0.25 : -0.25, (o & 4) ? 0.25 : -0.25};
        Harm R = regular(p, t[0], t[1], t[2]);
        size_t base = (size_t)o * nc * nc;
        for (int l = 0; l <= p; l++)
            for (int m = 0; m <= l; m++)
                for (int j = 0; j <= l; j++)
                    for (int k = -j; k <= j; k++) {
                        if (std::abs(m - k) > l - j) continue;
                        double c = std::sqrt(binom(l + m, j + k) * binom(l - m, j - k));
                        // M2M: Mp_l^m += c 2^-j Mc_j^k conj(R_{l-j}^{m-k}(t))
                        add_entry(ops.m2m, base, nc, l, m, j, k,
                                  c * pw[j] * std::conj(R.get(l - j, m - k)));
                        // L2L: Lc_j^k += c 2^-j Lp_l^m R_{l-j}^{m-k}(t)
                        if (k >= 0)
                            add_entry(ops.l2l, base, nc, j, k, l, m, c * pw[j] * R.get(l - j, m - k));
                    }
        // L2L needs inputs of negative order too
        for (int l = 0; l <= p; l++)
            for (int m = -l; m < 0; m++)
                for (int j = 0; j <= l; j++)
                    for (int k = 0; k <= j; k++) {
                        if (std::abs(m - k) > l - j) continue;
                        double c = std::sqrt(binom(l + m, j + k) * binom(l - m, j - k));
                        add_entry(ops.l2l, base, nc, j, k, l, m, c * pw[j] * R.get(l - j, m - k));
                    }
    }
    // M2L per offset; offsets are source ijk minus target ijk, so the
    // translation vector (target centre minus source centre) is -offset.
    // With t = Rz(phi) t', t' = (rho, 0, tz) in the xz-plane,
    // I_n^q(t) = e^{i q phi} I_n^q(t') and I_n^q(t') is real, so
    //   L_j^k = e^{-i k phi} sum_{l,m} a_{jklm}(t') (e^{i m phi} M_l^m)
    // with real a: on the rotated coefficients the real parts map to real
    // parts and the imaginary parts to imaginary parts, each by a real
    // ncoef(pt)/2 square matrix (rows j,k; columns l,m >= 0). Stored per
    // offset: the real-part matrix, then the imaginary-part matrix.
    size_t tot = 0;
    for (int o = 0; o < 343; o++) {
        int i = o / 49 - 3, j = (o / 7) % 7 - 3, k = o % 7 - 3;
        int far = std::max(std::abs(i), std::max(std::abs(j), std::abs(k))) >= 2;
        int pt = far ? std::min(p, itable[o]) : 0;
        ops.m2l_pt[o] = far ? pt : -1;
        ops.m2l_off[o] = (int64_t)tot;
        ops.m2l_phi[o] = far ? std::atan2(-(double)j, -(double)i) : 0.0;
        if (far) tot += (size_t)2 * (ncoef(pt) / 2) * (ncoef(pt) / 2);
    }
    ops.m2l.assign(tot, 0.0);
    for (int o = 0; o < 343; o++) {
        int pt = ops.m2l_pt[o];
        if (pt < 0) continue;
        int i = o / 49 - 3, j = (o / 7) % 7 - 3, k = o % 7 - 3;
        int n = ncoef(pt), h = n / 2;
        double rho = std::sqrt((double)(i * i + j * j));
        Harm I = irregular(2 * pt, rho, 0.0, -(double)k);
        std::vector<double> D((size_t)n * n, 0.0);
        for (int jo = 0; jo <= pt; jo++)
            for (int ko = 0; ko <= jo; ko++)
                for (int l = 0; l <= pt; l++)
                    for (int m = -l; m <= l; m++) {
                        double c = std::sqrt(binom(l + jo - m + ko, jo + ko) *
                                             binom(l + jo + m - ko, jo - ko));
                        double s = ((jo + ko) & 1) ? -1.0 : 1.0;
                        add_entry(D, 0, n, jo, ko, l, m, s * c * I.get(l + jo, m - ko));
                    }
        double *Rm = &ops.m2l[(size_t)ops.m2l_off[o]], *Im = Rm + (size_t)h * h;
        for (int r = 0; r < h; r++)
            for (int c = 0; c < h; c++) {
                Rm[(size_t)r * h + c] = D[(size_t)(2 * r) * n + 2 * c];
                Im[(size_t)r * h + c] = D[(size_t)(2 * r + 1) * n + 2 * c + 1];
            }
    }
}

}  // namespace elfmm
