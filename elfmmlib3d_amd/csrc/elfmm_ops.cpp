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
#include <algorithm>
#include <cstring>
#include <stdexcept>
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
    int nc = ncoef(p), h = nc / 2;
    ops.nc = nc;
    // M2M / L2L per octant, rotated like M2L below: the child-centre offset
    // t = Rz(phi) t' with t' = (rho, 0, tz) in the xz-plane, where the
    // regular harmonics R_n^q(t') are real, so on coefficients rotated by
    // e^{i m phi} the translation is real: real parts to real parts and
    // imaginary parts to imaginary parts, two h x h matrices per octant
    // (real-part matrix, then imaginary-part matrix).
    ops.m2m.assign((size_t)8 * 2 * h * h, 0.0);
    ops.l2l.assign((size_t)8 * 2 * h * h, 0.0);
    std::vector<double> pw(p + 1);
    for (int l = 0; l <= p; l++) pw[l] = std::ldexp(1.0, -l);
    std::vector<double> Dm((size_t)nc * nc), Dl((size_t)nc * nc);
    for (int o = 0; o < 8; o++) {
        // child centre minus parent centre, in units of the parent side
        double t[3] = {(o & 1) ? 0.25 : -0.25, (o & 2) ? 0.25 : -0.25, (o & 4) ? 0.25 : -0.25};
        ops.oct_phi[o] = std::atan2(t[1], t[0]);
        Harm R = regular(p, std::sqrt(t[0] * t[0] + t[1] * t[1]), 0.0, t[2]);
        std::fill(Dm.begin(), Dm.end(), 0.0);
        std::fill(Dl.begin(), Dl.end(), 0.0);
        for (int l = 0; l <= p; l++)
            for (int m = 0; m <= l; m++)
                for (int j = 0; j <= l; j++)
                    for (int k = -j; k <= j; k++) {
                        if (std::abs(m - k) > l - j) continue;
                        double c = std::sqrt(binom(l + m, j + k) * binom(l - m, j - k));
                        // M2M: Mp_l^m += c 2^-j Mc_j^k conj(R_{l-j}^{m-k}(t))
                        add_entry(Dm, 0, nc, l, m, j, k, c * pw[j] * std::conj(R.get(l - j, m - k)));
                        // L2L: Lc_j^k += c 2^-j Lp_l^m R_{l-j}^{m-k}(t)
                        if (k >= 0) add_entry(Dl, 0, nc, j, k, l, m, c * pw[j] * R.get(l - j, m - k));
                    }
        // L2L needs inputs of negative order too
        for (int l = 0; l <= p; l++)
            for (int m = -l; m < 0; m++)
                for (int j = 0; j <= l; j++)
                    for (int k = 0; k <= j; k++) {
                        if (std::abs(m - k) > l - j) continue;
                        double c = std::sqrt(binom(l + m, j + k) * binom(l - m, j - k));
                        add_entry(Dl, 0, nc, j, k, l, m, c * pw[j] * R.get(l - j, m - k));
                    }
        for (int which = 0; which < 2; which++) {
            const std::vector<double> &D = which ? Dl : Dm;
            double *Rm = (which ? ops.l2l : ops.m2m).data() + (size_t)o * 2 * h * h, *Im = Rm + (size_t)h * h;
            double big = 0, cross = 0;
            for (int r = 0; r < h; r++)
                for (int c = 0; c < h; c++) {
                    Rm[(size_t)r * h + c] = D[(size_t)(2 * r) * nc + 2 * c];
                    Im[(size_t)r * h + c] = D[(size_t)(2 * r + 1) * nc + 2 * c + 1];
                    big = std::max(big, std::max(std::abs(Rm[(size_t)r * h + c]), std::abs(Im[(size_t)r * h + c])));
                    cross = std::max(cross, std::max(std::abs(D[(size_t)(2 * r) * nc + 2 * c + 1]),
                                                     std::abs(D[(size_t)(2 * r + 1) * nc + 2 * c])));
                }
            if (cross > 1e-13 * big) throw std::runtime_error("rotated M2M/L2L operator is not real");
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
    // The rotated operator depends on the offset only through rho^2 =
    // i^2 + j^2, k and pt: offsets that share them (up to eight, mirror
    // images in x, y and the diagonal) share one stored copy. The mirror
    // image in z shares it too: the irregular harmonics of the translation
    // satisfy I_n^0(rho, 0, k) = (-1)^n I_n^0(rho, 0, -k) term by term, so
    // R(i, j, -k) = S R(i, j, k) S with S = diag((-1)^(n+m)) on the packed
    // coefficients (m2l_refl; the kernel applies S with the azimuth phase).
    size_t tot = 0;
    std::vector<int> owner(343, -1);
    for (int o = 0; o < 343; o++) {
        int i = o / 49 - 3, j = (o / 7) % 7 - 3, k = o % 7 - 3;
        int far = std::max(std::abs(i), std::max(std::abs(j), std::abs(k))) >= 2;
        int pt = far ? std::min(p, itable[o]) : 0;
        ops.m2l_pt[o] = far ? pt : -1;
        ops.m2l_phi[o] = far ? std::atan2(-(double)j, -(double)i) : 0.0;
        ops.m2l_off[o] = (int64_t)tot;
        ops.m2l_refl[o] = 0;
        if (!far) continue;
        for (int q = 0; q < o; q++) {
            int iq = q / 49 - 3, jq = (q / 7) % 7 - 3, kq = q % 7 - 3;
            if (owner[q] < 0 && ops.m2l_pt[q] == pt && std::abs(kq) == std::abs(k) &&
                iq * iq + jq * jq == i * i + j * j) {
                owner[o] = q;
                ops.m2l_off[o] = ops.m2l_off[q];
                ops.m2l_refl[o] = kq != k ? 1 : 0;
                break;
            }
        }
        if (owner[o] < 0) tot += (size_t)2 * (ncoef(pt) / 2) * (ncoef(pt) / 2);
    }
    ops.m2l.assign(tot, 0.0);
    for (int o = 0; o < 343; o++) {
        int pt = ops.m2l_pt[o];
        if (pt < 0 || owner[o] >= 0) continue;
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

namespace {

// Real linear form over the doubles of a packed expansion.
struct LForm {
    std::vector<std::pair<int, double>> t;
    void add(const LForm &o, double s)
    {
        for (auto &e : o.t) t.push_back({e.first, e.second * s});
    }
};
struct CForm {
    LForm re, im;
};

CForm cscale(const CForm &a, double s)
{
    CForm r;
    r.re.add(a.re, s);
    r.im.add(a.im, s);
    return r;
}

// coefficient (l, m) of a real-field expansion of order pe whose m >= 0
// coefficients are the forms f(l, m): c^{-m} = (-1)^m conj(c^m), zero
// outside 0 <= |m| <= l <= pe
template <class F>
CForm sym_get(int pe, int l, int m, F &&f)
{
    int am = std::abs(m);
    if (l < 0 || l > pe || am > l) return CForm();
    CForm c = f(l, am);
    if (m < 0) {
        double sr = (am & 1) ? -1.0 : 1.0;
        CForm r;
        r.re.add(c.re, sr);
        r.im.add(c.im, -sr);
        return r;
    }
    return c;
}

// gradient (x, y, z) at (n, k) from the coefficients c (order k), a (k-1)
// and b (k+1) at degree n+1 (regular) or n-1 (irregular)
void grad_comb(bool irr, int n, int k, const CForm &c, const CForm &a, const CForm &b, CForm g[3])
{
    auto sq = [](int v) { return std::sqrt((double)v); };
    double fz, fa, fb;
    if (!irr) {
        fz = sq((n + 1 - k) * (n + 1 + k));
        fa = sq((n + 2 - k) * (n + 1 - k));
        fb = -sq((n + 2 + k) * (n + 1 + k));
    } else {
        fz = -sq((n - k) * (n + k));
        fa = sq(std::max(n + k - 1, 0) * (n + k));
        fb = -sq(std::max(n - k - 1, 0) * (n - k));
    }
    g[2] = cscale(c, fz);
    CForm A = cscale(a, fa), B = cscale(b, fb);
    g[0] = CForm();
    g[0].re.add(A.re, 0.5);
    g[0].re.add(B.re, 0.5);
    g[0].im.add(A.im, 0.5);
    g[0].im.add(B.im, 0.5);
    // y = (A - B) / (2i)
    g[1] = CForm();
    g[1].re.add(A.im, 0.5);
    g[1].re.add(B.im, -0.5);
    g[1].im.add(A.re, -0.5);
    g[1].im.add(B.re, 0.5);
}

}  // namespace

void build_field_table(int p, bool irr, std::vector<FieldTerm> &tab)
{
    int ph = irr ? p + 2 : p, p1 = irr ? p + 1 : p - 1, p2 = irr ? p + 2 : p - 2;
    int ncf = (ph + 1) * (ph + 2) / 2;
    int npos = ncf + (ncf & 1);
    tab.assign((size_t)npos * 2 * 9, FieldTerm{{0, 0, 0, 0}, {0.0, 0.0, 0.0, 0.0}});
    std::vector<double> fact(2 * ph + 2, 1.0);
    for (size_t k = 1; k < fact.size(); k++) fact[k] = fact[k - 1] * (double)k;
    auto e0 = [&](int l, int m) {
        CForm c;
        int j = 2 * (l * (l + 1) / 2 + m);
        c.re.t.push_back({j, 1.0});
        c.im.t.push_back({j + 1, 1.0});
        return c;
    };
    auto E = [&](int l, int m) { return sym_get(p, l, m, e0); };
    // gradient of the expansion at (l, m >= 0)
    auto grad = [&](int l, int m, CForm g[3]) {
        int d = irr ? l - 1 : l + 1;
        grad_comb(irr, l, m, E(d, m), E(d, m - 1), E(d, m + 1), g);
    };
    for (int m = 0; m <= ph; m++)
        for (int l = m; l <= ph; l++) {
            int pos = m * (ph + 1) - m * (m - 1) / 2 + (l - m);
            CForm f[9];
            if (l <= p) f[0] = E(l, m);
            if (l <= p1) grad(l, m, &f[1]);
            if (l <= p2) {
                int d = irr ? l - 1 : l + 1;
                CForm g[3][3];     // gradient at (d, m), (d, m-1), (d, m+1)
                for (int u = 0; u < 3; u++) {
                    int mm = u == 0 ? m : (u == 1 ? m - 1 : m + 1);
                    for (int v = 0; v < 3; v++) {
                        g[u][v] = sym_get(p1, d, mm, [&](int ll, int am) {
                            CForm gg[3];
                            grad(ll, am, gg);
                            return gg[v];
                        });
                    }
                }
                CForm hx[3], hy[3];
                grad_comb(irr, l, m, g[0][0], g[1][0], g[2][0], hx);   // of d/dx: xx, xy, xz
                grad_comb(irr, l, m, g[0][1], g[1][1], g[2][1], hy);   // of d/dy: yx, yy, yz
                f[4] = hx[0];
                f[6] = hx[1];
                f[7] = hx[2];
                f[5] = hy[1];
                f[8] = hy[2];
            }
            double w = (m == 0 ? 1.0 : 2.0) / std::sqrt(fact[l + m] * fact[l - m]);
            for (int q = 0; q < 9; q++)
                for (int part = 0; part < 2; part++) {
                    const LForm &lf = part ? f[q].im : f[q].re;
                    double sg = part ? -w : w;
                    // merge equal indices, drop exact zeros
                    std::vector<std::pair<int, double>> mt;
                    for (auto &e : lf.t) {
                        bool found = false;
                        for (auto &x : mt)
                            if (x.first == e.first) {
                                x.second += e.second;
                                found = true;
                            }
                        if (!found) mt.push_back(e);
                    }
                    FieldTerm &ft = tab[((size_t)pos * 2 + part) * 9 + q];
                    int nt = 0;
                    for (auto &x : mt) {
                        if (x.second == 0.0) continue;
                        if (nt == 3) throw std::runtime_error("field table: more than three terms");
                        ft.idx[nt] = x.first;
                        ft.c[nt] = x.second * sg;
                        nt++;
                    }
                }
        }
}

}  // namespace elfmm
