// Mindlin harmonic-image far field: the operations of the reference's
// lfmm3dmindlintriatarg (src/lfmm3dtriamindlin_iter_1.f:41) that are not
// plain Laplace translations, on expansions in this library's packed layout.
// Compiled for the gfx950 kernels and, for CPU tests, for the host.
//
// Expansions: packed complex coefficients (l, 0 <= m <= l) at doubles
// 2*(l(l+1)/2 + m), of a box of side s. With the reference's scaling
// parameter rscale = 1/s (src/lfmm3dtriamindlin_iter_1.f:268-271) its
// multipole and local coefficients are exactly these (Racah-normalised
// solid harmonics, multipole / s^l, local * s^l), so the reference's
// coefficient formulas below apply to them unchanged.
//
// The three expansions of a box: B (the harmonic part U = R3 log(R+R3) - R
// of the image source, carried as its second z-derivative, a Laplace
// multipole), C1 and C2 (the 1/R part and its z-weighted companion).
// Far list-2 pairs (|dz| >= 1.5 box sides) translate through plane waves
// (src/lfmm3droutsmindlin.f:329-345): multipole -> Fourier modes of the
// exponential representation (mpoletoexp, src/mtxdown.f:2), -> physical
// samples (ftophys, src/ftophys.f:2), B divided by lambda^2 (its two
// z-antiderivatives, src/mindlinbformmp.f:318-386), shifted (xyzshift_add,
// src/xyzshift.f:75), summed per target box, back to Fourier modes (phystof,
// src/phystof.f:2) and to local coefficients (exptolocald_add,
// src/xtmdown.f:130). Near pairs take B through proxy charges on two rings
// (mptoslpborder, src/mptoslpborder.f:24) that carry its sectoral terms,
// the rest of B integrated twice in z (mpdzminus2, src/mindlinmpeval.f:284)
// and translated exactly, plus the rings' U-field projected on the target
// sphere (l3dformtaintkerbf90sc, src/mindlinformta.f:980; projloc3d,
// src/projloc.f:21).
#pragma once

#include "elfmm_math.h"
#include "elfmm_pwtab.h"

namespace elfmm {

ELFMM_HD int pk(int l, int m) { return l * (l + 1) / 2 + m; }

// Plane-wave table for a precision flag (iprec < 0 uses the iprec-0 table,
// src/lfmm3dtriamindlin_iter_1.f:401-402).
ELFMM_HD const PwTable &pw_table(int iprec)
{
    int k = iprec < 0 ? 0 : (iprec > 3 ? 3 : iprec);
    return kPwTables[k];
}
// numfour / numphys of the reference: the table counts + 2
ELFMM_HD int pw_nfour(const PwTable &T, int i) { return T.numhalf[i] + 2; }
ELFMM_HD int pw_nphys(const PwTable &T, int i) { return T.numfour[i] + 2; }
ELFMM_HD int pw_total_phys(const PwTable &T)
{
    int s = 0;
    for (int i = 0; i < T.nlams; i++) s += pw_nphys(T, i);
    return s;
}

// sqrt(k!) for k <= 2p (the reference's facts table)
ELFMM_HD double sqrt_fact(int k)
{
    double f = 1.0;
    for (int j = 1; j <= k; j++) f *= sqrt((double)j);
    return f;
}

// Multipole mt (order p) -> physical plane-wave samples phys (complex,
// pw_total_phys entries), the reference's mpoletoexp + ftophys.
ELFMM_HD void pw_from_multipole(const PwTable &T, int p, const double *mt, double *phys)
{
    int nphys_off = 0;
    for (int nl = 0; nl < T.nlams; nl++) {
        double lam = T.rlams[nl];
        int nf = pw_nfour(T, nl), np = pw_nphys(T, nl);
        // Fourier modes f_mth = (-i)^mth sum_n (-lam)^n / sqrt((n-m)!(n+m)!) M_n^mth
        double fre[80], fim[80];
        double zr = 1.0, zi = 0.0;                    // (-i)^mth
        for (int mth = 0; mth < nf; mth++) {
            double s1r = 0, s1i = 0, s2r = 0, s2i = 0;
            for (int n = mth; n <= p; n++) {
                double c = pow(-lam, (double)n) / (sqrt_fact(n - mth) * sqrt_fact(n + mth));
                const double *M = mt + 2 * pk(n, mth);
                if (((n - mth) & 1) == 0) {
                    s1r += c * M[0];
                    s1i += c * M[1];
                } else {
                    s2r += c * M[0];
                    s2i += c * M[1];
                }
            }
            double sr = s1r + s2r, si = s1i + s2i;
            fre[mth] = sr * zr - si * zi;
            fim[mth] = sr * zi + si * zr;
            double t = zi;                            // multiply by -i
            zi = -zr;
            zr = t;
        }
        for (int j = 0; j < np; j++) {
            double alpha = j * (2 * M_PI / np);
            double pr = fre[0], pi = fim[0];
            double sgn = -1.0;
            for (int mm = 1; mm < nf; mm++) {
                double cr = cos(mm * alpha), ci = sin(mm * alpha);
                double tr = cr * fre[mm] - ci * fim[mm], ti = cr * fim[mm] + ci * fre[mm];
                pr += tr + sgn * tr;
                pi += ti - sgn * ti;
                sgn = -sgn;
            }
            phys[2 * (nphys_off + j)] = pr;
            phys[2 * (nphys_off + j) + 1] = pi;
        }
        nphys_off += np;
    }
}

// B's plane waves: the two z-antiderivatives and the B scale,
// (1 - alpha) / alpha / rscale / lambda^2 (mindlinbpwdiagall).
ELFMM_HD void pw_scale_b(const PwTable &T, double lam, double mu, double rscale, double *phys)
{
    double al = (lam + mu) / (lam + 2 * mu);
    double sc = (1.0 - al) / al / rscale;
    int ii = 0;
    for (int nl = 0; nl < T.nlams; nl++) {
        double z = 1.0 / (T.rlams[nl] * T.rlams[nl]);
        for (int j = 0; j < pw_nphys(T, nl); j++, ii++) {
            phys[2 * ii] = phys[2 * ii] * z * sc;
            phys[2 * ii + 1] = phys[2 * ii + 1] * z * sc;
        }
    }
}

// dst += src shifted by (xd, yd, zd) box sides (xyzshift_add).
ELFMM_HD void pw_shift_add(const PwTable &T, const double *src, double xd, double yd, double zd, double *dst)
{
    int ii = 0;
    for (int nl = 0; nl < T.nlams; nl++) {
        double lam = T.rlams[nl];
        int np = pw_nphys(T, nl);
        double r0 = exp(lam * zd);
        for (int j = 0; j < np; j++, ii++) {
            double u = j * (2 * M_PI / np);
            double ph = lam * (xd * cos(u) + yd * sin(u));
            double mr = r0 * cos(ph), mi = r0 * sin(ph);
            double a = src[2 * ii], b = src[2 * ii + 1];
            dst[2 * ii] += a * mr - b * mi;
            dst[2 * ii + 1] += a * mi + b * mr;
        }
    }
}

// Local coefficients (order p, box of side s) from physical samples,
// added to lo (phystof + exptolocald_add with boxsize * scale = 1).
// four: scratch of 2 * sum(numfour) doubles.
ELFMM_HD void pw_to_local_add(const PwTable &T, int p, double side, const double *phys, double *four, double *lo)
{
    // Fourier modes per lambda: f_m = (1/np) sum_j e^{-i m alpha_j} phys_j
    int off = 0, foff = 0;
    for (int nl = 0; nl < T.nlams; nl++) {
        int nf = pw_nfour(T, nl), np = pw_nphys(T, nl);
        for (int m = 0; m < nf; m++) {
            double fr = 0, fi = 0;
            for (int j = 0; j < np; j++) {
                double a = phys[2 * (off + j)], b = phys[2 * (off + j) + 1];
                if (m == 0) {
                    fr += a;
                    fi += b;
                } else {
                    double al = j * (2 * M_PI / np);
                    double cr = cos(m * al), ci = -sin(m * al);
                    fr += cr * a - ci * b;
                    fi += cr * b + ci * a;
                }
            }
            four[2 * (foff + m)] = fr / np;
            four[2 * (foff + m) + 1] = fi / np;
        }
        off += np;
        foff += nf;
    }
    for (int n = 0; n <= p; n++)
        for (int m = 0; m <= n; m++) {
            double ar = 0, ai = 0;
            foff = 0;
            for (int nl = 0; nl < T.nlams; nl++) {
                int nf = pw_nfour(T, nl);
                if (m <= nf - 1) {
                    double rl = T.whts[nl] * pow(T.rlams[nl], (double)n);
                    ar += rl * four[2 * (foff + m)];
                    ai += rl * four[2 * (foff + m) + 1];
                }
                foff += nf;
            }
            // times (-i)^m / sqrt((n-m)!(n+m)!) / side
            double zr, zi;
            switch (m & 3) {
            case 0: zr = 1; zi = 0; break;
            case 1: zr = 0; zi = -1; break;
            case 2: zr = -1; zi = 0; break;
            default: zr = 0; zi = 1; break;
            }
            double sc = 1.0 / (sqrt_fact(n - m) * sqrt_fact(n + m)) / side;
            lo[2 * pk(n, m)] += (ar * zr - ai * zi) * sc;
            lo[2 * pk(n, m) + 1] += (ar * zi + ai * zr) * sc;
        }
}

// Normalised Legendre functions of the reference (ylgndr,
// src/yrecursion.f:222): y[n*(p+1)+m] = sqrt(2n+1) sqrt((n-m)!/(n+m)!) P_n^m(x)
// with the Condon-Shortley phase.
ELFMM_HD void ylg(int p, double x, double *y)
{
    double u = -sqrt((1 - x) * (1 + x));
    y[0] = 1;
    for (int m = 0; m <= p; m++) {
        if (m > 0) y[m * (p + 1) + m] = y[(m - 1) * (p + 1) + m - 1] * u * sqrt((2 * m - 1.0) / (2 * m));
        if (m < p) y[(m + 1) * (p + 1) + m] = x * y[m * (p + 1) + m] * sqrt(2 * m + 1.0);
        for (int n = m + 2; n <= p; n++)
            y[n * (p + 1) + m] = ((2 * n - 1) * x * y[(n - 1) * (p + 1) + m] -
                                  sqrt((n + m - 1.0) * (n - m - 1.0)) * y[(n - 2) * (p + 1) + m]) /
                                 sqrt((n - m + 0.0) * (n + m));
    }
    for (int n = 0; n <= p; n++)
        for (int m = 0; m <= n; m++) y[n * (p + 1) + m] *= sqrt(2 * n + 1.0);
}

// Proxy charges of B's sectoral terms on two rings of radius r sin(theta),
// cos(theta) = +-1/sqrt(3), about the box centre (mptoslpborder): charge
// values q[k] (real: the rings carry a real field) and positions xs[3k],
// k < 2*nq; the first nq on the upper ring. ywork: (p+1)^2 doubles.
ELFMM_HD void skeleton_from_b(int p, const double *mt, const double *c, double rscale, double radius, int nq,
                              double *q, double *xs, double *ywork)
{
    const double ct = 1.0 / sqrt(3.0), st = sqrt(2.0 / 3.0);
    ylg(p, ct, ywork);
    double rfac = 1.0 / (radius * rscale), rmul = 1.0;
    // chat1/chat2 for k >= 0 (k < 0: conjugates)
    double c1r[64], c1i[64], c2r[64], c2i[64];
    for (int j = 0; j <= p; j++) {
        const double *Mjj = mt + 2 * pk(j, j);
        double d = ywork[j * (p + 1) + j] * st;
        double f1r = Mjj[0] / d * rmul, f1i = Mjj[1] / d * rmul;
        rmul *= rfac;
        f1r *= sqrt(2 * j + 1.0);
        f1i *= sqrt(2 * j + 1.0);
        double f2r = 0, f2i = 0;
        if (j + 1 <= p) {
            const double *M = mt + 2 * pk(j + 1, j);
            double d2 = ywork[(j + 1) * (p + 1) + j] * st;
            f2r = M[0] / d2 * rmul * sqrt(2 * (j + 1) + 1.0);
            f2i = M[1] / d2 * rmul * sqrt(2 * (j + 1) + 1.0);
        }
        c1r[j] = (f1r + f2r) / 2;
        c1i[j] = (f1i + f2i) / 2;
        c2r[j] = (f1r - f2r) / 2;
        c2i[j] = (f1i - f2i) / 2;
    }
    double h = 2 * M_PI / nq;
    for (int j = 0; j < nq; j++) {
        double phi = j * h;
        // sum_k chat(k) e^{ik phi}: chat(0) + 2 Re sum_{k>0} chat(k) e^{ik phi}
        double s1 = c1r[0], s2 = c2r[0];
        for (int k = 1; k <= p; k++) {
            double cr = cos(k * phi), si = sin(k * phi);
            s1 += 2 * (c1r[k] * cr - c1i[k] * si);
            s2 += 2 * (c2r[k] * cr - c2i[k] * si);
        }
        q[j] = s1 * st / nq;
        q[j + nq] = s2 * st / nq;
        double x = c[0] + radius * st * cos(phi), y = c[1] + radius * st * sin(phi);
        xs[3 * j] = x;
        xs[3 * j + 1] = y;
        xs[3 * j + 2] = c[2] + radius * ct;
        xs[3 * (j + nq)] = x;
        xs[3 * (j + nq) + 1] = y;
        xs[3 * (j + nq) + 2] = c[2] - radius * ct;
    }
}

// mt = mpdzminus2(mt - ms): the remainder of B after the proxy charges'
// multipole ms, integrated twice in z (rows p-1, p become zero).
ELFMM_HD void b_remainder_dz2(int p, const double *ms, double *mt)
{
    for (int k = 0; k < (p + 1) * (p + 2); k++) mt[k] = mt[k] - ms[k];
    for (int m = 0; m <= p - 2; m++)
        for (int i = m + 2; i <= p; i++) {
            double rs = 1.0 / sqrt((i - m) * (i - m - 1.0) * (i + m) * (i + m - 1.0));
            mt[2 * pk(i - 2, m)] = mt[2 * pk(i, m)] * rs;
            mt[2 * pk(i - 2, m) + 1] = mt[2 * pk(i, m) + 1] * rs;
        }
    for (int i = p - 1; i <= p; i++)
        for (int m = 0; m <= i; m++) mt[2 * pk(i, m)] = mt[2 * pk(i, m) + 1] = 0.0;
}

// Gauss-Legendre nodes and weights on [-1, 1] (legewhts, src/legewhts.f:33:
// Newton from cos((2i-1) pi / 2n), descending order).
ELFMM_HD void gauss_legendre(int n, double *x, double *w)
{
    double h = M_PI / (2 * n);
    for (int i = 1; i <= n; i++) x[n - i] = cos((2 * i - 1) * h);
    if (n & 1) x[n / 2] = 0.0;
    auto pol = [&](double t, double &pn, double &der, double &sum) {
        double pkm1 = 1, pk1 = t;
        sum = 0.5 + 1.5 * t * t;   // sum_k (k + 1/2) P_k^2 up to k = n
        for (int k = 1; k < n; k++) {
            double pkp1 = ((2 * k + 1) * t * pk1 - k * pkm1) / (k + 1);
            sum += (k + 1.5) * pkp1 * pkp1;
            pkm1 = pk1;
            pk1 = pkp1;
        }
        pn = pk1;
        der = n * (t * pk1 - pkm1) / (t * t - 1);
    };
    for (int i = 0; i < n / 2; i++) {
        double t = x[i];
        int done = 0;
        for (int it = 0; it < 10; it++) {
            double pn, der, s;
            pol(t, pn, der, s);
            double d = -pn / der;
            t += d;
            if (fabs(d) < 1e-14) done++;
            if (done == 3) break;
        }
        x[i] = t;
        x[n - 1 - i] = -t;
    }
    for (int i = 0; i < (n + 1) / 2; i++) {
        double pn, der, s;
        pol(x[i], pn, der, s);
        w[i] = 1.0 / s;
        w[n - 1 - i] = w[i];
    }
}

// Local expansion (order p, about ct, box scale rscale) of the U-field of
// proxy charges (q, xs; nx of them) of a source box of scale rscale_src,
// added to lo: the field (1 - alpha)/alpha * rscale_src *
// sum q (-r3 log(|r| - r3) - |r|), r = x - xs, tabulated at nquad Gauss-
// Legendre latitudes x nquad longitudes 2 pi k / nquad (k = 1..nquad) on
// the sphere of radius `radius` about ct, projected on the normalised
// harmonics and rescaled (l3dformtaintkerbf90sc with projloc3d and
// l3drescale2). glx, glw: the nquad Gauss-Legendre nodes and weights;
// work: nquad * nquad + (p+1)^2 doubles.
ELFMM_HD void proxy_local_add(int p, int nx, const double *q, const double *xs, const double *ct, double radius,
                              double rscale, double rscale_src, double lam, double mu, int nquad, const double *glx,
                              const double *glw, double *work, double *lo)
{
    double al = (lam + mu) / (lam + 2 * mu);
    double sc = (1.0 - al) / al * rscale_src;
    double *phi = work, *y = work + nquad * nquad;
    for (int jj = 0; jj < nquad; jj++) {
        double cth = glx[jj], sth = sqrt(1.0 - cth * cth);
        for (int kk = 0; kk < nquad; kk++) {
            double ph = 2 * M_PI * (kk + 1) / nquad;
            double tx = ct[0] + radius * sth * cos(ph), ty = ct[1] + radius * sth * sin(ph), tz = ct[2] + radius * cth;
            double v = 0.0;
            for (int i = 0; i < nx; i++) {
                double r1 = tx - xs[3 * i], r2 = ty - xs[3 * i + 1], r3 = tz - xs[3 * i + 2];
                double rr = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
                v += q[i] * (-r3 * log(rr - r3) - rr);
            }
            phi[jj * nquad + kk] = v * sc;
        }
    }
    // project: local(l, m) = sum_jj w_jj Y_l^m(x_jj) (1/(2 nquad)) sum_kk phi e^{-i m phi_kk}
    double rs = 1.0 / (rscale * radius);
    for (int jj = 0; jj < nquad; jj++) {
        ylg(p, glx[jj], y);
        for (int m = 0; m <= p; m++) {
            double sr = 0, si = 0;
            for (int kk = 0; kk < nquad; kk++) {
                double ph = 2 * M_PI * (kk + 1) / nquad;
                sr += phi[jj * nquad + kk] * cos(m * ph);
                si -= phi[jj * nquad + kk] * sin(m * ph);
            }
            double zr = sr / (2 * nquad) * glw[jj], zi = si / (2 * nquad) * glw[jj];
            double rmul = pow(rs, (double)m);
            for (int l = m; l <= p; l++) {
                double f = y[l * (p + 1) + m] * sqrt(2 * l + 1.0) * rmul;
                lo[2 * pk(l, m)] += zr * f;
                lo[2 * pk(l, m) + 1] += zi * f;
                rmul *= rs;
            }
        }
    }
}

}  // namespace elfmm
