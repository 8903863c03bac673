// Half-space (Mindlin) image kernels and the adaptive triangle quadrature of
// the image near field; compiled for the gfx950 kernels and, for CPU tests,
// for the host.
//
// The half-space Green's function is the free-space Kelvin solution plus an
// image part; the image part is what these kernels evaluate, for a point
// source x on a triangle (z < 0) and a target y (z < 0), as the reference's
// green3eluh_image_eval (single layer, src/el3greenh.f:643-777) and
// green3elth_image_eval (double layer, src/el3greenh.f:118-263):
//   * the anti-image: the Kelvin-type field of x seen from the mirrored
//     target y* = (y1, y2, -y3), with every y3-derivative sign-flipped
//     (src/el3green.f:749-877, :879-1092; single layer with (lam+4mu, -mu)
//     and the sign of the result flipped);
//   * the harmonic parts B and C of the image source x* = (x1, x2, -x3)
//     (src/mindlinpieces.f:29-367 and :368-769), built here from the
//     derivatives of U = R3 log(|R| + R3) - |R| (B) and V = 1/|R| (C) with
//     respect to R = (y1 - x1, y2 - x2, -(y3 + x3)), up to fourth order, in
//     closed form: a derivative with two or more R3 indices is a derivative
//     of V, one with a single R3 index a derivative of log(|R| + R3), one
//     with none a derivative of -R_a / (|R| + R3).
// Outputs are f[12]: displacement (3) then the symmetric strain (3x3, row
// major), packed like the reference's integrand (src/elhfmm3dtria_iter.f:
// 2901-3027): displacement first if requested, strain next if requested,
// unused entries zero.
#pragma once

#include "elfmm_math.h"

namespace elfmm {

// Scalars of one image evaluation point: R, |R|, t = |R| + R3 and the
// inverse powers the closed forms use.
struct ImgR {
    double R[3];
    double r, t;
    double i1, i3, i5, i7, i9;        // 1/r^k
    double e1, e2, e3;                // log(r + R3) series (one R3 index)
    double f1, f2, f3;                // -R_a/(r + R3) series (no R3 index)
};

ELFMM_HD void img_setup(double R0, double R1, double R2, ImgR &g)
{
    g.R[0] = R0;
    g.R[1] = R1;
    g.R[2] = R2;
    g.r = sqrt(R0 * R0 + R1 * R1 + R2 * R2);
    g.t = g.r + R2;
    double i1 = 1.0 / g.r, i2 = i1 * i1;
    g.i1 = i1;
    g.i3 = i1 * i2;
    g.i5 = g.i3 * i2;
    g.i7 = g.i5 * i2;
    g.i9 = g.i7 * i2;
    double it = 1.0 / g.t, it2 = it * it;
    g.e1 = i1 * it;
    g.e2 = g.i3 * it + i2 * it2;
    g.e3 = 3 * g.i5 * it + 3 * i2 * i2 * it2 + 2 * g.i3 * it2 * it;
    g.f1 = i1 * it2;
    g.f2 = g.i3 * it2 + 2 * i2 * it2 * it;
    g.f3 = 3 * g.i5 * it2 + 6 * i2 * i2 * it2 * it + 6 * g.i3 * it2 * it2;
}

ELFMM_HD double dlt(int a, int b) { return a == b ? 1.0 : 0.0; }

// derivatives of V = 1/|R|
ELFMM_HD double V1(const ImgR &g, int a) { return -g.R[a] * g.i3; }
ELFMM_HD double V2(const ImgR &g, int a, int b) { return 3 * g.R[a] * g.R[b] * g.i5 - dlt(a, b) * g.i3; }
ELFMM_HD double V3(const ImgR &g, int a, int b, int c)
{
    const double *R = g.R;
    return -15 * R[a] * R[b] * R[c] * g.i7 + 3 * (dlt(a, b) * R[c] + dlt(a, c) * R[b] + dlt(b, c) * R[a]) * g.i5;
}
ELFMM_HD double V4(const ImgR &g, int a, int b, int c, int d)
{
    const double *R = g.R;
    return 105 * R[a] * R[b] * R[c] * R[d] * g.i9 -
           15 *
               (dlt(a, b) * R[c] * R[d] + dlt(a, c) * R[b] * R[d] + dlt(a, d) * R[b] * R[c] +
                dlt(b, c) * R[a] * R[d] + dlt(b, d) * R[a] * R[c] + dlt(c, d) * R[a] * R[b]) *
               g.i7 +
           3 * (dlt(a, b) * dlt(c, d) + dlt(a, c) * dlt(b, d) + dlt(a, d) * dlt(b, c)) * g.i5;
}

// derivatives of U = R3 log(|R| + R3) - |R| (index 2 is R3)
ELFMM_HD double U2(const ImgR &g, int a, int b)
{
    const double *R = g.R;
    int n3 = (a == 2) + (b == 2);
    if (n3 == 2) return g.i1;
    if (n3 == 1) return R[a == 2 ? b : a] * g.e1;
    return -dlt(a, b) / g.t + R[a] * R[b] * g.f1;
}
ELFMM_HD double U3(const ImgR &g, int a, int b, int c)
{
    const double *R = g.R;
    int n3 = (a == 2) + (b == 2) + (c == 2);
    if (n3 == 3) return V1(g, 2);
    if (n3 == 2) return V1(g, a != 2 ? a : (b != 2 ? b : c));
    if (n3 == 1) {
        int p = a == 2 ? b : a, q = (a == 2 || b == 2) ? c : b;
        return dlt(p, q) * g.e1 - R[p] * R[q] * g.e2;
    }
    return (dlt(a, b) * R[c] + dlt(a, c) * R[b] + dlt(b, c) * R[a]) * g.f1 - R[a] * R[b] * R[c] * g.f2;
}
ELFMM_HD double U4(const ImgR &g, int a, int b, int c, int d)
{
    const double *R = g.R;
    int idx[4] = {a, b, c, d};
    int n3 = (a == 2) + (b == 2) + (c == 2) + (d == 2);
    if (n3 >= 2) {
        // drop two R3 indices, a second derivative of V remains
        int rest[2] = {2, 2}, m = 0, dropped = 0;
        for (int k = 0; k < 4; k++) {
            if (idx[k] == 2 && dropped < 2) dropped++;
            else rest[m++] = idx[k];
        }
        return V2(g, rest[0], rest[1]);
    }
    if (n3 == 1) {
        int q[3], m = 0;
        for (int k = 0; k < 4; k++)
            if (idx[k] != 2) q[m++] = idx[k];
        return -(dlt(q[0], q[1]) * R[q[2]] + dlt(q[0], q[2]) * R[q[1]] + dlt(q[1], q[2]) * R[q[0]]) * g.e2 +
               R[q[0]] * R[q[1]] * R[q[2]] * g.e3;
    }
    return (dlt(a, b) * dlt(c, d) + dlt(a, c) * dlt(b, d) + dlt(a, d) * dlt(b, c)) * g.f1 -
           (dlt(a, b) * R[c] * R[d] + dlt(a, c) * R[b] * R[d] + dlt(a, d) * R[b] * R[c] + dlt(b, c) * R[a] * R[d] +
            dlt(b, d) * R[a] * R[c] + dlt(c, d) * R[a] * R[b]) *
               g.f2 +
           R[a] * R[b] * R[c] * R[d] * g.f3;
}

// sign of d/dx_l against d/dR_l (R3 = -(y3 + x3)), and of the R3 index in
// the reference's image formulas
ELFMM_HD double sgn3(int a) { return a == 2 ? -1.0 : 1.0; }

// Pack displacement u and gradient G (G[i][l] = du_i/dy_l) into f[12].
ELFMM_HD void img_pack(const double u[3], double G[3][3], int ifpt, int ifst, double *f)
{
    for (int k = 0; k < 12; k++) f[k] = 0.0;
    int k = 0;
    if (ifpt)
        for (int i = 0; i < 3; i++) f[k++] = u[i];
    if (ifst)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) f[k++] = (G[i][j] + G[j][i]) / 2;
}

// Image field of a point force sg at x (single layer), at y.
ELFMM_HD void image_sl(double lam, double mu, const double *x, const double *sg, const double *y, int ifpt,
                       int ifst, double *f)
{
    double u[3] = {0, 0, 0}, G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    bool grad = ifst != 0;
    // anti-image: Kelvin field with (lam + 4 mu, -mu) at the mirrored target,
    // result sign-flipped; the 1/(2 mu') scaling with the flip is 1/(2 mu)
    {
        double d[3] = {y[0] - x[0], y[1] - x[1], -y[2] - x[2]};
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double ir = 1.0 / r, ir3 = ir * ir * ir, ir5 = ir3 * ir * ir;
        double cb = (lam + 3 * mu) / (lam + 2 * mu), ca = 2 - cb;
        double fd = sg[0] * d[0] + sg[1] * d[1] + sg[2] * d[2];
        double s = 1.0 / (2 * mu);
#pragma unroll
        for (int i = 0; i < 3; i++) u[i] += (ca * sg[i] * ir + cb * fd * d[i] * ir3) * s;
        if (grad)
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int l = 0; l < 3; l++)
                    G[i][l] += sgn3(l) * s *
                               (-ca * sg[i] * d[l] * ir3 +
                                cb * (sg[l] * d[i] * ir3 + fd * (dlt(i, l) * ir3 - 3 * d[i] * d[l] * ir5)));
    }
    // harmonic parts of the image source x* = (x1, x2, -x3)
    ImgR g;
    img_setup(y[0] - x[0], y[1] - x[1], -(y[2] + x[2]), g);
    double al = (lam + mu) / (lam + 2 * mu);
    double sb = (1 - al) / al;          // B-part scale
    double s2 = -(2 - al);              // C-part scale
    double xi3 = -x[2];                 // image source height
    double imu = 1.0 / mu;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        // B: pot = sb (-s1 U_i1 - s2 U_i2 + s3 U_i3)
        double pot = sb * (-sg[0] * U2(g, i, 0) - sg[1] * U2(g, i, 1) + sg[2] * U2(g, i, 2));
        u[i] += pot * imu;
        if (grad)
#pragma unroll
            for (int l = 0; l < 3; l++) {
                double fld = sb * sgn3(l) * (sg[0] * U3(g, i, 0, l) + sg[1] * U3(g, i, 1, l) - sg[2] * U3(g, i, 2, l));
                G[i][l] -= fld * imu;
            }
        // C: multiplied by y3
        double potc;
        if (i < 2)
            potc = s2 * sg[2] * V1(g, i) + xi3 * al * (sg[0] * V2(g, i, 0) + sg[1] * V2(g, i, 1) + sg[2] * V2(g, i, 2));
        else
            potc = s2 * (V1(g, 0) * sg[0] + V1(g, 1) * sg[1]) -
                   xi3 * al * (sg[0] * V2(g, 2, 0) + sg[1] * V2(g, 2, 1) + sg[2] * V2(g, 2, 2));
        u[i] += y[2] * potc * imu;
        if (grad)
#pragma unroll
            for (int l = 0; l < 3; l++) {
                double fld = 0;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    double A = -al * xi3 * sgn3(i) * V3(g, i, j, l);
                    if (i != 2 && j == 2) A -= s2 * V2(g, i, l);
                    if (i == 2 && j != 2) A -= s2 * V2(g, j, l);
                    fld += A * sg[j];
                }
                fld *= sgn3(l);
                G[i][l] += -y[2] * fld * imu + (l == 2 ? potc * imu : 0.0);
            }
    }
    img_pack(u, G, ifpt, ifst, f);
}

// Image field of a displacement jump b across a surface element with unit
// normal nv at x (double layer), at y.
ELFMM_HD void image_dl(double lam, double mu, const double *x, const double *b, const double *nv, const double *y,
                       int ifpt, int ifst, double *f)
{
    double u[3] = {0, 0, 0}, G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    bool grad = ifst != 0;
    double al = (lam + mu) / (lam + 2 * mu);
    double bn = b[0] * nv[0] + b[1] * nv[1] + b[2] * nv[2];
    // anti-image
    {
        double d[3] = {y[0] - x[0], y[1] - x[1], -y[2] - x[2]};
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double ir = 1.0 / r, ir3 = ir * ir * ir, ir5 = ir3 * ir * ir, ir7 = ir5 * ir * ir;
        double c1 = al - 1, c2 = 2 - al;
        double dn = d[0] * nv[0] + d[1] * nv[1] + d[2] * nv[2];
        double db = d[0] * b[0] + d[1] * b[1] + d[2] * b[2];
        double ft[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            ft[i] = -bn * d[i] + dn * b[i] + db * nv[i];
            u[i] += ft[i] * c1 * ir3 - 2 * bn * d[i] * ir3 + 3 * c2 * db * dn * d[i] * ir5;
        }
        if (grad)
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int l = 0; l < 3; l++)
                    G[i][l] += sgn3(l) * (-3 * ft[i] * c1 * d[l] * ir5 +
                                          c1 * ir3 * (-bn * dlt(i, l) + b[i] * nv[l] + b[l] * nv[i]) -
                                          2 * bn * (dlt(i, l) * ir3 - 3 * d[i] * d[l] * ir5) +
                                          3 * c2 *
                                              ((b[l] * dn + nv[l] * db) * d[i] * ir5 +
                                               db * dn * (dlt(i, l) * ir5 - 5 * d[i] * d[l] * ir7)));
    }
    ImgR g;
    img_setup(y[0] - x[0], y[1] - x[1], -(y[2] + x[2]), g);
    double sb = (1 - al) / al, s2 = -(2 - al), xi3 = -x[2], imu = 1.0 / mu;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        // B part
        double pot = 0, uuu = 0;
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) pot += mu * sb * U3(g, i, j, k) * (sgn3(j) + sgn3(k)) * b[j] * nv[k];
#pragma unroll
        for (int m = 0; m < 3; m++) uuu += sb * U3(g, i, m, m) * sgn3(m);
        pot += lam * uuu * bn;
        u[i] += pot * imu;
        if (grad)
#pragma unroll
            for (int l = 0; l < 3; l++) {
                double acc = 0, uu4 = 0;
#pragma unroll
                for (int j = 0; j < 3; j++)
#pragma unroll
                    for (int k = 0; k < 3; k++) acc += mu * U4(g, i, j, k, l) * (sgn3(j) + sgn3(k)) * b[j] * nv[k];
#pragma unroll
                for (int m = 0; m < 3; m++) uu4 += U4(g, i, m, m, l) * sgn3(m);
                double fld = -sgn3(l) * sb * (acc + lam * uu4 * bn);
                G[i][l] -= fld * imu;
            }
        // C part: A_ijk with the reference's single- and double-layer
        // adjustments (src/mindlinpieces.f:520-618)
        auto A = [&](int j, int k) {
            double a = -al * xi3 * sgn3(i) * V3(g, i, j, k);
            if (i != 2 && j == 2) a -= s2 * V2(g, i, k);
            if (i == 2 && j != 2) a -= s2 * V2(g, j, k);
            if (i != 2 && k == 2) a -= al * V2(g, i, j);
            if (i == 2 && k == 2) a += al * V2(g, 2, j);
            return a;
        };
        double potc = 0, uc = 0;
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) potc += mu * (A(j, k) + A(k, j)) * b[j] * nv[k];
#pragma unroll
        for (int m = 0; m < 3; m++) uc += A(m, m);
        potc += lam * uc * bn;
        u[i] += y[2] * potc * imu;
        if (grad)
#pragma unroll
            for (int l = 0; l < 3; l++) {
                double ud = 0;
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    ud += -V4(g, i, l, m, m) * al * xi3 * sgn3(i);
                    if (i == 2) ud -= s2 * V3(g, m, m, l);
                }
                if (i != 2) ud += -s2 * V3(g, i, 2, l) - al * V3(g, i, 2, l);
                else ud += s2 * V3(g, 2, 2, l) + al * V3(g, 2, 2, l);
                if (l != 2) ud = -ud;
                double fld = 0;
#pragma unroll
                for (int j = 0; j < 3; j++)
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        double w = b[j] * nv[k];
                        double base = -V4(g, i, l, j, k) * al * xi3 * sgn3(i);
                        double u1 = base, u2 = base;
                        if (i != 2) {
                            if (j == 2) u1 -= s2 * V3(g, i, k, l);
                            if (k == 2) u2 -= s2 * V3(g, i, j, l);
                            if (k == 2) u1 -= al * V3(g, i, j, l);
                            if (j == 2) u2 -= al * V3(g, i, k, l);
                        } else {
                            if (j != 2) u1 -= s2 * V3(g, j, k, l);
                            if (k != 2) u2 -= s2 * V3(g, j, k, l);
                            if (k == 2) u1 += al * V3(g, 2, j, l);
                            if (j == 2) u2 += al * V3(g, 2, k, l);
                        }
                        if (l != 2) {
                            u1 = -u1;
                            u2 = -u2;
                        }
                        fld += mu * (u1 + u2) * w + (j == k ? lam * ud * w : 0.0);
                    }
                G[i][l] += -y[2] * fld * imu + (l == 2 ? potc * imu : 0.0);
            }
    }
    img_pack(u, G, ifpt, ifst, f);
}

// ------------------------------------------------------------------------
// Adaptive quadrature over one triangle (the reference's tria3adam with
// nrec = 200, relative test, 12-node order-6 rule: src/tria3adamar.f:16-424,
// triadiv src/triagauc.f:654-717, rule placement src/triasymq.f:98-135).
// Sub-triangles live in the (u, v) parameter plane of the triangle; a
// sub-triangle's integral is sum_q f(X_q) w_q 2 area, w_q the order-6 weight
// of node q scaled to the sub-triangle's (u, v) area (exact: the areas are
// powers of two), area that of the triangle in space. A sub-triangle is
// split into four (corners first, then the middle one); the split is
// accepted when no output changes by more than eps * (largest output of the
// parent), else the four children go on the stack, the last one on top.
// ------------------------------------------------------------------------

constexpr int ADAP_NREC = 200;        // depth of subdivision
constexpr int ADAP_STACK = 200;       // stack entries (the reference's maxdepth)
constexpr int ADAP_MAXIT = 100000;    // subdivisions per integral
constexpr int ADAP_ENTRY = 20;        // doubles per stack entry: uv (6), values (12), depth, pad
// the reference sets eps = 1e-6 from a single-precision literal
constexpr double ADAP_EPS = (double)1e-6f;

// Children of a (u, v) sub-triangle z[6] = (z1, z2, z3), as triadiv.
ELFMM_HD void adap_child(const double *z, int c, double *out)
{
    double w12[2] = {(z[0] + z[2]) / 2, (z[1] + z[3]) / 2};
    double w13[2] = {(z[0] + z[4]) / 2, (z[1] + z[5]) / 2};
    double w23[2] = {(z[2] + z[4]) / 2, (z[3] + z[5]) / 2};
    const double *v[4][3] = {{z, w12, w13}, {z + 2, w23, w12}, {z + 4, w13, w23}, {w12, w23, w13}};
    for (int k = 0; k < 3; k++) {
        out[2 * k] = v[c][k][0];
        out[2 * k + 1] = v[c][k][1];
    }
}

// Node q of the order-6 rule on the (u, v) sub-triangle z, mapped onto the
// triangle tri (vertices in tri[0..2], [3..5], [6..8]): point X and the
// weight w (including 2 * (u, v) area). Rounding as in the reference's
// trianmap / tria3inm0 (no contraction).
ELFMM_HD void adap_node(const double *tri, const double *z, double qu, double qv, double qw, double *X, double &w)
{
#pragma clang fp contract(off)
    double a = (z[2] - z[0]) * (z[5] - z[1]) - (z[4] - z[0]) * (z[3] - z[1]);
    double area = fabs(a) * 0.5;
    double u = (z[2] - z[0]) * qu + (z[4] - z[0]) * qv + z[0];
    double v = (z[3] - z[1]) * qu + (z[5] - z[1]) * qv + z[1];
    for (int d = 0; d < 3; d++) X[d] = tri[d] + u * (tri[3 + d] - tri[d]) + v * (tri[6 + d] - tri[d]);
    w = qw * (2 * area);
}

// area of the triangle in space (tria3area)
ELFMM_HD double adap_area(const double *tri)
{
#pragma clang fp contract(off)
    double x[3] = {tri[3] - tri[0], tri[4] - tri[1], tri[5] - tri[2]};
    double y[3] = {tri[6] - tri[0], tri[7] - tri[1], tri[8] - tri[2]};
    double z0 = x[1] * y[2] - x[2] * y[1], z1 = x[2] * y[0] - x[0] * y[2], z2 = x[0] * y[1] - x[1] * y[0];
    return sqrt(z0 * z0 + z1 * z1 + z2 * z2) * 0.5;
}

// The integrand of one image-near-field integral: layer 0 = single layer
// (density dens), layer 1 = double layer (jump dens, normal nv).
struct AdapTask {
    double tri[9], dens[3], nv[3], y[3];
    double lam, mu;
    int layer, ifpt, ifst;
};

ELFMM_HD void adap_eval(const AdapTask &T, const double *X, double *f)
{
    if (T.layer == 0) image_sl(T.lam, T.mu, X, T.dens, T.y, T.ifpt, T.ifst, f);
    else image_dl(T.lam, T.mu, X, T.dens, T.nv, T.y, T.ifpt, T.ifst, f);
}

// Sequential restatement (host tests). rule[12][3]: the order-6 rule on the
// unit triangle (u, v, weight). Returns the reference's ier (0, 4, 8, 16).
inline int adap_integrate_host(const AdapTask &T, const double (*rule)[3], double *rint, int *nsub = nullptr)
{
    struct Ent {
        double z[6], val[12];
        int depth;
    };
    static thread_local Ent st[ADAP_STACK + 4];
    double area3 = adap_area(T.tri);
    auto integ = [&](const double *z, double *out) {
#pragma clang fp contract(off)
        for (int k = 0; k < 12; k++) out[k] = 0.0;
        for (int q = 0; q < 12; q++) {
            double X[3], w, f[12];
            adap_node(T.tri, z, rule[q][0], rule[q][1], rule[q][2], X, w);
            adap_eval(T, X, f);
            for (int k = 0; k < 12; k++) out[k] = out[k] + f[k] * w * 2 * area3;
        }
    };
    double z0[6] = {0, 0, 1, 0, 0, 1};
    for (int k = 0; k < 6; k++) st[0].z[k] = z0[k];
    integ(z0, st[0].val);
    st[0].depth = 1;
    for (int k = 0; k < 12; k++) rint[k] = 0.0;
    int j = 0, ier = 0;
    for (int it = 0; it < ADAP_MAXIT; it++) {
        if (nsub) *nsub = it + 1;
        Ent &e = st[j];
        if (e.depth >= ADAP_NREC) {
            ier = 4;
            for (int k = 0; k < 12; k++) rint[k] += e.val[k];
            if (--j < 0) return ier;
            continue;
        }
        double cz[4][6], cv[4][12];
        for (int c = 0; c < 4; c++) {
            adap_child(e.z, c, cz[c]);
            integ(cz[c], cv[c]);
        }
        double dmax = 0;
        for (int k = 0; k < 12; k++) dmax = fabs(e.val[k]) > dmax ? fabs(e.val[k]) : dmax;
        bool done = true;
        {
#pragma clang fp contract(off)
            for (int k = 0; k < 12; k++)
                if (fabs(cv[0][k] + cv[1][k] + cv[2][k] + cv[3][k] - e.val[k]) > ADAP_EPS * dmax) done = false;
        }
        if (done) {
#pragma clang fp contract(off)
            for (int k = 0; k < 12; k++) rint[k] = rint[k] + cv[0][k] + cv[1][k] + cv[2][k] + cv[3][k];
            if (--j < 0) return ier;
            continue;
        }
        int d = e.depth + 1;
        for (int c = 0; c < 4; c++) {
            Ent &o = st[j + c];
            for (int k = 0; k < 6; k++) o.z[k] = cz[c][k];
            for (int k = 0; k < 12; k++) o.val[k] = cv[c][k];
            o.depth = d;
        }
        j += 3;
        if (j + 4 > ADAP_STACK) return 8;
    }
    return 16;
}

}  // namespace elfmm
