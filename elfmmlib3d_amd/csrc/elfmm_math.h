// Device-side mathematics shared by the gfx950 kernels (also compiled for
// the host so the planner can use the same record layout).
//
// Near field: the elastostatic single- and double-layer quadratures of the
// reference's pair routines elust3triadirecttarg_one and
// eltst3triadirecttarg_one (src/elfmm3dtria.f:1689-2009, 2368-2905). The
// reference builds 24 (SL) and 45 (DL) plane-triangle integrals of
// derivatives of r and 1/r with separate closed forms (src/triahquad.f).
// Here all of them come from one pass over the three edges: in-plane
// derivatives are boundary integrals (divergence theorem), z-derivatives use
// Laplace(1/r) = 0 and Laplace(r) = 2/r, and the area terms are
// F = int 1/r dA and the signed solid angle. Every boundary integral is a
// combination of int t^m / (t^2+rho^2)^(q/2) dt (m <= 3, q in {1,3,5})
// evaluated in cancellation-free forms.
#pragma once

#ifndef ELFMM_HD
#if defined(__HIPCC__)
#define ELFMM_HD __host__ __device__ __forceinline__
#else
#define ELFMM_HD inline
#endif
#endif

#include <math.h>

namespace elfmm {

// Per-triangle near-field record (tree order):
//   [0..8]  frame U (columns e1,e2,e3; e3 = unit normal), U[3*col+row]
//   [9..11] first vertex
//   [12..14] a, b, c: local vertices (0,0), (a,0), (b,c)
//   [15..17] single-layer density in the local frame
//   [18..20] double-layer density in the local frame
//   [21..23] padding
constexpr int NF_REC = 24;

// Frame of the reference's rotshift3d (src/trilib.f:453-533).
ELFMM_HD void nf_make_record(const double *tri, const double *ssl, const double *sdl,
                             double *rec)
{
    double d1[3], d2[3], e3[3], e2[3];
    for (int i = 0; i < 3; i++) {
        d1[i] = tri[3 + i] - tri[i];
        d2[i] = tri[6 + i] - tri[i];
    }
    double dd = sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
    for (int i = 0; i < 3; i++) d1[i] /= dd;
    e3[0] = d1[1] * d2[2] - d1[2] * d2[1];
    e3[1] = d1[2] * d2[0] - d1[0] * d2[2];
    e3[2] = d1[0] * d2[1] - d1[1] * d2[0];
    double d3 = sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
    for (int i = 0; i < 3; i++) e3[i] /= d3;
    e2[0] = e3[1] * d1[2] - e3[2] * d1[1];
    e2[1] = e3[2] * d1[0] - e3[0] * d1[2];
    e2[2] = e3[0] * d1[1] - e3[1] * d1[0];
    for (int i = 0; i < 3; i++) {
        rec[i] = d1[i];
        rec[3 + i] = e2[i];
        rec[6 + i] = e3[i];
        rec[9 + i] = tri[i];
    }
    rec[12] = dd;
    rec[13] = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
    rec[14] = e2[0] * d2[0] + e2[1] * d2[1] + e2[2] * d2[2];
    for (int j = 0; j < 3; j++) {
        const double *e = rec + 3 * j;
        rec[15 + j] = ssl ? e[0] * ssl[0] + e[1] * ssl[1] + e[2] * ssl[2] : 0.0;
        rec[18 + j] = sdl ? e[0] * sdl[0] + e[1] * sdl[1] + e[2] * sdl[2] : 0.0;
    }
    rec[21] = rec[22] = rec[23] = 0.0;
}

struct EdgePrims {
    double k01, k11, k03, k13, k23, k05, k15, k25, k35;
};

// int_{t1}^{t2} t^m / (t^2 + rho2)^(q/2) dt. LOW selects the primitives of
// the orders <= 3 tables (q in {1,3}); otherwise those of the order-4 tables
// (q in {3,5}), which need no logarithm.
template <bool LOW>
ELFMM_HD void edge_prims(double t1, double t2, double rho2, EdgePrims &p)
{
    bool flip = t2 <= 0.0;
    if (flip) {
        double tmp = -t2;
        t2 = -t1;
        t1 = tmp;
    }
    double r1 = sqrt(t1 * t1 + rho2), r2 = sqrt(t2 * t2 + rho2);
    double k11 = (t2 - t1) * (t2 + t1) / (r1 + r2);
    double ir1 = 1.0 / r1, ir2 = 1.0 / r2;
    double k13 = k11 * ir1 * ir2;
    double k01 = 0.0, k03, k05 = 0.0, k15 = 0.0, k25 = 0.0;
    double r13 = r1 * r1 * r1, r23 = r2 * r2 * r2;
    if (!LOW) k15 = k11 * (r2 * r2 + r1 * r2 + r1 * r1) * (ir1 * ir1 * ir1) * (ir2 * ir2 * ir2) * (1.0 / 3.0);
    if (t1 >= 0.0) {
        if (LOW) k01 = log((t2 + r2) / (t1 + r1));
        k03 = 1.0 / ((t1 + r1) * r1) - 1.0 / ((t2 + r2) * r2);
        if (!LOW) {
            double ph1 = (3.0 * t1 * t1 + 4.0 * rho2) / ((2.0 * t1 * t1 * t1 + 3.0 * rho2 * t1 + 2.0 * r13) * r13);
            double ph2 = (3.0 * t2 * t2 + 4.0 * rho2) / ((2.0 * t2 * t2 * t2 + 3.0 * rho2 * t2 + 2.0 * r23) * r23);
            k05 = (ph1 - ph2) * (1.0 / 3.0);
            double c1 = (3.0 * t1 * t1 * t1 * t1 + 3.0 * t1 * t1 * rho2 + rho2 * rho2) / ((t1 * t1 * t1 + r13) * r13);
            double c2 = (3.0 * t2 * t2 * t2 * t2 + 3.0 * t2 * t2 * rho2 + rho2 * rho2) / ((t2 * t2 * t2 + r23) * r23);
            k25 = (c1 - c2) * (1.0 / 3.0);
        }
    } else {
        double irho2 = 1.0 / rho2;
        if (LOW) k01 = log((t2 + r2) * (r1 - t1) * irho2);
        k03 = (t2 * ir2 - t1 * ir1) * irho2;
        if (!LOW) {
            double ir13 = ir1 * ir1 * ir1, ir23 = ir2 * ir2 * ir2;
            k05 = (t2 * (2.0 * t2 * t2 + 3.0 * rho2) * ir23 - t1 * (2.0 * t1 * t1 + 3.0 * rho2) * ir13) *
                  (irho2 * irho2 * (1.0 / 3.0));
            k25 = (t2 * t2 * t2 * ir23 - t1 * t1 * t1 * ir13) * (irho2 * (1.0 / 3.0));
        }
    }
    double k23 = LOW ? k01 - rho2 * k03 : 0.0;
    double k35 = LOW ? 0.0 : k13 - rho2 * k15;
    if (flip) {
        k11 = -k11;
        k13 = -k13;
        k15 = -k15;
        k35 = -k35;
    }
    p.k01 = k01; p.k11 = k11; p.k03 = k03; p.k13 = k13; p.k23 = k23;
    p.k05 = k05; p.k15 = k15; p.k25 = k25; p.k35 = k35;
}

// Symmetric tensors: order-2 components xx xy xz yy yz zz, order-3
// components xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz, order-4 xxxx xxxy xxxz
// xxyy xxyz xxzz xyyy xyyz xyzz xzzz yyyy yyyz yyzz yzzz zzzz.
struct TablesLow {
    double F, r1[3];      // int 1/r; int d_i (1/r)
    double b2[6];         // int d_i d_j r
    double b3[10];        // int d_i d_j d_k r
};
struct TablesHigh {
    double r2[6];         // int d_i d_j (1/r)
    double b4[15];        // int d_i d_j d_k d_l r
};

// Geometry of edge e of the triangle (0,0),(a,0),(b,c) seen from (x0,y0):
// unit tangent s, outward normal n, foot parameter t1 and signed distance.
ELFMM_HD void nf_edge(int e, double a, double b, double c, double x0, double y0, double &sx,
                      double &sy, double &L, double &t1, double &del)
{
    double ax = e == 0 ? 0.0 : (e == 1 ? a : b), ay = e == 2 ? c : 0.0;
    double bx = e == 0 ? a : (e == 1 ? b : 0.0), by = e == 1 ? c : 0.0;
    double dx = bx - ax, dy = by - ay;
    L = sqrt(dx * dx + dy * dy);
    double iL = 1.0 / L;
    sx = dx * iL;
    sy = dy * iL;
    // outward normal (sy, -sx)
    t1 = (ax - x0) * sx + (ay - y0) * sy;
    del = (ax - x0) * sy - (ay - y0) * sx;
}

// Orders <= 3 of r and <= 1 of 1/r, in the reference's derivative
// convention (derivatives with respect to the source point,
// src/triahquad.f:680-1014). iself selects the principal value for the
// triangle's own centroid.
ELFMM_HD void nf_tables_low(double a, double b, double c, double x0, double y0, double z0, int iself,
                            TablesLow &T)
{
    double w = iself ? 0.0 : -z0;
    double zz = iself ? 0.0 : z0;
    double sumF = 0.0, r1x = 0.0, r1y = 0.0;
    double b1x[3] = {0, 0, 0}, b1yy = 0.0, b1yz = 0.0;      // n_x E[D1 r], n_y E[D1 r]_{y,z}
    double d2x[6] = {0, 0, 0, 0, 0, 0}, d2y[3] = {0, 0, 0}; // n_x E[D2 r], n_y E[D2 r]_{yy,yz,zz}
#pragma nounroll
    for (int e = 0; e < 3; e++) {
        double sx, sy, L, t1, del;
        nf_edge(e, a, b, c, x0, y0, sx, sy, L, t1, del);
        double nx = sy, ny = -sx;
        EdgePrims p;
        edge_prims<true>(t1, t1 + L, del * del + w * w, p);
        double s[3] = {sx, sy, 0.0};
        double v[3] = {del * nx, del * ny, w};
        sumF += del * p.k01;
        r1x += nx * p.k01;
        r1y += ny * p.k01;
        double U1[3];
#pragma unroll
        for (int i = 0; i < 3; i++) U1[i] = s[i] * p.k11 + v[i] * p.k01;
#pragma unroll
        for (int i = 0; i < 3; i++) b1x[i] += nx * U1[i];
        b1yy += ny * U1[1];
        b1yz += ny * U1[2];
        int idx = 0;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i; j < 3; j++) {
                double uu = s[i] * s[j] * p.k23 + (s[i] * v[j] + v[i] * s[j]) * p.k13 + v[i] * v[j] * p.k03;
                double d2 = (i == j ? p.k01 : 0.0) - uu;
                d2x[idx] += nx * d2;
                if (idx >= 3) d2y[idx - 3] += ny * d2;
                idx++;
            }
    }
    double wz = 0.0;
    if (!(iself || z0 == 0.0)) {
        double A0 = -x0, A1 = -y0, A2 = -z0;
        double B0 = a - x0, B1 = -y0, B2 = -z0;
        double C0 = b - x0, C1 = c - y0, C2 = -z0;
        double la = sqrt(A0 * A0 + A1 * A1 + A2 * A2);
        double lb = sqrt(B0 * B0 + B1 * B1 + B2 * B2);
        double lc = sqrt(C0 * C0 + C1 * C1 + C2 * C2);
        double ab = A0 * B0 + A1 * B1 + A2 * B2;
        double ac = A0 * C0 + A1 * C1 + A2 * C2;
        double bc = B0 * C0 + B1 * C1 + B2 * C2;
        double den = la * lb * lc + ab * lc + ac * lb + bc * la;
        wz = 2.0 * atan2(z0 * a * c, den);
    }
    T.F = sumF - zz * wz;
    T.r1[0] = r1x;
    T.r1[1] = r1y;
    T.r1[2] = wz;
    T.b2[0] = b1x[0];
    T.b2[1] = b1x[1];
    T.b2[2] = b1x[2];
    T.b2[3] = b1yy;
    T.b2[4] = b1yz;
    T.b2[5] = 2.0 * T.F - T.b2[0] - T.b2[3];
    // x-leading entries from n_x, the y/z-only ones from n_y
#pragma unroll
    for (int i = 0; i < 6; i++) T.b3[i] = d2x[i];
    T.b3[6] = d2y[0];
    T.b3[7] = d2y[1];
    T.b3[8] = d2y[2];
    T.b3[9] = 2.0 * T.r1[2] - T.b3[2] - T.b3[7];
}

// Order 2 of 1/r and order 4 of r.
ELFMM_HD void nf_tables_high(double a, double b, double c, double x0, double y0, double z0, int iself,
                             TablesHigh &T)
{
    double w = iself ? 0.0 : -z0;
    double e1x[3] = {0, 0, 0}, e1yy = 0.0, e1yz = 0.0;
    double d3x[10], d3y[4] = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 10; i++) d3x[i] = 0.0;
#pragma nounroll
    for (int e = 0; e < 3; e++) {
        double sx, sy, L, t1, del;
        nf_edge(e, a, b, c, x0, y0, sx, sy, L, t1, del);
        double nx = sy, ny = -sx;
        EdgePrims p;
        edge_prims<false>(t1, t1 + L, del * del + w * w, p);
        double s[3] = {sx, sy, 0.0};
        double v[3] = {del * nx, del * ny, w};
        double U3[3];
#pragma unroll
        for (int i = 0; i < 3; i++) U3[i] = s[i] * p.k13 + v[i] * p.k03;
#pragma unroll
        for (int i = 0; i < 3; i++) e1x[i] -= nx * U3[i];
        e1yy -= ny * U3[1];
        e1yz -= ny * U3[2];
        int idx = 0;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i; j < 3; j++)
#pragma unroll
                for (int k = j; k < 3; k++) {
                    double sss = s[i] * s[j] * s[k];
                    double ssv = s[i] * s[j] * v[k] + s[i] * v[j] * s[k] + v[i] * s[j] * s[k];
                    double svv = s[i] * v[j] * v[k] + v[i] * s[j] * v[k] + v[i] * v[j] * s[k];
                    double vvv = v[i] * v[j] * v[k];
                    double uuu = sss * p.k35 + ssv * p.k25 + svv * p.k15 + vvv * p.k05;
                    double tr = (i == j ? U3[k] : 0.0) + (i == k ? U3[j] : 0.0) + (j == k ? U3[i] : 0.0);
                    double d3 = 3.0 * uuu - tr;
                    d3x[idx] += nx * d3;
                    if (idx >= 6) d3y[idx - 6] += ny * d3;
                    idx++;
                }
    }
    T.r2[0] = e1x[0];
    T.r2[1] = e1x[1];
    T.r2[2] = e1x[2];
    T.r2[3] = e1yy;
    T.r2[4] = e1yz;
    T.r2[5] = -T.r2[0] - T.r2[3];
#pragma unroll
    for (int i = 0; i < 10; i++) T.b4[i] = d3x[i];
    T.b4[10] = d3y[0];
    T.b4[11] = d3y[1];
    T.b4[12] = d3y[2];
    T.b4[13] = d3y[3];
    T.b4[14] = 2.0 * T.r2[5] - T.b4[5] - T.b4[12];
}

// component index helpers for derivative multi-indices
ELFMM_HD int nf_i2(int i, int j)
{
    if (i > j) { int t = i; i = j; j = t; }
    return i == 0 ? j : (i == 1 ? 2 + j : 5);
}
ELFMM_HD int nf_i3(int i, int j, int k)
{
    // sort
    if (i > j) { int t = i; i = j; j = t; }
    if (j > k) { int t = j; j = k; k = t; }
    if (i > j) { int t = i; i = j; j = t; }
    // lexicographic rank among i<=j<=k
    const int base[3] = {0, 6, 9};
    const int off0[3] = {0, 3, 5};   // j offsets inside i=0
    if (i == 0) return off0[j] + (k - j);
    if (i == 1) return 6 + (j == 1 ? (k - 1) : 2);
    (void)base;
    return 9;
}
ELFMM_HD int nf_i4(int i, int j, int k, int l)
{
    int v[4] = {i, j, k, l};
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++)
            if (v[b] < v[a]) { int t = v[a]; v[a] = v[b]; v[b] = t; }
    // count x's, y's
    int nx = 0, ny = 0;
    for (int a = 0; a < 4; a++) {
        nx += v[a] == 0;
        ny += v[a] == 1;
    }
    // ordering of b4: by number of x descending, then y descending
    // nx=4:0; nx=3: y->1 z->2; nx=2: yy3 yz4 zz5; nx=1: yyy6 yyz7 yzz8 zzz9;
    // nx=0: yyyy10 yyyz11 yyzz12 yzzz13 zzzz14
    const int start[5] = {10, 6, 3, 1, 0};
    int rest = 4 - nx;
    return start[nx] + (rest - ny);
}

// Assembly of the elastic fields from the integral tables, in the
// triangle's local frame (z along the normal), for local densities sg.
// Single layer, displacement and strain (src/elfmm3dtria.f:1762-1840).
ELFMM_HD void nf_sl_low(const TablesLow &T, const double *sg, double c1, double *disp, double *H)
{
    // disp_i = -c1 sum_k s_k b(i,k) + 2 s_i F
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) acc += sg[k] * T.b2[nf_i2(i, k)];
        disp[i] += -c1 * acc + 2.0 * sg[i] * T.F;
    }
    // strain_ij = -[ -c1 sum_k s_k b(i,j,k) + s_i r(j) + s_j r(i) ]
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) acc += sg[k] * T.b3[nf_i3(i, j, k)];
            H[nf_i2(i, j)] -= -c1 * acc + sg[i] * T.r1[j] + sg[j] * T.r1[i];
        }
}

// Double layer, displacement. Reference assembly (src/elfmm3dtria.f:
// 2597-2634): with u(a,k,b) = -c1 b(a,b,k) + 2 d_ak r(b) and
// T(a,k) = lam d_k3 sum_l u(l,a,l) + mu (u(k,a,3) + u(3,a,k)),
// disp_a = sum_k s_k T(a,k). Summing over k first:
// disp_a = lam s_z (2 r(a) - c1 sum_l b(l,l,a))
//        + mu (-2 c1 sum_k s_k b(a,k,z) + 2 s_a r(z) + 2 d_az (s.r))
ELFMM_HD void nf_dl_low(const TablesLow &T, const double *sg, double c1, double rlam, double rmu,
                        double *disp)
{
    double sr = sg[0] * T.r1[0] + sg[1] * T.r1[1] + sg[2] * T.r1[2];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double lap = T.b3[nf_i3(0, 0, a)] + T.b3[nf_i3(1, 1, a)] + T.b3[nf_i3(2, 2, a)];
        double q = sg[0] * T.b3[nf_i3(a, 0, 2)] + sg[1] * T.b3[nf_i3(a, 1, 2)] + sg[2] * T.b3[nf_i3(a, 2, 2)];
        disp[a] += rlam * sg[2] * (2.0 * T.r1[a] - c1 * lap) +
                   rmu * (-2.0 * c1 * q + 2.0 * sg[a] * T.r1[2] + (a == 2 ? 2.0 * sr : 0.0));
    }
}

// Double layer, strain. Reference assembly (src/elfmm3dtria.f:2636-2660,
// 2692-2714): with d(a,k,b,m) = -c1 b(a,b,k,m) + 2 d_ak r(b,m) and
// S(a,k,b) = lam d_k3 sum_l d(l,a,l,b) + mu (d(k,a,3,b) + d(3,a,k,b)),
// strain_ab = -sym_ab sum_k s_k S(a,k,b). Summing over k first:
// lam s_z (2 r(a,b) - c1 sum_l b(l,l,a,b))
// + mu (-2 c1 sum_k s_k b(a,b,k,z) + s_a r(z,b) + s_b r(z,a)
//       + d_az (s.r)_b + d_bz (s.r)_a)
ELFMM_HD void nf_dl_high(const TablesHigh &T, const double *sg, double c1, double rlam, double rmu,
                         double *H)
{
    double sr[3];
#pragma unroll
    for (int b = 0; b < 3; b++)
        sr[b] = sg[0] * T.r2[nf_i2(0, b)] + sg[1] * T.r2[nf_i2(1, b)] + sg[2] * T.r2[nf_i2(2, b)];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = a; b < 3; b++) {
            double lap = T.b4[nf_i4(0, 0, a, b)] + T.b4[nf_i4(1, 1, a, b)] + T.b4[nf_i4(2, 2, a, b)];
            double q = sg[0] * T.b4[nf_i4(a, b, 0, 2)] + sg[1] * T.b4[nf_i4(a, b, 1, 2)] +
                       sg[2] * T.b4[nf_i4(a, b, 2, 2)];
            double v = rlam * sg[2] * (2.0 * T.r2[nf_i2(a, b)] - c1 * lap) +
                       rmu * (-2.0 * c1 * q + sg[a] * T.r2[nf_i2(2, b)] + sg[b] * T.r2[nf_i2(2, a)] +
                              (a == 2 ? sr[b] : 0.0) + (b == 2 ? sr[a] : 0.0));
            H[nf_i2(a, b)] -= v;
        }
}

// Local displacement and strain back to global coordinates, scaled by
// s: out[0..2] += s U disp, out[3..8] += s (U H U^T) as xx yy zz xy xz yz.
ELFMM_HD void nf_to_global(const double *U, const double *disp, const double *H, double s, double *out)
{
    double g[3];
#pragma unroll
    for (int i = 0; i < 3; i++) g[i] = U[i] * disp[0] + U[3 + i] * disp[1] + U[6 + i] * disp[2];
    double Hm[3][3] = {{H[0], H[1], H[2]}, {H[1], H[3], H[4]}, {H[2], H[4], H[5]}};
    double W[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) W[i][k] = U[i] * Hm[0][k] + U[3 + i] * Hm[1][k] + U[6 + i] * Hm[2][k];
    double G[6];
    const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        int i = ia[q], k = ib[q];
        G[q] = W[i][0] * U[k] + W[i][1] * U[3 + k] + W[i][2] * U[6 + k];
    }
    out[0] += s * g[0];
    out[1] += s * g[1];
    out[2] += s * g[2];
#pragma unroll
    for (int q = 0; q < 6; q++) out[3 + q] += s * G[q];
}

// Point coordinates in the triangle's local frame.
ELFMM_HD void nf_local_point(const double *rec, const double *P, double &x0, double &y0, double &z0)
{
    const double *U = rec;
    double q0 = P[0] - rec[9], q1 = P[1] - rec[10], q2 = P[2] - rec[11];
    x0 = U[0] * q0 + U[1] * q1 + U[2] * q2;
    y0 = U[3] * q0 + U[4] * q1 + U[5] * q2;
    z0 = U[6] * q0 + U[7] * q1 + U[8] * q2;
}

// One (target, triangle) pair: accumulate displacement (3) and strain
// (xx yy zz xy xz yz) in global coordinates into out[9]; c1 =
// (lam+mu)/(lam+2mu); ifsl/ifdl select the layers. Two passes over the
// edges keep the live state small: orders <= 3 (single layer, double-layer
// displacement), then order 4 (double-layer strain).
ELFMM_HD void nf_pair(const double *rec, const double *P, int iself, double rlam, double rmu,
                      double c1, int ifsl, int ifdl, double *out)
{
    double x0, y0, z0;
    nf_local_point(rec, P, x0, y0, z0);
    double disp[3] = {0, 0, 0};
    double H[6] = {0, 0, 0, 0, 0, 0};   // local xx xy xz yy yz zz
    {
        TablesLow T;
        nf_tables_low(rec[12], rec[13], rec[14], x0, y0, z0, iself, T);
        if (ifsl) nf_sl_low(T, rec + 15, c1, disp, H);
        if (ifdl) nf_dl_low(T, rec + 18, c1, rlam, rmu, disp);
    }
    if (ifdl) {
        TablesHigh T;
        nf_tables_high(rec[12], rec[13], rec[14], x0, y0, z0, iself, T);
        nf_dl_high(T, rec + 18, c1, rlam, rmu, H);
    }
    nf_to_global(rec, disp, H, 0.5 / rmu, out);
}

// The same pair for the unit densities e_k (global frame), k = 0..2:
// emit(layer, k, v) receives the 9 fields (displacement, strain xx yy zz
// xy xz yz) of layer 0 (single) or 1 (double) with density e_k. The tables
// are computed once; the fields are linear in the density, so
// sum_k sigma_k v[layer][k] is the pair's contribution.
template <class Emit>
ELFMM_HD void nf_pair_basis(const double *rec, const double *P, int iself, double rlam, double rmu,
                            double c1, int ifsl, int ifdl, Emit &&emit)
{
    double x0, y0, z0;
    nf_local_point(rec, P, x0, y0, z0);
    double dld[3][3];                   // double-layer local displacement per k
    {
        TablesLow T;
        nf_tables_low(rec[12], rec[13], rec[14], x0, y0, z0, iself, T);
        for (int k = 0; k < 3; k++) {
            double sg[3] = {rec[k], rec[3 + k], rec[6 + k]};   // e_k in the local frame
            if (ifsl) {
                double disp[3] = {0, 0, 0}, H[6] = {0, 0, 0, 0, 0, 0}, v[9];
                nf_sl_low(T, sg, c1, disp, H);
                for (int q = 0; q < 9; q++) v[q] = 0.0;
                nf_to_global(rec, disp, H, 0.5 / rmu, v);
                emit(0, k, v);
            }
            dld[k][0] = dld[k][1] = dld[k][2] = 0.0;
            if (ifdl) nf_dl_low(T, sg, c1, rlam, rmu, dld[k]);
        }
    }
    if (ifdl) {
        TablesHigh T;
        nf_tables_high(rec[12], rec[13], rec[14], x0, y0, z0, iself, T);
        for (int k = 0; k < 3; k++) {
            double sg[3] = {rec[k], rec[3 + k], rec[6 + k]};
            double H[6] = {0, 0, 0, 0, 0, 0}, v[9];
            nf_dl_high(T, sg, c1, rlam, rmu, H);
            for (int q = 0; q < 9; q++) v[q] = 0.0;
            nf_to_global(rec, dld[k], H, 0.5 / rmu, v);
            emit(1, k, v);
        }
    }
}

}  // namespace elfmm
