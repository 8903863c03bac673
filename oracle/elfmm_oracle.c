/*
 * CPU restatement of the reference's free-space triangle layer-potential
 * path. TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's
 * cpu_baseline leg load it as the reference to compare against; the product
 * never links it.
 *
 * Contents
 *  1. Plane-triangle potential tables (src/triahquad.f:680-1014). The
 *     reference evaluates each of its 51 kernel integrals by a separate
 *     closed form. This restatement derives all of them from two
 *     potentials, F = int 1/r dA and G = int r dA: in-plane derivatives
 *     become boundary integrals by the divergence theorem (the same
 *     "gradient identity" the reference uses, src/triahquad.f:210-226),
 *     z-derivatives follow from Laplace(1/r)=0 and Laplace(r)=2/r, and the
 *     remaining area terms are F and the signed solid angle
 *     (Van Oosterom & Strackee, IEEE TBME 30 (1983) 125). Every boundary
 *     integral is a combination of int t^m / (t^2+rho^2)^(q/2) dt,
 *     m<=3, q in {1,3,5}, evaluated in cancellation-free forms.
 *  2. Pair quadratures elust3triadirecttarg_one / eltst3triadirecttarg_one
 *     (src/elfmm3dtria.f:1689-2009, 2368-2905): the same assembly of the
 *     displacement and strain from the tables.
 *  3. Direct sum el3dtriadirecttarg (src/elfmm3dtria.f:1453-1675).
 *  4. Octree and lists of d3tstrcr (src/d3tstrcr.f:31-2350).
 *
 * Parity pins (tests/): items 1-3 against the compiled reference's
 * triabtable/triartable and pair routines and against brute-force
 * quadrature; item 4 bit-exactly against the compiled reference's boxes,
 * permutations and lists.
 */
#include "elfmm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* 1. plane-triangle tables                                            */
/* ------------------------------------------------------------------ */

typedef struct {
    double k01, k11, k03, k13, k23, k05, k15, k25, k35;
} prims_t;

/* int_{t1}^{t2} t^m / r^q dt with r = sqrt(t^2 + rho2). */
static void edge_prims(double t1, double t2, double rho2, prims_t *p)
{
    int flip = 0;
    if (t2 <= 0.0) {                 /* reflect: [t1,t2] -> [-t2,-t1] */
        double tmp = -t2;
        t2 = -t1;
        t1 = tmp;
        flip = 1;
    }
    double r1 = sqrt(t1 * t1 + rho2), r2 = sqrt(t2 * t2 + rho2);
    double dsq = (t2 - t1) * (t2 + t1);
    double k11 = dsq / (r1 + r2);                       /* r2 - r1 */
    double k13 = k11 / (r1 * r2);                       /* 1/r1 - 1/r2 */
    double r13 = r1 * r1 * r1, r23 = r2 * r2 * r2;
    double k15 = k11 * (r2 * r2 + r1 * r2 + r1 * r1) / (3.0 * r13 * r23);
    double k01, k03, k05, k25;
    if (t1 >= 0.0) {
        k01 = log((t2 + r2) / (t1 + r1));
        k03 = 1.0 / ((t1 + r1) * r1) - 1.0 / ((t2 + r2) * r2);
        double ph1 = (3.0 * t1 * t1 + 4.0 * rho2) /
                     ((2.0 * t1 * t1 * t1 + 3.0 * rho2 * t1 + 2.0 * r13) * r13);
        double ph2 = (3.0 * t2 * t2 + 4.0 * rho2) /
                     ((2.0 * t2 * t2 * t2 + 3.0 * rho2 * t2 + 2.0 * r23) * r23);
        k05 = (ph1 - ph2) / 3.0;
        double c1 = (3.0 * t1 * t1 * t1 * t1 + 3.0 * t1 * t1 * rho2 + rho2 * rho2) /
                    ((t1 * t1 * t1 + r13) * r13);
        double c2 = (3.0 * t2 * t2 * t2 * t2 + 3.0 * t2 * t2 * rho2 + rho2 * rho2) /
                    ((t2 * t2 * t2 + r23) * r23);
        k25 = (c1 - c2) / 3.0;
    } else {
        k01 = log((t2 + r2) * (r1 - t1) / rho2);
        k03 = (t2 / r2 - t1 / r1) / rho2;
        k05 = (t2 * (2.0 * t2 * t2 + 3.0 * rho2) / r23 -
               t1 * (2.0 * t1 * t1 + 3.0 * rho2) / r13) / (3.0 * rho2 * rho2);
        k25 = (t2 * t2 * t2 / r23 - t1 * t1 * t1 / r13) / (3.0 * rho2);
    }
    double k23 = k01 - rho2 * k03;
    double k35 = k13 - rho2 * k15;
    if (flip) {
        k11 = -k11;
        k13 = -k13;
        k15 = -k15;
        k35 = -k35;
    }
    p->k01 = k01; p->k11 = k11; p->k03 = k03; p->k13 = k13; p->k23 = k23;
    p->k05 = k05; p->k15 = k15; p->k25 = k25; p->k35 = k35;
}

#define BT(i, j, k) btab[25 * (i) + 5 * (j) + (k)]
#define RT(i, j, k) rtab[9 * (i) + 3 * (j) + (k)]

/* index of a symmetric-tensor component given a multi-index (i,j,k) */
static int sym2(int i, int j, int k)       /* order 2: xx xy xz yy yz zz */
{
    if (i == 2) return 0;
    if (i == 1 && j == 1) return 1;
    if (i == 1) return 2;
    if (j == 2) return 3;
    if (j == 1) return 4;
    (void)k;
    return 5;
}

static int sym3(int i, int j, int k)       /* order 3, lexicographic */
{
    /* components: xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz, indexed by the
     * number of x's (i) and y's (j) */
    static const int tab[4][4] = {{9, 8, 7, 6}, {5, 4, 3, -1}, {2, 1, -1, -1}, {0, -1, -1, -1}};
    (void)k;
    return tab[i][j];
}

void oracle_tables(double a, double b, double c, double x0, double y0,
                   double z0, int iself, double *btab, double *rtab)
{
    memset(btab, 0, 125 * sizeof(double));
    memset(rtab, 0, 27 * sizeof(double));
    double w = iself ? 0.0 : -z0;      /* u_z = z_source - z0 */
    double zz = iself ? 0.0 : z0;
    const double vx[3] = {0.0, a, b}, vy[3] = {0.0, 0.0, c};

    double sumF = 0.0;
    double r1x = 0, r1y = 0;                 /* rt(1,0,0), rt(0,1,0) */
    double e1[3][3] = {{0}};                 /* sum n_a * E[D1(1/r)]_b  (a in x,y) */
    double b1[2][3] = {{0}};                 /* sum n_a * E[D1 r]_b */
    double b2[2][6] = {{0}};                 /* sum n_a * E[D2 r] */
    double b3[2][10] = {{0}};                /* sum n_a * E[D3 r] */

    for (int e = 0; e < 3; e++) {
        double ax = vx[e], ay = vy[e];
        double bx = vx[(e + 1) % 3], by = vy[(e + 1) % 3];
        double dx = bx - ax, dy = by - ay;
        double L = sqrt(dx * dx + dy * dy);
        double sx = dx / L, sy = dy / L;
        double nx = sy, ny = -sx;                       /* outward */
        double t1 = (ax - x0) * sx + (ay - y0) * sy;
        double del = (ax - x0) * nx + (ay - y0) * ny;
        double t2 = t1 + L;
        double rho2 = del * del + w * w;
        prims_t p;
        edge_prims(t1, t2, rho2, &p);
        double s[3] = {sx, sy, 0.0};
        double v[3] = {del * nx, del * ny, w};
        double n2[2] = {nx, ny};

        sumF += del * p.k01;
        r1x += nx * p.k01;
        r1y += ny * p.k01;

        double U3[3], U1[3];
        for (int i = 0; i < 3; i++) {
            U3[i] = s[i] * p.k13 + v[i] * p.k03;      /* E[u_i/r^3] */
            U1[i] = s[i] * p.k11 + v[i] * p.k01;      /* E[u_i/r]   */
        }
        double D2[6];
        {
            int idx = 0;
            for (int i = 0; i < 3; i++)
                for (int j = i; j < 3; j++) {
                    double uu = s[i] * s[j] * p.k23 + (s[i] * v[j] + v[i] * s[j]) * p.k13 +
                                v[i] * v[j] * p.k03;
                    D2[idx++] = (i == j ? p.k01 : 0.0) - uu;
                }
        }
        double D3[10];
        {
            int idx = 0;
            for (int i = 0; i < 3; i++)
                for (int j = i; j < 3; j++)
                    for (int k = j; k < 3; k++) {
                        double sss = s[i] * s[j] * s[k];
                        double ssv = s[i] * s[j] * v[k] + s[i] * v[j] * s[k] + v[i] * s[j] * s[k];
                        double svv = s[i] * v[j] * v[k] + v[i] * s[j] * v[k] + v[i] * v[j] * s[k];
                        double vvv = v[i] * v[j] * v[k];
                        double uuu = sss * p.k35 + ssv * p.k25 + svv * p.k15 + vvv * p.k05;
                        double tr = (i == j ? U3[k] : 0.0) + (i == k ? U3[j] : 0.0) +
                                    (j == k ? U3[i] : 0.0);
                        D3[idx++] = 3.0 * uuu - tr;
                    }
        }
        for (int q = 0; q < 2; q++) {
            for (int i = 0; i < 3; i++) {
                e1[q][i] += n2[q] * (-U3[i]);
                b1[q][i] += n2[q] * U1[i];
            }
            for (int i = 0; i < 6; i++) b2[q][i] += n2[q] * D2[i];
            for (int i = 0; i < 10; i++) b3[q][i] += n2[q] * D3[i];
        }
    }

    /* signed solid angle: int z0/r^3 dA */
    double wz = 0.0;
    if (!(iself || z0 == 0.0)) {
        double A[3] = {-x0, -y0, -z0}, B[3] = {a - x0, -y0, -z0}, Cc[3] = {b - x0, c - y0, -z0};
        double la = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
        double lb = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
        double lc = sqrt(Cc[0] * Cc[0] + Cc[1] * Cc[1] + Cc[2] * Cc[2]);
        double ab = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
        double ac = A[0] * Cc[0] + A[1] * Cc[1] + A[2] * Cc[2];
        double bc = B[0] * Cc[0] + B[1] * Cc[1] + B[2] * Cc[2];
        double den = la * lb * lc + ab * lc + ac * lb + bc * la;
        wz = 2.0 * atan2(z0 * a * c, den);
    }
    double F = sumF - zz * wz;

    /* 1/r tables */
    RT(0, 0, 0) = F;
    RT(1, 0, 0) = r1x;
    RT(0, 1, 0) = r1y;
    RT(0, 0, 1) = wz;
    RT(2, 0, 0) = e1[0][0];
    RT(1, 1, 0) = e1[0][1];
    RT(1, 0, 1) = e1[0][2];
    RT(0, 2, 0) = e1[1][1];
    RT(0, 1, 1) = e1[1][2];
    RT(0, 0, 2) = -RT(2, 0, 0) - RT(0, 2, 0);

    /* r tables: i>=1 from n_x, i==0 && j>=1 from n_y */
    for (int ord = 2; ord <= 4; ord++) {
        for (int i = 0; i <= ord; i++)
            for (int j = 0; i + j <= ord; j++) {
                int k = ord - i - j;
                double val;
                if (i >= 1 || j >= 1) {
                    int q = (i >= 1) ? 0 : 1;
                    int ii = i - (q == 0), jj = j - (q == 1);
                    if (ord == 2) {
                        int comp = ii == 1 ? 0 : (jj == 1 ? 1 : 2);
                        val = b1[q][comp];
                    } else if (ord == 3) {
                        val = b2[q][sym2(ii, jj, k)];
                    } else {
                        val = b3[q][sym3(ii, jj, k)];
                    }
                    BT(i, j, k) = val;
                }
            }
    }
    BT(0, 0, 2) = 2.0 * RT(0, 0, 0) - BT(2, 0, 0) - BT(0, 2, 0);
    BT(0, 0, 3) = 2.0 * RT(0, 0, 1) - BT(2, 0, 1) - BT(0, 2, 1);
    BT(0, 0, 4) = 2.0 * RT(0, 0, 2) - BT(2, 0, 2) - BT(0, 2, 2);
    /* orders 0 and 1 of r (not used by the pair quadratures) */
    BT(1, 0, 0) = 0.0;
    BT(0, 1, 0) = 0.0;
    BT(0, 0, 1) = -zz * F;
}

/* ------------------------------------------------------------------ */
/* 2. pair quadratures                                                 */
/* ------------------------------------------------------------------ */

/* Frame of rotshift3d (src/trilib.f:453-533): e1 along v1->v2,
 * e3 = e1 x (v3-v1) normalised, e2 = e3 x e1; U columns e1,e2,e3. */
static void tri_frame(const double *tri, double U[3][3], double *a, double *b, double *c)
{
    double d1[3], d2[3], e3[3], e2[3];
    for (int i = 0; i < 3; i++) {
        d1[i] = tri[3 + i] - tri[i];
        d2[i] = tri[6 + i] - tri[i];
    }
    double dd = sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
    *a = dd;
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
        U[i][0] = d1[i];
        U[i][1] = e2[i];
        U[i][2] = e3[i];
    }
    *b = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
    *c = e2[0] * d2[0] + e2[1] * d2[1] + e2[2] * d2[2];
}

static void to_local(double U[3][3], const double *x, const double *o, double *y)
{
    for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int i = 0; i < 3; i++) s += U[i][j] * (x[i] - (o ? o[i] : 0.0));
        y[j] = s;
    }
}

/* accumulate U*v and U*H*U^T */
static void rot_back_vec(double U[3][3], const double *v, double *out)
{
    for (int i = 0; i < 3; i++) out[i] = U[i][0] * v[0] + U[i][1] * v[1] + U[i][2] * v[2];
}

static void rot_back_sym(double U[3][3], double H[3][3], double out[3][3])
{
    double T[3][3];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++)
            T[i][k] = U[i][0] * H[0][k] + U[i][1] * H[1][k] + U[i][2] * H[2][k];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++)
            out[i][k] = T[i][0] * U[k][0] + T[i][1] * U[k][1] + T[i][2] * U[k][2];
}

static int ev(int i, int d) { return i == d; }

void oracle_pair_sl(double rlam, double rmu, const double *tri,
                    const double *sigma, int ifself, const double *target,
                    double *ptfrc, int ifstrain, double *strain)
{
    double U[3][3], a, b, c, P[3], sg[3];
    double btab[125], rtab[27];
    double c1 = (rlam + rmu) / (rlam + 2 * rmu);
    tri_frame(tri, U, &a, &b, &c);
    to_local(U, target, tri, P);
    to_local(U, sigma, NULL, sg);
    oracle_tables(a, b, c, P[0], P[1], P[2], ifself, btab, rtab);
    double dl[3] = {0, 0, 0}, H[3][3] = {{0}};
    for (int k = 0; k < 3; k++) {
        double val[3];
        for (int i = 0; i < 3; i++)
            val[i] = -BT(ev(i, 0) + ev(k, 0), ev(i, 1) + ev(k, 1), ev(i, 2) + ev(k, 2)) * c1 +
                     (i == k ? 2.0 * RT(0, 0, 0) : 0.0);
        for (int i = 0; i < 3; i++) dl[i] += sg[k] * val[i];
        if (ifstrain) {
            double S[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    S[i][j] = -BT(ev(i, 0) + ev(j, 0) + ev(k, 0), ev(i, 1) + ev(j, 1) + ev(k, 1),
                                  ev(i, 2) + ev(j, 2) + ev(k, 2)) * c1;
                    if (i == k) S[i][j] += 2.0 * RT(ev(j, 0), ev(j, 1), ev(j, 2));
                }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) H[i][j] += sg[k] * 0.5 * (S[i][j] + S[j][i]);
        }
    }
    double pg[3];
    rot_back_vec(U, dl, pg);
    for (int i = 0; i < 3; i++) ptfrc[i] = pg[i] / (2 * rmu);
    if (ifstrain) {
        double Hg[3][3];
        rot_back_sym(U, H, Hg);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) strain[3 * j + i] = -Hg[i][j] / (2 * rmu);
    }
}

void oracle_pair_dl(double rlam, double rmu, const double *tri,
                    const double *sigma, int ifself, const double *target,
                    double *ptfrc, int ifstrain, double *strain)
{
    double U[3][3], a, b, c, P[3], sg[3];
    double btab[125], rtab[27];
    double c1 = (rlam + rmu) / (rlam + 2 * rmu);
    tri_frame(tri, U, &a, &b, &c);
    to_local(U, target, tri, P);
    to_local(U, sigma, NULL, sg);
    oracle_tables(a, b, c, P[0], P[1], P[2], ifself, btab, rtab);
    /* um[a][k][b]: derivative (a,b) of the displacement for density e_k */
    double um[3][3][3], dm[3][3][3][3];
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                um[i][k][j] = -BT(ev(i, 0) + ev(j, 0) + ev(k, 0), ev(i, 1) + ev(j, 1) + ev(k, 1),
                                  ev(i, 2) + ev(j, 2) + ev(k, 2)) * c1 +
                              (i == k ? 2.0 * RT(ev(j, 0), ev(j, 1), ev(j, 2)) : 0.0);
                for (int m = 0; m < 3; m++)
                    dm[i][k][j][m] =
                        -BT(ev(i, 0) + ev(j, 0) + ev(k, 0) + ev(m, 0),
                            ev(i, 1) + ev(j, 1) + ev(k, 1) + ev(m, 1),
                            ev(i, 2) + ev(j, 2) + ev(k, 2) + ev(m, 2)) * c1 +
                        (i == k ? 2.0 * RT(ev(j, 0) + ev(m, 0), ev(j, 1) + ev(m, 1), ev(j, 2) + ev(m, 2))
                                : 0.0);
            }
    double T[3][3], S[3][3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            T[i][j] = (j == 2 ? rlam * (um[0][i][0] + um[1][i][1] + um[2][i][2]) : 0.0) +
                      rmu * (um[j][i][2] + um[2][i][j]);
            for (int m = 0; m < 3; m++)
                S[i][j][m] = (j == 2 ? rlam * (dm[0][i][0][m] + dm[1][i][1][m] + dm[2][i][2][m]) : 0.0) +
                             rmu * (dm[j][i][2][m] + dm[2][i][j][m]);
        }
    double dl[3] = {0, 0, 0}, H[3][3] = {{0}};
    for (int k = 0; k < 3; k++) {
        for (int i = 0; i < 3; i++) dl[i] += sg[k] * T[i][k];
        if (ifstrain)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) H[i][j] += sg[k] * 0.5 * (S[i][k][j] + S[j][k][i]);
    }
    double pg[3];
    rot_back_vec(U, dl, pg);
    for (int i = 0; i < 3; i++) ptfrc[i] = pg[i] / (2 * rmu);
    if (ifstrain) {
        double Hg[3][3];
        rot_back_sym(U, H, Hg);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) strain[3 * j + i] = -Hg[i][j] / (2 * rmu);
    }
}

/* ------------------------------------------------------------------ */
/* 3. direct sum                                                       */
/* ------------------------------------------------------------------ */

static void direct_one(double rlam, double rmu, const double *tri, int nsource,
                       int ifsingle, const double *ssl, int ifdouble, const double *sdl,
                       const double *x, int iself, int ifstrain, double *pt, double *st)
{
    double p0[3], s0[9];
    for (int i = 0; i < 3; i++) pt[i] = 0.0;
    for (int i = 0; i < 9; i++) st[i] = 0.0;
    for (int j = 0; j < nsource; j++) {
        int ifself = (j == iself);
        if (ifsingle) {
            oracle_pair_sl(rlam, rmu, tri + 9 * j, ssl + 3 * j, ifself, x, p0, ifstrain, s0);
            for (int i = 0; i < 3; i++) pt[i] += p0[i];
            if (ifstrain)
                for (int i = 0; i < 9; i++) st[i] += s0[i];
        }
        if (ifdouble) {
            oracle_pair_dl(rlam, rmu, tri + 9 * j, sdl + 3 * j, ifself, x, p0, ifstrain, s0);
            for (int i = 0; i < 3; i++) pt[i] += p0[i];
            if (ifstrain)
                for (int i = 0; i < 9; i++) st[i] += s0[i];
        }
    }
}

void oracle_direct(double rlam, double rmu, const double *tri,
                   const double *trinorm, int nsource, const double *source,
                   int ifsingle, const double *sigma_sl, int ifdouble,
                   const double *sigma_dl, int ifptfrc, double *ptfrc,
                   int ifstrain, double *strain, int ntarget,
                   const double *target, int ifptfrctarg, double *ptfrctarg,
                   int ifstraintarg, double *straintarg)
{
    (void)trinorm;
    if (ifptfrc || ifstrain) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int j = 0; j < nsource; j++) {
            double pt[3], st[9];
            direct_one(rlam, rmu, tri, nsource, ifsingle, sigma_sl, ifdouble, sigma_dl,
                       source + 3 * j, j, ifstrain, pt, st);
            if (ifptfrc)
                for (int i = 0; i < 3; i++) ptfrc[3 * j + i] = pt[i];
            if (ifstrain)
                for (int i = 0; i < 9; i++) strain[9 * j + i] = st[i];
        }
    }
    if (ifptfrctarg || ifstraintarg) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int j = 0; j < ntarget; j++) {
            double pt[3], st[9];
            direct_one(rlam, rmu, tri, nsource, ifsingle, sigma_sl, ifdouble, sigma_dl,
                       target + 3 * j, -1, ifstraintarg, pt, st);
            if (ifptfrctarg)
                for (int i = 0; i < 3; i++) ptfrctarg[3 * j + i] = pt[i];
            if (ifstraintarg)
                for (int i = 0; i < 9; i++) straintarg[9 * j + i] = st[i];
        }
    }
}

/* ------------------------------------------------------------------ */
/* 4. octree and lists                                                 */
/* ------------------------------------------------------------------ */

typedef struct {
    int *v;
    int n, cap;
} ivec_t;

static void ivec_push(ivec_t *a, int x)
{
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 8;
        a->v = (int *)realloc(a->v, sizeof(int) * a->cap);
    }
    a->v[a->n++] = x;
}

/* stable two-way split of idx[0..n) by coordinate `dim` against thresh
 * (src/d3tstrcr.f:1867-1897): points with coord <= thresh keep their order
 * in front, the rest follow in order. Returns the count of the first part. */
static int sepa0(const double *z, int *idx, int n, int dim, double thresh, int *work)
{
    int i1 = 0, i2 = 0;
    for (int i = 0; i < n; i++) {
        int j = idx[i];
        if (z[3 * (j - 1) + dim] > thresh)
            work[i2++] = j;
        else
            idx[i1++] = j;
    }
    for (int i = 0; i < i2; i++) idx[i1 + i] = work[i];
    return i1;
}

/* eight-way split in z, then y, then x (src/d3tstrcr.f:1712-1860) */
static void sepa1(const double *cent, const double *z, int *idx, int n, int *work, int *is, int *ns)
{
    int n1234 = sepa0(z, idx, n, 2, cent[2], work);
    int n5678 = n - n1234;
    int n12 = n1234 ? sepa0(z, idx, n1234, 1, cent[1], work) : 0;
    int n34 = n1234 - n12;
    int n56 = n5678 ? sepa0(z, idx + n1234, n5678, 1, cent[1], work) : 0;
    int n78 = n5678 - n56;
    int c[8];
    c[0] = n12 ? sepa0(z, idx, n12, 0, cent[0], work) : 0;
    c[1] = n12 - c[0];
    c[2] = n34 ? sepa0(z, idx + n12, n34, 0, cent[0], work) : 0;
    c[3] = n34 - c[2];
    c[4] = n56 ? sepa0(z, idx + n1234, n56, 0, cent[0], work) : 0;
    c[5] = n56 - c[4];
    c[6] = n78 ? sepa0(z, idx + n1234 + n56, n78, 0, cent[0], work) : 0;
    c[7] = n78 - c[6];
    is[0] = 1;
    for (int i = 0; i < 8; i++) {
        ns[i] = c[i];
        if (i) is[i] = is[i - 1] + ns[i - 1];
    }
}

/* box corners (src/d3tstrcr.f:1596-1676) reduced to min/max per axis */
static void box_minmax(const double *center0, double size, const int *box, double *mn, double *mx)
{
    int level = box[0];
    double side = size / (double)(1 << level);
    double side2 = side / 2;
    double x00 = center0[0] - size / 2, y00 = center0[1] - size / 2, z00 = center0[2] - size / 2;
    double c[3];
    c[0] = x00 + (box[1] - 1) * side + side2;
    c[1] = y00 + (box[2] - 1) * side + side2;
    c[2] = z00 + (box[3] - 1) * side + side2;
    for (int i = 0; i < 3; i++) {
        mn[i] = c[i] - side / 2;
        mx[i] = mn[i] + side;
    }
}

/* touch test with a relative tolerance (src/d3tstrcr.f:1082-1179) */
static int ifint(const double *mn1, const double *mx1, const double *mn2, const double *mx2)
{
    double eps = mx1[0] - mn1[0];
    if (eps > mx2[0] - mn2[0]) eps = mx2[0] - mn2[0];
    if (eps > mx2[1] - mn2[1]) eps = mx2[1] - mn2[1];
    if (eps > mx2[2] - mn2[2]) eps = mx2[2] - mn2[2];
    if (eps > mx1[1] - mn1[1]) eps = mx1[1] - mn1[1];
    if (eps > mx1[2] - mn1[2]) eps = mx1[2] - mn1[2];
    eps = eps / 10000;
    for (int d = 0; d < 3; d++) {
        if (mn1[d] > mx2[d] + eps) return 0;
        if (mn2[d] > mx1[d] + eps) return 0;
    }
    return 1;
}

int oracle_tree(const double *src, int n, const double *targ, int nt,
                int nbox, int imindlin, int maxboxes, int *nboxes_out,
                int *boxes_out, double *centers, int *iz, int *iztarg,
                int *laddr, int *nlev_out, double *center0, double *size_out,
                int *lstart, int *litems, long *nitems)
{
    if (n < 1) return 128;
    /* root box: bounding cube of sources and targets (src/d3tstrcr.f:1364-1402) */
    double xmin = src[0], xmax = src[0], ymin = src[1], ymax = src[1], zmin = src[2], zmax = src[2];
    for (int i = 0; i < n; i++) {
        const double *p = src + 3 * i;
        if (p[0] < xmin) xmin = p[0];
        if (p[0] > xmax) xmax = p[0];
        if (p[1] < ymin) ymin = p[1];
        if (p[1] > ymax) ymax = p[1];
        if (p[2] < zmin) zmin = p[2];
        if (p[2] > zmax) zmax = p[2];
    }
    for (int i = 0; i < nt; i++) {
        const double *p = targ + 3 * i;
        if (p[0] < xmin) xmin = p[0];
        if (p[0] > xmax) xmax = p[0];
        if (p[1] < ymin) ymin = p[1];
        if (p[1] > ymax) ymax = p[1];
        if (p[2] < zmin) zmin = p[2];
        if (p[2] > zmax) zmax = p[2];
    }
    double size = xmax - xmin, sy = ymax - ymin, sz = zmax - zmin;
    if (sy > size) size = sy;
    if (sz > size) size = sz;
    center0[0] = (xmin + xmax) / 2;
    center0[1] = (ymin + ymax) / 2;
    center0[2] = (zmin + zmax) / 2;
    if (imindlin) {
        /* half-space root (src/d3tstrcr_mindlin.f:1380-1383) */
        center0[2] = 0;
        double zm = fabs(zmin) > fabs(zmax) ? fabs(zmin) : fabs(zmax);
        if (size < 2 * zm) size = 2 * zm;
    }
    *size_out = size;

    int cap = maxboxes > 0 ? maxboxes : 8 * (n + nt) / (nbox > 0 ? nbox : 1) + 1024;
    int *B = (int *)calloc((size_t)20 * cap, sizeof(int));
    int *work = (int *)malloc(sizeof(int) * (n + nt + 4));
    for (int i = 0; i < n; i++) iz[i] = i + 1;
    for (int i = 0; i < nt; i++) iztarg[i] = i + 1;
    for (int i = 0; i < 400; i++) laddr[i] = 0;
    B[0] = 0; B[1] = 1; B[2] = 1; B[3] = 1; B[4] = 0;
    B[13] = 1; B[14] = n; B[15] = 1; B[16] = nt;
    B[17] = n > 0 ? 1 : 0;
    B[18] = nt > 0 ? 1 : 0;
    laddr[0] = 1;
    laddr[1] = 1;
    int ison = 1, nlev = 0;
    static const int iis[8] = {1, 2, 1, 2, 1, 2, 1, 2}, jjs[8] = {1, 1, 2, 2, 1, 1, 2, 2},
                     kks[8] = {1, 1, 1, 1, 2, 2, 2, 2};
    for (int level = 0; level < 100; level++) {
        laddr[2 * (level + 1)] = laddr[2 * level] + laddr[2 * level + 1];
        int nlevson = 0;
        int idad0 = laddr[2 * level], idad1 = idad0 + laddr[2 * level + 1] - 1;
        for (int idad = idad0; idad <= idad1; idad++) {
            int *D = B + 20 * (idad - 1);
            if (D[14] <= nbox && D[16] <= nbox) continue;
            if (D[17] == 1) D[17] = 2;
            if (D[18] == 1) D[18] = 2;
            int ii = D[1], jj = D[2], kk = D[3];
            double side = size / (double)(1 << level), side2 = side / 2;
            double cen[3] = {center0[0] - size / 2 + (ii - 1) * side + side2,
                             center0[1] - size / 2 + (jj - 1) * side + side2,
                             center0[2] - size / 2 + (kk - 1) * side + side2};
            int is[8], ns[8], ist[8], nst[8];
            int iiz = D[13], nz = D[14], iizt = D[15], nzt = D[16];
            sepa1(cen, src, iz + iiz - 1, nz, work, is, ns);
            sepa1(cen, targ, iztarg + iizt - 1, nzt, work, ist, nst);
            int idadson = 5;
            for (int i = 0; i < 8; i++) {
                if (ns[i] == 0 && nst[i] == 0) continue;
                nlevson++;
                ison++;
                nlev = level + 1;
                if (ison > cap) {
                    free(B);
                    free(work);
                    return 4;
                }
                int *S = B + 20 * (ison - 1);
                memset(S, 0, 20 * sizeof(int));
                S[0] = level + 1;
                S[1] = (ii - 1) * 2 + iis[i];
                S[2] = (jj - 1) * 2 + jjs[i];
                S[3] = (kk - 1) * 2 + kks[i];
                S[4] = idad;
                S[13] = is[i] + iiz - 1;
                S[14] = ns[i];
                S[15] = ist[i] + iizt - 1;
                S[16] = nst[i];
                S[17] = ns[i] > 0 ? 1 : 0;
                S[18] = nst[i] > 0 ? 1 : 0;
                D = B + 20 * (idad - 1);
                D[idadson++] = ison;
            }
        }
        laddr[2 * (level + 1) + 1] = nlevson;
        if (nlevson == 0) break;
    }
    int nboxes = ison;
    free(work);
    *nboxes_out = nboxes;
    *nlev_out = nlev;

    double *mn = (double *)malloc(sizeof(double) * 3 * nboxes);
    double *mx = (double *)malloc(sizeof(double) * 3 * nboxes);
    for (int i = 0; i < nboxes; i++) {
        box_minmax(center0, size, B + 20 * i, mn + 3 * i, mx + 3 * i);
        if (centers)
            for (int d = 0; d < 3; d++) centers[3 * i + d] = (mn[3 * i + d] + mx[3 * i + d]) / 2;
    }
    if (centers) {
        /* exact centre formula of d3tcentc */
        for (int i = 0; i < nboxes; i++) {
            int level = B[20 * i];
            double side = size / (double)(1 << level), side2 = side / 2;
            for (int d = 0; d < 3; d++)
                centers[3 * i + d] = center0[d] - size / 2 + (B[20 * i + 1 + d] - 1) * side + side2;
        }
    }

    /* lists, stored per (type, box) in insertion order (src/d3tstrcr.f:702-913) */
    ivec_t *L = (ivec_t *)calloc((size_t)5 * nboxes, sizeof(ivec_t));
#define LST(t, b) (&L[(size_t)((t)-1) * nboxes + ((b)-1)])
    int *collkids = (int *)malloc(sizeof(int) * 8 * 64);
    for (int ibox = 2; ibox <= nboxes; ibox++) {
        int idad = B[20 * (ibox - 1) + 4];
        ivec_t *c5 = LST(5, idad);
        int ncolls = c5->n + 1;
        int nk = 0;
        for (int i = 0; i < ncolls; i++) {
            int icoll = i == 0 ? idad : c5->v[i - 1];
            for (int j = 0; j < 8; j++) {
                int kid = B[20 * (icoll - 1) + 5 + j];
                if (kid <= 0) break;
                if (kid == ibox) continue;
                collkids[nk++] = kid;
            }
        }
        for (int i = 0; i < nk; i++) {
            int kid = collkids[i];
            if (ifint(mn + 3 * (kid - 1), mx + 3 * (kid - 1), mn + 3 * (ibox - 1), mx + 3 * (ibox - 1)))
                ivec_push(LST(5, ibox), kid);
            else
                ivec_push(LST(2, ibox), kid);
        }
    }
    free(collkids);
    /* lists 1 and 3 by depth-first descent into each colleague (src/d3tstrcr.f:919-1076) */
    int *stk = (int *)malloc(sizeof(int) * 2 * 256);
    for (int i = 1; i <= nboxes; i++) {
        if (B[20 * (i - 1) + 5] > 0) continue;
        if (B[20 * (i - 1)] == 0) continue;
        ivec_t *l5 = LST(5, i);
        for (int j = 0; j < l5->n; j++) {
            int jbox0 = l5->v[j];
            int jbox = jbox0, istack = 0;
            int nsons = 0;
            for (int q = 5; q < 13; q++)
                if (B[20 * (jbox - 1) + q] > 0) nsons++;
            stk[0] = jbox;
            stk[1] = nsons;
            for (;;) {
                const double *mnj = mn + 3 * (jbox - 1), *mxj = mx + 3 * (jbox - 1);
                if (!ifint(mn + 3 * (i - 1), mx + 3 * (i - 1), mnj, mxj)) {
                    ivec_push(LST(3, i), jbox);
                    istack--;
                    stk[2 * istack + 1]--;
                    jbox = stk[2 * istack];
                    continue;
                }
                if (B[20 * (jbox - 1) + 5] == 0) {
                    ivec_push(LST(1, i), jbox);
                    if (B[20 * (jbox - 1)] != B[20 * (i - 1)]) ivec_push(LST(1, jbox), i);
                    if (jbox == jbox0) break;
                    istack--;
                    stk[2 * istack + 1]--;
                    jbox = stk[2 * istack];
                    continue;
                }
                nsons = stk[2 * istack + 1];
                if (nsons == 0) {
                    if (jbox == jbox0) break;
                    istack--;
                    stk[2 * istack + 1]--;
                    jbox = stk[2 * istack];
                    continue;
                }
                jbox = B[20 * (jbox - 1) + 4 + nsons];
                istack++;
                nsons = 0;
                for (int q = 5; q < 13; q++)
                    if (B[20 * (jbox - 1) + q] > 0) nsons++;
                stk[2 * istack] = jbox;
                stk[2 * istack + 1] = nsons;
            }
        }
    }
    free(stk);
    /* list 4 is the dual of list 3 */
    for (int ibox = 1; ibox <= nboxes; ibox++) {
        ivec_t *l3 = LST(3, ibox);
        for (int j = 0; j < l3->n; j++) ivec_push(LST(4, l3->v[j]), ibox);
    }
    long tot = 0;
    for (int t = 1; t <= 5; t++)
        for (int b = 1; b <= nboxes; b++) tot += LST(t, b)->n;
    *nitems = tot;
    if (litems && lstart) {
        long pos = 0;
        for (int t = 1; t <= 5; t++) {
            for (int b = 1; b <= nboxes; b++) {
                lstart[(t - 1) * (nboxes + 1) + (b - 1)] = (int)pos;
                ivec_t *l = LST(t, b);
                for (int q = 0; q < l->n; q++) litems[pos++] = l->v[q];
            }
            lstart[(t - 1) * (nboxes + 1) + nboxes] = (int)pos;
        }
    }
    if (boxes_out) memcpy(boxes_out, B, sizeof(int) * 20 * (size_t)nboxes);
    for (int t = 0; t < 5 * nboxes; t++) free(L[t].v);
    free(L);
#undef LST
    free(mn);
    free(mx);
    free(B);
    return 0;
}
