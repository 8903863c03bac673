// Mindlin harmonic-image far field on gfx950 (lfmm3dmindlintriatarg,
// src/lfmm3dtriamindlin_iter_1.f:41; passes src/lfmm3droutsmindlin.f:28).
// The kernels here are the steps that are not plain Laplace translations;
// M2M, L2L, the near list-2 M2L and L2T run on the Laplace kernels of
// elfmm_gpu.hip with the three expansions B, C1, C2 as right-hand sides.
// One workgroup per box (or per target box) throughout: the per-box work is
// O(p^3) with p = 26 and the boxes are few (leaves of up to 240 sources).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "elfmm_half.h"
#include "elfmm_internal.h"
#include "elfmm_mindlin.h"

namespace elfmm {

constexpr int MW = 256;                      // threads per workgroup
constexpr int MPX = 1024;                    // threads per workgroup of the proxy projection
constexpr int MP_MAX = 39;                   // largest supported order (iprec 3)
constexpr int MH = (MP_MAX + 1) * (MP_MAX + 2) / 2;
constexpr int MPW_FOUR = 544;                // Fourier modes of the largest plane-wave table (525)
constexpr size_t MIND_P2M_LDS = 100 * 1024;   // LDS of the P2M moment pass

// rows 0-2: order 2, rows 3-8: order 4, rows 9-20: order 6
__constant__ double c_mrule[21][3] = {ELFMM_RULE2, ELFMM_RULE4, ELFMM_RULE6};

// Regular harmonics R_l^m(x) (Racah normalisation, Condon-Shortley phase)
// of one point for l <= L, 0 <= m <= l, into H[pk(l, m)] (re, im): thread m
// computes column m. Caller synchronises.
__device__ void harm_regular(int L, double x, double y, double z, double2 *H)
{
    int m = threadIdx.x;
    if (m > L) return;
    double r2 = x * x + y * y + z * z;
    double dr = 1.0, di = 0.0;
    for (int k = 1; k <= m; k++) {
        double f = sqrt((2.0 * k - 1) / (2.0 * k));
        double nr = -(x * dr - y * di) * f, ni = -(x * di + y * dr) * f;
        dr = nr;
        di = ni;
    }
    double ar = dr, ai = di, br = 0, bi = 0;
    H[pk(m, m)] = make_double2(ar, ai);
    if (m + 1 <= L) {
        double f = sqrt(2.0 * m + 1) * z;
        br = f * ar;
        bi = f * ai;
        H[pk(m + 1, m)] = make_double2(br, bi);
    }
    for (int l = m + 2; l <= L; l++) {
        double f1 = (2.0 * l - 1) * z, f2 = sqrt((l - 1.0 + m) * (l - 1.0 - m)) * r2, d = sqrt((l + m) * (double)(l - m));
        double cr = (f1 * br - f2 * ar) / d, ci = (f1 * bi - f2 * ai) / d;
        H[pk(l, m)] = make_double2(cr, ci);
        ar = br;
        ai = bi;
        br = cr;
        bi = ci;
    }
}

// R_l^m for any m (R^{-m} = (-1)^m conj(R^m)); zero outside 0 <= |m| <= l.
__device__ __forceinline__ double2 hget(const double2 *H, int l, int m)
{
    if (l < 0 || m > l || -m > l) return make_double2(0.0, 0.0);
    if (m >= 0) return H[pk(l, m)];
    double2 c = H[pk(l, -m)];
    return (m & 1) ? make_double2(-c.x, c.y) : make_double2(c.x, -c.y);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Ladders of the regular harmonics: d_z R_l^m = sqrt((l+m)(l-m)) R_{l-1}^m,
// (d_x + i d_y) R_l^m = sqrt((l-m)(l-m-1)) R_{l-1}^{m+1},
// (d_x - i d_y) R_l^m = -sqrt((l+m)(l+m-1)) R_{l-1}^{m-1}.
__device__ __forceinline__ double lz(int l, int m) { return (l + m) * (double)(l - m) > 0 ? sqrt((l + m) * (double)(l - m)) : 0.0; }
__device__ __forceinline__ double lp(int l, int m)
{
    double v = (l - m) * (double)(l - m - 1);
    return v > 0 ? sqrt(v) : 0.0;
}
__device__ __forceinline__ double lmn(int l, int m)
{
    double v = (l + m) * (double)(l + m - 1);
    return v > 0 ? -sqrt(v) : 0.0;
}

// P2M of B, C1, C2 (l3dtriaformmp_mindlinb_trunc / _mindlinc_trunc,
// src/mindlinformmptri.f:305, :449 with formmp_mindlinb_trunc
// src/mindlinbformmp.f and formmp_mindlinc_trunc src/mindlincformmp.f) from
// the nodes of the order-nqtri rule on a leaf's (mirrored) triangles (order
// 6 from iprec 1 up, 4 at iprec 0, 2 below,
// src/lfmm3dtriamindlin_iter_1.f:679-693), densities times weight * 2 *
// area. Every charge, dipole and quadrupole strength the three expansions
// take at a node is a fixed combination of 18 real columns:
//   0-2 sigma_sl, 3-5 y3 sigma_sl, 6-11 q, 12-17 y3 q,
// q = (b1 n1, b2 n2, b3 n3, b1 n2 + b2 n1, b1 n3 + b3 n1, b2 n3 + b3 n2)
// with b = sigma_dl and n the original normal, y3 the node's (mirrored)
// height. Pass 1 sums, per box and column, the column times the regular
// harmonics of the node about the box centre (moments G_col(l, m)); pass 2
// forms each coefficient from the moments through the harmonics' ladders.

constexpr int MCOLS = 18;

// Pass 1: one workgroup per box, thread = coefficient (l, m). The columns
// are per-triangle values times per-node factors: a triangle's moments are
// its 9 values (sigma_sl, q) times U = sum_nodes w R(x) and U' = sum_nodes
// w y3 R(x), w = rule weight * 2 * area. Nodes go through LDS nch at a time
// (their harmonics by column recurrences with the constants tabulated in
// LDS); each thread folds a finished triangle into its 18 moments.
// NT >= (p+1)(p+2)/2.
template <int NT>
__global__ __launch_bounds__(NT) void k_mind_moments(MindArgs a, const int *__restrict__ boxes, int nch,
                                                     double2 *__restrict__ G)
{
    extern __shared__ double smem[];
    int p = a.p, nown = (p + 1) * (p + 2) / 2;
    double2 *Hs = (double2 *)smem;                          // [nch][nown] harmonics
    double2 *Rc = Hs + (size_t)nch * nown;                  // [nown] (a, b) of R_l = a z R_{l-1} - b r^2 R_{l-2}
    double *Dg = (double *)(Rc + nown);                     // [p + 1] sqrt((2k-1)/(2k))
    double *Ws = Dg + (p + 1);                              // [nch][2] w, w y3
    double *Tv = Ws + 2 * nch;                              // [nch][9] triangle values at the node
    double *Xs = Tv + 9 * nch;                              // [nch][3] node - centre
    int *Is = (int *)(Xs + 3 * nch);                        // [nch] triangle of the node (in the box)
    int tid = threadIdx.x;
    for (int k = tid; k < nown; k += NT) {
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= k) l++;
        int m = k - l * (l + 1) / 2;
        double ca = 0.0, cb = 0.0;
        if (l == m + 1) {
            ca = sqrt(2.0 * m + 1);
        } else if (l >= m + 2) {
            double d = 1.0 / sqrt((l + m) * (double)(l - m));
            ca = (2.0 * l - 1) * d;
            cb = sqrt((l - 1.0 + m) * (l - 1.0 - m)) * d;
        }
        Rc[k] = make_double2(ca, cb);
    }
    for (int k = tid; k <= p; k += NT) Dg[k] = k ? sqrt((2.0 * k - 1) / (2.0 * k)) : 1.0;
    int b = boxes[blockIdx.x];
    const double *c = a.centers + 3 * (size_t)b;
    int q0 = a.nqtri == 2 ? 0 : a.nqtri == 4 ? 3 : 9, q1 = a.nqtri == 2 ? 3 : a.nqtri == 4 ? 9 : 21;
    int nqn = q1 - q0;
    int s0 = a.src_start[b];
    int K = a.src_count[b] * nqn;
    double2 g[MCOLS];
#pragma unroll
    for (int j = 0; j < MCOLS; j++) g[j] = make_double2(0.0, 0.0);
    double2 U = make_double2(0.0, 0.0), U2 = make_double2(0.0, 0.0);
    double sv[9];
#pragma unroll
    for (int j = 0; j < 9; j++) sv[j] = 0.0;
    int tc = -1;
    auto flush = [&]() {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            g[k].x += sv[k] * U.x;
            g[k].y += sv[k] * U.y;
            g[3 + k].x += sv[k] * U2.x;
            g[3 + k].y += sv[k] * U2.y;
        }
#pragma unroll
        for (int j = 0; j < 6; j++) {
            g[6 + j].x += sv[3 + j] * U.x;
            g[6 + j].y += sv[3 + j] * U.y;
            g[12 + j].x += sv[3 + j] * U2.x;
            g[12 + j].y += sv[3 + j] * U2.y;
        }
    };
    for (int k0 = 0; k0 < K; k0 += nch) {
        int nk = min(nch, K - k0);
        __syncthreads();
        if (tid < nk) {
            int tl = (k0 + tid) / nqn, qn = q0 + (k0 + tid) % nqn, ti = s0 + tl;
            const double *t = a.tri + 9 * (size_t)ti;
            double ax = t[3] - t[0], ay = t[4] - t[1], az = t[5] - t[2];
            double bx = t[6] - t[0], by = t[7] - t[1], bz = t[8] - t[2];
            double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
            double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
            double u = c_mrule[qn][0], v = c_mrule[qn][1];
            double X[3];
            for (int d = 0; d < 3; d++) X[d] = t[d] + u * (t[3 + d] - t[d]) + v * (t[6 + d] - t[d]);
            double w = c_mrule[qn][2] * area * 2;
            Ws[2 * tid] = w;
            Ws[2 * tid + 1] = w * X[2];
            double *tv = Tv + 9 * tid;
            for (int k = 0; k < 3; k++) tv[k] = a.ifsl ? a.ssl[3 * (size_t)ti + k] : 0.0;
            double q[6] = {0, 0, 0, 0, 0, 0};
            if (a.ifdl) {
                const double *nv = a.nrm + 3 * (size_t)ti;
                const double *bv = a.sdl + 3 * (size_t)ti;
                q[0] = bv[0] * nv[0];
                q[1] = bv[1] * nv[1];
                q[2] = bv[2] * nv[2];
                q[3] = bv[0] * nv[1] + bv[1] * nv[0];
                q[4] = bv[0] * nv[2] + bv[2] * nv[0];
                q[5] = bv[1] * nv[2] + bv[2] * nv[1];
            }
            for (int k = 0; k < 6; k++) tv[3 + k] = q[k];
            for (int d = 0; d < 3; d++) Xs[3 * tid + d] = X[d] - c[d];
            Is[tid] = tl;
        }
        __syncthreads();
        // harmonics of the chunk's nodes, one column (node, m) per task
        for (int task = tid; task < nk * (p + 1); task += NT) {
            int k = task / (p + 1), m = task % (p + 1);
            double x = Xs[3 * k], y = Xs[3 * k + 1], z = Xs[3 * k + 2];
            double2 *H = Hs + (size_t)k * nown;
            double r2 = x * x + y * y + z * z;
            double dr = 1.0, di = 0.0;
            for (int j = 1; j <= m; j++) {
                double f = Dg[j];
                double nr = -(x * dr - y * di) * f, ni = -(x * di + y * dr) * f;
                dr = nr;
                di = ni;
            }
            double ar = dr, ai = di;
            H[pk(m, m)] = make_double2(ar, ai);
            if (m + 1 <= p) {
                double f = Rc[pk(m + 1, m)].x * z;
                double br = f * ar, bi = f * ai;
                H[pk(m + 1, m)] = make_double2(br, bi);
                for (int l = m + 2; l <= p; l++) {
                    double2 cc = Rc[pk(l, m)];
                    double f1 = cc.x * z, f2 = cc.y * r2;
                    double cr = f1 * br - f2 * ar, ci = f1 * bi - f2 * ai;
                    H[pk(l, m)] = make_double2(cr, ci);
                    ar = br;
                    ai = bi;
                    br = cr;
                    bi = ci;
                }
            }
        }
        __syncthreads();
        if (tid < nown) {
            for (int k = 0; k < nk; k++) {
                int t = Is[k];
                if (t != tc) {
                    if (tc >= 0) flush();
                    tc = t;
#pragma unroll
                    for (int j = 0; j < 9; j++) sv[j] = Tv[9 * k + j];
                    U = make_double2(0.0, 0.0);
                    U2 = make_double2(0.0, 0.0);
                }
                double2 v = Hs[(size_t)k * nown + tid];
                double w = Ws[2 * k], wy = Ws[2 * k + 1];
                U.x += w * v.x;
                U.y += w * v.y;
                U2.x += wy * v.x;
                U2.y += wy * v.y;
            }
        }
    }
    if (tid < nown) {
        if (tc >= 0) flush();
#pragma unroll
        for (int j = 0; j < MCOLS; j++) G[((size_t)blockIdx.x * MCOLS + j) * nown + tid] = g[j];
    }
}

// Moment of column col at (l, m), any m (G^{-m} = (-1)^m conj(G^m): the
// columns are real); zero outside 0 <= |m| <= l.
__device__ __forceinline__ double2 gmom(const double2 *G, int nown, int col, int l, int m)
{
    if (l < 0 || m > l || -m > l) return make_double2(0.0, 0.0);
    double2 v = G[(size_t)col * nown + pk(l, m < 0 ? -m : m)];
    if (m >= 0) return v;
    return (m & 1) ? make_double2(-v.x, v.y) : make_double2(v.x, -v.y);
}

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cscl(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
// a - i b and a + i b
__device__ __forceinline__ double2 cmi(double2 a, double2 b) { return make_double2(a.x + b.y, a.y - b.x); }
__device__ __forceinline__ double2 cpi(double2 a, double2 b) { return make_double2(a.x - b.y, a.y + b.x); }

// (D R_l^m) summed over the nodes for the operator D with charge moment Mq,
// dipole moments Mx, My, Mz and quadrupole moments Q0..Q5 (xx yy zz xy xz
// yz, each pair once), each a function (l', m') -> complex moment:
//   d.grad = d_z d_z + ((d_x - i d_y) d_+ + (d_x + i d_y) d_-) / 2,
//   sum Q d d = A d_+^2 + B d_-^2 + C d_z^2 + E d_+ d_z + F d_- d_z
// (d_+ d_- = -d_z^2 on harmonic functions), with the ladders
//   d_z R_l^m = sqrt((l+m)(l-m)) R_{l-1}^m,
//   d_+ R_l^m = sqrt((l-m)(l-m-1)) R_{l-1}^{m+1},
//   d_- R_l^m = -sqrt((l+m)(l+m-1)) R_{l-1}^{m-1}.
template <class FQ, class FX, class FY, class FZ, class F0, class F1, class F2, class F3, class F4, class F5>
__device__ double2 apply_moments(int l, int m, FQ Mq, FX Mx, FY My, FZ Mz, F0 Q0, F1 Q1, F2 Q2, F3 Q3, F4 Q4, F5 Q5)
{
    double2 acc = Mq(l, m);
    acc = cadd(acc, cscl(lz(l, m), Mz(l - 1, m)));
    acc = cadd(acc, cscl(0.5 * lp(l, m), cmi(Mx(l - 1, m + 1), My(l - 1, m + 1))));
    acc = cadd(acc, cscl(0.5 * lmn(l, m), cpi(Mx(l - 1, m - 1), My(l - 1, m - 1))));
    if (l >= 2) {
        double2 d01 = cadd(Q0(l - 2, m + 2), cscl(-1.0, Q1(l - 2, m + 2)));
        acc = cadd(acc, cscl(0.25 * lp(l, m) * lp(l - 1, m + 1), cmi(d01, Q3(l - 2, m + 2))));
        d01 = cadd(Q0(l - 2, m - 2), cscl(-1.0, Q1(l - 2, m - 2)));
        acc = cadd(acc, cscl(0.25 * lmn(l, m) * lmn(l - 1, m - 1), cpi(d01, Q3(l - 2, m - 2))));
        double2 cz = cadd(Q2(l - 2, m), cscl(-0.5, cadd(Q0(l - 2, m), Q1(l - 2, m))));
        acc = cadd(acc, cscl(lz(l, m) * lz(l - 1, m), cz));
        acc = cadd(acc, cscl(0.5 * lz(l, m) * lp(l - 1, m), cmi(Q4(l - 2, m + 1), Q5(l - 2, m + 1))));
        acc = cadd(acc, cscl(0.5 * lz(l, m) * lmn(l - 1, m), cpi(Q4(l - 2, m - 1), Q5(l - 2, m - 1))));
    }
    return acc;
}

// Pass 2: one thread per (box, coefficient): B, C1, C2 from the moments,
// conjugated and scaled by side^-l, added to the box's expansions.
__global__ void k_mind_p2m_combine(MindArgs a, const int *__restrict__ boxes, int nbx, const double2 *__restrict__ G)
{
    int p = a.p, nown = (p + 1) * (p + 2) / 2;
    size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)nbx * nown) return;
    int i = (int)(gid / nown), idx = (int)(gid % nown);
    int b = boxes[i];
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= idx) l++;
    int m = idx - l * (l + 1) / 2;
    const double2 *Gb = G + (size_t)i * MCOLS * nown;
    double side = a.tsize / (double)(1 << a.level[b]);
    double al = (a.lam + a.mu) / (a.lam + 2 * a.mu), lam = a.lam, mu = a.mu;
    auto zero = [](int, int) { return make_double2(0.0, 0.0); };
    auto lin = [&](double w0, int c0) { return [=](int ll, int mm) { return cscl(w0, gmom(Gb, nown, c0, ll, mm)); }; };
    // B: dipole side * sigma_sl; quadrupole of the double layer
    double rf = 2.0 * mu * side;
    auto bq2 = [=](int ll, int mm) {
        double2 t = cadd(cadd(gmom(Gb, nown, 6, ll, mm), gmom(Gb, nown, 7, ll, mm)), gmom(Gb, nown, 8, ll, mm));
        return cadd(cscl(-rf, gmom(Gb, nown, 8, ll, mm)), cscl(-2.0 * lam * side, t));
    };
    double2 vb = apply_moments(l, m, zero, lin(side, 0), lin(side, 1), lin(side, 2), lin(rf, 6), lin(rf, 7), bq2,
                               lin(rf, 9), zero, zero);
    // C1: charge -(2-al) s3, dipole -al y3 (s1, s2, -s3) and the double
    // layer's (-r2 q5, -r2 q6, 2 r2 q3 - 2 lam (al-1) tr q), quadrupole
    // -2 al mu y3 (q1, q2, q3, q4, -q5, -q6)
    double r2 = (2.0 - 2.0 * al) * mu, c1q = 2.0 * al * mu;
    auto c1x = [=](int ll, int mm) {
        return cadd(cscl(-al, gmom(Gb, nown, 3, ll, mm)), cscl(-r2, gmom(Gb, nown, 10, ll, mm)));
    };
    auto c1y = [=](int ll, int mm) {
        return cadd(cscl(-al, gmom(Gb, nown, 4, ll, mm)), cscl(-r2, gmom(Gb, nown, 11, ll, mm)));
    };
    auto c1z = [=](int ll, int mm) {
        double2 t = cadd(cadd(gmom(Gb, nown, 6, ll, mm), gmom(Gb, nown, 7, ll, mm)), gmom(Gb, nown, 8, ll, mm));
        return cadd(cadd(cscl(al, gmom(Gb, nown, 5, ll, mm)), cscl(2.0 * r2, gmom(Gb, nown, 8, ll, mm))),
                    cscl(-2.0 * lam * (al - 1.0), t));
    };
    double2 vc1 = apply_moments(l, m, lin(-(2.0 - al), 2), c1x, c1y, c1z, lin(-c1q, 12), lin(-c1q, 13),
                                lin(-c1q, 14), lin(-c1q, 15), lin(c1q, 16), lin(c1q, 17));
    // C2: dipole -(2-al) (s1, s2, -s3); quadrupole -r3 (q1, q2, q3, q4, -q5, -q6)
    double r3 = 2.0 * (2.0 - al) * mu;
    double2 vc2 = apply_moments(l, m, zero, lin(-(2.0 - al), 0), lin(-(2.0 - al), 1), lin(2.0 - al, 2), lin(-r3, 6),
                                lin(-r3, 7), lin(-r3, 8), lin(-r3, 9), lin(r3, 10), lin(r3, 11));
    double sc = pow(side, -(double)l);
    double2 v[3] = {vb, vc1, vc2};
    for (int e = 0; e < 3; e++) {
        double *o = a.mp + ((size_t)b * 4 + e) * a.nc + 2 * idx;
        o[0] += v[e].x * sc;
        o[1] -= v[e].y * sc;                 // the multipole takes conj(D R)
    }
}

// B of boxes with sources and kids times 2 after their M2M (mprstemp with
// the ratio of the children's and the parent's scales,
// src/lfmm3droutsmindlin.f:206-210).
__global__ void k_mind_bscale(const int *__restrict__ boxes, int n, int nc, double *mp)
{
    int i = blockIdx.x;
    if (i >= n) return;
    double *o = mp + (size_t)boxes[i] * 4 * nc;
    for (int k = threadIdx.x; k < nc; k += blockDim.x) o[k] *= 2.0;
}


// Per source box of level >= 2: the outgoing plane waves of B, C1, C2
// (mpoletoexp + ftophys; B scaled by (1-alpha)/alpha * side / lambda^2),
// B's proxy charges (mptoslpborder) and B replaced by its remainder after
// the proxies' multipole, integrated twice in z and times
// (1-alpha)/alpha * side, the form the near list-2 M2L uses
// (src/lfmm3droutsmindlin.f:257-279, mindlinbmploc src/mindlinmpeval.f:320).
// pwo: [slot][3][nexptotp] complex; skel: [slot][2 nq][4] (x, y, z, q);
// ycst: ylgndr(p, 1/sqrt(3)), (p+1)^2.
__global__ __launch_bounds__(MW) void k_mind_bmod(MindArgs a, MindPw pw, const int *__restrict__ boxes,
                                                  const int *__restrict__ slot, const double *__restrict__ ycst,
                                                  int nq, double *__restrict__ pwo, double *__restrict__ skel)
{
    constexpr int NBUF = 3 * MPW_FOUR > 2 * MH ? 3 * MPW_FOUR : 2 * MH;
    __shared__ double2 H[MH];
    __shared__ double2 buf[NBUF];                         // Fourier modes, later the rings' harmonics
    __shared__ double2 ch[2][MP_MAX + 1];
    __shared__ double2 S[2][MP_MAX + 1];                 // per ring: sum_j q_j e^{-i m phi_j}
    __shared__ double isf[2 * MP_MAX + 2];                // 1 / sqrt(k!)
    double2(*F)[MPW_FOUR] = (double2(*)[MPW_FOUR])buf;   // [3][MPW_FOUR]
    double2(*A)[MH] = (double2(*)[MH])buf;               // [2][MH]: per ring, R_l^m at phi = 0
    int b = boxes[blockIdx.x];
    int p = a.p, nown = (p + 1) * (p + 2) / 2;
    double side = a.tsize / (double)(1 << a.level[b]);
    const double *c = a.centers + 3 * (size_t)b;
    double al = (a.lam + a.mu) / (a.lam + 2 * a.mu);
    double *Mb = a.mp + (size_t)b * 4 * a.nc;
    double *po = pwo + (size_t)slot[b] * 3 * 2 * pw.nexptotp;
    if (threadIdx.x == 0) {
        double f = 1.0;
        isf[0] = 1.0;
        for (int k = 1; k <= 2 * p + 1; k++) {
            f *= sqrt((double)k);
            isf[k] = 1.0 / f;
        }
    }
    __syncthreads();
    // Fourier modes of the three expansions: f_{nl,m} = (-i)^m sum_n (-lam)^n / sqrt((n-m)!(n+m)!) M_n^m
    for (int t = threadIdx.x; t < 3 * pw.nexptot; t += MW) {
        int e = t / pw.nexptot, f = t % pw.nexptot;
        int nl = 0;
        while (nl + 1 < pw.nlams && pw.foff[nl + 1] <= f) nl++;
        int mth = f - pw.foff[nl];
        double lam = pw.rlams[nl];
        double sr = 0, si = 0;
        const double *M = Mb + (size_t)e * a.nc;
        double cc = pow(-lam, (double)mth) * isf[2 * mth];
        for (int n = mth; n <= p; n++) {
            const double *Mm = M + 2 * pk(n, mth);
            sr += cc * Mm[0];
            si += cc * Mm[1];
            cc *= -lam / sqrt((n + 1.0 - mth) * (n + 1.0 + mth));
        }
        double zr, zi;
        switch (mth & 3) {
        case 0: zr = 1; zi = 0; break;
        case 1: zr = 0; zi = -1; break;
        case 2: zr = -1; zi = 0; break;
        default: zr = 0; zi = 1; break;
        }
        F[e][f] = make_double2(sr * zr - si * zi, sr * zi + si * zr);
    }
    __syncthreads();
    // physical samples: f_0 + 2 Re-like sums over the modes with the sign
    // alternation of ftophys, e^{i m alpha} by rotation
    double scb = (1.0 - al) / al * side;
    for (int t = threadIdx.x; t < 3 * pw.nexptotp; t += MW) {
        int e = t / pw.nexptotp, j = t % pw.nexptotp;
        int nl = pw.jnl[j];
        int np = pw.nphys[nl], nf = pw.nfour[nl];
        double alpha = (j - pw.poff[nl]) * (2 * M_PI / np);
        const double2 *f = &F[e][pw.foff[nl]];
        double pr = f[0].x, pi = f[0].y;
        double sgn = -1.0;
        double c1, s1;
        sincos(alpha, &s1, &c1);
        double cr = c1, ci = s1;
        for (int mm = 1; mm < nf; mm++) {
            double tr = cr * f[mm].x - ci * f[mm].y, ti = cr * f[mm].y + ci * f[mm].x;
            pr += tr + sgn * tr;
            pi += ti - sgn * ti;
            sgn = -sgn;
            double nr = cr * c1 - ci * s1;
            ci = cr * s1 + ci * c1;
            cr = nr;
        }
        if (e == 0) {
            double z = 1.0 / (pw.rlams[nl] * pw.rlams[nl]);
            pr = pr * z * scb;
            pi = pi * z * scb;
        }
        po[2 * ((size_t)e * pw.nexptotp + j)] = pr;
        po[2 * ((size_t)e * pw.nexptotp + j) + 1] = pi;
    }
    // proxy rings of B's sectoral terms
    const double ct = 1.0 / sqrt(3.0), st = sqrt(2.0 / 3.0);
    double radius = side * sqrt(3.0) / 2;
    if ((int)threadIdx.x <= p) {
        int j = threadIdx.x;
        double rfac = 1.0 / (radius / side);           // 1 / (radius * rscale)
        double rmulj = pow(rfac, (double)j), rmulj1 = rmulj * rfac;
        const double *Mjj = Mb + 2 * pk(j, j);
        double d = ycst[j * (p + 1) + j] * st;
        double f1r = Mjj[0] / d * rmulj * sqrt(2 * j + 1.0), f1i = Mjj[1] / d * rmulj * sqrt(2 * j + 1.0);
        double f2r = 0, f2i = 0;
        if (j + 1 <= p) {
            const double *M2 = Mb + 2 * pk(j + 1, j);
            double d2 = ycst[(j + 1) * (p + 1) + j] * st;
            f2r = M2[0] / d2 * rmulj1 * sqrt(2 * (j + 1) + 1.0);
            f2i = M2[1] / d2 * rmulj1 * sqrt(2 * (j + 1) + 1.0);
        }
        ch[0][j] = make_double2((f1r + f2r) / 2, (f1i + f2i) / 2);
        ch[1][j] = make_double2((f1r - f2r) / 2, (f1i - f2i) / 2);
    }
    __syncthreads();
    double *sk = skel + (size_t)slot[b] * 2 * nq * 4;
    for (int j = threadIdx.x; j < 2 * nq; j += MW) {
        int ring = j / nq, jj = j % nq;
        double phi = jj * (2 * M_PI / nq);
        double c1, s1;
        sincos(phi, &s1, &c1);
        double cr = c1, ci = s1;
        double s = ch[ring][0].x;
        for (int k = 1; k <= p; k++) {
            s += 2 * (ch[ring][k].x * cr - ch[ring][k].y * ci);
            double nr = cr * c1 - ci * s1;
            ci = cr * s1 + ci * c1;
            cr = nr;
        }
        sk[4 * j] = c[0] + radius * st * cos(phi);
        sk[4 * j + 1] = c[1] + radius * st * sin(phi);
        sk[4 * j + 2] = c[2] + (ring ? -radius * ct : radius * ct);
        sk[4 * j + 3] = s * st / nq;
    }
    __threadfence_block();
    __syncthreads();
    // multipole of the proxy charges: a ring's charges share R_l^m up to
    // e^{i m phi_j}, so M_l^m = sum_rings R_l^m(ring, phi = 0) S_ring(m)
    // with S_ring(m) = sum_j q_j e^{-i m phi_j}
    for (int t = threadIdx.x; t < 2 * (p + 1); t += MW) {
        int ring = t / (p + 1), m = t % (p + 1);
        double sr = 0, si = 0;
        for (int jj = 0; jj < nq; jj++) {
            double q = sk[4 * (ring * nq + jj) + 3];
            double sn, cs;
            sincos(m * (jj * (2 * M_PI / nq)), &sn, &cs);
            sr += q * cs;
            si -= q * sn;
        }
        S[ring][m] = make_double2(sr, si);
    }
    harm_regular(p, radius * st, 0.0, radius * ct, A[0]);
    harm_regular(p, radius * st, 0.0, -radius * ct, A[1]);
    __syncthreads();
    // remainder of B (scaled proxies' multipole subtracted), dz^-2, scale
    for (int k = 0; k < 2; k++) {
        int idx = threadIdx.x + k * MW;
        if (idx >= nown) break;
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= idx) l++;
        int m = idx - l * (l + 1) / 2;
        double2 mp = make_double2(0.0, 0.0);
        for (int ring = 0; ring < 2; ring++) {
            double2 r = A[ring][idx], sv = S[ring][m];
            // conj(R_l^m(x_j)) q_j = conj(A e^{i m phi_j}) q_j
            mp.x += r.x * sv.x + r.y * sv.y;
            mp.y += r.x * sv.y - r.y * sv.x;
        }
        double sc = pow(side, -(double)l);
        H[idx] = make_double2(Mb[2 * idx] - mp.x * sc, Mb[2 * idx + 1] - mp.y * sc);
    }
    __syncthreads();
    for (int k = 0; k < 2; k++) {
        int idx = threadIdx.x + k * MW;
        if (idx >= nown) break;
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= idx) l++;
        int m = idx - l * (l + 1) / 2;
        double vr = 0, vi = 0;
        if (l <= p - 2) {
            int i = l + 2;
            double rs = 1.0 / sqrt((i - m) * (i - m - 1.0) * (i + m) * (i + m - 1.0));
            double2 v = H[pk(i, m)];
            vr = v.x * rs * scb;
            vi = v.y * rs * scb;
        }
        Mb[2 * idx] = vr;
        Mb[2 * idx + 1] = vi;
    }
}

// Per target box: the U-field of the proxy rings of each near list-2 source
// box, tabulated on the sphere about the target centre and projected on
// the local expansion of B (l3dformtaintkerbf90sc, src/mindlinformta.f:980).
// gl: Gauss-Legendre nodes (nquad) then weights (nquad); ytab: ylgndr at the
// nodes, [nquad][(p+1)^2]; eph: e^{-i m 2 pi (k+1)/nquad}, [m][nquad].
__global__ __launch_bounds__(MPX) void k_mind_proxy(MindArgs a, const int *__restrict__ tboxes,
                                                   const int *__restrict__ tstart, const int *__restrict__ titems,
                                                   const int *__restrict__ slot, const double *__restrict__ skel,
                                                   int nq, int nquad, const double *__restrict__ gl,
                                                   const double *__restrict__ ytab, const double2 *__restrict__ eph)
{
    extern __shared__ double smem[];
    double *phi = smem;                                   // nquad * nquad
    double2 *marr = (double2 *)(smem + nquad * nquad);    // nquad * (p + 1)
    double *sk = (double *)(marr + nquad * (a.p + 1));    // 2 nq * 4
    int t = tboxes[blockIdx.x];
    int p = a.p, nown = (p + 1) * (p + 2) / 2;
    double side = a.tsize / (double)(1 << a.level[t]);
    const double *c = a.centers + 3 * (size_t)t;
    double al = (a.lam + a.mu) / (a.lam + 2 * a.mu);
    double radius = side * sqrt(3.0) / 2;
    double rs = side / radius;                            // 1 / (rscale * radius)
    double2 acc[2] = {make_double2(0, 0), make_double2(0, 0)};
    for (int it = tstart[blockIdx.x]; it < tstart[blockIdx.x + 1]; it++) {
        int s = titems[it];
        double sside = a.tsize / (double)(1 << a.level[s]);
        double sc = (1.0 - al) / al / sside;              // times the source box's scale
        const double *ks = skel + (size_t)slot[s] * 2 * nq * 4;
        __syncthreads();
        for (int k = threadIdx.x; k < 2 * nq * 4; k += MPX) sk[k] = ks[k];
        __syncthreads();
        for (int g = threadIdx.x; g < nquad * nquad; g += MPX) {
            int jj = g / nquad, kk = g % nquad;
            double cth = gl[jj], sth = sqrt(1.0 - cth * cth);
            double ph = 2 * M_PI * (kk + 1) / nquad;
            double tx = c[0] + radius * sth * cos(ph), ty = c[1] + radius * sth * sin(ph), tz = c[2] + radius * cth;
            double v = 0.0;
            for (int i = 0; i < 2 * nq; i++) {
                double r1 = tx - sk[4 * i], r2 = ty - sk[4 * i + 1], r3 = tz - sk[4 * i + 2];
                double rr = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
                v += sk[4 * i + 3] * (-r3 * log(rr - r3) - rr);
            }
            phi[g] = v * sc;
        }
        __syncthreads();
        for (int g = threadIdx.x; g < nquad * (p + 1); g += MPX) {
            int jj = g / (p + 1), m = g % (p + 1);
            double sr = 0, si = 0;
            for (int kk = 0; kk < nquad; kk++) {
                double2 e = eph[m * nquad + kk];
                sr += phi[jj * nquad + kk] * e.x;
                si += phi[jj * nquad + kk] * e.y;
            }
            marr[jj * (p + 1) + m] = make_double2(sr / (2 * nquad) * gl[nquad + jj], si / (2 * nquad) * gl[nquad + jj]);
        }
        __syncthreads();
        for (int k = 0; k < 2; k++) {
            int idx = threadIdx.x + k * MPX;
            if (idx >= nown) break;
            int l = 0;
            while ((l + 1) * (l + 2) / 2 <= idx) l++;
            int m = idx - l * (l + 1) / 2;
            double f0 = sqrt(2 * l + 1.0) * pow(rs, (double)l);
            double sr = 0, si = 0;
            for (int jj = 0; jj < nquad; jj++) {
                double y = ytab[(size_t)jj * (p + 1) * (p + 1) + l * (p + 1) + m];
                double2 mv = marr[jj * (p + 1) + m];
                sr += mv.x * y;
                si += mv.y * y;
            }
            acc[k].x += sr * f0;
            acc[k].y += si * f0;
        }
    }
    for (int k = 0; k < 2; k++) {
        int idx = threadIdx.x + k * MPX;
        if (idx >= nown) break;
        double *o = a.lo + (size_t)t * 4 * a.nc + 2 * idx;
        o[0] += acc[k].x;
        o[1] += acc[k].y;
    }
}

// Per target box: the plane waves of its far list-2 sources shifted and
// summed (xyzshift_add; a box that also has sources starts from its own
// outgoing waves, as the reference's shared array does), then back to
// Fourier modes (phystof) and to the local expansions of B, C1, C2
// (exptolocald_add). wl: per (lambda, degree) whts * lam^n, [nl][p+1].
__global__ __launch_bounds__(MW) void k_mind_far(MindArgs a, MindPw pw, const int *__restrict__ tboxes,
                                                 const int *__restrict__ tstart, const int *__restrict__ titems,
                                                 const int *__restrict__ slot, const double *__restrict__ pwo,
                                                 const double *__restrict__ wl)
{
    extern __shared__ double smem[];
    double2 *acc = (double2 *)smem;                       // [3][nexptotp]
    double2 *F = acc + 3 * pw.nexptotp;                   // [3][nexptot]
    int t = tboxes[blockIdx.x];
    int p = a.p, nown = (p + 1) * (p + 2) / 2;
    double side = a.tsize / (double)(1 << a.level[t]);
    const double *c = a.centers + 3 * (size_t)t;
    int own = slot[t];
    for (int k = threadIdx.x; k < 3 * pw.nexptotp; k += MW) {
        if (own >= 0) {
            const double *s = pwo + (size_t)own * 3 * 2 * pw.nexptotp + 2 * k;
            acc[k] = make_double2(s[0], s[1]);
        } else {
            acc[k] = make_double2(0.0, 0.0);
        }
    }
    for (int it = tstart[blockIdx.x]; it < tstart[blockIdx.x + 1]; it++) {
        int s = titems[it];
        double sside = a.tsize / (double)(1 << a.level[s]);
        const double *cs = a.centers + 3 * (size_t)s;
        double xd = (c[0] - cs[0]) / sside, yd = (c[1] - cs[1]) / sside, zd = (c[2] - cs[2]) / sside;
        const double *src = pwo + (size_t)slot[s] * 3 * 2 * pw.nexptotp;
        for (int j = threadIdx.x; j < pw.nexptotp; j += MW) {
            const double *tb = pw.tab + 4 * j;
            double lam = tb[0];
            double r0 = exp(lam * zd);
            double ph = lam * (xd * tb[1] + yd * tb[2]);
            double mr = r0 * cos(ph), mi = r0 * sin(ph);
            for (int e = 0; e < 3; e++) {
                double x = src[2 * (e * pw.nexptotp + j)], y = src[2 * (e * pw.nexptotp + j) + 1];
                double2 &o = acc[e * pw.nexptotp + j];
                o.x += x * mr - y * mi;
                o.y += x * mi + y * mr;
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 3 * pw.nexptot; k += MW) {
        int e = k / pw.nexptot, f = k % pw.nexptot;
        int nl = 0;
        while (nl + 1 < pw.nlams && pw.foff[nl + 1] <= f) nl++;
        int m = f - pw.foff[nl], np = pw.nphys[nl];
        double fr = 0, fi = 0;
        for (int j = 0; j < np; j++) {
            double2 v = acc[e * pw.nexptotp + pw.poff[nl] + j];
            if (m == 0) {
                fr += v.x;
                fi += v.y;
            } else {
                double al = j * (2 * M_PI / np);
                double cr = cos(m * al), ci = -sin(m * al);
                fr += cr * v.x - ci * v.y;
                fi += cr * v.y + ci * v.x;
            }
        }
        F[k] = make_double2(fr / np, fi / np);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 3 * nown; k += MW) {
        int e = k / nown, idx = k % nown;
        int n = 0;
        while ((n + 1) * (n + 2) / 2 <= idx) n++;
        int m = idx - n * (n + 1) / 2;
        double ar = 0, ai = 0;
        for (int nl = 0; nl < pw.nlams; nl++) {
            if (m > pw.nfour[nl] - 1) continue;
            double2 f = F[e * pw.nexptot + pw.foff[nl] + m];
            double r = wl[nl * (p + 1) + n];
            ar += r * f.x;
            ai += r * f.y;
        }
        double zr, zi;
        switch (m & 3) {
        case 0: zr = 1; zi = 0; break;
        case 1: zr = 0; zi = -1; break;
        case 2: zr = -1; zi = 0; break;
        default: zr = 0; zi = 1; break;
        }
        double sc = 1.0 / (sqrt_fact(n - m) * sqrt_fact(n + m)) / side;
        double *o = a.lo + ((size_t)t * 4 + e) * a.nc + 2 * idx;
        o[0] += (ar * zr - ai * zi) * sc;
        // The m = 0 coefficients come out complex. Their imaginary parts
        // carry only the imaginary part of the potential, which the
        // reference drops when it returns the real displacement and strain;
        // the packed (m >= 0) storage assumes a real potential, so they are
        // dropped here.
        if (m > 0) o[1] += (ar * zi + ai * zr) * sc;
    }
}

// Targets: displacement and strain of the Mindlin harmonic part from the
// L2T fields of B, C1, C2 (lap: per point 4 x (pot, grad, xx yy zz xy xz yz))
// (mindlintaevalb/c_trunc, src/mindlinformta.f:830-978; combined as
// src/lfmm3dtriamindlin_iter_1.f:1130-1169 and added as
// src/elhfmm3dtria_iter.f:806-831: displacement + pot/mu, strain -
// sym(field)/(2 mu)).
__global__ void k_mind_assemble(int n, int nt, const int *__restrict__ izt, const double *__restrict__ pts,
                                const double *__restrict__ lap, double mu, int ifpt, double *__restrict__ ptt,
                                int ifst, double *__restrict__ stt)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nt) return;
    int ip = n + j;
    const double *f = lap + (size_t)ip * 40;
    double z = pts[3 * (size_t)ip + 2];
    // B: fld = (d1 pot, d2 pot, -d3 pot)
    const double *B = f, *C1 = f + 10, *C2 = f + 20;
    double fb[3] = {B[1], B[2], -B[3]};
    double hb[6] = {B[4], B[5], B[6], B[7], B[8], B[9]};   // xx yy zz xy xz yz
    double gb[3][3] = {{-hb[0], -hb[3], hb[4]}, {-hb[3], -hb[1], hb[5]}, {-hb[4], -hb[5], hb[2]}};
    // C: fld = grad pot1 - (0, 0, pot2)
    double fc[3] = {C1[1], C1[2], C1[3] - C2[0]};
    double hc[6] = {C1[4], C1[5], C1[6], C1[7], C1[8], C1[9]};
    double f2[3] = {-C2[1], -C2[2], -C2[3]};               // -grad pot2
    double gc[3][3] = {{-hc[0], -hc[3], -hc[4] - f2[0]},
                       {-hc[3], -hc[1], -hc[5] - f2[1]},
                       {-hc[4], -hc[5], -hc[2] - f2[2]}};
    double u[3], H[3][3];
    for (int k = 0; k < 3; k++) u[k] = fb[k] + z * fc[k];
    for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) H[k][l] = gb[k][l] + z * gc[k][l];
    for (int l = 0; l < 3; l++) H[2][l] -= fc[l];
    int o = izt[j];
    if (ifpt)
        for (int k = 0; k < 3; k++) ptt[3 * (size_t)o + k] += u[k] / mu;
    if (ifst)
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++) stt[9 * (size_t)o + 3 * l + k] -= (H[k][l] + H[l][k]) / (2 * mu);
}

// ---- launchers ----

// LDS of the moment pass: nch nodes of (harmonics, 2 weights, 9 triangle
// values, 3 coordinates, triangle index) and the recurrence tables
size_t mind_p2m_lds(int p, int nch)
{
    size_t nown = (size_t)(p + 1) * (p + 2) / 2;
    return (size_t)nch * (nown * 16 + (2 + 9 + 3) * 8 + 4) + nown * 16 + (size_t)(p + 1) * 8;
}

size_t mind_p2m_nch(int p)
{
    size_t nown = (size_t)(p + 1) * (p + 2) / 2, per = nown * 16 + (2 + 9 + 3) * 8 + 4;
    size_t fixed = nown * 16 + (size_t)(p + 1) * 8;
    return std::min<size_t>(32, (MIND_P2M_LDS - fixed) / per);
}

size_t mind_p2m_scratch_bytes(int p, int nbx)
{
    return (size_t)nbx * MCOLS * ((size_t)(p + 1) * (p + 2) / 2) * 16;
}

hipError_t launch_mind_p2m(const MindDev &d, const int *boxes, int nbx, hipStream_t st)
{
    if (nbx <= 0) return hipSuccess;
    const MindArgs &a = d.args;
    int nown = (a.p + 1) * (a.p + 2) / 2;
    int nch = (int)mind_p2m_nch(a.p);
    size_t sh = mind_p2m_lds(a.p, nch);
    if (nown <= 512)
        hipLaunchKernelGGL(k_mind_moments<512>, dim3(nbx), dim3(512), sh, st, a, boxes, nch, (double2 *)d.gmom);
    else
        hipLaunchKernelGGL(k_mind_moments<1024>, dim3(nbx), dim3(1024), sh, st, a, boxes, nch, (double2 *)d.gmom);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t nth = (size_t)nbx * nown;
    hipLaunchKernelGGL(k_mind_p2m_combine, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0, st, a, boxes, nbx,
                       (const double2 *)d.gmom);
    return hipGetLastError();
}

hipError_t launch_mind_bscale(const MindDev &d, const int *boxes, int nbx, hipStream_t st)
{
    if (nbx <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mind_bscale, dim3(nbx), dim3(256), 0, st, boxes, nbx, d.args.nc, d.args.mp);
    return hipGetLastError();
}

hipError_t launch_mind_bmod(const MindDev &d, const int *boxes, int nbx, hipStream_t st)
{
    if (nbx <= 0) return hipSuccess;
    const MindArgs &a = d.args;
    const MindPw &pw = d.pw;
    hipLaunchKernelGGL(k_mind_bmod, dim3(nbx), dim3(MW), 0, st, a, pw, boxes, d.slot, d.ycst, d.nq, d.pwo, d.skel);
    return hipGetLastError();
}

hipError_t launch_mind_proxy(const MindDev &d, const int *tb, const int *ts, const int *ti, int nbx, hipStream_t st)
{
    if (nbx <= 0) return hipSuccess;
    const MindArgs &a = d.args;
    size_t sh = (size_t)d.nquad * d.nquad * 8 + (size_t)d.nquad * (a.p + 1) * 16 + (size_t)2 * d.nq * 4 * 8;
    hipLaunchKernelGGL(k_mind_proxy, dim3(nbx), dim3(MPX), sh, st, a, tb, ts, ti, d.slot, d.skel, d.nq, d.nquad, d.gl,
                       d.ytab, (const double2 *)d.eph);
    return hipGetLastError();
}

hipError_t launch_mind_far(const MindDev &d, const int *tb, const int *ts, const int *ti, int nbx, hipStream_t st)
{
    if (nbx <= 0) return hipSuccess;
    const MindArgs &a = d.args;
    const MindPw &pw = d.pw;
    size_t sh = (size_t)3 * (pw.nexptotp + pw.nexptot) * 16;
    hipLaunchKernelGGL(k_mind_far, dim3(nbx), dim3(MW), sh, st, a, pw, tb, ts, ti, d.slot, d.pwo, d.wl);
    return hipGetLastError();
}

hipError_t launch_mind_assemble(int n, int nt, const int *izt, const double *pts, const double *lap, double mu,
                                int ifpt, double *ptt, int ifst, double *stt, hipStream_t st)
{
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mind_assemble, dim3((nt + 255) / 256), dim3(256), 0, st, n, nt, izt, pts, lap, mu, ifpt, ptt,
                       ifst, stt);
    return hipGetLastError();
}

size_t mind_proxy_lds(int p, int nquad, int nq)
{
    return (size_t)nquad * nquad * 8 + (size_t)nquad * (p + 1) * 16 + (size_t)2 * nq * 4 * 8;
}

size_t mind_far_lds(int nexptotp, int nexptot) { return (size_t)3 * (nexptotp + nexptot) * 16; }

void mind_set_lds_limits(size_t proxy, size_t far)
{
    (void)hipFuncSetAttribute((const void *)k_mind_proxy, hipFuncAttributeMaxDynamicSharedMemorySize, (int)proxy);
    (void)hipFuncSetAttribute((const void *)k_mind_far, hipFuncAttributeMaxDynamicSharedMemorySize, (int)far);
    (void)hipFuncSetAttribute((const void *)k_mind_moments<512>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)MIND_P2M_LDS);
    (void)hipFuncSetAttribute((const void *)k_mind_moments<1024>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)MIND_P2M_LDS);
    (void)hipGetLastError();
}

}  // namespace elfmm
