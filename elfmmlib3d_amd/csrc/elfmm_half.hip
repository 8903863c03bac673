// Half-space (Mindlin) image near field on gfx950: the adaptive triangle
// quadrature of the image kernels (elfmm_image.h) as a work queue of
// integrals, one wave per integral.
//
// Work queue: a persistent grid of one-wave workgroups takes integrals from
// an atomic counter until the queue is empty. An integral is the reference's
// tria3adam (src/tria3adamar.f:16-424) run by one wave: each subdivision
// step evaluates the four children's 48 quadrature nodes on 48 lanes (the
// first step also the parent's 12), sums each child's nodes in the
// reference's order, decides acceptance on the uniform values and pushes or
// pops the wave's explicit stack, which lives in device memory (200 entries,
// the reference's bound; depth 200 and 100000 subdivisions bound the work).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "elfmm_image.h"
#include "elfmm_half.h"
#include "elfmm_internal.h"

namespace elfmm {

__constant__ double c_rule6[12][3] = {ELFMM_RULE6};

// LDS of one wave: node values (parent + 4 children) and per-triangle sums.
struct AdapLds {
    double fv[60][13];
    double cs[5][12];
};

// One integral by the calling wave; every lane returns the 12 results.
// stk: this wave's stack, ADAP_STACK entries of ADAP_ENTRY doubles.
__device__ void wave_adap(const AdapTask &T, double *__restrict__ stk, AdapLds &L, double rint[12])
{
    const int lane = threadIdx.x & 63;
    const double area3 = adap_area(T.tri);
    const double z0[6] = {0, 0, 1, 0, 0, 1};
    const int c = lane / 12, qn = lane % 12;
    // nodes of child c (lanes 0..47) of z, or of z itself (lanes 48..59 when
    // with_parent): weighted values into L.fv[lane]
    auto eval_nodes = [&](const double *z, bool with_parent) {
        if (lane < 48 || (with_parent && lane < 60)) {
            double zz[6];
            if (lane < 48) adap_child(z, c, zz);
            else
                for (int k = 0; k < 6; k++) zz[k] = z[k];
            double X[3], w, f[12];
            adap_node(T.tri, zz, c_rule6[qn][0], c_rule6[qn][1], c_rule6[qn][2], X, w);
            adap_eval(T, X, f);
            {
#pragma clang fp contract(off)
                for (int k = 0; k < 12; k++) L.fv[lane][k] = f[k] * w * 2 * area3;
            }
        }
        __syncthreads();
        if (lane < (with_parent ? 60 : 48)) {
#pragma clang fp contract(off)
            int tr = lane / 12, o = lane % 12;
            double s = 0.0;
            for (int q = 0; q < 12; q++) s = s + L.fv[tr * 12 + q][o];
            L.cs[tr][o] = s;
        }
        __syncthreads();
    };
    eval_nodes(z0, true);
    double top_z[6], top_v[12];
    int depth = 1;
    for (int k = 0; k < 6; k++) top_z[k] = z0[k];
    for (int k = 0; k < 12; k++) {
        top_v[k] = L.cs[4][k];
        rint[k] = 0.0;
    }
    int j = 0;                        // stack top; entry j is (top_z, top_v, depth)
    bool have_children = true;        // L.cs[0..3] hold the children of the top entry
    for (int it = 0; it < ADAP_MAXIT; it++) {
        if (!have_children) {
            const double *e = stk + (size_t)j * ADAP_ENTRY;
            for (int k = 0; k < 6; k++) top_z[k] = e[k];
            for (int k = 0; k < 12; k++) top_v[k] = e[6 + k];
            depth = (int)e[18];
            if (depth >= ADAP_NREC) {     // deepest level: take the value as it is
                for (int k = 0; k < 12; k++) rint[k] += top_v[k];
                if (--j < 0) break;
                continue;
            }
            eval_nodes(top_z, false);
        }
        have_children = false;
        double dmax = 0.0;
        for (int k = 0; k < 12; k++) dmax = fmax(dmax, fabs(top_v[k]));
        bool done = true;
        {
#pragma clang fp contract(off)
            for (int k = 0; k < 12; k++)
                if (fabs(L.cs[0][k] + L.cs[1][k] + L.cs[2][k] + L.cs[3][k] - top_v[k]) > ADAP_EPS * dmax)
                    done = false;
        }
        if (done) {
            {
#pragma clang fp contract(off)
                for (int k = 0; k < 12; k++) rint[k] = rint[k] + L.cs[0][k] + L.cs[1][k] + L.cs[2][k] + L.cs[3][k];
            }
            if (--j < 0) break;
            continue;
        }
        // push the four children at j .. j + 3 (the last on top)
        if (lane < 48) {
            int o = lane % 12;
            double *e = stk + (size_t)(j + c) * ADAP_ENTRY;
            e[6 + o] = L.cs[c][o];
            if (o < 6) {
                double zz[6];
                adap_child(top_z, c, zz);
                e[o] = zz[o];
            } else if (o == 6) {
                e[18] = (double)(depth + 1);
            }
        }
        __threadfence_block();
        __syncthreads();
        j += 3;
        if (j + 4 > ADAP_STACK) break;    // the reference stops here too (ier = 8)
    }
    __syncthreads();                      // L is reused by the next integral
}

// Take the next item of the queue; uniform across the wave.
__device__ __forceinline__ int64_t queue_next(unsigned long long *ctr)
{
    __shared__ unsigned long long tk;
    if ((threadIdx.x & 63) == 0) tk = atomicAdd(ctr, 1ull);
    __syncthreads();
    unsigned long long t = tk;
    __syncthreads();
    return (int64_t)t;
}

// Triangle (tree order) at position t of near box q's concatenated list.
__device__ __forceinline__ int near_tri(int q, int t, const int *__restrict__ lstart, const int *__restrict__ litems,
                                        const int *__restrict__ src_start, const int *__restrict__ src_count)
{
    for (int bb = lstart[q]; bb < lstart[q + 1]; bb++) {
        int sbox = litems[bb], sn = src_count[sbox];
        if (t < sn) return src_start[sbox] + t;
        t -= sn;
    }
    return -1;
}

__device__ __forceinline__ void fill_task(AdapTask &T, int i, const double *__restrict__ tri_s,
                                          const double *__restrict__ nrm_s, const double *__restrict__ dens,
                                          int unit, int layer, const double *y, double lam, double mu, int ifpt,
                                          int ifst)
{
    for (int k = 0; k < 9; k++) T.tri[k] = tri_s[9 * (size_t)i + k];
    for (int k = 0; k < 3; k++) {
        T.nv[k] = nrm_s[3 * (size_t)i + k];
        T.dens[k] = unit >= 0 ? (k == unit ? 1.0 : 0.0) : dens[3 * (size_t)i + k];
        T.y[k] = y[k];
    }
    T.lam = lam;
    T.mu = mu;
    T.layer = layer;
    T.ifpt = ifpt;
    T.ifst = ifst;
}

// Image near field, one integral per (point, triangle, layer[, unit density])
// of the near boxes. Points are targets in tree order; pts holds them
// mirrored (the tree of the image near field is built on the mirrored
// targets), so the evaluation point is y = (p1, p2, -p3).
// units = 3: the unit densities e_k (icomp 2/5), results into the store in
// the layout of the free-space store (StoreLayout, store_idx); units = 1:
// the caller's densities (icomp 1), 9 words per (pair, layer) into tmp
// (displacement, strain xx yy zz xy xz yz) with the same tiling, W = 9 * NL.

__device__ __forceinline__ int64_t st_idx(int k, int t, int w, int ntri, int W)
{
    int t0 = t & ~63;
    int tw = ntri - t0 < 64 ? ntri - t0 : 64;
    return ((int64_t)k * ntri + t0) * W + (int64_t)w * tw + (t - t0);
}

__global__ __launch_bounds__(64) void k_hnear(HalfNearArgs H, int64_t ntask, unsigned long long *ctr,
                                              double *__restrict__ stack, double *__restrict__ out)
{
    __shared__ AdapLds L;
    const int lane = threadIdx.x;
    double *stk = stack + (size_t)blockIdx.x * ADAP_STACK * ADAP_ENTRY;
    int nl = H.ifsl + H.ifdl;
    int per = nl * H.units;
    for (;;) {
        int64_t t = queue_next(ctr);
        if (t >= ntask) break;
        int lo = 0, hi = H.nq - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (H.tb[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        int q = lo;
        int64_t loc = t - H.tb[q];
        int ntri = H.ntri_q[q];
        int64_t pair = loc / per;
        int sub = (int)(loc % per);
        int slot = sub / H.units, unit = H.units == 3 ? sub % 3 : -1;
        int k = (int)(pair / ntri), tt = (int)(pair % ntri);
        int layer = (H.ifsl && slot == 0) ? 0 : 1;
        int tbox = H.boxes[q];
        int gp = H.n + H.tgt_start[tbox] + k;
        double y[3] = {H.pts[3 * (size_t)gp], H.pts[3 * (size_t)gp + 1], -H.pts[3 * (size_t)gp + 2]};
        int i = near_tri(q, tt, H.lstart, H.litems, H.src_start, H.src_count);
        AdapTask T;
        fill_task(T, i, H.tri_s, H.nrm_s, layer ? H.sdl_s : H.ssl_s, unit, layer, y, H.lam, H.mu, H.ifpt, H.ifst);
        double r[12];
        wave_adap(T, stk, L, r);
        // r: displacement (if ifpt) then the 3x3 strain (if ifst)
        double *o = out + H.sbase[q];
        int ks = H.ifpt ? 3 : 0;
        if (H.units == 3) {
            int nf = 3 * H.ifpt + 6 * H.ifst;
            int w0 = (slot * 3 + unit) * nf;
            if (lane < nf) {
                double v;
                if (H.ifpt && lane < 3) v = r[lane];
                else {
                    const int si[6] = {0, 4, 8, 1, 2, 5};
                    v = r[ks + si[lane - (H.ifpt ? 3 : 0)]];
                }
                o[st_idx(k, tt, w0 + lane, ntri, H.W)] = v;
            }
        } else if (lane < 9) {
            double v = 0.0;
            if (lane < 3) v = H.ifpt ? r[lane] : 0.0;
            else if (H.ifst) {
                const int si[6] = {0, 4, 8, 1, 2, 5};
                v = r[ks + si[lane - 3]];
            }
            o[st_idx(k, tt, slot * 9 + lane, ntri, H.W)] = v;
        }
    }
}

// icomp 1: the near field of each point, the sum over its (pair, layer)
// words in tmp; one wave per point, as the replay.
__global__ __launch_bounds__(256) void k_hnear_sum(int n, const int *__restrict__ boxes,
                                                   const int *__restrict__ tgt_start,
                                                   const int64_t *__restrict__ sbase, const int *__restrict__ ntri_q,
                                                   const int *__restrict__ pstart, int nq, int nl,
                                                   const double *__restrict__ tmp, double *__restrict__ nf)
{
    int lane = threadIdx.x & 63;
    int g = __builtin_amdgcn_readfirstlane((int)blockIdx.x * 4 + (int)(threadIdx.x >> 6));
    if (g >= pstart[nq]) return;
    int lo = 0, hi = nq - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pstart[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    int q = lo, k = g - pstart[q], ntri = ntri_q[q], W = 9 * nl;
    const double *S = tmp + sbase[q];
    double acc[9];
    for (int c = 0; c < 9; c++) acc[c] = 0.0;
    for (int t = lane; t < ntri; t += 64)
        for (int a = 0; a < nl; a++)
            for (int c = 0; c < 9; c++) acc[c] += S[st_idx(k, t, a * 9 + c, ntri, W)];
    for (int c = 0; c < 9; c++)
        for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off);
    if (lane == 0) {
        int gp = n + tgt_start[boxes[q]] + k;
        for (int c = 0; c < 9; c++) nf[(size_t)gp * 9 + c] = acc[c];
    }
}

// Direct image sum (the image half of elh3dtriadirecttarg): every triangle
// at every point, chunks of HD_CHUNK triangles per queue item, the caller's
// densities, both layers added; partial sums per (point, chunk) in part.
constexpr int HD_CHUNK = 16;
__global__ __launch_bounds__(64) void k_hdirect(int64_t ntask, int nchunk, int n, const double *__restrict__ pts,
                                                const double *__restrict__ tri, const double *__restrict__ nrm,
                                                const double *__restrict__ ssl, const double *__restrict__ sdl,
                                                int ifsl, int ifdl, int ifpt, int ifst, double lam, double mu,
                                                unsigned long long *ctr, double *__restrict__ stack,
                                                double *__restrict__ part)
{
    __shared__ AdapLds L;
    const int lane = threadIdx.x;
    double *stk = stack + (size_t)blockIdx.x * ADAP_STACK * ADAP_ENTRY;
    for (;;) {
        int64_t t = queue_next(ctr);
        if (t >= ntask) break;
        int ip = (int)(t / nchunk), ch = (int)(t % nchunk);
        double y[3] = {pts[3 * (size_t)ip], pts[3 * (size_t)ip + 1], pts[3 * (size_t)ip + 2]};
        double acc[12];
        for (int k = 0; k < 12; k++) acc[k] = 0.0;
        int i1 = min(n, (ch + 1) * HD_CHUNK);
        for (int i = ch * HD_CHUNK; i < i1; i++)
            for (int layer = 0; layer < 2; layer++) {
                if (!(layer ? ifdl : ifsl)) continue;
                AdapTask T;
                fill_task(T, i, tri, nrm, layer ? sdl : ssl, -1, layer, y, lam, mu, ifpt, ifst);
                double r[12];
                wave_adap(T, stk, L, r);
                for (int k = 0; k < 12; k++) acc[k] += r[k];
            }
        if (lane < 12) part[(size_t)t * 12 + lane] = acc[lane];
    }
}

// part (point, chunk) sums added to the caller's displacement (3 per
// point) and strain (3x3 per point) arrays.
__global__ void k_hdirect_sum(int npt, int nchunk, int ifpt, int ifst, const double *__restrict__ part,
                              double *__restrict__ pt, double *__restrict__ st)
{
    int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= npt) return;
    double s[12];
    for (int k = 0; k < 12; k++) s[k] = 0.0;
    for (int ch = 0; ch < nchunk; ch++)
        for (int k = 0; k < 12; k++) s[k] += part[((size_t)ip * nchunk + ch) * 12 + k];
    int ks = ifpt ? 3 : 0;
    if (ifpt)
        for (int k = 0; k < 3; k++) pt[3 * (size_t)ip + k] += s[k];
    if (ifst)
        for (int k = 0; k < 9; k++) st[9 * (size_t)ip + k] += s[ks + k];   // symmetric: either order
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// out[i] = (x, y, -z) of in[i]
__global__ void k_mirror(int n, const double *__restrict__ in, double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[3 * (size_t)i] = in[3 * (size_t)i];
    out[3 * (size_t)i + 1] = in[3 * (size_t)i + 1];
    out[3 * (size_t)i + 2] = -in[3 * (size_t)i + 2];
}

// ---- host launchers ----

int adap_queue_waves()
{
    static int v = 0;
    if (!v) {
        int dev = 0, cus = 256;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, dev) == hipSuccess) cus = pr.multiProcessorCount;
        v = cus * 16;
    }
    return v;
}

size_t adap_stack_bytes(int64_t ntask)
{
    int64_t w = ntask < adap_queue_waves() ? ntask : adap_queue_waves();
    return (size_t)(w > 0 ? w : 1) * ADAP_STACK * ADAP_ENTRY * 8 + 256;
}

// scratch: adap_stack_bytes(ntask) bytes; the first 256 hold the counter
hipError_t launch_hnear(const HalfNearArgs &a, int64_t ntask, void *scratch, double *out, hipStream_t st)
{
    if (ntask <= 0) return hipSuccess;
    auto *ctr = (unsigned long long *)scratch;
    double *stack = (double *)((char *)scratch + 256);
    int waves = (int)(ntask < adap_queue_waves() ? ntask : adap_queue_waves());
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, ctr, 1);
    hipLaunchKernelGGL(k_hnear, dim3(waves), dim3(64), 0, st, a, ntask, ctr, stack, out);
    return hipGetLastError();
}

hipError_t launch_hnear_sum(int n, const int *boxes, const int *tgt_start, const int64_t *sbase, const int *ntri_q,
                            const int *pstart, int nq, int npts, int nl, const double *tmp, double *nf,
                            hipStream_t st)
{
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_hnear_sum, dim3((npts + 3) / 4), dim3(256), 0, st, n, boxes, tgt_start, sbase, ntri_q,
                       pstart, nq, nl, tmp, nf);
    return hipGetLastError();
}

size_t hdirect_scratch_bytes(int npt, int n)
{
    int nchunk = (n + HD_CHUNK - 1) / HD_CHUNK;
    int64_t ntask = (int64_t)npt * nchunk;
    return adap_stack_bytes(ntask) + (size_t)ntask * 12 * 8 + 256;
}

hipError_t launch_hdirect(int npt, const double *pts, int n, const double *tri, const double *nrm, const double *ssl,
                          const double *sdl, int ifsl, int ifdl, int ifpt, int ifst, double lam, double mu,
                          void *scratch, double *pt_out, double *st_out, hipStream_t st)
{
    int nchunk = (n + HD_CHUNK - 1) / HD_CHUNK;
    int64_t ntask = (int64_t)npt * nchunk;
    if (ntask <= 0 || !(ifsl || ifdl)) return hipSuccess;
    auto *ctr = (unsigned long long *)scratch;
    double *stack = (double *)((char *)scratch + 256);
    double *part = (double *)((char *)scratch + ((adap_stack_bytes(ntask) + 255) & ~(size_t)255));
    int waves = (int)(ntask < adap_queue_waves() ? ntask : adap_queue_waves());
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, ctr, 1);
    hipLaunchKernelGGL(k_hdirect, dim3(waves), dim3(64), 0, st, ntask, nchunk, n, pts, tri, nrm, ssl, sdl, ifsl,
                       ifdl, ifpt, ifst, lam, mu, ctr, stack, part);
    hipLaunchKernelGGL(k_hdirect_sum, dim3((npt + 255) / 256), dim3(256), 0, st, npt, nchunk, ifpt, ifst, part,
                       pt_out, st_out);
    return hipGetLastError();
}

hipError_t launch_mirror(int n, const double *in, double *out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mirror, dim3((n + 255) / 256), dim3(256), 0, st, n, in, out);
    return hipGetLastError();
}

// Image part of the point-source half-space direct sum (elh3dpartdirecttarg,
// src/elhfmm3dpart.f:961-1178): at every evaluation point (the sources, when
// ifpt or ifst, then the targets; caller's order) the image fields of all
// sources, the point's own included as in the reference (image_sl with
// sigma_sl, image_dl with sigma_dl and orientation sigma_dv), added to the
// outputs. One thread per point; the sources go through LDS in tiles.
constexpr int PI_THREADS = 128, PI_TILE = 128, PI_REC = 12;
__global__ __launch_bounds__(PI_THREADS) void k_pimage(int ne, int n, int nt, const double *__restrict__ src,
                                                      const double *__restrict__ tgt, const double *__restrict__ ssl,
                                                      const double *__restrict__ sdl, const double *__restrict__ sdv,
                                                      int ifsl, int ifdl, double lam, double mu, int ifpt,
                                                      double *__restrict__ pt, int ifst, double *__restrict__ st,
                                                      int ifptt, double *__restrict__ ptt, int ifstt,
                                                      double *__restrict__ stt)
{
    __shared__ double S[PI_TILE * PI_REC];
    int ip = blockIdx.x * PI_THREADS + threadIdx.x;
    bool has = ip < ne + nt;
    bool is_src = ip < ne;
    const double *y = has ? (is_src ? src + 3 * (size_t)ip : tgt + 3 * (size_t)(ip - ne)) : nullptr;
    double Y[3] = {0.0, 0.0, -1.0};
    if (has) Y[0] = y[0], Y[1] = y[1], Y[2] = y[2];
    double u[3] = {0, 0, 0}, g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int t0 = 0; t0 < n; t0 += PI_TILE) {
        int m = min(PI_TILE, n - t0);
        __syncthreads();
        for (int e = threadIdx.x; e < m * PI_REC; e += PI_THREADS) {
            int sj = e / PI_REC, c = e - sj * PI_REC;
            size_t j = (size_t)(t0 + sj);
            double v = c < 3 ? src[3 * j + c]
                             : c < 6 ? (ifsl ? ssl[3 * j + c - 3] : 0.0)
                                     : c < 9 ? (ifdl ? sdl[3 * j + c - 6] : 0.0) : (ifdl ? sdv[3 * j + c - 9] : 0.0);
            S[e] = v;
        }
        __syncthreads();
        if (!has) continue;
        for (int sj = 0; sj < m; sj++) {
            const double *r = S + sj * PI_REC;
            double f[12];
            if (ifsl) {
                image_sl(lam, mu, r, r + 3, Y, 1, 1, f);
                for (int d = 0; d < 3; d++) u[d] += f[d];
                for (int d = 0; d < 9; d++) g[d] += f[3 + d];
            }
            if (ifdl) {
                image_dl(lam, mu, r, r + 6, r + 9, Y, 1, 1, f);
                for (int d = 0; d < 3; d++) u[d] += f[d];
                for (int d = 0; d < 9; d++) g[d] += f[3 + d];
            }
        }
    }
    if (!has) return;
    double *po = is_src ? (ifpt ? pt + 3 * (size_t)ip : nullptr) : (ifptt ? ptt + 3 * (size_t)(ip - ne) : nullptr);
    double *so = is_src ? (ifst ? st + 9 * (size_t)ip : nullptr) : (ifstt ? stt + 9 * (size_t)(ip - ne) : nullptr);
    if (po)
        for (int d = 0; d < 3; d++) po[d] += u[d];
    if (so)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) so[3 * b + a] += g[3 * a + b];
}

hipError_t launch_pimage(int ne, int n, int nt, const double *src, const double *tgt, const double *ssl,
                         const double *sdl, const double *sdv, int ifsl, int ifdl, double lam, double mu, int ifpt,
                         double *pt, int ifst, double *st, int ifptt, double *ptt, int ifstt, double *stt,
                         hipStream_t s)
{
    int npt = ne + nt;
    if (npt <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pimage, dim3((npt + PI_THREADS - 1) / PI_THREADS), dim3(PI_THREADS), 0, s, ne, n, nt, src,
                       tgt, ssl, sdl, sdv, ifsl, ifdl, lam, mu, ifpt, pt, ifst, st, ifptt, ptt, ifstt, stt);
    return hipGetLastError();
}

}  // namespace elfmm
