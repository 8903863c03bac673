/*
 * CPU restatement of the reference's free-space triangle layer-potential
 * path. TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's
 * cpu_baseline leg load this to compare against; the product never links it.
 *
 * Arrays use the Fortran layouts of the reference: triangle(3,3,n) is
 * tri[9*k + 3*v + c], vectors (3,n) are x[3*k + c], strain(3,3,n) is
 * st[9*k + 3*col + row].
 */
#ifndef ELFMM_ORACLE_H
#define ELFMM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Integrals over a plane triangle (vertices (0,0),(a,0),(b,c), c>0) of
 * derivatives of r and 1/r at target (x0,y0,z0): btab[i][j][k] for i+j+k<=4
 * (r), rtab[i][j][k] for i+j+k<=2 (1/r), in the reference's convention of
 * triabtable/triartable (src/triahquad.f:680,797). iself=1 selects the
 * principal value for a target on the triangle. Flat arrays of 5*5*5 and
 * 3*3*3 doubles, index [25*i+5*j+k] and [9*i+3*j+k]. */
void oracle_tables(double a, double b, double c, double x0, double y0,
                   double z0, int iself, double *btab, double *rtab);

/* elust3triadirecttarg_one (src/elfmm3dtria.f:1689) */
void oracle_pair_sl(double rlam, double rmu, const double *tri,
                    const double *sigma, int ifself, const double *target,
                    double *ptfrc, int ifstrain, double *strain);

/* eltst3triadirecttarg_one (src/elfmm3dtria.f:2368) */
void oracle_pair_dl(double rlam, double rmu, const double *tri,
                    const double *sigma, int ifself, const double *target,
                    double *ptfrc, int ifstrain, double *strain);

/* el3dtriadirecttarg (src/elfmm3dtria.f:1453), OpenMP over targets. */
void oracle_direct(double rlam, double rmu, const double *tri,
                   const double *trinorm, int nsource, const double *source,
                   int ifsingle, const double *sigma_sl, int ifdouble,
                   const double *sigma_dl, int ifptfrc, double *ptfrc,
                   int ifstrain, double *strain, int ntarget,
                   const double *target, int ifptfrctarg, double *ptfrctarg,
                   int ifstraintarg, double *straintarg);

/* Octree and lists of d3tstrcr (src/d3tstrcr.f:31) with the free-space root
 * rule (imindlin=0) or the half-space root rule of
 * src/d3tstrcr_mindlin.f:1380-1383 (imindlin=1).
 * Outputs (1-based box numbers and point indices, as the reference):
 *   boxes[20*nboxes], centers[3*nboxes], iz[n], iztarg[nt], laddr[2*200],
 *   *nlev, center0[3], *size,
 *   lists: for itype 1..5, CSR arrays lstart[5*(nboxes+1)], litems[...].
 * Call once with litems=NULL to get *nboxes and *nitems, then again with
 * storage. Returns 0 or a reference-style error code. */
int oracle_tree(const double *src, int n, const double *targ, int nt,
                int nbox, int imindlin, int maxboxes, int *nboxes,
                int *boxes, double *centers, int *iz, int *iztarg,
                int *laddr, int *nlev, double *center0, double *size,
                int *lstart, int *litems, long *nitems);

#ifdef __cplusplus
}
#endif
#endif
