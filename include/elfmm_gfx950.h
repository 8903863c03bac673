/*
 * elfmm_gfx950: MI355X (gfx950) elastostatic triangle layer-potential FMM.
 *
 * C-ABI of libelfmm_gfx950.so. Two families of entry points:
 *
 *  1. Drop-in replacements of the reference's Fortran subroutines, with the
 *     reference's symbol convention (lowercase, trailing underscore, every
 *     argument by reference, Fortran column-major arrays, host memory). A
 *     Fortran program that linked the reference's objects links this
 *     library instead; iso_c_binding interfaces are in
 *     elfmmlib3d_amd/fortran/elfmm_gfx950.f90.
 *
 *  2. Device-resident entry points (same argument meaning, device pointers
 *     and a HIP stream) for callers that keep their data in HBM.
 *
 * Array layouts (Fortran): triangle(3,3,n) = vertex coordinates, trinorm(3,n)
 * unit normals, source(3,n) centroids, sigma_sl/sigma_dl(3,n) densities,
 * target(3,nt), ptfrc(3,*) displacement, strain(3,3,*). Outputs are not
 * scaled by 1/(4 pi), as in the reference.
 *
 * Error codes (ier): 0 ok; 4 tree construction failed; 8 device memory
 * allocation failed; 16 unsupported precision flag on this device path
 * (iprec outside -2..3); 32 HIP runtime error. Errors never abort the process.
 */
#ifndef ELFMM_GFX950_H
#define ELFMM_GFX950_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces elfmm3dtria (src/elfmm3dtria.f:71-74): evaluation at the
 * triangle centroids only. */
void elfmm3dtria_(int *ier, const int *iprec, const double *rlam, const double *rmu,
                  const double *triangle, const double *trinorm, const int *nparts,
                  const double *source, const int *ifsingle, const double *sigma_sl,
                  const int *ifdouble, const double *sigma_dl, const int *ifptfrc,
                  double *ptfrc, const int *ifstrain, double *strain);

/* Replaces elfmm3dtriatarg (src/elfmm3dtria.f:162-167). */
void elfmm3dtriatarg_(int *ier, const int *iprec, const double *rlam, const double *rmu,
                      const double *triangle, const double *trinorm, const int *nparts,
                      const double *source, const int *ifsingle, const double *sigma_sl,
                      const int *ifdouble, const double *sigma_dl, const int *ifptfrc,
                      double *ptfrc, const int *ifstrain, double *strain, const int *ntargs,
                      const double *target, const int *ifptfrctarg, double *ptfrctarg,
                      const int *ifstraintarg, double *straintarg);

/* Replaces el3dtriadirecttarg (src/elfmm3dtria.f:1453-1458): the O(N^2)
 * direct sum, evaluated on the GPU with the same pair quadratures. */
void el3dtriadirecttarg_(const double *rlam, const double *rmu, const double *triangle,
                         const double *trinorm, const int *nsource, const double *source,
                         const int *ifsingle, const double *sigma_sl, const int *ifdouble,
                         const double *sigma_dl, const int *ifptfrc, double *ptfrc,
                         const int *ifstrain, double *strain, const int *ntarget,
                         const double *target, const int *ifptfrctarg, double *ptfrctarg,
                         const int *ifstraintarg, double *straintarg);

/* Replaces elfmm3dtriatargiter (src/elfmm3dtria_iter_1.f:160-166), the iter
 * engine (leaf size twice elfmm3dtriatarg's, lists 3 and 4 in the near
 * field, half-space root rule of the link set the reference builds it in).
 * icomp_type: 1 compute; 2 compute and keep the near-field interactions; 3
 * compute the far field and replay the kept interactions with the current
 * densities (same geometry and flags as the icomp 2 call); 4 plan only, set
 * lused; 5 keep the interactions and add the near field to the outputs
 * (no far field); 0 nothing (lused = the reference's work-array size).
 * The interactions are kept in device memory; wsave (lwsave >= lused words,
 * lused = 4 for icomp 1-5) carries their handle between calls. ier = 64:
 * icomp 3 found no interactions matching wsave and computed the near field
 * directly. lwsave below lused with icomp 2/3/5 also computes directly. */
void elfmm3dtriatargiter_(int *ier, const int *iprec, const double *rlam, const double *rmu,
                          const double *triangle, const double *trinorm, const int *nparts,
                          const double *source, const int *ifsingle, const double *sigma_sl,
                          const int *ifdouble, const double *sigma_dl, const int *ifptfrc,
                          double *ptfrc, const int *ifstrain, double *strain, const int *ntargs,
                          const double *target, const int *ifptfrctarg, double *ptfrctarg,
                          const int *ifstraintarg, double *straintarg, const int *icomp_type,
                          double *wsave, const int *lwsave, int *lused);

/* Replaces elfmm3dtriaiter (src/elfmm3dtria_iter_1.f:43-47): as above at the
 * centroids only. */
void elfmm3dtriaiter_(int *ier, const int *iprec, const double *rlam, const double *rmu,
                      const double *triangle, const double *trinorm, const int *nparts,
                      const double *source, const int *ifsingle, const double *sigma_sl,
                      const int *ifdouble, const double *sigma_dl, const int *ifptfrc,
                      double *ptfrc, const int *ifstrain, double *strain, const int *icomp_type,
                      double *wsave, const int *lwsave, int *lused);

/* Device-resident elfmm3dtriatargiter (device arrays, host wsave), same
 * semantics; returns ier. No counterpart in the reference. */
int elfmm_gfx950_triatargiter_dev(int iprec, double rlam, double rmu, int nparts,
                                  const double *d_triangle, const double *d_trinorm,
                                  const double *d_source, int ifsingle, const double *d_sigma_sl,
                                  int ifdouble, const double *d_sigma_dl, int ifptfrc,
                                  double *d_ptfrc, int ifstrain, double *d_strain, int ntargs,
                                  const double *d_target, int ifptfrctarg, double *d_ptfrctarg,
                                  int ifstraintarg, double *d_straintarg, int icomp_type,
                                  double *wsave, int lwsave, int *lused, void *stream);

/* Replays (icomp_type 3) of a store for nrhs density sets in one call: the
 * near field of all sets in one pass over the store per 4 (both layers) or
 * 8 sets, the far field of each set on the plan kept with the store.
 * Densities and outputs are nrhs consecutive blocks of the single-set
 * layout (d_sigma_dl[r * 3 * nparts ...], d_ptfrctarg[r * 3 * ntargs ...]).
 * Returns 64 without a store of these coordinates (nothing computed). No
 * counterpart in the reference (its callers loop over icomp 3 calls). */
int elfmm_gfx950_triatargiter_multi_dev(int iprec, double rlam, double rmu, int nparts,
                                        const double *d_triangle, const double *d_trinorm,
                                        const double *d_source, int nrhs, int ifsingle,
                                        const double *d_sigma_sl, int ifdouble, const double *d_sigma_dl,
                                        int ifptfrc, double *d_ptfrc, int ifstrain, double *d_strain,
                                        int ntargs, const double *d_target, int ifptfrctarg,
                                        double *d_ptfrctarg, int ifstraintarg, double *d_straintarg,
                                        double *wsave, int lwsave, void *stream);
/* The same for the half-space image part (elhfmm3dtriatargiter's store). */
int elfmm_gfx950_htriatargiter_multi_dev(int iprec, double rlam, double rmu, int nparts,
                                         const double *d_triangle, const double *d_trinorm,
                                         const double *d_source, int nrhs, int ifsingle,
                                         const double *d_sigma_sl, int ifdouble, const double *d_sigma_dl,
                                         int ntargs, const double *d_target, int ifptfrctarg,
                                         double *d_ptfrctarg, int ifstraintarg, double *d_straintarg,
                                         double *wsave, int lwsave, void *stream);

/* Frees the device-resident interactions named by wsave (returns 1 if
 * freed). No counterpart in the reference, whose wsave is the caller's
 * memory.
 * Sharing a wsave between threads: replays (icomp_type 3) of one store may
 * run concurrently; a release, or an icomp_type 2/5 call that replaces the
 * store, while other calls use it frees the old store only when they have
 * finished. Two icomp_type 2/5 calls on one wsave must not overlap (each
 * writes its handle into wsave; the store whose handle is lost is never
 * freed). An icomp_type 3 call replays a store only if the coordinates of
 * its sources and targets equal those the store was built on; otherwise it
 * computes directly and returns ier = 64. */
int elfmm_gfx950_iter_release(const double *wsave, int lwsave);

/* Expansion order the library uses at this iprec (-2..3), or -1. No
 * counterpart in the reference (l3dterms, src/l3dterms.f:42-85, gives the
 * same order). */
int elfmm_gfx950_expansion_order(int iprec);

/* Replaces elhfmm3dtriatargiter (src/elhfmm3dtria_iter.f:46-51): the
 * elastostatic layer potentials of the Mindlin half space z < 0, image part
 * only (the reference's caller adds the free-space "direct arrival" from
 * elfmm3dtriatargiter with the same tree rule; sources are evaluated by
 * passing them as targets). Result = anti-image far field (subtracted) +
 * Mindlin harmonic far field + image near field of lists 1, 3, 4 and self,
 * integrated adaptively (1e-6 relative, depth <= 200) on the GPU.
 * icomp_type as elfmm3dtriatargiter_ (1 compute, 2 compute and keep the
 * image near field on the device, 3 replay it, 4 plan only, 5 keep it and
 * return the near field only, 0 zero outputs); wsave (lwsave >= 4) carries
 * the handle, released by elfmm_gfx950_iter_release. Deviation: the
 * reference's integrand leaves the output entries it was not asked for
 * unset (src/elhfmm3dtria_iter.f:2901-3027) and its subdivision test reads
 * them; here they are zero, so with only one of ifptfrctarg/ifstraintarg
 * set the subdivision can differ from a given run of the reference. */
void elhfmm3dtriatargiter_(int *ier, const int *iprec, const double *rlam, const double *rmu,
                           const double *triangle, const double *trinorm, const int *nparts,
                           const double *source, const int *ifsingle, const double *sigma_sl,
                           const int *ifdouble, const double *sigma_dl, const int *ntargs,
                           const double *target, const int *ifptfrctarg, double *ptfrctarg,
                           const int *ifstraintarg, double *straintarg, const int *icomp_type,
                           double *wsave, const int *lwsave, int *lused);

/* Device-resident elhfmm3dtriatargiter (device arrays, host wsave, a HIP
 * stream); returns ier. No counterpart in the reference. */
int elfmm_gfx950_htriatargiter_dev(int iprec, double rlam, double rmu, int nparts,
                                   const double *d_triangle, const double *d_trinorm,
                                   const double *d_source, int ifsingle, const double *d_sigma_sl,
                                   int ifdouble, const double *d_sigma_dl, int ntargs,
                                   const double *d_target, int ifptfrctarg, double *d_ptfrctarg,
                                   int ifstraintarg, double *d_straintarg, int icomp_type,
                                   double *wsave, int lwsave, int *lused, void *stream);

/* Replaces elh3dtriadirecttarg (src/elh3dtria.f:53-58): the O(N^2) half-space
 * direct sum (free-space pair quadratures plus the adaptive image integral
 * of every pair), on the GPU. */
void elh3dtriadirecttarg_(const double *rlam, const double *rmu, const double *triangle,
                          const double *trinorm, const int *nsource, const double *source,
                          const int *ifsingle, const double *sigma_sl, const int *ifdouble,
                          const double *sigma_dl, const int *ifptfrc, double *ptfrc,
                          const int *ifstrain, double *strain, const int *ntarget,
                          const double *target, const int *ifptfrctarg, double *ptfrctarg,
                          const int *ifstraintarg, double *straintarg);

/* Replaces elfmm3dparttarg (src/elfmm3dpart.f:104-109): the elastostatic
 * FMM of point sources (point forces sigma_sl, force dipoles sigma_dl with
 * orientations sigma_dv) as four Laplace FMMs on the particle FMM's octree
 * (lfmm3dparthesstarg, leaf size src/lfmm3dparthess.f:272-280), the self
 * pair excluded; evaluated at the sources (ptfrc, strain) and targets. Same
 * error codes as elfmm3dtriatarg_. */
void elfmm3dparttarg_(int *ier, const int *iprec, const double *rlam, const double *rmu, const int *nparts,
                      const double *source, const int *ifsingle, const double *sigma_sl, const int *ifdouble,
                      const double *sigma_dl, const double *sigma_dv, const int *ifptfrc, double *ptfrc,
                      const int *ifstrain, double *strain, const int *ntargs, const double *target,
                      const int *ifptfrctarg, double *ptfrctarg, const int *ifstraintarg, double *straintarg);

/* Replaces elfmm3dpart (src/elfmm3dpart.f:52-55): elfmm3dparttarg_
 * without targets. */
void elfmm3dpart_(int *ier, const int *iprec, const double *rlam, const double *rmu, const int *nparts,
                  const double *source, const int *ifsingle, const double *sigma_sl, const int *ifdouble,
                  const double *sigma_dl, const double *sigma_dv, const int *ifptfrc, double *ptfrc,
                  const int *ifstrain, double *strain);

/* Replaces el3dpartdirecttarg (src/elfmm3dpart.f:506-511): the O(N^2)
 * point-source sum without self pairs, on the GPU (the near-field kernel of
 * the point FMM over one box holding every point). */
void el3dpartdirecttarg_(const double *rlam, const double *rmu, const int *nsource, const double *source,
                         const int *ifsingle, const double *sigma_sl, const int *ifdouble, const double *sigma_dl,
                         const double *sigma_dv, const int *ifptfrc, double *ptfrc, const int *ifstrain,
                         double *strain, const int *ntarget, const double *target, const int *ifptfrctarg,
                         double *ptfrctarg, const int *ifstraintarg, double *straintarg);

/* Replaces elh3dpartdirecttarg (src/elhfmm3dpart.f:961-966): the O(N^2)
 * point-source sum with the Mindlin half-space Green's function (points in
 * z < 0): the free-space sum without self pairs plus the image part of
 * every source, its own point included, on the GPU. */
void elh3dpartdirecttarg_(const double *rlam, const double *rmu, const int *nsource, const double *source,
                          const int *ifsingle, const double *sigma_sl, const int *ifdouble, const double *sigma_dl,
                          const double *sigma_dv, const int *ifptfrc, double *ptfrc, const int *ifstrain,
                          double *strain, const int *ntarget, const double *target, const int *ifptfrctarg,
                          double *ptfrctarg, const int *ifstraintarg, double *straintarg);

/* Device-resident elfmm3dparttarg: device pointers with the Fortran layouts
 * above, enqueued on `stream` (a hipStream_t, NULL for the default stream),
 * returns after the outputs are written; sigma_dv is read when ifdouble is
 * 1. No counterpart in the reference: the same operation for callers whose
 * data already lives in HBM. */
int elfmm_gfx950_parttarg_dev(int iprec, double rlam, double rmu, int nparts, const double *d_source, int ifsingle,
                              const double *d_sigma_sl, int ifdouble, const double *d_sigma_dl,
                              const double *d_sigma_dv, int ifptfrc, double *d_ptfrc, int ifstrain, double *d_strain,
                              int ntargs, const double *d_target, int ifptfrctarg, double *d_ptfrctarg,
                              int ifstraintarg, double *d_straintarg, void *stream);

/* Diagnostic: which parts of the half-space apply run (bit 0 image near
 * field, bit 1 anti-image far field, bit 2 Mindlin harmonic far field;
 * default 7), process-wide. Lets a test compare each part with the
 * reference's routine for that part. No counterpart in the reference. */
void elfmm_gfx950_half_parts(int mask);

/* el3dtriadirecttarg on device arrays without any allocation: no counterpart
 * in the reference; for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results as el3dtriadirecttarg_ (device
 * pointers, Fortran layouts), enqueued on `stream` (a hipStream_t, NULL for
 * the default stream); the call returns at once and the outputs are ready
 * when the stream has run it. The workspace is a static 46 MiB device array
 * held by every device the library runs on until the process ends:
 * nsource <= 131072 and nsource + ntarget <= 262144, else it returns 128
 * and enqueues nothing. All calls on one device share that array: calls
 * that overlap on one device (different streams, or threads) race. Returns
 * 0, 128 (sizes above the limits) or 32 (HIP error). */
int elfmm_gfx950_el3dtriadirecttarg_noalloc(double rlam, double rmu, int nsource,
                                            const double *d_triangle, const double *d_trinorm,
                                            const double *d_source, int ifsingle,
                                            const double *d_sigma_sl, int ifdouble,
                                            const double *d_sigma_dl, int ifptfrc, double *d_ptfrc,
                                            int ifstrain, double *d_strain, int ntarget,
                                            const double *d_target, int ifptfrctarg,
                                            double *d_ptfrctarg, int ifstraintarg,
                                            double *d_straintarg, void *stream);

/* Device-resident elfmm3dtriatarg: all arrays are device pointers with the
 * Fortran layouts above; the call is enqueued on `stream` (a hipStream_t,
 * NULL for the default stream) and returns after the outputs are written.
 * Calls from several host threads, each on its own stream, may run at
 * once: their host planning takes turns, and their far fields run on the
 * device in the order the calls finish planning, while near fields run
 * beside them. No counterpart in the reference: it is the same operation
 * for callers whose data already lives in HBM. */
int elfmm_gfx950_triatarg_dev(int iprec, double rlam, double rmu, int nparts,
                              const double *d_triangle, const double *d_trinorm,
                              const double *d_source, int ifsingle, const double *d_sigma_sl,
                              int ifdouble, const double *d_sigma_dl, int ifptfrc,
                              double *d_ptfrc, int ifstrain, double *d_strain, int ntargs,
                              const double *d_target, int ifptfrctarg, double *d_ptfrctarg,
                              int ifstraintarg, double *d_straintarg, void *stream);

/* Batch of meshes: nmesh independent free-space applies, each the
 * elfmm3dtriatarg of one mesh (src/elfmm3dtria.f:57-61 per mesh, same
 * precision flag, Lame constants, layers and output flags for all), on
 * geometries fixed at creation. Arrays are packed mesh after mesh in the
 * caller's order: mesh m owns nparts[m] triangles (d_triangle 9 doubles,
 * d_trinorm, d_source, densities and d_ptfrc 3, d_strain 9 per triangle)
 * and ntargs[m] targets (ntargs may be NULL: no targets). The geometry is
 * copied; octrees, task lists and device buffers are built once. Two
 * schedules: the forest (all meshes' octrees planned as one FMM whose task
 * lists never cross meshes: one launch per phase for the whole batch) and
 * branches (each mesh its own apply, meshes spread over concurrent chains).
 * width in 1..16 fixes the number of branches and replays one graph built
 * for them; -16..-1 fixes -width branches launched directly on streams;
 * 100 / -100 the forest as a graph / direct launches; 0 lets the tuner time
 * the forest and widths 1, 2, 4, 8, 16 both ways and keep the fastest
 * (the buffers of the other schedule are released). Returns ier
 * (16 bad arguments, 4 tree failure, 8 host memory, 32 device error); on
 * success *batch is the handle. No counterpart in the reference, whose
 * callers loop over meshes. */
int elfmm_gfx950_batch_create(int iprec, double rlam, double rmu, int nmesh, const int *nparts,
                              const double *d_triangle, const double *d_trinorm,
                              const double *d_source, int ifsingle, int ifdouble, int ifptfrc,
                              int ifstrain, const int *ntargs, const double *d_target,
                              int ifptfrctarg, int ifstraintarg, int width, void **batch);
/* The batch's own packed density and output arrays (NULL where a layer or
 * output is off): densities written there need no copy in, outputs read
 * there need no copy out. */
int elfmm_gfx950_batch_buffers(void *batch, double **d_sigma_sl, double **d_sigma_dl,
                               double **d_ptfrc, double **d_strain, double **d_ptfrctarg,
                               double **d_straintarg);
/* One apply of every mesh, enqueued on stream and returning at once: the
 * densities (packed; NULL or the batch's own array: no copy) are copied
 * in, all meshes evaluated, and the outputs copied to the given arrays
 * (NULL or the batch's own: no copy). Applies of one batch share its
 * buffers: issue them on one stream (or synchronize between them). Must be
 * called with the batch's device current. Returns ier. */
int elfmm_gfx950_batch_apply(void *batch, const double *d_sigma_sl, const double *d_sigma_dl,
                             double *d_ptfrc, double *d_strain, double *d_ptfrctarg,
                             double *d_straintarg, void *stream);
/* The schedule in use (branches, 0 for the forest; 1 = graph replay / 0 =
 * direct launches) and up to ntune (graph, width, milliseconds) triples of
 * the schedules the tuner timed; returns the number of schedules timed. */
int elfmm_gfx950_batch_schedule(void *batch, int *width, int *graph, int ntune, double *tune);
/* Releases the batch (waits for the device first). */
int elfmm_gfx950_batch_destroy(void *batch);

/* The adaptive octree and interaction lists of the reference's d3tstrcr
 * (src/d3tstrcr.f:31-266, lists :702-1179; imindlin != 0: the half-space
 * root of src/d3tstrcr_mindlin.f:1380-1383) for nsource sources and ntarget
 * targets (device arrays, 3 doubles per point; a box splits above nbox
 * sources or targets), built on the GPU: the applies build their trees this
 * way. Returns the number of boxes, or -4 when a box below level 21 would
 * split (the applies then build that tree on the host), -16 bad arguments,
 * -32 device error. sizes (7 ints, may be NULL): boxes, levels, entries of
 * lists 1..5. The tree is kept for elfmm_gfx950_octree_get on this thread. */
int elfmm_gfx950_octree_dev(const double *d_source, int nsource, const double *d_target, int ntarget,
                            int nbox, int imindlin, int *sizes, void *stream);
/* The last octree of this thread, 0-based: boxes (18 ints per box: level,
 * i, j, k (1-based cube coordinates), parent (-1 for the root), number of
 * children, first source, sources, first target, targets, the 8 child
 * slots, -1 where empty), centers (3 doubles per box), iz / iztarg (input
 * index of each point in tree order), lstart ((boxes + 1) ints per list,
 * lists 1..5 in turn) and litems (their entries, lists 1..5 in turn). */
void elfmm_gfx950_octree_get(int *boxes, double *centers, int *iz, int *iztarg, int *lstart, int *litems);

/* Per-phase times (milliseconds) of the last call on this thread:
 * [0] tree+plan (host), [1] upload, [2] P2M, [3] M2M, [4] M2L, [5] P2L,
 * [6] L2L, [7] L2T, [8] M2T, [9] near field, [10] assembly, [11] total
 * wall time of the call, [12] octree and lists (part of [0]), [13]
 * interaction plan (part of [0]); after a half-space call [0]-[13] are
 * those of its anti-image and image near-field pass, [14] the time of its
 * Mindlin far field (wall time after the anti-image pass; in a replay of
 * one set, where it runs beside that pass, its device time) and [15] the
 * wall time of the whole call. [0],
 * [12], [13] are host times; [1]-[10] are HIP-event times on the stream
 * each phase runs on (the near field overlaps [2]-[8]). */
void elfmm_gfx950_last_timings(double *ms16);

/* Work counts of the last call on this thread: [0] boxes, [1] levels,
 * [2] expansion order p, [3] M2L pairs, [4] M2L flops as executed (dense
 * operators, four right-hand sides), [5] M2L flops of the reference's
 * rotate/shift/rotate scheme for the same pairs, [6] near-field (point,
 * triangle) pairs, [7] M2L tiles, [8] M2M pairs, [9] L2L pairs, [10] P2M
 * boxes, [11] P2L (box, source box) pairs, [12] M2T (box, source box) pairs,
 * [13] L2T boxes, [14] iter: near-field store size in the reference's
 * word layout (its lused), [15] iter: this library's store size in words. */
void elfmm_gfx950_last_stats(double *st16);

/* Library version string. */
const char *elfmm_gfx950_version(void);

#ifdef __cplusplus
}
#endif
#endif
