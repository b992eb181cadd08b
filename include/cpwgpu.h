/*
 * cpwgpu -- MI355X-native H|psi> for cpw2000.
 *
 * C ABI of the shared library libcpwgpu.so.  Plain pointers and sizes only.
 * Complex arrays are interleaved (re, im) pairs: Fortran complex(REAL64)
 * (16 bytes) for the double-precision arrays, column-major with the leading
 * dimensions the reference uses.
 *
 * Two layers:
 *   1. cpwgpu_hk_psi_c16  -- the drop-in.  Same argument list, meaning and
 *      by-reference passing as the Fortran subroutine it replaces, host
 *      arrays in and out.  The Fortran shim hk_psi_c16 (fortran/hk_psi_c16.f90)
 *      forwards to it through iso_c_binding.
 *   2. cpwgpu_plan_*      -- the device-resident path: set up a k-point once
 *      (index maps, potential, projectors), then apply H to device arrays
 *      on a caller stream.  The drop-in is built on it.
 *
 * Every function returns CPWGPU_OK (0) or one of the CPWGPU_ERR_* codes;
 * cpwgpu_last_error() gives a message for the calling thread.
 */
#ifndef CPWGPU_H
#define CPWGPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference stops the program in these situations; the
 * Fortran shim turns a nonzero code into a message on unit 6 and STOP. */
#define CPWGPU_OK 0
#define CPWGPU_ERR_NG 1          /* ng > mxdgve                  Src/hk_psi/reference/hk_psi_c16.f90:108     */
#define CPWGPU_ERR_MXDSCR 2      /* FFT work size > mxdscr       Src/hk_psi/reference/hk_psi_c16.f90:112-119 */
#define CPWGPU_ERR_WRAP 3        /* sphere wraps the FFT mesh    Src/hk_psi/reference/hk_psi_c16.f90:162-172 */
#define CPWGPU_ERR_MTXD 4        /* mtxd > mxddim or mtxd < 0                                                 */
#define CPWGPU_ERR_MODE 5        /* bad precision / option value                                              */
#define CPWGPU_ERR_SIZE 6        /* FFT mesh size not supported  Src/fft/_cfft/cfft_wf_c16.f:95-102         */
#define CPWGPU_ERR_ARG 7         /* other invalid argument (null pointer, nanl > mxdanl, ...)                 */
#define CPWGPU_ERR_NOTPD 8       /* overlap matrix of the subspace not positive definite                      */
#define CPWGPU_ERR_HIP 10        /* HIP runtime failure                                                       */
#define CPWGPU_ERR_NOMEM 11      /* device allocation failed                                                  */

/* Arithmetic of the FFT / pointwise / projector stages. */
#define CPWGPU_PREC_FP64 0       /* complex128 throughout                                                     */
#define CPWGPU_PREC_FP32 1       /* complex64 FFTs, V(r) product and projector GEMMs with FP32 accumulation;
                                    kinetic term, psi and H|psi> stay FP64                                  */

typedef struct cpwgpu_plan_s *cpwgpu_plan_t;

/* ---- library ------------------------------------------------------------ */

/* Select the device used by later calls from this process (default 0). */
int cpwgpu_init(int device);
/* Free every cached plan and device buffer. */
int cpwgpu_finalize(void);
/* Message describing the last error of the calling thread ("" if none). */
const char *cpwgpu_last_error(void);
/* Precision of cpwgpu_hk_psi_c16 (CPWGPU_PREC_*); default FP64. */
int cpwgpu_set_precision(int prec);
/* 1 (default): the non-local GEMMs run on a second stream beside the FFT
 * kernels (fork/join with events on the caller's stream); 0: every kernel on
 * the caller's stream, one after the other.  Same work and result. */
int cpwgpu_set_overlap(int on);
/* Iteration of cpwgpu_plan_ditsp / cpwgpu_ditsp_c16 (no counterpart in the
 * reference, whose Src/ditsp/ditsp_c16_loop.f90 relaxes band by band):
 * CPWGPU_DITSP_CHEB (default) = Chebyshev-filtered subspace iteration (the
 * filter degree follows the spectrum bounds, 4..24; CPWGPU_DITSP_DEGREE in
 * the environment fixes it); CPWGPU_DITSP_BLOCK = locally optimal block
 * preconditioned iteration with locking.  Same start, convergence test and
 * outputs. */
#define CPWGPU_DITSP_BLOCK 0
#define CPWGPU_DITSP_CHEB 1
int cpwgpu_set_ditsp_method(int method);

/* ---- drop-in ------------------------------------------------------------ */

/* Replaces Src/hk_psi/reference/hk_psi_c16.f90:28-241 (subroutine hk_psi_c16).
 * Arguments are the Fortran dummies, by reference:
 *   mtxd, neig          basis size and number of bands
 *   psi(mxddim,neig)    complex(8) in          hpsi(mxddim,neig) complex(8) out
 *                       (rows 1..mtxd written, rows > mtxd untouched)
 *   lnewanl             logical, read only (never written: callers pass literals)
 *   ng, kgv(3,mxdgve)   G list;  isort(mxddim) 1-based row -> G index
 *   ekpg(mxddim)        1/2|k+G|^2;  vscr(mxdscr) V(r), index (k3*n2+k2)*id+k1
 *   kmscr(7)            (4:6) = n1,n2,n3 mesh, (7) = id leading dimension
 *   anlga(mxddim,mxdanl), xnlkb(mxdanl), nanl    KB projectors
 * neig = 0 is a no-op; nanl = 0 skips the non-local term. */
int cpwgpu_hk_psi_c16(const int *mtxd, const int *neig, const double *psi, double *hpsi,
                      const int *lnewanl, const int *ng, const int *kgv, const double *ekpg,
                      const int *isort, const double *vscr, const int *kmscr, const double *anlga,
                      const double *xnlkb, const int *nanl, const int *mxddim, const int *mxdbnd,
                      const int *mxdanl, const int *mxdgve, const int *mxdscr);

/* Replaces Src/hk_psi/reference/hk_psi_spin_c16.f90:28-307 (subroutine
 * hk_psi_spin_c16), the two-component version: same dummies, by reference,
 *   psi_sp(2*mxddim, neig), hpsi_sp(2*mxddim, neig) complex(8): row 2i-1 / 2i
 *                       (1-based) = spin up / down coefficient of plane wave i
 *   vscr_sp(mxdscr, mxdnsp) real(8): V(r) components (1, sigma_z, sigma_x,
 *                       sigma_y); nsp = 1, 2 or 4 of them are used
 *   anlsp(2*mxddim, mxdasp) complex(8), xnlkbsp(mxdasp), nanlsp: projectors on
 *                       the interleaved rows
 * The reference's stops (ng > mxdgve, nsp > mxdnsp, nsp not 1/2/4, mxdwrk >
 * mxdscr, wrap-around) return CPWGPU_ERR_NG / _ARG / _MODE / _MXDSCR / _WRAP.
 * Precision as set by cpwgpu_set_precision. */
int cpwgpu_hk_psi_spin_c16(const int *mtxd, const int *neig, const double *psi_sp, double *hpsi_sp,
                           const int *lnewanl, const int *ng, const int *kgv, const double *ekpg, const int *isort,
                           const double *vscr_sp, const int *kmscr, const int *nsp, const double *anlsp,
                           const double *xnlkbsp, const int *nanlsp, const int *mxddim, const int *mxdbnd,
                           const int *mxdasp, const int *mxdgve, const int *mxdscr, const int *mxdnsp);

/* Single-precision drop-in, the complex*8 interface of the reference's
 * Src/hk_psi/cuda/hk_psi_c8.f90 (hk_psi_c8) brought to the current argument
 * list: the dummies of hk_psi_c16 with complex(4) psi, hpsi, anlga and real(4)
 * ekpg, vscr, xnlkb.  Computed on the FP32-mixed path (inputs widened
 * exactly, H psi rounded to complex(4) on the way out); same checks and
 * status codes as cpwgpu_hk_psi_c16.  The reference's stale version takes
 * qmod (|k+G|) and kmscr(3) instead of ekpg and kmscr(7). */
int cpwgpu_hk_psi_c8(const int *mtxd, const int *neig, const float *psi, float *hpsi, const int *lnewanl,
                     const int *ng, const int *kgv, const float *ekpg, const int *isort, const float *vscr,
                     const int *kmscr, const float *anlga, const float *xnlkb, const int *nanl, const int *mxddim,
                     const int *mxdbnd, const int *mxdanl, const int *mxdgve, const int *mxdscr);

/* Drop-in for proj_nl_kb_c16 (Src/omp/proj_nl_kb_c16.f90:23-28, the
 * reference's argument list, all by reference, host arrays):
 *   rkpt(3), mtxd, isort(mxddim), nanl, ng, kgv(3,mxdgve), nqnl(mxdtyp),
 *   delqnl(mxdtyp), vkb(-2:mxdlqp,0:3,-1:1,mxdtyp), nkb(0:3,-1:1,mxdtyp),
 *   ntype, natom(mxdtyp), rat(3,mxdatm,mxdtyp), adot(3,3) in;
 *   anlga(mxddim,mxdanl) complex(8), xnlkb(mxdanl) out (rows 1..mtxd and
 *   columns 1..nanl written, the rest untouched);
 *   mxdtyp, mxdatm, mxdlqp, mxddim, mxdanl, mxdgve dimensions.
 * The projectors are computed on the GPU and copied back.  ng > mxdgve is
 * CPWGPU_ERR_NG (:117), nanl inconsistent with nkb/natom or > mxdanl
 * CPWGPU_ERR_ARG (:143-157), with the conditions of
 * cpwgpu_plan_build_projectors_kb below. */
int cpwgpu_proj_nl_kb_c16(const double *rkpt, const int *mtxd, const int *isort, const int *nanl, const int *ng,
                          const int *kgv, const int *nqnl, const double *delqnl, const double *vkb, const int *nkb,
                          const int *ntype, const int *natom, const double *rat, const double *adot, double *anlga,
                          double *xnlkb, const int *mxdtyp, const int *mxdatm, const int *mxdlqp, const int *mxddim,
                          const int *mxdanl, const int *mxdgve);

/* ---- device-resident plan ------------------------------------------------ */

/* Builds the index maps of one k-point from host arrays (the hk_psi_c16
 * inputs kgv, isort, ekpg, kmscr; same meaning and checks as the drop-in,
 * Src/hk_psi/reference/hk_psi_c16.f90:105-172). */
int cpwgpu_plan_create(cpwgpu_plan_t *plan, int mtxd, int ng, const int *kgv, const double *ekpg,
                       const int *isort, const int *kmscr, int mxddim, int mxdgve);
int cpwgpu_plan_destroy(cpwgpu_plan_t plan);
/* V(r) from host memory (mxdscr doubles, layout of vscr). */
int cpwgpu_plan_set_potential(cpwgpu_plan_t plan, const double *vscr, int mxdscr);
/* Projectors from host memory: anlga(mxddim, nanl) complex(8), xnlkb(nanl). */
int cpwgpu_plan_set_projectors(cpwgpu_plan_t plan, const double *anlga, const double *xnlkb,
                               int nanl, int mxddim);
/* Same from device memory (anlga complex128 with leading dimension lda). */
int cpwgpu_plan_set_projectors_device(cpwgpu_plan_t plan, const double *anlga_dev, long lda,
                                      const double *xnlkb_dev, int nanl);
/* Builds the projectors of the plan's k-point on the device, as
 * proj_nl_kb_c16 (Src/omp/proj_nl_kb_c16.f90:23-275) computes anlga and xnlkb
 * for the plan's plane waves: rkpt(3), nqnl(mxdtyp), delqnl(mxdtyp),
 * vkb(-2:mxdlqp, 0:3, -1:1, mxdtyp), nkb(0:3, -1:1, mxdtyp), natom(mxdtyp),
 * rat(3, mxdatm, mxdtyp), adot(3, 3) in the reference's layouts (host memory).
 * Nothing of size mtxd * nanl crosses PCIe.  A nanl that does not match nkb
 * and natom, or a cell with a non-positive metric, is CPWGPU_ERR_ARG (the
 * reference's stops, proj_nl_kb_c16.f90:143-157 and Src/crystal/adot_to_avec.f90);
 * so is nqnl(type) > mxdlqp + 1 or delqnl(type) <= 0 for a type with
 * projectors (a table read beyond vkb's bounds in the reference).  Returns
 * when the projectors are built; as with cpwgpu_plan_set_projectors, applies
 * of this plan still running on other streams must have finished first. */
int cpwgpu_plan_build_projectors_kb(cpwgpu_plan_t plan, const double *rkpt, int nanl, const int *nqnl,
                                    const double *delqnl, const double *vkb, const int *nkb, int ntype,
                                    const int *natom, const double *rat, const double *adot, int mxdtyp,
                                    int mxdatm, int mxdlqp);
/* H|psi> on device arrays: psi_dev(ldp, neig), hpsi_dev(ldh, neig) complex128,
 * rows 0..mtxd-1 of hpsi written.  stream is a hipStream_t (NULL = the null
 * stream).  Returns when the work is enqueued. */
int cpwgpu_plan_apply(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp,
                      double *hpsi_dev, long ldh, int prec, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same H|psi> as cpwgpu_plan_apply, with the z-column
 * buffer, split-K partials and D<beta|psi> in a fixed device array of
 * CPWGPU_NOALLOC_SCRATCH_MIB MiB per device (held until the process ends)
 * instead of the plan's buffers.  Bands go through in chunks that fit; a
 * plan whose single band needs more returns CPWGPU_ERR_SIZE.  No hipMalloc,
 * hipFree or allocating library call.  FP32 needs the plan's FP32 projector
 * copy, made by one FP32 cpwgpu_plan_apply after the projectors are set
 * (else CPWGPU_ERR_ARG).  Every call on a device shares that device's
 * array: calls that overlap on one GPU (other streams, other threads) race. */
#define CPWGPU_NOALLOC_SCRATCH_MIB 64
int cpwgpu_plan_apply_noalloc(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp,
                              double *hpsi_dev, long ldh, int prec, void *stream);
/* Two-component bands on the same plan (the device side of
 * cpwgpu_hk_psi_spin_c16): the potential components vscr_sp(mxdscr, nsp) and
 * the projectors anlsp(2*mxddim, nanlsp) from host memory, then H psi_sp on
 * device arrays psi_dev(ldp, neig), hpsi_dev(ldh, neig) with ldp, ldh >=
 * 2*mtxd (interleaved up/down rows).  nsp = 4 (non-collinear) needs one of
 * the specialised meshes. */
int cpwgpu_plan_set_potential_spin(cpwgpu_plan_t plan, const double *vscr_sp, int mxdscr, int nsp);
int cpwgpu_plan_set_projectors_spin(cpwgpu_plan_t plan, const double *anlsp, const double *xnlkbsp, int nanlsp,
                                    int mxddim);
int cpwgpu_plan_apply_spin(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp, double *hpsi_dev,
                           long ldh, int prec, void *stream);
/* Sizes of the plan: out[0..9] = mtxd, n1, n2, n3, id, kd1, kd2, kd3, ncol, nanl. */
int cpwgpu_plan_info(cpwgpu_plan_t plan, int *out);

/* ---- kernel variants and autotuning (no counterpart in the reference) ------- */

/* Tunable choices of a plan, per precision (CPWGPU_PREC_*):
 *   CPWGPU_TUNE_PLANE   kernel of the y/x plane stage: 0 = LDS two-pass lines,
 *                       1..3 = register-resident lines (DPP butterflies) in
 *                       three shapes, 4..6 = LDS two-pass lines exchanged
 *                       within one wavefront (256, 192, 128 threads per
 *                       block; where the plane fits in LDS), 7 = register
 *                       x rows of shape 1 with the y lines through LDS
 *                       exchanged within one wavefront, 8..11 = register
 *                       shapes that neither load, transform nor store the
 *                       register slots outside the sphere box (where the box
 *                       lies inside |x|, |y| <= N/4); -1 = the default for
 *                       the mesh
 *   CPWGPU_TUNE_SPLITK  target number of blocks of the split-K projection
 *                       GEMM (64..4096); -1 = default (256)
 *   CPWGPU_TUNE_PROJ    kernel of the projection GEMM: 0 = LDS-staged tiles,
 *                       1..4 = direct register loads in four tile shapes,
 *                       5 = LDS-staged with half the wave tile;
 *                       -1 = default (0)
 *   CPWGPU_TUNE_BACK    kernel of the back-projection GEMM: 0 = LDS-staged D,
 *                       1..3 = direct register loads in three tile shapes,
 *                       4 = LDS-staged with half-size projector chunks;
 *                       -1 = default (0)
 *   CPWGPU_TUNE_KIN     where ekpg*psi is added when projectors are set: 0 = in
 *                       the back-projection epilogue, 1 = in the gather of the
 *                       forward column kernel (moves work between the two
 *                       streams); -1 = default (0)
 *   CPWGPU_TUNE_SPIN    plane kernel of cpwgpu_plan_apply_spin with nsp = 4: 0 =
 *                       LDS two-pass lines, 1 = register-resident lines (built
 *                       when 8 divides the mesh length); -1 = default (1
 *                       where built, else 0)
 *   CPWGPU_TUNE_COL     z-column kernels: 0 = LDS two-pass lines, 1 =
 *                       register-resident lines (built where a column fits in
 *                       the registers of 8 or 4 lanes); -1 = default (0)
 *   CPWGPU_TUNE_SCHED   stream schedule with overlap on: 0 = the non-local
 *                       term (projection, D scaling, back-projection) on the
 *                       second stream beside the local term, the gather adding
 *                       the local term to it; 1 = the projection and D scaling
 *                       beside the local term, the gather writing H psi, then
 *                       the back-projection adding to it on the caller's
 *                       stream; 2 = the inverse column and plane kernels on
 *                       the second stream, the whole non-local term on the
 *                       caller's stream, then the gather there; -1 = default (0) */
#define CPWGPU_TUNE_PLANE 0
#define CPWGPU_TUNE_SPLITK 1
#define CPWGPU_TUNE_PROJ 2
#define CPWGPU_TUNE_BACK 3
#define CPWGPU_TUNE_KIN 4
#define CPWGPU_TUNE_SPIN 5
#define CPWGPU_TUNE_COL 6
#define CPWGPU_TUNE_SCHED 7
int cpwgpu_plan_set_variant(cpwgpu_plan_t plan, int prec, int what, int value);
/* Current choice (the value used by the next apply, defaults resolved). */
int cpwgpu_plan_get_variant(cpwgpu_plan_t plan, int prec, int what, int *value);
/* Autotuner: times cpwgpu_plan_apply for every plane variant built for the
 * mesh, then both projection kernels with a few split-K targets, then both
 * back-projection kernels, on the caller's arrays (neig bands, stream), and
 * keeps the fastest in the plan.  hpsi_dev is overwritten with H psi.
 * If the environment variable CPWGPU_TUNE_CACHE names a file, choices are
 * read from it for meshes already tuned (no timing) and appended after
 * tuning; the key is (precision, n1, n2, n3, nx, ncol, nanl, neig). */
int cpwgpu_plan_tune(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp, double *hpsi_dev,
                     long ldh, int prec, void *stream);
/* Autotuner of the two-component path (cpwgpu_plan_apply_spin arguments; the
 * potential and projectors of cpwgpu_plan_set_*_spin must be set): times the
 * coupled plane kernels (nsp = 4), the projection kernel with a few split-K
 * targets and the back-projection kernel on the spinor GEMM shapes, and keeps
 * the fastest in the plan (the GEMM choices are shared with the scalar path).
 * hpsi_dev is overwritten with H psi.  No cache file. */
int cpwgpu_plan_tune_spin(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp, double *hpsi_dev,
                          long ldh, int prec, void *stream);

/* ---- every k-point of a self-consistency step in one call (host arrays) ------ */

/* H|psi> for the nrk k-points of the irk loop of Src/cpw/cpw_scf.f90:488-544,
 * with its all-k arrays (all by reference, Fortran layout):
 *   psi_allk, hpsi_allk (mxddim, mxdbnd, mxdnrk) complex(8); k-point k holds
 *     nband(k) bands of mtxd_allk(k) plane waves (ragged)
 *   isort_allk (mxddim, mxdnrk), ekpg_allk (mxddim, mxdnrk)
 *   anlga_allk (mxddim, mxdanl, mxdnrk) complex(8), xnlkb_allk (mxdanl, mxdnrk),
 *     nanl_allk (mxdnrk): the projectors of each k-point
 *   ng, kgv, vscr (mxdscr), kmscr(7): shared by all k-points, as in hk_psi_c16.
 * Rows beyond mtxd_allk(k) and columns beyond nband(k) of hpsi_allk are not
 * written.  One plan per k-point is kept between calls while its plane-wave
 * set is unchanged; psi, V and the projectors cross PCIe on every call (the
 * device-resident batch is cpwgpu_graph_*).  Precision: cpwgpu_set_precision.
 * Errors: as cpwgpu_hk_psi_c16, per k-point. */
int cpwgpu_hk_psi_allk_c16(const int *nrk, const int *mtxd_allk, const int *nband, const double *psi_allk,
                           double *hpsi_allk, const int *ng, const int *kgv, const double *ekpg_allk,
                           const int *isort_allk, const double *vscr, const int *kmscr, const double *anlga_allk,
                           const double *xnlkb_allk, const int *nanl_allk, const int *mxddim, const int *mxdbnd,
                           const int *mxdanl, const int *mxdgve, const int *mxdscr, const int *mxdnrk);

/* ---- several k-points in one graph launch -------------------------------------- */

typedef struct cpwgpu_graph_s *cpwgpu_graph_t;
/* Records H|psi> for nk k-points (plans[k], neig[k] bands psi[k](ldp[k]) ->
 * hpsi[k](ldh[k]) on device, each k with its own plane-wave count) into one
 * HIP graph (each k-point's non-local GEMMs overlap its FFT chain, as in
 * cpwgpu_plan_apply).  The batch is applied once eagerly first (buffers,
 * kernel attributes), then captured; the device arrays must stay valid while
 * the graph exists, and the graph must be destroyed before its plans.
 * Replaces the k-loop of Src/cpw/cpw_scf.f90:488-544 around hk_psi_c16 for a
 * batch of k-points.
 * The graph reads the plans' potential and projectors where they live, so a
 * replay sees new values set with the same sizes (cpwgpu_plan_set_potential,
 * or projectors of the same count from any cpwgpu_plan_set_projectors* /
 * cpwgpu_plan_build_projectors_kb call: the FP32 copy is converted again in
 * place by those calls). */
int cpwgpu_graph_create(cpwgpu_graph_t *graph, int nk, const cpwgpu_plan_t *plans, const int *neig,
                        const double *const *psi, const long *ldp, double *const *hpsi,
                        const long *ldh, int prec);
/* One launch of the recorded graph on stream (NULL = null stream).  Returns
 * CPWGPU_ERR_ARG without launching when a plan changed in a way the recording
 * cannot follow: another projector count, or a plan workspace reallocated by
 * a larger apply since the capture; create the graph again then. */
int cpwgpu_graph_launch(cpwgpu_graph_t graph, void *stream);
int cpwgpu_graph_destroy(cpwgpu_graph_t graph);

/* ---- FP64 subspace step ----------------------------------------------------- */

/* Rayleigh-Ritz of nb vectors in FP64 on device arrays psi_dev(ldp, nb) and
 * hpsi_dev(ldh, nb) = H psi (complex128): S = psi^H psi and M = psi^H hpsi on
 * the matrix cores, S = L L^H (Cholesky), L^-1 M L^-H = U diag(eig) U^H
 * (Jacobi), C = L^-H U.  If rotate != 0, psi <- psi C and hpsi <- hpsi C (the
 * Ritz vectors, orthonormal, and H applied to them).  eig (nb, ascending) and
 * res (nb, may be NULL: res_n = ||hpsi_n - eig_n psi_n||^2 after the
 * rotation) are host arrays.  Returns when done.  Replaces the FP64 steps of
 * Src/ditsp/ditsp_c16_ritz.f90:157-287 (orthonormalisation, H_sub, zheevd,
 * rotation) and the residual of Src/ditsp/ditsp_error.f90:70-76.  nb <= 2048. */
int cpwgpu_rayleigh_ritz(int mtxd, int nb, double *psi_dev, long ldp, double *hpsi_dev, long ldh,
                         int rotate, double *eig, double *res, void *stream);

/* ---- iterative diagonalisation on the device --------------------------------- */

/* Lowest neig eigenpairs of the plan's H, everything on the device; replaces
 * Src/ditsp/ditsp_c16.f90 (with ditsp_c16_start / _loop / _ritz / ditsp_error)
 * for a caller whose arrays live on the GPU:
 *   prec        CPWGPU_PREC_FP64, or CPWGPU_PREC_FP32 = H in FP32-mixed until
 *               the residuals reach ~1e-5, then FP64 H to the end (the result
 *               is always to FP64 accuracy)
 *   icmax       maximum number of iterations (each adds a block of relaxed
 *               residuals and makes one Rayleigh-Ritz step)
 *   iguess      1: psi_dev holds neig guess vectors; otherwise the start is
 *               the eigenvectors of the small matrix (leading mtxds x mtxds
 *               block of H), padded with zeros (ditsp_c16_start.f90:116,209-217)
 *   epspsi      convergence: max_n ||H psi_n - e_n psi_n||^2 < epspsi^2 over
 *               the neig bands (ditsp_c16.f90:375-377)
 *   mtxds       size of the small matrix; <= 0: chosen from diag(H) by the
 *               rule of Src/ALFA/size_mtxds.f90; >= mtxd: H is diagonalised
 *               directly (no iterations)
 *   hamsm_dev   device complex128 small matrix (ldhs, mtxds) or NULL: then
 *               built on the device as the leading block of H
 *   hdiag_dev   device diag(H) (mtxd) for the preconditioner, or NULL: the
 *               exact diagonal (ekpg + V(G=0) + sum_a xnlkb_a |anlga_ia|^2)
 *   psi_dev     (ldp, neig) in (guesses, iguess = 1) / out: orthonormal
 *               eigenvectors;  hpsi_dev (ldh, neig) out: H psi in FP64
 *   ei          host (neig) out: eigenvalues, ascending
 *   ifail       host out: 0 converged, else 1 + int(-log10(max residual^2))
 *               as ditsp_c16.f90:408-409 (may be NULL)
 *   stats       host int[5] out (may be NULL): iterations, H|psi> products,
 *               vectors iterated (neig + buffer), mtxds used, iteration at
 *               which FP64 H took over (-1: FP64 throughout)
 * Up to max(8, neig/8) buffer vectors beyond neig are iterated and not
 * required to converge.  The small eigenproblems use rocSOLVER (zheevd,
 * zpotrf) and rocBLAS (ztrsm) on the calling stream.  Returns when done;
 * not converging within icmax is reported in ifail, not as an error. */
int cpwgpu_plan_ditsp(cpwgpu_plan_t plan, int prec, int icmax, int iguess, double epspsi, int neig,
                      int mtxds, const double *hamsm_dev, long ldhs, const double *hdiag_dev,
                      double *psi_dev, long ldp, double *hpsi_dev, long ldh, double *ei, int *ifail,
                      int *stats, void *stream);

/* Drop-in for the reference's subroutine ditsp_c16 (Src/ditsp/ditsp_c16.f90:20-27):
 * the same argument list, by reference, host arrays (column-major, leading
 * dimensions mxddim for psi/hpsi/anlga and mxdsml for hamsm).  Precision as
 * set by cpwgpu_set_precision.  psi/hpsi/ei are copied once per call; the H
 * products of the whole iteration stay on the GPU.  The argument checks of
 * ditsp_c16.f90:151-212 return CPWGPU_ERR_ARG / _MTXD / _NG (the reference
 * stops); neig < 1 prints the reference's warning and returns.  When mtxds >=
 * mtxd hpsi is also written (the reference leaves it unset there).  lnewanl
 * is never written; neig is not changed. */
int cpwgpu_ditsp_c16(const int *ipr, int *ifail, const int *icmax, const int *iguess, const double *epspsi,
                     const int *lnewanl, int *neig, const int *mtxd, const int *mtxds, double *psi,
                     double *hpsi, double *ei, const double *ekpg, const int *isort, const double *vscr,
                     const int *kmscr, const int *ng, const int *kgv, const double *anlga, const double *xnlkb,
                     const int *nanl, const double *hamsm, const double *hdiag, const int *mxddim,
                     const int *mxdsml, const int *mxdbnd, const int *mxdgve, const int *mxdscr,
                     const int *mxdanl);

/* ---- electron density ---------------------------------------------------------- */

/* rho(r) = sum_n occ_n |psi_n(r)|^2 on the plan's FFT mesh, psi_n(r) =
 * sum_G psi_n(G) exp(+i G.r) (the unnormalised transform of cfft_wf_c16 mode
 * -1), FP64: the band loop of Src/omp/charge_by_fft.f90:170-197 (array
 * rhomsh).  psi_dev (ldp, neig) device complex128; occ host (neig) or NULL
 * (= 1); rho_dev device double in the vscr layout rho[(k3*n2 + k2)*id + k1]
 * (id*n2*n3 values, zero for k1 >= n1).  The mesh must be one of the
 * specialised sizes.  Enqueued on stream. */
int cpwgpu_plan_density(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp, const double *occ,
                        double *rho_dev, void *stream);

/* ---- atomic forces ----------------------------------------------------------- */

/* Forces on the atoms from E = sum_n occ_n <psi_n| V_loc + V_NL |psi_n> at
 * fixed psi, FP64, replacing the force parts of Src/ALFA/for_str_nl_kb.f90
 * (:286-398) and Src/ALFA/for_str_local_force.f90 (:102-145) for this energy:
 *   psi_dev(ldp, neig)   device complex128 bands; occ (host, neig) or NULL = 1
 *   natom, rat(3,natom)  fractional positions; avec(3,3) lattice vectors as
 *                        columns (bohr)
 *   proj_atom(nanl)      atom (0-based) of each projector column of the plan
 *   nq, qv(3,nq)         q vectors (integers) of the position-dependent local
 *                        potential V(q) = sum_s fq(q,s) sum_{atoms of s} exp(-2 pi i q.rat)
 *   nspec, species(natom) (0-based), fq(nq, nspec) real form factors
 * Outputs (host): force(3, natom) Cartesian Ha/bohr, enl = E_NL (may be NULL).
 * The plan's mesh must be one of the specialised FFT sizes. */
int cpwgpu_plan_forces(cpwgpu_plan_t plan, int neig, const double *psi_dev, long ldp,
                       const double *occ, int natom, const double *rat, const double *avec,
                       const int *proj_atom, int nq, const int *qv, int nspec, const int *species,
                       const double *fq, double *force, double *enl);

/* ---- measurement hook (no counterpart in the reference) ------------------- */

/* Kernels of one H|psi> application, in launch order. */
#define CPWGPU_K_COL_BWD 0       /* scatter + inverse FFT along z                  */
#define CPWGPU_K_PLANE 1         /* y/x FFTs with the V(r) product                 */
#define CPWGPU_K_COL_FWD 2       /* forward FFT along z + gather + kinetic         */
#define CPWGPU_K_PROJ 3          /* <beta|psi> split-K GEMM                        */
#define CPWGPU_K_PROJ_REDUCE 4   /* split-K sum and D scaling                      */
#define CPWGPU_K_BACK 5          /* hpsi += beta D<beta|psi> GEMM                  */
#define CPWGPU_K_PROJ_KB 6       /* KB projector construction                      */
#define CPWGPU_NKERNEL 7
/* When on, each later launch of the kernels selected by mask (bit k = kernel
 * k) carries a HIP start/stop event pair in its dispatch on the launch stream
 * (the kernel's own start and end). */
int cpwgpu_profile(int on, unsigned mask);
/* Waits for the recorded events; ms[k], launches[k] (k < n) = summed
 * durations and launch counts per kernel since the previous read. */
int cpwgpu_profile_read(double *ms, long *launches, int n);

#ifdef __cplusplus
}
#endif
#endif
