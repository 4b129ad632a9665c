/*
 * sincdvr.h -- C-ABI of the MI355X (gfx950) sinc-DVR orbital-operator library.
 *
 * This is the drop-in boundary for the SINCDVR project routines of
 * LBNL-AMO-MCTDHF.  The reference's Fortran driver calls, unchanged, a thin
 * iso_c_binding shim (v1-1_amd/fortran/sincdvr_shim.f90) that forwards to the
 * entry points below.  Arrays are the reference's DATATYPE arrays: complex*16,
 * column-major (first index fastest), interleaved (re, im) doubles, passed as
 * plain pointers.  Sizes follow the reference's explicit-shape dummies.
 *
 * Every entry point returns SDV_OK (0) or a negative SDV_ERR_* code; the
 * message is available from sdv_last_error().  The reference has no return
 * codes: its fatal conditions print and call mpistop (CFLST,
 * S/Definitions.ALL:45); the Fortran shim maps a nonzero code to the same stop.
 *
 * Host-pointer entry points are synchronous (all device work done on return).
 * The *_dev variants take device pointers and a hipStream_t (as void*) and
 * return after enqueueing; the caller synchronises the stream.
 */
#ifndef SINCDVR_H
#define SINCDVR_H

#ifdef __cplusplus
extern "C" {
#endif

#define SDV_OK 0
#define SDV_ERR_ARG (-1)         /* bad argument (size, range, null)          */
#define SDV_ERR_HIP (-2)         /* HIP runtime error                         */
#define SDV_ERR_UNSUPPORTED (-3) /* grid size / option not built              */
#define SDV_ERR_NOMEM (-4)       /* device allocation failed                  */
#define SDV_ERR_NODEVICE (-5)    /* no gfx950 device visible                  */

typedef struct sdv_ctx sdv_ctx; /* opaque: one grid, resident kernel + KE tables */

/* Project parameters: the subset of `myparams` (SD/sinc_params.f90:6-102)
 * that the hot path reads, plus the build flavour. */
typedef struct sdv_config {
  int numpoints;          /* n: points per axis, cube only (numpoints(1:3), :40)        */
  double spacing;         /* grid spacing in bohr (:41)                                 */
  int toothnbig;          /* 240 by default (:33)                                       */
  int toothnsmall;        /* 40 by default (:32)                                        */
  int numspf;             /* orbitals per call (myparams numspf)                        */
  int cnormflag;          /* 0 hermnorm (CONJUGATE=conjg), 1 ECS c-norm (CONJUGATE=)    */
  int scalingflag;        /* 1 = smooth exterior complex scaling (:54)                  */
  double scalingtheta;    /* :57                                                        */
  double scalingstretch;  /* :58                                                        */
  double scalingdistance; /* :55                                                        */
  double smoothness;      /* :56                                                        */
  int notwoflag;          /* call_twoe_matel00 returns zeros (SD/coreproject.f90:1021)  */
  int debugflag;          /* 4040/4242/4343: op_tinv returns zeros (:260-263)           */
  int precision;          /* 0: FP64 plain drop-in; 1: FP32 butterflies (mixed form)    */
  int maskflag;           /* 1 = kinetic apply through the edge mask when scalingflag = 0
                             (SD/coreproject.f90:1355-1376, sinc_params.f90:73)          */
  int masknumpoints;      /* mask width in points per end (SD/init_sinc_new.f90:580-599) */
  int velflag;            /* pulse gauge of sdv_actreduced00 (S/derivs.f90:742-750): 0 length
                             (lenmultiply), else velocity (velmultiply,
                             SD/coreproject.f90:2119-2134)                             */
} sdv_config;

/* Fill `cfg` with the reference defaults (SD/sinc_params.f90). */
void sdv_config_default(sdv_config* cfg);

/* Replaces the project-side set-up of myprojectalloc / init_project /
 * get_twoe_new (SD/sincproject.f90:46-113,311-338; SD/init_sinc_new.f90:395-529):
 * builds on the device the Coulomb kernel threed_two, its circulant spectrum,
 * the sinc KE stencil and the ECS Jacobian tables. */
int sdv_init(const sdv_config* cfg, sdv_ctx** ctx);
int sdv_free(sdv_ctx* ctx);
const char* sdv_last_error(void);
int sdv_set_precision(sdv_ctx* ctx, int precision);

/* Nuclear attraction potential of get_3dpoisson / get_twoe_new
 * (SD/sincproject.f90:192-250,311-338): pot(totpoints) real, for numcenters
 * nuclei at centershift(3, numcenters) half-spacing offsets with charges
 * nuccharges(numcenters).  realkind4 = 1 evaluates the far-field sqrt in
 * 4-byte REAL as builds without -r8 do (:220); 0 in 8-byte REAL. */
int sdv_nuclear_potential(sdv_ctx* ctx, int numcenters, const int* centershift, const double* nuccharges,
                          int realkind4, double* pot);

/* op_tinv (module tinvsubmod, SD/coreproject.f90:250-258):
 *   twoeden03(totpoints, allsize) in, twoereduced(totpoints, allsize) out.
 *   allsize must be a multiple of circsize (:387-391). */
int sdv_op_tinv(sdv_ctx* ctx, const double* twoeden03, double* twoereduced, int allsize, int circsize);
int sdv_op_tinv_dev(sdv_ctx* ctx, const void* twoeden03, void* twoereduced, int allsize, int circsize,
                    void* stream);
/* Allocation-free op_tinv.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * sdv_op_tinv_dev (bitwise: the same kernels per density); it calls no
 * hipMalloc/hipFree.  Its work memory is one library-owned 128 MiB device
 * array per GPU, held until the process ends; the densities are convolved in
 * chunks that fit it, and a grid whose single convolution does not fit is
 * refused with SDV_ERR_UNSUPPORTED (every supported grid fits: 17 MB per
 * density at 81^3 in FP64).  Every call on one GPU shares that array, so
 * calls that overlap on one GPU (different streams or host threads) race. */
int sdv_op_tinv_noalloc_dev(sdv_ctx* ctx, const void* twoeden03, void* twoereduced, int allsize, int circsize,
                            void* stream);

/* call_twoe_matel00 (SD/coreproject.f90:945-955):
 *   inspfs10, inspfs20 (totpoints, numspf) in;
 *   twoematel(numspf, numspf, numspf, lowspf:highspf) out;
 *   twoereduced(totpoints, numspf, lowspf:highspf) out.
 * When inspfs10 == inspfs20 (the all_matel call, S/matel.f90:759) each
 * unordered pair is convolved once and the mirror filled by symmetry. */
int sdv_call_twoe_matel00(sdv_ctx* ctx, int lowspf, int highspf, const double* inspfs10,
                          const double* inspfs20, double* twoematel, double* twoereduced);
int sdv_call_twoe_matel00_dev(sdv_ctx* ctx, int lowspf, int highspf, const void* inspfs10,
                              const void* inspfs20, void* twoematel, void* twoereduced, void* stream);

/* mult_ke (SD/coreproject.f90:2104-2116): in/out (totpoints, howmany). */
int sdv_mult_ke(sdv_ctx* ctx, const double* in, double* out, int howmany);
int sdv_mult_ke_dev(sdv_ctx* ctx, const void* in, void* out, int howmany, void* stream);

/* mult_reducedpot (SD/coreproject.f90:1200-1205):
 *   inspfs(totpoints, numspf), reducedpot(totpoints, numspf, firstspf:lastspf) in;
 *   outspfs(totpoints, firstspf:lastspf) out. */
int sdv_mult_reducedpot(sdv_ctx* ctx, int firstspf, int lastspf, const double* inspfs, double* outspfs,
                        const double* reducedpot);
int sdv_mult_reducedpot_dev(sdv_ctx* ctx, int firstspf, int lastspf, const void* inspfs, void* outspfs,
                            const void* reducedpot, void* stream);

/* get_reducedpot0 ZGEMM (S/mean.f90:556-558):
 *   reducedpot(totpoints, numspf^2) = twoereduced(totpoints, numspf^2) . twoden(numspf^2, numspf^2) */
int sdv_get_reducedpot(sdv_ctx* ctx, const double* twoereduced, const double* twoden, double* reducedpot);
int sdv_get_reducedpot_dev(sdv_ctx* ctx, const void* twoereduced, const void* twoden, void* reducedpot,
                           void* stream);

/* Pair potentials only: the convolution half of call_twoe_matel00
 * (SD/coreproject.f90:1039-1081), filling twoereduced(totpoints, numspf,
 * lowspf:highspf) without the matrix-element reduction. */
int sdv_twoereduced(sdv_ctx* ctx, int lowspf, int highspf, const double* inspfs10, const double* inspfs20,
                    double* twoereduced);
int sdv_twoereduced_dev(sdv_ctx* ctx, int lowspf, int highspf, const void* inspfs10, const void* inspfs20,
                        void* twoereduced, void* stream);

/* Number of convolutions sdv_twoereduced / call_twoe_matel00 run for
 * (lowspf, highspf) with one array as bra and ket (same = 1) or two (same =
 * 0).  No counterpart in the reference, which convolves every ordered pair
 * (SD/coreproject.f90:1074-1079); here each unordered pair once, and for
 * hermnorm two diagonal pairs (real densities |phi_a|^2) share one complex
 * convolution.  For throughput accounting. */
int sdv_twoe_convolutions(sdv_ctx* ctx, int lowspf, int highspf, int same);

/* One-electron matrix elements (S/matel.f90:241-293 sparseops_matel,
 * :113-169 pot_matel), the rows after mult_ke in the driver:
 *   opmatel(numspf, lowspf:highspf)  = inspfs1^H . mult_ke(inspfs2(:, lowspf:highspf))
 *   potmatel(numspf, lowspf:highspf) = inspfs1^H . (pot * inspfs2(:, lowspf:highspf))
 * (^H = conjugate transpose for hermnorm, MYGEMM CNORMCHAR 'C'; plain
 * transpose for cnorm, 'T').  pot is the real nuclear potential (totpoints),
 * e.g. from sdv_nuclear_potential; the contact, hydrogen-atom and frozen-core
 * terms of mult_pot/pot_matel are zero for the sinc-DVR project without
 * frozen orbitals and are not applied.  The kinetic apply runs in the
 * context's precision.  Grid sums are split over workgroups and reduced in a
 * fixed order (deterministic). */
int sdv_sparseops_matel(sdv_ctx* ctx, int lowspf, int highspf, const double* inspfs1, const double* inspfs2,
                        double* opmatel);
int sdv_sparseops_matel_dev(sdv_ctx* ctx, int lowspf, int highspf, const void* inspfs1, const void* inspfs2,
                            void* opmatel, void* stream);
int sdv_pot_matel(sdv_ctx* ctx, int lowspf, int highspf, const double* inspfs1, const double* inspfs2,
                  const double* pot, double* potmatel);
int sdv_pot_matel_dev(sdv_ctx* ctx, int lowspf, int highspf, const void* inspfs1, const void* inspfs2,
                      const double* pot, void* potmatel, void* stream);

/* The orbital operator of the propagator's Krylov matvec, actreduced00
 * (S/derivs.f90:629-852) for the sinc-DVR project (one bond point, no frozen
 * orbitals, scalar or constraint terms), over all numspf orbitals:
 *   w       = inspfs . denmat                        (spfinvr = spfinvrsq = spfr)
 *   spfmult = mult_ke(w) + pot * w + pulse(w) + mult_reducedpot(inspfs, reducedpot)
 *             with the pulse values tdpot = {x.re, x.im, y.re, y.im, z.re, z.im}
 *             (host, or null for no field; what vectdpot returns for the gauge):
 *             velflag = 0: lenmultiply (S/proputils.f90:157-182), the field
 *             values times the dipoles (grid coordinates, complex-scaled under
 *             ECS, SD/init_sinc_new.f90:527,619-645);
 *             velflag != 0: velmultiply (SD/coreproject.f90:2119-2134), the
 *             first-derivative sinc matrix fdtot per axis (SD/SineDVR.f90:25-33)
 *             times -i A_d, through mult_general option 2 (:1321-1379):
 *             invsqrtjacobian on both sides under ECS (SD/init_sinc_new.f90:504),
 *             the edge mask under maskflag
 *   dentimeflag: spfmult = timefac * spfmult . invdenmat
 *   projflag:    out = spfmult - projspfs . <projspfs | spfmult>   (project00,
 *                S/proputils.f90:10-60, flavour dot), else out = spfmult.
 * denmat, invdenmat: numspf x numspf complex, column-major; reducedpot:
 * (totpoints, numspf, numspf); pot: real (totpoints) or null. */
int sdv_actreduced00(sdv_ctx* ctx, const double* inspfs, const double* projspfs, const double* denmat,
                     const double* invdenmat, const double* reducedpot, const double* pot, const double* tdpot,
                     int dentimeflag, int projflag, double timefac_re, double timefac_im, double* outspfs);
int sdv_actreduced00_dev(sdv_ctx* ctx, const void* inspfs, const void* projspfs, const void* denmat,
                         const void* invdenmat, const void* reducedpot, const double* pot, const double* tdpot,
                         int dentimeflag, int projflag, double timefac_re, double timefac_im, void* outspfs,
                         void* stream);

/* Tuning (no counterpart in the reference, whose circulant edge is fixed at
 * 2*numpoints, SD/coreproject.f90:389-391).  Any circulant edge M = 2m with
 * m >= numpoints gives the same linear convolution; sdv_circulant_sizes lists
 * the supported half-edges m (returns their count, writes at most cap).
 * sdv_set_circulant rebuilds the kernel spectrum for half-edge m (0 = the
 * smallest supported m >= numpoints); a batch built on the context before the
 * change keeps the spectrum it was built with (it holds its own reference and
 * its own pair lists, so the context may also be freed while the batch
 * lives).  sdv_set_pair_batch sets how many pair potentials
 * one launch of the convolution pipeline carries (work buffer = pairs x 2m x
 * numpoints^2 complex); 0 (the default) = every pair of the call, up to an
 * 8 GiB work buffer. */
int sdv_circulant_sizes(int* half_edges, int cap);
int sdv_set_circulant(sdv_ctx* ctx, int half_edge);
int sdv_set_pair_batch(sdv_ctx* ctx, int pairs);

/* Batched form: the pair potentials of many independent orbital sets, of
 * uneven grid sizes and orbital counts, as one step of a few launches (see
 * the schedules below), capturable as one hipGraph.  Each set is a context (its grid, kernel
 * spectrum, numspf) with device arrays inspfs(totpoints, numspf) and
 * twoereduced(totpoints, numspf, numspf); the pair-symmetric path of
 * call_twoe_matel00 (one array as bra and ket, lowspf..highspf = 1..numspf)
 * is computed for every set.  No counterpart in the reference (which runs
 * one orbital set per process); the per-set result equals that of
 * sdv_twoereduced_dev on the same set.
 * sdv_batch_capture records one step into a hipGraph on `stream`; later
 * sdv_batch_run calls on that stream replay it (a schedule change drops the
 * graph).  sdv_batch_set_schedule picks the schedule (autotuning):
 *   SDV_BATCH_GROUPED (default): one launch per stage per circulant size,
 *     each with the block shape and LDS of its size; sizes run on up to
 *     `nstreams` forked streams (parallel branches of the captured graph);
 *   SDV_BATCH_FUSED: one persistent launch per stage over every set, grid of
 *     grid_mult x (compute units) workgroups. */
typedef struct sdv_batch sdv_batch;
typedef struct sdv_batch_set {
  sdv_ctx* ctx;
  int numspf;
  const void* inspfs;
  void* twoereduced;
} sdv_batch_set;
int sdv_batch_create(int nsets, const sdv_batch_set* sets, int precision, sdv_batch** batch);
#define SDV_BATCH_FUSED 0
#define SDV_BATCH_GROUPED 1
int sdv_batch_set_schedule(sdv_batch* batch, int mode, int grid_mult, int nstreams);
int sdv_batch_run(sdv_batch* batch, void* stream);
int sdv_batch_capture(sdv_batch* batch, void* stream);
/* npairs: pair potentials delivered per step (the unordered pairs of every
 * set); work_bytes: the convolution work buffer. */
int sdv_batch_stats(sdv_batch* batch, long long* npairs, long long* work_bytes);
int sdv_batch_free(sdv_batch* batch);
/* Per-stage timing of sdv_batch_run (stages XFWD, YZ, XINV), as sdv_profile /
 * sdv_profile_read below; graph replays are not timed per stage. */
int sdv_batch_profile(sdv_batch* batch, int enable);
int sdv_batch_profile_read(sdv_batch* batch, double* ms, int* count);

/* Serving: a request engine for many concurrent propagations (no
 * counterpart in the reference, which serves one propagation per process).
 * Requests: SDV_REQ_COULOMB = the pair potentials twoereduced(totpoints,
 * numspf, numspf) of one orbital set (as call_twoe_matel00's pair-symmetric
 * path), SDV_REQ_KE = mult_ke of its numspf orbitals; inspfs is a device
 * array (totpoints, numspf) that must stay unchanged until the request is
 * done.  hbm_budget is the server's whole device footprint: two tick work
 * buffers of tick_bytes and, in the rest, one arena for the outputs (the
 * inputs and the kernel spectra belong to the caller's contexts); a queued
 * request is admitted (first in, first out) when its output fits, and its
 * block returns to the arena at sdv_server_release.  Admitted pair work is
 * cut into ticks of tick_bytes of convolution work buffer; one tick mixes
 * pairs of several requests in one grouped launch (continuous batching), two
 * ticks are in flight on two streams (the device starts the next while the
 * last workgroups of the previous drain).  precision 1 = FP32 butterflies,
 * 0 = FP64.
 * sdv_server_poll retires finished ticks, admits and launches (wait = 1:
 * until every admitted and admissible request is done); sdv_server_result
 * gives the state (0 queued, 1 admitted, 2 done), the device output when done
 * and the submit-to-done and submit-to-admit times in milliseconds.  A
 * request's result is the same alone and in any mix.  Not thread-safe; the
 * contexts must outlive their requests. */
#define SDV_REQ_COULOMB 0
#define SDV_REQ_KE 1
typedef struct sdv_server sdv_server;
int sdv_server_create(long long hbm_budget, long long tick_bytes, int precision, sdv_server** server);
int sdv_server_submit(sdv_server* server, sdv_ctx* ctx, int kind, const void* inspfs, long long* request_id);
int sdv_server_poll(sdv_server* server, int wait, int* ncompleted);
int sdv_server_result(sdv_server* server, long long request_id, void** out, int* state, double* latency_ms,
                      double* queue_ms);
/* copy a done request's output (out_bytes of its kind) to device memory dst,
 * enqueued on `stream`; a release right after it is safe: the output's block
 * returns to the arena only once the copy has finished.  Other readers of the
 * pointer that sdv_server_result returns must finish before the release. */
int sdv_server_copy_result(sdv_server* server, long long request_id, void* dst, void* stream);
int sdv_server_release(sdv_server* server, long long request_id);
int sdv_server_stats(sdv_server* server, long long* queued, long long* in_flight, long long* arena_used,
                     long long* arena_peak, long long* ticks);
int sdv_server_free(sdv_server* server);

/* Per-kernel timing with HIP events recorded on the launch stream, the
 * library's counterpart of the reference's twoematel.time.dat / zke.time.dat
 * counters (SD/coreproject.f90:1114-1142,1499-1517).  sdv_profile(ctx, 1)
 * clears and starts recording; sdv_profile_read synchronises and returns, per
 * stage, the summed milliseconds and the launch count (arrays of
 * SDV_NSTAGES). */
#define SDV_STAGE_XFWD 0
#define SDV_STAGE_YZ 1
#define SDV_STAGE_XINV 2
#define SDV_STAGE_MATEL_ASSEMBLE 3
#define SDV_STAGE_MATEL_GEMM 4
#define SDV_STAGE_MATEL_REDUCE 5
#define SDV_STAGE_KE 6
#define SDV_NSTAGES 7
int sdv_profile(sdv_ctx* ctx, int enable);
int sdv_profile_read(sdv_ctx* ctx, double* ms, int* count);

/* Inspection (tests): the kernel threed_two on its first octant, n^3 reals
 * K(i,j,k) for i,j,k in [0,n) (SD/sincproject.f90:182-187); the 1-D KE
 * stencil ketot(1-n:n-1) (SD/SineDVR.f90:32); the circulant edge M in use. */
int sdv_get_kernel_octant(sdv_ctx* ctx, double* koct);
int sdv_get_ke_stencil(sdv_ctx* ctx, double* ketot);
int sdv_circulant_edge(sdv_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* SINCDVR_H */
