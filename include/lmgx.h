/* lmgx - MI355X-native pressure-Poisson multigrid for LESter's multiblock layout.
 *
 * C ABI of liblmgx.so.  Plain pointers and sizes only.  Every entry point returns 0 on success
 * or a negative LMGX_E* code; lmgx_last_error() gives a message for the calling thread.
 *
 * Data layout (same as the reference): multi-level arrays concatenate the blocks of level 1,
 * then the blocks of level 2, ... (reference code/set_offset.F:166-234); inside a block the
 * index is (i-1) + nip1*(j-1) + nip1*njp1*(k-1), i fastest, one halo layer per side
 * (code/norms_block.F:36-45).  Level and block indices in this API are 0-based.
 *
 * Threading: one host thread per handle at a time.  A handle owns its device memory and one
 * HIP stream; host pointers are borrowed for the duration of a call only.
 *
 * Device arrays (the *_device entry points) are read and written on the handle's own stream,
 * which is synchronised before the call returns.  Work the caller has queued on other streams
 * is not waited for: the caller makes its inputs ready first (synchronises the stream that
 * produced them; the Python binding does this for torch's current stream).
 */
#ifndef LMGX_H
#define LMGX_H

#ifdef __cplusplus
extern "C" {
#endif

#define LMGX_OK 0
#define LMGX_E_ARG (-1)      /* bad argument / topology */
#define LMGX_E_LEVELS (-2)   /* more levels requested than the block dims allow */
#define LMGX_E_FACE (-3)     /* interface faces do not match */
#define LMGX_E_HIP (-4)      /* HIP runtime error */
#define LMGX_E_STATE (-5)    /* call out of order (e.g. solve before set_operator) */
#define LMGX_E_NUMERIC (-6)  /* NaN/Inf in the residual norms */
#define LMGX_E_UNSUPPORTED (-7)
#define LMGX_E_FILE (-8)     /* field dump: cannot open / read / write, or the file does not fit */

/* boundary codes, reference code/bc_type.inc:17-27 (only these two change the solver) */
#define LMGX_BC_PERIODIC 12
#define LMGX_BC_INTERFACE 13

typedef struct lmgx_handle lmgx_handle;

typedef struct lmgx_topology {
    int nblk;              /* blocks (reference numbl) */
    int nlev;              /* levels wanted, 0 = as many as the dims allow (nummgl = 0) */
    const int *nijkp1;     /* [nblk][3] level-1 dims incl. halos (bl_nijkp1) */
    const int *bc;         /* [nblk][6] side codes (bl_bc) */
    const int *bl2bl;      /* [nblk][6] partner block, 1-based, 0 = none (bl2bl) */
    const int *bl2side;    /* [nblk][6] partner side 1..6 (bl2side) */
} lmgx_topology;

/* smoother */
#define LMGX_SMOOTH_EXACT 0   /* reference order: 3 lexicographic SOR sweeps per iteration */
#define LMGX_SMOOTH_LINE 1      /* zebra y-line relaxation (lines coloured by i+k), colour-split layout */
#define LMGX_SMOOTH_SPORLX2 2   /* reference order, meth_smooth = 2: 7 natural-order SOR sweeps per
                                   iteration, no jumble (code/sporlx2.F:11-161) */
/* precision */
#define LMGX_PREC_FP64 0
#define LMGX_PREC_MIXED 1     /* FP32 V-cycle inside FP64 iterative refinement (needs SMOOTH_LINE) */

typedef struct lmgx_params {
    int cycle;            /* 1 = V-cycle (meth_cycle), 2 = F-cycle (exact smoother only) */
    int numpre, numpost;  /* smoothing iterations per level (numpre/numpost) */
    int numcits;          /* smoothing iterations on the coarsest level (numcits) */
    int max_cycles;       /* upper bound on V-cycles (nummgits) */
    double omega;         /* relax_pmg */
    double omega_c;       /* relax_pmgc */
    double tolf;          /* reference stop test enorm/vnorm <= tolf (lmg_solve.F:82) */
    double rtol;          /* extra stop test rnorm_k/rnorm_0 <= rtol, 0 = off */
    int smoother;         /* LMGX_SMOOTH_*: EXACT and SPORLX2 are the reference's meth_smooth 1
                             and 2 (code/smooth.F:62-73); another value returns
                             LMGX_E_UNSUPPORTED where the reference stops (smooth.F:77-79) */
    int precision;        /* LMGX_PREC_* (MIXED needs LMGX_SMOOTH_LINE) */
    int check_every;      /* fast path: test convergence every N cycles (>= 1) */
    int sweeps_per_iter;  /* fast path: zebra line sweeps per smoothing iteration (reference SOR: 3) */
    int coarse_op;        /* fast path: 0 = coarse levels as given (re-discretised, as coeff_p),
                             1 = built from level 1 (half the piecewise-constant Galerkin product) */
    int coarse_solve;     /* fast path: 0 = numcits relaxations on the coarsest level (solversc),
                             1 = exact solve (Cholesky inverse) when the coarsest operator is SPD */
    int fine_post;        /* fast path: post-smoothing line sweeps on level 1 in a "lean" cycle,
                             0 = numpost * sweeps_per_iter (every cycle alike).  With fewer sweeps
                             than that, a cycle is lean unless the stop test (rtol) is predicted
                             to end the solve with that cycle only when it is done in full: the
                             post-smoothing of level 1 mostly cleans the residual the stop test
                             measures, which the next cycle's pre-smoothing cleans anyway. */
} lmgx_params;

/* Default parameters: the reference's hard-wired knobs (code/defaults.F:59-71). */
void lmgx_default_params(lmgx_params *p);

const char *lmgx_last_error(void);

/* Replaces the table set-up of code/set_offset.F:8-263, code/set_loop_bound.F:11-125 and
 * code/set_interfaces.F:6-256 (level count, per-level per-block offsets, interior ranges, face
 * copy tables) and allocates the device mirrors. */
int lmgx_init(const lmgx_topology *topo, lmgx_handle **out);
int lmgx_free(lmgx_handle *h);

/* Level tables: nlev, total points over all levels, level-1 points, and per level/block the
 * dims and offsets (arrays sized [LMGX_MAXL][nblk]...; any pointer may be NULL). */
#define LMGX_MAXL 8
int lmgx_levels(const lmgx_handle *h, int *nlev, long *ntot, long *n1, int *nijkp1_all,
                long *off_all);

/* Stencil coefficients ap..ar for every level (host arrays, ntot doubles each), as produced by
 * code/coeff_p.F:5-78 and consumed by code/lmg_solve.F:5 (its first seven arguments). */
int lmgx_set_operator(lmgx_handle *h, const double *ap, const double *aw, const double *ae,
                      const double *as, const double *an, const double *al, const double *ar);
/* Same, from device pointers (e.g. torch tensors). */
int lmgx_set_operator_device(lmgx_handle *h, const double *const d_coeff[7]);

/* The solve.  Replaces code/lmg_solve.F:5-172 with everything it calls: smooth.F, sporlx.F,
 * solversc.F, norms.F/norms_block.F, restriction.F/rest_block.F, prolongation.F/prol_block.F,
 * swap_sc0.F, init_array.F.  f: initial guess in (level 1), solution out (all levels); b: source
 * in (level 1), coarse right-hand sides out; residu: residuals out (all levels).  Host arrays of
 * ntot doubles.  history (may be NULL): [max_cycles][nlev][4] = (rnorm, fnorm, enorm, vnorm) after
 * each cycle, as in /cbmgdata2/ (cb.control.inc:156-164); the fast path (smoother =
 * LMGX_SMOOTH_LINE) measures level 1 only and reports zeros for the coarser levels;
 * ncycles_out: cycles run. */
int lmgx_solve(lmgx_handle *h, double *f, double *b, double *residu, const lmgx_params *prm,
               double *history, int *ncycles_out);
/* Same on caller-owned device arrays (ntot doubles each, stream-ordered on the handle's stream,
 * synchronised before return).  The three arrays must not overlap (LMGX_E_ARG): the solve reads b
 * in place while it writes f and residu. */
int lmgx_solve_device(lmgx_handle *h, double *d_f, double *d_b, double *d_res,
                      const lmgx_params *prm, double *history, int *ncycles_out);

/* Face-Dirichlet sides ([nblk][6] flags, NULL = none) of an operator given through
 * lmgx_set_operator (lmgx_build_operator records them itself).  The reference has none (every
 * physical side is Neumann, code/coeff_p_block.F:160-228); the fast path uses them to extend the
 * coarse correction oddly (face value 0) instead of the reference's zero-gradient copy
 * (code/prol_block.F:103-135).  The reference-order path ignores them. */
int lmgx_set_dirichlet(lmgx_handle *h, const int *flags);

/* The SAVEd sweep-order counter of code/sporlx.F:58 (persists across solves, like LESter). */
int lmgx_get_jumble(const lmgx_handle *h);
int lmgx_set_jumble(lmgx_handle *h, int jumble);

/* After-solve steps of code/step.F:376-380: code/extrap_p.F:5-48 (zero-gradient halos on
 * physical sides, then halo swap) and code/p_ref_subtract.F:6-100 (subtract p at cell (2,2,2)
 * of block 1 from every point of level 1).  Device array of ntot (or n1) doubles. */
int lmgx_extrap_p_device(lmgx_handle *h, double *d_p);
int lmgx_p_ref_subtract_device(lmgx_handle *h, double *d_p);

/* code/global_sum.F:105-147: sum, max and min of f over the interior of level l. */
int lmgx_global_sum_device(lmgx_handle *h, int level, const double *d_f, double out[3]);

/* Mean-pressure gauge, the alternative to code/p_ref_subtract.F:6-100 (which subtracts the
 * value at one cell): subtract the interior mean of level 1 (the FP64 sum of lmgx_global_sum
 * divided by the interior count npts_all(1), code/set_loop_bound.F:101-102) from every point of
 * level 1, halos included.  The reference has no such mode.  mean_out (may be NULL): the mean
 * subtracted.  Device array of n1 (or ntot) doubles. */
int lmgx_p_mean_subtract_device(lmgx_handle *h, double *d_p, double *mean_out);

/* The same mean-pressure gauge for callers that cannot allocate memory (code under graph
 * capture, tight serving loops, harnesses that forbid allocation); no counterpart in the
 * reference.  It calls no hipMalloc/hipFree and no allocating library routine: the workgroup
 * partials live in a static device array of 8192 partials (about 192 KiB) that every GPU the
 * library runs on holds until the process ends.  Level 1 may need at most 8192 partials (one per
 * 64 x 4 column tile of each block's interior), else LMGX_E_UNSUPPORTED.  d_mean: device
 * pointer receiving the mean, or NULL.  Same result as lmgx_p_mean_subtract_device, bit for
 * bit.  All calls on one GPU share the array: calls that overlap on one GPU (other host
 * threads, other handles) race and must be serialised by the caller. */
int lmgx_p_mean_subtract_device_noalloc(lmgx_handle *h, double *d_p, double *d_mean);

/* Single passes (device arrays, level 0-based), exposed for testing and for callers that drive
 * their own schedule: code/norms.F:5-126, code/sporlx.F:5-203 (one call = 3 sweeps),
 * code/swap_sc0.F:6-293, code/restriction.F:5-68 (+ init_array of the coarse f),
 * code/prolongation.F:5-75.  norms_out = (rnorm, fnorm, enorm, vnorm). */
int lmgx_norms_device(lmgx_handle *h, int level, const double *d_x, const double *d_b,
                      double *d_res, double norms_out[4]);
int lmgx_sporlx_device(lmgx_handle *h, int level, double *d_f, const double *d_b, double omega);
/* code/sporlx2.F:11-161 (one call = 7 natural-order sweeps, jumble untouched) */
int lmgx_sporlx2_device(lmgx_handle *h, int level, double *d_f, const double *d_b, double omega);
int lmgx_swap_device(lmgx_handle *h, int level, double *d_var);
int lmgx_restrict_device(lmgx_handle *h, int coarse_level, const double *d_res, double *d_b,
                         double *d_f);
int lmgx_prolong_device(lmgx_handle *h, int fine_level, double *d_f);

/* Operator set-up on the device (reference code/coeff_p_block.F:69-228 on every level and
 * code/sour_p_block.F:89-167 on level 1, with the metric pipeline of code/mg_xyz.F,
 * code/metric_block2.F, code/comp_metric.F in front).  xv1/yv1: level-1 interior vertices of
 * every block ([j=2..njp1][i=2..nip1] per block, blocks in order), zv1: [k=2..nkp1] per block.
 * rho: cell densities on level 1 incl. halos (n1 doubles, NULL = 1); when given, each face
 * coefficient is divided by the face density 0.5*(rho_P + rho_N) and coarse levels use 8-child
 * means.  dirichlet: [nblk][6] flags, 1 = face-Dirichlet p = 0 on that physical side (NULL =
 * none).  The operator is stored in the handle (as lmgx_set_operator would). */
int lmgx_build_operator(lmgx_handle *h, const double *xv1, const double *yv1, const double *zv1,
                        const double *rho, const int *dirichlet);
/* Per-step density change (the variable-density LES step): the coefficients of every level
 * rebuilt on the device from the metrics kept by lmgx_build_operator, for a new level-1 density
 * d_rho1 (device, n1 doubles with halos; NULL = 1), with the face-Dirichlet sides already set.
 * The same operator as lmgx_build_operator with that density, without the geometry pass and
 * the host transfer (reference: coeff_p, code/coeff_p.F:5-78, is re-run whenever the grid
 * changes; the density scaling has no reference counterpart). */
int lmgx_update_density_device(lmgx_handle *h, const double *d_rho1);
/* Copy the handle's operator to host (7 x ntot) - for checking and for hand-off to LESter. */
int lmgx_get_operator(const lmgx_handle *h, double *ap, double *aw, double *ae, double *as,
                      double *an, double *al, double *ar);
/* Level-1 source b = sour_p_block(u, v, w, rho, rhonp1, dtime) into d_b (device, n1 doubles);
 * requires a prior lmgx_build_operator (metrics). */
int lmgx_source_device(lmgx_handle *h, const double *d_u, const double *d_v, const double *d_w,
                       const double *d_rho, const double *d_rhonp1, double dtime, double *d_b);

/* Velocity correction after the pressure solve; replaces reference code/uvw_adv.F:5
 * uvw_adv(ivar1, diu, rp, dkw, u, u_int, rho, p, wk, offwork, step) with its block routine
 * code/uvw_adv_block.F:5 and the swap_sc of u at code/uvw_adv.F:59.  For ivar = 1 (U), 2 (V) or
 * 3 (W): u = u_int - 1.5 dtime (p_P - p_N) / d / (0.5 (rho_P + rho_N)) on the component's level-1
 * range, then the interface/periodic halos of u are swapped.  As at code/step.F:418-425, diu is
 * the U/W metric and the caller passes djv in its place for V; rho is rho^{n+1}.  u, u_int, rho:
 * level-1 layout (n1 doubles); p: level-1 part of the multi-level pressure; diu, rp: the
 * multi-level 2-D arrays (level-1 part read), dkw: the multi-level 1-D k array.  u keeps its
 * values outside the corrected range.
 * lmgx_uvw_adv: host arrays (copied over PCIe each call).
 * lmgx_uvw_adv_device: device arrays; d_diu, d_rp, d_dkw may all be NULL to use the metrics of
 * lmgx_build_operator (diu, or djv for V). */
int lmgx_uvw_adv(lmgx_handle *h, int ivar, const double *diu, const double *rp, const double *dkw,
                 double *u, const double *u_int, const double *rho, const double *p, double dtime);
int lmgx_uvw_adv_device(lmgx_handle *h, int ivar, const double *d_diu, const double *d_rp,
                        const double *d_dkw, double *d_u, const double *d_u_int,
                        const double *d_rho, const double *d_p, double dtime);

/* Field dumps in the reference's 3-D file format (the restart / warm-start files of LESter).
 * lmgx_file_3_device replaces code/file_3.F:5 file_3(dev, filename, filefrmt, f, ilev,
 * mg_or_sg, ierr): writes the device field d_f (multi-level natural layout) of level `level`
 * (0-based; mg_or_sg = 2, a single-grid field, means level 0) to `filename` as a Fortran
 * sequential file - filefrmt 0: unformatted records (4-byte length markers), numbl, the
 * (nip1, njp1, nkp1) of every block, then one record per (j, k) row of nip1 values, halos
 * included, as REAL*4 (real_kind 4, the reference's default) or REAL*8 (real_kind 8, its
 * -DWRITE_DOUBLE build); filefrmt 1: the same as text, each row written '(8(1x,g15.8))'.
 * lmgx_read_3_device replaces code/read_3.F:5 read_3(dev, filename, filefrmt, f, work, ierr):
 * reads such a file into level 0 of d_f.  A missing file gives zeros and *found = 0, as the
 * reference does; a block count or block dimension that differs from the handle's, or a short
 * record, returns LMGX_E_FILE (the message names the reference's ierr 3, 4 or 5). */
int lmgx_file_3_device(lmgx_handle *h, const char *filename, int filefrmt, const double *d_f,
                       int level, int mg_or_sg, int real_kind);
int lmgx_read_3_device(lmgx_handle *h, const char *filename, int filefrmt, double *d_f,
                       int real_kind, int *found);

/* Adjoint (no counterpart in the reference; for gradients of pressure functionals).
 *
 * lmgx_solve_adjoint_device: A^T lambda = g with the same multigrid run on the explicitly
 * transposed operator (row P's coupling to N = row N's coupling to P; across blocks, the
 * partner side's coefficient).  Same parameters, arrays (device, ntot doubles) and history as
 * lmgx_solve_device; lambda: in = initial guess on level 1, out = solution.
 * lmgx_transpose_operator_device: the transposed operator (7 device arrays of ntot).
 * lmgx_rho_adjoint_device: dJ/dlog(rho) for every level-1 entry of the density given to
 * lmgx_build_operator, from the converged p (A p = b) and lambda (A^T lambda = dJ/dp); level-1
 * device arrays (n1 doubles).  Entries the operator does not read get 0. */
int lmgx_solve_adjoint_device(lmgx_handle *h, double *d_lambda, double *d_g, double *d_res,
                              const lmgx_params *prm, double *history, int *ncycles_out);
int lmgx_transpose_operator_device(lmgx_handle *h, double *const d_out[7]);
int lmgx_rho_adjoint_device(lmgx_handle *h, const double *d_lambda, const double *d_p,
                            double *d_glogrho);

/* Line kernels chosen by the fast path's autotuner: out[precision][LMGX_MAXL], precision 0 =
 * FP64, 1 = FP32; 64 / 128 = two-ended line solve with that many lines per wavefront row, 16 =
 * segmented solve in 16-wavefront workgroups, 8 = segmented solve in 8-wavefront workgroups; 0
 * for levels that do not exist or before the first fast solve. */
int lmgx_tile_widths(const lmgx_handle *h, int *out);

/* The line-kernel autotuner's choices, to reuse them (no counterpart in the reference):
 * out[2][2][LMGX_MAXL] = [precision 0 FP64 / 1 FP32][0: line kernel as lmgx_tile_widths, 1: halo
 * stores as lmgx_fused_halos][level]; zeros for a precision not tuned yet.
 * lmgx_set_tuning installs such a table: every precision with a nonzero kernel entry is not timed
 * at its next fast solve but takes these choices (before a handle's first solve, or later, for
 * the next solve).  The segmented solve and the halo-store choice change the rounding, so
 * handles or processes that must give bitwise the same solve share one table; the autotuner,
 * which times the candidates, can choose differently from run to run.  A choice that a level
 * cannot run (the segmented solve on a non-symmetric level or on lines it does not hold, halo
 * stores where the topology allows none) makes that solve return LMGX_E_ARG. */
int lmgx_get_tuning(const lmgx_handle *h, int *out);
int lmgx_set_tuning(lmgx_handle *h, const int *in);

/* Levels whose face-interior halos the fast path's line kernels store themselves instead of a
 * swap launch after each half-sweep: out[precision][LMGX_MAXL] like lmgx_tile_widths, 1 = stored
 * by the line kernel, 0 = swapped.  A level qualifies when every swapped face is an i- or k-face
 * copying a whole plane of its partner with the same in-plane axes and the same colour; the
 * autotuner then times each line kernel both ways (a pinned kernel stores them).  All 0 before
 * the first fast solve of a precision, or with LMGX_NO_FUSED_SWAP=1 in the environment at the
 * first fast solve. */
int lmgx_fused_halos(const lmgx_handle *h, int *out);

/* Levels whose FP32 segmented line relaxation (mixed precision) reads binary16 copies of the
 * couplings, scaled by a power of two per level, and forms the diagonal from them plus the
 * level's Dirichlet excess ap - sum(a_nb): out[LMGX_MAXL], 1 = binary16, 0 = FP32.  A level
 * qualifies when its couplings are non-negative and, with the diagonal excess, span at most 2^28.
 * The relaxation then works on an operator whose edge weights are within 2^-11 of the true ones
 * (symmetric, zero row sums kept), which the FP64 refinement absorbs.  All 0 before the first
 * mixed solve or with LMGX_COEF16=0 in the environment at the first mixed solve after an operator
 * is set. */
int lmgx_coef16_levels(const lmgx_handle *h, int *out);

/* *out = 1 when the last fast solve ended on a cycle that wrote the returned solution and
 * residual itself: with lean cycles (lmgx_params.fine_post), the cycles the prediction runs in
 * full are those expected to end the solve, and their level-1 residual pass also stores the
 * residual and the solution into the caller's natural-layout arrays, so that a solve ending on
 * one skips the final conversion and residual pass (the same values, bit for bit).
 * LMGX_NO_FIN=1 in the environment at a solve keeps the final passes.  No counterpart in the
 * reference. */
int lmgx_finished_in_cycle(const lmgx_handle *h, int *out);

/* Whether level 1 takes the builder's operator in its metric form, out[2]:
 * out[1] = 1 when the last fast solve's level-1 FP64 residual passes formed the couplings of the
 * cells off the blocks' boundary planes from the densities (FP64) and the builder's face factors,
 * with its operations in its order (bitwise its operator), instead of reading them: for an
 * operator of lmgx_build_operator / lmgx_update_density_device built with densities, when
 * LMGX_METRIC64=1 in the environment at the first solve after an operator change (off by default).
 * out[0] = 1 when level 1's FP32 segmented line relaxation and restriction residual (mixed
 * precision) took it in metric form as well: when the last mixed solve recomputed
 * each cell's couplings from binary16 copies of the densities (times a power of two) and the
 * grid's face factors (binary16, one 64-bit word per cell), g dk / ((rho_a + rho_b) / 2) as
 * lmgx_build_operator builds them, instead of reading stored couplings.  Asked for with
 * LMGX_METRIC=1 in the environment at the first mixed solve after an operator change (the stored
 * binary16 couplings are the default), and only for the operator of lmgx_build_operator /
 * lmgx_update_density_device, when the densities are positive and their range and the factors'
 * fit binary16 (2^28).  No counterpart in the reference (an extension of its
 * coeff_p_block.F:104-137 coefficients with the 1/rho_face scaling). */
int lmgx_metric_form(const lmgx_handle *h, int *out);

/* The fast path's coarsest-level solve: out[0] = 1 when it is the exact solve (the coarsest
 * operator factored, its inverse applied per cycle), out[1] = 1 when the factor and inverse came
 * from the banded Cholesky (components in reverse Cuthill-McKee order, half-bandwidth out[2] at
 * most 63) rather than dense rocSOLVER, out[3] = connected components of blocks.  Zeros (and
 * out[2] = -1) before the first fast solve.  LMGX_COARSE_DENSE=1 in the environment at the
 * handle's first fast solve keeps the dense path. */
int lmgx_coarsest_info(const lmgx_handle *h, int *out);

/* Kernel timing of the last solve: total duration (ms), count and updated cells of the dominant
 * smoother kernel, measured with HIP events on the handle's stream.  The fast path times the
 * finest level's plain half-sweeps (not the two zero-start half-sweeps that open each cycle's
 * correction, which read fewer words). */
int lmgx_profile(lmgx_handle *h, int enable);
int lmgx_profile_read(const lmgx_handle *h, double *smoother_ms_total, long *smoother_launches,
                      double *smoother_cells_total);

#ifdef __cplusplus
}
#endif
#endif
