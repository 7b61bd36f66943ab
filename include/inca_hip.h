/*
 * inca_hip.h -- C-ABI of libinca_hip.so, the MI355X (gfx950) engine behind
 * the INCA intracule drop-in.
 *
 * The reference (markelgtx/INCA_parallel) has no FFI: its hot path is the
 * Fortran module procedure intracule::calculate_intracule
 * (src/mod_intracule.f90:18), which reads module globals and writes files.
 * The project's Fortran host keeps that procedure and its globals and calls
 * the entry points below through ISO_C_BINDING interfaces; each entry point
 * names the reference code whose work it takes over.
 *
 * Conventions
 *  - Plain C types only; every function returns 0 on success or a negative
 *    INCA_E* code, and inca_last_error() gives the message.
 *  - Arrays use the reference's Fortran layouts: r(3,G) is point-major
 *    (x,y,z of point g at r[3g..3g+2]); tmn(nprim,3) is column-major.
 *  - Logical flags cross as int32 (0 = false).
 *  - Host arrays are owned by the caller and only read during the call;
 *    device memory is owned by the context until inca_destroy().
 *  - One context serves one host thread; it is not thread-safe.
 */
#ifndef INCA_HIP_H
#define INCA_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct inca_ctx inca_ctx;
typedef struct inca_batch inca_batch;

enum {
  INCA_OK = 0,
  INCA_EINVAL = -1,   /* bad argument or call order */
  INCA_EHIP = -2,     /* HIP runtime error */
  INCA_EIO = -3,      /* file could not be opened or read */
  INCA_EFORMAT = -4,  /* unsupported .dm2p format (mod_intracule.f90:192-194) */
  INCA_ENOMEM = -5,   /* device allocation failed */
  INCA_ERANGE = -6    /* size or angular momentum beyond what is supported */
};

enum { INCA_FP64 = 0, INCA_MIXED = 1 };

/* Create a context on HIP device `device` (one stream, default priority). */
int inca_create(int32_t device, inca_ctx **ctx);
int inca_destroy(inca_ctx *ctx);
/* Message of the last failing call (or of inca_create when ctx is NULL). */
const char *inca_last_error(const inca_ctx *ctx);

/* Primitive tables: geninfo Alpha, TMN, Xn/Yn/Zn (mod_geninfo.f90:16-19).
 * normalize != 0 multiplies every DM2 value by N_i N_j N_k N_l with the
 * normalisation of mod_intracule.f90:145-152 (calculate_intracule's
 * normalize_dm2p argument).  Also builds the primitive-pair table used by
 * screen 1 (JA8, mod_intrastuff.f90:32-123, called at
 * mod_intracule.f90:297-305). */
int inca_set_primitives(inca_ctx *ctx, int32_t nprim, const double *alpha,
                        const int32_t *tmn, const double *x, const double *y,
                        const double *z, int32_t normalize);

/* Multicentre radial x Lebedev grid with TFVC weights: gridpoints + becke
 * (mod_build_grid.f90:13-277, 474-575), called at mod_intracule.f90:91-97.
 * cent(3,nquad); nradc/nangc/sfalpha/Ps(nquad).  Returns the reference's
 * counters maxgrid, rgrid, rrgrid (mod_intracule.f90:416-418). */
int inca_grid_multicenter(inca_ctx *ctx, int32_t nquad, const double *cent,
                          const int32_t *nradc, const int32_t *nangc,
                          const double *sfalpha, const double *Ps,
                          int32_t nosym, int64_t *maxgrid, int64_t *rgrid,
                          int64_t *rrgrid);

/* Radial-scan grid: gridpoints2 (mod_build_grid.f90:281-408), called at
 * mod_intracule.f90:99-104.  tart(2,nblock) = (start,end) per block.
 * Returns the number of radii and of kept points. */
int inca_grid_radial_plot(inca_ctx *ctx, int32_t nblock, const double *tart,
                          const double *step, const int32_t *nang,
                          int32_t nosym, int64_t *nradi, int64_t *ngrid);

/* Cube grid: gridpoints3 (mod_build_grid.f90:412-440), called at
 * mod_intracule.f90:106-112.  Point order: x outer, y middle, z inner. */
int inca_grid_cube(inca_ctx *ctx, const double *center, const double *step,
                   const int32_t *np, int64_t *ngrid);

/* Explicit points r(3,n): intracule_at_zero / intracule_two_points
 * (mod_intracule.f90:114-129). */
int inca_grid_points(inca_ctx *ctx, int64_t n, const double *r);

/* Copy the current grid to the host (any pointer may be NULL):
 * r(3,G) = rrrg/rpg/rg, w = rweight (multicentre) or w_ang (radial scan),
 * w_vee = rweight_vee, center_map = point_center_map (1-based)
 * (mod_quadratures.f90:5-21). */
int inca_get_grid(inca_ctx *ctx, double *r, double *w, double *w_vee,
                  int32_t *center_map);
/* Radial-scan layout: radi(nradi) and smn(nradi) (mod_quadratures.f90:16-18) */
int inca_get_radial_layout(inca_ctx *ctx, double *radi, int32_t *smn);

/* Read a .dm2p file into device memory, with the reference's format
 * detection (mod_intracule.f90:164-195): flat 24-byte records, or Fortran
 * sequential records with 4-byte markers.  8-byte-marker and zero-marker
 * files fail with INCA_EFORMAT.  *format = 0 flat, 1 sequential.
 * *nrec = records stored (reading stops at EOF or at a bad marker; index
 * validity is applied by inca_evaluate). */
int inca_load_dm2p(inca_ctx *ctx, const char *path, int64_t *nrec,
                   int32_t *format);
/* Same records from memory: n packed 24-byte records (int32 i,j,k,l;
 * double D).  on_device != 0: `records` is a device pointer that is copied
 * device-to-device (the caller keeps ownership); the call first waits for
 * all work queued on the device, so records still being written by a
 * kernel on any stream are copied complete. */
int inca_set_quartets(inca_ctx *ctx, int64_t n, const void *records,
                      int32_t on_device);

/* Free the DM2 records and the scratch of the pre-pass and of past grid
 * loops, keeping the grid and the prepared quartets: for holding many
 * prepared molecules at once (a batch being assembled).  Later grid loops
 * allocate their scratch again; inca_prepare needs new records
 * (inca_load_dm2p / inca_set_quartets) and inca_autotune a new
 * inca_prepare.  No reference counterpart. */
int inca_release_scratch(inca_ctx *ctx);

/* Records per chunk of the pre-pass (default 2^26, at least 1024).  The
 * records are decoded, screened and fitted a chunk at a time and each
 * chunk's survivors appended in file order, so the pre-pass scratch is
 * bounded by the chunk (stride doubles per record) whatever the length of
 * the stream; results do not depend on the chunk size (the term-integral
 * sum is grouped by chunk).  The reference reads 40,960 records per batch
 * (mod_intracule.f90:66, 229-262); no counterpart as a setting. */
int inca_set_prepare_chunk(inca_ctx *ctx, int64_t records);

/* The quartet loop of calculate_intracule (mod_intracule.f90:198-388):
 * decode + normalise (:276-285), screen 1 (:297-305), prefactors and
 * polycoef on three axes (:307-341, mod_intrastuff.f90:127-194), screen 2
 * (:342-344) and the grid loop (:345-358) summed over all quartets.
 *   precision: INCA_FP64 (reference arithmetic) or INCA_MIXED (FP32
 *              Gaussian factors and polynomials, compensated FP32 tile sums
 *              folded into FP64).
 *   thresh:    the namelist thresh; lim = thresh / (nprim(nprim+1)/2).
 *   ivec:      optional host copy of Ivec(G) (may be NULL).
 *   counts:    [read, after screen 1, after screen 2, polycoef failures]
 *              (mod_intracule.f90:419-421; the last is the number of
 *              per-axis fits the reference zeroes at mod_intrastuff.f90:170).
 * Records are consumed up to the first invalid index (i outside 1..nprim,
 * or j,k,l outside 1..nprim) as the reference's reader does (:239-240). */
int inca_evaluate(inca_ctx *ctx, int32_t precision, double thresh,
                  double *ivec, int64_t *counts);

/* The pre-pass of inca_evaluate alone (decode, screens, prefactors,
 * polynomials, compaction: mod_intracule.f90:229-344) without the grid
 * loop; the prepared quartets stay on the device for inca_batch_create.
 * counts as in inca_evaluate.  No reference counterpart as a separate
 * call. */
int inca_prepare(inca_ctx *ctx, int32_t precision, double thresh,
                 int64_t *counts);

/* The grid loop alone (mod_intracule.f90:345-381) over the quartets the
 * last inca_prepare / inca_evaluate prepared, on the current grid: for
 * evaluating one DM2 on several grids.  INCA_MIXED needs quartets prepared
 * with INCA_MIXED (the FP64/FP32 split is made in the pre-pass). */
int inca_evaluate_prepared(inca_ctx *ctx, int32_t precision, double *ivec);

/* inca_evaluate_prepared without device allocation.  No reference
 * counterpart: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc, hipFree or allocating library routine
 * and does not wait: the grid loop over the quartets the last inca_prepare /
 * inca_evaluate prepared, on the current grid, is enqueued on `stream` (a
 * hipStream_t; NULL: the context's stream), and its result lands in ivec, a
 * DEVICE pointer to G doubles owned by the caller.  Every grid kind is
 * evaluated record by record and point by point with the per-point cutoff
 * off (cube grids are not contracted per axis, shell-quartet groups are not
 * used), and the FP32 range of INCA_MIXED always runs the mixed kernel: the
 * result is that of inca_evaluate_prepared after inca_set_point_cutoff(ctx,
 * 0), inca_set_cube_contraction(ctx, 0), inca_set_grouping(ctx, 0) and
 * non-zero FP32 tiles, up to summation order.  A mixed preparation whose
 * record split has not been made yet (grouped preparations make it at the
 * first record-by-record mixed evaluation or inca_get_mixed_split) runs
 * every record in FP64.  A grouped preparation of one pre-pass chunk keeps
 * its records uncompacted until their first record-by-record use (any of
 * the calls that read records, e.g. inca_get_prepared with n > 0); before
 * that, this call fails with INCA_EINVAL.  The tile records and partial
 * sums live in a static 32 MiB array per device, held until the process
 * ends: the grid must fit one partial row of G doubles per precision range
 * plus the tile records (about 4 million points in FP64, 2 million mixed);
 * a larger grid fails with INCA_ERANGE before anything is enqueued.  Every
 * call on a device shares that array and leaves its contents behind, so
 * calls that overlap on one device (any contexts, any streams) race. */
int inca_evaluate_prepared_noalloc(inca_ctx *ctx, int32_t precision, double *ivec,
                                   void *stream);

/* Integration centres of the multicentre grid, replacing centercalc
 * (mod_autogrid.f90:12-131): the origin, +-(R_j - R_k) for heavy-atom pairs
 * (all pairs when nohydro = 0), z snapped to 0 below 0.5 bohr, centres
 * within 0.1 bohr merged.  cartes(natoms,3) column-major as geninfo::cartes
 * (bohr).  radius: per-atom radii in bohr, or NULL for the Bragg-Slater
 * table of get_bs_radius (mod_autogrid.f90:142-185).  *nquad receives the
 * number of centres; if cap < *nquad nothing is written and INCA_ERANGE is
 * returned (call with cap = 0 to size the arrays).  cent(3,nquad), Ps and
 * sfalpha(nquad) as intrainfo::cent/Ps/sfalpha. */
int inca_autogrid(int32_t natoms, const int32_t *an, const double *cartes,
                  int32_t nohydro, int32_t betaone, const double *radius,
                  int32_t cap, int32_t *nquad, double *cent, double *Ps,
                  double *sfalpha);

/* One-electron functions on the current grid: the cube files of &cube_job
 * (cubegen/fwrite, mod_cubefile.f90:16-128) with the functions of
 * mod_functions.f90, on the primitives of inca_set_primitives (centres x, y,
 * z; the normalisation is not applied, as in Prim):
 *   kind 1: primitive `index` (1-based), Prim (:17-31);
 *   kind 3: orbital `index`, MoOr (:199-210): sum_p T(index,p) Prim_p over
 *           |T| > 1e-10;
 *   kind 5: density sum_i w_i MoOr_i^2, Density / Dens_a / Dens_b
 *           (:312-352; w_i = 2 for the closed shell, 1 per spin);
 *   kind 6: the reference's Laplacian sum_i 2 w_i MoOr_i sum_p T(i,p) sd_p,
 *           Lapl (:351-372; sd the analytic Laplacian of a primitive,
 *           :113-152).
 * T(nmo, nprim) column-major (wfxinfo T: coefficients of the unnormalised
 * primitives), w(nmo); both unused for kind 1.  values: host array of G.
 * kind 2 (atomic orbitals) needs the .log reader and is not provided. */
int inca_onee(inca_ctx *ctx, int32_t kind, int32_t index, int32_t nmo, const double *T,
              const double *w, double *values);

/* The Bragg-Slater radius in bohr of each atomic number an[0..n): Slater's
 * radii in angstrom times 1.8897, get_bs_radius (mod_autogrid.f90:142-185);
 * the table inca_autogrid uses when radius is NULL.  Host code. */
int inca_bragg_slater(int32_t n, const int32_t *an, double *radius);

/* Batched evaluation of many molecules (no reference counterpart: the
 * reference runs one molecule per process, inca_main.f90:76).  Every
 * context must be on `device` and hold a grid and prepared quartets
 * (inca_prepare or inca_evaluate); the batch copies both into its own
 * arena, so the contexts may be destroyed afterwards.  Quartets are sorted
 * by class (Lsum, Lx, Ly, Lz) per molecule; work items (molecule, 256 P
 * point tile, single-class chunk) are ordered largest first and consumed by
 * a persistent kernel from an atomic work counter; the whole evaluation is
 * captured into a hipGraph at creation and replayed by inca_batch_evaluate.
 * Each molecule runs in the precision its quartets were prepared for:
 * INCA_FP64, or INCA_MIXED (its FP64 range with the FP64 body, its FP32
 * range with the mixed one; the FP64 and FP32 ranges are sorted
 * separately).  On failure *out may still be set: read
 * inca_batch_last_error and call inca_batch_destroy. */
int inca_batch_create(int32_t device, int32_t nmol, inca_ctx *const *mols,
                      inca_batch **out);
int inca_batch_destroy(inca_batch *b);
const char *inca_batch_last_error(const inca_batch *b);
/* One replay of the captured graph: integral[m] = sum w I and vee[m] =
 * sum w/|u| I per molecule (mod_intracule.f90:441-457); either may be
 * NULL. */
int inca_batch_evaluate(inca_batch *b, double *integral, double *vee);
/* Ivec of molecule `mol` after the last inca_batch_evaluate. */
int inca_batch_get_ivec(inca_batch *b, int32_t mol, double *ivec);
/* Times points-per-lane 1, 2, 4, 8 on the batch, keeps the fastest in
 * *best and re-captures the graph with it. */
int inca_batch_autotune(inca_batch *b, int32_t *best);
/* Re-captures the graph with points-per-lane P (1, 2, 4 or 8): a value
 * the autotune found on a like batch, without timing all four again. */
int inca_batch_set_points(inca_batch *b, int32_t points_per_lane);
/* Work items, points per lane, persistent workgroups, graph replays and
 * the device time of the last replay in ms. */
int inca_batch_info(inca_batch *b, int64_t *nitems, int32_t *points_per_lane,
                    int32_t *blocks, int64_t *replays, double *last_ms);

/* Learned autogrid (no reference counterpart).  A per-atom network
 * 16 -> 64 -> 64 -> 64 -> 9 (GELU, FP32; hand-written HIP kernels for
 * inference and training) predicts a factor on the Bragg-Slater radius that
 * get_bs_radius would return (mod_autogrid.f90:36, 61-63) and classes for
 * nrad in {50, 75, 100, 150} and nang in {110, 194, 302, 590}, which replace
 * the one nrad/nang broadcast to every centre (:127-131).  Parameters are a
 * flat FP32 array of INCA_MLP_NPARAM values: W1 (16x64, [in][out]), b1, W2
 * (64x64), b2, W3 (64x64), b3, W4 (64x9), b4. */
#define INCA_MLP_IN 16
#define INCA_MLP_OUT 9
#define INCA_MLP_NPARAM 9993
typedef struct inca_mlp inca_mlp;
/* Deterministic initial weights from `seed` (uniform, variance-scaled). */
int inca_mlp_create(int32_t device, uint64_t seed, inca_mlp **out);
int inca_mlp_destroy(inca_mlp *m);
const char *inca_mlp_last_error(const inca_mlp *m);
int inca_mlp_get_params(inca_mlp *m, float *params);
/* Also resets the optimiser state. */
int inca_mlp_set_params(inca_mlp *m, const float *params);
/* Parameters from a .npy file (little-endian float32, shape
 * (INCA_MLP_NPARAM,), as numpy.save writes them); resets the optimiser.
 * INCA_EIO if the file cannot be read, INCA_EFORMAT if it is not such an
 * array. */
int inca_mlp_load(inca_mlp *m, const char *path);
/* out(n,9) = network(x(n,16)), row-major. */
int inca_mlp_forward(inca_mlp *m, int32_t n, const float *x, float *out);
/* Loss over n atoms with targets (n,3) = [log radius factor, nrad class
 * 0..3, nang class 0..3]: mean of (out0 - t0)^2 + the two softmax
 * cross-entropies.  Computes the gradient on the device (kept for
 * inca_mlp_adam_step) and copies it to `grad` (INCA_MLP_NPARAM) unless
 * NULL. */
int inca_mlp_loss_grad(inca_mlp *m, int32_t n, const float *x, const float *target,
                       double *loss, float *grad);
/* One Adam step (beta 0.9/0.999, eps 1e-8, bias-corrected) with the last
 * gradient. */
int inca_mlp_adam_step(inca_mlp *m, float lr);
/* The network's 16 inputs per atom: feat(natoms,16) row-major.  an,
 * cartes(natoms,3) column-major in bohr; primitives as geninfo::Alpha,
 * TMN(nprim,3) column-major and Ra (1-based atom of each primitive). */
int inca_atom_features(int32_t natoms, const int32_t *an, const double *cartes, int32_t nprim,
                       const double *alpha, const int32_t *tmn, const int32_t *prim_atom,
                       float *feat);
/* Centres as inca_autogrid with the predicted radii (betaone off), plus
 * per-centre nradc/nangc: a centre takes the largest classes among the
 * atoms it stands for (the origin: all atoms; a pair centre: its pair).
 * radius (natoms, may be NULL) receives the radii used.  cap/ERANGE as in
 * inca_autogrid. */
int inca_learned_autogrid(inca_mlp *m, int32_t natoms, const int32_t *an, const double *cartes,
                          int32_t nprim, const double *alpha, const int32_t *tmn,
                          const int32_t *prim_atom, int32_t nohydro, int32_t cap, int32_t *nquad,
                          double *cent, double *Ps, double *sfalpha, int32_t *nradc,
                          int32_t *nangc, double *radius);

/* Mixed precision by magnitude: after screening, each quartet's term
 * A exp(-|rp|^2) Vx Vy Vz gets a bound B on its L1 norm over all space
 * (|A| prod_axis sum_k |C_k| Gamma((k+1)/2) / sqe) and its exact integral M.
 * The lightest quartets go to the FP32 kernel as long as their summed B
 * stays within tail * |sum M| (sum M = the integral of I over all space), so
 * the FP32 roundoff is budgeted against the result, not the magnitudes of
 * terms that cancel; the rest run in FP64.  tail >= 1 evaluates everything
 * in FP32; the default is 0.02.  No reference counterpart (the reference is
 * FP64 only). */
int inca_set_mixed_tail(inca_ctx *ctx, double tail);
/* Split used by the last INCA_MIXED evaluation: quartets in FP64 and in
 * FP32, and the L1-bound threshold B between them. */
int inca_get_mixed_split(inca_ctx *ctx, int64_t *n_fp64, int64_t *n_fp32,
                         double *tau);

/* Per-point cutoff.  The reference evaluates every quartet at every grid
 * point (mod_intracule.f90:345-358).  With log2_rel > 0 each prepared
 * quartet gets a radius in rp beyond which its term is provably below
 * 2^-log2_rel of its peak bound (|A| prod_axis sum_k |C_k| max_y |y|^k
 * e^-y^2, using |term| <= |A| e^-s prod_axis sum_k |C_k| s^(k/2) at
 * s = |rp|^2), and a wavefront skips the quartet when all of its 64 P grid
 * points lie beyond that radius (a bit mask per 64 quartets and wave tile,
 * built before the grid loop).  The skipped contributions sum to at most
 * 2^-log2_rel times the summed peak bounds.  0 turns the cutoff off (the
 * reference's loop); otherwise 64 <= log2_rel <= 1000; default 100.  Takes
 * effect at the next inca_evaluate / inca_prepare.  No reference
 * counterpart. */
int inca_set_point_cutoff(inca_ctx *ctx, double log2_rel);

/* Cube grids by the per-axis contraction.  A cube (mod_build_grid.f90:
 * 412-440) is a tensor product of nx x-, ny y- and nz z-coordinates, and
 * exp(-|rp|^2) is the product of its per-axis factors, so every term is
 * A Fx(x) Fy(y) Fz(z) with F(c) = exp(-rp_c^2) P_c(rp_c).  The factors are
 * formed once per quartet and coordinate (nx + ny + nz of them) and
 * Ivec(i,j,k) = sum_q Fx_q(i) Fy_q(j) Fz_q(k) is contracted: every quartet
 * at every voxel, with no cutoff.  For each chunk of quartets the
 * contraction is the matrix product of the x factors with the row-wise
 * Kronecker product of the y and z factors.  on = 2 (the default) runs it on
 * the matrix cores with factor tiles staged in LDS (v_mfma_f64_16x16x4f64;
 * 0.67 s on the 80^3 benchmark cube), on = 1 on the vector units (0.86 s);
 * both agree with the term-by-term FP64 evaluation to 1e-11 relative.  A
 * mixed evaluation contracts its FP32 range on FP32 factors and the FP32
 * matrix cores (v_mfma_f32_16x16x4f32), sums of 64 records folded into FP64,
 * whatever `on` is (1 or 2).  0 evaluates cubes like other grids: by
 * shell-quartet groups where the preparation made them (FP64 on ray tiles
 * along z with the per-point cutoff, 0.53 s on the benchmark cube), else
 * term by term.  Takes effect at the next evaluation.  No reference counterpart:
 * the reference evaluates each voxel's terms directly
 * (mod_intracule.f90:345-358). */
int inca_set_cube_contraction(inca_ctx *ctx, int32_t on);
/* Evaluations of the last grid loop: reference = quartets x points (the
 * reference's count), executed = the (record, point) pairs the cutoff kept
 * (equal when off), or, when the loop ran by shell-quartet groups, the
 * (group, point) pairs it kept. */
int inca_get_eval_counts(inca_ctx *ctx, int64_t *reference, int64_t *executed);

/* Evaluation by shell-quartet groups.  The records of one shell quartet
 * (and any others with the same sqe and sqe*(P_ik - P_jl), compared bit for
 * bit) share rp and exp(-|rp|^2) in the grid loop (mod_intracule.f90:
 * 345-358), so the pre-pass adds their A-scaled axis polynomials into one
 * three-variable polynomial per group, K_abc = sum_q (A Cx_q)[a] Cy_q[b]
 * Cz_q[c] in record order, and the grid loop evaluates one exponential and
 * one polynomial per group and point (at most 165 coefficients for s/p/d
 * against up to 1296 records).  on = 1 (default) builds the groups in every
 * inca_prepare / inca_evaluate and evaluates radial scans, multicentre grids
 * and explicit points by groups; on = 0 evaluates record by record (the
 * reference's loop order).  Cube grids under the per-axis contraction and
 * inca_evaluate_prepared_noalloc always run record by record.  The results
 * differ from the record loop only by summation order.  A mixed preparation
 * splits the groups by their summed L1 bounds as it splits records (see
 * inca_set_mixed_tail).  Takes effect at the next preparation (turning it
 * off also at the next evaluation).  No reference counterpart. */
int inca_set_grouping(inca_ctx *ctx, int32_t on);
/* Groups of the last preparation: their count, those evaluated in FP64 in
 * mixed mode (the FP64 range), and the coefficients in the pool (0 when the
 * preparation did not group). */
int inca_get_groups(inca_ctx *ctx, int64_t *ngroups, int64_t *ngroups_fp64,
                    int64_t *ncoefficients);
/* Free the prepared records of a grouped preparation, keeping its groups,
 * grid and counters: for callers that keep many prepared molecules (the
 * batch takes groups).  Afterwards only evaluations by groups work; any
 * record-by-record use (grouping off, cube contraction, the no-allocation
 * entry point, inca_get_prepared, inca_get_mixed_split, autotune of the
 * record kernels) fails with INCA_EINVAL until the next preparation.  Fails
 * when the preparation made no groups.  No reference counterpart. */
int inca_release_records(inca_ctx *ctx);
/* keep = 0: the next preparations build the shell-quartet groups chunk by
 * chunk of the pre-pass (inca_set_prepare_chunk) and keep no prepared
 * records, so a molecule whose records would not fit the device is still
 * prepared (the state inca_release_records leaves, reached without ever
 * holding every record).  A shell quartet split between two chunks becomes
 * two groups.  Bases beyond d functions, or grouping off, keep the records
 * whatever the setting.  Default 1.  No reference counterpart. */
int inca_set_keep_records(inca_ctx *ctx, int32_t keep);
/* FP64-equivalent arithmetic of the last grid loop when it ran by groups:
 * the sum over the (group, point) pairs it evaluated of 14 + 2 M (M the
 * group's coefficient count; rp 6, |rp|^2 5, the polynomial 2 M, product and
 * sum 3, exp not counted), 0 otherwise. */
int inca_get_eval_flops(inca_ctx *ctx, double *flops);

/* Sum over the quartets prepared by the last inca_evaluate of each term's
 * integral over all space, prod_axis sum_{k even} C_k Gamma((k+1)/2) A /
 * sqe^3, times 1/2 (the grid weights carry 2 pi for the 4 pi of the volume
 * element, mod_build_grid.f90:132): the analytic value of the integral that
 * inca_integrate takes on the grid.  With exact polynomials it equals the
 * pair sum rule N(N-1)/2 up to the screened terms.  No reference
 * counterpart. */
int inca_get_analytic_integral(inca_ctx *ctx, double *total);

/* Axis polynomials.  mode 0 (default): the reference's Gauss-Hermite +
 * Vandermonde fit in rp = sqe (r + P_ik - P_jl) (mod_intrastuff.f90:127-194),
 * roundoff-dominated where the Vandermonde matrix is ill-conditioned (large
 * exponents times high angular momentum).  mode 1: the same polynomials
 * expanded exactly in rp (each linear factor of the product written in rp;
 * no Vandermonde solve), well conditioned.  The screens
 * and counters are those of the reference in both modes.  No reference
 * counterpart for mode 1. */
int inca_set_polynomials(inca_ctx *ctx, int32_t mode);

/* Evaluation tiles: grid points per lane (1, 2, 4, 6, 8 or 10) of the FP64 and
 * of the FP32 kernel (on radial scans and cube grids, the FP64 kernel's
 * lanes are rays of that many points).  p32 = 0 runs the FP32 range of a
 * mixed evaluation on the FP64 ray tiles instead, where the grid has them
 * (more accurate; faster wherever the ray kernel is); inca_autotune picks it
 * when it times faster.  The same values set the points per lane of the
 * shell-quartet group kernels (10 runs 8 there; a p32 of 0 leaves the
 * group FP32 choice as it was).  inca_autotune times each choice on an evenly strided
 * sample of the quartets prepared by the last inca_evaluate or inca_prepare
 * (about 2^36 evaluations per launch), on the current grid and with the
 * current cutoff, keeps the
 * fastest and returns it in best[0] (FP64) and best[1] (FP32) -- the group
 * kernels' choices (1, 2, 4, 6 or 8, from a sample of the groups) when the
 * current grid is evaluated by groups.  It
 * overwrites the partial sums, so evaluate again before reading Ivec.  No
 * reference counterpart. */
int inca_set_tiles(inca_ctx *ctx, int32_t points_fp64, int32_t points_fp32);
int inca_autotune(inca_ctx *ctx, int32_t *best);

/* Diagnostics: copy prepared quartet records [first, first+n) of the last
 * inca_evaluate (FP64 range first, then the FP32 range in mixed mode).
 * Record layout: [0] sqe, [1..3] sqe*(P_ik - P_jl), [4] class bits
 * Lx | Ly<<8 | Lz<<16 | (axes of nonzero degree)<<24 (as integer bits),
 * [5] B (the L1 bound of the term, see inca_set_mixed_tail), [6] the cutoff
 * radius in rp (+inf when off, see inca_set_point_cutoff), [7]
 * exp(-2 sqe^2 h^2) for the ray step h of the current radial-scan or cube
 * grid (0 until such a grid is evaluated), then A*Cx(0..Lx), Cy(0..Ly),
 * Cz(0..Lz); *stride receives the record length in doubles.
 * No reference counterpart. */
int inca_get_prepared(inca_ctx *ctx, int64_t first, int64_t n, double *out,
                      int32_t *stride);

/* Copy Ivec(G) of the last evaluation. */
int inca_get_ivec(inca_ctx *ctx, double *ivec);

/* FP64 grid reductions of mod_intracule.f90:441-455 on the device:
 * integral = sum rweight*Ivec, vee = sum rweight_vee*Ivec, and the TFVC
 * per-centre split (each of length nquad; may be NULL).  Multicentre grid
 * only. */
int inca_integrate(inca_ctx *ctx, double *integral, double *vee,
                   double *int_per_centre, double *vee_per_centre);

/* Radial scan values rintra(i) = sum_k w_ang*Ivec over radius i
 * (mod_intracule.f90:522-530).  rintra has nradi entries. */
int inca_radial_values(inca_ctx *ctx, double *rintra);

/* radial_smearing (mod_intracule.f90:576-649) on the multicentre grid and
 * the last Ivec: u, P, Iavg each n_bins long. */
int inca_smearing(inca_ctx *ctx, double u_max, int32_t n_bins, double sigma,
                  double *u, double *P, double *Iavg);

/* Kernel time of the last inca_evaluate, in milliseconds, measured with
 * HIP events on the context's stream: t[0] quartet preparation (decode,
 * screens, polycoef, compaction), t[1] grid evaluation, t[2] reduction.
 * Replaces the reference's omp_get_wtime stamps (mod_intracule.f90:88-503). */
int inca_get_timings(inca_ctx *ctx, double *t);

/* Histogram of Lsum = Lx+Ly+Lz over the quartets that passed both screens in
 * the last inca_evaluate: hist[L] for L = 0..nbins-1 (the per-eval work is
 * 16 + 2 Lsum FLOP + 1 exp, so this gives the algorithmic FLOP count).
 * No reference counterpart: a measurement aid. */
int inca_class_histogram(inca_ctx *ctx, int32_t nbins, int64_t *hist);

/* Lebedev-Laikov rule with n points as unit vectors and weights summing to
 * one (lebedev.f LDnnnn).  Returns INCA_ERANGE for an unsupported n. */
int inca_lebedev(int32_t n, double *x, double *y, double *z, double *w);

/* Gauss-Legendre nodes and weights on [-1,1] (mod_quadratures.f90:37-86). */
int inca_gauleg(int32_t n, double *t, double *w);

#ifdef __cplusplus
}
#endif
#endif
