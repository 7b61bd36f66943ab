/*
 * inca_oracle.h -- CPU FP64 restatement of the INCA intracule hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libinca_hip.so + the Fortran host) never links or calls it.
 *
 * Every function cites the reference routine it restates
 * (markelgtx/INCA_parallel, src/ paths).  Arrays follow the reference's
 * Fortran layouts so that a C caller can pass Fortran arrays directly:
 *   r(3,G)      -> r[3*g + axis]
 *   tmn(nprim,3)-> tmn[i + nprim*axis]    (column-major, as geninfo::TMN)
 *   cent(3,nq)  -> cent[3*c + axis]
 * Indices inside DM2 records are 1-based, as in the .dm2p files.
 */
#ifndef INCA_ORACLE_H
#define INCA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mod_numbers.f90:14-27 */
int orc_dfact(int a);
/* mod_intracule.f90:145-152 */
double orc_normprim(double alpha, int t, int m, int n);
/* mod_intrastuff.f90:198-341 (nodes from the Hermite roots, weights by the
 * closed-form Gauss-Hermite weight formula) */
int orc_gauherm(int L, double *rh, double *wh);
/* mod_intrastuff.f90:32-123 */
double orc_ja8(int ti, int tk, int mi, int mk, int ni, int nk,
               double alfi, double alfk, double xi, double xk, double yi,
               double yk, double zi, double zk, double rik2, double aik,
               double eik);
/* mod_intrastuff.f90:127-194 (+ LAPACK dgesv, partial-pivot LU).
 * Returns 0 on success, 1 if the reference would zero the coefficients
 * (INFO != 0 or any |C| > 1e3).  C has np entries, C[0] = highest power. */
int orc_polycoef(double ri, double rj, double rk, double rl, double rik,
                 double rjl, double rijkl, int ti, int tj, int tk, int tl,
                 int np, double sqe, double invz, double alfijkl, double *C);

/* Per-quartet prepared data (what the reference holds in registers just
 * before its grid loop, mod_intracule.f90:307-344). */
typedef struct {
  double A, sqe, dx, dy, dz;   /* d = P_ik - P_jl */
  double screen1, screen2;
  int L[3];
  int pass1, pass2, polyfail;
  double C[3][22];   /* fitted coefficients in rp (reference) */
  double E[3][22];   /* exact coefficients in rp (see exact_poly) */
} orc_quartet;

/* mod_intracule.f90:282-344 for one record (1-based i,j,k,l). */
void orc_prepare_quartet(int nprim, const double *alpha, const int32_t *tmn,
                         const double *x, const double *y, const double *z,
                         const double *normprim, double lim, int i, int j,
                         int k, int l, double D, orc_quartet *q);

/* The whole quartet loop, mod_intracule.f90:229-381.
 *   rec_ijkl : 4*nrec int32 (i,j,k,l per record), rec_d : nrec doubles
 *   r        : 3*G doubles, ivec : G doubles (out)
 *   counts   : [read, after screen 1, after screen 2, polycoef failures]
 * Records are read until the first invalid index (mod_intracule.f90:239-240;
 * j,k,l > nprim is also treated as the end of valid data).
 * nthreads <= 0 uses the OpenMP default; dynamic != 0 uses
 * SCHEDULE(DYNAMIC) like the reference (summation order then varies). */
int orc_intracule(int nprim, const double *alpha, const int32_t *tmn,
                  const double *x, const double *y, const double *z,
                  int normalize, double thresh, int64_t nrec,
                  const int32_t *rec_ijkl, const double *rec_d, int64_t G,
                  const double *r, double *ivec, int64_t *counts,
                  int nthreads, int dynamic);

/* Same loop; exact != 0 evaluates the exact axis polynomials. */
int orc_intracule_ex(int nprim, const double *alpha, const int32_t *tmn,
                     const double *x, const double *y, const double *z,
                     int normalize, double thresh, int64_t nrec,
                     const int32_t *rec_ijkl, const double *rec_d, int64_t G,
                     const double *r, double *ivec, int64_t *counts,
                     int nthreads, int dynamic, int exact);

/* mod_quadratures.f90:37-86 on [-1,1] */
void orc_gauleg(int n, double *t, double *w);

/* mod_build_grid.f90:13-277 (+ becke :474-575).
 * ang: per centre, nangc[c] points as (x,y,z,w) quadruples, concatenated.
 * Outputs sized by cap (>= maxgrid).  Returns 0, or -1 if cap too small. */
int orc_gridpoints(int nquad, const double *cent, const int32_t *nradc,
                   const int32_t *nangc, const double *sfalpha,
                   const double *Ps, int nosym, const double *ang,
                   int64_t *maxgrid, int64_t *rgrid, int64_t *rrgrid,
                   double *r_out, double *w_out, double *wvee_out,
                   int32_t *map_out, int64_t cap);

/* mod_build_grid.f90:474-575 on explicit points, owner[i] 0-based centre.
 * Writes the normalised TFVC weight of the owner centre. */
void orc_tfvc(int64_t npts, const double *pts, const int32_t *owner,
              int nquad, const double *cent, const double *Ps, double *wout);

/* mod_build_grid.f90:281-408.  tart: 2*nblock (start,end), ang per block as
 * in orc_gridpoints.  radi/smn sized nradi (query with radi=NULL first),
 * r_out/w_out sized by cap.  Returns number of kept points, or -1. */
int64_t orc_gridpoints2(int nblock, const double *tart, const double *step,
                        const int32_t *nang, int nosym, const double *ang,
                        int64_t *nradi, double *radi, int32_t *smn,
                        double *r_out, double *w_out, int64_t cap);

/* mod_build_grid.f90:412-440 */
int64_t orc_gridpoints3(const double *center, const double *step,
                        const int32_t *np, double *r_out);

/* mod_autogrid.f90:12-185.  cartes: natoms x 3 column-major (as geninfo).
 * Outputs sized natoms*(natoms-1)+1.  Returns nquad. */
int orc_centercalc(int natoms, const int32_t *an, const double *cartes,
                   int nohydro, int betaone, double *cent, double *Ps,
                   double *sfalpha);
/* mod_autogrid.f90:142-185 */
double orc_bs_radius(int iz);

/* mod_intracule.f90:576-649 */
void orc_smearing(int64_t G, const double *coords, const double *weights,
                  const double *density, double u_max, int n_bins,
                  double sigma, double *u, double *P, double *Iavg);

/* One-electron functions at points r(3,npts) (mod_functions.f90: Prim,
 * MoOr, Density, Lapl): kind 1 primitive `index`, 3 orbital `index`,
 * 5 density sum_i w_i phi_i^2, 6 the reference's Laplacian
 * sum_i 2 w_i phi_i sum_p T(i,p) lap(Prim_p).  T(nmo,nprim) column-major;
 * cx/cy/cz the primitive centres. */
void orc_onee(int kind, int index, int nprim, const double *alpha, const int32_t *tmn,
              const double *cx, const double *cy, const double *cz, int nmo, const double *T,
              const double *w, int64_t npts, const double *r, double *out);

/* Arithmetic of powers, Gauss-Hermite roots and weights: 0 (default)
 * repeated multiplication and correctly rounded roots; 1 the reference's
 * literal libm forms (see inca_oracle.c).  Global; set before a run. */
void orc_set_arith(int mode);
int orc_get_arith(void);

/* Inputs of every polynomial fit of the records passing screen 1 (axes of
 * degree >= 1), ORC_FIT_ROW doubles each; and the fits of such rows in the
 * current arithmetic mode.  Test infrastructure for the roundoff audit of
 * mod_intrastuff.f90:127-175. */
#define ORC_FIT_ROW 17
int64_t orc_fit_rows(int nprim, const double *alpha, const int32_t *tmn, const double *x,
                     const double *y, const double *z, int normalize, double thresh,
                     int64_t nrec, const int32_t *rec_ijkl, const double *rec_d, double *rows,
                     int64_t cap);
int64_t orc_polycoef_rows(int64_t n, const double *rows, double *C, int32_t *status);
/* The rows whose fit is zeroed or has max |C| >= cmin in mode 0 or 1
 * (OpenMP over records, rows in record order, at most cap written). */
int64_t orc_fit_census(int nprim, const double *alpha, const int32_t *tmn, const double *x,
                       const double *y, const double *z, int normalize, double thresh,
                       int64_t nrec, const int32_t *rec_ijkl, const double *rec_d, double cmin,
                       double *rows, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
