// Internal declarations shared by the libinca_hip.so translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/inca_hip.h"

namespace inca {

// Largest per-axis polynomial degree the Gauss-Hermite tables cover
// (mod_intrastuff.f90:11-15: nodes for L <= 21).
constexpr int kLmaxAxis = 21;
constexpr int kNpMax = kLmaxAxis + 1;
// Prepared-quartet record: [0] sqe, [1..3] sqe*(P_ik - P_jl), [4] class bits
// Lx | Ly<<8 | Lz<<16 | (axes of nonzero degree)<<24, [5] B = |A| prod_axis
// sum_k |C_k| Gamma((k+1)/2) / sqe, a bound on the L1 norm of the term over
// all space, [6] rcut: beyond |rp| = rcut the term stays below 2^-cut_log2
// of its peak bound (+inf when the cutoff is off), [7] rho = exp(-2 sqe^2
// h^2) for the ray step h of the current radial-scan or cube grid (set by
// k_ray_rho, 0 before), then the coefficients A*Cx(0..Lx), Cy(0..Ly),
// Cz(0..Lz) (C(0) = highest power).
constexpr int kQHead = 8;
// Word 4 bit 30 of a pre-pass slot whose record failed screen 2: its words
// 0-3 are set (so it sorts with its shell quartet) and nothing else counts.
// Such rows exist only while a grouped preparation's rows are uncompacted
// (rows_sparse); every record-by-record use compacts them first.
constexpr unsigned kQDropped = 1u << 30;
// Shell-quartet group header (group.hip): words 0-7 as a prepared record's,
// [8] coefficient pool offset, [9] coefficient count, [10] member count,
// [11] FP64 operations per grid point.
constexpr int kGStride = 12;

enum GridKind { GRID_NONE = 0, GRID_MULTI = 1, GRID_RADIAL = 2, GRID_CUBE = 3, GRID_POINTS = 4 };

// Per primitive pair (a,b), the quantities the reference recomputes per
// record: JA8 bound, Gaussian-product exponent and centre
// (mod_intracule.f90:297-311).
struct PairEntry {
  double J;     // JA8(a as i, b as k)
  double aik;   // alpha_a + alpha_b
  double eR2;   // (alpha_a alpha_b / aik) |R_a - R_b|^2
  double P[3];  // Gaussian-product centre
  double pad[2];
};

struct DeviceBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

}  // namespace inca

struct inca_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // primitives
  int nprim = 0;
  int normalize = 1;
  int lmax_axis = 0;  // 4 * max(t) over primitives
  int lmax_prim = 0;  // max(t + m + n) over primitives
  double *d_prim = nullptr;  // SoA: alpha | x | y | z | norm  (5*nprim)
  int32_t *d_tmn = nullptr;  // axis-major t: tmn[ax*nprim + i]
  inca::PairEntry *d_pair = nullptr;

  // grid
  int grid_kind = inca::GRID_NONE;
  int64_t G = 0;
  double *d_gx = nullptr, *d_gy = nullptr, *d_gz = nullptr;  // SoA points
  double *d_w = nullptr, *d_wvee = nullptr;
  int32_t *d_map = nullptr;       // 1-based centre per point
  int nquad = 0;
  std::vector<int64_t> centre_first;  // first point of each centre (+ end)
  std::vector<double> radi;           // radial scan radii
  std::vector<int32_t> smn;           // kept points per radius
  int64_t grid_serial = 0;            // bumped by every new grid
  int32_t cube_n[3] = {0, 0, 0};      // cube grid: points per axis (z fastest)
  double cube_step[3] = {0, 0, 0};

  // ray tiles of a radial scan (index list per points-per-lane choice)
  int64_t ray_serial = -1;    // grid the ray analysis belongs to
  double ray_h = 0.0;         // step along the rays (0: no rays)
  int ray_kind = 0;           // 1: radial-scan shells, 2: cube lines along z
  std::vector<int32_t> ray_runs;  // (first shell, shell count) of each ray run
  std::vector<int32_t> dup_of;    // per point: an earlier point at the same place, or -1
  int32_t *d_dup = nullptr;       // (point, its copy source) pairs not evaluated
  size_t dup_cap = 0;
  int64_t ndup = 0;
  int ray_P = 0;              // points per lane of the ray tiles (0: no list)
  int64_t ray_ntile = 0;      // ray wave tiles (multiple of 4)
  int32_t *d_ridx = nullptr;  // [tile][k][lane] grid index of the ray tiles (-1: none)
  size_t ridx_cap = 0;
  double rho_h = 0.0;         // step whose rho the prepared records hold in word 7

  // DM2 records (24 B each)
  int64_t nrec = 0;
  char *d_rec = nullptr;
  size_t rec_cap = 0;
  // .dm2p reader: two pinned staging buffers (hipHostMalloc), kept
  char *h_pin[2] = {nullptr, nullptr};
  size_t pin_bytes = 0;

  // prepared quartets
  int qstride = 0;       // doubles per prepared record
  int prep_precision = 0; // precision the records were prepared for (mixed: split done)
  int64_t q2 = 0;
  int64_t q_fp64 = 0;         // mixed mode: leading records evaluated in FP64
  double mixed_tail = 0.02;   // FP32 tail budget: sum of its L1 bounds / |integral|
  double mixed_tau = 0.0;     // L1-bound threshold chosen for the split
  int rec_split_pending = 0;  // mixed preparation whose record split is not made yet
  int64_t rows_sparse = 0;    // > 0: d_q holds that many pre-pass slots, screen-2
                              // failures marked (kQDropped), not yet compacted
  double term_integral = 0.0; // sum of the prepared terms' integrals over all space
  int tile_p[2] = {4, 4};     // grid points per lane: FP64 kernel, mixed kernel
  int poly_exact = 0;         // 1: exact axis polynomials instead of the fits
  int64_t prep_chunk = (int64_t)1 << 26;  // records per pre-pass chunk
  double cut_log2 = 100.0;    // per-point cutoff 2^-cut_log2 of a term's peak (0: off)
  int cube_contract = 2;      // cube grids by the per-axis contraction (2: matrix cores, 1: vector units, 0: off)
  double *d_cf = nullptr;     // per-axis factors of a batch of quartets (cube)
  size_t cf_cap = 0;
  double *d_q = nullptr;      // compacted prepared records
  size_t q_cap = 0;
  double *d_qtmp = nullptr;   // uncompacted records (slot = record index)
  size_t qtmp_cap = 0;
  float *d_qf = nullptr;      // FP32 copy for the mixed kernel
  size_t qf_cap = 0;
  // shell-quartet groups of the prepared records (group.hip): headers of
  // kGStride doubles (FP64 range first), one coefficient pool, its FP32 copy
  int grouping = 1;            // evaluate by groups where the grid loop allows
  int keep_records = 1;        // 0: a grouped preparation keeps only its groups
  int64_t ng = 0, ng_fp64 = 0; // groups, of them evaluated in FP64
  double *d_gh = nullptr;
  size_t gh_cap = 0;
  double *d_gpool = nullptr;
  size_t gpool_cap = 0;
  float *d_gpool32 = nullptr;
  size_t gpool32_cap = 0;
  int64_t gpool_n = 0;
  double gflop_fp64 = 0.0, gflop_fp32 = 0.0;  // FP64 operations per point, summed per range
  double flops_executed = 0.0;                // of the last grid loop (group evaluation)
  int grouped_last = 0;                       // the last grid loop ran by groups
  int gtile_p[2] = {4, 4};                    // points per lane of the group kernels
  double grho_h = 0.0;                        // ray step whose rho the group headers hold
  char *d_gscr = nullptr;                     // grouping scratch (sort, scan, maps)
  size_t gscr_cap = 0;
  double *d_ghtmp = nullptr;                  // headers during the mixed split
  size_t ghtmp_cap = 0;
  double *d_wflop = nullptr;                  // per 64-group word: operations kept
  size_t wflop_cap = 0;
  double *d_flops = nullptr;                  // their sum over the last grid loop
  size_t flops_cap = 0;

  // scratch
  uint8_t *d_flag = nullptr;
  size_t flag_cap = 0;
  int64_t *d_idx = nullptr;
  size_t idx_cap = 0;
  int64_t *d_bc = nullptr;
  size_t bc_cap = 0;
  unsigned long long *d_cnt = nullptr;  // device counters
  double *d_tb = nullptr;     // per wave tile: anchor xyz, rho min/max, flag
  size_t tb_cap = 0;
  unsigned long long *d_mask = nullptr;  // cutoff skip bits [64-quartet word][wave tile]
  size_t mask_cap = 0;
  unsigned long long *d_evals = nullptr;  // executed evaluations of the last grid loop
  size_t evals_cap = 0;
  unsigned long long *d_wcount = nullptr; // per 64-quartet word: evaluations kept
  size_t wcount_cap = 0;
  int64_t evals_reference = 0, evals_executed = 0;
  double *d_gcent = nullptr;  // multicentre grid: centre coordinates (3*nquad)
  double *d_partial = nullptr;
  size_t partial_cap = 0;
  double *d_red = nullptr;  // slice sums of the chunk partials
  size_t red_cap = 0;
  // device scratch of one API call (every call that uses it synchronises
  // before returning), and the capacities of the grid arrays
  char *d_tmp = nullptr;
  size_t tmp_cap = 0;
  size_t grid_cap = 0, gcent_cap = 0;
  double *d_ivec = nullptr;
  size_t ivec_cap = 0;
  int have_ivec = 0;

  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double t_ms[3] = {0, 0, 0};
};

namespace inca {

int set_err(inca_ctx *c, int code, const char *fmt, ...);
int hip_fail(inca_ctx *c, hipError_t e, const char *what, int line);

#define INCA_HIP(call)                                              \
  do {                                                              \
    hipError_t e_ = (call);                                         \
    if (e_ != hipSuccess) return ::inca::hip_fail(ctx, e_, #call, __LINE__); \
  } while (0)

// grow-only device buffer
template <class T>
int ensure(inca_ctx *ctx, T *&ptr, size_t &cap, size_t n) {
  if (n <= cap && ptr) return 0;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
  size_t bytes = (n ? n : 1) * sizeof(T);
  hipError_t e = hipMalloc((void **)&ptr, bytes);
  if (e != hipSuccess) {
    ptr = nullptr;
    return set_err(ctx, INCA_ENOMEM, "hipMalloc of %zu bytes failed: %s", bytes,
                   hipGetErrorString(e));
  }
  cap = n;
  return 0;
}

// Offsets of sub-buffers carved from one scratch block, 256-byte aligned:
// first add() every buffer, then ensure the block and take at().
struct Carve {
  size_t size = 0;
  template <class T>
  size_t add(int64_t n) {
    const size_t o = size;
    size += ((size_t)(n > 0 ? n : 1) * sizeof(T) + 255) & ~(size_t)255;
    return o;
  }
  template <class T>
  static T *at(char *base, size_t off) {
    return reinterpret_cast<T *>(base + off);
  }
};

// Deterministic stream compaction of a 0/1 byte flag array:
// idx[0..total) = positions of set flags, in increasing order.
int compact_flags(inca_ctx *ctx, const uint8_t *d_flag, int64_t n,
                  int64_t *d_idx, int64_t *total);

// host-side tables
void gauleg(int n, double *t, double *w);
int lebedev(int n, double *x, double *y, double *z, double *w);

// launchers (grid.hip)
int grid_multicenter(inca_ctx *ctx, int nquad, const double *cent,
                     const int32_t *nradc, const int32_t *nangc,
                     const double *sfalpha, const double *Ps, int nosym,
                     int64_t *maxgrid, int64_t *rgrid, int64_t *rrgrid);
int grid_radial_plot(inca_ctx *ctx, int nblock, const double *tart,
                     const double *step, const int32_t *nang, int nosym,
                     int64_t *nradi, int64_t *ngrid);
int grid_cube(inca_ctx *ctx, const double *center, const double *step,
              const int32_t *np, int64_t *ngrid);
int grid_points(inca_ctx *ctx, int64_t n, const double *r);
int grid_integrate(inca_ctx *ctx, double *integral, double *vee,
                   double *ipc, double *vpc);
int grid_radial_values(inca_ctx *ctx, double *rintra);
int grid_smearing(inca_ctx *ctx, double u_max, int n_bins, double sigma,
                  double *u, double *P, double *Iavg);

// launchers (quartet.hip)
int build_pair_table(inca_ctx *ctx);
int prepare_quartets(inca_ctx *ctx, double thresh, int precision,
                     int64_t *counts);
// the records' FP64/FP32 split of a mixed preparation, made on first need
int split_records(inca_ctx *ctx);
// INCA_EINVAL when the prepared records were released (inca_release_records)
int need_records(inca_ctx *ctx);
// group.hip: groups of all prepared records (build_groups), or built chunk
// by chunk (reset_groups, group_rows per chunk, finish_groups); heads, when
// given, holds the rows' group starts (1 where a row's group words differ
// from the previous row's), else they are found from the rows
int build_groups(inca_ctx *ctx, int precision, const uint8_t *heads = nullptr);
bool groupable(const inca_ctx *ctx);
void reset_groups(inca_ctx *ctx);
int group_rows(inca_ctx *ctx, const double *q, int64_t n, const uint8_t *heads = nullptr);
int finish_groups(inca_ctx *ctx, int precision);
// the pre-pass's stable radix sort of its slots by group key
size_t sort_slots_bytes(int64_t n);
int sort_slots(inca_ctx *ctx, void *temp, size_t bytes, uint32_t *kin, uint32_t *kout,
               uint32_t *vin, uint32_t *vout, int64_t n, unsigned key_bits = 32);

// autogrid.hip
double bragg_slater_bohr(int z);
int autogrid(int natoms, const int32_t *an, const double *cartes, int nohydro, int betaone,
             const double *radius, std::vector<double> &cent, std::vector<double> &Ps,
             std::vector<double> &sfalpha);
void atom_features(int natoms, const int32_t *an, const double *cartes, int nprim,
                   const double *alpha, const int32_t *tmn, const int32_t *prim_atom, float *f);

// launchers (eval.hip)
int evaluate_grid(inca_ctx *ctx, int precision);
int class_histogram(inca_ctx *ctx, int nbins, int64_t *hist);
int autotune(inca_ctx *ctx, int32_t *best);
// one-electron functions on the current grid (onee.hip)
int onee(inca_ctx *ctx, int kind, int index, int nmo, const double *T, const double *w,
         double *values);
// no-allocation evaluation (eval.hip): per-device scratch set-up, called by
// inca_create with the device current, and the grid loop on that scratch
hipError_t noalloc_scratch_init(int device);
int evaluate_noalloc(inca_ctx *ctx, int precision, double *ivec, hipStream_t st);

}  // namespace inca
