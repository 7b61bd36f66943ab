// The hot grid loop (mod_intracule.f90:345-358) on the device:
//   Ivec(g) += A exp(-|rp|^2) Vx(rp_x) Vy(rp_y) Vz(rp_z),  rp = sqe (r_g + P_ik - P_jl)
// summed over every prepared quartet.
//
// Mapping.  A block of 256 lanes owns a tile of 256*P grid points (P points
// per lane in registers) and one chunk of the quartet list.  Every lane
// walks the same quartets, so a quartet's data (sqe, sqe*d, class, the
// A-scaled coefficients) is wave-uniform: it comes in through scalar loads
// and sits in SGPRs, there is no divergence whatever the angular-momentum
// class, and the Horner loops have uniform trip counts.  Each (tile, chunk)
// writes its sums to partial[chunk][g]; k_reduce adds the chunks in a fixed
// order, so results are bitwise reproducible.
//
// Arithmetic.  The quartet x point loop is exp-bound VALU work; only the
// per-axis contraction of cube grids is a matrix product (vector FMAs by
// default, FP64 MFMA as an option).  FP64 path: rp by one FMA per axis, exp(-s) by
// reduction by ln2/256 to |r| <= ln2/512, a degree-4 Taylor polynomial of
// e^r (truncation error 3.8e-17), a 256-entry table of 2^(j/256) in LDS and
// an exponent shift (eval_core.hpp).  Mixed path: rp in FP64 (the cancellation-prone step),
// everything after it in FP32 -- hardware exp2, FP32 Horner, FP32 products
// -- and every term added to an FP64 accumulator.
#include <array>
#include <cmath>
#include <map>
#include <mutex>

#include "common.hpp"
#include "eval_core.hpp"

namespace inca {
namespace {

template <int P, bool CUT>
__global__ __launch_bounds__(kT) void k_eval_fp64(const double *__restrict__ q, int qstride,
                                                  int64_t nq, int64_t qchunk,
                                                  const double *__restrict__ gx,
                                                  const double *__restrict__ gy,
                                                  const double *__restrict__ gz, int64_t G,
                                                  const double *__restrict__ tb,
                                                  const unsigned long long *__restrict__ mask,
                                                  int64_t ntile, int64_t tiles, int64_t nch,
                                                  int xcd, double *__restrict__ partial) {
  const BlockTask task = block_task(tiles, xcd);
  if (task.chunk >= nch) return;  // padding block of the XCD order (whole block)
  const int64_t tile = task.tile_block * (kT / 64) +
                       __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t base = tile * (64 * P) + (threadIdx.x & 63);
  const bool empty = !(tb[tile * kTB + 5] > 0.0);
  __shared__ double s_exp2[kExpTable];
  stage_exp_table(s_exp2);
  double x[P], y[P], z[P], acc[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t g = base + (int64_t)k * 64;
    const bool ok = g < G;
    x[k] = ok ? gx[g] : 0.0;
    y[k] = ok ? gy[g] : 0.0;
    z[k] = ok ? gz[g] : 0.0;
    acc[k] = 0.0;
  }
  const int64_t q0 = task.chunk * qchunk;
  const int64_t q1 = empty ? q0 : (q0 + qchunk < nq ? q0 + qchunk : nq);
  for_live(CUT ? mask : nullptr, ntile, tile, q0, q1, [&](int64_t qi) {
    const double *__restrict__ rec = q + qi * qstride;
    accumulate_fp64<P>(rec, load_header(rec), x, y, z, acc, s_exp2);
  });
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t g = base + (int64_t)k * 64;
    if (g < G) partial[task.chunk * G + g] = acc[k];
  }
}

// Shell-quartet groups (group.hip) with the mapping of k_eval_fp64: a
// group's exponential and polynomial per point in place of its records'
// terms.  MIXED: the FP32 range (group_mixed).
template <int P, bool CUT, bool MIXED>
__global__ __launch_bounds__(kT) void k_geval(const double *__restrict__ gh, int64_t ng,
                                              int64_t gchunk, const double *__restrict__ pool,
                                              const float *__restrict__ pool32,
                                              const double *__restrict__ gx,
                                              const double *__restrict__ gy,
                                              const double *__restrict__ gz, int64_t G,
                                              const double *__restrict__ tb,
                                              const unsigned long long *__restrict__ mask,
                                              int64_t ntile, int64_t tiles, int64_t nch, int xcd,
                                              double *__restrict__ partial) {
  const BlockTask task = block_task(tiles, xcd);
  if (task.chunk >= nch) return;  // padding block of the XCD order (whole block)
  const int64_t tile = task.tile_block * (kT / 64) +
                       __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t base = tile * (64 * P) + (threadIdx.x & 63);
  const bool empty = !(tb[tile * kTB + 5] > 0.0);
  __shared__ double s_exp2[kExpTable];
  if (!MIXED) stage_exp_table(s_exp2);
  double x[P], y[P], z[P], acc[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t g = base + (int64_t)k * 64;
    const bool ok = g < G;
    x[k] = ok ? gx[g] : 0.0;
    y[k] = ok ? gy[g] : 0.0;
    z[k] = ok ? gz[g] : 0.0;
    acc[k] = 0.0;
  }
  const int64_t q0 = task.chunk * gchunk;
  const int64_t q1 = empty ? q0 : (q0 + gchunk < ng ? q0 + gchunk : ng);
  for_live(CUT ? mask : nullptr, ntile, tile, q0, q1, [&](int64_t gi) {
    if (MIXED)
      group_mixed<P>(gh + gi * kGStride, pool32, x, y, z, acc);
    else
      group_fp64<P>(gh + gi * kGStride, pool, x, y, z, acc, s_exp2);
  });
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t g = base + (int64_t)k * 64;
    if (g < G) partial[task.chunk * G + g] = acc[k];
  }
}

// Shell-quartet groups on ray tiles (FP64): the lanes of k_eval_rays (one
// direction at P consecutive radii of a radial scan, or P consecutive
// points of a cube line along z), each group by group_ray.  Eight points per lane ask for occupancy 4 (128 VGPRs, 44
// B/lane saved to scratch): with the polynomial's values at t = 4..7 taken
// from its difference table for total degrees up to 5, 28.2 ms on the
// radial-plot step, against 38.3 ms at four points per lane.
template <int P, bool CUT, bool ZLINE>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(P == 8 ? 4 : 1))) void
k_geval_rays(const double *__restrict__ gh, int64_t ng,
                                                   int64_t gchunk, const double *__restrict__ pool,
                                                   const double *__restrict__ gx,
                                                   const double *__restrict__ gy,
                                                   const double *__restrict__ gz, int64_t G,
                                                   const double *__restrict__ tb,
                                                   const unsigned long long *__restrict__ mask,
                                                   int64_t ntile, int64_t tiles, int64_t nch,
                                                   int xcd, const int32_t *__restrict__ idx,
                                                   double h, double *__restrict__ partial) {
  const BlockTask task = block_task(tiles, xcd);
  if (task.chunk >= nch) return;  // padding block of the XCD order (whole block)
  const int64_t tile = task.tile_block * (kT / 64) +
                       __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t base = tile * (64 * P) + (threadIdx.x & 63);
  const bool empty = !(tb[tile * kTB + 5] > 0.0);
  __shared__ double s_exp2[kExpTable];
  stage_exp_table(s_exp2);
  const int32_t g0 = idx[base];
  double o[3] = {0.0, 0.0, 0.0}, hw[3] = {0.0, 0.0, 0.0};
  if (g0 >= 0) {
    o[0] = gx[g0];
    o[1] = gy[g0];
    o[2] = gz[g0];
    if (ZLINE) {  // cube line along z
      hw[2] = h;
    } else {  // radial ray: w = o / |o|
      const double r0 = sqrt(fma(o[0], o[0], fma(o[1], o[1], o[2] * o[2])));
      const double f = r0 > 0.0 ? h / r0 : 0.0;
      hw[0] = f * o[0];
      hw[1] = f * o[1];
      hw[2] = f * o[2];
    }
  }
  const double span = (P - 1) * h;
  double acc[P];
#pragma unroll
  for (int k = 0; k < P; ++k) acc[k] = 0.0;
  const int64_t q0 = task.chunk * gchunk;
  const int64_t q1 = empty ? q0 : (q0 + gchunk < ng ? q0 + gchunk : ng);
  for_live(CUT ? mask : nullptr, ntile, tile, q0, q1, [&](int64_t gi) {
    group_ray<P, !ZLINE>(gh + gi * kGStride, pool, o, hw, span, h * h, acc, s_exp2);
  });
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int32_t g = idx[base + k * 64];
    if (g >= 0) partial[task.chunk * G + g] = acc[k];
  }
}

// Ray tiles: lane `lane` of tile t holds the points idx[t][k][lane], k < P
// (-1: none), equally spaced by h w from its first point o (or o alone):
// one direction w = o / |o| at consecutive radii of a radial scan (kind 1),
// or consecutive points of a cube line along z (kind 2).  The lane keeps o
// and h w, and the quartets are evaluated by ray_any.
//
// Occupancy 5 is asked for up to 8 points per lane, 4 beyond.  At 8 points
// the kernel fits 96 VGPRs with two values saved once per lane outside the
// quartet loop (radial-plot workload 368 ms, against 417 ms at occupancy 4;
// occupancy 6 spills inside the loop and takes 828 ms).
template <int P, bool CUT, bool ZLINE>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(P >= 10 ? 4 : 5))) void k_eval_rays(const double *__restrict__ q, int qstride,
                                                  int64_t nq, int64_t qchunk,
                                                  const double *__restrict__ gx,
                                                  const double *__restrict__ gy,
                                                  const double *__restrict__ gz, int64_t G,
                                                  const double *__restrict__ tb,
                                                  const unsigned long long *__restrict__ mask,
                                                  int64_t ntile, int64_t tiles, int64_t nch,
                                                  int xcd, const int32_t *__restrict__ idx,
                                                  double h, double *__restrict__ partial) {
  const BlockTask task = block_task(tiles, xcd);
  if (task.chunk >= nch) return;  // padding block of the XCD order (whole block)
  const int64_t tile = task.tile_block * (kT / 64) +
                       __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t base = tile * (64 * P) + (threadIdx.x & 63);
  const bool empty = !(tb[tile * kTB + 5] > 0.0);
  __shared__ double s_exp2[kExpTable];
  stage_exp_table(s_exp2);
  const int32_t g0 = idx[base];
  double o[3] = {0.0, 0.0, 0.0}, hw[3] = {0.0, 0.0, 0.0};
  if (g0 >= 0) {
    o[0] = gx[g0];
    o[1] = gy[g0];
    o[2] = gz[g0];
    if (ZLINE) {  // cube line along z
      hw[2] = h;
    } else {          // radial ray: w = o / |o|
      const double r0 = sqrt(fma(o[0], o[0], fma(o[1], o[1], o[2] * o[2])));
      const double f = r0 > 0.0 ? h / r0 : 0.0;
      hw[0] = f * o[0];
      hw[1] = f * o[1];
      hw[2] = f * o[2];
    }
  }
  const double span = (P - 1) * h;
  double acc[P];
#pragma unroll
  for (int k = 0; k < P; ++k) acc[k] = 0.0;
  const int64_t q0 = task.chunk * qchunk;
  const int64_t q1 = empty ? q0 : (q0 + qchunk < nq ? q0 + qchunk : nq);
  for_live(CUT ? mask : nullptr, ntile, tile, q0, q1, [&](int64_t qi) {
    const double *__restrict__ rec = q + qi * qstride;
    ray_any<P, ZLINE>(rec, load_header(rec), o, hw, span, h * h, acc, s_exp2);
  });
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int32_t g = idx[base + k * 64];
    if (g >= 0) partial[task.chunk * G + g] = acc[k];
  }
}

// Cube contraction, per-axis factors: F[r][a] for the n records of a batch
// and the cube coordinates a: x (nx of them, zero up to nxp), then y, then
// z (the cube's point g = (i ny + j) nz + k sits at (gx[i ny nz], gy[j nz],
// gz[k])).  F = exp(-rp^2) P(rp), rp = sqe c + (the record's offset on
// that axis), P the axis polynomial (the x one carries A).
__global__ __launch_bounds__(kT) void k_cube_factors(const double *__restrict__ q, int qstride,
                                                     int64_t n, int nx, int nxp, int ny, int nz,
                                                     const double *__restrict__ gx,
                                                     const double *__restrict__ gy,
                                                     const double *__restrict__ gz,
                                                     double *__restrict__ F) {
  __shared__ double s_exp2[kExpTable];
  stage_exp_table(s_exp2);
  const int64_t nf = (int64_t)nxp + ny + nz;
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= n * nf) return;
  const int64_t r = t / nf;
  const int a = (int)(t - r * nf);
  if (a >= nx && a < nxp) {  // padding of the x factors
    F[t] = 0.0;
    return;
  }
  const double *__restrict__ rec = q + r * qstride;
  const QHeader h = load_header(rec);
  const double *__restrict__ c = rec + kQHead;
  int L;
  double coord, off;
  if (a < nx) {
    L = h.Lx;
    coord = gx[(int64_t)a * ny * nz];
    off = h.ax;
  } else if (a < nxp + ny) {
    c += h.Lx + 1;
    L = h.Ly;
    coord = gy[(int64_t)(a - nxp) * nz];
    off = h.ay;
  } else {
    c += h.Lx + h.Ly + 2;
    L = h.Lz;
    coord = gz[a - nxp - ny];
    off = h.az;
  }
  const double rp = fma(h.sqe, coord, off);
  const double poly = (L == 0) ? c[0] : horner1(c, L, rp);
  F[t] = exp_neg(rp * rp, s_exp2) * poly;
}

// Cube contraction: partial[c][(i ny + j) nz + k] (=, or += when
// `accumulate`) sum over the records of chunk c (qch records each) of
// Fx(i) Fy(j) Fz(k).  Block x: an i block (IB consecutive i) and a 16 x 16
// tile of (j, k), each wave 8 x 8 of it, so that a wave reads 8 y and 8 z
// factors per record (its vector-cache traffic, not its FMAs, bounds the
// loop otherwise); block y: the chunk.  Each lane holds one (j, k) and IB
// sums; per record it forms Fy Fz once and adds Fx(i) times it for its IB
// values of i, which are the same for the whole wave (scalar loads; the x
// factors are zero-padded to nxp = ni IB, so no i needs a test).
template <int IB>
__global__ __launch_bounds__(kT) void k_cube_contract(const double *__restrict__ F, int64_t n,
                                                      int64_t qch, int nx, int ny, int nz,
                                                      int64_t ni, int accumulate,
                                                      double *__restrict__ partial, int64_t G) {
  const int nxp = (int)(ni * IB);
  const int64_t nf = (int64_t)nxp + ny + nz;
  const int64_t njk = (int64_t)ny * nz;
  const int njt = (ny + 15) / 16;
  const int64_t bi = blockIdx.x % ni, bt = blockIdx.x / ni;
  const int bj = (int)(bt % njt), bk = (int)(bt / njt);
  const int64_t chunk = blockIdx.y;
  const int64_t q0 = chunk * qch;
  const int64_t q1 = (q0 + qch < n) ? q0 + qch : n;
  const int i0 = (int)(bi * IB);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int jj = bj * 16 + (w >> 1) * 8 + (lane >> 3);
  const int kk = bk * 16 + (w & 1) * 8 + (lane & 7);
  const bool valid = jj < ny && kk < nz;
  const int j = valid ? jj : 0, k = valid ? kk : 0;
  const int64_t jk = (int64_t)j * nz + k;
  double acc[IB];
#pragma unroll
  for (int ii = 0; ii < IB; ++ii) acc[ii] = 0.0;
  for (int64_t qi = q0; qi < q1; ++qi) {
    const double *__restrict__ Fq = F + qi * nf;
    const double t = Fq[nxp + j] * Fq[nxp + ny + k];
#pragma unroll
    for (int ii = 0; ii < IB; ++ii) acc[ii] = fma(Fq[i0 + ii], t, acc[ii]);
  }
  if (!valid) return;
  double *__restrict__ row = partial + chunk * G;
#pragma unroll
  for (int ii = 0; ii < IB; ++ii) {
    const int i = i0 + ii;
    if (i < nx) {
      const int64_t g = (int64_t)i * njk + jk;
      row[g] = accumulate ? row[g] + acc[ii] : acc[ii];
    }
  }
}

// Per-axis factors in FP32 for the FP32 range of a mixed evaluation: rp is
// formed in FP64 (sqe c + offset cancels), then exp(-rp^2) by the hardware
// exp2 with -log2(e) split in two floats, and the axis polynomial by FP32
// Horner (the x factor carries A), as the mixed point kernel does.
__global__ __launch_bounds__(kT) void k_cube_factors_f32(const double *__restrict__ q, int qstride,
                                                         int64_t n, int nx, int nxp, int ny,
                                                         int nz, const double *__restrict__ gx,
                                                         const double *__restrict__ gy,
                                                         const double *__restrict__ gz,
                                                         float *__restrict__ F) {
  const float kNL2eHi = -1.44269502162933349609375f;
  const float kNL2eLo = -1.925963033500011e-08f;
  const int64_t nf = (int64_t)nxp + ny + nz;
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= n * nf) return;
  const int64_t r = t / nf;
  const int a = (int)(t - r * nf);
  if (a >= nx && a < nxp) {  // padding of the x factors
    F[t] = 0.0f;
    return;
  }
  const double *__restrict__ rec = q + r * qstride;
  const QHeader h = load_header(rec);
  const double *__restrict__ c = rec + kQHead;
  int L;
  double coord, off;
  if (a < nx) {
    L = h.Lx;
    coord = gx[(int64_t)a * ny * nz];
    off = h.ax;
  } else if (a < nxp + ny) {
    c += h.Lx + 1;
    L = h.Ly;
    coord = gy[(int64_t)(a - nxp) * nz];
    off = h.ay;
  } else {
    c += h.Lx + h.Ly + 2;
    L = h.Lz;
    coord = gz[a - nxp - ny];
    off = h.az;
  }
  const float rp = (float)fma(h.sqe, coord, off);
  float poly = (float)c[0];
  for (int j = 1; j <= L; ++j) poly = fmaf(poly, rp, (float)c[j]);
  const float s = rp * rp;
  F[t] = __builtin_amdgcn_exp2f(fmaf(s, kNL2eHi, s * kNL2eLo)) * poly;
}

// The contraction on the matrix cores, factor tiles staged in LDS.  Over the
// records of a chunk it is the matrix product
//   C[i][(j,k)] = sum_q Fx[q][i] (Fy[q][j] Fz[q][k]).
// A block takes 16 x values (i tile) and a 16 x 16 (j, k) tile; wave w holds
// the four 16 x 16 output tiles of j = j0 + 4w + c (c = 0..3), k = k0..k0+15.
// Per step of four records lane l supplies A = Fx[q + l/16][i0 + l%16] and,
// per tile c, B = Fy[q + l/16][j0 + 4w + c] Fz[q + l/16][k0 + l%16]; the 16
// records of a stage are loaded into LDS by the whole block (zeros past the
// chunk and outside the cube, and the x factors are zero-padded to whole i
// tiles, so no operand needs a test).
//   T = double: v_mfma_f64_16x16x4f64, FP64 accumulators (mode 2);
//   T = float:  v_mfma_f32_16x16x4f32 on FP32 factors (the FP32 range of a
//               mixed evaluation), FP32 sums of 64 records at a time folded
//               into FP64 accumulators.
// Output register r of lane l holds row Mma<T>::row(l, r), column l%16 (the
// instructions' lane layouts on gfx950).
template <typename T>
struct Mma;
template <>
struct Mma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
  static constexpr int kFold = 0;  // accumulate in FP64 throughout
};
template <>
struct Mma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
  static constexpr int kFold = 2;  // stages of 32 records per FP32 run
};

constexpr int kStageQ = 32;  // records per LDS stage

template <typename T>
__global__ __launch_bounds__(kT) void k_cube_contract_mma(const T *__restrict__ F, int64_t n,
                                                          int64_t qch, int nx, int nxp, int ny,
                                                          int nz, int accumulate,
                                                          double *__restrict__ partial,
                                                          int64_t G) {
  typedef typename Mma<T>::v4 v4;
  __shared__ T sx[kStageQ][16], sy[kStageQ][16], sz[kStageQ][16];
  const int64_t nf = (int64_t)nxp + ny + nz;
  const int nit = nxp / 16, njt = (ny + 15) / 16;
  const int it = (int)(blockIdx.x % (unsigned)nit);
  const int jkt = (int)(blockIdx.x / (unsigned)nit);
  const int i0 = it * 16, j0 = (jkt % njt) * 16, k0 = (jkt / njt) * 16;
  const int64_t chunk = blockIdx.y;
  const int64_t q0 = chunk * qch;
  const int64_t q1 = (q0 + qch < n) ? q0 + qch : n;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int kr = lane >> 4, col = lane & 15;
  v4 acc[4];
  double acc64[4][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    acc[c] = v4{0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < 4; ++r) acc64[c][r] = 0.0;
  }
  // loader: thread t stages column t%16 of records t/16 and t/16 + 16 of
  // each 16-wide tile; the next stage is fetched into registers while the
  // current one is consumed
  const int lq = threadIdx.x >> 4, lc = threadIdx.x & 15;
  const bool jok = j0 + lc < ny, kok = k0 + lc < nz;
  T pre[2][3];
  auto fetch = [&](int64_t qb) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t qq = qb + lq + 16 * h;
      const bool ok = qq < q1;
      const T *__restrict__ row = F + (ok ? qq : q0) * nf;
      pre[h][0] = ok ? row[i0 + lc] : T(0);
      pre[h][1] = (ok && jok) ? row[nxp + j0 + lc] : T(0);
      pre[h][2] = (ok && kok) ? row[nxp + ny + k0 + lc] : T(0);
    }
  };
  if (q0 < q1) fetch(q0);
  int stage = 0;
  for (int64_t qb = q0; qb < q1; qb += kStageQ, ++stage) {
    __syncthreads();  // the previous stage has been consumed
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      sx[lq + 16 * h][lc] = pre[h][0];
      sy[lq + 16 * h][lc] = pre[h][1];
      sz[lq + 16 * h][lc] = pre[h][2];
    }
    __syncthreads();
    if (qb + kStageQ < q1) fetch(qb + kStageQ);
#pragma unroll
    for (int s = 0; s < kStageQ / 4; ++s) {
      const int rq = 4 * s + kr;
      const T a = sx[rq][col];
      const T fz = sz[rq][col];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] = Mma<T>::step(a, sy[rq][4 * w + c] * fz, acc[c]);
    }
    if (Mma<T>::kFold && (stage + 1) % (Mma<T>::kFold > 0 ? Mma<T>::kFold : 1) == 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[c][r] += (double)acc[c][r];
        acc[c] = v4{0, 0, 0, 0};
      }
    }
  }
  double *__restrict__ out = partial + chunk * G;
  const int64_t njk = (int64_t)ny * nz;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int j = j0 + 4 * w + c, k = k0 + col;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + Mma<T>::row(lane, r);
      const double v = Mma<T>::kFold ? acc64[c][r] + (double)acc[c][r] : (double)acc[c][r];
      if (i < nx && j < ny && k < nz) {
        const int64_t g = (int64_t)i * njk + (int64_t)j * nz + k;
        out[g] = accumulate ? out[g] + v : v;
      }
    }
  }
}

// rho = exp(-2 sqe^2 h^2) of every prepared record into word 7 (ray tiles).
__global__ __launch_bounds__(kT) void k_ray_rho(double *__restrict__ q, int qstride, int64_t nq,
                                                double h2) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nq) return;
  const double sqe = q[i * qstride];
  q[i * qstride + 7] = exp(-2.0 * sqe * sqe * h2);
}

template <int P, bool CUT>
__global__ __launch_bounds__(kT) void k_eval_mixed(const double *__restrict__ q, int qstride,
                                                   int64_t nq, int64_t qchunk,
                                                   const double *__restrict__ gx,
                                                   const double *__restrict__ gy,
                                                   const double *__restrict__ gz, int64_t G,
                                                   const double *__restrict__ tb,
                                                   const unsigned long long *__restrict__ mask,
                                                   int64_t ntile, int64_t tiles, int64_t nch,
                                                   int xcd, double *__restrict__ partial) {
  const BlockTask task = block_task(tiles, xcd);
  if (task.chunk >= nch) return;  // padding block of the XCD order (whole block)
  const int64_t tile = task.tile_block * (kT / 64) +
                       __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t base = tile * (64 * P) + (threadIdx.x & 63);
  const bool empty = !(tb[tile * kTB + 5] > 0.0);
  double x[P], y[P], z[P], acc[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t g = base + (int64_t)k * 64;
    const bool ok = g < G;
    x[k] = ok ? gx[g] : 0.0;
    y[k] = ok ? gy[g] : 0.0;
    z[k] = ok ? gz[g] : 0.0;
    acc[k] = 0.0;
  }
  const int64_t q0 = task.chunk * qchunk;
  const int64_t q1 = empty ? q0 : (q0 + qchunk < nq ? q0 + qchunk : nq);
  for_live(CUT ? mask : nullptr, ntile, tile, q0, q1, [&](int64_t qi) {
    const double *__restrict__ rec = q + qi * qstride;
    accumulate_mixed_any<P>(rec, load_header(rec), x, y, z, acc);
  });
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t g = base + (int64_t)k * 64;
    if (g < G) partial[task.chunk * G + g] = acc[k];
  }
}

// First stage of the chunk reduction when there are many chunks: slice s
// (blockIdx.y) adds chunks [nchunks s / nslice, nchunks (s + 1) / nslice)
// of every point in order; k_reduce then adds the slices in order.
__global__ __launch_bounds__(kT) void k_reduce_slices(const double *__restrict__ partial,
                                                      int nchunks, int nslice, int64_t G,
                                                      double *__restrict__ red) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= G) return;
  const int s = blockIdx.y;
  const int c0 = (int)((int64_t)nchunks * s / nslice);
  const int c1 = (int)((int64_t)nchunks * (s + 1) / nslice);
  double acc = 0.0;
  for (int c = c0; c < c1; ++c) acc += partial[(int64_t)c * G + g];
  red[(int64_t)s * G + g] = acc;
}

__global__ __launch_bounds__(kT) void k_reduce(const double *__restrict__ partial, int nchunks,
                                               int64_t G, double *__restrict__ ivec) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= G) return;
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += partial[(int64_t)c * G + g];
  ivec[g] = s;
}

// Ivec of the points evaluated through another point at the same place.
__global__ __launch_bounds__(kT) void k_copy_dups(const int32_t *__restrict__ dup, int64_t n,
                                                  double *__restrict__ ivec) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) ivec[dup[2 * i]] = ivec[dup[2 * i + 1]];
}

__global__ __launch_bounds__(kT) void k_class_hist(const double *__restrict__ q, int qstride,
                                                   int64_t nq, int nbins,
                                                   unsigned long long *__restrict__ hist) {
  __shared__ unsigned int h[64];
  if (threadIdx.x < 64) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < nq; i += (int64_t)gridDim.x * kT) {
    const unsigned cls = (unsigned)__double_as_longlong(q[i * qstride + 4]);
    int L = (int)((cls & 0xff) + ((cls >> 8) & 0xff) + ((cls >> 16) & 0xff));
    if (L >= nbins) L = nbins - 1;
    atomicAdd(&h[L], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 64 && (int)threadIdx.x < nbins && h[threadIdx.x])
    atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// dst record r = src record floor(r nq / ns), whole records of `stride` doubles
__global__ __launch_bounds__(kT) void k_gather_strided(const double *__restrict__ src, int64_t nq,
                                                       int64_t ns, int stride,
                                                       double *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= ns * stride) return;
  const int64_t r = t / stride;
  const int w = (int)(t - r * stride);
  dst[t] = src[((r * nq) / ns) * stride + w];
}

__global__ void k_zero_u64(unsigned long long *p) {
  if (threadIdx.x == 0) p[0] = 0ull;
}

__global__ void k_zero(double *p, int64_t n) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g < n) p[g] = 0.0;
}

}  // namespace

int class_histogram(inca_ctx *ctx, int nbins, int64_t *hist) {
  if (int rc = need_records(ctx)) return rc;
  if (nbins > 64) nbins = 64;
  unsigned long long *d_h = nullptr;
  INCA_HIP(hipMalloc((void **)&d_h, nbins * sizeof(unsigned long long)));
  std::vector<unsigned long long> h(nbins, 0);
  hipError_t e = hipMemcpy(d_h, h.data(), nbins * sizeof(unsigned long long), hipMemcpyHostToDevice);
  if (e == hipSuccess && ctx->q2 > 0) {
    hipLaunchKernelGGL(k_class_hist, dim3(1024), dim3(kT), 0,
                       ctx->stream, ctx->d_q, ctx->qstride, ctx->q2, nbins, d_h);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  if (e == hipSuccess)
    e = hipMemcpy(h.data(), d_h, nbins * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  (void)hipFree(d_h);
  if (e != hipSuccess) return hip_fail(ctx, e, "class histogram", __LINE__);
  for (int i = 0; i < nbins; ++i) hist[i] = (int64_t)h[i];
  return 0;
}

namespace {

// ---- ray tiles of a radial scan or a cube grid ----
//
// A cube grid is lines of points along z a step apart (mod_build_grid.f90:
// 412-440, z fastest): each line is cut into lanes of P consecutive points.
//
// A radial scan is shells of points s_j w_d (mod_build_grid.f90:306-326):
// within one block of the scan, the shells after s = 0 keep the same
// directions and are s_0 + j step apart.  Runs of consecutive shells with
// equal direction lists and one common step become rays: P consecutive
// shells of one direction are the P points of one lane, 64 such lanes a
// wave tile.  Every other point (the s = 0 shell, shells left over at the
// end of a run, runs of another step) goes to ordinary tiles of 64 P points
// in grid order.  Both kinds are described by one index list, so the skip
// mask and the kernel treat them alike.

// Runs (first shell, shell count) and their common step, once per grid.
int ray_analyse(inca_ctx *ctx) {
  if (ctx->ray_serial == ctx->grid_serial) return 0;
  ctx->ray_serial = ctx->grid_serial;
  ctx->ray_h = 0.0;
  ctx->ray_P = 0;
  ctx->ray_runs.clear();
  ctx->dup_of.clear();
  ctx->ndup = 0;
  ctx->ray_kind = 0;
  const int64_t G = ctx->G;
  if (ctx->grid_kind == GRID_CUBE) {
    // lines along z (the fastest index), points step_z apart
    if (G > 0 && G <= INT32_MAX && ctx->cube_n[2] >= 2 && ctx->cube_step[2] > 0.0) {
      ctx->ray_h = ctx->cube_step[2];
      ctx->ray_kind = 2;
    }
    return 0;
  }
  const int nsh = (int)ctx->radi.size();
  if (ctx->grid_kind != GRID_RADIAL || G == 0 || G > INT32_MAX || nsh < 2) return 0;
  std::vector<int64_t> off(nsh + 1, 0);
  for (int j = 0; j < nsh; ++j) off[j + 1] = off[j] + ctx->smn[j];
  if (off[nsh] != G) return 0;
  std::vector<double> x(G), y(G), z(G);
  INCA_HIP(hipMemcpyAsync(x.data(), ctx->d_gx, G * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
  INCA_HIP(hipMemcpyAsync(y.data(), ctx->d_gy, G * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
  INCA_HIP(hipMemcpyAsync(z.data(), ctx->d_gz, G * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  // points at the same place as an earlier one (the s = 0 shell holds one
  // point per direction, all at the origin) are evaluated once
  ctx->dup_of.assign(G, -1);
  {
    std::map<std::array<double, 3>, int32_t> seen;
    for (int64_t g = 0; g < G; ++g) {
      auto it = seen.emplace(std::array<double, 3>{x[g], y[g], z[g]}, (int32_t)g);
      if (!it.second) ctx->dup_of[g] = it.first->second;
    }
  }
  const std::vector<double> &r = ctx->radi;
  // shells j and j+1 on the same directions: s_(j+1) x_j = s_j x_(j+1)
  auto same_dirs = [&](int j) {
    if (!(r[j] > 0.0) || !(r[j + 1] > r[j]) || ctx->smn[j] != ctx->smn[j + 1] ||
        ctx->smn[j] == 0)
      return false;
    const double tol = 1e-13 * r[j] * r[j + 1];
    for (int64_t d = 0; d < ctx->smn[j]; ++d) {
      const int64_t a = off[j] + d, b = off[j + 1] + d;
      if (std::fabs(r[j + 1] * x[a] - r[j] * x[b]) > tol ||
          std::fabs(r[j + 1] * y[a] - r[j] * y[b]) > tol ||
          std::fabs(r[j + 1] * z[a] - r[j] * z[b]) > tol)
        return false;
    }
    return true;
  };
  struct Run {
    int j0, len;
    double h;
  };
  std::vector<Run> runs;
  for (int j = 0; j + 1 < nsh;) {
    if (!same_dirs(j)) {
      ++j;
      continue;
    }
    const double h = r[j + 1] - r[j];
    int e = j + 1;  // last shell of the run
    while (e + 1 < nsh && same_dirs(e) && std::fabs((r[e + 1] - r[e]) - h) <= 1e-12 * h) ++e;
    runs.push_back(Run{j, e - j + 1, h});
    j = e + 1;
  }
  if (runs.empty()) return 0;
  // the step covering the most points; runs of other steps stay ordinary
  double best_pts = -1.0, h_main = 0.0;
  for (const Run &a : runs) {
    double pts = 0.0;
    for (const Run &b : runs)
      if (std::fabs(b.h - a.h) <= 1e-12 * a.h) pts += (double)b.len * ctx->smn[b.j0];
    if (pts > best_pts) {
      best_pts = pts;
      h_main = a.h;
    }
  }
  for (const Run &a : runs)
    if (std::fabs(a.h - h_main) <= 1e-12 * h_main) {
      ctx->ray_runs.push_back(a.j0);
      ctx->ray_runs.push_back(a.len);
    }
  ctx->ray_h = h_main;
  ctx->ray_kind = 1;
  return 0;
}

// Index list for P points per ray lane.  A run of shells gives each of its
// directions lanes of P consecutive shells (the last lane of a run may hold
// fewer); every other point is a lane of its own (its first slot), and a
// point at the same place as an earlier one is not evaluated but copied
// after the reduction.  Unused slots hold -1.
int ray_plan(inca_ctx *ctx, int P) {
  int rc;
  if ((rc = ray_analyse(ctx))) return rc;
  if (ctx->ray_h <= 0.0 || P < 2) return 0;
  if (ctx->ray_P == P) return 0;
  const int64_t G = ctx->G;
  std::vector<int32_t> lanes;  // P slots per lane
  std::vector<int32_t> dups;   // (copy, source) pairs
  if (ctx->ray_kind == 2) {
    // cube: each z line cut into lanes of P consecutive points
    const int64_t nz = ctx->cube_n[2], nline = G / nz;
    for (int64_t l = 0; l < nline; ++l)
      for (int64_t gs = 0; gs < nz; gs += P)
        for (int k = 0; k < P; ++k) lanes.push_back(gs + k < nz ? (int32_t)(l * nz + gs + k) : -1);
  }
  const int nsh = ctx->ray_kind == 1 ? (int)ctx->radi.size() : 0;
  std::vector<int64_t> off(nsh + 1, 0);
  for (int j = 0; j < nsh; ++j) off[j + 1] = off[j] + ctx->smn[j];
  std::vector<uint8_t> used(ctx->ray_kind == 1 ? G : 0, 0);
  for (size_t i = 0; i < ctx->ray_runs.size(); i += 2) {
    const int j0 = ctx->ray_runs[i], len = ctx->ray_runs[i + 1];
    const int64_t nd = ctx->smn[j0];
    for (int gs = 0; gs < len; gs += P)
      for (int64_t d = 0; d < nd; ++d)
        for (int k = 0; k < P; ++k) {
          if (gs + k < len) {
            const int64_t g = off[j0 + gs + k] + d;
            lanes.push_back((int32_t)g);
            used[g] = 1;
          } else {
            lanes.push_back(-1);
          }
        }
  }
  for (int64_t g = 0; ctx->ray_kind == 1 && g < G; ++g) {
    if (used[g]) continue;
    if (ctx->dup_of[g] >= 0) {
      dups.push_back((int32_t)g);
      dups.push_back(ctx->dup_of[g]);
      continue;
    }
    lanes.push_back((int32_t)g);
    for (int k = 1; k < P; ++k) lanes.push_back(-1);
  }
  const int64_t nlanes = (int64_t)lanes.size() / P;
  if (nlanes == 0) return 0;
  const int64_t nray = (nlanes + 255) / 256 * 4;  // whole blocks of 4 waves
  std::vector<int32_t> idx((size_t)nray * 64 * P, -1);
  for (int64_t l = 0; l < nlanes; ++l) {
    const int64_t t = l / 64, lane = l % 64;
    for (int k = 0; k < P; ++k) idx[(t * P + k) * 64 + lane] = lanes[l * P + k];
  }
  if ((rc = ensure(ctx, ctx->d_ridx, ctx->ridx_cap, idx.size()))) return rc;
  INCA_HIP(hipMemcpy(ctx->d_ridx, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  ctx->ndup = (int64_t)dups.size() / 2;
  if (ctx->ndup > 0) {
    if ((rc = ensure(ctx, ctx->d_dup, ctx->dup_cap, dups.size()))) return rc;
    INCA_HIP(hipMemcpy(ctx->d_dup, dups.data(), dups.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice));
  }
  ctx->ray_P = P;
  ctx->ray_ntile = nray;
  return 0;
}

// rho of the grid's ray step in word 7 of every prepared record.
int ensure_rho(inca_ctx *ctx) {
  int rc;
  if ((rc = ray_analyse(ctx))) return rc;
  if (ctx->ray_h <= 0.0 || ctx->q2 == 0 || ctx->rho_h == ctx->ray_h) return 0;
  hipLaunchKernelGGL(k_ray_rho, dim3((unsigned)((ctx->q2 + kT - 1) / kT)), dim3(kT), 0,
                     ctx->stream, ctx->d_q, ctx->qstride, ctx->q2, ctx->ray_h * ctx->ray_h);
  INCA_HIP(hipGetLastError());
  ctx->rho_h = ctx->ray_h;
  return 0;
}

// Blocks of 4 wave tiles for one kernel with P points per lane, and whether
// the tiles come from the ray index lists (FP64 kernel on a radial scan).
int tile_layout(inca_ctx *ctx, bool mixed, int P, int64_t *tiles, bool *rays) {
  *rays = false;
  *tiles = (ctx->G + (int64_t)kT * P - 1) / ((int64_t)kT * P);
  if (mixed || P < 2 || (ctx->grid_kind != GRID_RADIAL && ctx->grid_kind != GRID_CUBE)) return 0;
  int rc;
  if ((rc = ray_plan(ctx, P))) return rc;
  if (ctx->ray_P != P) return 0;
  *rays = true;
  *tiles = ctx->ray_ntile / 4;
  return 0;
}

// Launch one evaluation over quartets [q, q+n) in `nch` chunks of `qch`
// quartets with P points per lane; partial sums go to partial[chunk][g].
// The FP64 evaluation of a radial scan runs on its ray tiles
// (k_eval_rays).
int launch_eval(inca_ctx *ctx, bool mixed, int P, const double *q, int64_t n, int64_t qch,
                int64_t nch, double *partial) {
  const int64_t G = ctx->G;
  int64_t tiles;
  bool rays;
  int rc;
  if ((rc = tile_layout(ctx, mixed, P, &tiles, &rays))) return rc;
  const int32_t *ridx = rays ? ctx->d_ridx : nullptr;
  const int xcd = nch >= kXcds;
  const int64_t nblocks = tiles * (xcd ? (nch + kXcds - 1) / kXcds * kXcds : nch);
  dim3 grid((unsigned)nblocks);
  hipStream_t st = ctx->stream;
  // bounds of the wave tiles of this P (a few microseconds; redone per launch)
  const bool cut = ctx->cut_log2 > 0.0;
  const int64_t ntile = tiles * (kT / 64);
  if ((rc = ensure(ctx, ctx->d_tb, ctx->tb_cap, (size_t)ntile * kTB))) return rc;
  if (cut) {
    const int mode = ctx->grid_kind == GRID_RADIAL ? 0 : (ctx->grid_kind == GRID_MULTI ? 1 : 2);
    hipLaunchKernelGGL(k_tile_bounds, dim3((unsigned)tiles), dim3(kT), 0, st, ctx->d_gx,
                       ctx->d_gy, ctx->d_gz, G, P, mode, ctx->d_map, ctx->d_gcent, ntile,
                       ctx->d_tb, ridx);
    const int64_t nwords = (n + 63) / 64;
    if ((rc = ensure(ctx, ctx->d_mask, ctx->mask_cap, (size_t)(nwords > 0 ? nwords : 1) * ntile)))
      return rc;
    if ((rc = ensure(ctx, ctx->d_wcount, ctx->wcount_cap, (size_t)(nwords > 0 ? nwords : 1))))
      return rc;
    // the executed-evaluation counter is added to here; the grid loop clears
    // it first, and autotune launches (which may come before any grid loop)
    // only need it to exist
    if ((rc = ensure(ctx, ctx->d_evals, ctx->evals_cap, 1))) return rc;
    if (nwords > 0) {
      hipLaunchKernelGGL(k_skip_mask, dim3((unsigned)((nwords + 3) / 4)), dim3(kT), 0, st, q,
                         ctx->qstride, n, ctx->d_tb, ntile, nwords, ctx->d_mask, ctx->d_wcount);
      hipLaunchKernelGGL(k_sum_u64, dim3(1), dim3(kT), 0, st, ctx->d_wcount, nwords, ctx->d_evals);
    }
  } else {
    // only the point-count word is read: mark every tile non-empty
    hipLaunchKernelGGL(k_fill_tiles, dim3((unsigned)((ntile + kT - 1) / kT)), dim3(kT), 0, st,
                       ntile, ctx->d_tb);
  }
  INCA_HIP(hipGetLastError());
  const double h = ctx->ray_h;
#define INCA_EVAL2(PP, CC)                                                                     \
  if (mixed) {                                                                                 \
    hipLaunchKernelGGL((k_eval_mixed<PP, CC>), grid, dim3(kT), 0, st, q, ctx->qstride, n, qch, \
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, G, ctx->d_tb, ctx->d_mask, ntile,     \
                       tiles, nch, xcd, partial);                                              \
  } else if (rays) {                                                                           \
    if constexpr (PP >= 2) {                                                                   \
      if (ctx->ray_kind == 2)                                                                  \
        hipLaunchKernelGGL((k_eval_rays<PP, CC, true>), grid, dim3(kT), 0, st, q, ctx->qstride,\
                           n, qch, ctx->d_gx, ctx->d_gy, ctx->d_gz, G, ctx->d_tb, ctx->d_mask, \
                           ntile, tiles, nch, xcd, ridx, h, partial);                          \
      else                                                                                     \
        hipLaunchKernelGGL((k_eval_rays<PP, CC, false>), grid, dim3(kT), 0, st, q,             \
                           ctx->qstride, n, qch, ctx->d_gx, ctx->d_gy, ctx->d_gz, G, ctx->d_tb,\
                           ctx->d_mask, ntile, tiles, nch, xcd, ridx, h, partial);             \
    }                                                                                          \
  } else {                                                                                     \
    hipLaunchKernelGGL((k_eval_fp64<PP, CC>), grid, dim3(kT), 0, st, q, ctx->qstride, n, qch,  \
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, G, ctx->d_tb, ctx->d_mask, ntile,     \
                       tiles, nch, xcd, partial);                                              \
  }
#define INCA_EVAL(PP)      \
  if (cut) {               \
    INCA_EVAL2(PP, true)   \
  } else {                 \
    INCA_EVAL2(PP, false)  \
  }
  switch (P) {
    case 1: INCA_EVAL(1) break;
    case 2: INCA_EVAL(2) break;
    case 6: INCA_EVAL(6) break;
    case 8: INCA_EVAL(8) break;
    case 10: INCA_EVAL(10) break;
    default: INCA_EVAL(4) break;
  }
#undef INCA_EVAL
#undef INCA_EVAL2
  INCA_HIP(hipGetLastError());
  return 0;
}

// Wave tiles of the group kernels: the ray tiles of a radial scan for the
// FP64 kernel at P >= 2 points per lane (the record kernel's plan), plain
// tiles of 64 P consecutive points otherwise.
int group_tiles(inca_ctx *ctx, bool mixed, int P, int64_t *tiles, bool *rays) {
  *rays = false;
  *tiles = (ctx->G + (int64_t)kT * P - 1) / ((int64_t)kT * P);
  if (mixed || P < 2 || (ctx->grid_kind != GRID_RADIAL && ctx->grid_kind != GRID_CUBE)) return 0;
  int rc;
  if ((rc = ray_plan(ctx, P))) return rc;
  if (ctx->ray_P != P || (ctx->ray_kind != 1 && ctx->ray_kind != 2)) return 0;
  *rays = true;
  *tiles = ctx->ray_ntile / 4;
  return 0;
}

// rho = exp(-2 sqe^2 h^2) of every group header (word 7) for the scan's
// ray step h, once per step.
int ensure_grho(inca_ctx *ctx) {
  int rc;
  if ((rc = ray_analyse(ctx))) return rc;
  if (ctx->ray_h <= 0.0 || ctx->ng == 0 || ctx->grho_h == ctx->ray_h) return 0;
  hipLaunchKernelGGL(k_ray_rho, dim3((unsigned)((ctx->ng + kT - 1) / kT)), dim3(kT), 0,
                     ctx->stream, ctx->d_gh, kGStride, ctx->ng, ctx->ray_h * ctx->ray_h);
  INCA_HIP(hipGetLastError());
  ctx->grho_h = ctx->ray_h;
  return 0;
}

// Launch the group kernel over groups [g0, g0 + n) (the headers at
// ctx->d_gh) in `nch` chunks of `gch` groups, P points per lane; partial
// sums to partial[chunk][g].  With the cutoff, the skip mask also counts the
// FP64 operations the launch executes (header word 11) into ctx->d_flops.
int launch_geval_part(inca_ctx *ctx, bool mixed, int P, const double *gh, int64_t n, int64_t gch,
                      int64_t nch, double *partial) {
  const int64_t G = ctx->G;
  int64_t tiles;
  bool rays;
  int rc;
  if ((rc = group_tiles(ctx, mixed, P, &tiles, &rays))) return rc;
  const int32_t *ridx = rays ? ctx->d_ridx : nullptr;
  const int xcd = nch >= kXcds;
  const int64_t nblocks = tiles * (xcd ? (nch + kXcds - 1) / kXcds * kXcds : nch);
  dim3 grid((unsigned)nblocks);
  hipStream_t st = ctx->stream;
  const bool cut = ctx->cut_log2 > 0.0;
  const int64_t ntile = tiles * (kT / 64);
  if ((rc = ensure(ctx, ctx->d_tb, ctx->tb_cap, (size_t)ntile * kTB))) return rc;
  if ((rc = ensure(ctx, ctx->d_evals, ctx->evals_cap, 1))) return rc;
  if ((rc = ensure(ctx, ctx->d_flops, ctx->flops_cap, 1))) return rc;
  if (cut) {
    const int mode = ctx->grid_kind == GRID_RADIAL ? 0 : (ctx->grid_kind == GRID_MULTI ? 1 : 2);
    hipLaunchKernelGGL(k_tile_bounds, dim3((unsigned)tiles), dim3(kT), 0, st, ctx->d_gx,
                       ctx->d_gy, ctx->d_gz, G, P, mode, ctx->d_map, ctx->d_gcent, ntile,
                       ctx->d_tb, ridx);
    const int64_t nwords = (n + 63) / 64;
    const size_t nw = (size_t)(nwords > 0 ? nwords : 1);
    if ((rc = ensure(ctx, ctx->d_mask, ctx->mask_cap, nw * ntile))) return rc;
    if ((rc = ensure(ctx, ctx->d_wcount, ctx->wcount_cap, nw))) return rc;
    if ((rc = ensure(ctx, ctx->d_wflop, ctx->wflop_cap, nw))) return rc;
    if (nwords > 0) {
      hipLaunchKernelGGL(k_skip_mask, dim3((unsigned)((nwords + 3) / 4)), dim3(kT), 0, st, gh,
                         kGStride, n, ctx->d_tb, ntile, nwords, ctx->d_mask, ctx->d_wcount, 11,
                         ctx->d_wflop);
      hipLaunchKernelGGL(k_sum_u64, dim3(1), dim3(kT), 0, st, ctx->d_wcount, nwords, ctx->d_evals);
      hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(kT), 0, st, ctx->d_wflop, nwords, ctx->d_flops);
    }
  } else {
    hipLaunchKernelGGL(k_fill_tiles, dim3((unsigned)((ntile + kT - 1) / kT)), dim3(kT), 0, st,
                       ntile, ctx->d_tb);
  }
  INCA_HIP(hipGetLastError());
#define INCA_GEVAL2(PP, CC)                                                                    \
  if (rays) {                                                                                  \
    if constexpr (PP >= 2) {                                                                   \
      if (ctx->ray_kind == 2)                                                                  \
        hipLaunchKernelGGL((k_geval_rays<PP, CC, true>), grid, dim3(kT), 0, st, gh, n, gch,    \
                           ctx->d_gpool, ctx->d_gx, ctx->d_gy, ctx->d_gz, G, ctx->d_tb,        \
                           ctx->d_mask, ntile, tiles, nch, xcd, ridx, ctx->ray_h, partial);    \
      else                                                                                     \
        hipLaunchKernelGGL((k_geval_rays<PP, CC, false>), grid, dim3(kT), 0, st, gh, n, gch,   \
                           ctx->d_gpool, ctx->d_gx, ctx->d_gy, ctx->d_gz, G, ctx->d_tb,        \
                           ctx->d_mask, ntile, tiles, nch, xcd, ridx, ctx->ray_h, partial);    \
    }                                                                                          \
  } else if (mixed)                                                                            \
    hipLaunchKernelGGL((k_geval<PP, CC, true>), grid, dim3(kT), 0, st, gh, n, gch,             \
                       ctx->d_gpool, ctx->d_gpool32, ctx->d_gx, ctx->d_gy, ctx->d_gz, G,       \
                       ctx->d_tb, ctx->d_mask, ntile, tiles, nch, xcd, partial);               \
  else                                                                                         \
    hipLaunchKernelGGL((k_geval<PP, CC, false>), grid, dim3(kT), 0, st, gh, n, gch,            \
                       ctx->d_gpool, ctx->d_gpool32, ctx->d_gx, ctx->d_gy, ctx->d_gz, G,       \
                       ctx->d_tb, ctx->d_mask, ntile, tiles, nch, xcd, partial);
#define INCA_GEVAL(PP)     \
  if (cut) {               \
    INCA_GEVAL2(PP, true)  \
  } else {                 \
    INCA_GEVAL2(PP, false) \
  }
  switch (P) {
    case 1: INCA_GEVAL(1) break;
    case 2: INCA_GEVAL(2) break;
    case 6: INCA_GEVAL(6) break;
    case 8: INCA_GEVAL(8) break;
    default: INCA_GEVAL(4) break;
  }
#undef INCA_GEVAL
#undef INCA_GEVAL2
  INCA_HIP(hipGetLastError());
  return 0;
}

// The skip mask of a launch holds one word per 64 groups and wave tile.
// Where that passes kMaskBudget bytes (large molecules on large grids), the
// chunks run in consecutive parts, each with the mask of its own groups
// and writing its own rows of the partial sums.
constexpr size_t kMaskBudget = (size_t)2 << 30;
int launch_geval(inca_ctx *ctx, bool mixed, int P, const double *gh, int64_t n, int64_t gch,
                 int64_t nch, double *partial) {
  if (P != 1 && P != 2 && P != 6 && P != 8) P = 4;
  const int64_t G = ctx->G;
  const int64_t ntile = (G + (int64_t)kT * P - 1) / ((int64_t)kT * P) * (kT / 64);
  const size_t per_chunk = (size_t)((gch + 63) / 64) * (size_t)ntile * sizeof(unsigned long long);
  int64_t cpp = nch;  // chunks per part
  if (ctx->cut_log2 > 0.0 && per_chunk * (size_t)nch > kMaskBudget)
    cpp = std::max<int64_t>(1, (int64_t)(kMaskBudget / std::max<size_t>(per_chunk, 1)));
  for (int64_t c0 = 0; c0 < nch; c0 += cpp) {
    const int64_t c1 = std::min(nch, c0 + cpp);
    const int64_t g0 = c0 * gch, g1 = std::min(n, c1 * gch);
    if (g1 <= g0) break;
    int rc = launch_geval_part(ctx, mixed, P, gh + g0 * kGStride, g1 - g0, gch, c1 - c0,
                               partial + c0 * G);
    if (rc) return rc;
  }
  return 0;
}

// Chunking of n quartets into `share` of a budget of (tile, chunk) blocks.
// Small chunks keep the quartet stream of the blocks resident on one XCD
// inside its L2 (measured on the radial-plot benchmark: 8192 blocks 689 ms,
// 131072 599 ms, 262144 582 ms + 3 ms of chunk reduction, 524288 no
// better); chunks keep at least kMinChunk quartets so that small systems
// do not end up with one partial row per quartet.
constexpr int64_t kBlockBudget = 262144;
constexpr int64_t kMinChunk = 512;
void chunking(int64_t n, int64_t tiles, double share, int64_t *qch, int64_t *nch) {
  int64_t want = (int64_t)((double)kBlockBudget * share) / tiles + 1;
  const int64_t cap = (n + kMinChunk - 1) / kMinChunk;
  if (want > cap) want = cap;
  if (want > (1 << 20)) want = 1 << 20;
  if (want < 1) want = 1;
  *qch = (n + want - 1) / want;
  *nch = (n + *qch - 1) / *qch;
}

}  // namespace

// The cube contraction over every prepared quartet: factors for batches of
// quartets (at most 2^28 doubles of them at a time), each batch contracted
// into the same `*nrows` partial rows, one per chunk, added in batch order.
// The FP64 range (all quartets in FP64; the leading q_fp64 of a mixed
// evaluation) runs on the vector kernel (mode 1) or the FP64 matrix cores
// (mode 2); the FP32 range of a mixed evaluation runs on FP32 factors and
// the FP32 matrix cores, folded into FP64.
int cube_contract(inca_ctx *ctx, int precision, int64_t *nrows) {
  hipStream_t st = ctx->stream;
  const int nx = ctx->cube_n[0], ny = ctx->cube_n[1], nz = ctx->cube_n[2];
  const int64_t G = ctx->G, nq = ctx->q2;
  const int64_t njk = (int64_t)ny * nz;
  if ((int64_t)nx * njk != G) return set_err(ctx, INCA_EINVAL, "cube grid of unexpected size");
  const int64_t n64 = (precision == INCA_MIXED) ? ctx->q_fp64 : nq;
  const int64_t n32 = nq - n64;
  const bool mma64 = ctx->cube_contract == 2;
  int ncu = 0, per_cu = 0;
  INCA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  hipError_t oe = hipSuccess;
  // matrix-core tiles: 16 x values by a 16 x 16 (j, k) tile per block
  const int nx16 = (nx + 15) / 16 * 16;
  const int64_t per_mma = (int64_t)(nx16 / 16) * ((ny + 15) / 16) * ((nz + 15) / 16);
  int IB = 8, nxp = nx16;
  int64_t ni = 1, per_chunk = per_mma;
  if (n64 > 0 && !mma64) {
    // vector units: x values per lane, the choice that wastes the fewest,
    // the larger on ties
    const int ibs[5] = {8, 16, 24, 32, 40};
    int64_t best = INT64_MAX;
    for (int c : ibs) {
      const int64_t span = (nx + c - 1) / c * (int64_t)c;
      if (span <= best) {
        best = span;
        IB = c;
      }
    }
    ni = (nx + IB - 1) / IB;
    const int64_t njkb = (int64_t)((ny + 15) / 16) * ((nz + 15) / 16);  // 16 x 16 (j, k) tiles
    nxp = (int)(ni * IB);  // x factors zero-padded to whole i blocks
    per_chunk = ni * njkb;
    switch (IB) {
      case 8: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cube_contract<8>, kT, 0); break;
      case 16: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cube_contract<16>, kT, 0); break;
      case 24: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cube_contract<24>, kT, 0); break;
      case 32: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cube_contract<32>, kT, 0); break;
      default: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cube_contract<40>, kT, 0); break;
    }
  } else if (n64 > 0) {
    oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cube_contract_mma<double>, kT, 0);
  } else {
    oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cube_contract_mma<float>, kT, 0);
  }
  if (oe != hipSuccess) return hip_fail(ctx, oe, "occupancy query", __LINE__);
  const int64_t nfp = (int64_t)nxp + ny + nz;     // FP64 factors per quartet
  const int64_t nfp32 = (int64_t)nx16 + ny + nz;  // FP32 factors per quartet
  // chunks: the blocks of a batch fill two rounds of the device (every block
  // does the same work, so a partial last round would idle most CUs)
  const int64_t resident = (int64_t)std::max(1, ncu) * std::max(1, per_cu);
  // quartets per factor batch: within 2^28 doubles (2 GB; 2^26 ran as
  // fast, 2^30 slower as each chunk's rows left the caches), and at most half
  // of them beyond 8192 (no batch need hold every quartet)
  auto batch = [&](int64_t n, int64_t per) {
    return std::max<int64_t>(1, std::min<int64_t>({n, ((int64_t)1 << 28) / per,
                                                   std::max<int64_t>(4096, (n + 1) / 2)}));
  };
  const int64_t nb = n64 > 0 ? batch(n64, nfp) : 0;
  const int64_t nb32 = n32 > 0 ? batch(n32, (nfp32 + 1) / 2) : 0;  // two floats per double
  int64_t cpb = std::max<int64_t>(1, 2 * resident / per_chunk);
  cpb = std::min<int64_t>(cpb, std::max<int64_t>(1, (std::max(nb, nb32) + 63) / 64));
  if (cpb > 65535) cpb = 65535;
  int rc;
  const size_t cf_doubles = std::max<size_t>((size_t)(nb * nfp), (size_t)((nb32 * nfp32 + 1) / 2));
  if ((rc = ensure(ctx, ctx->d_cf, ctx->cf_cap, cf_doubles))) return rc;
  if ((rc = ensure(ctx, ctx->d_partial, ctx->partial_cap, (size_t)(cpb * G)))) return rc;
  for (int64_t qb = 0; qb < n64; qb += nb) {
    const int64_t n = std::min(nb, n64 - qb);
    const int64_t nt = n * nfp;
    hipLaunchKernelGGL(k_cube_factors, dim3((unsigned)((nt + kT - 1) / kT)), dim3(kT), 0, st,
                       ctx->d_q + qb * ctx->qstride, ctx->qstride, n, nx, nxp, ny, nz, ctx->d_gx,
                       ctx->d_gy, ctx->d_gz, ctx->d_cf);
    // every batch fills all cpb rows (a chunk may be empty: it writes zeros)
    const int64_t qch = (n + cpb - 1) / cpb;
    const int accumulate = qb > 0;
    if (mma64) {
      hipLaunchKernelGGL(k_cube_contract_mma<double>, dim3((unsigned)per_mma, (unsigned)cpb),
                         dim3(kT), 0, st, ctx->d_cf, n, qch, nx, nxp, ny, nz, accumulate,
                         ctx->d_partial, G);
    } else {
      const dim3 grid((unsigned)per_chunk, (unsigned)cpb);
#define INCA_CUBE(IBV)                                                                         \
  hipLaunchKernelGGL(k_cube_contract<IBV>, grid, dim3(kT), 0, st, ctx->d_cf, n, qch, nx, ny, nz, \
                     ni, accumulate, ctx->d_partial, G)
      switch (IB) {
        case 8: INCA_CUBE(8); break;
        case 16: INCA_CUBE(16); break;
        case 24: INCA_CUBE(24); break;
        case 32: INCA_CUBE(32); break;
        default: INCA_CUBE(40); break;
      }
#undef INCA_CUBE
    }
    INCA_HIP(hipGetLastError());
  }
  float *cf32 = reinterpret_cast<float *>(ctx->d_cf);
  for (int64_t qb = 0; qb < n32; qb += nb32) {
    const int64_t n = std::min(nb32, n32 - qb);
    const int64_t nt = n * nfp32;
    hipLaunchKernelGGL(k_cube_factors_f32, dim3((unsigned)((nt + kT - 1) / kT)), dim3(kT), 0, st,
                       ctx->d_q + (n64 + qb) * ctx->qstride, ctx->qstride, n, nx, nx16, ny, nz,
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, cf32);
    const int64_t qch = (n + cpb - 1) / cpb;
    const int accumulate = n64 > 0 || qb > 0;
    hipLaunchKernelGGL(k_cube_contract_mma<float>, dim3((unsigned)per_mma, (unsigned)cpb),
                       dim3(kT), 0, st, (const float *)cf32, n, qch, nx, nx16, ny, nz, accumulate,
                       ctx->d_partial, G);
    INCA_HIP(hipGetLastError());
  }
  *nrows = cpb;
  return 0;
}

int evaluate_grid(inca_ctx *ctx, int precision) {
  hipStream_t st = ctx->stream;
  const int64_t G = ctx->G, nq = ctx->q2;
  int rc;
  if ((rc = ensure(ctx, ctx->d_ivec, ctx->ivec_cap, (size_t)(G > 0 ? G : 1)))) return rc;
  INCA_HIP(hipEventRecord(ctx->ev[1], st));
  if (G == 0 || nq == 0) {
    if (G > 0) {
      hipLaunchKernelGGL(k_zero, dim3((unsigned)((G + kT - 1) / kT)), dim3(kT), 0, st,
                         ctx->d_ivec, G);
      INCA_HIP(hipGetLastError());
    }
    INCA_HIP(hipEventRecord(ctx->ev[2], st));
    INCA_HIP(hipEventRecord(ctx->ev[3], st));
    ctx->have_ivec = 1;
    return 0;
  }
  if ((rc = ensure(ctx, ctx->d_evals, ctx->evals_cap, 1))) return rc;
  if ((rc = ensure(ctx, ctx->d_flops, ctx->flops_cap, 1))) return rc;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, ctx->d_evals);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(kT), 0, st, ctx->d_flops, (int64_t)1);
  const bool contracted = ctx->grid_kind == GRID_CUBE && ctx->cube_contract;
  // shell-quartet groups where they exist (every grid the contraction does
  // not take); FP64 groups first, then the FP32 range of a mixed preparation
  const bool grouped = !contracted && ctx->grouping && ctx->ng > 0;
  ctx->grouped_last = grouped;
  if (!grouped && (rc = need_records(ctx))) return rc;
  if (precision == INCA_MIXED && !grouped && (rc = split_records(ctx))) return rc;
  int64_t nchunks = 0;
  bool rays_used = false;  // a range ran on the ray lists (duplicate points skipped)
  if (contracted) {
    if ((rc = cube_contract(ctx, precision, &nchunks))) return rc;
  } else if (grouped) {
    const int64_t ng = ctx->ng;
    const int64_t n64 = (precision == INCA_MIXED) ? ctx->ng_fp64 : ng;
    const int64_t nr[2] = {n64, ng - n64};
    int64_t nch[2] = {0, 0}, gch[2] = {1, 1};
    for (int r = 0; r < 2; ++r) {
      if (nr[r] == 0) continue;
      int64_t tiles;
      bool rays;
      if ((rc = group_tiles(ctx, r == 1, ctx->gtile_p[r], &tiles, &rays))) return rc;
      if (rays && (rc = ensure_grho(ctx))) return rc;
      rays_used = rays_used || rays;
      chunking(nr[r], tiles, (double)nr[r] / (double)ng, &gch[r], &nch[r]);
    }
    nchunks = nch[0] + nch[1];
    if ((rc = ensure(ctx, ctx->d_partial, ctx->partial_cap, (size_t)(nchunks * G)))) return rc;
    if (nch[0] > 0 && (rc = launch_geval(ctx, false, ctx->gtile_p[0], ctx->d_gh, nr[0], gch[0],
                                         nch[0], ctx->d_partial)))
      return rc;
    if (nch[1] > 0 &&
        (rc = launch_geval(ctx, true, ctx->gtile_p[1], ctx->d_gh + n64 * kGStride, nr[1], gch[1],
                           nch[1], ctx->d_partial + nch[0] * G)))
      return rc;
  } else {
    // ranges: [0, n64) in FP64, [n64, nq) in the mixed kernel
    const int64_t n64 = (precision == INCA_MIXED) ? ctx->q_fp64 : nq;
    const int64_t nr[2] = {n64, nq - n64};
    // the kernel of each range: FP64 with tile_p[0]; the FP32 range with the
    // mixed kernel, or (tile_p[1] == 0) on the FP64 ray tiles where the grid has
    // them
    bool mixedk[2] = {false, true};
    int Pk[2] = {ctx->tile_p[0], ctx->tile_p[1]};
    if (Pk[1] == 0) {
      int64_t t;
      bool ry;
      if ((rc = tile_layout(ctx, false, Pk[0], &t, &ry))) return rc;
      if (ry) {
        mixedk[1] = false;
        Pk[1] = Pk[0];
      } else {
        Pk[1] = 8;
      }
    }
    int64_t nch[2] = {0, 0}, qch[2] = {1, 1};
    for (int r = 0; r < 2; ++r) {
      if (nr[r] == 0) continue;
      int64_t tiles;
      bool rays;
      if ((rc = tile_layout(ctx, mixedk[r], Pk[r], &tiles, &rays))) return rc;
      if (rays && (rc = ensure_rho(ctx))) return rc;
      rays_used = rays_used || rays;
      chunking(nr[r], tiles, (double)nr[r] / (double)nq, &qch[r], &nch[r]);
    }
    nchunks = nch[0] + nch[1];
    if ((rc = ensure(ctx, ctx->d_partial, ctx->partial_cap, (size_t)(nchunks * G)))) return rc;
    if (nch[0] > 0 &&
        (rc = launch_eval(ctx, mixedk[0], Pk[0], ctx->d_q, nr[0], qch[0], nch[0], ctx->d_partial)))
      return rc;
    if (nch[1] > 0 &&
        (rc = launch_eval(ctx, mixedk[1], Pk[1], ctx->d_q + n64 * ctx->qstride, nr[1], qch[1],
                          nch[1], ctx->d_partial + nch[0] * G)))
      return rc;
  }
  INCA_HIP(hipEventRecord(ctx->ev[2], st));
  {
    // Chunk partials to Ivec in a fixed order.  With few points per chunk
    // row, one thread per point would leave most CUs idle, so the chunks are
    // first summed in slices (about 2048 blocks in all).
    const int64_t pblk = (G + kT - 1) / kT;
    int64_t nslice = std::min<int64_t>(nchunks, std::max<int64_t>(1, 2048 / pblk));
    if (nslice > 65535) nslice = 65535;
    const double *src = ctx->d_partial;
    int nsum = (int)nchunks;
    if (nslice > 1 && nchunks >= 4 * nslice) {
      if ((rc = ensure(ctx, ctx->d_red, ctx->red_cap, (size_t)(nslice * G)))) return rc;
      hipLaunchKernelGGL(k_reduce_slices, dim3((unsigned)pblk, (unsigned)nslice), dim3(kT), 0, st,
                         ctx->d_partial, (int)nchunks, (int)nslice, G, ctx->d_red);
      src = ctx->d_red;
      nsum = (int)nslice;
    }
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)pblk), dim3(kT), 0, st, src, nsum, G,
                       ctx->d_ivec);
  }
  if (rays_used && ctx->ndup > 0)
    hipLaunchKernelGGL(k_copy_dups, dim3((unsigned)((ctx->ndup + kT - 1) / kT)), dim3(kT), 0, st,
                       ctx->d_dup, ctx->ndup, ctx->d_ivec);
  INCA_HIP(hipGetLastError());
  INCA_HIP(hipEventRecord(ctx->ev[3], st));
  ctx->have_ivec = 1;
  ctx->evals_reference = (int64_t)nq * G;
  // executed evaluations: (record, point) pairs, or (group, point) pairs
  // when the loop ran by groups
  ctx->evals_executed = grouped ? ctx->ng * G : ctx->evals_reference;
  ctx->flops_executed = 0.0;
  if (grouped)
    ctx->flops_executed = (double)G * (ctx->gflop_fp64 + ctx->gflop_fp32);
  if (ctx->cut_log2 > 0.0 && !contracted) {
    unsigned long long ex = 0;
    double fl = 0.0;
    INCA_HIP(hipMemcpyAsync(&ex, ctx->d_evals, sizeof(ex), hipMemcpyDeviceToHost, st));
    INCA_HIP(hipMemcpyAsync(&fl, ctx->d_flops, sizeof(fl), hipMemcpyDeviceToHost, st));
    INCA_HIP(hipStreamSynchronize(st));
    ctx->evals_executed = (int64_t)ex;
    if (grouped) ctx->flops_executed = fl;
  }
  return 0;
}

// Time every points-per-lane choice for both kernels on a slice of the
// prepared quartets of the current grid and keep the fastest.
// The group kernels' points per lane, on an evenly strided sample of the
// groups of each range (FP64 groups, and the FP32 ones of a mixed
// preparation), timed like the record kernels below.
int autotune_groups(inca_ctx *ctx, int32_t *best) {
  const int64_t G = ctx->G;
  const int cand[] = {1, 2, 4, 6, 8};
  hipEvent_t a, b;
  INCA_HIP(hipEventCreate(&a));
  INCA_HIP(hipEventCreate(&b));
  int rc = 0;
  for (int kind = 0; kind < 2 && !rc; ++kind) {
    const int64_t g0 = kind == 0 ? 0 : ctx->ng_fp64;
    const int64_t nr = kind == 0 ? ctx->ng_fp64 : ctx->ng - ctx->ng_fp64;
    if (nr == 0 || (kind == 1 && !ctx->d_gpool32)) {
      if (best) best[kind] = ctx->gtile_p[kind];
      continue;
    }
    int64_t n = ((int64_t)1 << 33) / (G > 0 ? G : 1);
    n = n < 1024 ? 1024 : (n > (1 << 20) ? (1 << 20) : n);
    if (n > nr) n = nr;
    if ((rc = ensure(ctx, ctx->d_qtmp, ctx->qtmp_cap, (size_t)n * kGStride))) break;
    if (ctx->grid_kind == GRID_RADIAL && (rc = ensure_grho(ctx))) break;  // the sample's word 7
    hipLaunchKernelGGL(k_gather_strided, dim3((unsigned)((n * kGStride + kT - 1) / kT)), dim3(kT),
                       0, ctx->stream, ctx->d_gh + g0 * kGStride, nr, n, kGStride, ctx->d_qtmp);
    INCA_HIP(hipGetLastError());
    float best_ms = 1e30f;
    int best_p = 4;
    for (int c = 0; c < (int)(sizeof(cand) / sizeof(cand[0])) && !rc; ++c) {
      const int P = cand[c];
      int64_t tiles;
      bool rays;
      if ((rc = group_tiles(ctx, kind == 1, P, &tiles, &rays))) break;
      int64_t qch, nch;
      chunking(nr, tiles, (double)nr / (double)ctx->ng, &qch, &nch);
      nch = (n + qch - 1) / qch;
      if ((rc = ensure(ctx, ctx->d_partial, ctx->partial_cap, (size_t)(nch * G)))) break;
      float ms = 1e30f;
      for (int rep = 0; rep < 3 && !rc; ++rep) {
        float t = 0.f;
        INCA_HIP(hipEventRecord(a, ctx->stream));
        rc = launch_geval(ctx, kind == 1, P, ctx->d_qtmp, n, qch, nch, ctx->d_partial);
        INCA_HIP(hipEventRecord(b, ctx->stream));
        INCA_HIP(hipEventSynchronize(b));
        INCA_HIP(hipEventElapsedTime(&t, a, b));
        if (rep > 0 && t < ms) ms = t;
      }
      if (ms < best_ms) {
        best_ms = ms;
        best_p = P;
      }
    }
    ctx->gtile_p[kind] = best_p;
    if (best) best[kind] = best_p;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  ctx->have_ivec = 0;
  return rc;
}

int autotune(inca_ctx *ctx, int32_t *best) {
  const int64_t G = ctx->G, nq = ctx->q2;
  if (G == 0 || nq == 0) return set_err(ctx, INCA_EINVAL, "autotune needs a grid and quartets");
  if (ctx->grouping && ctx->ng > 0 && !(ctx->grid_kind == GRID_CUBE && ctx->cube_contract))
    return autotune_groups(ctx, best);
  if (int rc = need_records(ctx)) return rc;
  const int cand[] = {1, 2, 4, 6, 8, 10};
  // an evenly strided sample of the prepared quartets (file order groups
  // quartets by primitive, and with the cutoff their cost depends on where
  // they sit), gathered into the pre-pass scratch, which is free after it;
  // about 2^36 quartet x point evaluations per timed launch (4096 to 2^22
  // quartets), enough blocks for the launch to run at its steady occupancy
  int64_t n = ((int64_t)1 << 36) / (G > 0 ? G : 1);
  n = n < 4096 ? 4096 : (n > (1 << 22) ? (1 << 22) : n);
  if (n > nq) n = nq;
  if ((size_t)n * ctx->qstride > ctx->qtmp_cap)
    return set_err(ctx, INCA_EINVAL, "autotune needs prepared quartets");
  int rc = ensure_rho(ctx);  // the sample carries word 7
  if (rc) return rc;
  hipLaunchKernelGGL(k_gather_strided, dim3((unsigned)((n * ctx->qstride + kT - 1) / kT)), dim3(kT),
                     0, ctx->stream, ctx->d_q, nq, n, ctx->qstride, ctx->d_qtmp);
  INCA_HIP(hipGetLastError());
  hipEvent_t a, b;
  INCA_HIP(hipEventCreate(&a));
  INCA_HIP(hipEventCreate(&b));
  for (int kind = 0; kind < 2 && !rc; ++kind) {
    float best_ms = 1e30f;
    int best_p = 4;
    for (int c = 0; c < (int)(sizeof(cand) / sizeof(cand[0])) && !rc; ++c) {
      const int P = cand[c];
      int64_t tiles;
      bool rays;
      if ((rc = tile_layout(ctx, kind == 1, P, &tiles, &rays))) break;
      // chunks of the size the whole evaluation would use, so that the
      // sample sees the same L2 reuse per chunk
      int64_t qch, nch;
      chunking(nq, tiles, 1.0, &qch, &nch);
      nch = (n + qch - 1) / qch;
      if ((rc = ensure(ctx, ctx->d_partial, ctx->partial_cap, (size_t)(nch * G)))) break;
      float ms = 1e30f;
      for (int rep = 0; rep < 3 && !rc; ++rep) {  // first launch warms up; best of two
        float t = 0.f;
        INCA_HIP(hipEventRecord(a, ctx->stream));
        rc = launch_eval(ctx, kind == 1, P, ctx->d_qtmp, n, qch, nch, ctx->d_partial);
        INCA_HIP(hipEventRecord(b, ctx->stream));
        INCA_HIP(hipEventSynchronize(b));
        INCA_HIP(hipEventElapsedTime(&t, a, b));
        if (rep > 0 && t < ms) ms = t;
      }
      if (ms < best_ms) {
        best_ms = ms;
        best_p = P;
      }
    }
    if (kind == 1 && !rc) {
      // the FP32 range on the FP64 ray tiles (tile_p[0], tuned above)
      int64_t tiles;
      bool rays;
      if ((rc = tile_layout(ctx, false, ctx->tile_p[0], &tiles, &rays))) break;
      if (rays) {
        int64_t qch, nch;
        chunking(nq, tiles, 1.0, &qch, &nch);
        nch = (n + qch - 1) / qch;
        if ((rc = ensure(ctx, ctx->d_partial, ctx->partial_cap, (size_t)(nch * G)))) break;
        float ms = 1e30f;
        for (int rep = 0; rep < 3 && !rc; ++rep) {
          float t = 0.f;
          INCA_HIP(hipEventRecord(a, ctx->stream));
          rc = launch_eval(ctx, false, ctx->tile_p[0], ctx->d_qtmp, n, qch, nch, ctx->d_partial);
          INCA_HIP(hipEventRecord(b, ctx->stream));
          INCA_HIP(hipEventSynchronize(b));
          INCA_HIP(hipEventElapsedTime(&t, a, b));
          if (rep > 0 && t < ms) ms = t;
        }
        if (ms < best_ms) best_p = 0;
      }
    }
    ctx->tile_p[kind] = best_p;
    if (best) best[kind] = best_p;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  ctx->have_ivec = 0;  // the partial buffer no longer matches Ivec
  return rc;
}

}  // namespace inca

// ---- evaluation without device allocation (inca_evaluate_prepared_noalloc) ----
//
// The call's device memory beyond its inputs (the context's grid and
// prepared quartets) and its output (the caller's Ivec) is the tile records
// and the per-chunk partial rows of the grid loop.  Both live in `scratch`,
// a file-scope array of which every device holds its own copy from module
// load to process end.  Host code never passes the array's name to a kernel
// (on the host it is only a stand-in symbol): inca_create calls
// noalloc_scratch_init once per device, with the device current, and the
// device address hipGetSymbolAddress returns is what the launches take.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;  // 32 MiB per device
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace inca {
namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // per device: the device address of `scratch`

double *noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

// Point-by-point launch (no cutoff, no ray tiles) of one precision range.
template <int P>
void launch_plain(bool mixed, dim3 grid, hipStream_t st, const double *q, int qstride, int64_t n,
                  int64_t qch, const inca_ctx *ctx, const double *tb, int64_t ntile,
                  int64_t tiles, int64_t nch, int xcd, double *partial) {
  if (mixed)
    hipLaunchKernelGGL((k_eval_mixed<P, false>), grid, dim3(kT), 0, st, q, qstride, n, qch,
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, ctx->G, tb,
                       (const unsigned long long *)nullptr, ntile, tiles, nch, xcd, partial);
  else
    hipLaunchKernelGGL((k_eval_fp64<P, false>), grid, dim3(kT), 0, st, q, qstride, n, qch,
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, ctx->G, tb,
                       (const unsigned long long *)nullptr, ntile, tiles, nch, xcd, partial);
}
}  // namespace

hipError_t noalloc_scratch_init(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < (int)g_scratch_ptr.size() && g_scratch_ptr[device]) return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  g_scratch_ptr[device] = static_cast<double *>(p);
  return hipSuccess;
}

// The grid loop of evaluate_grid, point by point with the cutoff off, its
// tile records and partial rows in `scratch`, reduced in a fixed order into
// the caller's device array ivec.  Enqueued on `st`; nothing waits.
int evaluate_noalloc(inca_ctx *ctx, int precision, double *ivec, hipStream_t st) {
  if (ctx->rows_sparse)  // their compaction would allocate and wait
    return set_err(ctx, INCA_EINVAL,
                   "the records of this grouped preparation are not compacted yet: make one "
                   "record-by-record use first (e.g. inca_get_prepared), or prepare with "
                   "grouping off");
  if (int rc = need_records(ctx)) return rc;
  double *base = noalloc_scratch(ctx->device);
  if (!base) return set_err(ctx, INCA_EINVAL, "no-allocation scratch not set up on device %d",
                            ctx->device);
  const int64_t G = ctx->G, nq = ctx->q2;
  const int64_t cap = (int64_t)(kNoAllocScratchBytes / sizeof(double));
  if (G == 0) return 0;
  if (nq == 0) {
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((G + kT - 1) / kT)), dim3(kT), 0, st, ivec, G);
    INCA_HIP(hipGetLastError());
    return 0;
  }
  // ranges: [0, n64) FP64, [n64, nq) FP32 (mixed); points per lane from the
  // context's tiles (a ray-tile choice of the FP32 range runs 8 per lane)
  auto valid_p = [](int p) { return p == 1 || p == 2 || p == 4 || p == 6 || p == 8 || p == 10; };
  // a mixed preparation whose record split is not made yet (it allocates)
  // runs every record in FP64
  const int64_t n64 = (precision == INCA_MIXED && !ctx->rec_split_pending) ? ctx->q_fp64 : nq;
  const int64_t nr[2] = {n64, nq - n64};
  const int Pk[2] = {valid_p(ctx->tile_p[0]) ? ctx->tile_p[0] : 4,
                     valid_p(ctx->tile_p[1]) ? ctx->tile_p[1] : 8};
  int64_t tiles[2] = {0, 0}, ntile_max = 0, nranges = 0;
  for (int r = 0; r < 2; ++r) {
    if (nr[r] == 0) continue;
    ++nranges;
    tiles[r] = (G + (int64_t)kT * Pk[r] - 1) / ((int64_t)kT * Pk[r]);
    ntile_max = std::max<int64_t>(ntile_max, tiles[r] * (kT / 64));
  }
  const int64_t tb_d = (ntile_max * kTB + 31) / 32 * 32;  // 256-byte aligned rows follow
  if (tb_d + nranges * G > cap)
    return set_err(ctx, INCA_ERANGE,
                   "no-allocation evaluation: a grid of %lld points needs %lld bytes of "
                   "scratch, more than the %zu it holds",
                   (long long)G, (long long)((tb_d + nranges * G) * 8), kNoAllocScratchBytes);
  const int64_t rows = (cap - tb_d) / G;
  int64_t qch[2] = {1, 1}, nch[2] = {0, 0};
  for (int r = 0; r < 2; ++r)
    if (nr[r] > 0) chunking(nr[r], tiles[r], (double)nr[r] / (double)nq, &qch[r], &nch[r]);
  if (nch[0] + nch[1] > rows) {
    // fewer, longer chunks: rows shared in proportion to the quartets
    // (the floors sum to at most `rows`, and nranges <= rows)
    for (int r = 0; r < 2; ++r) {
      if (nr[r] == 0) continue;
      const int64_t budget =
          std::max<int64_t>(1, (int64_t)((double)rows * (double)nr[r] / (double)nq));
      qch[r] = (nr[r] + budget - 1) / budget;
      nch[r] = (nr[r] + qch[r] - 1) / qch[r];
    }
  }
  double *tb = base;
  double *partial = base + tb_d;
  hipLaunchKernelGGL(k_fill_tiles, dim3((unsigned)((ntile_max + kT - 1) / kT)), dim3(kT), 0, st,
                     ntile_max, tb);
  INCA_HIP(hipGetLastError());
  int64_t row0 = 0;
  for (int r = 0; r < 2; ++r) {
    if (nr[r] == 0) continue;
    const int xcd = nch[r] >= kXcds;
    const int64_t nblocks = tiles[r] * (xcd ? (nch[r] + kXcds - 1) / kXcds * kXcds : nch[r]);
    const dim3 grid((unsigned)nblocks);
    const double *q = ctx->d_q + (r == 0 ? 0 : n64 * ctx->qstride);
    const int64_t ntile = tiles[r] * (kT / 64);
    double *prow = partial + row0 * G;
    const bool mixed = r == 1;
    switch (Pk[r]) {
      case 1: launch_plain<1>(mixed, grid, st, q, ctx->qstride, nr[r], qch[r], ctx, tb, ntile, tiles[r], nch[r], xcd, prow); break;
      case 2: launch_plain<2>(mixed, grid, st, q, ctx->qstride, nr[r], qch[r], ctx, tb, ntile, tiles[r], nch[r], xcd, prow); break;
      case 6: launch_plain<6>(mixed, grid, st, q, ctx->qstride, nr[r], qch[r], ctx, tb, ntile, tiles[r], nch[r], xcd, prow); break;
      case 8: launch_plain<8>(mixed, grid, st, q, ctx->qstride, nr[r], qch[r], ctx, tb, ntile, tiles[r], nch[r], xcd, prow); break;
      case 10: launch_plain<10>(mixed, grid, st, q, ctx->qstride, nr[r], qch[r], ctx, tb, ntile, tiles[r], nch[r], xcd, prow); break;
      default: launch_plain<4>(mixed, grid, st, q, ctx->qstride, nr[r], qch[r], ctx, tb, ntile, tiles[r], nch[r], xcd, prow); break;
    }
    INCA_HIP(hipGetLastError());
    row0 += nch[r];
  }
  // chunk rows to Ivec in a fixed order, through slices where the scratch
  // has room for them
  const int64_t nchunks = row0;
  const int64_t pblk = (G + kT - 1) / kT;
  int64_t nslice = std::min<int64_t>({nchunks, std::max<int64_t>(1, 2048 / pblk), 65535,
                                      rows - nchunks});
  if (nslice > 1 && nchunks >= 4 * nslice) {
    double *red = partial + nchunks * G;
    hipLaunchKernelGGL(k_reduce_slices, dim3((unsigned)pblk, (unsigned)nslice), dim3(kT), 0, st,
                       partial, (int)nchunks, (int)nslice, G, red);
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)pblk), dim3(kT), 0, st, red, (int)nslice, G, ivec);
  } else {
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)pblk), dim3(kT), 0, st, partial, (int)nchunks, G,
                       ivec);
  }
  INCA_HIP(hipGetLastError());
  return 0;
}

}  // namespace inca
