// One-electron functions on the current grid: the cube files of the
// reference's &cube_job (cubegen/fwrite, mod_cubefile.f90:16-128) evaluated
// on the device with the functions of mod_functions.f90:
//   Prim    (:17-31)   x^t y^m z^n exp(-alpha r^2) about the primitive's
//                      centre, unnormalised;
//   sd      (:113-152) its Laplacian;
//   MoOr    (:199-210) sum over primitives with |T(mo,p)| > 1e-10;
//   Density (:312-329) sum_i w_i MoOr_i^2 (w_i = 2 for a closed shell, 1 per
//                      spin for an open one);
//   Lapl    (:351-372) sum_i 2 w_i MoOr_i sum_p T(i,p) sd_p (all p).
// One thread per point; the coefficient of a primitive is the same for every
// lane (scalar loads), and the orbitals of a density or Laplacian are
// accumulated kMoPass at a time in registers, the primitive values being
// recomputed for each pass.  These are small, exp-bound elementwise jobs
// (points x primitives x orbitals), not matrix products.
#include <cmath>

#include "common.hpp"

namespace inca {
namespace {

constexpr int kT = 256;
constexpr int kMoPass = 16;

__device__ __forceinline__ double ipow_d(double x, int n) {
  double r = 1.0;
  for (int k = 0; k < n; ++k) r *= x;
  return r;
}

// value and (LAP) Laplacian of primitive p at (x, y, z)
template <bool LAP>
__device__ __forceinline__ double prim_at(int p, int nprim, const double *__restrict__ prim,
                                          const int32_t *__restrict__ tmn, double x, double y,
                                          double z, double *lap) {
  const double a = prim[p];
  const double d[3] = {x - prim[nprim + p], y - prim[2 * nprim + p], z - prim[3 * nprim + p]};
  const int t[3] = {tmn[p], tmn[nprim + p], tmn[2 * nprim + p]};
  const double e = exp(-a * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
  double f[3], f2[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const double xn = ipow_d(d[ax], t[ax]);
    f[ax] = xn;
    if (LAP) {
      // d2/dx2 of x^n exp(-a x^2) without the exponential
      double v = (4.0 * a * a * d[ax] * d[ax] - 2.0 * a * (2 * t[ax] + 1)) * xn;
      if (t[ax] >= 2) v += (double)(t[ax] * (t[ax] - 1)) * ipow_d(d[ax], t[ax] - 2);
      f2[ax] = v;
    }
  }
  if (LAP) *lap = (f2[0] * f[1] * f[2] + f[0] * f2[1] * f[2] + f[0] * f[1] * f2[2]) * e;
  return f[0] * f[1] * f[2] * e;
}

__global__ __launch_bounds__(kT) void k_onee_prim(int p, int nprim, const double *__restrict__ prim,
                                                  const int32_t *__restrict__ tmn,
                                                  const double *__restrict__ gx,
                                                  const double *__restrict__ gy,
                                                  const double *__restrict__ gz, int64_t G,
                                                  double *__restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= G) return;
  out[g] = prim_at<false>(p, nprim, prim, tmn, gx[g], gy[g], gz[g], nullptr);
}

// orbitals [m0, m0 + nm) of T(nmo, nprim) (column-major); KIND 3: the
// orbital itself (nm = 1), 5: sum w phi^2, 6: sum 2 w phi lap(phi)
template <int KIND>
__global__ __launch_bounds__(kT) void k_onee_mo(int nprim, const double *__restrict__ prim,
                                                const int32_t *__restrict__ tmn, int nmo,
                                                const double *__restrict__ T,
                                                const double *__restrict__ w, int m0, int nm,
                                                const double *__restrict__ gx,
                                                const double *__restrict__ gy,
                                                const double *__restrict__ gz, int64_t G,
                                                int accumulate, double *__restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= G) return;
  const double x = gx[g], y = gy[g], z = gz[g];
  double phi[kMoPass], psi[kMoPass];
#pragma unroll
  for (int m = 0; m < kMoPass; ++m) phi[m] = psi[m] = 0.0;
  for (int p = 0; p < nprim; ++p) {
    double lap = 0.0;
    const double v = prim_at<KIND == 6>(p, nprim, prim, tmn, x, y, z, &lap);
    const double *__restrict__ c = T + (int64_t)nmo * p + m0;
#pragma unroll
    for (int m = 0; m < kMoPass; ++m) {
      if (m < nm) {
        const double cm = c[m];
        if (fabs(cm) > 1e-10) phi[m] = fma(cm, v, phi[m]);
        if (KIND == 6) psi[m] = fma(cm, lap, psi[m]);
      }
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < kMoPass; ++m) {
    if (m < nm) {
      if (KIND == 3) acc = phi[m];
      if (KIND == 5) acc += w[m0 + m] * phi[m] * phi[m];
      if (KIND == 6) acc += 2.0 * w[m0 + m] * phi[m] * psi[m];
    }
  }
  out[g] = accumulate ? out[g] + acc : acc;
}

}  // namespace

int onee(inca_ctx *ctx, int kind, int index, int nmo, const double *T, const double *w,
         double *values) {
  const int64_t G = ctx->G;
  const int nprim = ctx->nprim;
  if (G == 0) return 0;
  hipStream_t st = ctx->stream;
  // scratch block: values, then T and w
  Carve cv;
  const size_t o_val = cv.add<double>(G), o_T = cv.add<double>((int64_t)nmo * nprim),
               o_w = cv.add<double>(nmo);
  int rc;
  if ((rc = ensure(ctx, ctx->d_tmp, ctx->tmp_cap, cv.size))) return rc;
  double *d_val = Carve::at<double>(ctx->d_tmp, o_val);
  const unsigned nb = (unsigned)((G + kT - 1) / kT);
  if (kind == 1) {
    hipLaunchKernelGGL(k_onee_prim, dim3(nb), dim3(kT), 0, st, index - 1, nprim, ctx->d_prim,
                       ctx->d_tmn, ctx->d_gx, ctx->d_gy, ctx->d_gz, G, d_val);
  } else {
    double *d_T = Carve::at<double>(ctx->d_tmp, o_T);
    double *d_w = Carve::at<double>(ctx->d_tmp, o_w);
    INCA_HIP(hipMemcpyAsync(d_T, T, (size_t)nmo * nprim * sizeof(double), hipMemcpyHostToDevice,
                            st));
    if (kind != 3)
      INCA_HIP(hipMemcpyAsync(d_w, w, (size_t)nmo * sizeof(double), hipMemcpyHostToDevice, st));
    if (kind == 3) {
      hipLaunchKernelGGL(k_onee_mo<3>, dim3(nb), dim3(kT), 0, st, nprim, ctx->d_prim, ctx->d_tmn,
                         nmo, d_T, d_w, index - 1, 1, ctx->d_gx, ctx->d_gy, ctx->d_gz, G, 0,
                         d_val);
    } else {
      for (int m0 = 0; m0 < nmo; m0 += kMoPass) {
        const int nm = nmo - m0 < kMoPass ? nmo - m0 : kMoPass;
        if (kind == 5)
          hipLaunchKernelGGL(k_onee_mo<5>, dim3(nb), dim3(kT), 0, st, nprim, ctx->d_prim,
                             ctx->d_tmn, nmo, d_T, d_w, m0, nm, ctx->d_gx, ctx->d_gy, ctx->d_gz,
                             G, m0 > 0, d_val);
        else
          hipLaunchKernelGGL(k_onee_mo<6>, dim3(nb), dim3(kT), 0, st, nprim, ctx->d_prim,
                             ctx->d_tmn, nmo, d_T, d_w, m0, nm, ctx->d_gx, ctx->d_gy, ctx->d_gz,
                             G, m0 > 0, d_val);
      }
    }
  }
  INCA_HIP(hipGetLastError());
  INCA_HIP(hipMemcpyAsync(values, d_val, (size_t)G * sizeof(double), hipMemcpyDeviceToHost, st));
  INCA_HIP(hipStreamSynchronize(st));
  return 0;
}

}  // namespace inca
