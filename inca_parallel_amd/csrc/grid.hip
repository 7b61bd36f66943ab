// Grid generation and grid reductions on the device.
//   multicentre radial x Lebedev grid + TFVC   mod_build_grid.f90:13-277, 474-575
//   radial-scan grid                            mod_build_grid.f90:281-408
//   cube grid                                   mod_build_grid.f90:412-440
//   integral / Vee / per-centre split           mod_intracule.f90:441-455
//   radial-scan values                          mod_intracule.f90:522-530
//   Gaussian smearing                           mod_intracule.f90:576-649
// Points are stored SoA (x[], y[], z[]) so the evaluation kernels load them
// coalesced; weights and the centre map sit beside them.  Only the small
// per-centre tables (Gauss-Legendre radii, Lebedev unit vectors) are built
// on the host; every per-point quantity is computed here.
#include <cmath>

#include "common.hpp"

// Grid arithmetic keeps separate multiply and add roundings: the TFVC step
// function cancels near nu = 1, and the w > 1e-106 filter must decide on the
// same values as the FP64 restatement.
#pragma clang fp contract(off)

namespace inca {
namespace {

constexpr int kT = 256;

// ---- multicentre grid ----------------------------------------------------

struct CentreInfo {
  double cx, cy, cz;
  int64_t pfirst;  // first raw point of this centre
  int32_t nrad, nang;
  int32_t rfirst, afirst;  // offsets into radial / angular tables
};

__global__ __launch_bounds__(kT) void k_multi_expand(
    const CentreInfo *__restrict__ ci, int nquad, int64_t npts,
    const double *__restrict__ rad, const double *__restrict__ wl,
    const double4 *__restrict__ ang, int nosym, double *__restrict__ px,
    double *__restrict__ py, double *__restrict__ pz, double *__restrict__ pw,
    int32_t *__restrict__ own, uint8_t *__restrict__ flag) {
  int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= npts) return;
  int c = 0;
  while (c + 1 < nquad && ci[c + 1].pfirst <= g) ++c;
  const CentreInfo C = ci[c];
  int64_t loc = g - C.pfirst;
  int ir = (int)(loc / C.nang), ia = (int)(loc % C.nang);
  double r = rad[C.rfirst + ir], w_r = wl[C.rfirst + ir];
  double4 a = ang[C.afirst + ia];
  double x = r * a.x + C.cx, y = r * a.y + C.cy, z = r * a.z + C.cz;
  const double tol = 1.0e-100;
  double w;
  uint8_t keep = 1;
  if (nosym) {
    w = a.w * w_r;
  } else {
    keep = (z >= -tol);
    w = (fabs(z) < tol) ? a.w * w_r : 2.0 * a.w * w_r;
  }
  px[g] = x;
  py[g] = y;
  pz[g] = z;
  pw[g] = w;
  own[g] = c;
  flag[g] = keep;
}

template <class T>
__global__ __launch_bounds__(kT) void k_gather(const T *__restrict__ src,
                                               const int64_t *__restrict__ idx,
                                               int64_t n, T *__restrict__ dst) {
  int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g < n) dst[g] = src[idx[g]];
}

// TFVC weight of the owning centre times the quadrature weight, then the
// keep flag w > 1e-106 (mod_build_grid.f90:229-247).  Per point: P_i =
// prod_{j != i} s_ij with s = (1 - f4(nu))/2, nu = (1+mu-X(1-mu))/(1+mu+X(1-mu)),
// mu = (r_i - r_j)/R_ij; normalised by sum_i P_i when that exceeds 1e-15.
__device__ __forceinline__ double tfvc_fk(double v) {
#pragma unroll
  for (int i = 0; i < 4; ++i) v = 1.5 * v - 0.5 * (v * v * v);
  return v;
}

__global__ __launch_bounds__(kT) void k_tfvc(
    int64_t n, const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, double *__restrict__ pw,
    const int32_t *__restrict__ own, const double *__restrict__ cent,
    const double *__restrict__ Rij, const double *__restrict__ Xi, int nquad,
    uint8_t *__restrict__ flag) {
  int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= n) return;
  const double x = px[g], y = py[g], z = pz[g];
  const int o = own[g];
  double Ptot = 0.0, Pown = 0.0;
  for (int i = 0; i < nquad; ++i) {
    double dx = x - cent[3 * i], dy = y - cent[3 * i + 1], dz = z - cent[3 * i + 2];
    const double ri = sqrt(dx * dx + dy * dy + dz * dz);
    double Pi = 1.0;
    for (int j = 0; j < nquad; ++j) {
      if (j == i) continue;
      double ex = x - cent[3 * j], ey = y - cent[3 * j + 1], ez = z - cent[3 * j + 2];
      const double rj = sqrt(ex * ex + ey * ey + ez * ez);
      const double R = Rij[i * nquad + j];
      const double mu = (R > 1e-12) ? (ri - rj) * (1.0 / R) : 0.0;
      const double term = Xi[i * nquad + j] * (1.0 - mu);
      const double v = (1.0 + mu - term) / (1.0 + mu + term);
      Pi *= 0.5 * (1.0 - tfvc_fk(v));
    }
    Ptot += Pi;
    if (i == o) Pown = Pi;
  }
  double wb = Pown;
  if (Ptot > 1e-15) wb *= 1.0 / Ptot;
  double w = pw[g] * wb;
  pw[g] = w;
  flag[g] = (w > 1.0e-106);
}

__global__ __launch_bounds__(kT) void k_multi_final(
    int64_t n, const int64_t *__restrict__ idx, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz,
    const double *__restrict__ sw, const int32_t *__restrict__ so,
    double *__restrict__ gx, double *__restrict__ gy, double *__restrict__ gz,
    double *__restrict__ gw, double *__restrict__ gwv,
    int32_t *__restrict__ gmap) {
  int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= n) return;
  int64_t s = idx ? idx[g] : g;
  double x = sx[s], y = sy[s], z = sz[s], w = sw[s];
  double d = sqrt(x * x + y * y + z * z);
  gx[g] = x;
  gy[g] = y;
  gz[g] = z;
  gw[g] = w;
  gwv[g] = w * (1.0 / d);
  gmap[g] = so[s] + 1;
}

__global__ void k_fill_u8(uint8_t *f, int64_t n, uint8_t v) {
  int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g < n) f[g] = v;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

struct Scratch {
  std::vector<void *> ptrs;
  ~Scratch() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int alloc(inca_ctx *ctx, T *&p, int64_t n) {
    p = nullptr;
    size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(T);
    hipError_t e = hipMalloc((void **)&p, bytes);
    if (e != hipSuccess)
      return set_err(ctx, INCA_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    ptrs.push_back(p);
    return 0;
  }
};

// Grid arrays for G points; arrays large enough from an earlier grid are
// kept.
int alloc_grid(inca_ctx *ctx, int64_t G) {
  ++ctx->grid_serial;
  const size_t n = (size_t)(G > 0 ? G : 1);
  double **arrs[5] = {&ctx->d_gx, &ctx->d_gy, &ctx->d_gz, &ctx->d_w,
                      &ctx->d_wvee};
  if (n > ctx->grid_cap || !ctx->d_gx || !ctx->d_map) {
    for (double **a : arrs) {
      if (*a) (void)hipFree(*a);
      *a = nullptr;
    }
    if (ctx->d_map) (void)hipFree(ctx->d_map);
    ctx->d_map = nullptr;
    ctx->grid_cap = 0;
    for (double **a : arrs)
      if (hipMalloc((void **)a, n * sizeof(double)) != hipSuccess)
        return set_err(ctx, INCA_ENOMEM, "grid allocation of %lld points failed",
                       (long long)G);
    if (hipMalloc((void **)&ctx->d_map, n * sizeof(int32_t)) != hipSuccess)
      return set_err(ctx, INCA_ENOMEM, "grid allocation failed");
    ctx->grid_cap = n;
  }
  ctx->G = G;
  ctx->have_ivec = 0;
  return 0;
}

}  // namespace

int grid_multicenter(inca_ctx *ctx, int nquad, const double *cent,
                     const int32_t *nradc, const int32_t *nangc,
                     const double *sfalpha, const double *Ps, int nosym,
                     int64_t *maxgrid, int64_t *rgrid, int64_t *rrgrid) {
  if (nquad < 1) return set_err(ctx, INCA_EINVAL, "nquad must be >= 1");
  const double pi = 4.0 * std::atan(1.0);
  std::vector<CentreInfo> ci(nquad);
  std::vector<double> rad, wl;
  std::vector<double4> ang;
  int64_t np = 0;
  for (int c = 0; c < nquad; ++c) {
    const int nrad = nradc[c], nang = nangc[c];
    if (nrad < 1) return set_err(ctx, INCA_EINVAL, "nradc(%d) = %d", c + 1, nrad);
    std::vector<double> lx(nang), ly(nang), lz(nang), lw(nang);
    if (lebedev(nang, lx.data(), ly.data(), lz.data(), lw.data()))
      return set_err(ctx, INCA_ERANGE, "no Lebedev rule with %d points", nang);
    std::vector<double> xl(nrad), w(nrad);
    gauleg(nrad, xl.data(), w.data());
    ci[c] = CentreInfo{cent[3 * c], cent[3 * c + 1], cent[3 * c + 2], np,
                       nrad, nang, (int32_t)rad.size(), (int32_t)ang.size()};
    // mod_build_grid.f90:125-134
    for (int ir = 0; ir < nrad; ++ir) {
      double r = (1.0 + xl[ir]) / (1.0 - xl[ir]) * sfalpha[c];
      double wr;
      if (nquad > 1 && r > 15.0) {
        wr = 0.0;
      } else {
        double om = 1.0 - xl[ir];
        wr = 2 * pi * (2.0 * sfalpha[c] / (om * om)) * w[ir] * r * r;
      }
      rad.push_back(r);
      wl.push_back(wr);
    }
    for (int ia = 0; ia < nang; ++ia) ang.push_back(make_double4(lx[ia], ly[ia], lz[ia], lw[ia]));
    np += (int64_t)nrad * nang;
  }
  *maxgrid = np;
  // every temporary from the context's scratch block (the symmetry-kept
  // arrays at their upper bound np); the call synchronises before returning
  Carve cv;
  const size_t o_ci = cv.add<CentreInfo>(nquad), o_rad = cv.add<double>(rad.size()),
               o_wl = cv.add<double>(wl.size()), o_ang = cv.add<double4>(ang.size()),
               o_rx = cv.add<double>(np), o_ry = cv.add<double>(np), o_rz = cv.add<double>(np),
               o_rw = cv.add<double>(np), o_rown = cv.add<int32_t>(np),
               o_fl = cv.add<uint8_t>(np), o_idx = cv.add<int64_t>(np),
               o_cent = cv.add<double>(3 * nquad), o_R = cv.add<double>((int64_t)nquad * nquad),
               o_X = cv.add<double>((int64_t)nquad * nquad), o_kx = cv.add<double>(np),
               o_ky = cv.add<double>(np), o_kz = cv.add<double>(np), o_kw = cv.add<double>(np),
               o_kown = cv.add<int32_t>(np);
  int rc = 0;
  if ((rc = ensure(ctx, ctx->d_tmp, ctx->tmp_cap, cv.size))) return rc;
  char *tb = ctx->d_tmp;
  CentreInfo *d_ci = Carve::at<CentreInfo>(tb, o_ci);
  double *d_rad = Carve::at<double>(tb, o_rad), *d_wl = Carve::at<double>(tb, o_wl);
  double4 *d_ang = Carve::at<double4>(tb, o_ang);
  double *rx = Carve::at<double>(tb, o_rx), *ry = Carve::at<double>(tb, o_ry),
         *rz = Carve::at<double>(tb, o_rz), *rw = Carve::at<double>(tb, o_rw);
  int32_t *rown = Carve::at<int32_t>(tb, o_rown);
  uint8_t *fl = Carve::at<uint8_t>(tb, o_fl);
  int64_t *idx = Carve::at<int64_t>(tb, o_idx);
  double *d_cent = Carve::at<double>(tb, o_cent), *d_R = Carve::at<double>(tb, o_R),
         *d_X = Carve::at<double>(tb, o_X);
  hipStream_t st = ctx->stream;
  INCA_HIP(hipMemcpyAsync(d_ci, ci.data(), nquad * sizeof(CentreInfo), hipMemcpyHostToDevice, st));
  INCA_HIP(hipMemcpyAsync(d_rad, rad.data(), rad.size() * sizeof(double), hipMemcpyHostToDevice, st));
  INCA_HIP(hipMemcpyAsync(d_wl, wl.data(), wl.size() * sizeof(double), hipMemcpyHostToDevice, st));
  INCA_HIP(hipMemcpyAsync(d_ang, ang.data(), ang.size() * sizeof(double4), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_multi_expand, dim3(nblk(np)), dim3(kT), 0, st, d_ci, nquad, np,
                     d_rad, d_wl, d_ang, nosym, rx, ry, rz, rw, rown, fl);
  INCA_HIP(hipGetLastError());
  int64_t nk = 0;
  if ((rc = compact_flags(ctx, fl, np, idx, &nk))) return rc;
  *rgrid = nk;
  // gather the symmetry-kept points
  double *kx = Carve::at<double>(tb, o_kx), *ky = Carve::at<double>(tb, o_ky),
         *kz = Carve::at<double>(tb, o_kz), *kw = Carve::at<double>(tb, o_kw);
  int32_t *kown = Carve::at<int32_t>(tb, o_kown);
  if (nk > 0) {
    hipLaunchKernelGGL(k_gather<double>, dim3(nblk(nk)), dim3(kT), 0, st, rx, idx, nk, kx);
    hipLaunchKernelGGL(k_gather<double>, dim3(nblk(nk)), dim3(kT), 0, st, ry, idx, nk, ky);
    hipLaunchKernelGGL(k_gather<double>, dim3(nblk(nk)), dim3(kT), 0, st, rz, idx, nk, kz);
    hipLaunchKernelGGL(k_gather<double>, dim3(nblk(nk)), dim3(kT), 0, st, rw, idx, nk, kw);
    hipLaunchKernelGGL(k_gather<int32_t>, dim3(nblk(nk)), dim3(kT), 0, st, rown, idx, nk, kown);
    INCA_HIP(hipGetLastError());
  }
  int64_t nf = nk;
  if (nquad > 1 && nk > 0) {
    // centre tables (mod_build_grid.f90:496-508)
    std::vector<double> R((size_t)nquad * nquad), X((size_t)nquad * nquad);
    for (int i = 0; i < nquad; ++i)
      for (int j = 0; j < nquad; ++j) {
        double dx = cent[3 * i] - cent[3 * j], dy = cent[3 * i + 1] - cent[3 * j + 1],
               dz = cent[3 * i + 2] - cent[3 * j + 2];
        R[(size_t)i * nquad + j] = std::sqrt(dx * dx + dy * dy + dz * dz);
        X[(size_t)i * nquad + j] = (Ps[j] > 1e-12) ? std::pow(Ps[i] / Ps[j], 1.0 / 3.0) : 1.0;
      }
    INCA_HIP(hipMemcpyAsync(d_cent, cent, 3 * nquad * sizeof(double), hipMemcpyHostToDevice, st));
    INCA_HIP(hipMemcpyAsync(d_R, R.data(), R.size() * sizeof(double), hipMemcpyHostToDevice, st));
    INCA_HIP(hipMemcpyAsync(d_X, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_tfvc, dim3(nblk(nk)), dim3(kT), 0, st, nk, kx, ky, kz, kw, kown,
                       d_cent, d_R, d_X, nquad, fl);
    INCA_HIP(hipGetLastError());
    if ((rc = compact_flags(ctx, fl, nk, idx, &nf))) return rc;
  } else if (nk > 0) {
    hipLaunchKernelGGL(k_fill_u8, dim3(nblk(nk)), dim3(kT), 0, st, fl, nk, (uint8_t)1);
  }
  if ((rc = alloc_grid(ctx, nf))) return rc;
  if (nf > 0) {
    hipLaunchKernelGGL(k_multi_final, dim3(nblk(nf)), dim3(kT), 0, st, nf,
                       (nquad > 1) ? idx : (const int64_t *)nullptr, kx, ky, kz, kw, kown,
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, ctx->d_w, ctx->d_wvee, ctx->d_map);
    INCA_HIP(hipGetLastError());
  }
  // centre segments (points stay grouped by centre through both compactions)
  std::vector<int32_t> map(nf > 0 ? nf : 1);
  if (nf > 0)
    INCA_HIP(hipMemcpyAsync(map.data(), ctx->d_map, nf * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  INCA_HIP(hipStreamSynchronize(st));
  ctx->centre_first.assign(nquad + 1, nf);
  for (int64_t g = nf - 1; g >= 0; --g) ctx->centre_first[map[g] - 1] = g;
  for (int c = nquad - 1; c >= 0; --c)
    if (ctx->centre_first[c] > ctx->centre_first[c + 1]) ctx->centre_first[c] = ctx->centre_first[c + 1];
  // centre coordinates: the anchors of the evaluation's tile bounds
  if ((rc = ensure(ctx, ctx->d_gcent, ctx->gcent_cap, 3 * (size_t)nquad))) return rc;
  INCA_HIP(hipMemcpyAsync(ctx->d_gcent, cent, 3 * (size_t)nquad * sizeof(double),
                          hipMemcpyHostToDevice, st));
  INCA_HIP(hipStreamSynchronize(st));
  *rrgrid = nf;
  ctx->nquad = nquad;
  ctx->grid_kind = GRID_MULTI;
  return 0;
}

// ---- radial scan ----------------------------------------------------------

namespace {
struct ShellInfo {
  double s;
  int64_t pfirst;  // first raw point
  int32_t nang, afirst;
};

__global__ __launch_bounds__(kT) void k_radial_expand(
    const ShellInfo *__restrict__ sh, int nsh, int64_t npts,
    const double4 *__restrict__ ang, double *__restrict__ px,
    double *__restrict__ py, double *__restrict__ pz, double *__restrict__ pw,
    uint8_t *__restrict__ flag, double tol) {
  int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= npts) return;
  int lo = 0, hi = nsh - 1;
  while (lo < hi) {  // last shell with pfirst <= g
    int mid = (lo + hi + 1) >> 1;
    if (sh[mid].pfirst <= g) lo = mid; else hi = mid - 1;
  }
  const ShellInfo S = sh[lo];
  double4 a = ang[S.afirst + (int)(g - S.pfirst)];
  double z = S.s * a.z;
  px[g] = S.s * a.x;
  py[g] = S.s * a.y;
  pz[g] = z;
  pw[g] = (fabs(z) < tol) ? a.w : 2.0 * a.w;
  flag[g] = (z >= -tol);
}

__global__ __launch_bounds__(kT) void k_shell_count(
    const ShellInfo *__restrict__ sh, int nsh, const uint8_t *__restrict__ flag,
    int32_t *__restrict__ smn) {
  int r = blockIdx.x * kT + threadIdx.x;
  if (r >= nsh) return;
  int c = 0;
  for (int k = 0; k < sh[r].nang; ++k) c += flag[sh[r].pfirst + k];
  smn[r] = c;
}

__global__ __launch_bounds__(kT) void k_radial_final(
    int64_t n, const int64_t *__restrict__ idx, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz,
    const double *__restrict__ sw, double *__restrict__ gx,
    double *__restrict__ gy, double *__restrict__ gz, double *__restrict__ gw,
    double *__restrict__ gwv, int32_t *__restrict__ gmap) {
  int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= n) return;
  int64_t s = idx[g];
  gx[g] = sx[s];
  gy[g] = sy[s];
  gz[g] = sz[s];
  gw[g] = sw[s];
  gwv[g] = 0.0;
  gmap[g] = 1;
}
}  // namespace

int grid_radial_plot(inca_ctx *ctx, int nblock, const double *tart,
                     const double *step, const int32_t *nang, int nosym,
                     int64_t *nradi, int64_t *ngrid) {
  if (nblock < 1) return set_err(ctx, INCA_EINVAL, "nblock must be >= 1");
  std::vector<ShellInfo> sh;
  std::vector<double4> ang;
  int64_t np = 0;
  ctx->radi.clear();
  for (int b = 0; b < nblock; ++b) {
    if (!(step[b] > 0.0)) return set_err(ctx, INCA_EINVAL, "scan_step(%d) must be > 0", b + 1);
    const int na = nang[b];
    std::vector<double> lx(na), ly(na), lz(na), lw(na);
    if (lebedev(na, lx.data(), ly.data(), lz.data(), lw.data()))
      return set_err(ctx, INCA_ERANGE, "no Lebedev rule with %d points", na);
    const int afirst = (int)ang.size();
    for (int ia = 0; ia < na; ++ia) ang.push_back(make_double4(lx[ia], ly[ia], lz[ia], lw[ia]));
    // mod_build_grid.f90:306-326
    const int npb = (int)((tart[2 * b + 1] - tart[2 * b]) / step[b]) + 1;
    for (int j = 0; j < npb; ++j) {
      double s = tart[2 * b] + step[b] * (double)j;
      ctx->radi.push_back(s);
      sh.push_back(ShellInfo{s, np, na, afirst});
      np += na;
    }
  }
  const int nsh = (int)sh.size();
  *nradi = nsh;
  ctx->smn.assign(nsh, 0);
  int rc;
  if (nosym || nsh == 0) {
    // reference quirk: the nosym branch of gridpoints2 keeps no point
    // (smn and rgrid never incremented, mod_build_grid.f90:374-379)
    if ((rc = alloc_grid(ctx, 0))) return rc;
    *ngrid = 0;
    ctx->grid_kind = GRID_RADIAL;
    return 0;
  }
  Scratch S;
  ShellInfo *d_sh;
  double4 *d_ang;
  double *rx, *ry, *rz, *rw;
  uint8_t *fl;
  int64_t *idx;
  int32_t *d_smn;
  if ((rc = S.alloc(ctx, d_sh, nsh)) || (rc = S.alloc(ctx, d_ang, ang.size())) ||
      (rc = S.alloc(ctx, rx, np)) || (rc = S.alloc(ctx, ry, np)) ||
      (rc = S.alloc(ctx, rz, np)) || (rc = S.alloc(ctx, rw, np)) ||
      (rc = S.alloc(ctx, fl, np)) || (rc = S.alloc(ctx, idx, np)) ||
      (rc = S.alloc(ctx, d_smn, nsh)))
    return rc;
  hipStream_t st = ctx->stream;
  INCA_HIP(hipMemcpyAsync(d_sh, sh.data(), nsh * sizeof(ShellInfo), hipMemcpyHostToDevice, st));
  INCA_HIP(hipMemcpyAsync(d_ang, ang.data(), ang.size() * sizeof(double4), hipMemcpyHostToDevice, st));
  const double tol = std::sqrt(2.220446049250313e-16);  // sqrt(epsilon(1d0))
  hipLaunchKernelGGL(k_radial_expand, dim3(nblk(np)), dim3(kT), 0, st, d_sh, nsh, np, d_ang,
                     rx, ry, rz, rw, fl, tol);
  hipLaunchKernelGGL(k_shell_count, dim3(nblk(nsh)), dim3(kT), 0, st, d_sh, nsh, fl, d_smn);
  INCA_HIP(hipGetLastError());
  int64_t nk = 0;
  if ((rc = compact_flags(ctx, fl, np, idx, &nk))) return rc;
  if ((rc = alloc_grid(ctx, nk))) return rc;
  if (nk > 0)
    hipLaunchKernelGGL(k_radial_final, dim3(nblk(nk)), dim3(kT), 0, st, nk, idx, rx, ry, rz, rw,
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, ctx->d_w, ctx->d_wvee, ctx->d_map);
  INCA_HIP(hipGetLastError());
  INCA_HIP(hipMemcpyAsync(ctx->smn.data(), d_smn, nsh * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  INCA_HIP(hipStreamSynchronize(st));
  *ngrid = nk;
  ctx->grid_kind = GRID_RADIAL;
  return 0;
}

// ---- cube and explicit points --------------------------------------------

namespace {
__global__ __launch_bounds__(kT) void k_cube(int64_t n, int ny, int nz, double x0,
                                             double y0, double z0, double sx,
                                             double sy, double sz, double *gx,
                                             double *gy, double *gz, double *gw,
                                             double *gwv, int32_t *gmap) {
  int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= n) return;
  int64_t i = g / ((int64_t)ny * nz);
  int64_t rem = g - i * (int64_t)ny * nz;
  int j = (int)(rem / nz), k = (int)(rem % nz);
  gx[g] = x0 + sx * (double)i;
  gy[g] = y0 + sy * (double)j;
  gz[g] = z0 + sz * (double)k;
  gw[g] = 0.0;
  gwv[g] = 0.0;
  gmap[g] = 1;
}

__global__ __launch_bounds__(kT) void k_split_points(int64_t n, const double *__restrict__ r,
                                                     double *gx, double *gy, double *gz,
                                                     double *gw, double *gwv, int32_t *gmap) {
  int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= n) return;
  gx[g] = r[3 * g];
  gy[g] = r[3 * g + 1];
  gz[g] = r[3 * g + 2];
  gw[g] = 0.0;
  gwv[g] = 0.0;
  gmap[g] = 1;
}
}  // namespace

int grid_cube(inca_ctx *ctx, const double *center, const double *step,
              const int32_t *np, int64_t *ngrid) {
  if (np[0] < 1 || np[1] < 1 || np[2] < 1)
    return set_err(ctx, INCA_EINVAL, "np_i must be positive");
  double xm[3];
  for (int i = 0; i < 3; ++i) {  // mod_build_grid.f90:422-428
    if (std::fmod((double)np[i], 2.0) == 0.0)
      xm[i] = center[i] - (step[i] / 2.0) - (((double)np[i] - 2.0) / 2.0) * step[i];
    else
      xm[i] = center[i] - (((double)np[i] - 1.0) / 2.0) * step[i];
  }
  int64_t n = (int64_t)np[0] * np[1] * np[2];
  int rc;
  if ((rc = alloc_grid(ctx, n))) return rc;
  hipLaunchKernelGGL(k_cube, dim3(nblk(n)), dim3(kT), 0, ctx->stream, n, np[1], np[2], xm[0],
                     xm[1], xm[2], step[0], step[1], step[2], ctx->d_gx, ctx->d_gy, ctx->d_gz,
                     ctx->d_w, ctx->d_wvee, ctx->d_map);
  INCA_HIP(hipGetLastError());
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  *ngrid = n;
  ctx->grid_kind = GRID_CUBE;
  for (int i = 0; i < 3; ++i) {
    ctx->cube_n[i] = np[i];
    ctx->cube_step[i] = step[i];
  }
  return 0;
}

int grid_points(inca_ctx *ctx, int64_t n, const double *r) {
  if (n < 0) return set_err(ctx, INCA_EINVAL, "negative point count");
  int rc;
  if ((rc = alloc_grid(ctx, n))) return rc;
  if (n == 0) {
    ctx->grid_kind = GRID_POINTS;
    return 0;
  }
  Scratch S;
  double *d_r;
  if ((rc = S.alloc(ctx, d_r, 3 * n))) return rc;
  INCA_HIP(hipMemcpyAsync(d_r, r, 3 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_split_points, dim3(nblk(n)), dim3(kT), 0, ctx->stream, n, d_r, ctx->d_gx,
                     ctx->d_gy, ctx->d_gz, ctx->d_w, ctx->d_wvee, ctx->d_map);
  INCA_HIP(hipGetLastError());
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  ctx->grid_kind = GRID_POINTS;
  return 0;
}

// ---- reductions -----------------------------------------------------------

namespace {
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kT / 64];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kT / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

// one block per segment [first[b], first[b+1]): sum a*I and b*I
__global__ __launch_bounds__(kT) void k_segment_dot2(
    const int64_t *__restrict__ first, const double *__restrict__ a,
    const double *__restrict__ b, const double *__restrict__ I,
    double *__restrict__ outa, double *__restrict__ outb) {
  const int64_t lo = first[blockIdx.x], hi = first[blockIdx.x + 1];
  double sa = 0.0, sb = 0.0;
  for (int64_t g = lo + threadIdx.x; g < hi; g += kT) {
    sa += a[g] * I[g];
    if (b) sb += b[g] * I[g];
  }
  sa = block_sum(sa);
  sb = block_sum(sb);
  if (threadIdx.x == 0) {
    outa[blockIdx.x] = sa;
    if (outb) outb[blockIdx.x] = sb;
  }
}

// one thread per bin; loops over all points with the reference's +-4 sigma
// bin window (mod_intracule.f90:611-630)
__global__ __launch_bounds__(kT) void k_smear(int64_t G, const double *__restrict__ gx,
                                              const double *__restrict__ gy,
                                              const double *__restrict__ gz,
                                              const double *__restrict__ w,
                                              const double *__restrict__ I, int nb,
                                              double bw, double sigma, double pref,
                                              double *__restrict__ P) {
  int j = blockIdx.x * kT + threadIdx.x;  // 0-based bin
  if (j >= nb) return;
  const double uj = (double)j * bw;
  double acc = 0.0;
  for (int64_t i = 0; i < G; ++i) {
    double x = gx[i], y = gy[i], z = gz[i];
    double ui = sqrt(x * x + y * y + z * z);
    int jmin = (int)((ui - 4.0 * sigma) / bw) + 1;
    int jmax = (int)((ui + 4.0 * sigma) / bw) + 1;
    if (jmin < 1) jmin = 1;
    if (jmax > nb) jmax = nb;
    if (j + 1 < jmin || j + 1 > jmax) continue;
    double dist = uj - ui;
    double gs = pref * exp(-(dist * dist) / (2.0 * sigma * sigma));
    acc += (w[i] * I[i]) * gs;
  }
  P[j] = acc;
}
}  // namespace

int grid_integrate(inca_ctx *ctx, double *integral, double *vee, double *ipc,
                   double *vpc) {
  if (ctx->grid_kind != GRID_MULTI)
    return set_err(ctx, INCA_EINVAL, "inca_integrate needs a multicentre grid");
  if (!ctx->have_ivec) return set_err(ctx, INCA_EINVAL, "no Ivec: call inca_evaluate first");
  const int nq = ctx->nquad;
  Scratch S;
  int64_t *d_first;
  double *d_a, *d_b;
  int rc;
  if ((rc = S.alloc(ctx, d_first, nq + 1)) || (rc = S.alloc(ctx, d_a, nq)) ||
      (rc = S.alloc(ctx, d_b, nq)))
    return rc;
  INCA_HIP(hipMemcpyAsync(d_first, ctx->centre_first.data(), (nq + 1) * sizeof(int64_t),
                          hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_segment_dot2, dim3(nq), dim3(kT), 0, ctx->stream, d_first, ctx->d_w,
                     ctx->d_wvee, ctx->d_ivec, d_a, d_b);
  INCA_HIP(hipGetLastError());
  std::vector<double> a(nq), b(nq);
  INCA_HIP(hipMemcpyAsync(a.data(), d_a, nq * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  INCA_HIP(hipMemcpyAsync(b.data(), d_b, nq * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  double ti = 0.0, tv = 0.0;
  for (int c = 0; c < nq; ++c) {
    ti += a[c];
    tv += b[c];
    if (ipc) ipc[c] = a[c];
    if (vpc) vpc[c] = b[c];
  }
  if (integral) *integral = ti;
  if (vee) *vee = tv;
  return 0;
}

int grid_radial_values(inca_ctx *ctx, double *rintra) {
  if (ctx->grid_kind != GRID_RADIAL)
    return set_err(ctx, INCA_EINVAL, "inca_radial_values needs a radial-scan grid");
  const int nr = (int)ctx->radi.size();
  if (nr == 0) return 0;
  if (ctx->G == 0) {
    for (int i = 0; i < nr; ++i) rintra[i] = 0.0;
    return 0;
  }
  if (!ctx->have_ivec) return set_err(ctx, INCA_EINVAL, "no Ivec: call inca_evaluate first");
  std::vector<int64_t> first(nr + 1, 0);
  for (int i = 0; i < nr; ++i) first[i + 1] = first[i] + ctx->smn[i];
  Scratch S;
  int64_t *d_first;
  double *d_a;
  int rc;
  if ((rc = S.alloc(ctx, d_first, nr + 1)) || (rc = S.alloc(ctx, d_a, nr))) return rc;
  INCA_HIP(hipMemcpyAsync(d_first, first.data(), (nr + 1) * sizeof(int64_t),
                          hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_segment_dot2, dim3(nr), dim3(kT), 0, ctx->stream, d_first, ctx->d_w,
                     (const double *)nullptr, ctx->d_ivec, d_a, (double *)nullptr);
  INCA_HIP(hipGetLastError());
  INCA_HIP(hipMemcpyAsync(rintra, d_a, nr * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int grid_smearing(inca_ctx *ctx, double u_max, int n_bins, double sigma,
                  double *u, double *P, double *Iavg) {
  if (ctx->grid_kind != GRID_MULTI)
    return set_err(ctx, INCA_EINVAL, "inca_smearing needs a multicentre grid");
  if (!ctx->have_ivec) return set_err(ctx, INCA_EINVAL, "no Ivec: call inca_evaluate first");
  if (n_bins < 2) return set_err(ctx, INCA_EINVAL, "n_bins must be >= 2");
  const double pi = 4.0 * std::atan(1.0);
  const double bw = u_max / (double)(n_bins - 1);
  const double pref = 1.0 / (sigma * std::sqrt(2.0 * pi));
  Scratch S;
  double *d_P;
  int rc;
  if ((rc = S.alloc(ctx, d_P, n_bins))) return rc;
  hipLaunchKernelGGL(k_smear, dim3((n_bins + kT - 1) / kT), dim3(kT), 0, ctx->stream, ctx->G,
                     ctx->d_gx, ctx->d_gy, ctx->d_gz, ctx->d_w, ctx->d_ivec, n_bins, bw, sigma,
                     pref, d_P);
  INCA_HIP(hipGetLastError());
  INCA_HIP(hipMemcpyAsync(P, d_P, n_bins * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  for (int j = 0; j < n_bins; ++j) {
    u[j] = (double)j * bw;
    Iavg[j] = (u[j] > 1e-6) ? P[j] * (1.0 / (4.0 * pi * u[j] * u[j])) : 0.0;
  }
  return 0;
}

}  // namespace inca
