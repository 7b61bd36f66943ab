// Error reporting, deterministic stream compaction, and the host-side
// quadrature tables (Gauss-Legendre nodes, Lebedev-Laikov spheres).
#include <cmath>
#include <cstdarg>

#include "common.hpp"
#include "lebedev_orbits.h"

namespace inca {

int set_err(inca_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int hip_fail(inca_ctx *c, hipError_t e, const char *what, int line) {
  return set_err(c, INCA_EHIP, "HIP error '%s' in %s (line %d)",
                 hipGetErrorString(e), what, line);
}

// ---------------------------------------------------------------------------
// Compaction: count per block, scan block counts in one block, scatter.
// Each block covers kCompTile = 256 threads x 4 flags.

constexpr int kCompThreads = 256;
constexpr int kCompPer = 4;
constexpr int kCompTile = kCompThreads * kCompPer;

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int64_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// exclusive scan of v across the block; returns block total via *tot
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t *tot) {
  __shared__ int64_t wsum[kCompThreads / 64];
  const int lane = __lane_id(), wid = threadIdx.x >> 6;
  int64_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int64_t base = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kCompThreads / 64; ++w) {
    if (w < wid) base += wsum[w];
    all += wsum[w];
  }
  __syncthreads();
  *tot = all;
  return base + inc - v;
}

__global__ __launch_bounds__(kCompThreads) void k_flag_count(
    const uint8_t *__restrict__ f, int64_t n, int64_t *__restrict__ bc) {
  int64_t base = (int64_t)blockIdx.x * kCompTile + threadIdx.x * kCompPer;
  int64_t c = 0;
#pragma unroll
  for (int k = 0; k < kCompPer; ++k)
    if (base + k < n) c += f[base + k];
  int64_t tot;
  (void)block_excl_scan(c, &tot);
  if (threadIdx.x == 0) bc[blockIdx.x] = tot;
}

// single block: exclusive scan of bc[0..nb) in place, total into bc[nb]
__global__ __launch_bounds__(kCompThreads) void k_scan_counts(int64_t *bc,
                                                              int64_t nb) {
  int64_t per = (nb + kCompThreads - 1) / kCompThreads;
  int64_t lo = threadIdx.x * per, hi = lo + per < nb ? lo + per : nb;
  int64_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += bc[i];
  int64_t tot;
  int64_t ex = block_excl_scan(s, &tot);
  for (int64_t i = lo; i < hi; ++i) {
    int64_t v = bc[i];
    bc[i] = ex;
    ex += v;
  }
  if (threadIdx.x == 0) bc[nb] = tot;
}

__global__ __launch_bounds__(kCompThreads) void k_flag_scatter(
    const uint8_t *__restrict__ f, int64_t n, const int64_t *__restrict__ bc,
    int64_t *__restrict__ idx) {
  int64_t base = (int64_t)blockIdx.x * kCompTile + threadIdx.x * kCompPer;
  uint8_t v[kCompPer];
  int64_t c = 0;
#pragma unroll
  for (int k = 0; k < kCompPer; ++k) {
    v[k] = (base + k < n) ? f[base + k] : 0;
    c += v[k];
  }
  int64_t tot;
  int64_t pos = bc[blockIdx.x] + block_excl_scan(c, &tot);
#pragma unroll
  for (int k = 0; k < kCompPer; ++k)
    if (v[k]) idx[pos++] = base + k;
}

int compact_flags(inca_ctx *ctx, const uint8_t *d_flag, int64_t n,
                  int64_t *d_idx, int64_t *total) {
  *total = 0;
  if (n <= 0) return 0;
  int64_t nb = (n + kCompTile - 1) / kCompTile;
  if (int rc = ensure(ctx, ctx->d_bc, ctx->bc_cap, (size_t)nb + 1)) return rc;
  hipLaunchKernelGGL(k_flag_count, dim3((unsigned)nb), dim3(kCompThreads), 0,
                     ctx->stream, d_flag, n, ctx->d_bc);
  INCA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kCompThreads), 0,
                     ctx->stream, ctx->d_bc, nb);
  INCA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_flag_scatter, dim3((unsigned)nb), dim3(kCompThreads), 0,
                     ctx->stream, d_flag, n, ctx->d_bc, d_idx);
  INCA_HIP(hipGetLastError());
  INCA_HIP(hipMemcpyAsync(total, ctx->d_bc + nb, sizeof(int64_t),
                          hipMemcpyDeviceToHost, ctx->stream));
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// Gauss-Legendre on [-1,1] by Newton iteration on P_n from the Chebyshev
// guess cos(pi (i - 1/4)/(n + 1/2)), stopping at |dz| <= 1e-14 -- the
// method of mod_quadratures.f90:37-86 (Numerical Recipes gauleg).
void gauleg(int n, double *t, double *w) {
  const double pi = std::acos(-1.0);
  const int m = (n + 1) / 2;
  for (int i = 1; i <= m; ++i) {
    double z = std::cos(pi * ((double)i - 0.25) / ((double)n + 0.5));
    double z1, p1, p2, pp;
    do {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - ((double)j - 1.0) * p3) / (double)j;
      }
      pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (std::fabs(z - z1) > 1.0e-14);
    t[i - 1] = -z;
    t[n - i] = z;
    w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

// Expand one Oh orbit into its points (all sign and permutation images of
// the generator); returns the number written.
static int expand_orbit(const LebOrbit &o, double *x, double *y, double *z,
                        double *w) {
  int k = 0;
  auto put = [&](double a, double b, double c) {
    x[k] = a;
    y[k] = b;
    z[k] = c;
    w[k] = o.v;
    ++k;
  };
  const double s[2] = {1.0, -1.0};
  switch (o.code) {
    case 1:
      for (int ax = 0; ax < 3; ++ax)
        for (int sg = 0; sg < 2; ++sg) {
          double v[3] = {0.0, 0.0, 0.0};
          v[ax] = s[sg];
          put(v[0], v[1], v[2]);
        }
      break;
    case 2: {  // zero coordinate on one axis, +-a on the other two
      const double a = std::sqrt(0.5);
      for (int ax = 0; ax < 3; ++ax)
        for (int s1 = 0; s1 < 2; ++s1)
          for (int s2 = 0; s2 < 2; ++s2) {
            double v[3];
            v[ax] = 0.0;
            v[(ax + 1) % 3] = s[s2] * a;
            v[(ax + 2) % 3] = s[s1] * a;
            put(v[0], v[1], v[2]);
          }
      break;
    }
    case 3: {
      const double a = std::sqrt(1.0 / 3.0);
      for (int s1 = 0; s1 < 2; ++s1)
        for (int s2 = 0; s2 < 2; ++s2)
          for (int s3 = 0; s3 < 2; ++s3) put(s[s3] * a, s[s2] * a, s[s1] * a);
      break;
    }
    case 4: {  // (a,a,b): b on each axis in turn
      const double a = o.a, b = std::sqrt(1.0 - 2.0 * a * a);
      for (int ax = 2; ax >= 0; --ax)
        for (int s1 = 0; s1 < 2; ++s1)
          for (int s2 = 0; s2 < 2; ++s2)
            for (int s3 = 0; s3 < 2; ++s3) {
              double v[3] = {a, a, a};
              v[ax] = b;
              put(s[s3] * v[0], s[s2] * v[1], s[s1] * v[2]);
            }
      break;
    }
    case 5: {  // (a,b,0) and (b,a,0) on each coordinate plane
      const double a = o.a, b = std::sqrt(1.0 - a * a);
      for (int ax = 2; ax >= 0; --ax)
        for (int sw = 0; sw < 2; ++sw)
          for (int s1 = 0; s1 < 2; ++s1)
            for (int s2 = 0; s2 < 2; ++s2) {
              double p = sw ? b : a, q = sw ? a : b;
              double v[3];
              v[ax] = 0.0;
              v[(ax + 1) % 3] = s[s2] * p;
              v[(ax + 2) % 3] = s[s1] * q;
              put(v[0], v[1], v[2]);
            }
      break;
    }
    case 6: {  // all permutations of (a,b,c) with all signs
      const double a = o.a, b = o.b, c = std::sqrt(1.0 - a * a - b * b);
      const double g[3] = {a, b, c};
      const int perm[6][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1},
                              {2, 1, 0}, {0, 2, 1}, {1, 2, 0}};
      for (int p = 0; p < 6; ++p)
        for (int s1 = 0; s1 < 2; ++s1)
          for (int s2 = 0; s2 < 2; ++s2)
            for (int s3 = 0; s3 < 2; ++s3)
              put(s[s3] * g[perm[p][0]], s[s2] * g[perm[p][1]],
                  s[s1] * g[perm[p][2]]);
      break;
    }
    default:
      break;
  }
  return k;
}

int lebedev(int n, double *x, double *y, double *z, double *w) {
  for (int r = 0; r < LEB_NRULES; ++r) {
    if (LEB_SIZE[r] != n) continue;
    int k = 0;
    for (int o = LEB_FIRST[r]; o < LEB_FIRST[r + 1]; ++o)
      k += expand_orbit(LEB_ORBITS[o], x + k, y + k, z + k, w + k);
    return k == n ? 0 : -1;
  }
  return -1;
}

}  // namespace inca
