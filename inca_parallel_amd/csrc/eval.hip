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
// Arithmetic.  This is exp-bound VALU work (no MFMA: nothing here is a
// matrix product).  FP64 path: rp by one FMA per axis, exp(-s) by
// reduction by ln2/256 to |r| <= ln2/512, a degree-4 Taylor polynomial of
// e^r (truncation error 3.8e-17), a 256-entry table of 2^(j/256) in LDS and
// an exponent shift (eval_core.hpp).  Mixed path: rp in FP64 (the cancellation-prone step),
// everything after it in FP32 -- hardware exp2, FP32 Horner -- with a
// compensated (Kahan) FP32 sum over tiles of quartets that is folded into
// an FP64 accumulator at the end of every tile.
#include <cmath>

#include "common.hpp"
#include "eval_core.hpp"

namespace inca {
namespace {

// Tile bounds.  Wave w of block b owns the 64 P consecutive grid points
// from (4 b + w) 64 P on.  For each such tile: an anchor c and the range
// [rho_min, rho_max] of |r - c| over its points (words 0..4), and whether
// it holds any point (word 5).  Anchors: the origin for radial scans (the
// points lie on spheres about it), the owning centre for multicentre grids
// when the whole tile belongs to one centre, else the tile's centroid.
constexpr int kTB = 8;  // doubles per tile record

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__global__ __launch_bounds__(kT) void k_tile_bounds(const double *__restrict__ gx,
                                                    const double *__restrict__ gy,
                                                    const double *__restrict__ gz, int64_t G,
                                                    int P, int mode,
                                                    const int32_t *__restrict__ map,
                                                    const double *__restrict__ gcent,
                                                    int64_t ntile, double *__restrict__ tb) {
  const int lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  if (t >= ntile) return;  // uniform per wave
  const int64_t g0 = t * 64 * P;
  const int first = (mode == 1 && g0 < G) ? map[g0] : 0;
  double sx = 0.0, sy = 0.0, sz = 0.0, n = 0.0;
  int same = 1;
  for (int k = 0; k < P; ++k) {
    const int64_t g = g0 + (int64_t)k * 64 + lane;
    if (g < G) {
      sx += gx[g];
      sy += gy[g];
      sz += gz[g];
      n += 1.0;
      if (mode == 1 && map[g] != first) same = 0;
    }
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  sz = wave_sum(sz);
  n = wave_sum(n);
  same = __all(same);
  double cx = 0.0, cy = 0.0, cz = 0.0;
  if (mode == 1 && same && first > 0) {
    cx = gcent[3 * (first - 1)];
    cy = gcent[3 * (first - 1) + 1];
    cz = gcent[3 * (first - 1) + 2];
  } else if (mode != 0 && n > 0.0) {
    cx = sx / n;
    cy = sy / n;
    cz = sz / n;
  }
  double lo = HUGE_VAL, hi = 0.0;
  for (int k = 0; k < P; ++k) {
    const int64_t g = g0 + (int64_t)k * 64 + lane;
    if (g < G) {
      const double dx = gx[g] - cx, dy = gy[g] - cy, dz = gz[g] - cz;
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      lo = fmin(lo, d);
      hi = fmax(hi, d);
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, d, 64));
    hi = fmax(hi, __shfl_xor(hi, d, 64));
  }
  if (lane == 0) {
    double *o = tb + t * kTB;
    o[0] = cx;
    o[1] = cy;
    o[2] = cz;
    // widened by far more than the roundoff of the distances
    o[3] = n > 0.0 ? lo * (1.0 - 1e-12) : 0.0;
    o[4] = n > 0.0 ? hi * (1.0 + 1e-12) + 1e-300 : 0.0;
    o[5] = n;
    o[6] = 0.0;
    o[7] = 0.0;
  }
}

__global__ __launch_bounds__(kT) void k_fill_tiles(int64_t ntile, double *__restrict__ tb) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= ntile) return;
  for (int w = 0; w < kTB; ++w) tb[t * kTB + w] = (w == 5) ? 1.0 : 0.0;
}

// True when every point of the tile lies beyond the quartet's cutoff radius:
// |rp| = |sqe (r - c) + u|, u = sqe c + sqe d, is at least
// max(|u| - sqe rho_max, sqe rho_min - |u|) on the tile.
__device__ __forceinline__ bool tile_far(const QHeader &h, double cx, double cy, double cz,
                                         double rlo, double rhi) {
  const double ux = fma(h.sqe, cx, h.ax), uy = fma(h.sqe, cy, h.ay), uz = fma(h.sqe, cz, h.az);
  const double u2 = fma(ux, ux, fma(uy, uy, uz * uz));
  const double o = fma(h.sqe, rhi, h.rcut);
  const double i = fma(h.sqe, rlo, -h.rcut);
  return (u2 >= o * o) || (i > 0.0 && u2 <= i * i);
}

// Skip mask: bit i of mask[w][t] is set when quartet 64 w + i is beyond its
// cutoff radius at every point of wave tile t.  One wave per 64-quartet word:
// lane i holds quartet 64 w + i, a ballot per tile gives the word, and the
// words of 64 consecutive tiles are stored by the 64 lanes at once.
__global__ __launch_bounds__(kT) void k_skip_mask(const double *__restrict__ q, int qstride,
                                                  int64_t nq, const double *__restrict__ tb,
                                                  int64_t ntile, int64_t nwords,
                                                  unsigned long long *__restrict__ mask) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  if (w >= nwords) return;  // uniform per wave
  const int64_t qi = w * 64 + lane;
  const bool valid = qi < nq;
  QHeader h{};
  if (valid) h = load_header(q + qi * qstride);
  for (int64_t t0 = 0; t0 < ntile; t0 += 64) {
    unsigned long long keep = 0;
    const int nt = (int)((ntile - t0) < 64 ? (ntile - t0) : 64);
    for (int j = 0; j < nt; ++j) {
      const double *__restrict__ tbw = tb + (t0 + j) * kTB;
      const bool far = valid && tbw[5] > 0.0 && tile_far(h, tbw[0], tbw[1], tbw[2], tbw[3], tbw[4]);
      const unsigned long long m = __ballot(far);
      if (lane == j) keep = m;
    }
    if (lane < nt) mask[w * ntile + t0 + lane] = keep;
  }
}

// Executed evaluations under the skip mask: for every word and tile, the
// kept quartets times the tile's points (integer atomics: exact, order-free).
__global__ __launch_bounds__(kT) void k_mask_evals(const unsigned long long *__restrict__ mask,
                                                   const double *__restrict__ tb, int64_t ntile,
                                                   int64_t nwords, int64_t nq,
                                                   unsigned long long *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  unsigned long long v = 0;
  if (i < nwords * ntile) {
    const int64_t w = i / ntile, t = i - w * ntile;
    unsigned long long live = ~mask[i];
    const int64_t left = nq - w * 64;
    if (left < 64) live &= (1ull << left) - 1ull;
    v = (unsigned long long)__popcll(live) * (unsigned long long)tb[t * kTB + 5];
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, v);
}

// Calls f(qi) for the quartets of [a, b) that the skip mask keeps for
// `tile` (all of them when mask is null).  The mask words are wave-uniform,
// so the iteration runs on scalar bit scans.
template <class F>
__device__ __forceinline__ void for_live(const unsigned long long *__restrict__ mask, int64_t ntile,
                                         int64_t tile, int64_t a, int64_t b, F &&f) {
  if (!mask) {
    for (int64_t qi = a; qi < b; ++qi) f(qi);
    return;
  }
  for (int64_t wd = a >> 6; (wd << 6) < b; ++wd) {
    const int64_t base = wd << 6;
    unsigned long long live = ~mask[wd * ntile + tile];
    if (base < a) live &= ~0ull << (a - base);
    if (b - base < 64) live &= (1ull << (b - base)) - 1ull;
    while (live) {
      const int i = __builtin_ctzll(live);
      live &= live - 1ull;
      f(base + i);
    }
  }
}

template <int P, bool CUT>
__global__ __launch_bounds__(kT) void k_eval_fp64(const double *__restrict__ q, int qstride,
                                                  int64_t nq, int64_t qchunk,
                                                  const double *__restrict__ gx,
                                                  const double *__restrict__ gy,
                                                  const double *__restrict__ gz, int64_t G,
                                                  const double *__restrict__ tb,
                                                  const unsigned long long *__restrict__ mask,
                                                  int64_t ntile, double *__restrict__ partial) {
  const int64_t tile = (int64_t)blockIdx.x * (kT / 64) +
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
  const int64_t q0 = (int64_t)blockIdx.y * qchunk;
  const int64_t q1 = empty ? q0 : (q0 + qchunk < nq ? q0 + qchunk : nq);
  for_live(CUT ? mask : nullptr, ntile, tile, q0, q1, [&](int64_t qi) {
    const double *__restrict__ rec = q + qi * qstride;
    const QHeader h = load_header(rec);
    accumulate_fp64<P>(rec, h, x, y, z, acc, s_exp2);
  });
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t g = base + (int64_t)k * 64;
    if (g < G) partial[(int64_t)blockIdx.y * G + g] = acc[k];
  }
}

// Mixed precision: FP64 rp, FP32 exp/polynomials, Kahan FP32 per tile of
// kMixTile quartets folded into FP64.
constexpr int kMixTile = 128;

template <int P, bool CUT>
__global__ __launch_bounds__(kT) void k_eval_mixed(const double *__restrict__ q, int qstride,
                                                   int64_t nq, int64_t qchunk,
                                                   const double *__restrict__ gx,
                                                   const double *__restrict__ gy,
                                                   const double *__restrict__ gz, int64_t G,
                                                   const double *__restrict__ tb,
                                                   const unsigned long long *__restrict__ mask,
                                                   int64_t ntile, double *__restrict__ partial) {
  const int64_t tile = (int64_t)blockIdx.x * (kT / 64) +
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
  const int64_t q0 = (int64_t)blockIdx.y * qchunk;
  const int64_t q1 = empty ? q0 : (q0 + qchunk < nq ? q0 + qchunk : nq);
  // -log2(e) split so that s*log2(e) is formed to FP32 rounding without the
  // 1.3e-8 relative bias of a single-float constant
  const float kNL2eHi = -1.44269502162933349609375f;
  const float kNL2eLo = -1.925963033500011e-08f;
  for (int64_t t0 = q0; t0 < q1; t0 += kMixTile) {
    const int64_t t1 = t0 + kMixTile < q1 ? t0 + kMixTile : q1;
    float sum[P], comp[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      sum[k] = 0.0f;
      comp[k] = 0.0f;
    }
    for_live(CUT ? mask : nullptr, ntile, tile, t0, t1, [&](int64_t qi) {
      const double *__restrict__ rec = q + qi * qstride;
      const QHeader h = load_header(rec);
      const double *__restrict__ cx = rec + kQHead;
      const double *__restrict__ cy = cx + h.Lx + 1;
      const double *__restrict__ cz = cy + h.Ly + 1;
      float px[P], py[P], pz[P], vx[P], vy[P], vz[P];
      const float c0x = (float)cx[0], c0y = (float)cy[0], c0z = (float)cz[0];
#pragma unroll
      for (int k = 0; k < P; ++k) {
        px[k] = (float)fma(h.sqe, x[k], h.ax);
        py[k] = (float)fma(h.sqe, y[k], h.ay);
        pz[k] = (float)fma(h.sqe, z[k], h.az);
        vx[k] = c0x;
        vy[k] = c0y;
        vz[k] = c0z;
      }
      for (int p = 1; p <= h.Lx; ++p) {
        const float c = (float)cx[p];
#pragma unroll
        for (int k = 0; k < P; ++k) vx[k] = fmaf(vx[k], px[k], c);
      }
      for (int p = 1; p <= h.Ly; ++p) {
        const float c = (float)cy[p];
#pragma unroll
        for (int k = 0; k < P; ++k) vy[k] = fmaf(vy[k], py[k], c);
      }
      for (int p = 1; p <= h.Lz; ++p) {
        const float c = (float)cz[p];
#pragma unroll
        for (int k = 0; k < P; ++k) vz[k] = fmaf(vz[k], pz[k], c);
      }
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const float s = fmaf(px[k], px[k], fmaf(py[k], py[k], pz[k] * pz[k]));
        const float e = __builtin_amdgcn_exp2f(fmaf(s, kNL2eHi, s * kNL2eLo));
        const float term = e * vx[k] * vy[k] * vz[k];
        const float yk = term - comp[k];
        const float tk = sum[k] + yk;
        comp[k] = (tk - sum[k]) - yk;
        sum[k] = tk;
      }
    });
#pragma unroll
    for (int k = 0; k < P; ++k) acc[k] += (double)sum[k] - (double)comp[k];
  }
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t g = base + (int64_t)k * 64;
    if (g < G) partial[(int64_t)blockIdx.y * G + g] = acc[k];
  }
}

__global__ __launch_bounds__(kT) void k_reduce(const double *__restrict__ partial, int nchunks,
                                               int64_t G, double *__restrict__ ivec) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= G) return;
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += partial[(int64_t)c * G + g];
  ivec[g] = s;
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
  if (threadIdx.x < 64 && threadIdx.x < nbins && h[threadIdx.x])
    atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
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

// Launch one evaluation kernel over quartets [q, q+n) in `nch` chunks of
// `qch` quartets with P points per lane; partial sums go to partial[chunk][g].
int launch_eval(inca_ctx *ctx, bool mixed, int P, const double *q, int64_t n, int64_t qch,
                int64_t nch, double *partial) {
  const int64_t G = ctx->G;
  const int64_t tiles = (G + (int64_t)kT * P - 1) / ((int64_t)kT * P);
  dim3 grid((unsigned)tiles, (unsigned)nch);
  hipStream_t st = ctx->stream;
  // bounds of the wave tiles of this P (a few microseconds; redone per launch)
  const bool cut = ctx->cut_log2 > 0.0;
  const int64_t ntile = tiles * (kT / 64);
  int rc;
  if ((rc = ensure(ctx, ctx->d_tb, ctx->tb_cap, (size_t)ntile * kTB))) return rc;
  if (cut) {
    const int mode = ctx->grid_kind == GRID_RADIAL ? 0 : (ctx->grid_kind == GRID_MULTI ? 1 : 2);
    hipLaunchKernelGGL(k_tile_bounds, dim3((unsigned)tiles), dim3(kT), 0, st, ctx->d_gx,
                       ctx->d_gy, ctx->d_gz, G, P, mode, ctx->d_map, ctx->d_gcent, ntile,
                       ctx->d_tb);
    const int64_t nwords = (n + 63) / 64;
    if ((rc = ensure(ctx, ctx->d_mask, ctx->mask_cap, (size_t)(nwords > 0 ? nwords : 1) * ntile)))
      return rc;
    if (nwords > 0) {
      hipLaunchKernelGGL(k_skip_mask, dim3((unsigned)((nwords + 3) / 4)), dim3(kT), 0, st, q,
                         ctx->qstride, n, ctx->d_tb, ntile, nwords, ctx->d_mask);
      hipLaunchKernelGGL(k_mask_evals, dim3((unsigned)((nwords * ntile + kT - 1) / kT)), dim3(kT),
                         0, st, ctx->d_mask, ctx->d_tb, ntile, nwords, n, ctx->d_evals);
    }
  } else {
    // only the point-count word is read: mark every tile non-empty
    hipLaunchKernelGGL(k_fill_tiles, dim3((unsigned)((ntile + kT - 1) / kT)), dim3(kT), 0, st,
                       ntile, ctx->d_tb);
  }
  INCA_HIP(hipGetLastError());
#define INCA_EVAL2(PP, CC)                                                                     \
  if (mixed)                                                                                   \
    hipLaunchKernelGGL((k_eval_mixed<PP, CC>), grid, dim3(kT), 0, st, q, ctx->qstride, n, qch, \
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, G, ctx->d_tb, ctx->d_mask, ntile,     \
                       partial);                                                                \
  else                                                                                         \
    hipLaunchKernelGGL((k_eval_fp64<PP, CC>), grid, dim3(kT), 0, st, q, ctx->qstride, n, qch,  \
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, G, ctx->d_tb, ctx->d_mask, ntile,     \
                       partial);
#define INCA_EVAL(PP)      \
  if (cut) {               \
    INCA_EVAL2(PP, true)   \
  } else {                 \
    INCA_EVAL2(PP, false)  \
  }
  switch (P) {
    case 1: INCA_EVAL(1) break;
    case 2: INCA_EVAL(2) break;
    case 8: INCA_EVAL(8) break;
    default: INCA_EVAL(4) break;
  }
#undef INCA_EVAL
#undef INCA_EVAL2
  INCA_HIP(hipGetLastError());
  return 0;
}

// chunking of n quartets: enough (tile, chunk) blocks to fill 256 CUs
// several times over (`share` of the 8192-block budget)
void chunking(int64_t n, int64_t tiles, double share, int64_t *qch, int64_t *nch) {
  int64_t want = (int64_t)(8192.0 * share) / tiles + 1;
  if (want > n) want = n;
  if (want > 65535) want = 65535;
  if (want < 1) want = 1;
  *qch = (n + want - 1) / want;
  *nch = (n + *qch - 1) / *qch;
}

}  // namespace

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
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, ctx->d_evals);
  // ranges: [0, n64) in FP64, [n64, nq) in the mixed kernel
  const int64_t n64 = (precision == INCA_MIXED) ? ctx->q_fp64 : nq;
  const int64_t nr[2] = {n64, nq - n64};
  int64_t nch[2] = {0, 0}, qch[2] = {1, 1};
  for (int r = 0; r < 2; ++r) {
    if (nr[r] == 0) continue;
    const int P = ctx->tile_p[r];
    const int64_t tiles = (G + (int64_t)kT * P - 1) / ((int64_t)kT * P);
    chunking(nr[r], tiles, (double)nr[r] / (double)nq, &qch[r], &nch[r]);
  }
  const int64_t nchunks = nch[0] + nch[1];
  if ((rc = ensure(ctx, ctx->d_partial, ctx->partial_cap, (size_t)(nchunks * G)))) return rc;
  if (nch[0] > 0 &&
      (rc = launch_eval(ctx, false, ctx->tile_p[0], ctx->d_q, nr[0], qch[0], nch[0],
                        ctx->d_partial)))
    return rc;
  if (nch[1] > 0 &&
      (rc = launch_eval(ctx, true, ctx->tile_p[1], ctx->d_q + n64 * ctx->qstride, nr[1], qch[1],
                        nch[1], ctx->d_partial + nch[0] * G)))
    return rc;
  INCA_HIP(hipEventRecord(ctx->ev[2], st));
  hipLaunchKernelGGL(k_reduce, dim3((unsigned)((G + kT - 1) / kT)), dim3(kT), 0, st,
                     ctx->d_partial, (int)nchunks, G, ctx->d_ivec);
  INCA_HIP(hipGetLastError());
  INCA_HIP(hipEventRecord(ctx->ev[3], st));
  ctx->have_ivec = 1;
  ctx->evals_reference = (int64_t)nq * G;
  ctx->evals_executed = ctx->evals_reference;
  if (ctx->cut_log2 > 0.0) {
    unsigned long long ex = 0;
    INCA_HIP(hipMemcpyAsync(&ex, ctx->d_evals, sizeof(ex), hipMemcpyDeviceToHost, st));
    INCA_HIP(hipStreamSynchronize(st));
    ctx->evals_executed = (int64_t)ex;
  }
  return 0;
}

// Time every points-per-lane choice for both kernels on a slice of the
// prepared quartets of the current grid and keep the fastest.
int autotune(inca_ctx *ctx, int32_t *best) {
  const int64_t G = ctx->G, nq = ctx->q2;
  if (G == 0 || nq == 0) return set_err(ctx, INCA_EINVAL, "autotune needs a grid and quartets");
  const int cand[4] = {1, 2, 4, 8};
  const int64_t n = nq < 262144 ? nq : 262144;
  hipEvent_t a, b;
  INCA_HIP(hipEventCreate(&a));
  INCA_HIP(hipEventCreate(&b));
  int rc = 0;
  for (int kind = 0; kind < 2 && !rc; ++kind) {
    float best_ms = 1e30f;
    int best_p = 4;
    for (int c = 0; c < 4 && !rc; ++c) {
      const int P = cand[c];
      const int64_t tiles = (G + (int64_t)kT * P - 1) / ((int64_t)kT * P);
      int64_t qch, nch;
      chunking(n, tiles, 1.0, &qch, &nch);
      if ((rc = ensure(ctx, ctx->d_partial, ctx->partial_cap, (size_t)(nch * G)))) break;
      float ms = 0.f;
      for (int rep = 0; rep < 2 && !rc; ++rep) {  // first launch warms up
        INCA_HIP(hipEventRecord(a, ctx->stream));
        rc = launch_eval(ctx, kind == 1, P, ctx->d_q, n, qch, nch, ctx->d_partial);
        INCA_HIP(hipEventRecord(b, ctx->stream));
        INCA_HIP(hipEventSynchronize(b));
        INCA_HIP(hipEventElapsedTime(&ms, a, b));
      }
      if (ms < best_ms) {
        best_ms = ms;
        best_p = P;
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
