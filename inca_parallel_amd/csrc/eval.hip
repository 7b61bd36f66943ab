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
// everything after it in FP32 -- hardware exp2, FP32 Horner, FP32 products
// -- and every term added to an FP64 accumulator.
#include <cmath>

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
    const QHeader h = load_header(rec);
    accumulate_fp64<P>(rec, h, x, y, z, acc, s_exp2);
  });
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t g = base + (int64_t)k * 64;
    if (g < G) partial[task.chunk * G + g] = acc[k];
  }
}

// Mixed precision: rp in FP64 (sqe r + sqe d cancels), everything after it
// in FP32 -- the hardware exp2 and FP32 Horner -- and each term added to an
// FP64 accumulator.  The body is specialised on the axes of nonzero degree
// like the FP64 kernel's.
template <int P, bool MX, bool MY, bool MZ>
__device__ __forceinline__ void accumulate_mixed(const double *__restrict__ rec, const QHeader &h,
                                                 const double (&x)[P], const double (&y)[P],
                                                 const double (&z)[P], double (&acc)[P]) {
  // -log2(e) split so that s*log2(e) is formed to FP32 rounding without the
  // 1.3e-8 relative bias of a single-float constant
  const float kNL2eHi = -1.44269502162933349609375f;
  const float kNL2eLo = -1.925963033500011e-08f;
  const double *__restrict__ cx = rec + kQHead;
  const double *__restrict__ cy = cx + h.Lx + 1;
  const double *__restrict__ cz = cy + h.Ly + 1;
  float px[P], py[P], pz[P], vx[P], vy[P], vz[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    px[k] = (float)fma(h.sqe, x[k], h.ax);
    py[k] = (float)fma(h.sqe, y[k], h.ay);
    pz[k] = (float)fma(h.sqe, z[k], h.az);
  }
  auto horner32 = [&](const double *__restrict__ c, int L, const float(&p)[P], float(&v)[P]) {
    const float c0 = (float)c[0], c1 = (float)c[1];
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = fmaf(p[k], c0, c1);
    for (int j = 2; j <= L; ++j) {
      const float cj = (float)c[j];
#pragma unroll
      for (int k = 0; k < P; ++k) v[k] = fmaf(v[k], p[k], cj);
    }
  };
  if (MX) horner32(cx, h.Lx, px, vx);
  if (MY) horner32(cy, h.Ly, py, vy);
  if (MZ) horner32(cz, h.Lz, pz, vz);
  float S = 1.0f;
  if (!MX) S = (float)cx[0];
  if (!MY) S = MX ? (float)cy[0] : S * (float)cy[0];
  if (!MZ) S = (MX && MY) ? (float)cz[0] : S * (float)cz[0];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const float s = fmaf(px[k], px[k], fmaf(py[k], py[k], pz[k] * pz[k]));
    float f = __builtin_amdgcn_exp2f(fmaf(s, kNL2eHi, s * kNL2eLo));
    if (!(MX && MY && MZ)) f *= S;
    if (MX) f *= vx[k];
    if (MY) f *= vy[k];
    if (MZ) f *= vz[k];
    acc[k] += (double)f;
  }
}

template <int P>
__device__ __forceinline__ void accumulate_mixed_any(const double *__restrict__ rec,
                                                     const QHeader &h, const double (&x)[P],
                                                     const double (&y)[P], const double (&z)[P],
                                                     double (&acc)[P]) {
  const int mask = (h.Lx > 0) | ((h.Ly > 0) << 1) | ((h.Lz > 0) << 2);
  switch (mask) {
    case 0: accumulate_mixed<P, false, false, false>(rec, h, x, y, z, acc); break;
    case 1: accumulate_mixed<P, true, false, false>(rec, h, x, y, z, acc); break;
    case 2: accumulate_mixed<P, false, true, false>(rec, h, x, y, z, acc); break;
    case 3: accumulate_mixed<P, true, true, false>(rec, h, x, y, z, acc); break;
    case 4: accumulate_mixed<P, false, false, true>(rec, h, x, y, z, acc); break;
    case 5: accumulate_mixed<P, true, false, true>(rec, h, x, y, z, acc); break;
    case 6: accumulate_mixed<P, false, true, true>(rec, h, x, y, z, acc); break;
    default: accumulate_mixed<P, true, true, true>(rec, h, x, y, z, acc); break;
  }
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
    const QHeader h = load_header(rec);
    accumulate_mixed_any<P>(rec, h, x, y, z, acc);
  });
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t g = base + (int64_t)k * 64;
    if (g < G) partial[task.chunk * G + g] = acc[k];
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
  const int xcd = nch >= kXcds;
  const int64_t nblocks = tiles * (xcd ? (nch + kXcds - 1) / kXcds * kXcds : nch);
  dim3 grid((unsigned)nblocks);
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
    if ((rc = ensure(ctx, ctx->d_wcount, ctx->wcount_cap, (size_t)(nwords > 0 ? nwords : 1))))
      return rc;
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
#define INCA_EVAL2(PP, CC)                                                                     \
  if (mixed)                                                                                   \
    hipLaunchKernelGGL((k_eval_mixed<PP, CC>), grid, dim3(kT), 0, st, q, ctx->qstride, n, qch, \
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, G, ctx->d_tb, ctx->d_mask, ntile,     \
                       tiles, nch, xcd, partial);                                              \
  else                                                                                         \
    hipLaunchKernelGGL((k_eval_fp64<PP, CC>), grid, dim3(kT), 0, st, q, ctx->qstride, n, qch,  \
                       ctx->d_gx, ctx->d_gy, ctx->d_gz, G, ctx->d_tb, ctx->d_mask, ntile,     \
                       tiles, nch, xcd, partial);
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
  // an evenly strided sample of the prepared quartets (file order groups
  // quartets by primitive, and with the cutoff their cost depends on where
  // they sit), gathered into the pre-pass scratch, which is free after it
  const int64_t n = nq < (1 << 20) ? nq : (1 << 20);
  if ((size_t)n * ctx->qstride > ctx->qtmp_cap)
    return set_err(ctx, INCA_EINVAL, "autotune needs prepared quartets");
  hipLaunchKernelGGL(k_gather_strided, dim3((unsigned)((n * ctx->qstride + kT - 1) / kT)), dim3(kT),
                     0, ctx->stream, ctx->d_q, nq, n, ctx->qstride, ctx->d_qtmp);
  INCA_HIP(hipGetLastError());
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
        rc = launch_eval(ctx, kind == 1, P, ctx->d_qtmp, n, qch, nch, ctx->d_partial);
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
