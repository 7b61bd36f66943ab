// Shell-quartet groups of the prepared records (no reference counterpart;
// the sum they evaluate is the grid loop of mod_intracule.f90:345-358).
//
// A prepared record's term is A exp(-|rp|^2) Vx(rp_x) Vy(rp_y) Vz(rp_z) with
// rp = sqe r + sqe (P_ik - P_jl).  sqe and the offset depend only on the
// exponents and centres of the four primitives, so the Cartesian components
// of one shell quartet -- and any other records that happen to agree -- share
// rp and the Gaussian, and their terms add into one:
//   sum_q A_q exp(-|rp|^2) Vx_q Vy_q Vz_q = exp(-|rp|^2) sum_abc K_abc rpx^a rpy^b rpz^c,
//   K_abc = sum_q (A Cx_q)[a] Cy_q[b] Cz_q[c],
// one exponential and one three-variable polynomial per group and grid point
// in place of one exponential, three axis polynomials and their product per
// record and point.  A shell quartet of d functions holds up to 1296 records
// and needs at most 165 coefficients.
//
// Building the groups (after the pre-pass, deterministic).  The pre-pass
// has already ordered its slots by a 32-bit hash of the bits of (sqe,
// offset) with a stable radix sort (sort_slots below), so the rows of a
// group are adjacent and keep their file order:
//   group starts    wherever the four doubles change (a hash collision only
//                   splits a group, never merges two): marked by the
//                   pre-pass's sorted-record copy, or by k_group_heads
//                   from the rows when those marks are not at hand;
//   k_group_header  one wave per group: degrees, bound, cutoff radius;
//   exclusive scan  of the coefficient counts: the coefficient pool offsets;
//   k_group_coeffs  one wave per group, lanes over coefficients (and over
//                   member slices for groups of up to 64 coefficients), the
//                   members added in a fixed order.
// Group header (kGStride doubles), words 0-7 as a prepared record's:
//   [0] sqe, [1..3] offsets, [4] class Lx | Ly<<8 | Lz<<16 | axes<<24 |
//   Ls<<32 (Ls the members' largest total degree, Lx = Ly = Lz = Ls),
//   [5] B: the members' L1 bounds summed, [6] rcut: their largest cutoff
//   radius, [7] rho (ray step), [8] pool offset, [9] coefficient count M,
//   [10] member count, [11] FP64 operations per point, 14 + 2 M.
// Coefficients: every monomial of total degree <= Ls, for a = Ls..0,
// b = Ls - a..0, c = Ls - a - b..0 (nested Horner order, highest powers
// first); (Ls + 1)(Ls + 2)(Ls + 3) / 6 of them.
#include <algorithm>
#include <cmath>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"

namespace inca {
namespace {

constexpr int kT = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

__device__ __forceinline__ bool same_gaussian(const double *a, const double *b) {
  bool s = true;
  for (int w = 0; w < 4; ++w) s = s && __double_as_longlong(a[w]) == __double_as_longlong(b[w]);
  return s;
}

// Rows arrive sorted by group (the pre-pass orders its slots by a hash of
// the four words): a group starts wherever the words change.
__global__ __launch_bounds__(kT) void k_group_heads(const double *__restrict__ q, int stride,
                                                    int64_t n, uint8_t *__restrict__ head) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  head[i] = (uint8_t)(i == 0 || !same_gaussian(q + i * stride, q + (i - 1) * stride));
}

__host__ __device__ __forceinline__ int monomials(int Lx, int Ly, int Lz, int Ls) {
  int m = 0;
  for (int a = Lx; a >= 0; --a) {
    const int bm = Ly < Ls - a ? Ly : Ls - a;
    for (int b = bm; b >= 0; --b) m += (Lz < Ls - a - b ? Lz : Ls - a - b) + 1;
  }
  return m;
}

__device__ __forceinline__ void record_degrees(const double *r, int &Lx, int &Ly, int &Lz) {
  const unsigned cls = (unsigned)__double_as_longlong(r[4]);
  Lx = cls & 0xff;
  Ly = (cls >> 8) & 0xff;
  Lz = (cls >> 16) & 0xff;
}

// Members of group g: rows first[g] .. first[g+1] (the last group ends at n).
// One wave per group: lane j takes members j, j + 64, ...; the degrees and
// the cutoff radius reduce by max, B by a fixed-order butterfly sum.
__global__ __launch_bounds__(kT) void k_group_header(const double *__restrict__ q, int stride,
                                                     const int64_t *__restrict__ first, int64_t ng,
                                                     int64_t n, double *__restrict__ gh,
                                                     int64_t *__restrict__ mcount) {
  const int64_t g = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  if (g >= ng) return;  // whole wave
  const int lane = threadIdx.x & 63;
  const int64_t m0 = first[g], m1 = (g + 1 < ng) ? first[g + 1] : n;
  int Ls = 0, live = 0;
  double B = 0.0, rcut = 0.0;
  for (int64_t m = m0 + lane; m < m1; m += 64) {
    const double *r = q + m * stride;
    if ((unsigned)__double_as_longlong(r[4]) & kQDropped) continue;  // failed screen 2
    int lx, ly, lz;
    record_degrees(r, lx, ly, lz);
    Ls = max(Ls, lx + ly + lz);
    B += r[5];
    rcut = fmax(rcut, r[6]);
    ++live;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    Ls = max(Ls, __shfl_xor(Ls, d, 64));
    live += __shfl_xor(live, d, 64);
    B += __shfl_xor(B, d, 64);
    rcut = fmax(rcut, __shfl_xor(rcut, d, 64));
  }
  if (lane != 0) return;
  const double *r0 = q + m0 * stride;
  // every monomial of total degree <= Ls is laid out (groups are built for
  // total degrees up to 8, and the evaluation unrolls per degree)
  const int Lx = Ls, Ly = Ls, Lz = Ls;
  const int M = monomials(Lx, Ly, Lz, Ls);
  const int axes = (Lx > 0) | ((Ly > 0) << 1) | ((Lz > 0) << 2);
  double *o = gh + g * kGStride;
  o[0] = r0[0];
  o[1] = r0[1];
  o[2] = r0[2];
  o[3] = r0[3];
  o[4] = __longlong_as_double((long long)((uint64_t)(Lx | (Ly << 8) | (Lz << 16) | (axes << 24)) |
                                          ((uint64_t)Ls << 32)));
  o[5] = B;
  o[6] = rcut;
  o[7] = r0[7];
  o[8] = 0.0;  // pool offset, after the scan
  o[9] = __longlong_as_double((long long)M);
  o[10] = __longlong_as_double((long long)live);
  o[11] = 14.0 + 2.0 * M;
  mcount[g] = M;
}

__global__ __launch_bounds__(kT) void k_group_offsets(const int64_t *__restrict__ off, int64_t ng,
                                                      double *__restrict__ gh) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g < ng) gh[g * kGStride + 8] = __longlong_as_double((long long)off[g]);
}

// One wave per group.  Groups of more than 64 coefficients: lane j of pass
// p owns coefficient 64 p + j and adds the members' products into it in
// member order: (A Cx[a] Cy[b]) Cz[c].
// The members' rows are contiguous.  Per block of 64 members, lane r reads
// member r's degrees once; the member loop takes them back as wave-uniform
// values (readlane) and each lane reads its three factors straight from the
// row (the same two or three cache lines for the whole wave), so a member
// costs a lane three compares, three loads and the product, with the loads
// of consecutive members in flight together.
constexpr int kMaxStride = 36;  // s/p/d records: 8 + 3 (8 + 1) doubles, rounded even

template <int NP>
__device__ __forceinline__ void group_coeffs(const double *__restrict__ q, int stride, int64_t m0,
                                             int64_t m1, int Ls, int M, double *__restrict__ K,
                                             int lane) {
  // this lane's (a, b, c) per pass, in the storage order (every monomial of
  // total degree <= Ls: a = Ls..0, b = Ls - a..0, c = Ls - a - b..0)
  int la[NP], lb[NP], lc[NP];
  double acc[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    la[p] = -1;
    lb[p] = lc[p] = 0;
    acc[p] = 0.0;
    const int idx = 64 * p + lane;
    if (idx >= M) continue;
    int at = 0;
    for (int aa = Ls; aa >= 0 && la[p] < 0; --aa) {
      for (int bb = Ls - aa; bb >= 0; --bb) {
        const int cm = Ls - aa - bb;
        if (idx <= at + cm) {
          la[p] = aa;
          lb[p] = bb;
          lc[p] = cm - (idx - at);
          break;
        }
        at += cm + 1;
      }
    }
  }
  for (int64_t s0 = m0; s0 < m1; s0 += 64) {
    const int ns = (int)((m1 - s0) < 64 ? (m1 - s0) : 64);
    const double *rows = q + s0 * stride;
    // lane r: member r's degrees Lx | Ly << 8 | Lz << 16 (-1: failed screen 2)
    int info = -1;
    if (lane < ns) {
      const unsigned c = (unsigned)__double_as_longlong(rows[(int64_t)lane * stride + 4]);
      info = (c & kQDropped) ? -1 : (int)(c & 0xffffffu);
    }
#pragma unroll 4
    for (int r = 0; r < ns; ++r) {
      const int inf = __builtin_amdgcn_readlane(info, r);
      if (inf < 0) continue;
      const int lx = inf & 0xff, ly = (inf >> 8) & 0xff, lz = (inf >> 16) & 0xff;
      const double *cx = rows + (int64_t)r * stride + kQHead;
      const double *cy = cx + lx + 1;
      const double *cz = cy + ly + 1;
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const int a = la[p], b = lb[p], c = lc[p];
        if (a >= 0 && a <= lx && b <= ly && c <= lz)
          acc[p] += (cx[lx - a] * cy[ly - b]) * cz[lz - c];
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int idx = 64 * p + lane;
    if (idx < M) K[idx] = acc[p];
  }
}

// Groups of at most 64 coefficients: W = M rounded up to a power of two
// lanes per coefficient set and 64 / W slices, slice s adding members s,
// s + 64/W, ... (so a group of 10 coefficients takes four members at a
// time), then a fixed-order tree over the slices.
// Degrees of member m0 + lane (lane < 64, below m1): Lx | Ly << 8 | Lz << 16,
// or -1 (no member, or failed screen 2).
__device__ __forceinline__ int member_info(const double *__restrict__ q, int stride, int64_t m0,
                                           int64_t m1, int lane) {
  if (m0 + lane >= m1) return -1;
  const unsigned c = (unsigned)__double_as_longlong(q[(m0 + lane) * stride + 4]);
  return (c & kQDropped) ? -1 : (int)(c & 0xffffffu);
}

// info0: member_info of the first block of 64, read by the caller.
__device__ __forceinline__ void group_coeffs_sliced(const double *__restrict__ q, int stride,
                                                    int64_t m0, int64_t m1, int Ls, int M,
                                                    double *__restrict__ K, int lane, int info0) {
  int W = 1;
  while (W < M) W <<= 1;
  const int S = 64 / W;
  const int idx = lane & (W - 1), slice = lane / W;
  int la = -1, lb = 0, lc = 0;
  if (idx < M) {
    int at = 0;
    for (int aa = Ls; aa >= 0 && la < 0; --aa) {
      for (int bb = Ls - aa; bb >= 0; --bb) {
        const int cm = Ls - aa - bb;
        if (idx <= at + cm) {
          la = aa;
          lb = bb;
          lc = cm - (idx - at);
          break;
        }
        at += cm + 1;
      }
    }
  }
  double acc = 0.0;
  for (int64_t s0 = m0; s0 < m1; s0 += 64) {
    const int ns = (int)((m1 - s0) < 64 ? (m1 - s0) : 64);
    const double *rows = q + s0 * stride;
    // lane r: member r's degrees Lx | Ly << 8 | Lz << 16 (-1: failed screen 2)
    const int info = s0 == m0 ? info0 : member_info(q, stride, s0, m1, lane);
#pragma unroll 4
    for (int r0 = 0; r0 < ns; r0 += S) {
      const int r = r0 + slice;
      const int inf = __shfl(info, r < ns ? r : 0, 64);
      if (r < ns && inf >= 0) {
        const int lx = inf & 0xff, ly = (inf >> 8) & 0xff, lz = (inf >> 16) & 0xff;
        const double *cx = rows + (int64_t)r * stride + kQHead;
        const double *cy = cx + lx + 1;
        const double *cz = cy + ly + 1;
        if (la >= 0 && la <= lx && lb <= ly && lc <= lz)
          acc += (cx[lx - la] * cy[ly - lb]) * cz[lz - lc];
      }
    }
  }
  for (int d = 32; d >= W; d >>= 1) acc += __shfl_down(acc, d, 64);
  if (lane < M) K[lane] = acc;
}

__global__ __launch_bounds__(kT) void k_group_coeffs(const double *__restrict__ q, int stride,
                                                     const int64_t *__restrict__ first, int64_t ng,
                                                     int64_t n, const double *__restrict__ gh,
                                                     double *__restrict__ pool) {
  const int64_t g = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  if (g >= ng) return;  // whole wave
  const int lane = threadIdx.x & 63;
  const double *h = gh + g * kGStride;
  const int Ls = (int)(((uint64_t)__double_as_longlong(h[4]) >> 32) & 0xff);
  const int M = (int)__double_as_longlong(h[9]);
  double *K = pool + __double_as_longlong(h[8]);
  const int64_t m0 = first[g], m1 = (g + 1 < ng) ? first[g + 1] : n;
  // M <= 165 (total degree 8): at most three passes of the wave
  if (M <= 64)
    group_coeffs_sliced(q, stride, m0, m1, Ls, M, K, lane, member_info(q, stride, m0, m1, lane));
  else if (M <= 128) group_coeffs<2>(q, stride, m0, m1, Ls, M, K, lane);
  else group_coeffs<3>(q, stride, m0, m1, Ls, M, K, lane);
}

// Operations per point summed over the FP64 range [0, ng64) and the rest,
// in a fixed order: block b sums its contiguous slice (part[2 b], part[2 b +
// 1]), then one block adds the slices in order.
constexpr int kFlopBlocks = 1024;
__global__ __launch_bounds__(kT) void k_group_flops(const double *__restrict__ gh, int64_t ng,
                                                    int64_t ng64, double *__restrict__ part) {
  __shared__ double red[2][kT];
  const int64_t per = (ng + kFlopBlocks - 1) / kFlopBlocks;
  const int64_t g0 = (int64_t)blockIdx.x * per, g1 = g0 + per < ng ? g0 + per : ng;
  double s[2] = {0.0, 0.0};
  for (int64_t g = g0 + threadIdx.x; g < g1; g += kT) s[g < ng64 ? 0 : 1] += gh[g * kGStride + 11];
  red[0][threadIdx.x] = s[0];
  red[1][threadIdx.x] = s[1];
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0];
    part[2 * blockIdx.x + 1] = red[1][0];
  }
}

__global__ void k_group_flops_total(const double *__restrict__ part, double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int i = 0; i < kFlopBlocks; ++i) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  out[0] = a;
  out[1] = b;
}

constexpr int kMagBins = 2100;  // ilogb in [-1075, 1024]
constexpr int kMagOff = 1075;

__global__ __launch_bounds__(kT) void k_group_mag_hist(const double *__restrict__ gh, int64_t ng,
                                                       unsigned long long *__restrict__ hist) {
  __shared__ unsigned int h[kMagBins];
  for (int b = threadIdx.x; b < kMagBins; b += kT) h[b] = 0;
  __syncthreads();
  for (int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x; g < ng; g += (int64_t)gridDim.x * kT) {
    const double u = gh[g * kGStride + 5];
    int b = (u > 0.0) ? ilogb(u) + kMagOff : 0;
    b = b < 0 ? 0 : (b >= kMagBins ? kMagBins - 1 : b);
    atomicAdd(&h[b], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kMagBins; b += kT)
    if (h[b]) atomicAdd(&hist[b], (unsigned long long)h[b]);
}

__global__ __launch_bounds__(kT) void k_group_split(const double *__restrict__ gh, int64_t ng,
                                                    double tau, int heavy,
                                                    uint8_t *__restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= ng) return;
  const bool big = gh[g * kGStride + 5] >= tau;
  out[g] = (uint8_t)(heavy ? big : !big);
}

__global__ __launch_bounds__(kT) void k_group_gather(const double *__restrict__ src,
                                                     const int64_t *__restrict__ idx, int64_t n,
                                                     double *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= n * kGStride) return;
  const int64_t r = t / kGStride;
  dst[t] = src[idx[r] * kGStride + (t - r * kGStride)];
}

__global__ __launch_bounds__(kT) void k_to_float(const double *__restrict__ x, int64_t n,
                                                 float *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

}  // namespace

// Device buffer growth keeping the first `used` elements (device to
// device), to at least `need` (1.5x the old capacity when larger).
template <class T>
int grow_dev(inca_ctx *ctx, T *&ptr, size_t &cap, size_t need, size_t used) {
  if (need <= cap && ptr) return 0;
  const size_t n = std::max(need, cap + cap / 2);
  T *p = nullptr;
  hipError_t e = hipMalloc((void **)&p, (n ? n : 1) * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_err(ctx, INCA_ENOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
  }
  if (used && ptr) INCA_HIP(hipMemcpyAsync(p, ptr, used * sizeof(T), hipMemcpyDeviceToDevice,
                                           ctx->stream));
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  if (ptr) (void)hipFree(ptr);
  ptr = p;
  cap = n;
  return 0;
}

bool groupable(const inca_ctx *ctx) {
  // total degree of a group at most 8 (s, p and d primitives): at most 165
  // coefficients, which k_group_coeffs covers in three passes of a wave
  return ctx->grouping && 4 * ctx->lmax_prim <= 8 && ctx->qstride <= kMaxStride;
}

void reset_groups(inca_ctx *ctx) {
  ctx->ng = ctx->ng_fp64 = 0;
  ctx->gpool_n = 0;
  ctx->grho_h = 0.0;  // word 7 of the new headers holds no rho yet
  ctx->gflop_fp64 = ctx->gflop_fp32 = 0.0;
}

// Groups of the prepared rows q[0..n), which arrive sorted by group
// (adjacent, file order within a group; rows marked kQDropped failed screen
// 2 and count for nothing), appended to the context's groups:
// their headers after the existing ones, their coefficients after the
// existing pool.  heads: the rows' group starts when the caller has them
// (null: k_group_heads compares the rows).
int group_rows(inca_ctx *ctx, const double *q, int64_t n, const uint8_t *heads) {
  if (n == 0) return 0;
  hipStream_t st = ctx->stream;
  const int stride = ctx->qstride;
  // temporaries carved from the context's grouping scratch (kept between
  // preparations, so that small molecules pay no allocation)
  size_t scan_bytes = 0;
  {
    int64_t *a = nullptr, *b = nullptr;
    INCA_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, a, b, (int64_t)0, (size_t)n,
                                     rocprim::plus<int64_t>(), st));
  }
  Carve cv;
  const size_t o_head = cv.add<uint8_t>(n), o_first = cv.add<int64_t>(n),
               o_m = cv.add<int64_t>(n), o_off = cv.add<int64_t>(n + 1),
               o_scan = cv.add<char>((int64_t)scan_bytes);
  int rc;
  if ((rc = ensure(ctx, ctx->d_gscr, ctx->gscr_cap, cv.size))) return rc;
  char *base = ctx->d_gscr;
  uint8_t *d_head = Carve::at<uint8_t>(base, o_head);
  int64_t *d_first = Carve::at<int64_t>(base, o_first);
  int64_t *d_m = Carve::at<int64_t>(base, o_m), *d_off = Carve::at<int64_t>(base, o_off);
  char *d_scan = Carve::at<char>(base, o_scan);
  if (!heads) {
    hipLaunchKernelGGL(k_group_heads, dim3(nblk(n)), dim3(kT), 0, st, q, stride, n, d_head);
    INCA_HIP(hipGetLastError());
    heads = d_head;
  }
  int64_t ng = 0;
  if ((rc = compact_flags(ctx, heads, n, d_first, &ng))) return rc;
  const int64_t ng0 = ctx->ng, pool0 = ctx->gpool_n;
  if ((rc = grow_dev(ctx, ctx->d_gh, ctx->gh_cap, (size_t)(ng0 + ng) * kGStride,
                     (size_t)ng0 * kGStride)))
    return rc;
  double *gh = ctx->d_gh + ng0 * kGStride;
  const unsigned gblk = (unsigned)((ng + kT / 64 - 1) / (kT / 64));  // a wave per group
  hipLaunchKernelGGL(k_group_header, dim3(gblk), dim3(kT), 0, st, q, stride, d_first, ng, n, gh,
                     d_m);
  INCA_HIP(hipGetLastError());
  INCA_HIP(rocprim::exclusive_scan(d_scan, scan_bytes, d_m, d_off, pool0, (size_t)ng,
                                   rocprim::plus<int64_t>(), st));
  int64_t last[2] = {0, 0};
  INCA_HIP(hipMemcpyAsync(&last[0], d_off + ng - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  INCA_HIP(hipMemcpyAsync(&last[1], d_m + ng - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  INCA_HIP(hipStreamSynchronize(st));
  const int64_t pool_end = last[0] + last[1];
  hipLaunchKernelGGL(k_group_offsets, dim3(nblk(ng)), dim3(kT), 0, st, d_off, ng, gh);
  INCA_HIP(hipGetLastError());
  if ((rc = grow_dev(ctx, ctx->d_gpool, ctx->gpool_cap, (size_t)pool_end, (size_t)pool0))) return rc;
  hipLaunchKernelGGL(k_group_coeffs, dim3(gblk), dim3(kT), 0, st, q, stride, d_first, ng, n, gh,
                     ctx->d_gpool);
  INCA_HIP(hipGetLastError());
  ctx->ng = ng0 + ng;
  ctx->gpool_n = pool_end;
  return 0;
}

// Stable sort of n (key, value) pairs by the low key_bits bits of the key
// (rocPRIM radix sort); temp of sort_slots_bytes(n) bytes (enough for any
// key_bits).
size_t sort_slots_bytes(int64_t n) {
  size_t bytes = 0;
  uint32_t *k = nullptr, *v = nullptr;
  if (rocprim::radix_sort_pairs(nullptr, bytes, k, k, v, v, (size_t)n, 0u, 32u) != hipSuccess)
    return 0;
  return bytes;
}

int sort_slots(inca_ctx *ctx, void *temp, size_t bytes, uint32_t *kin, uint32_t *kout,
               uint32_t *vin, uint32_t *vout, int64_t n, unsigned key_bits) {
  if (n == 0) return 0;
  INCA_HIP(rocprim::radix_sort_pairs(temp, bytes, kin, kout, vin, vout, (size_t)n, 0u, key_bits,
                                     ctx->stream));
  return 0;
}

// After the last group_rows: the precision split of a mixed preparation
// (the groups by their summed L1 bounds, the way records are split: the
// lightest to FP32 while their bounds stay within mixed_tail * |sum M|; the
// FP64 groups first) and the operation counts of the two ranges.
int finish_groups(inca_ctx *ctx, int precision) {
  hipStream_t st = ctx->stream;
  const int64_t ng = ctx->ng, npool = ctx->gpool_n;
  ctx->ng_fp64 = ng;
  if (ng == 0) return 0;
  Carve cv;
  const size_t o_hist = cv.add<unsigned long long>(kMagBins),
               o_fl2 = cv.add<double>(2 + 2 * kFlopBlocks), o_fl = cv.add<uint8_t>(ng),
               o_idx = cv.add<int64_t>(ng);
  int rc;
  if ((rc = ensure(ctx, ctx->d_gscr, ctx->gscr_cap, cv.size))) return rc;
  char *base = ctx->d_gscr;
  int64_t ng64 = ng;
  if (precision == INCA_MIXED && ctx->mixed_tail > 0.0) {
    unsigned long long *d_h = Carve::at<unsigned long long>(base, o_hist);
    std::vector<unsigned long long> h(kMagBins, 0);
    INCA_HIP(hipMemcpyAsync(d_h, h.data(), kMagBins * sizeof(unsigned long long),
                            hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_group_mag_hist, dim3(1024), dim3(kT), 0, st, ctx->d_gh, ng, d_h);
    INCA_HIP(hipGetLastError());
    INCA_HIP(hipMemcpyAsync(h.data(), d_h, kMagBins * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, st));
    INCA_HIP(hipStreamSynchronize(st));
    std::vector<double> mass(kMagBins);
    double total = 0.0;
    for (int b = 0; b < kMagBins; ++b) {
      mass[b] = h[b] ? (double)h[b] * std::ldexp(2.0, std::min(b - kMagOff, 1000)) : 0.0;
      total += mass[b];
    }
    const double budget = ctx->mixed_tail >= 1.0
                              ? HUGE_VAL
                              : ctx->mixed_tail * std::min(std::fabs(ctx->term_integral), total);
    double cum = 0.0;
    int cut = 0;
    while (cut < kMagBins && cum + mass[cut] <= budget) cum += mass[cut++];
    const double tau = std::ldexp(1.0, cut - kMagOff);
    // stable partition: heavy groups (FP64) first, then the light ones
    if ((rc = ensure(ctx, ctx->d_ghtmp, ctx->ghtmp_cap, (size_t)ng * kGStride))) return rc;
    double *d_new = ctx->d_ghtmp;
    uint8_t *d_fl = Carve::at<uint8_t>(base, o_fl);
    int64_t *d_idx = Carve::at<int64_t>(base, o_idx);
    int64_t nh = 0, nl = 0;
    hipLaunchKernelGGL(k_group_split, dim3(nblk(ng)), dim3(kT), 0, st, ctx->d_gh, ng, tau, 1, d_fl);
    INCA_HIP(hipGetLastError());
    if ((rc = compact_flags(ctx, d_fl, ng, d_idx, &nh))) return rc;
    if (nh > 0)
      hipLaunchKernelGGL(k_group_gather, dim3(nblk(nh * kGStride)), dim3(kT), 0, st, ctx->d_gh,
                         d_idx, nh, d_new);
    hipLaunchKernelGGL(k_group_split, dim3(nblk(ng)), dim3(kT), 0, st, ctx->d_gh, ng, tau, 0, d_fl);
    INCA_HIP(hipGetLastError());
    if ((rc = compact_flags(ctx, d_fl, ng, d_idx, &nl))) return rc;
    if (nl > 0)
      hipLaunchKernelGGL(k_group_gather, dim3(nblk(nl * kGStride)), dim3(kT), 0, st, ctx->d_gh,
                         d_idx, nl, d_new + nh * kGStride);
    INCA_HIP(hipGetLastError());
    INCA_HIP(hipMemcpyAsync(ctx->d_gh, d_new, (size_t)ng * kGStride * sizeof(double),
                            hipMemcpyDeviceToDevice, st));
    ng64 = nh;
    if ((rc = ensure(ctx, ctx->d_gpool32, ctx->gpool32_cap, (size_t)npool))) return rc;
    hipLaunchKernelGGL(k_to_float, dim3(nblk(npool)), dim3(kT), 0, st, ctx->d_gpool, npool,
                       ctx->d_gpool32);
    INCA_HIP(hipGetLastError());
  }
  // operations per point of the FP64 and FP32 ranges (header word 11)
  double *d_fl2 = Carve::at<double>(base, o_fl2);
  hipLaunchKernelGGL(k_group_flops, dim3(kFlopBlocks), dim3(kT), 0, st, ctx->d_gh, ng, ng64,
                     d_fl2 + 2);
  hipLaunchKernelGGL(k_group_flops_total, dim3(1), dim3(64), 0, st, d_fl2 + 2, d_fl2);
  INCA_HIP(hipGetLastError());
  double fl[2] = {0.0, 0.0};
  INCA_HIP(hipMemcpyAsync(fl, d_fl2, sizeof(fl), hipMemcpyDeviceToHost, st));
  INCA_HIP(hipStreamSynchronize(st));
  ctx->ng_fp64 = ng64;
  ctx->gflop_fp64 = fl[0];
  ctx->gflop_fp32 = fl[1];
  return 0;
}


// Group the prepared records (after prepare_quartets keeps them all, sorted
// by group within each pre-pass chunk; a group split between chunks becomes
// two).  Leaves ctx->ng = 0 (the evaluation then runs record by record) when
// there is nothing to group or primitives beyond d.
int build_groups(inca_ctx *ctx, int precision, const uint8_t *heads) {
  reset_groups(ctx);
  if (!groupable(ctx) || ctx->q2 == 0) return 0;
  int rc;
  if ((rc = group_rows(ctx, ctx->d_q, ctx->rows_sparse ? ctx->rows_sparse : ctx->q2, heads)))
    return rc;
  return finish_groups(ctx, precision);
}

}  // namespace inca
