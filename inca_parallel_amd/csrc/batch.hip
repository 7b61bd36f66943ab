// Batched evaluation: many molecules' intracules in one persistent launch.
//
// The reference evaluates one molecule per run (inca_main.f90:76 ->
// calculate_intracule).  A batch takes contexts whose grids and quartets are
// already prepared (inca_prepare) and lays them out in one HBM arena:
//   * each molecule's quartets, stably sorted by angular-momentum class
//     (Lsum, Lz, Ly, Lx; the FP32 range of a mixed preparation after the
//     FP64 one), so that every work item holds quartets of a single class
//     and its cost is known in advance.  One device radix sort over
//     (molecule, precision, class) keys orders the whole batch, and a
//     run-length encoding of the sorted keys gives the class segments;
//   * all grid points back to back, with per-molecule offsets;
//   * for contexts prepared with shell-quartet groups (group.hip), their
//     group headers take the place of records (the class is then the
//     group's total degree) and their coefficients are copied into one
//     pool, the headers' pool offsets moved accordingly.
// Work items are (molecule, block tile of 256 P points, class chunk of at most
// 4096 quartets).  All items of one chunk go to one of 8 queues (chunk mod
// 8), each ordered by the chunks' estimated cost, largest first, and
// expanded on the device from the host's chunk table; a persistent grid of
// workgroups pulls from the queues through atomic counters, block L serving
// the queue of its own XCD (L mod 8, the round-robin dispatch) first and
// the others once it is empty, so a chunk's records are read into one L2 and
// small and large molecules share the device without one launch per
// molecule.  Each item writes its own partial-sum slot and a
// fixed-order reduction follows, so results do not depend on which
// workgroup took which item.  The per-item arithmetic is the single-context
// FP64 kernel's, or its group evaluation's (eval_core.hpp).  Points per lane are autotuned over
// {1, 2, 4, 8}; the launch sequence (counter reset, persistent kernel,
// reduction, per-molecule integrals) is captured once into a hipGraph and
// replayed by every inca_batch_evaluate.  The counter is cleared by a kernel
// node, not a memset node.  When every context had the per-point cutoff on,
// the batch builds each molecule's skip mask (eval_core.hpp) whenever the
// items are rebuilt, outside the graph: grids and quartets are fixed.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>

#include "common.hpp"
#include "eval_core.hpp"

struct inca_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int nmol = 0;
  int P = 4;
  int nblocks = 0;  // persistent grid size
  // per molecule
  std::vector<int64_t> G, g_off, nq, q_off, part_off, nch;
  std::vector<int32_t> stride;
  std::vector<std::vector<std::array<int64_t, 3>>> segs;  // (first, count, class key)
  int64_t Gtot = 0, Qdoubles = 0;
  // molecules evaluated by shell-quartet groups (their headers take the
  // place of records, their coefficients sit in one pool)
  std::vector<int> grouped;
  std::vector<int64_t> pool_off;
  int64_t pool_total = 0;
  double *d_pool = nullptr;
  float *d_pool32 = nullptr;
  // device arena
  double *d_q = nullptr, *d_gx = nullptr, *d_gy = nullptr, *d_gz = nullptr;
  double *d_w = nullptr, *d_wvee = nullptr, *d_ivec = nullptr, *d_res = nullptr;
  double *d_partial = nullptr;
  size_t partial_cap = 0;
  int64_t *d_items = nullptr;
  size_t items_cap = 0;
  int64_t *d_chunks = nullptr;  // chunk table the items are expanded from
  size_t chunks_cap = 0;
  int64_t *d_mol = nullptr;   // per molecule: kMol words (see build_items)
  int32_t *d_pmol = nullptr;  // molecule of every point
  // per-point cutoff (when every context had it on): tile anchors come from
  // each molecule's grid kind, centre map and centres, as in one context
  bool cut = false;
  std::vector<int> mode;
  std::vector<int64_t> c_off, ntile, nword, moff;
  int32_t *d_pmap = nullptr;  // molecule-local 1-based centre of every point
  double *d_cent = nullptr;   // all molecules' centres (3 per centre)
  double *d_tb = nullptr;
  size_t tb_cap = 0;
  unsigned long long *d_mask = nullptr;
  size_t mask_cap = 0;
  unsigned int *d_counter = nullptr;  // one work counter per XCD queue
  int64_t *d_queues = nullptr;        // per XCD queue: first item, item count
  int64_t nitems = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t replays = 0;
  double last_ms = 0.0;
  hipEvent_t ev[2] = {nullptr, nullptr};
};

namespace inca {
namespace {

// int64 words per item: mol, tile, qoff, n, stride, partial, local q, and
// the kind of its quartets: bit 0 the precision (0: FP64, 1: the FP32 range
// of a molecule prepared for INCA_MIXED), bit 1 set for shell-quartet
// groups (headers of kGStride doubles, coefficients in the batch's pool)
constexpr int kItem = 8;
constexpr int kMol = 6;   // int64 words per molecule: g_off, G, part_off, nch, mask off, tiles
constexpr int64_t kMaxChunk = 4096;  // quartets per work item at most

int bfail(inca_batch *b, int code, const std::string &msg) {
  b->err = msg;
  return code;
}

#define B_HIP(call)                                                                   \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return bfail(b, INCA_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

__global__ void k_batch_reset(unsigned int *counter) {
  if (threadIdx.x < kXcds) counter[threadIdx.x] = 0u;
}

template <int P>
__global__ __launch_bounds__(kT) void k_batch_eval(const int64_t *__restrict__ items,
                                                   int64_t nitems, const double *__restrict__ q,
                                                   const double *__restrict__ pool,
                                                   const float *__restrict__ pool32,
                                                   const double *__restrict__ gx,
                                                   const double *__restrict__ gy,
                                                   const double *__restrict__ gz,
                                                   const int64_t *__restrict__ mol,
                                                   const unsigned long long *__restrict__ mask,
                                                   const int64_t *__restrict__ queues,
                                                   double *__restrict__ partial,
                                                   unsigned int *__restrict__ counter) {
  __shared__ int s_next[2];  // queue, item within it (-1: every queue is empty)
  __shared__ double s_exp2[kExpTable];
  stage_exp_table(s_exp2);
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // block L runs on XCD L mod 8 (round-robin dispatch): it serves that XCD's
  // queue first, then the others in turn once it is empty
  int home = (int)(blockIdx.x % kXcds), tried = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      int found = -1;
      unsigned int pos = 0;
      while (tried < kXcds) {
        pos = atomicAdd(&counter[home], 1u);
        if ((int64_t)pos < queues[2 * home + 1]) {
          found = home;
          break;
        }
        home = (home + 1) % kXcds;
        ++tried;
      }
      s_next[0] = found;
      s_next[1] = (int)pos;
    }
    __syncthreads();
    const int qf = __builtin_amdgcn_readfirstlane(s_next[0]);
    const int64_t pos = (int64_t)__builtin_amdgcn_readfirstlane(s_next[1]);
    __syncthreads();  // read by every wave before the next fetch
    if (qf < 0) break;  // uniform over the block: every wave leaves together
    const int64_t it = queues[2 * qf] + pos;
    const int64_t *__restrict__ d = items + it * kItem;
    const int64_t m = d[0], bt = d[1], qo = d[2], n = d[3], pb = d[5], lq = d[6];
    const bool f32 = (d[7] & 1) != 0, grp = (d[7] & 2) != 0;
    const int stride = (int)d[4];
    const int64_t goff = mol[kMol * m], G = mol[kMol * m + 1];
    const int64_t moff = mol[kMol * m + 4], nt = mol[kMol * m + 5];
    const int64_t base = bt * (kT * P) + (int64_t)wv * (64 * P) + lane;
    double x[P], y[P], z[P], acc[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int64_t g = base + (int64_t)k * 64;
      const bool ok = g < G;
      x[k] = ok ? gx[goff + g] : 0.0;
      y[k] = ok ? gy[goff + g] : 0.0;
      z[k] = ok ? gz[goff + g] : 0.0;
      acc[k] = 0.0;
    }
    const double *__restrict__ q0 = q + qo - lq * stride;  // the molecule's first quartet
    if (grp && f32) {
      for_live(mask ? mask + moff : nullptr, nt, bt * (kT / 64) + wv, lq, lq + n, [&](int64_t l) {
        group_mixed<P>(q0 + l * stride, pool32, x, y, z, acc);
      });
    } else if (grp) {
      for_live(mask ? mask + moff : nullptr, nt, bt * (kT / 64) + wv, lq, lq + n, [&](int64_t l) {
        group_fp64<P>(q0 + l * stride, pool, x, y, z, acc, s_exp2);
      });
    } else if (f32) {
      for_live(mask ? mask + moff : nullptr, nt, bt * (kT / 64) + wv, lq, lq + n, [&](int64_t l) {
        const double *__restrict__ rec = q0 + l * stride;
        accumulate_mixed_any<P>(rec, load_header(rec), x, y, z, acc);
      });
    } else {
      for_live(mask ? mask + moff : nullptr, nt, bt * (kT / 64) + wv, lq, lq + n, [&](int64_t l) {
        const double *__restrict__ rec = q0 + l * stride;
        accumulate_fp64<P>(rec, load_header(rec), x, y, z, acc, s_exp2);
      });
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int64_t g = base + (int64_t)k * 64;
      if (g < G) partial[pb + g] = acc[k];
    }
  }
}

// Ivec of every point: its molecule's chunk partials in chunk order.
__global__ __launch_bounds__(kT) void k_batch_reduce(const double *__restrict__ partial,
                                                     const int32_t *__restrict__ pmol,
                                                     const int64_t *__restrict__ mol,
                                                     int64_t Gtot, double *__restrict__ ivec) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= Gtot) return;
  const int m = pmol[g];
  const int64_t goff = mol[kMol * m], G = mol[kMol * m + 1], po = mol[kMol * m + 2],
                nc = mol[kMol * m + 3];
  const int64_t local = g - goff;
  double s = 0.0;
  for (int64_t c = 0; c < nc; ++c) s += partial[po + c * G + local];
  ivec[g] = s;
}

// One block per molecule: sum w I and sum w_vee I in a fixed order
// (mod_intracule.f90:441-457).
__global__ __launch_bounds__(kT) void k_batch_integrate(const double *__restrict__ ivec,
                                                        const double *__restrict__ w,
                                                        const double *__restrict__ wvee,
                                                        const int64_t *__restrict__ mol,
                                                        double *__restrict__ res) {
  __shared__ double r1[kT], r2[kT];
  const int64_t m = blockIdx.x;
  const int64_t goff = mol[kMol * m], G = mol[kMol * m + 1];
  double s1 = 0.0, s2 = 0.0;
  for (int64_t g = threadIdx.x; g < G; g += kT) {
    s1 += w[goff + g] * ivec[goff + g];
    s2 += wvee[goff + g] * ivec[goff + g];
  }
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      r1[threadIdx.x] += r1[threadIdx.x + h];
      r2[threadIdx.x] += r2[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    res[2 * m] = r1[0];
    res[2 * m + 1] = r2[0];
  }
}

// dst[r] = src[perm[r]] for whole records of `stride` doubles
__global__ __launch_bounds__(kT) void k_batch_gather(const double *__restrict__ src,
                                                     const int64_t *__restrict__ perm, int64_t n,
                                                     int stride, double *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= n * stride) return;
  const int64_t r = t / stride;
  const int w = (int)(t - r * stride);
  dst[t] = src[perm[r] * stride + w];
}

// Sort key of each of the n prepared records of molecule `mol` and its
// index: molecule, then the FP32 range of a mixed preparation (records from
// nfp64 on) after the FP64 one, then Lsum, Lz, Ly, Lx (header word 4).
constexpr int kKeyMolShift = 41;
__global__ __launch_bounds__(kT) void k_class_keys(const double *__restrict__ q, int64_t n,
                                                   int stride, int64_t nfp64, uint64_t mol,
                                                   uint64_t *__restrict__ key,
                                                   int64_t *__restrict__ idx) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r >= n) return;
  const uint64_t bits = (uint64_t)__double_as_longlong(q[r * stride + 4]);
  const uint64_t lsum = (bits & 0xff) + ((bits >> 8) & 0xff) + ((bits >> 16) & 0xff);
  const uint64_t f32 = r >= nfp64 ? 1u : 0u;
  key[r] = (mol << kKeyMolShift) | (f32 << 40) | (lsum << 24) | (bits & 0xffffff);
  idx[r] = r;
}

// Work items from the chunk table: row c (kChunkWords words: mol, qoff, n,
// stride, partial, local q, precision, first item, tiles) gives one item per
// block tile, items first .. first + tiles - 1.
constexpr int kChunkWords = 9;
__global__ __launch_bounds__(kT) void k_expand_items(const int64_t *__restrict__ tab,
                                                     int64_t nchunk, int64_t *__restrict__ items) {
  const int64_t c = blockIdx.x;
  if (c >= nchunk) return;
  const int64_t *__restrict__ r = tab + c * kChunkWords;
  const int64_t first = r[7], tiles = r[8];
  for (int64_t t = threadIdx.x; t < tiles; t += kT) {
    int64_t *__restrict__ it = items + (first + t) * kItem;
    it[0] = r[0];
    it[1] = t;
    it[2] = r[1];
    it[3] = r[2];
    it[4] = r[3];
    it[5] = r[4];
    it[6] = r[5];
    it[7] = r[6];
  }
}

__global__ __launch_bounds__(kT) void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t < n) p[t] = v;
}

// Group headers copied into the batch: their pool offsets (word 8) move by
// the molecule's place in the batch's pool.
__global__ __launch_bounds__(kT) void k_rebase_pool(double *__restrict__ gh, int64_t n,
                                                    int64_t by) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g < n)
    gh[g * kGStride + 8] =
        __longlong_as_double((long long)(__double_as_longlong(gh[g * kGStride + 8]) + by));
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

template <class T>
int bensure(inca_batch *b, T *&ptr, size_t &cap, size_t n) {
  if (n <= cap && ptr) return 0;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
  if (hipMalloc((void **)&ptr, n * sizeof(T)) != hipSuccess)
    return bfail(b, INCA_ENOMEM, "batch allocation of " + std::to_string(n * sizeof(T)) + " bytes failed");
  cap = n;
  return 0;
}

// Items for points-per-lane P: chunks of at most qchunk quartets that never
// straddle a class boundary, one item per (chunk, block tile).  Each XCD
// queue takes every 8th chunk; its chunks are ordered by estimated cost
// (quartets x (16 + 2 Lsum) x points of a full tile, FP32 at half), largest
// first, and a chunk's items stay together in tile order.  Only the chunk
// table is built on the host; the items are expanded on the device.
int build_items(inca_batch *b, int P) {
  const int64_t tile_pts = (int64_t)kT * P;
  double work = 0.0;
  for (int m = 0; m < b->nmol; ++m)
    work += (double)b->nq[m] * (double)((b->G[m] + tile_pts - 1) / tile_pts);
  // chunks of at most kMaxChunk quartets (their records stay within an
  // XCD's L2 while the blocks of one chunk run) and at least 16 items per
  // persistent block
  const double target = 16.0 * std::max(1, b->nblocks);
  int64_t qchunk = (int64_t)std::ceil(work / target);
  qchunk = std::max<int64_t>(32, std::min<int64_t>(qchunk, kMaxChunk));
  struct Chunk {
    int64_t m, qo, n, part, lq, f32, tiles;
    double cost;
  };
  std::vector<Chunk> ch;
  std::vector<int64_t> order[kXcds];  // a chunk's items share one XCD queue
  int64_t part = 0, chunk_id = 0;
  for (int m = 0; m < b->nmol; ++m) {
    const int64_t tiles = (b->G[m] + tile_pts - 1) / tile_pts;
    b->part_off[m] = part;
    int64_t c = 0;
    for (const auto &sg : b->segs[m]) {
      const int lsum = (int)((sg[2] >> 24) & 0xffff);
      const int64_t f32 = (sg[2] >> 40) & 1;
      // operations per point: 16 + 2 Lsum for a record; 14 + 2 M for a group
      // of total degree Ls (Lx = Ly = Lz = Ls in its class bits)
      const int ls = lsum / 3;
      const double ops = b->grouped[m] ? 14.0 + (ls + 1) * (ls + 2) * (ls + 3) / 3.0
                                       : 16.0 + 2.0 * lsum;
      for (int64_t s = 0; s < sg[1]; s += qchunk, ++c, ++chunk_id) {
        const int64_t n = std::min(qchunk, sg[1] - s);
        const int64_t qo = b->q_off[m] + (sg[0] + s) * b->stride[m];
        // FP32 terms cost about half of the FP64 ones
        const double cost = (double)n * ops * (double)std::min(tile_pts, b->G[m]) *
                            (f32 ? 0.5 : 1.0);
        order[chunk_id % kXcds].push_back((int64_t)ch.size());
        ch.push_back({m, qo, n, part + c * b->G[m], sg[0] + s, f32 | (b->grouped[m] ? 2 : 0),
                      tiles, cost});
      }
    }
    b->nch[m] = c;
    part += c * b->G[m];
  }
  b->nitems = 0;
  int64_t qinfo[2 * kXcds];
  std::vector<int64_t> tab(std::max<size_t>(1, ch.size()) * kChunkWords);
  size_t row = 0;
  for (int x = 0; x < kXcds; ++x) {
    std::stable_sort(order[x].begin(), order[x].end(),
                     [&](int64_t a, int64_t c) { return ch[a].cost > ch[c].cost; });
    qinfo[2 * x] = b->nitems;
    for (int64_t ci : order[x]) {
      const Chunk &k = ch[ci];
      const int64_t w[kChunkWords] = {k.m,   k.qo,  k.n,         b->stride[k.m], k.part,
                                      k.lq,  k.f32, b->nitems,   k.tiles};
      std::copy(w, w + kChunkWords, tab.begin() + row * kChunkWords);
      ++row;
      b->nitems += k.tiles;
    }
    qinfo[2 * x + 1] = b->nitems - qinfo[2 * x];
  }
  int rc;
  if ((rc = bensure(b, b->d_items, b->items_cap,
                    (size_t)std::max<int64_t>(1, b->nitems) * kItem)) ||
      (rc = bensure(b, b->d_chunks, b->chunks_cap, tab.size())))
    return rc;
  B_HIP(hipMemcpy(b->d_queues, qinfo, sizeof(qinfo), hipMemcpyHostToDevice));
  if ((rc = bensure(b, b->d_partial, b->partial_cap, (size_t)std::max<int64_t>(1, part)))) return rc;
  B_HIP(hipMemcpy(b->d_chunks, tab.data(), tab.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  if (!ch.empty()) {
    hipLaunchKernelGGL(k_expand_items, dim3((unsigned)ch.size()), dim3(kT), 0, b->stream,
                       b->d_chunks, (int64_t)ch.size(), b->d_items);
    B_HIP(hipGetLastError());
  }
  // cutoff masks for this P: wave tiles of 64 P points per molecule
  int64_t tb_total = 0, mask_total = 0;
  for (int m = 0; m < b->nmol; ++m) {
    b->ntile[m] = ((b->G[m] + tile_pts - 1) / tile_pts) * (kT / 64);
    b->nword[m] = (b->nq[m] + 63) / 64;
    b->moff[m] = mask_total;
    tb_total += b->ntile[m];
    mask_total += b->ntile[m] * b->nword[m];
  }
  if (b->cut) {
    if ((rc = bensure(b, b->d_tb, b->tb_cap, (size_t)std::max<int64_t>(1, tb_total) * kTB)) ||
        (rc = bensure(b, b->d_mask, b->mask_cap, (size_t)std::max<int64_t>(1, mask_total))))
      return rc;
    int64_t toff = 0;
    for (int m = 0; m < b->nmol; ++m) {
      const int64_t nt = b->ntile[m];
      if (nt > 0 && b->G[m] > 0) {
        hipLaunchKernelGGL(k_tile_bounds, dim3((unsigned)(nt / (kT / 64))), dim3(kT), 0, b->stream,
                           b->d_gx + b->g_off[m], b->d_gy + b->g_off[m], b->d_gz + b->g_off[m],
                           b->G[m], P, b->mode[m], b->d_pmap + b->g_off[m],
                           b->d_cent + 3 * b->c_off[m], nt, b->d_tb + toff * kTB, nullptr);
        if (b->nword[m] > 0)
          hipLaunchKernelGGL(k_skip_mask, dim3((unsigned)((b->nword[m] + 3) / 4)), dim3(kT), 0,
                             b->stream, b->d_q + b->q_off[m], b->stride[m], b->nq[m],
                             b->d_tb + toff * kTB, nt, b->nword[m], b->d_mask + b->moff[m],
                             nullptr);
      }
      toff += nt;
    }
    B_HIP(hipGetLastError());
    B_HIP(hipStreamSynchronize(b->stream));
  }
  std::vector<int64_t> mi(kMol * (size_t)b->nmol);
  for (int m = 0; m < b->nmol; ++m) {
    mi[kMol * m] = b->g_off[m];
    mi[kMol * m + 1] = b->G[m];
    mi[kMol * m + 2] = b->part_off[m];
    mi[kMol * m + 3] = b->nch[m];
    mi[kMol * m + 4] = b->moff[m];
    mi[kMol * m + 5] = b->ntile[m];
  }
  B_HIP(hipMemcpy(b->d_mol, mi.data(), mi.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  return 0;
}

// the evaluation sequence on b->stream (captured into the graph)
int enqueue(inca_batch *b, int P) {
  hipStream_t st = b->stream;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, b->d_counter);
  if (b->nitems > 0) {
#define INCA_BEVAL(PP)                                                                        \
  hipLaunchKernelGGL(k_batch_eval<PP>, dim3(b->nblocks), dim3(kT), 0, st, b->d_items,         \
                     b->nitems, b->d_q, b->d_pool, b->d_pool32, b->d_gx, b->d_gy, b->d_gz,     \
                     b->d_mol,                                                                 \
                     b->cut ? b->d_mask : nullptr, b->d_queues, b->d_partial, b->d_counter)
    switch (P) {
      case 1: INCA_BEVAL(1); break;
      case 2: INCA_BEVAL(2); break;
      case 8: INCA_BEVAL(8); break;
      default: INCA_BEVAL(4); break;
    }
#undef INCA_BEVAL
  }
  if (b->Gtot > 0)
    hipLaunchKernelGGL(k_batch_reduce, dim3(nblk(b->Gtot)), dim3(kT), 0, st, b->d_partial,
                       b->d_pmol, b->d_mol, b->Gtot, b->d_ivec);
  hipLaunchKernelGGL(k_batch_integrate, dim3(b->nmol), dim3(kT), 0, st, b->d_ivec, b->d_w,
                     b->d_wvee, b->d_mol, b->d_res);
  B_HIP(hipGetLastError());
  return 0;
}

int persistent_blocks(inca_batch *b, int P) {
  int dev = b->device, ncu = 0, per = 0;
  B_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  hipError_t e = hipSuccess;
  switch (P) {
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_eval<1>, kT, 0); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_eval<2>, kT, 0); break;
    case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_eval<8>, kT, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_eval<4>, kT, 0); break;
  }
  B_HIP(e);
  b->nblocks = std::max(kXcds, ncu * std::max(1, per) / kXcds * kXcds);
  return 0;
}

int capture(inca_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  B_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue(b, b->P);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  B_HIP(e);
  b->graph = g;
  B_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return 0;
}

// Device buffers freed on every return (hipFree waits for the device).
struct DeviceTemps {
  std::vector<void *> p;
  ~DeviceTemps() {
    for (void *q : p)
      if (q) (void)hipFree(q);
  }
  template <class T>
  hipError_t get(T *&ptr, size_t n) {
    ptr = nullptr;
    hipError_t e = hipMalloc((void **)&ptr, std::max<size_t>(1, n) * sizeof(T));
    if (e == hipSuccess) p.push_back(ptr);
    return e;
  }
};

int setup(inca_batch *b, int nmol, inca_ctx *const *mols) {
  b->nmol = nmol;
  b->G.assign(nmol, 0);
  b->g_off.assign(nmol, 0);
  b->nq.assign(nmol, 0);
  b->q_off.assign(nmol, 0);
  b->part_off.assign(nmol, 0);
  b->nch.assign(nmol, 0);
  b->stride.assign(nmol, 0);
  b->segs.assign(nmol, {});
  b->mode.assign(nmol, 2);
  b->c_off.assign(nmol, 0);
  b->ntile.assign(nmol, 0);
  b->nword.assign(nmol, 0);
  b->moff.assign(nmol, 0);
  b->grouped.assign(nmol, 0);
  b->pool_off.assign(nmol, 0);
  b->pool_total = 0;
  b->cut = true;
  int64_t ncent = 0;
  for (int m = 0; m < nmol; ++m) {
    const inca_ctx *c = mols[m];
    if (!c) return bfail(b, INCA_EINVAL, "molecule " + std::to_string(m) + ": null context");
    // shell-quartet groups where the context has them, else records (a
    // mixed preparation's record split is made now)
    b->grouped[m] = c->grouping && c->ng > 0;
    if (!b->grouped[m] && c->device == b->device &&
        (need_records(mols[m]) || split_records(mols[m])))
      return bfail(b, INCA_EINVAL, "molecule " + std::to_string(m) + ": " + c->err);
    if (c->device != b->device)
      return bfail(b, INCA_EINVAL, "molecule " + std::to_string(m) + " is on another device");
    if (c->grid_kind == GRID_NONE)
      return bfail(b, INCA_EINVAL, "molecule " + std::to_string(m) + ": no grid");
    if (c->qstride == 0)
      return bfail(b, INCA_EINVAL, "molecule " + std::to_string(m) + ": quartets not prepared");
    b->G[m] = c->G;
    b->g_off[m] = b->Gtot;
    b->Gtot += c->G;
    b->nq[m] = b->grouped[m] ? c->ng : c->q2;
    b->stride[m] = b->grouped[m] ? kGStride : c->qstride;
    b->q_off[m] = b->Qdoubles;
    b->Qdoubles += b->nq[m] * b->stride[m];
    if (b->grouped[m]) {
      b->pool_off[m] = b->pool_total;
      b->pool_total += c->gpool_n;
    }
    b->cut = b->cut && c->cut_log2 > 0.0;
    b->mode[m] = c->grid_kind == GRID_RADIAL ? 0 : (c->grid_kind == GRID_MULTI ? 1 : 2);
    b->c_off[m] = ncent;
    if (b->mode[m] == 1) ncent += c->nquad;
  }
  const size_t gt = (size_t)std::max<int64_t>(1, b->Gtot);
  const size_t qd = (size_t)std::max<int64_t>(1, b->Qdoubles);
  size_t cap = 0;
  int rc;
  double **garr[6] = {&b->d_gx, &b->d_gy, &b->d_gz, &b->d_w, &b->d_wvee, &b->d_ivec};
  for (double **a : garr)
    if ((cap = 0, rc = bensure(b, *a, cap, gt))) return rc;
  if ((cap = 0, rc = bensure(b, b->d_q, cap, qd))) return rc;
  if (b->pool_total > 0) {
    bool any32 = false;
    for (int m = 0; m < nmol; ++m)
      any32 = any32 || (b->grouped[m] && mols[m]->prep_precision == INCA_MIXED);
    if ((cap = 0, rc = bensure(b, b->d_pool, cap, (size_t)b->pool_total))) return rc;
    if (any32 && (cap = 0, rc = bensure(b, b->d_pool32, cap, (size_t)b->pool_total))) return rc;
  }
  if ((cap = 0, rc = bensure(b, b->d_res, cap, 2 * (size_t)nmol))) return rc;
  if ((cap = 0, rc = bensure(b, b->d_mol, cap, kMol * (size_t)nmol))) return rc;
  if ((cap = 0, rc = bensure(b, b->d_pmol, cap, gt))) return rc;
  if ((cap = 0, rc = bensure(b, b->d_counter, cap, kXcds))) return rc;
  if ((cap = 0, rc = bensure(b, b->d_queues, cap, 2 * kXcds))) return rc;
  if ((cap = 0, rc = bensure(b, b->d_pmap, cap, gt))) return rc;
  if ((cap = 0, rc = bensure(b, b->d_cent, cap, 3 * (size_t)std::max<int64_t>(1, ncent)))) return rc;
  hipStream_t st = b->stream;
  DeviceTemps tmp;  // temporaries of the class sort
  int64_t Qtot = 0;
  std::vector<int64_t> qstart(nmol, 0);
  for (int m = 0; m < nmol; ++m) {
    qstart[m] = Qtot;
    Qtot += b->nq[m];
    B_HIP(hipStreamSynchronize(mols[m]->stream));  // the context's own work is done
  }
  if (Qtot > (int64_t)UINT32_MAX)
    return bfail(b, INCA_EINVAL, "a batch holds at most 2^32 - 1 prepared quartets");
  for (int m = 0; m < nmol; ++m) {
    const inca_ctx *c = mols[m];
    const int64_t G = c->G;
    if (G > 0) {
      const double *src[5] = {c->d_gx, c->d_gy, c->d_gz, c->d_w, c->d_wvee};
      double *dst[5] = {b->d_gx, b->d_gy, b->d_gz, b->d_w, b->d_wvee};
      for (int a = 0; a < 5; ++a)
        B_HIP(hipMemcpyAsync(dst[a] + b->g_off[m], src[a], G * sizeof(double),
                             hipMemcpyDeviceToDevice, st));
      hipLaunchKernelGGL(k_fill_i32, dim3(nblk(G)), dim3(kT), 0, st, b->d_pmol + b->g_off[m], G,
                         (int32_t)m);
      B_HIP(hipGetLastError());
      if (b->mode[m] == 1) {
        B_HIP(hipMemcpyAsync(b->d_pmap + b->g_off[m], c->d_map, G * sizeof(int32_t),
                             hipMemcpyDeviceToDevice, st));
        B_HIP(hipMemcpyAsync(b->d_cent + 3 * b->c_off[m], c->d_gcent, 3 * c->nquad * sizeof(double),
                             hipMemcpyDeviceToDevice, st));
      }
    }
  }
  if (Qtot > 0) {
    // Every prepared record's (molecule, precision, class) key, one stable
    // radix sort over the batch, and the runs of equal keys: each molecule's
    // records in class order, the FP32 range of a mixed preparation after
    // the FP64 one, with the file order kept within a class.
    uint64_t *d_kin, *d_kout, *d_uniq;
    int64_t *d_vin, *d_vout;
    unsigned int *d_cnt, *d_nruns;
    B_HIP(tmp.get(d_kin, Qtot));
    B_HIP(tmp.get(d_kout, Qtot));
    B_HIP(tmp.get(d_vin, Qtot));
    B_HIP(tmp.get(d_vout, Qtot));
    B_HIP(tmp.get(d_uniq, Qtot));
    B_HIP(tmp.get(d_cnt, Qtot));
    B_HIP(tmp.get(d_nruns, 1));
    for (int m = 0; m < nmol; ++m) {
      const inca_ctx *c = mols[m];
      const int64_t n = b->nq[m];
      if (n == 0) continue;
      const int64_t nfp64 = c->prep_precision != INCA_MIXED ? n
                            : (b->grouped[m] ? c->ng_fp64 : c->q_fp64);
      hipLaunchKernelGGL(k_class_keys, dim3(nblk(n)), dim3(kT), 0, st,
                         b->grouped[m] ? c->d_gh : c->d_q, n, b->stride[m], nfp64, (uint64_t)m,
                         d_kin + qstart[m], d_vin + qstart[m]);
    }
    B_HIP(hipGetLastError());
    unsigned int end_bit = kKeyMolShift + 1;
    while (end_bit < 64 && ((uint64_t)(nmol - 1) >> (end_bit - kKeyMolShift)) != 0) ++end_bit;
    size_t sort_bytes = 0, rle_bytes = 0;
    B_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, d_kin, d_kout, d_vin, d_vout,
                                    (size_t)Qtot, 0u, end_bit, st));
    B_HIP(rocprim::run_length_encode(nullptr, rle_bytes, d_kout, (unsigned int)Qtot, d_uniq,
                                     d_cnt, d_nruns, st));
    char *d_temp;
    B_HIP(tmp.get(d_temp, std::max(sort_bytes, rle_bytes)));
    B_HIP(rocprim::radix_sort_pairs(d_temp, sort_bytes, d_kin, d_kout, d_vin, d_vout,
                                    (size_t)Qtot, 0u, end_bit, st));
    B_HIP(rocprim::run_length_encode(d_temp, rle_bytes, d_kout, (unsigned int)Qtot, d_uniq,
                                     d_cnt, d_nruns, st));
    for (int m = 0; m < nmol; ++m) {
      const inca_ctx *c = mols[m];
      const int64_t n = b->nq[m];
      if (n == 0) continue;
      hipLaunchKernelGGL(k_batch_gather, dim3(nblk(n * b->stride[m])), dim3(kT), 0, st,
                         b->grouped[m] ? c->d_gh : c->d_q, d_vout + qstart[m], n, b->stride[m],
                         b->d_q + b->q_off[m]);
      if (b->grouped[m]) {
        hipLaunchKernelGGL(k_rebase_pool, dim3(nblk(n)), dim3(kT), 0, st, b->d_q + b->q_off[m], n,
                           b->pool_off[m]);
        B_HIP(hipMemcpyAsync(b->d_pool + b->pool_off[m], c->d_gpool,
                             (size_t)c->gpool_n * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (c->prep_precision == INCA_MIXED && c->d_gpool32)
          B_HIP(hipMemcpyAsync(b->d_pool32 + b->pool_off[m], c->d_gpool32,
                               (size_t)c->gpool_n * sizeof(float), hipMemcpyDeviceToDevice, st));
      }
    }
    B_HIP(hipGetLastError());
    unsigned int nruns = 0;
    B_HIP(hipMemcpyAsync(&nruns, d_nruns, sizeof(nruns), hipMemcpyDeviceToHost, st));
    B_HIP(hipStreamSynchronize(st));
    std::vector<uint64_t> uniq(nruns);
    std::vector<unsigned int> cnt(nruns);
    B_HIP(hipMemcpyAsync(uniq.data(), d_uniq, nruns * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    B_HIP(hipMemcpyAsync(cnt.data(), d_cnt, nruns * sizeof(unsigned int), hipMemcpyDeviceToHost,
                         st));
    B_HIP(hipStreamSynchronize(st));
    std::vector<int64_t> at(nmol, 0);  // next local record of each molecule
    for (unsigned int r = 0; r < nruns; ++r) {
      const int m = (int)(uniq[r] >> kKeyMolShift);
      const int64_t key = (int64_t)(uniq[r] & ((1ull << kKeyMolShift) - 1));
      if (m < 0 || m >= nmol || at[m] + (int64_t)cnt[r] > b->nq[m])
        return bfail(b, INCA_EHIP, "class sort: inconsistent runs");
      b->segs[m].push_back({at[m], (int64_t)cnt[r], key});
      at[m] += cnt[r];
    }
    for (int m = 0; m < nmol; ++m)
      if (at[m] != b->nq[m]) return bfail(b, INCA_EHIP, "class sort: inconsistent runs");
  }
  B_HIP(hipStreamSynchronize(st));
  if ((rc = persistent_blocks(b, b->P))) return rc;
  if ((rc = build_items(b, b->P))) return rc;
  return capture(b);
}

}  // namespace
}  // namespace inca

using namespace inca;

extern "C" {

int inca_batch_create(int32_t device, int32_t nmol, inca_ctx *const *mols, inca_batch **out) {
  if (!out) return INCA_EINVAL;
  *out = nullptr;
  if (nmol < 1 || !mols) return INCA_EINVAL;
  inca_batch *b = new inca_batch();
  b->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreate(&b->stream);
  if (e == hipSuccess) e = hipEventCreate(&b->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&b->ev[1]);
  int rc = (e == hipSuccess) ? setup(b, nmol, mols)
                             : bfail(b, INCA_EHIP, std::string("batch set-up: ") + hipGetErrorString(e));
  *out = b;  // returned even on failure so that the message can be read
  return rc;
}

int inca_batch_destroy(inca_batch *b) {
  if (!b) return 0;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  void *bufs[] = {b->d_q,     b->d_pool,  b->d_pool32, b->d_gx,  b->d_gy,   b->d_gz,    b->d_w,
                  b->d_wvee,  b->d_ivec,  b->d_res,  b->d_partial, b->d_items,
                  b->d_mol,   b->d_pmol,  b->d_counter, b->d_pmap, b->d_cent,
                  b->d_tb,    b->d_mask,  b->d_queues, b->d_chunks};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  for (auto &ev : b->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

const char *inca_batch_last_error(const inca_batch *b) { return b ? b->err.c_str() : ""; }

int inca_batch_evaluate(inca_batch *b, double *integral, double *vee) {
  if (!b || !b->exec) return INCA_EINVAL;
  B_HIP(hipSetDevice(b->device));
  B_HIP(hipEventRecord(b->ev[0], b->stream));
  B_HIP(hipGraphLaunch(b->exec, b->stream));
  B_HIP(hipEventRecord(b->ev[1], b->stream));
  std::vector<double> res(2 * (size_t)b->nmol);
  B_HIP(hipMemcpyAsync(res.data(), b->d_res, res.size() * sizeof(double), hipMemcpyDeviceToHost,
                       b->stream));
  B_HIP(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  B_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  b->last_ms = ms;
  ++b->replays;
  for (int m = 0; m < b->nmol; ++m) {
    if (integral) integral[m] = res[2 * m];
    if (vee) vee[m] = res[2 * m + 1];
  }
  return 0;
}

int inca_batch_get_ivec(inca_batch *b, int32_t mol, double *ivec) {
  if (!b || !ivec) return INCA_EINVAL;
  if (mol < 0 || mol >= b->nmol) return bfail(b, INCA_EINVAL, "molecule index out of range");
  if (b->G[mol] == 0) return 0;
  B_HIP(hipSetDevice(b->device));
  B_HIP(hipMemcpy(ivec, b->d_ivec + b->g_off[mol], b->G[mol] * sizeof(double),
                  hipMemcpyDeviceToHost));
  return 0;
}

int inca_batch_autotune(inca_batch *b, int32_t *best) {
  if (!b) return INCA_EINVAL;
  B_HIP(hipSetDevice(b->device));
  const int cand[4] = {1, 2, 4, 8};
  float best_ms = 1e30f;
  int best_p = b->P, rc;
  for (int c = 0; c < 4; ++c) {
    const int P = cand[c];
    if ((rc = persistent_blocks(b, P)) || (rc = build_items(b, P))) return rc;
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up
      B_HIP(hipEventRecord(b->ev[0], b->stream));
      if ((rc = enqueue(b, P))) return rc;
      B_HIP(hipEventRecord(b->ev[1], b->stream));
      B_HIP(hipEventSynchronize(b->ev[1]));
      B_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    }
    if (ms < best_ms) {
      best_ms = ms;
      best_p = P;
    }
  }
  b->P = best_p;
  if ((rc = persistent_blocks(b, b->P)) || (rc = build_items(b, b->P)) || (rc = capture(b)))
    return rc;
  if (best) *best = best_p;
  return 0;
}

int inca_batch_set_points(inca_batch *b, int32_t points_per_lane) {
  if (!b) return INCA_EINVAL;
  const int P = points_per_lane;
  if (P != 1 && P != 2 && P != 4 && P != 8)
    return bfail(b, INCA_EINVAL, "points per lane " + std::to_string(P) + ": 1, 2, 4 or 8");
  B_HIP(hipSetDevice(b->device));
  b->P = P;
  int rc;
  if ((rc = persistent_blocks(b, b->P)) || (rc = build_items(b, b->P)) || (rc = capture(b)))
    return rc;
  return 0;
}

int inca_batch_info(inca_batch *b, int64_t *nitems, int32_t *points_per_lane, int32_t *blocks,
                    int64_t *replays, double *last_ms) {
  if (!b) return INCA_EINVAL;
  if (nitems) *nitems = b->nitems;
  if (points_per_lane) *points_per_lane = b->P;
  if (blocks) *blocks = b->nblocks;
  if (replays) *replays = b->replays;
  if (last_ms) *last_ms = b->last_ms;
  return 0;
}

}  // extern "C"
