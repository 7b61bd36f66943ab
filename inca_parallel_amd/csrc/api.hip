// extern "C" entry points of libinca_hip.so (declared in include/inca_hip.h).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.hpp"

using namespace inca;

namespace {
std::string g_create_err;

#define CHECK_CTX(c) \
  do {               \
    if (!(c)) return INCA_EINVAL; \
  } while (0)

// normalisation of a primitive (mod_intracule.f90:148-150)
double normprim(double alpha, int t, int m, int n) {
  auto dfact = [](int a) {  // mod_numbers.f90:14-27
    int d = 1;
    if (a % 2 == 0)
      for (int i = 1; i <= a / 2; ++i) d *= 2 * i;
    else
      for (int i = 1; i <= (a + 1) / 2; ++i) d *= 2 * i - 1;
    return d;
  };
  const double pi = 4.0 * std::atan(1.0);
  double df = (double)(dfact(2 * t - 1) * dfact(2 * m - 1) * dfact(2 * n - 1));
  return std::pow(2.0 * alpha / pi, 0.75) *
         std::sqrt(std::pow(4.0 * alpha, (double)(t + m + n)) * (1.0 / df));
}
}  // namespace

extern "C" {

int inca_create(int32_t device, inca_ctx **out) {
  if (!out) return INCA_EINVAL;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    g_create_err = std::string("no HIP device available: ") + hipGetErrorString(e);
    return INCA_EHIP;
  }
  if (device < 0 || device >= n) {
    g_create_err = "device index out of range";
    return INCA_EINVAL;
  }
  if ((e = hipSetDevice(device)) != hipSuccess) {
    g_create_err = std::string("hipSetDevice failed: ") + hipGetErrorString(e);
    return INCA_EHIP;
  }
  if ((e = noalloc_scratch_init(device)) != hipSuccess) {
    g_create_err = std::string("address of the no-allocation scratch: ") + hipGetErrorString(e);
    return INCA_EHIP;
  }
  inca_ctx *c = new inca_ctx();
  c->device = device;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc((void **)&c->d_cnt, 4 * sizeof(unsigned long long))) != hipSuccess) {
    g_create_err = std::string("context set-up failed: ") + hipGetErrorString(e);
    delete c;
    return INCA_EHIP;
  }
  for (auto &ev : c->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess) {
      g_create_err = "hipEventCreate failed";
      delete c;
      return INCA_EHIP;
    }
  *out = c;
  return 0;
}

int inca_destroy(inca_ctx *c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void *bufs[] = {c->d_prim, c->d_tmn, c->d_pair, c->d_gx, c->d_gy, c->d_gz, c->d_w,
                  c->d_wvee, c->d_map, c->d_rec, c->d_q, c->d_qtmp, c->d_qf, c->d_flag,
                  c->d_idx, c->d_bc, c->d_cnt, c->d_partial, c->d_ivec, c->d_tb,
                  c->d_gcent, c->d_mask, c->d_evals, c->d_wcount, c->d_ridx, c->d_dup,
                  c->d_red, c->d_tmp, c->d_cf, c->d_gh, c->d_gpool, c->d_gpool32,
                  c->d_wflop, c->d_flops, c->d_gscr, c->d_ghtmp};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  for (auto &ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (char *h : c->h_pin)
    if (h) (void)hipHostFree(h);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

const char *inca_last_error(const inca_ctx *c) {
  return c ? c->err.c_str() : g_create_err.c_str();
}

int inca_set_primitives(inca_ctx *ctx, int32_t nprim, const double *alpha, const int32_t *tmn,
                        const double *x, const double *y, const double *z, int32_t normalize) {
  CHECK_CTX(ctx);
  if (nprim < 1 || !alpha || !tmn || !x || !y || !z)
    return set_err(ctx, INCA_EINVAL, "inca_set_primitives: bad arguments");
  INCA_HIP(hipSetDevice(ctx->device));
  int tmax = 0;
  for (int64_t i = 0; i < 3LL * nprim; ++i) {
    if (tmn[i] < 0) return set_err(ctx, INCA_EINVAL, "negative angular exponent");
    if (tmn[i] > tmax) tmax = tmn[i];
  }
  if (4 * tmax > kLmaxAxis)
    return set_err(ctx, INCA_ERANGE,
                   "per-axis angular momentum 4*%d exceeds the Gauss-Hermite tables (L <= 21)",
                   tmax);
  for (int i = 0; i < nprim; ++i)
    if (!(alpha[i] > 0.0)) return set_err(ctx, INCA_EINVAL, "non-positive exponent %d", i + 1);
  std::vector<double> prim(5 * (size_t)nprim);
  for (int i = 0; i < nprim; ++i) {
    prim[i] = alpha[i];
    prim[nprim + i] = x[i];
    prim[2 * nprim + i] = y[i];
    prim[3 * nprim + i] = z[i];
    prim[4 * nprim + i] = normprim(alpha[i], tmn[i], tmn[nprim + i], tmn[2 * nprim + i]);
  }
  if (ctx->d_prim) (void)hipFree(ctx->d_prim);
  if (ctx->d_tmn) (void)hipFree(ctx->d_tmn);
  ctx->d_prim = nullptr;
  ctx->d_tmn = nullptr;
  INCA_HIP(hipMalloc((void **)&ctx->d_prim, prim.size() * sizeof(double)));
  INCA_HIP(hipMalloc((void **)&ctx->d_tmn, 3 * (size_t)nprim * sizeof(int32_t)));
  INCA_HIP(hipMemcpyAsync(ctx->d_prim, prim.data(), prim.size() * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
  INCA_HIP(hipMemcpyAsync(ctx->d_tmn, tmn, 3 * (size_t)nprim * sizeof(int32_t),
                          hipMemcpyHostToDevice, ctx->stream));
  ctx->nprim = nprim;
  ctx->normalize = normalize ? 1 : 0;
  ctx->lmax_axis = 4 * tmax;
  ctx->lmax_prim = 0;
  for (int i = 0; i < nprim; ++i)
    ctx->lmax_prim = std::max(ctx->lmax_prim, tmn[i] + tmn[nprim + i] + tmn[2 * nprim + i]);
  int rc = build_pair_table(ctx);
  if (rc) return rc;
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int inca_grid_multicenter(inca_ctx *ctx, int32_t nquad, const double *cent,
                          const int32_t *nradc, const int32_t *nangc, const double *sfalpha,
                          const double *Ps, int32_t nosym, int64_t *maxgrid, int64_t *rgrid,
                          int64_t *rrgrid) {
  CHECK_CTX(ctx);
  if (!cent || !nradc || !nangc || !sfalpha || !Ps)
    return set_err(ctx, INCA_EINVAL, "inca_grid_multicenter: null argument");
  INCA_HIP(hipSetDevice(ctx->device));
  int64_t a = 0, b = 0, c = 0;
  int rc = grid_multicenter(ctx, nquad, cent, nradc, nangc, sfalpha, Ps, nosym, &a, &b, &c);
  if (maxgrid) *maxgrid = a;
  if (rgrid) *rgrid = b;
  if (rrgrid) *rrgrid = c;
  return rc;
}

int inca_grid_radial_plot(inca_ctx *ctx, int32_t nblock, const double *tart, const double *step,
                          const int32_t *nang, int32_t nosym, int64_t *nradi, int64_t *ngrid) {
  CHECK_CTX(ctx);
  if (!tart || !step || !nang) return set_err(ctx, INCA_EINVAL, "null argument");
  INCA_HIP(hipSetDevice(ctx->device));
  int64_t a = 0, b = 0;
  int rc = grid_radial_plot(ctx, nblock, tart, step, nang, nosym, &a, &b);
  if (nradi) *nradi = a;
  if (ngrid) *ngrid = b;
  return rc;
}

int inca_grid_cube(inca_ctx *ctx, const double *center, const double *step, const int32_t *np,
                   int64_t *ngrid) {
  CHECK_CTX(ctx);
  if (!center || !step || !np) return set_err(ctx, INCA_EINVAL, "null argument");
  INCA_HIP(hipSetDevice(ctx->device));
  int64_t n = 0;
  int rc = grid_cube(ctx, center, step, np, &n);
  if (ngrid) *ngrid = n;
  return rc;
}

int inca_grid_points(inca_ctx *ctx, int64_t n, const double *r) {
  CHECK_CTX(ctx);
  if (n > 0 && !r) return set_err(ctx, INCA_EINVAL, "null points");
  INCA_HIP(hipSetDevice(ctx->device));
  return grid_points(ctx, n, r);
}

int inca_get_grid(inca_ctx *ctx, double *r, double *w, double *w_vee, int32_t *center_map) {
  CHECK_CTX(ctx);
  if (ctx->grid_kind == GRID_NONE) return set_err(ctx, INCA_EINVAL, "no grid");
  const int64_t G = ctx->G;
  if (G == 0) return 0;
  hipStream_t st = ctx->stream;
  if (r) {
    std::vector<double> x(G), y(G), z(G);
    INCA_HIP(hipMemcpyAsync(x.data(), ctx->d_gx, G * sizeof(double), hipMemcpyDeviceToHost, st));
    INCA_HIP(hipMemcpyAsync(y.data(), ctx->d_gy, G * sizeof(double), hipMemcpyDeviceToHost, st));
    INCA_HIP(hipMemcpyAsync(z.data(), ctx->d_gz, G * sizeof(double), hipMemcpyDeviceToHost, st));
    INCA_HIP(hipStreamSynchronize(st));
    for (int64_t g = 0; g < G; ++g) {
      r[3 * g] = x[g];
      r[3 * g + 1] = y[g];
      r[3 * g + 2] = z[g];
    }
  }
  if (w) INCA_HIP(hipMemcpyAsync(w, ctx->d_w, G * sizeof(double), hipMemcpyDeviceToHost, st));
  if (w_vee)
    INCA_HIP(hipMemcpyAsync(w_vee, ctx->d_wvee, G * sizeof(double), hipMemcpyDeviceToHost, st));
  if (center_map)
    INCA_HIP(hipMemcpyAsync(center_map, ctx->d_map, G * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  INCA_HIP(hipStreamSynchronize(st));
  return 0;
}

int inca_get_radial_layout(inca_ctx *ctx, double *radi, int32_t *smn) {
  CHECK_CTX(ctx);
  if (ctx->grid_kind != GRID_RADIAL) return set_err(ctx, INCA_EINVAL, "no radial-scan grid");
  for (size_t i = 0; i < ctx->radi.size(); ++i) {
    if (radi) radi[i] = ctx->radi[i];
    if (smn) smn[i] = ctx->smn[i];
  }
  return 0;
}

int inca_load_dm2p(inca_ctx *ctx, const char *path, int64_t *nrec, int32_t *format) {
  CHECK_CTX(ctx);
  if (!path) return set_err(ctx, INCA_EINVAL, "null path");
  INCA_HIP(hipSetDevice(ctx->device));
  FILE *f = std::fopen(path, "rb");
  if (!f) return set_err(ctx, INCA_EIO, "cannot open %s", path);
  std::fseek(f, 0, SEEK_END);
  const long long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  int32_t head[2] = {0, 0};
  if (size >= 8) {
    if (std::fread(head, 4, 2, f) != 2) {
      std::fclose(f);
      return set_err(ctx, INCA_EIO, "cannot read %s", path);
    }
    std::fseek(f, 0, SEEK_SET);
  }
  // mod_intracule.f90:167-184
  int fmt = 0;
  if (head[0] == 24 && head[1] == 0) fmt = 3;
  else if (head[0] == 24) fmt = 1;
  else if (head[0] == 0 && size >= 8) fmt = 2;
  if (fmt == 3 || fmt == 2) {
    std::fclose(f);
    return set_err(ctx, INCA_EFORMAT,
                   "Unsupported file format with markers. Please convert to stream format.");
  }
  const int64_t cap = fmt == 1 ? size / 32 : size / 24;
  if (ensure(ctx, ctx->d_rec, ctx->rec_cap, (size_t)(cap > 0 ? cap : 1) * 24)) {
    std::fclose(f);
    return INCA_ENOMEM;
  }
  // Double-buffered: chunk c is read (and, with markers, repacked) into
  // pinned buffer c % 2 while the copy of chunk c - 1 to the device runs on
  // the context's stream; a buffer is refilled once its last copy is done.
  const size_t chunk_rec = 1 << 21;  // 2M records = 48 MiB per staging copy
  const size_t pin = chunk_rec * 24;
  if (ctx->pin_bytes < pin) {
    for (char *&h : ctx->h_pin) {
      if (h) (void)hipHostFree(h);
      h = nullptr;
    }
    ctx->pin_bytes = 0;
    for (char *&h : ctx->h_pin)
      if (hipHostMalloc((void **)&h, pin, hipHostMallocDefault) != hipSuccess) {
        std::fclose(f);
        (void)hipGetLastError();
        return set_err(ctx, INCA_ENOMEM, "hipHostMalloc of %zu bytes failed", pin);
      }
    ctx->pin_bytes = pin;
  }
  std::vector<char> buf(fmt == 1 ? chunk_rec * 32 : 0);
  hipEvent_t done[2] = {nullptr, nullptr};
  for (auto &ev : done)
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      std::fclose(f);
      for (auto &e2 : done)
        if (e2) (void)hipEventDestroy(e2);
      return set_err(ctx, INCA_EHIP, "hipEventCreate failed");
    }
  bool used[2] = {false, false};
  int64_t got = 0;
  bool stop = false;
  hipError_t e = hipSuccess;
  for (int c = 0; !stop && got < cap && e == hipSuccess; ++c) {
    char *stage = ctx->h_pin[c & 1];
    if (used[c & 1] && (e = hipEventSynchronize(done[c & 1])) != hipSuccess) break;
    size_t want = (size_t)std::min<int64_t>(chunk_rec, cap - got);
    size_t keep = 0;
    if (fmt == 1) {
      const size_t n = std::fread(buf.data(), 32, want, f);
      if (n < want) stop = true;
      for (size_t r = 0; r < n; ++r) {
        int32_t m0, m1;
        std::memcpy(&m0, buf.data() + 32 * r, 4);
        std::memcpy(&m1, buf.data() + 32 * r + 28, 4);
        if (m0 != 24 || m1 != 24) {  // a record that is not 24 bytes ends the read
          stop = true;
          break;
        }
        std::memcpy(stage + 24 * keep, buf.data() + 32 * r + 4, 24);
        ++keep;
      }
    } else {
      keep = std::fread(stage, 24, want, f);
      if (keep < want) stop = true;
    }
    if (keep) {
      e = hipMemcpyAsync((char *)ctx->d_rec + 24 * got, stage, 24 * keep, hipMemcpyHostToDevice,
                         ctx->stream);
      if (e == hipSuccess) e = hipEventRecord(done[c & 1], ctx->stream);
      used[c & 1] = true;
    }
    got += (int64_t)keep;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  for (auto &ev : done) (void)hipEventDestroy(ev);
  if (e != hipSuccess) {
    std::fclose(f);
    return hip_fail(ctx, e, "copy of .dm2p records", __LINE__);
  }
  std::fclose(f);
  ctx->nrec = got;
  if (nrec) *nrec = got;
  if (format) *format = fmt;
  return 0;
}

int inca_set_quartets(inca_ctx *ctx, int64_t n, const void *records, int32_t on_device) {
  CHECK_CTX(ctx);
  if (n < 0 || (n > 0 && !records)) return set_err(ctx, INCA_EINVAL, "bad record array");
  INCA_HIP(hipSetDevice(ctx->device));
  if (int rc = ensure(ctx, ctx->d_rec, ctx->rec_cap, (size_t)(n > 0 ? n : 1) * 24))
    return rc;
  // Device records may still be in flight on the producer's stream (the
  // context's own stream does not order itself after the default stream):
  // wait for all work queued on the device before copying them.
  if (n > 0 && on_device) INCA_HIP(hipDeviceSynchronize());
  if (n > 0)
    INCA_HIP(hipMemcpyAsync(ctx->d_rec, records, 24 * (size_t)n,
                            on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                            ctx->stream));
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  ctx->nrec = n;
  return 0;
}

int inca_evaluate(inca_ctx *ctx, int32_t precision, double thresh, double *ivec,
                  int64_t *counts) {
  CHECK_CTX(ctx);
  if (ctx->nprim == 0) return set_err(ctx, INCA_EINVAL, "primitives not set");
  if (ctx->grid_kind == GRID_NONE) return set_err(ctx, INCA_EINVAL, "grid not set");
  if (!ctx->d_rec) return set_err(ctx, INCA_EINVAL, "DM2 records not set");
  if (precision != INCA_FP64 && precision != INCA_MIXED)
    return set_err(ctx, INCA_EINVAL, "unknown precision %d", precision);
  INCA_HIP(hipSetDevice(ctx->device));
  int64_t cnt[4] = {0, 0, 0, 0};
  INCA_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
  int rc = prepare_quartets(ctx, thresh, precision, cnt);
  if (rc) return rc;
  if ((rc = evaluate_grid(ctx, precision))) return rc;
  INCA_HIP(hipEventSynchronize(ctx->ev[3]));
  float a = 0, b = 0, c = 0;
  INCA_HIP(hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]));
  INCA_HIP(hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]));
  INCA_HIP(hipEventElapsedTime(&c, ctx->ev[2], ctx->ev[3]));
  ctx->t_ms[0] = a;
  ctx->t_ms[1] = b;
  ctx->t_ms[2] = c;
  if (counts)
    for (int i = 0; i < 4; ++i) counts[i] = cnt[i];
  if (ivec && ctx->G > 0)
    INCA_HIP(hipMemcpy(ivec, ctx->d_ivec, ctx->G * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int inca_prepare(inca_ctx *ctx, int32_t precision, double thresh, int64_t *counts) {
  CHECK_CTX(ctx);
  if (ctx->nprim == 0) return set_err(ctx, INCA_EINVAL, "primitives not set");
  if (!ctx->d_rec) return set_err(ctx, INCA_EINVAL, "DM2 records not set");
  if (precision != INCA_FP64 && precision != INCA_MIXED)
    return set_err(ctx, INCA_EINVAL, "unknown precision %d", precision);
  INCA_HIP(hipSetDevice(ctx->device));
  int64_t cnt[4] = {0, 0, 0, 0};
  int rc = prepare_quartets(ctx, thresh, precision, cnt);
  if (rc) return rc;
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_ivec = 0;
  if (counts)
    for (int i = 0; i < 4; ++i) counts[i] = cnt[i];
  return 0;
}

int inca_evaluate_prepared(inca_ctx *ctx, int32_t precision, double *ivec) {
  CHECK_CTX(ctx);
  if (ctx->grid_kind == GRID_NONE) return set_err(ctx, INCA_EINVAL, "grid not set");
  if (ctx->qstride == 0) return set_err(ctx, INCA_EINVAL, "quartets not prepared");
  if (precision != INCA_FP64 && precision != INCA_MIXED)
    return set_err(ctx, INCA_EINVAL, "unknown precision %d", precision);
  if (precision == INCA_MIXED && ctx->prep_precision != INCA_MIXED)
    return set_err(ctx, INCA_EINVAL, "quartets were prepared for FP64: prepare with INCA_MIXED");
  INCA_HIP(hipSetDevice(ctx->device));
  INCA_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
  int rc = evaluate_grid(ctx, precision);
  if (rc) return rc;
  INCA_HIP(hipEventSynchronize(ctx->ev[3]));
  float b = 0, c = 0;
  INCA_HIP(hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]));
  INCA_HIP(hipEventElapsedTime(&c, ctx->ev[2], ctx->ev[3]));
  ctx->t_ms[0] = 0.0;
  ctx->t_ms[1] = b;
  ctx->t_ms[2] = c;
  if (ivec && ctx->G > 0)
    INCA_HIP(hipMemcpy(ivec, ctx->d_ivec, ctx->G * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int inca_evaluate_prepared_noalloc(inca_ctx *ctx, int32_t precision, double *ivec,
                                   void *stream) {
  CHECK_CTX(ctx);
  if (ctx->grid_kind == GRID_NONE) return set_err(ctx, INCA_EINVAL, "grid not set");
  if (ctx->qstride == 0) return set_err(ctx, INCA_EINVAL, "quartets not prepared");
  if (precision != INCA_FP64 && precision != INCA_MIXED)
    return set_err(ctx, INCA_EINVAL, "unknown precision %d", precision);
  if (precision == INCA_MIXED && ctx->prep_precision != INCA_MIXED)
    return set_err(ctx, INCA_EINVAL, "quartets were prepared for FP64: prepare with INCA_MIXED");
  if (!ivec && ctx->G > 0) return set_err(ctx, INCA_EINVAL, "null Ivec");
  INCA_HIP(hipSetDevice(ctx->device));
  return evaluate_noalloc(ctx, precision, ivec,
                          stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

int inca_autogrid(int32_t natoms, const int32_t *an, const double *cartes, int32_t nohydro,
                  int32_t betaone, const double *radius, int32_t cap, int32_t *nquad,
                  double *cent, double *Ps, double *sfalpha) {
  if (natoms < 1 || !an || !cartes || !nquad) return INCA_EINVAL;
  std::vector<double> c, p, a;
  const int n = autogrid(natoms, an, cartes, nohydro, betaone, radius, c, p, a);
  *nquad = n;
  if (cap < n) return INCA_ERANGE;
  if (!cent || !Ps || !sfalpha) return INCA_EINVAL;
  std::memcpy(cent, c.data(), 3 * (size_t)n * sizeof(double));
  std::memcpy(Ps, p.data(), (size_t)n * sizeof(double));
  std::memcpy(sfalpha, a.data(), (size_t)n * sizeof(double));
  return 0;
}

int inca_onee(inca_ctx *ctx, int32_t kind, int32_t index, int32_t nmo, const double *T,
              const double *w, double *values) {
  CHECK_CTX(ctx);
  if (ctx->nprim < 1) return set_err(ctx, INCA_EINVAL, "primitives not set");
  if (ctx->grid_kind == GRID_NONE) return set_err(ctx, INCA_EINVAL, "grid not set");
  if (!values && ctx->G > 0) return set_err(ctx, INCA_EINVAL, "null output");
  if (kind == 1) {
    if (index < 1 || index > ctx->nprim)
      return set_err(ctx, INCA_EINVAL, "primitive %d outside 1..%d", index, ctx->nprim);
  } else if (kind == 3 || kind == 5 || kind == 6) {
    if (nmo < 1 || !T) return set_err(ctx, INCA_EINVAL, "orbital coefficients missing");
    if (kind == 3 && (index < 1 || index > nmo))
      return set_err(ctx, INCA_EINVAL, "orbital %d outside 1..%d", index, nmo);
    if (kind != 3 && !w) return set_err(ctx, INCA_EINVAL, "orbital weights missing");
  } else {
    return set_err(ctx, INCA_EINVAL,
                   "one-electron function kind %d: 1 primitive, 3 orbital, 5 density, "
                   "6 Laplacian", kind);
  }
  INCA_HIP(hipSetDevice(ctx->device));
  return onee(ctx, kind, index, nmo, T, w, values);
}

int inca_bragg_slater(int32_t n, const int32_t *an, double *radius) {
  if (n < 0 || (n > 0 && (!an || !radius))) return INCA_EINVAL;
  for (int32_t i = 0; i < n; ++i) radius[i] = bragg_slater_bohr(an[i]);
  return 0;
}

int inca_release_scratch(inca_ctx *ctx) {
  CHECK_CTX(ctx);
  INCA_HIP(hipSetDevice(ctx->device));
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  auto drop = [](auto *&p, size_t &cap) {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  };
  for (char *&h : ctx->h_pin) {
    if (h) (void)hipHostFree(h);
    h = nullptr;
  }
  ctx->pin_bytes = 0;
  drop(ctx->d_rec, ctx->rec_cap);
  ctx->nrec = 0;
  drop(ctx->d_qtmp, ctx->qtmp_cap);
  drop(ctx->d_flag, ctx->flag_cap);
  drop(ctx->d_idx, ctx->idx_cap);
  drop(ctx->d_tmp, ctx->tmp_cap);
  drop(ctx->d_partial, ctx->partial_cap);
  drop(ctx->d_red, ctx->red_cap);
  drop(ctx->d_mask, ctx->mask_cap);
  drop(ctx->d_wcount, ctx->wcount_cap);
  drop(ctx->d_cf, ctx->cf_cap);
  drop(ctx->d_qf, ctx->qf_cap);
  drop(ctx->d_gscr, ctx->gscr_cap);
  drop(ctx->d_ghtmp, ctx->ghtmp_cap);
  drop(ctx->d_wflop, ctx->wflop_cap);
  return 0;
}

int inca_set_keep_records(inca_ctx *ctx, int32_t keep) {
  CHECK_CTX(ctx);
  if (keep != 0 && keep != 1) return set_err(ctx, INCA_EINVAL, "keep_records must be 0 or 1");
  ctx->keep_records = keep;
  return 0;
}

int inca_release_records(inca_ctx *ctx) {
  CHECK_CTX(ctx);
  if (ctx->ng == 0 || !ctx->grouping)
    return set_err(ctx, INCA_EINVAL, "no shell-quartet groups: the prepared records are needed");
  INCA_HIP(hipSetDevice(ctx->device));
  INCA_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->d_q) (void)hipFree(ctx->d_q);
  ctx->d_q = nullptr;
  ctx->q_cap = 0;
  ctx->rows_sparse = 0;
  ctx->rec_split_pending = 0;
  return 0;
}

int inca_set_prepare_chunk(inca_ctx *ctx, int64_t records) {
  CHECK_CTX(ctx);
  if (records < 1024) return set_err(ctx, INCA_EINVAL, "pre-pass chunk below 1024 records");
  ctx->prep_chunk = records;
  return 0;
}

int inca_set_mixed_tail(inca_ctx *ctx, double tail) {
  CHECK_CTX(ctx);
  if (!(tail >= 0.0 && tail <= 1.0)) return set_err(ctx, INCA_EINVAL, "tail must be in [0,1]");
  ctx->mixed_tail = tail;
  return 0;
}

int inca_get_mixed_split(inca_ctx *ctx, int64_t *n_fp64, int64_t *n_fp32, double *tau) {
  CHECK_CTX(ctx);
  INCA_HIP(hipSetDevice(ctx->device));
  if (int rc = split_records(ctx)) return rc;
  if (n_fp64) *n_fp64 = ctx->q_fp64;
  if (n_fp32) *n_fp32 = ctx->q2 - ctx->q_fp64;
  if (tau) *tau = ctx->mixed_tau;
  return 0;
}

int inca_set_point_cutoff(inca_ctx *ctx, double log2_rel) {
  CHECK_CTX(ctx);
  if (!(log2_rel == 0.0 || (log2_rel >= 64.0 && log2_rel <= 1000.0)))
    return set_err(ctx, INCA_EINVAL, "point cutoff must be 0 (off) or in [64, 1000]");
  ctx->cut_log2 = log2_rel;
  return 0;
}

int inca_set_cube_contraction(inca_ctx *ctx, int32_t on) {
  CHECK_CTX(ctx);
  if (on < 0 || on > 2) return set_err(ctx, INCA_EINVAL, "cube contraction must be 0, 1 or 2");
  ctx->cube_contract = on;
  return 0;
}

int inca_get_eval_counts(inca_ctx *ctx, int64_t *reference, int64_t *executed) {
  CHECK_CTX(ctx);
  if (reference) *reference = ctx->evals_reference;
  if (executed) *executed = ctx->evals_executed;
  return 0;
}

int inca_set_grouping(inca_ctx *ctx, int32_t on) {
  CHECK_CTX(ctx);
  if (on != 0 && on != 1) return set_err(ctx, INCA_EINVAL, "grouping must be 0 or 1");
  ctx->grouping = on;
  return 0;
}

int inca_get_groups(inca_ctx *ctx, int64_t *ngroups, int64_t *ngroups_fp64,
                    int64_t *ncoefficients) {
  CHECK_CTX(ctx);
  if (ngroups) *ngroups = ctx->ng;
  if (ngroups_fp64) *ngroups_fp64 = ctx->ng_fp64;
  if (ncoefficients) *ncoefficients = ctx->gpool_n;
  return 0;
}

int inca_get_eval_flops(inca_ctx *ctx, double *flops) {
  CHECK_CTX(ctx);
  if (!flops) return set_err(ctx, INCA_EINVAL, "null output");
  *flops = ctx->grouped_last ? ctx->flops_executed : 0.0;
  return 0;
}

int inca_get_analytic_integral(inca_ctx *ctx, double *total) {
  CHECK_CTX(ctx);
  if (!total) return set_err(ctx, INCA_EINVAL, "null output");
  *total = ctx->term_integral;
  return 0;
}

int inca_set_polynomials(inca_ctx *ctx, int32_t mode) {
  CHECK_CTX(ctx);
  if (mode != 0 && mode != 1) return set_err(ctx, INCA_EINVAL, "polynomial mode must be 0 or 1");
  ctx->poly_exact = mode;
  return 0;
}

int inca_set_tiles(inca_ctx *ctx, int32_t p64, int32_t p32) {
  CHECK_CTX(ctx);
  auto ok = [](int p) { return p == 1 || p == 2 || p == 4 || p == 6 || p == 8 || p == 10; };
  if (!ok(p64) || !(ok(p32) || p32 == 0))
    return set_err(ctx, INCA_EINVAL,
                   "points per lane must be 1, 2, 4, 6, 8 or 10 (or 0 for the FP32 range on ray tiles)");
  ctx->tile_p[0] = p64;
  ctx->tile_p[1] = p32;
  // the group kernels take 1, 2, 4, 6 or 8 (10 -> 8; 0 keeps their choice)
  ctx->gtile_p[0] = p64 == 10 ? 8 : p64;
  if (p32 != 0) ctx->gtile_p[1] = p32 == 10 ? 8 : p32;
  return 0;
}

int inca_autotune(inca_ctx *ctx, int32_t *best) {
  CHECK_CTX(ctx);
  INCA_HIP(hipSetDevice(ctx->device));
  return autotune(ctx, best);
}

int inca_get_prepared(inca_ctx *ctx, int64_t first, int64_t n, double *out, int32_t *stride) {
  CHECK_CTX(ctx);
  if (stride) *stride = ctx->qstride;
  if (first < 0 || n < 0 || first + n > ctx->q2)
    return set_err(ctx, INCA_EINVAL, "prepared-record range out of bounds");
  if (n > 0 && out) {
    INCA_HIP(hipSetDevice(ctx->device));
    if (int rc = need_records(ctx)) return rc;
    if (int rc = split_records(ctx)) return rc;  // the FP64 range first
  }
  if (n > 0 && out)
    INCA_HIP(hipMemcpy(out, ctx->d_q + first * ctx->qstride, n * ctx->qstride * sizeof(double),
                       hipMemcpyDeviceToHost));
  return 0;
}

int inca_get_ivec(inca_ctx *ctx, double *ivec) {
  CHECK_CTX(ctx);
  if (!ctx->have_ivec) return set_err(ctx, INCA_EINVAL, "no Ivec: call inca_evaluate first");
  if (ctx->G > 0)
    INCA_HIP(hipMemcpy(ivec, ctx->d_ivec, ctx->G * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int inca_integrate(inca_ctx *ctx, double *integral, double *vee, double *ipc, double *vpc) {
  CHECK_CTX(ctx);
  INCA_HIP(hipSetDevice(ctx->device));
  return grid_integrate(ctx, integral, vee, ipc, vpc);
}

int inca_radial_values(inca_ctx *ctx, double *rintra) {
  CHECK_CTX(ctx);
  INCA_HIP(hipSetDevice(ctx->device));
  return grid_radial_values(ctx, rintra);
}

int inca_smearing(inca_ctx *ctx, double u_max, int32_t n_bins, double sigma, double *u,
                  double *P, double *Iavg) {
  CHECK_CTX(ctx);
  if (!u || !P || !Iavg) return set_err(ctx, INCA_EINVAL, "null output");
  INCA_HIP(hipSetDevice(ctx->device));
  return grid_smearing(ctx, u_max, n_bins, sigma, u, P, Iavg);
}

int inca_class_histogram(inca_ctx *ctx, int32_t nbins, int64_t *hist) {
  CHECK_CTX(ctx);
  if (nbins < 1 || !hist) return set_err(ctx, INCA_EINVAL, "bad histogram arguments");
  INCA_HIP(hipSetDevice(ctx->device));
  return class_histogram(ctx, nbins, hist);
}

int inca_get_timings(inca_ctx *ctx, double *t) {
  CHECK_CTX(ctx);
  for (int i = 0; i < 3; ++i) t[i] = ctx->t_ms[i];
  return 0;
}

int inca_lebedev(int32_t n, double *x, double *y, double *z, double *w) {
  if (!x || !y || !z || !w) return INCA_EINVAL;
  return lebedev(n, x, y, z, w) ? INCA_ERANGE : 0;
}

int inca_gauleg(int32_t n, double *t, double *w) {
  if (n < 1 || !t || !w) return INCA_EINVAL;
  gauleg(n, t, w);
  return 0;
}

}  // extern "C"
