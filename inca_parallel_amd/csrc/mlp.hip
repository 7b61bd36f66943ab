// Learned autogrid: a per-atom multilayer perceptron whose outputs set the
// radial scale and the radial/angular orders of the multicentre grid.
//
// No reference counterpart.  The reference scales each centre's radial map
// by Bragg-Slater radii (mod_autogrid.f90:36, 61-63, get_bs_radius
// :142-185) and broadcasts one nrad/nang to every centre (:127-131).  Here
// the network predicts per atom
//   out[0]     log of a factor on the Bragg-Slater radius,
//   out[1..4]  logits over nrad in {50, 75, 100, 150},
//   out[5..8]  logits over nang in {110, 194, 302, 590},
// and the centres keep the reference's formulas (pair radius = mean of the
// two atoms' radii, merged centres population-weighted).
//
// Network: 16 -> 64 -> 64 -> 64 -> 9, GELU (exact, erf form) after the
// three hidden layers, FP32.  Parameters are one flat array: W1 (16x64,
// row-major [in][out]), b1, W2 (64x64), b2, W3 (64x64), b3, W4 (64x9), b4.
//
// Kernels (all deterministic, no atomics):
//   k_mlp_forward   one thread per atom; the whole parameter array (40 KB)
//                   is staged in LDS once per block and read as broadcasts;
//                   pre-activations and activations are kept for backward.
//   k_mlp_loss      squared error on out[0] + softmax cross-entropy on the
//                   two class heads, averaged over atoms; d loss / d out.
//   k_mlp_backward  one thread per atom: the per-atom deltas of all layers.
//   k_mlp_wgrad     one thread per parameter: sum over atoms in atom order.
//   k_mlp_adam      elementwise Adam with bias correction.
// The sizes are tiny (B x 64 x 64 with B a few thousand atoms), so plain
// FP32 FMAs from LDS-resident weights are used rather than matrix cores.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace inca {
namespace {

constexpr int kIn = INCA_MLP_IN, kH = 64, kOut = INCA_MLP_OUT;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kH;
constexpr int kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kH;
constexpr int kOffW4 = kOffB3 + kH, kOffB4 = kOffW4 + kH * kOut;
constexpr int kNParam = kOffB4 + kOut;
static_assert(kNParam == INCA_MLP_NPARAM, "parameter count");
constexpr int kTM = 128;  // threads per block

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_d(float z) {
  // d/dz [z Phi(z)] = Phi(z) + z phi(z)
  const float Phi = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float phi = 0.39894228040143268f * expf(-0.5f * z * z);
  return Phi + z * phi;
}

__device__ void stage_params(const float *__restrict__ p, float *lds) {
  for (int i = threadIdx.x; i < kNParam; i += blockDim.x) lds[i] = p[i];
  __syncthreads();
}

// z = W^T a + b for one atom (W row-major [nin][nout]); a and z in registers
template <int NIN, int NOUT>
__device__ __forceinline__ void dense(const float *W, const float *b, const float (&a)[NIN],
                                      float (&z)[NOUT]) {
#pragma unroll
  for (int o = 0; o < NOUT; ++o) z[o] = b[o];
#pragma unroll 4
  for (int i = 0; i < NIN; ++i) {
    const float ai = a[i];
#pragma unroll
    for (int o = 0; o < NOUT; ++o) z[o] = fmaf(ai, W[i * NOUT + o], z[o]);
  }
}

// act: per atom [z1 | a1 | z2 | a2 | z3 | a3] (6 x 64 floats)
__global__ __launch_bounds__(kTM) void k_mlp_forward(const float *__restrict__ params, int n,
                                                     const float *__restrict__ x,
                                                     float *__restrict__ out,
                                                     float *__restrict__ act) {
  extern __shared__ float sp[];
  stage_params(params, sp);
  const int a = blockIdx.x * kTM + threadIdx.x;
  if (a >= n) return;
  float in[kIn], z[kH], h[kH];
#pragma unroll
  for (int i = 0; i < kIn; ++i) in[i] = x[(size_t)a * kIn + i];
  float *ac = act ? act + (size_t)a * 6 * kH : nullptr;
  dense<kIn, kH>(sp + kOffW1, sp + kOffB1, in, z);
#pragma unroll
  for (int o = 0; o < kH; ++o) {
    h[o] = gelu(z[o]);
    if (ac) {
      ac[o] = z[o];
      ac[kH + o] = h[o];
    }
  }
  dense<kH, kH>(sp + kOffW2, sp + kOffB2, h, z);
#pragma unroll
  for (int o = 0; o < kH; ++o) {
    h[o] = gelu(z[o]);
    if (ac) {
      ac[2 * kH + o] = z[o];
      ac[3 * kH + o] = h[o];
    }
  }
  dense<kH, kH>(sp + kOffW3, sp + kOffB3, h, z);
#pragma unroll
  for (int o = 0; o < kH; ++o) {
    h[o] = gelu(z[o]);
    if (ac) {
      ac[4 * kH + o] = z[o];
      ac[5 * kH + o] = h[o];
    }
  }
  float y[kOut];
  dense<kH, kOut>(sp + kOffW4, sp + kOffB4, h, y);
#pragma unroll
  for (int o = 0; o < kOut; ++o) out[(size_t)a * kOut + o] = y[o];
}

// target per atom: [log-scale, nrad class, nang class]; per-block loss sums
__global__ __launch_bounds__(kTM) void k_mlp_loss(int n, const float *__restrict__ out,
                                                  const float *__restrict__ target,
                                                  float *__restrict__ dout,
                                                  double *__restrict__ block_loss) {
  __shared__ double red[kTM];
  const int a = blockIdx.x * kTM + threadIdx.x;
  double l = 0.0;
  if (a < n) {
    const float *y = out + (size_t)a * kOut;
    float *d = dout + (size_t)a * kOut;
    const float inv = 1.0f / (float)n;
    const float e = y[0] - target[(size_t)a * 3];
    l += (double)e * e;
    d[0] = 2.0f * e * inv;
    for (int head = 0; head < 2; ++head) {
      const int o0 = 1 + 4 * head;
      const int cls = (int)target[(size_t)a * 3 + 1 + head];
      float m = y[o0];
      for (int k = 1; k < 4; ++k) m = fmaxf(m, y[o0 + k]);
      float s = 0.0f;
      for (int k = 0; k < 4; ++k) s += expf(y[o0 + k] - m);
      const float lse = m + logf(s);
      l += (double)(lse - y[o0 + cls]);
      for (int k = 0; k < 4; ++k) d[o0 + k] = (expf(y[o0 + k] - lse) - (k == cls ? 1.0f : 0.0f)) * inv;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int h = kTM / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) block_loss[blockIdx.x] = red[0];
}

// delta per atom: [d1 | d2 | d3 | d4] with d_l = dL/dz_l (d4 = dL/dout)
__global__ __launch_bounds__(kTM) void k_mlp_backward(const float *__restrict__ params, int n,
                                                      const float *__restrict__ act,
                                                      const float *__restrict__ dout,
                                                      float *__restrict__ delta) {
  extern __shared__ float sp[];
  stage_params(params, sp);
  const int a = blockIdx.x * kTM + threadIdx.x;
  if (a >= n) return;
  const float *ac = act + (size_t)a * 6 * kH;
  float *de = delta + (size_t)a * (3 * kH + kOut);
  float d4[kOut], dh[kH], dz[kH];
#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    d4[o] = dout[(size_t)a * kOut + o];
    de[3 * kH + o] = d4[o];
  }
  // layer 3: dh3 = W4 d4
#pragma unroll 4
  for (int i = 0; i < kH; ++i) {
    float s = 0.0f;
#pragma unroll
    for (int o = 0; o < kOut; ++o) s = fmaf(sp[kOffW4 + i * kOut + o], d4[o], s);
    dz[i] = s * gelu_d(ac[4 * kH + i]);
    de[2 * kH + i] = dz[i];
  }
  // layer 2: dh2 = W3 dz3
#pragma unroll 4
  for (int i = 0; i < kH; ++i) {
    float s = 0.0f;
#pragma unroll
    for (int o = 0; o < kH; ++o) s = fmaf(sp[kOffW3 + i * kH + o], dz[o], s);
    dh[i] = s * gelu_d(ac[2 * kH + i]);
  }
#pragma unroll
  for (int i = 0; i < kH; ++i) de[kH + i] = dh[i];
  // layer 1: dh1 = W2 dz2
#pragma unroll 4
  for (int i = 0; i < kH; ++i) {
    float s = 0.0f;
#pragma unroll
    for (int o = 0; o < kH; ++o) s = fmaf(sp[kOffW2 + i * kH + o], dh[o], s);
    de[i] = s * gelu_d(ac[i]);
  }
}

// gradient of every parameter: sum over atoms (in atom order) of
// (input activation) x (delta) for weights, delta for biases
__global__ __launch_bounds__(kTM) void k_mlp_wgrad(int n, const float *__restrict__ x,
                                                   const float *__restrict__ act,
                                                   const float *__restrict__ delta,
                                                   float *__restrict__ grad) {
  const int p = blockIdx.x * kTM + threadIdx.x;
  if (p >= kNParam) return;
  const int dstride = 3 * kH + kOut;
  int layer, in_i = -1, out_o;
  int nout;
  if (p < kOffB1) { layer = 0; nout = kH; in_i = (p - kOffW1) / nout; out_o = (p - kOffW1) % nout; }
  else if (p < kOffW2) { layer = 0; nout = kH; out_o = p - kOffB1; }
  else if (p < kOffB2) { layer = 1; nout = kH; in_i = (p - kOffW2) / nout; out_o = (p - kOffW2) % nout; }
  else if (p < kOffW3) { layer = 1; nout = kH; out_o = p - kOffB2; }
  else if (p < kOffB3) { layer = 2; nout = kH; in_i = (p - kOffW3) / nout; out_o = (p - kOffW3) % nout; }
  else if (p < kOffW4) { layer = 2; nout = kH; out_o = p - kOffB3; }
  else if (p < kOffB4) { layer = 3; nout = kOut; in_i = (p - kOffW4) / nout; out_o = (p - kOffW4) % nout; }
  else { layer = 3; nout = kOut; out_o = p - kOffB4; }
  double s = 0.0;
  for (int a = 0; a < n; ++a) {
    const float d = delta[(size_t)a * dstride + layer * kH + out_o];
    float inp = 1.0f;
    if (in_i >= 0) {
      if (layer == 0)
        inp = x[(size_t)a * kIn + in_i];
      else  // activation of the previous hidden layer
        inp = act[(size_t)a * 6 * kH + (2 * (layer - 1) + 1) * kH + in_i];
    }
    s += (double)inp * (double)d;
  }
  grad[p] = (float)s;
}

__global__ __launch_bounds__(kTM) void k_mlp_adam(float *__restrict__ p,
                                                  const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v,
                                                  float lr, float c1, float c2) {
  const int i = blockIdx.x * kTM + threadIdx.x;
  if (i >= kNParam) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  m[i] = b1 * m[i] + (1.0f - b1) * g[i];
  v[i] = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
  p[i] -= lr * (m[i] * c1) / (sqrtf(v[i] * c2) + eps);
}

inline unsigned nb(int n) { return (unsigned)((n + kTM - 1) / kTM); }

}  // namespace
}  // namespace inca

struct inca_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
  float *d_x = nullptr, *d_out = nullptr, *d_act = nullptr, *d_dout = nullptr, *d_delta = nullptr;
  float *d_t = nullptr;
  double *d_bl = nullptr;
  int cap = 0;  // atoms the buffers hold
  int steps = 0;
};

using namespace inca;

namespace {

int mfail(inca_mlp *m, int code, const std::string &msg) {
  m->err = msg;
  return code;
}

#define M_HIP(call)                                                                   \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return mfail(m, INCA_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int reserve(inca_mlp *m, int n) {
  if (n <= m->cap) return 0;
  float **fb[] = {&m->d_x, &m->d_out, &m->d_act, &m->d_dout, &m->d_delta, &m->d_t};
  for (float **b : fb) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  if (m->d_bl) (void)hipFree(m->d_bl);
  m->d_bl = nullptr;
  const size_t sz[] = {(size_t)n * kIn, (size_t)n * kOut, (size_t)n * 6 * kH, (size_t)n * kOut,
                       (size_t)n * (3 * kH + kOut), (size_t)n * 3};
  for (int i = 0; i < 6; ++i)
    if (hipMalloc((void **)fb[i], sz[i] * sizeof(float)) != hipSuccess)
      return mfail(m, INCA_ENOMEM, "MLP buffers");
  if (hipMalloc((void **)&m->d_bl, nb(n) * sizeof(double)) != hipSuccess)
    return mfail(m, INCA_ENOMEM, "MLP buffers");
  m->cap = n;
  return 0;
}

int forward(inca_mlp *m, int n, const float *x, bool keep) {
  int rc;
  if ((rc = reserve(m, n))) return rc;
  M_HIP(hipMemcpyAsync(m->d_x, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice,
                       m->stream));
  hipLaunchKernelGGL(k_mlp_forward, dim3(nb(n)), dim3(kTM), kNParam * sizeof(float), m->stream,
                     m->d_p, n, m->d_x, m->d_out, keep ? m->d_act : nullptr);
  M_HIP(hipGetLastError());
  return 0;
}

// deterministic initial weights: uniform(-1,1) * sqrt(6 / (fan_in + fan_out)), zero biases
void init_params(uint64_t seed, std::vector<float> &p) {
  p.assign(kNParam, 0.0f);
  uint64_t s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
  auto uni = [&]() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  struct L { int off, nin, nout; } ls[4] = {{kOffW1, kIn, kH}, {kOffW2, kH, kH}, {kOffW3, kH, kH},
                                            {kOffW4, kH, kOut}};
  for (const L &l : ls) {
    const double a = std::sqrt(6.0 / (l.nin + l.nout));
    for (int i = 0; i < l.nin * l.nout; ++i) p[l.off + i] = (float)(a * uni());
  }
}

}  // namespace

extern "C" {

int inca_mlp_create(int32_t device, uint64_t seed, inca_mlp **out) {
  if (!out) return INCA_EINVAL;
  inca_mlp *m = new inca_mlp();
  *out = m;
  m->device = device;
  M_HIP(hipSetDevice(device));
  M_HIP(hipStreamCreate(&m->stream));
  float **pb[] = {&m->d_p, &m->d_g, &m->d_m, &m->d_v};
  for (float **b : pb) M_HIP(hipMalloc((void **)b, kNParam * sizeof(float)));
  std::vector<float> p, z(kNParam, 0.0f);
  init_params(seed, p);
  M_HIP(hipMemcpy(m->d_p, p.data(), kNParam * sizeof(float), hipMemcpyHostToDevice));
  M_HIP(hipMemcpy(m->d_m, z.data(), kNParam * sizeof(float), hipMemcpyHostToDevice));
  M_HIP(hipMemcpy(m->d_v, z.data(), kNParam * sizeof(float), hipMemcpyHostToDevice));
  M_HIP(hipMemcpy(m->d_g, z.data(), kNParam * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

int inca_mlp_destroy(inca_mlp *m) {
  if (!m) return 0;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  void *bufs[] = {m->d_p, m->d_g, m->d_m, m->d_v, m->d_x, m->d_out, m->d_act, m->d_dout,
                  m->d_delta, m->d_t, m->d_bl};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

const char *inca_mlp_last_error(const inca_mlp *m) { return m ? m->err.c_str() : ""; }

int inca_mlp_get_params(inca_mlp *m, float *p) {
  if (!m || !p) return INCA_EINVAL;
  M_HIP(hipSetDevice(m->device));
  M_HIP(hipMemcpy(p, m->d_p, kNParam * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int inca_mlp_set_params(inca_mlp *m, const float *p) {
  if (!m || !p) return INCA_EINVAL;
  M_HIP(hipSetDevice(m->device));
  M_HIP(hipMemcpy(m->d_p, p, kNParam * sizeof(float), hipMemcpyHostToDevice));
  std::vector<float> z(kNParam, 0.0f);
  M_HIP(hipMemcpy(m->d_m, z.data(), kNParam * sizeof(float), hipMemcpyHostToDevice));
  M_HIP(hipMemcpy(m->d_v, z.data(), kNParam * sizeof(float), hipMemcpyHostToDevice));
  m->steps = 0;
  return 0;
}

int inca_mlp_load(inca_mlp *m, const char *path) {
  if (!m || !path) return INCA_EINVAL;
  FILE *f = std::fopen(path, "rb");
  if (!f) return mfail(m, INCA_EIO, std::string("cannot open ") + path);
  // .npy: magic, version, header length, a dict header, then the raw array
  unsigned char pre[10];
  std::vector<float> p(kNParam);
  int rc = INCA_EFORMAT;
  std::string why = "not a .npy file";
  if (std::fread(pre, 1, 10, f) == 10 && std::memcmp(pre, "\x93NUMPY", 6) == 0) {
    size_t hlen = 0;
    bool ok = true;
    if (pre[6] == 1) {
      hlen = pre[8] | (pre[9] << 8);
    } else {
      unsigned char more[2];
      ok = std::fread(more, 1, 2, f) == 2;
      hlen = pre[8] | (pre[9] << 8) | ((size_t)more[0] << 16) | ((size_t)more[1] << 24);
    }
    std::string hdr(hlen, '\0');
    ok = ok && hlen < 65536 && std::fread(&hdr[0], 1, hlen, f) == hlen;
    auto has = [&](const char *t) { return hdr.find(t) != std::string::npos; };
    const std::string shape = "(" + std::to_string(kNParam) + ",)";
    if (!ok) {
      why = "truncated .npy header";
    } else if (!has("'descr': '<f4'") || !has("'fortran_order': False") ||
               !has(("'shape': " + shape).c_str())) {
      why = "expected a little-endian float32 array of shape " + shape;
    } else if (std::fread(p.data(), sizeof(float), kNParam, f) != (size_t)kNParam) {
      why = "truncated .npy data";
    } else {
      rc = 0;
    }
  }
  std::fclose(f);
  if (rc) return mfail(m, rc, std::string(path) + ": " + why);
  return inca_mlp_set_params(m, p.data());
}

int inca_mlp_forward(inca_mlp *m, int32_t n, const float *x, float *out) {
  if (!m || n < 0 || (n > 0 && (!x || !out))) return INCA_EINVAL;
  if (n == 0) return 0;
  M_HIP(hipSetDevice(m->device));
  int rc = forward(m, n, x, false);
  if (rc) return rc;
  M_HIP(hipMemcpyAsync(out, m->d_out, (size_t)n * kOut * sizeof(float), hipMemcpyDeviceToHost,
                       m->stream));
  M_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int inca_mlp_loss_grad(inca_mlp *m, int32_t n, const float *x, const float *target,
                       double *loss, float *grad) {
  if (!m || n < 1 || !x || !target) return INCA_EINVAL;
  for (int a = 0; a < n; ++a)
    for (int h = 0; h < 2; ++h) {
      const float c = target[(size_t)a * 3 + 1 + h];
      if (!(c >= 0.0f && c <= 3.0f && c == std::floor(c)))
        return mfail(m, INCA_EINVAL, "class targets must be 0..3");
    }
  M_HIP(hipSetDevice(m->device));
  int rc = forward(m, n, x, true);
  if (rc) return rc;
  M_HIP(hipMemcpyAsync(m->d_t, target, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice,
                       m->stream));
  hipLaunchKernelGGL(k_mlp_loss, dim3(nb(n)), dim3(kTM), 0, m->stream, n, m->d_out, m->d_t,
                     m->d_dout, m->d_bl);
  hipLaunchKernelGGL(k_mlp_backward, dim3(nb(n)), dim3(kTM), kNParam * sizeof(float), m->stream,
                     m->d_p, n, m->d_act, m->d_dout, m->d_delta);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(nb(kNParam)), dim3(kTM), 0, m->stream, n, m->d_x,
                     m->d_act, m->d_delta, m->d_g);
  M_HIP(hipGetLastError());
  std::vector<double> bl(nb(n));
  M_HIP(hipMemcpyAsync(bl.data(), m->d_bl, bl.size() * sizeof(double), hipMemcpyDeviceToHost,
                       m->stream));
  if (grad)
    M_HIP(hipMemcpyAsync(grad, m->d_g, kNParam * sizeof(float), hipMemcpyDeviceToHost,
                         m->stream));
  M_HIP(hipStreamSynchronize(m->stream));
  double s = 0.0;
  for (double v : bl) s += v;
  if (loss) *loss = s / n;
  return 0;
}

int inca_atom_features(int32_t natoms, const int32_t *an, const double *cartes, int32_t nprim,
                       const double *alpha, const int32_t *tmn, const int32_t *prim_atom,
                       float *feat) {
  if (natoms < 1 || !an || !cartes || !feat || nprim < 0) return INCA_EINVAL;
  if (nprim > 0 && (!alpha || !tmn || !prim_atom)) return INCA_EINVAL;
  for (int i = 0; i < nprim; ++i)
    if (prim_atom[i] < 1 || prim_atom[i] > natoms) return INCA_EINVAL;
  atom_features(natoms, an, cartes, nprim, alpha, tmn, prim_atom, feat);
  return 0;
}

int inca_learned_autogrid(inca_mlp *m, int32_t natoms, const int32_t *an, const double *cartes,
                          int32_t nprim, const double *alpha, const int32_t *tmn,
                          const int32_t *prim_atom, int32_t nohydro, int32_t cap, int32_t *nquad,
                          double *cent, double *Ps, double *sfalpha, int32_t *nradc,
                          int32_t *nangc, double *radius) {
  if (!m || !nquad) return INCA_EINVAL;
  std::vector<float> f((size_t)natoms * kIn), y((size_t)natoms * kOut);
  int rc = inca_atom_features(natoms, an, cartes, nprim, alpha, tmn, prim_atom, f.data());
  if (rc) return mfail(m, rc, "bad molecule description");
  if ((rc = inca_mlp_forward(m, natoms, f.data(), y.data()))) return rc;
  static const int kNrad[4] = {50, 75, 100, 150}, kNang[4] = {110, 194, 302, 590};
  std::vector<double> rad(natoms);
  std::vector<int> cr(natoms), ca(natoms);
  for (int a = 0; a < natoms; ++a) {
    const float *o = y.data() + (size_t)a * kOut;
    const double s = std::fmin(std::fmax((double)o[0], -1.2), 1.2);  // factor in [0.3, 3.3]
    rad[a] = bragg_slater_bohr(an[a]) * std::exp(s);
    int br = 0, bn = 0;
    for (int k = 1; k < 4; ++k) {
      if (o[1 + k] > o[1 + br]) br = k;
      if (o[5 + k] > o[5 + bn]) bn = k;
    }
    cr[a] = br;
    ca[a] = bn;
  }
  std::vector<double> c, p, al;
  const int nq = autogrid(natoms, an, cartes, nohydro, 0, rad.data(), c, p, al);
  *nquad = nq;
  if (cap < nq) return INCA_ERANGE;
  if (!cent || !Ps || !sfalpha || !nradc || !nangc) return INCA_EINVAL;
  // orders per centre: the largest class among the atoms the centre stands
  // for (the origin: all atoms; a pair centre: atoms j and k, merged centres
  // keep the first pair's atoms, which carry the same displacement)
  int gr = 0, ga = 0;
  for (int a = 0; a < natoms; ++a) {
    gr = std::max(gr, cr[a]);
    ga = std::max(ga, ca[a]);
  }
  std::memcpy(cent, c.data(), 3 * (size_t)nq * sizeof(double));
  std::memcpy(Ps, p.data(), (size_t)nq * sizeof(double));
  std::memcpy(sfalpha, al.data(), (size_t)nq * sizeof(double));
  for (int q = 0; q < nq; ++q) {
    int r = gr, aa = ga;
    if (q > 0) {  // find a heavy pair with this displacement
      r = 0;
      aa = 0;
      bool found = false;
      auto xyz = [&](int at, int ax) { return cartes[at + (size_t)natoms * ax]; };
      for (int j = 0; j + 1 < natoms && !found; ++j)
        for (int k = j + 1; k < natoms && !found; ++k) {
          if (nohydro && !(an[j] > 1 && an[k] > 1)) continue;
          for (int sgn = -1; sgn <= 1 && !found; sgn += 2) {
            double d[3];
            for (int ax = 0; ax < 3; ++ax) d[ax] = sgn * (xyz(j, ax) - xyz(k, ax));
            if (std::fabs(d[2]) < 0.5) d[2] = 0.0;
            bool same = true;
            for (int ax = 0; ax < 3; ++ax) same = same && std::fabs(d[ax] - c[3 * q + ax]) < 0.1;
            if (same) {
              found = true;
              r = std::max(cr[j], cr[k]);
              aa = std::max(ca[j], ca[k]);
            }
          }
        }
      if (!found) {
        r = gr;
        aa = ga;
      }
    }
    nradc[q] = kNrad[r];
    nangc[q] = kNang[aa];
  }
  if (radius)
    for (int a = 0; a < natoms; ++a) radius[a] = rad[a];
  return 0;
}

int inca_mlp_adam_step(inca_mlp *m, float lr) {
  if (!m || !(lr > 0.0f)) return INCA_EINVAL;
  M_HIP(hipSetDevice(m->device));
  ++m->steps;
  const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)m->steps));
  const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)m->steps));
  hipLaunchKernelGGL(k_mlp_adam, dim3(nb(kNParam)), dim3(kTM), 0, m->stream, m->d_p, m->d_g,
                     m->d_m, m->d_v, lr, c1, c2);
  M_HIP(hipGetLastError());
  M_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

}  // extern "C"
