// jms_nn.hip -- bf16 training and inference kernels (FP32 accumulation) of the
// detector-intensity surrogate network, and its C-ABI (include/jms_nn.h).
//
// Layouts: activations NCHW, one sample after the other; conv weights
// [co][ci][3][3]; dense weights [out][in].  bf16 is stored as the upper 16
// bits of the FP32 pattern (round to nearest even on conversion).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#include "jms_nn.h"

namespace {

constexpr int C0 = JMS_NN_IN_C, H0 = JMS_NN_IN_HW, CH = 32, HID = 64, NOUT = JMS_NN_OUT;
constexpr int NCONV = 4;
constexpr int STRIDE[NCONV] = {1, 2, 2, 2};
constexpr int CIN[NCONV] = {C0, CH, CH, CH};
constexpr int HIN[NCONV] = {H0, H0, H0 / 2, H0 / 4};
constexpr int HOUT[NCONV] = {H0, H0 / 2, H0 / 4, H0 / 8};
constexpr int FLAT = CH * (H0 / 8) * (H0 / 8);  // 512
constexpr int IN_SIZE = C0 * H0 * H0;

constexpr int conv_w(int l) { return CH * CIN[l] * 9; }
// parameter offsets
constexpr int off_cw(int l) { return l == 0 ? 0 : off_cw(l - 1) + conv_w(l - 1) + CH; }
constexpr int off_cb(int l) { return off_cw(l) + conv_w(l); }
constexpr int OFF_D1W = off_cb(NCONV - 1) + CH;
constexpr int OFF_D1B = OFF_D1W + HID * FLAT;
constexpr int OFF_D2W = OFF_D1B + HID;
constexpr int OFF_D2B = OFF_D2W + NOUT * HID;
constexpr int NPARAM = OFF_D2B + NOUT;

__device__ __forceinline__ float bf2f(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

}  // namespace

#include "jms_nn_mfma.hpp"

namespace {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
constexpr int GMAX = 8;  // channels per thread at most (a wavefront-uniform channel group)

// y[n][co][oy][ox] = relu(b[co] + sum_{ci,ky,kx} x[n][ci][oy*S+ky-1][ox*S+kx-1] w[co][ci][ky][kx]);
// thread (n, oy, ox) computes the G output channels of group blockIdx.y from
// one input read per tap
template <int S, int G>
__global__ __launch_bounds__(256) void k_conv_fwd(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
                                                  const float* __restrict__ b, uint16_t* __restrict__ y, int n,
                                                  int ci_n, int hin, int hout) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int co0 = blockIdx.y * G;
  float acc[G];
#pragma unroll
  for (int j = 0; j < G; ++j) acc[j] = b[co0 + j];
  const long long tot = (long long)n * hout * hout;
  if (i >= tot) return;
  const int ox = (int)(i % hout), oy = (int)((i / hout) % hout);
  const long long s = i / ((long long)hout * hout);
  const uint16_t* xs = x + s * ci_n * hin * hin;
  for (int ci = 0; ci < ci_n; ++ci) {
    const uint16_t* xc = xs + (long long)ci * hin * hin;
#pragma unroll
    for (int ky = 0; ky < 3; ++ky) {
      const int iy = oy * S + ky - 1;
#pragma unroll
      for (int kx = 0; kx < 3; ++kx) {
        const int ix = ox * S + kx - 1;
        const float xv = (iy >= 0 && iy < hin && ix >= 0 && ix < hin) ? bf2f(xc[iy * hin + ix]) : 0.f;
#pragma unroll
        for (int j = 0; j < G; ++j)
          acc[j] = fmaf(xv, bf2f(w[((co0 + j) * ci_n + ci) * 9 + ky * 3 + kx]), acc[j]);
      }
    }
  }
  const long long hw = (long long)hout * hout;
#pragma unroll
  for (int j = 0; j < G; ++j) y[(s * CH + co0 + j) * hw + oy * hout + ox] = f2bf(fmaxf(acc[j], 0.f));
}

// dx[n][ci][iy][ix] = [a[n][ci][iy][ix] > 0] * sum dy[n][co][oy][ox] w[co][ci][ky][kx]
// over iy = oy*S + ky - 1, ix = ox*S + kx - 1; thread (n, iy, ix) computes the
// G input channels of group blockIdx.y from one dy read per tap
template <int S, int G>
__global__ __launch_bounds__(256) void k_conv_bwd_data(const float* __restrict__ dy, const uint16_t* __restrict__ w,
                                                       const uint16_t* __restrict__ a, float* __restrict__ dx, int n,
                                                       int ci_n, int hin, int hout) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ci0 = blockIdx.y * G;
  const long long tot = (long long)n * hin * hin;
  if (i >= tot) return;
  const int ix = (int)(i % hin), iy = (int)((i / hin) % hin);
  const long long s = i / ((long long)hin * hin);
  const float* ds = dy + s * CH * hout * hout;
  float acc[G];
#pragma unroll
  for (int j = 0; j < G; ++j) acc[j] = 0.f;
#pragma unroll
  for (int ky = 0; ky < 3; ++ky) {
    const int ty = iy + 1 - ky;
    if (ty < 0 || ty % S) continue;
    const int oy = ty / S;
    if (oy >= hout) continue;
#pragma unroll
    for (int kx = 0; kx < 3; ++kx) {
      const int tx = ix + 1 - kx;
      if (tx < 0 || tx % S) continue;
      const int ox = tx / S;
      if (ox >= hout) continue;
      for (int co = 0; co < CH; ++co) {
        const float g = ds[((long long)co * hout + oy) * hout + ox];
#pragma unroll
        for (int j = 0; j < G; ++j) acc[j] = fmaf(g, bf2f(w[(co * ci_n + ci0 + j) * 9 + ky * 3 + kx]), acc[j]);
      }
    }
  }
  const long long hw = (long long)hin * hin;
#pragma unroll
  for (int j = 0; j < G; ++j) {
    const long long o = (s * ci_n + ci0 + j) * hw + iy * hin + ix;
    dx[o] = (bf2f(a[o]) > 0.f) ? acc[j] : 0.f;
  }
}

// block (ci, co group, part): the part-th share of the samples adds to
// dw[co][ci][ky][kx] = sum_{n,oy,ox} dy[n][co][oy][ox] x[n][ci][iy][ix] for
// the G channels co of the group (atomically; dw is cleared first); the
// ci = 0 blocks add db[co] = sum dy
template <int S, int G = GMAX>
__global__ __launch_bounds__(256) void k_conv_bwd_w(const float* __restrict__ dy, const uint16_t* __restrict__ x,
                                                    float* __restrict__ dw, float* __restrict__ db, int n,
                                                    int ci_n, int hin, int hout) {
  const int ci = blockIdx.x % ci_n, co0 = (blockIdx.x / ci_n) * G;
  float acc[G][10];
#pragma unroll
  for (int j = 0; j < G; ++j)
#pragma unroll
    for (int k = 0; k < 10; ++k) acc[j][k] = 0.f;
  const int per = hout * hout;
  const int s0 = (int)((long long)n * blockIdx.y / gridDim.y), s1 = (int)((long long)n * (blockIdx.y + 1) / gridDim.y);
  const long long tot = (long long)s1 * per;
  for (long long e = (long long)s0 * per + threadIdx.x; e < tot; e += blockDim.x) {
    const long long s = e / per;
    const int r = (int)(e % per), oy = r / hout, ox = r % hout;
    float g[G];
#pragma unroll
    for (int j = 0; j < G; ++j) g[j] = dy[(s * CH + co0 + j) * per + r];
    const uint16_t* xc = x + (s * ci_n + ci) * hin * hin;
#pragma unroll
    for (int j = 0; j < G; ++j) acc[j][9] += g[j];
#pragma unroll
    for (int ky = 0; ky < 3; ++ky) {
      const int iy = oy * S + ky - 1;
#pragma unroll
      for (int kx = 0; kx < 3; ++kx) {
        const int ix = ox * S + kx - 1;
        const float xv = (iy >= 0 && iy < hin && ix >= 0 && ix < hin) ? bf2f(xc[iy * hin + ix]) : 0.f;
#pragma unroll
        for (int j = 0; j < G; ++j) acc[j][ky * 3 + kx] = fmaf(g[j], xv, acc[j][ky * 3 + kx]);
      }
    }
  }
  __shared__ float red[G * 10][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < G; ++j)
#pragma unroll
    for (int k = 0; k < 10; ++k) {
      float v = acc[j][k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[j * 10 + k][wv] = v;
    }
  __syncthreads();
  if (threadIdx.x < G * 10) {
    const int j = threadIdx.x / 10, k = threadIdx.x % 10;
    const float v = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
    if (k < 9) atomicAdd(&dw[((long long)(co0 + j) * ci_n + ci) * 9 + k], v);
    else if (ci == 0) atomicAdd(&db[co0 + j], v);
  }
}

// y[n][o] = b[o] + sum_i x[n][i] w[o][i]; relu -> bf16 out, else fp32 out
// (nin a multiple of 8: eight bf16 per 16-byte load)
__global__ void k_dense_fwd(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
                            const float* __restrict__ b, int n, int nin, int nout, uint16_t* __restrict__ ybf,
                            float* __restrict__ yf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nout) return;
  const int o = i % nout, s = i / nout;
  const uint4* xs = reinterpret_cast<const uint4*>(x + (long long)s * nin);
  const uint4* wo = reinterpret_cast<const uint4*>(w + (long long)o * nin);
  float acc = b[o];
  for (int k = 0; k < nin / 8; ++k) {
    const uint4 xv = xs[k], wv = wo[k];
    const uint32_t xa[4] = {xv.x, xv.y, xv.z, xv.w}, wa[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc = fmaf(__uint_as_float(xa[j] << 16), __uint_as_float(wa[j] << 16), acc);
      acc = fmaf(__uint_as_float(xa[j] & 0xFFFF0000u), __uint_as_float(wa[j] & 0xFFFF0000u), acc);
    }
  }
  if (ybf) ybf[i] = f2bf(fmaxf(acc, 0.f));
  else yf[i] = acc;
}

// dx[n][i] = [a[n][i] > 0] * sum_o dy[n][o] w[o][i]
__global__ void k_dense_bwd_data(const float* __restrict__ dy, const uint16_t* __restrict__ w,
                                 const uint16_t* __restrict__ a, int n, int nin, int nout, float* __restrict__ dx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nin) return;
  const int k = i % nin, s = i / nin;
  if (bf2f(a[i]) <= 0.f) {
    dx[i] = 0.f;
    return;
  }
  float acc = 0.f;
  for (int o = 0; o < nout; ++o) acc = fmaf(dy[s * nout + o], bf2f(w[(long long)o * nin + k]), acc);
  dx[i] = acc;
}

// dw[o][i] += sum over this block's share of the samples of dy[n][o] x[n][i];
// threads i == nin add db[o] += sum dy[n][o]; dw and db are cleared first.
// grid: (ceil(nout (nin + 1) / 256), shares)
__global__ void k_dense_bwd_w(const float* __restrict__ dy, const uint16_t* __restrict__ x, int n, int nin,
                              int nout, float* __restrict__ dw, float* __restrict__ db) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nout * (nin + 1)) return;
  const int o = i / (nin + 1), k = i % (nin + 1);
  const int s0 = (int)((long long)n * blockIdx.y / gridDim.y), s1 = (int)((long long)n * (blockIdx.y + 1) / gridDim.y);
  float acc = 0.f;
  if (k < nin) {
    for (int s = s0; s < s1; ++s) acc = fmaf(dy[s * nout + o], bf2f(x[(long long)s * nin + k]), acc);
    atomicAdd(&dw[(long long)o * nin + k], acc);
  } else {
    for (int s = s0; s < s1; ++s) acc += dy[s * nout + o];
    atomicAdd(&db[o], acc);
  }
}

// dy = d/dy mean(((y - t)/scale)^2); per-block loss sums into loss_part[blockIdx.x]
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ y, const float* __restrict__ t,
                                              const float* __restrict__ scale, int n, float* __restrict__ dy,
                                              float* __restrict__ loss_part, float* __restrict__ loss_acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < n * NOUT) {
    const float sc = scale[i % NOUT];
    const float r = (y[i] - t[i]) / sc;
    l = r * r;
    dy[i] = 2.f * r / sc / (float)(n * NOUT);
  }
  __shared__ float red[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float v = red[0] + red[1] + red[2] + red[3];
    loss_part[blockIdx.x] = v;
    if (loss_acc) atomicAdd(loss_acc, v / (float)(n * NOUT));  // mean over the batch
  }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ pbf, int np, float lr, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float q = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
  p[i] = q;
  pbf[i] = f2bf(q);
}

__global__ void k_zero(float* __restrict__ p, int np) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) p[i] = 0.f;
}

__global__ void k_to_bf16(const float* __restrict__ p, uint16_t* __restrict__ q, int np) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) q[i] = f2bf(p[i]);
}

// batch gather: xb[j] = X[idx[j]], tb[j] = T[idx[j]]
__global__ void k_gather(const uint16_t* __restrict__ X, const float* __restrict__ T, const int* __restrict__ idx,
                         int nb, uint16_t* __restrict__ xb, float* __restrict__ tb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (long long)nb * IN_SIZE) {
    const long long j = i / IN_SIZE, e = i % IN_SIZE;
    xb[i] = X[(long long)idx[j] * IN_SIZE + e];
  }
  if (i < (long long)nb * NOUT) tb[i] = T[(long long)idx[i / NOUT] * NOUT + i % NOUT];
}

// input windows: thread per (sample, i, j); all depth bins of the pixel
__global__ void k_windows(const float* __restrict__ maps, const float* __restrict__ cells, int nt, int m, int n,
                          const int* __restrict__ cid, const float* __restrict__ px, const float* __restrict__ py,
                          const float* __restrict__ binw, float step, float scale, uint16_t* __restrict__ x) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * H0 * H0) return;
  const int jx = (int)(i % H0), iy = (int)((i / H0) % H0);
  const long long s = i / (H0 * H0);
  const int k = cid[s];
  const float cx = cells[2 * k], cy = cells[2 * k + 1];
  // grid coordinates of the pixel (periodic)
  const float u = (px[s] + (jx - H0 / 2) * step) / cx * m;
  const float v = (py[s] + (iy - H0 / 2) * step) / cy * m;
  const float fu0 = floorf(u), fv0 = floorf(v);
  const float fu = u - fu0, fv = v - fv0;
  const int u0 = (((int)fu0 % m) + m) % m, v0 = (((int)fv0 % m) + m) % m;
  const int u1 = (u0 + 1) % m, v1 = (v0 + 1) % m;
  float vt[4] = {0.f, 0.f, 0.f, 0.f};
  for (int t = 0; t < nt && t < 4; ++t) {
    const float* mp = maps + ((long long)k * nt + t) * m * m;
    vt[t] = (1.f - fv) * ((1.f - fu) * mp[v0 * m + u0] + fu * mp[v0 * m + u1]) +
            fv * ((1.f - fu) * mp[v1 * m + u0] + fu * mp[v1 * m + u1]);
  }
  for (int c = 0; c < C0; ++c) {
    float acc = 0.f;
    for (int t = 0; t < nt && t < 4; ++t) acc = fmaf(binw[c * nt + t], vt[t], acc);
    x[((s * C0 + c) * H0 + iy) * H0 + jx] = f2bf(acc * scale);
  }
}

unsigned int nblk(long long n, int bs) { return (unsigned int)((n + bs - 1) / bs); }

// channels per thread of a conv launch over px pixels and nch channels: as
// many as keep at least 64k threads in flight (8, 4, 2 or 1)
int group_size(long long px, int nch) {
  int g = GMAX;
  while (g > 1 && px * (nch / g) < 65536) g /= 2;
  return g;
}

}  // namespace

// ---------------------------------------------------------------------------
// model
// ---------------------------------------------------------------------------
struct jms_nn {
  int maxb = 0;
  hipStream_t st = nullptr;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* pbf = nullptr;
  uint16_t* act[NCONV + 2] = {};  // input, conv1..conv4, dense1
  float* out = nullptr;           // [maxb][3]
  float* gbuf[NCONV + 2] = {};    // gradient w.r.t. conv1..4 outputs, dense1 output, final output
  float* tb = nullptr;
  float* scale = nullptr;
  float* loss_part = nullptr;
  int* idx = nullptr;
  int* perm = nullptr;       // epoch permutation (device), capacity perm_cap
  int perm_cap = 0;
  float* eloss = nullptr;    // per-batch losses of an epoch (device), capacity eloss_cap
  int eloss_cap = 0;
  float* part = nullptr;     // per-sample weight-gradient partials of the matrix-core path
  bool mfma = true;          // convolutions on the matrix cores (else the FMA kernels)
  long long step = 0;
};

struct jms_nn_data {
  int n = 0;
  uint16_t* x = nullptr;
  float* t = nullptr;
};

namespace {

int act_size(int l) {  // elements per sample of act[l]
  if (l == 0) return IN_SIZE;
  if (l <= NCONV) return CH * HOUT[l - 1] * HOUT[l - 1];
  return HID;
}

template <class T>
hipError_t dalloc(T** p, size_t n) { return hipMalloc((void**)p, n * sizeof(T)); }

void release(jms_nn* m) {
  if (!m) return;
  (void)hipFree(m->p); (void)hipFree(m->g); (void)hipFree(m->m); (void)hipFree(m->v); (void)hipFree(m->pbf);
  for (auto* a : m->act) (void)hipFree(a);
  for (auto* a : m->gbuf) (void)hipFree(a);
  (void)hipFree(m->out); (void)hipFree(m->tb); (void)hipFree(m->scale); (void)hipFree(m->loss_part);
  (void)hipFree(m->idx); (void)hipFree(m->perm); (void)hipFree(m->eloss); (void)hipFree(m->part);
  if (m->st) (void)hipStreamDestroy(m->st);
  delete m;
}

// matrix-core convolution of layer l (one block per sample)
void conv_fwd_mfma(jms_nn* m, int l, int n, const uint16_t* in0) {
  const dim3 g(n), b(256);
  const uint16_t* x = (l == 0 && in0) ? in0 : m->act[l];
  const uint16_t* w = m->pbf + off_cw(l);
  const float* bias = m->p + off_cb(l);
  uint16_t* y = m->act[l + 1];
  switch (l) {
    case 0: hipLaunchKernelGGL((nnmfma::k_conv_fwd<1, C0, H0>), g, b, 0, m->st, x, w, bias, y); break;
    case 1: hipLaunchKernelGGL((nnmfma::k_conv_fwd<2, CH, H0>), g, b, 0, m->st, x, w, bias, y); break;
    case 2: hipLaunchKernelGGL((nnmfma::k_conv_fwd<2, CH, H0 / 2>), g, b, 0, m->st, x, w, bias, y); break;
    default: hipLaunchKernelGGL((nnmfma::k_conv_fwd<2, CH, H0 / 4>), g, b, 0, m->st, x, w, bias, y); break;
  }
}

// matrix-core gradients of layer l: weights (per-sample partials reduced in
// a fixed order), bias, and for l > 0 the gradient w.r.t. its input
void conv_bwd_mfma(jms_nn* m, int l, int n) {
  const dim3 g(n), b(256);
  const float* dy = m->gbuf[l];
  const uint16_t* x = m->act[l];
  switch (l) {
    case 0: hipLaunchKernelGGL((nnmfma::k_conv_bwd_w<1, C0, H0>), g, b, 0, m->st, dy, x, m->part); break;
    case 1: hipLaunchKernelGGL((nnmfma::k_conv_bwd_w<2, CH, H0>), g, b, 0, m->st, dy, x, m->part); break;
    case 2: hipLaunchKernelGGL((nnmfma::k_conv_bwd_w<2, CH, H0 / 2>), g, b, 0, m->st, dy, x, m->part); break;
    default: hipLaunchKernelGGL((nnmfma::k_conv_bwd_w<2, CH, H0 / 4>), g, b, 0, m->st, dy, x, m->part); break;
  }
  const int cnt = conv_w(l);
  const unsigned shares = (unsigned)std::max(1, std::min(n / 8, 32));
  hipLaunchKernelGGL(nnmfma::k_reduce_parts, dim3(nblk(cnt, 256), shares), dim3(256), 0, m->st, m->part, n, cnt,
                     m->g + off_cw(l));
  hipLaunchKernelGGL(nnmfma::k_bias_grad, dim3(CH, shares), dim3(256), 0, m->st, dy, n, HOUT[l] * HOUT[l],
                     m->g + off_cb(l));
  if (l == 0) return;
  const uint16_t* w = m->pbf + off_cw(l);
  float* dx = m->gbuf[l - 1];
  switch (l) {
    case 1: hipLaunchKernelGGL((nnmfma::k_conv_bwd_data<2, H0>), g, b, 0, m->st, dy, w, x, dx); break;
    case 2: hipLaunchKernelGGL((nnmfma::k_conv_bwd_data<2, H0 / 2>), g, b, 0, m->st, dy, w, x, dx); break;
    default: hipLaunchKernelGGL((nnmfma::k_conv_bwd_data<2, H0 / 4>), g, b, 0, m->st, dy, w, x, dx); break;
  }
}

// forward over act[0] (n samples) -> out
// forward pass of n samples: input act[0] (or in0: n windows in place),
// output [n][3] into out (default m->out)
hipError_t forward(jms_nn* m, int n, const uint16_t* in0 = nullptr, float* out = nullptr) {
  for (int l = 0; l < NCONV; ++l) {
    if (m->mfma) {
      conv_fwd_mfma(m, l, n, in0);
      continue;
    }
    const long long px = (long long)n * HOUT[l] * HOUT[l];
    const int gsz = group_size(px, CH);
    const dim3 g(nblk(px, 256), CH / gsz);
    const uint16_t* x = (l == 0 && in0) ? in0 : m->act[l];
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, g, dim3(256), 0, m->st, x, m->pbf + off_cw(l), m->p + off_cb(l), m->act[l + 1], n,
                         CIN[l], HIN[l], HOUT[l]);
    };
    if (STRIDE[l] == 1) {
      if (gsz == 8) go(k_conv_fwd<1, 8>); else if (gsz == 4) go(k_conv_fwd<1, 4>);
      else if (gsz == 2) go(k_conv_fwd<1, 2>); else go(k_conv_fwd<1, 1>);
    } else {
      if (gsz == 8) go(k_conv_fwd<2, 8>); else if (gsz == 4) go(k_conv_fwd<2, 4>);
      else if (gsz == 2) go(k_conv_fwd<2, 2>); else go(k_conv_fwd<2, 1>);
    }
  }
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk((long long)n * HID, 256)), dim3(256), 0, m->st, m->act[NCONV],
                     m->pbf + OFF_D1W, m->p + OFF_D1B, n, FLAT, HID, m->act[NCONV + 1], (float*)nullptr);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk((long long)n * NOUT, 256)), dim3(256), 0, m->st, m->act[NCONV + 1],
                     m->pbf + OFF_D2W, m->p + OFF_D2B, n, HID, NOUT, (uint16_t*)nullptr, out ? out : m->out);
  return hipGetLastError();
}

// loss and gradients of the batch in act[0] / tb (n samples) into m->g; the
// batch's mean loss is also added to *loss_acc when given
hipError_t backward(jms_nn* m, int n, float* loss_acc = nullptr) {
  const unsigned nb = nblk((long long)n * NOUT, 256);
  float* gout = m->gbuf[NCONV + 1];
  float* gh = m->gbuf[NCONV];
  // every gradient is accumulated over shares of the batch: clear them all
  hipLaunchKernelGGL(k_zero, dim3(nblk(NPARAM, 256)), dim3(256), 0, m->st, m->g, NPARAM);
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, m->st, m->out, m->tb, m->scale, n, gout, m->loss_part,
                     loss_acc);
  const unsigned shares = (unsigned)std::max(1, std::min(n / 16, 16));
  // dense 2
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((long long)NOUT * (HID + 1), 256), shares), dim3(256), 0, m->st,
                     gout, m->act[NCONV + 1], n, HID, NOUT, m->g + OFF_D2W, m->g + OFF_D2B);
  hipLaunchKernelGGL(k_dense_bwd_data, dim3(nblk((long long)n * HID, 256)), dim3(256), 0, m->st, gout,
                     m->pbf + OFF_D2W, m->act[NCONV + 1], n, HID, NOUT, gh);
  // dense 1
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((long long)HID * (FLAT + 1), 256), shares), dim3(256), 0, m->st,
                     gh, m->act[NCONV], n, FLAT, HID, m->g + OFF_D1W, m->g + OFF_D1B);
  hipLaunchKernelGGL(k_dense_bwd_data, dim3(nblk((long long)n * FLAT, 256)), dim3(256), 0, m->st, gh,
                     m->pbf + OFF_D1W, m->act[NCONV], n, FLAT, HID, m->gbuf[NCONV - 1]);
  // convolutions, last to first; gbuf[l] = gradient w.r.t. the output of conv l+1;
  // their weight gradients are accumulated atomically over shares of the batch
  if (m->mfma) {
    for (int l = NCONV - 1; l >= 0; --l) conv_bwd_mfma(m, l, n);
    return hipGetLastError();
  }
  for (int l = NCONV - 1; l >= 0; --l) {
    const float* dy = m->gbuf[l];
    const int nblk_w = CIN[l] * (CH / GMAX);
    const int parts = std::max(1, std::min(n, 2048 / nblk_w));
    if (STRIDE[l] == 1)
      hipLaunchKernelGGL(k_conv_bwd_w<1>, dim3(nblk_w, parts), dim3(256), 0, m->st, dy, m->act[l],
                         m->g + off_cw(l), m->g + off_cb(l), n, CIN[l], HIN[l], HOUT[l]);
    else
      hipLaunchKernelGGL(k_conv_bwd_w<2>, dim3(nblk_w, parts), dim3(256), 0, m->st, dy, m->act[l],
                         m->g + off_cw(l), m->g + off_cb(l), n, CIN[l], HIN[l], HOUT[l]);
    if (l > 0) {
      const long long px = (long long)n * HIN[l] * HIN[l];
      const int gsz = group_size(px, CIN[l]);
      const dim3 g(nblk(px, 256), CIN[l] / gsz);
      auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, g, dim3(256), 0, m->st, dy, m->pbf + off_cw(l), m->act[l], m->gbuf[l - 1], n,
                           CIN[l], HIN[l], HOUT[l]);
      };
      if (STRIDE[l] == 1) {
        if (gsz == 8) go(k_conv_bwd_data<1, 8>); else if (gsz == 4) go(k_conv_bwd_data<1, 4>);
        else if (gsz == 2) go(k_conv_bwd_data<1, 2>); else go(k_conv_bwd_data<1, 1>);
      } else {
        if (gsz == 8) go(k_conv_bwd_data<2, 8>); else if (gsz == 4) go(k_conv_bwd_data<2, 4>);
        else if (gsz == 2) go(k_conv_bwd_data<2, 2>); else go(k_conv_bwd_data<2, 1>);
      }
    }
  }
  return hipGetLastError();
}

float sum_loss(jms_nn* m, int n) {
  const unsigned nb = nblk((long long)n * NOUT, 256);
  std::vector<float> h(nb);
  if (hipMemcpyAsync(h.data(), m->loss_part, nb * sizeof(float), hipMemcpyDeviceToHost, m->st) != hipSuccess) return NAN;
  if (hipStreamSynchronize(m->st) != hipSuccess) return NAN;
  double s = 0;
  for (float x : h) s += x;
  return (float)(s / (double)(n * NOUT));
}

}  // namespace

extern "C" {

int jms_nn_num_params(void) { return NPARAM; }

int jms_nn_create(jms_nn** out, int max_batch, unsigned int seed) {
  if (!out || max_batch < 1) return 1;
  *out = nullptr;
  jms_nn* m = new jms_nn;
  m->maxb = max_batch;
  hipError_t e = hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = dalloc(&m->p, NPARAM);
  if (e == hipSuccess) e = dalloc(&m->g, NPARAM);
  if (e == hipSuccess) e = dalloc(&m->m, NPARAM);
  if (e == hipSuccess) e = dalloc(&m->v, NPARAM);
  if (e == hipSuccess) e = dalloc(&m->pbf, NPARAM);
  for (int l = 0; l < NCONV + 2 && e == hipSuccess; ++l) e = dalloc(&m->act[l], (size_t)max_batch * act_size(l));
  const int gsz[NCONV + 2] = {act_size(1), act_size(2), act_size(3), act_size(4), HID, NOUT};
  for (int l = 0; l < NCONV + 2 && e == hipSuccess; ++l) e = dalloc(&m->gbuf[l], (size_t)max_batch * gsz[l]);
  if (e == hipSuccess) e = dalloc(&m->out, (size_t)max_batch * NOUT);
  if (e == hipSuccess) e = dalloc(&m->tb, (size_t)max_batch * NOUT);
  if (e == hipSuccess) e = dalloc(&m->scale, NOUT);
  if (e == hipSuccess) e = dalloc(&m->loss_part, nblk((long long)max_batch * NOUT, 256));
  if (e == hipSuccess) e = dalloc(&m->idx, max_batch);
  if (e == hipSuccess) e = dalloc(&m->part, (size_t)max_batch * conv_w(1));  // the largest layer
  if (e != hipSuccess) {
    release(m);
    return 2;
  }
  // He-normal weights, zero biases
  std::vector<float> h(NPARAM, 0.f);
  std::mt19937 rng(seed);
  auto fill = [&](int off, int cnt, int fan_in) {
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)fan_in));
    for (int i = 0; i < cnt; ++i) h[off + i] = nd(rng);
  };
  for (int l = 0; l < NCONV; ++l) fill(off_cw(l), conv_w(l), CIN[l] * 9);
  fill(OFF_D1W, HID * FLAT, FLAT);
  fill(OFF_D2W, NOUT * HID, HID);
  *out = m;
  return jms_nn_set_params(m, h.data());
}

void jms_nn_destroy(jms_nn* m) { release(m); }

int jms_nn_set_kernels(jms_nn* m, int mfma) {
  if (!m) return 1;
  m->mfma = mfma != 0;
  return 0;
}

int jms_nn_set_params(jms_nn* m, const float* in) {
  if (!m || !in) return 1;
  hipError_t e = hipMemcpyAsync(m->p, in, NPARAM * sizeof(float), hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) e = hipMemsetAsync(m->m, 0, NPARAM * sizeof(float), m->st);
  if (e == hipSuccess) e = hipMemsetAsync(m->v, 0, NPARAM * sizeof(float), m->st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(NPARAM, 256)), dim3(256), 0, m->st, m->p, m->pbf, NPARAM);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(m->st);
  m->step = 0;
  return e == hipSuccess ? 0 : 2;
}

int jms_nn_get_params(jms_nn* m, float* out) {
  if (!m || !out) return 1;
  hipError_t e = hipMemcpyAsync(out, m->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->st);
  if (e == hipSuccess) e = hipStreamSynchronize(m->st);
  return e == hipSuccess ? 0 : 2;
}

int jms_nn_forward(jms_nn* m, const uint16_t* x, int n, float* out) {
  if (!m || !x || !out || n < 1 || n > m->maxb) return 1;
  hipError_t e = hipMemcpyAsync(m->act[0], x, (size_t)n * IN_SIZE * 2, hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) e = forward(m, n);
  if (e == hipSuccess) e = hipMemcpyAsync(out, m->out, (size_t)n * NOUT * 4, hipMemcpyDeviceToHost, m->st);
  if (e == hipSuccess) e = hipStreamSynchronize(m->st);
  return e == hipSuccess ? 0 : 2;
}

int jms_nn_gradients(jms_nn* m, const uint16_t* x, const float* t, const float* scale, int n, float* grads,
                     float* loss) {
  if (!m || !x || !t || !scale || !grads || n < 1 || n > m->maxb) return 1;
  hipError_t e = hipMemcpyAsync(m->act[0], x, (size_t)n * IN_SIZE * 2, hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) e = hipMemcpyAsync(m->tb, t, (size_t)n * NOUT * 4, hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) e = hipMemcpyAsync(m->scale, scale, NOUT * 4, hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) e = forward(m, n);
  if (e == hipSuccess) e = backward(m, n);
  if (e == hipSuccess) e = hipMemcpyAsync(grads, m->g, NPARAM * 4, hipMemcpyDeviceToHost, m->st);
  if (e != hipSuccess) return 2;
  const float l = sum_loss(m, n);
  if (loss) *loss = l;
  return std::isfinite(l) ? 0 : 3;
}

int jms_nn_data_create(jms_nn_data** out, const uint16_t* x, const float* t, int n) {
  if (!out || !x || !t || n < 1) return 1;
  *out = nullptr;
  jms_nn_data* d = new jms_nn_data;
  d->n = n;
  hipError_t e = dalloc(&d->x, (size_t)n * IN_SIZE);
  if (e == hipSuccess) e = dalloc(&d->t, (size_t)n * NOUT);
  if (e == hipSuccess) e = hipMemcpy(d->x, x, (size_t)n * IN_SIZE * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d->t, t, (size_t)n * NOUT * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    jms_nn_data_destroy(d);
    return 2;
  }
  *out = d;
  return 0;
}

int jms_nn_data_windows(jms_nn_data** out, int ncrys, int nt, int m, const float* maps, const float* cells,
                        int n, const int* cid, const float* px, const float* py, const float* binw, float step,
                        float scale, const float* t) {
  if (!out || !maps || !cells || !cid || !px || !py || !binw || n < 1 || ncrys < 1 || nt < 1 || nt > 4 || m < 2)
    return 1;
  for (int s = 0; s < n; ++s)
    if (cid[s] < 0 || cid[s] >= ncrys) return 1;
  for (int k = 0; k < 2 * ncrys; ++k)
    if (!(cells[k] > 0.f) || !std::isfinite(cells[k])) return 1;
  for (int s = 0; s < n; ++s)
    if (!std::isfinite(px[s]) || !std::isfinite(py[s])) return 1;
  *out = nullptr;
  jms_nn_data* d = new jms_nn_data;
  d->n = n;
  float *dmaps = nullptr, *dcells = nullptr, *dpx = nullptr, *dpy = nullptr, *dbw = nullptr;
  int* dcid = nullptr;
  const size_t nm = (size_t)ncrys * nt * m * m;
  hipError_t e = dalloc(&d->x, (size_t)n * IN_SIZE);
  if (e == hipSuccess) e = dalloc(&d->t, (size_t)n * NOUT);
  if (e == hipSuccess) e = dalloc(&dmaps, nm);
  if (e == hipSuccess) e = dalloc(&dcells, 2 * (size_t)ncrys);
  if (e == hipSuccess) e = dalloc(&dcid, n);
  if (e == hipSuccess) e = dalloc(&dpx, n);
  if (e == hipSuccess) e = dalloc(&dpy, n);
  if (e == hipSuccess) e = dalloc(&dbw, C0 * nt);
  if (e == hipSuccess) e = hipMemcpy(dmaps, maps, nm * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dcells, cells, 2 * (size_t)ncrys * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dcid, cid, (size_t)n * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dpx, px, (size_t)n * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dpy, py, (size_t)n * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dbw, binw, C0 * nt * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    if (t) e = hipMemcpy(d->t, t, (size_t)n * NOUT * 4, hipMemcpyHostToDevice);
    else e = hipMemset(d->t, 0, (size_t)n * NOUT * 4);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_windows, dim3(nblk((long long)n * H0 * H0, 256)), dim3(256), 0, 0, dmaps, dcells, nt, m,
                       n, dcid, dpx, dpy, dbw, step, scale, d->x);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(dmaps); (void)hipFree(dcells); (void)hipFree(dcid); (void)hipFree(dpx); (void)hipFree(dpy);
  (void)hipFree(dbw);
  if (e != hipSuccess) {
    jms_nn_data_destroy(d);
    return 2;
  }
  *out = d;
  return 0;
}

/* the windows of a set back on the host (bf16 bit patterns) */
int jms_nn_data_get(jms_nn_data* d, uint16_t* x) {
  if (!d || !x) return 1;
  return hipMemcpy(x, d->x, (size_t)d->n * IN_SIZE * 2, hipMemcpyDeviceToHost) == hipSuccess ? 0 : 2;
}

void jms_nn_data_destroy(jms_nn_data* d) {
  if (!d) return;
  (void)hipFree(d->x);
  (void)hipFree(d->t);
  delete d;
}

int jms_nn_train_epoch(jms_nn* m, jms_nn_data* d, int batch, float lr, const float* scale, unsigned int seed,
                       float* mean_loss) {
  if (!m || !d || !scale || batch < 1 || batch > m->maxb) return 1;
  const int nbatch = d->n / batch;  // whole batches only
  std::vector<int> perm(d->n);
  for (int i = 0; i < d->n; ++i) perm[i] = i;
  std::mt19937 rng(seed);
  std::shuffle(perm.begin(), perm.end(), rng);
  hipError_t e = hipSuccess;
  if (m->perm_cap < d->n) {
    (void)hipFree(m->perm);
    m->perm = nullptr;
    m->perm_cap = 0;
    e = dalloc(&m->perm, d->n);
    if (e == hipSuccess) m->perm_cap = d->n;
  }
  if (e == hipSuccess && m->eloss_cap < std::max(1, nbatch)) {
    (void)hipFree(m->eloss);
    m->eloss = nullptr;
    m->eloss_cap = 0;
    e = dalloc(&m->eloss, std::max(1, nbatch));
    if (e == hipSuccess) m->eloss_cap = std::max(1, nbatch);
  }
  // the whole epoch is enqueued without host synchronisation: the
  // permutation and the scale go up once, the batch losses come back once
  if (e == hipSuccess) e = hipMemcpyAsync(m->perm, perm.data(), (size_t)d->n * sizeof(int), hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess) e = hipMemcpyAsync(m->scale, scale, NOUT * 4, hipMemcpyHostToDevice, m->st);
  if (e == hipSuccess && nbatch > 0) {
    hipLaunchKernelGGL(k_zero, dim3(nblk(nbatch, 256)), dim3(256), 0, m->st, m->eloss, nbatch);
    e = hipGetLastError();
  }
  for (int b = 0; b < nbatch && e == hipSuccess; ++b) {
    hipLaunchKernelGGL(k_gather, dim3(nblk((long long)batch * IN_SIZE, 256)), dim3(256), 0, m->st, d->x, d->t,
                       m->perm + (size_t)b * batch, batch, m->act[0], m->tb);
    e = forward(m, batch);
    if (e == hipSuccess) e = backward(m, batch, m->eloss + b);
    if (e != hipSuccess) break;
    ++m->step;
    const float bc1 = 1.f - std::pow(0.9f, (float)m->step), bc2 = 1.f - std::pow(0.999f, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(NPARAM, 256)), dim3(256), 0, m->st, m->p, m->g, m->m, m->v, m->pbf,
                       NPARAM, lr, bc1, bc2);
    e = hipGetLastError();
  }
  std::vector<float> bl(std::max(1, nbatch), 0.f);
  if (e == hipSuccess && nbatch > 0)
    e = hipMemcpyAsync(bl.data(), m->eloss, (size_t)nbatch * 4, hipMemcpyDeviceToHost, m->st);
  if (e == hipSuccess) e = hipStreamSynchronize(m->st);  // perm (host) is released only after this
  double lsum = 0;
  for (int b = 0; b < nbatch; ++b) lsum += bl[b];
  if (mean_loss) *mean_loss = nbatch ? (float)(lsum / nbatch) : 0.f;
  return e == hipSuccess ? 0 : 2;
}

int jms_nn_predict(jms_nn* m, jms_nn_data* d, float* out) {
  if (!m || !d || !out) return 1;
  if (d->n < 1) return 0;
  // batches of up to max_batch windows read in place; every output goes to
  // one device array, copied back once
  float* dout = nullptr;
  hipError_t e = dalloc(&dout, (size_t)d->n * NOUT);
  for (int b0 = 0; b0 < d->n && e == hipSuccess; b0 += m->maxb) {
    const int nb = std::min(m->maxb, d->n - b0);
    e = forward(m, nb, d->x + (size_t)b0 * IN_SIZE, dout + (size_t)b0 * NOUT);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, (size_t)d->n * NOUT * 4, hipMemcpyDeviceToHost, m->st);
  if (e == hipSuccess) e = hipStreamSynchronize(m->st);
  if (dout) {
    (void)hipStreamSynchronize(m->st);
    (void)hipFree(dout);
  }
  return e == hipSuccess ? 0 : 2;
}

}  // extern "C"
