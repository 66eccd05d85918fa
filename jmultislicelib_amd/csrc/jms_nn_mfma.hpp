// jms_nn_mfma.hpp -- the surrogate's 3x3 convolutions as implicit GEMMs on
// the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16: bf16 operands, FP32
// accumulation), included by jms_nn.hip.
//
// One block of 256 threads (4 wavefronts) per sample.  The layer input (or,
// for the data gradient, the output gradient) is staged once into LDS in a
// channels-last layout with a one-pixel zero border, so that the 8 bf16
// values of an MFMA operand fragment -- 8 consecutive channels at one
// (pixel, tap) -- are one aligned 16-byte LDS read and the padding needs no
// bounds checks.  The GEMM K index runs over (tap, 8-channel group); weights
// are loaded once per wavefront into registers as A fragments.
//
// MFMA 16x16x32 operand layout (lane l of 64, g = l / 16): A fragment = row
// l % 16, K elements 8g .. 8g+7; B fragment = column l % 16, the same K
// elements; result = rows 4g .. 4g+3 of column l % 16.
#pragma once

namespace nnmfma {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int CH = 32;  // output (and hidden) channels of every convolution

__device__ __forceinline__ f32x4 mfma(const u16x8& a, const u16x8& b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                0, 0);
}

__device__ __forceinline__ uint16_t f2bf_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// y[n][co][oy][ox] = relu(b[co] + sum_{ci,ky,kx} x[n][ci][oy*S+ky-1][ox*S+kx-1] w[co][ci][ky][kx])
// grid: one block per sample
template <int S, int CI, int HIN>
__global__ __launch_bounds__(256) void k_conv_fwd(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
                                                  const float* __restrict__ b, uint16_t* __restrict__ y) {
  constexpr int HP = HIN + 2, HOUT = HIN / S, NPIX = HOUT * HOUT;
  constexpr int KG = 9 * CI / 8;          // K groups of 8 (tap, channel group)
  constexpr int KS = (KG + 3) / 4;        // K steps of 32
  constexpr int CG = CI / 8;
  __shared__ __align__(16) uint16_t xs[HP * HP * CI];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, g = lane >> 4, r = lane & 15;
  const long long n = blockIdx.x;
  // zero the whole staging area (border included), then the interior
  for (int i = tid; i < HP * HP * CI / 2; i += 256) reinterpret_cast<uint32_t*>(xs)[i] = 0u;
  __syncthreads();
  const uint16_t* xn = x + n * CI * HIN * HIN;
  for (int i = tid; i < CI * HIN * HIN; i += 256) {
    const int c = i / (HIN * HIN), p = i % (HIN * HIN), iy = p / HIN, ix = p % HIN;
    xs[((iy + 1) * HP + ix + 1) * CI + c] = xn[i];
  }
  // A fragments: w[co][ci0+j][tap] for co = mt*16 + r
  u16x8 af[2][KS];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int kg = ks * 4 + g;
      const int co = mt * 16 + r;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        uint16_t v = 0;
        if (kg < KG) {
          const int t = kg / CG, ci = (kg % CG) * 8 + j;
          v = w[(co * CI + ci) * 9 + t];
        }
        af[mt][ks][j] = v;
      }
    }
  __syncthreads();
  uint16_t* yn = y + n * CH * NPIX;
  for (int gi = wv; gi < NPIX / 16; gi += 4) {
    const int p = gi * 16 + r, oy = p / HOUT, ox = p % HOUT;
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int kg = ks * 4 + g;
      u16x8 bf = {0, 0, 0, 0, 0, 0, 0, 0};
      if (kg < KG) {
        const int t = kg / CG, c0 = (kg % CG) * 8, ky = t / 3, kx = t % 3;
        bf = *reinterpret_cast<const u16x8*>(&xs[((oy * S + ky) * HP + ox * S + kx) * CI + c0]);
      }
      acc0 = mfma(af[0][ks], bf, acc0);
      acc1 = mfma(af[1][ks], bf, acc1);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int co0 = 4 * g + i, co1 = 16 + 4 * g + i;
      yn[co0 * NPIX + p] = f2bf_rne(fmaxf(acc0[i] + b[co0], 0.f));
      yn[co1 * NPIX + p] = f2bf_rne(fmaxf(acc1[i] + b[co1], 0.f));
    }
  }
}

// dx[n][ci][iy][ix] = [a[n][ci][iy][ix] > 0] * sum_{co,ky,kx} dy[n][co][oy][ox] w[co][ci][ky][kx]
// over iy = oy*S + ky - 1, ix = ox*S + kx - 1 (CI = CH); dy is rounded to
// bf16 as it is staged.  grid: one block per sample
template <int S, int HIN>
__global__ __launch_bounds__(256) void k_conv_bwd_data(const float* __restrict__ dy, const uint16_t* __restrict__ w,
                                                       const uint16_t* __restrict__ a, float* __restrict__ dx) {
  constexpr int CI = CH, HOUT = HIN / S, HP = HOUT + 2, NPIN = HIN * HIN, NPOUT = HOUT * HOUT;
  constexpr int KS = 9;  // 36 K groups (tap, 8-channel group of co) in steps of 4
  __shared__ __align__(16) uint16_t gs[HP * HP * CH];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, g = lane >> 4, r = lane & 15;
  const long long n = blockIdx.x;
  for (int i = tid; i < HP * HP * CH / 2; i += 256) reinterpret_cast<uint32_t*>(gs)[i] = 0u;
  __syncthreads();
  const float* dyn = dy + n * CH * NPOUT;
  for (int i = tid; i < CH * NPOUT; i += 256) {
    const int c = i / NPOUT, p = i % NPOUT, oy = p / HOUT, ox = p % HOUT;
    gs[((oy + 1) * HP + ox + 1) * CH + c] = f2bf_rne(dyn[i]);
  }
  // A fragments: w[co0+j][ci][tap] for ci = mt*16 + r
  u16x8 af[2][KS];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int kg = ks * 4 + g, t = kg / 4, co0 = (kg % 4) * 8, ci = mt * 16 + r;
#pragma unroll
      for (int j = 0; j < 8; ++j) af[mt][ks][j] = w[((co0 + j) * CI + ci) * 9 + t];
    }
  __syncthreads();
  const uint16_t* an = a + n * CI * NPIN;
  float* dxn = dx + n * CI * NPIN;
  for (int gi = wv; gi < NPIN / 16; gi += 4) {
    const int p = gi * 16 + r, iy = p / HIN, ix = p % HIN;
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int kg = ks * 4 + g, t = kg / 4, co0 = (kg % 4) * 8, ky = t / 3, kx = t % 3;
      const int ty = iy + 1 - ky, tx = ix + 1 - kx;  // in [-1, HIN]
      int slot = 0;                                   // the zero border pixel (0, 0)
      if (S == 1 || ((ty & 1) == 0 && (tx & 1) == 0)) {
        const int oy = (S == 1) ? ty : (ty >> 1), ox = (S == 1) ? tx : (tx >> 1);  // in [-1, HOUT]
        slot = (oy + 1) * HP + ox + 1;
      }
      const u16x8 bf = *reinterpret_cast<const u16x8*>(&gs[slot * CH + co0]);
      acc0 = mfma(af[0][ks], bf, acc0);
      acc1 = mfma(af[1][ks], bf, acc1);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ci0 = 4 * g + i, ci1 = 16 + 4 * g + i;
      const long long o0 = (long long)ci0 * NPIN + p, o1 = (long long)ci1 * NPIN + p;
      dxn[o0] = (__uint_as_float((uint32_t)an[o0] << 16) > 0.f) ? acc0[i] : 0.f;
      dxn[o1] = (__uint_as_float((uint32_t)an[o1] << 16) > 0.f) ? acc1[i] : 0.f;
    }
  }
}

// per-sample partial weight gradient:
//   part[n][co][ci][tap] = sum_{oy,ox} dy[n][co][oy][ox] x[n][ci][oy*S+ky-1][ox*S+kx-1]
// (dy rounded to bf16).  The GEMM: rows co, columns c = ci*9 + tap, K the
// sample's output pixels.  grid: one block per sample
template <int S, int CI, int HIN>
__global__ __launch_bounds__(256) void k_conv_bwd_w(const float* __restrict__ dy, const uint16_t* __restrict__ x,
                                                    float* __restrict__ part) {
  constexpr int HP = HIN + 2, HOUT = HIN / S, NPOUT = HOUT * HOUT;
  constexpr int NCOL = CI * 9, NT = (NCOL + 15) / 16;  // column tiles
  constexpr int KS = (NPOUT + 31) / 32;                // K steps over the output pixels
  constexpr int NPAD = KS * 32;
  __shared__ __align__(16) uint16_t xs[CI * HP * HP];  // [ci][HP][HP], zero border
  __shared__ __align__(16) uint16_t gs[CH * NPAD];     // [co][p], bf16, zero beyond NPOUT
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, g = lane >> 4, r = lane & 15;
  const long long n = blockIdx.x;
  for (int i = tid; i < CI * HP * HP / 2; i += 256) reinterpret_cast<uint32_t*>(xs)[i] = 0u;
  for (int i = tid; i < CH * NPAD / 2; i += 256) reinterpret_cast<uint32_t*>(gs)[i] = 0u;
  __syncthreads();
  const uint16_t* xn = x + n * CI * HIN * HIN;
  for (int i = tid; i < CI * HIN * HIN; i += 256) {
    const int c = i / (HIN * HIN), p = i % (HIN * HIN), iy = p / HIN, ix = p % HIN;
    xs[(c * HP + iy + 1) * HP + ix + 1] = xn[i];
  }
  const float* dyn = dy + n * CH * NPOUT;
  for (int i = tid; i < CH * NPOUT; i += 256) gs[(i / NPOUT) * NPAD + i % NPOUT] = f2bf_rne(dyn[i]);
  __syncthreads();
  float* pn = part + n * CH * NCOL;
  for (int nt = wv; nt < NT; nt += 4) {
    const int c = nt * 16 + r;  // this lane's column
    const bool cok = c < NCOL;
    const int ci = cok ? c / 9 : 0, t = cok ? c % 9 : 0, ky = t / 3, kx = t % 3;
    const uint16_t* xc = xs + ci * HP * HP;
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < KS; ++ks) {
      const int k0 = ks * 32 + 8 * g;  // first output pixel of this lane's K elements
      const u16x8 a0 = *reinterpret_cast<const u16x8*>(&gs[r * NPAD + k0]);
      const u16x8 a1 = *reinterpret_cast<const u16x8*>(&gs[(16 + r) * NPAD + k0]);
      u16x8 bf;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int p = k0 + j;
        uint16_t v = 0;
        if (cok && p < NPOUT) {
          const int oy = p / HOUT, ox = p % HOUT;
          v = xc[(oy * S + ky) * HP + ox * S + kx];
        }
        bf[j] = v;
      }
      acc0 = mfma(a0, bf, acc0);
      acc1 = mfma(a1, bf, acc1);
    }
    if (cok) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        pn[(4 * g + i) * NCOL + c] = acc0[i];
        pn[(16 + 4 * g + i) * NCOL + c] = acc1[i];
      }
    }
  }
}

// dw[i] += sum over this block's share of the samples of part[n][i] (i <
// cnt); dw is cleared first.  grid: (ceil(cnt / 256), shares)
__global__ __launch_bounds__(256) void k_reduce_parts(const float* __restrict__ part, int nsamp, int cnt,
                                                      float* __restrict__ dw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const int k0 = (int)((long long)nsamp * blockIdx.y / gridDim.y);
  const int k1 = (int)((long long)nsamp * (blockIdx.y + 1) / gridDim.y);
  float s = 0.f;
  for (int k = k0; k < k1; ++k) s += part[(long long)k * cnt + i];
  atomicAdd(&dw[i], s);
}

// db[co] += sum over this block's share of the samples of dy[n][co][p];
// db is cleared first.  grid: (CH, shares)
__global__ __launch_bounds__(256) void k_bias_grad(const float* __restrict__ dy, int nsamp, int npix,
                                                   float* __restrict__ db) {
  const int co = blockIdx.x;
  const int k0 = (int)((long long)nsamp * blockIdx.y / gridDim.y);
  const int k1 = (int)((long long)nsamp * (blockIdx.y + 1) / gridDim.y);
  float s = 0.f;
  for (int k = k0; k < k1; ++k) {
    const float* d = dy + ((long long)k * CH + co) * npix;
    for (int p = threadIdx.x; p < npix; p += 256) s += d[p];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&db[co], red[0] + red[1] + red[2] + red[3]);
}

}  // namespace nnmfma
