// jms_kernels.hip -- size dispatch of the line-length templated kernels and
// the small non-templated kernels (readout combine, incident-wave offset,
// accumulation).
#include <hip/hip_runtime.h>
#include <mutex>
#include <vector>
#include <cstdint>
#include "jms_fft_device.hpp"
#include "jms_kernels.h"
#include "jms_sizes.h"
#include "jms_cfg.hpp"

namespace jms {

template <int N>
hipError_t launch_row_pass_n(int ny, int nwave, float2* wave, long long wave_stride,
                             const float2* pgr, long long slc_off, long long npg,
                             const int* var_tab, int var_ld, int jslc, const float2* twx,
                             const RowChanHost& rc, int mode, bool panel, hipStream_t st);
template <int N>
hipError_t launch_col_pass_n(int nwave, bool init, bool exit_, const ColArgsHost& h,
                             hipStream_t st);
template <int N>
hipError_t launch_row_fft_n(int ny, int nbatch, float2* data, long long stride, int dir,
                            const float2* twx, hipStream_t st);
template <int N>
hipError_t launch_col_fft_n(int nx, int nbatch, float2* data, long long stride, int dir,
                            const float2* twy, hipStream_t st);
template <int N>
hipError_t launch_row_pass_grp_n(int ny, int nwave, long long wave_stride, long long npg, const float2* twx,
                                 const GrpArgs& ga, bool panel, hipStream_t st);
template <int N>
hipError_t launch_col_pass_grp_n(int nx, int nwave, bool init, bool exit_, bool detect, long long part_dstride,
                                 const float2* twy, const GrpArgs& ga, bool panel, hipStream_t st);

// frequency index of position i on an n-point grid in FFT order
// (JMultiSlice.cpp:1699-1712)
__device__ __forceinline__ int fidx(int i, int n) {
  const int h = (n - (n & 1)) >> 1;
  return ((i + h) % n) - h;
}

// ---------------------------------------------------------------------------
// combine partial detector sums of a batch into per-position results:
//   out[(p0+p), idet, d] (+)= sum_r weight[r] * sum_tile partial[(p*npass+r), idet, d, tile]
// ---------------------------------------------------------------------------
__global__ void k_combine(const float* __restrict__ partial, int ntiles, int ndet, int ndetslc,
                          int npos, int npass, const float* __restrict__ weights,
                          float* __restrict__ out, int accumulate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = ndetslc * ndet;
  if (i >= npos * per) return;
  const int p = i / per, r0 = i % per;  // r0 = idet*ndet + d
  double acc = 0.0;
  for (int r = 0; r < npass; ++r) {
    const float* src = partial + ((long long)(p * npass + r) * per + r0) * ntiles;
    double s = 0.0;
    for (int k = 0; k < ntiles; ++k) s += (double)src[k];
    acc += (double)weights[r] * s;
  }
  out[i] = accumulate ? (float)((double)out[i] + acc) : (float)acc;
}

__global__ void k_zero_partial(float* __restrict__ partial, long long part_w, long long off, int len) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < len) partial[blockIdx.y * part_w + off + k] = 0.f;
}

// psi_out = psi0 * exp(-i [2 pi (kx dx + ky dy + k^2 fz) + cppl]) in Fourier space
// (OffsetIncomingWave GPU branch, JMultiSlice.cpp:2036-2077; ArrayOps.cu:390-422),
// with an optional cyclic beam-tilt shift of psi0 by (ish0, ish1) pixels
// (CShift2dKernel, ArrayOps.cu:340-352).
__global__ void k_offset_wave(float2* __restrict__ out, const float2* __restrict__ in, int nx,
                              int ny, float kscx, float kscy, float dx, float dy, float fz,
                              const float* __restrict__ cppl, int ish0, int ish1) {
  const long long n = (long long)nx * ny;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int i = (int)(idx % nx), j = (int)(idx / nx);
  const float kx = (float)fidx(i, nx) * kscx, ky = (float)fidx(j, ny) * kscy;
  const float ksqr = kx * kx + ky * ky;
  float chi = -6.28318531f * (dx * kx + dy * ky + fz * ksqr);
  if (cppl) chi -= cppl[idx];
  float sn, cs;
  sincosf(chi, &sn, &cs);
  int i0 = i - ish0, j0 = j - ish1;
  i0 %= nx; if (i0 < 0) i0 += nx;
  j0 %= ny; if (j0 < 0) j0 += ny;
  out[idx] = cmul(in[(long long)i0 + (long long)j0 * nx], make_float2(cs, sn));
}

// ---------------------------------------------------------------------------
// probe in Fourier space on the device: the aperture of ApertureFunctionA
// (JProbeGen.cpp:471-505) times exp(-i chi) with the axial aberration phase
// of AberrationFunction (JProbeGen.cpp:508-571), then normalised to unit
// intensity (CalculateProbeWaveFourier, JProbeGen.cpp:707-777).
// ---------------------------------------------------------------------------
__device__ float probe_chi(float qx, float qy, const ProbeArgsHost& a) {
  const float wx = qx * a.wl, wy = qy * a.wl;
  const float w = sqrtf(wx * wx + wy * wy);
  float px[9], py[9], pw[9];
  px[0] = py[0] = pw[0] = 1.f;
#pragma unroll
  for (int i = 1; i < 9; ++i) { px[i] = px[i - 1] * wx; py[i] = py[i - 1] * wy; pw[i] = pw[i - 1] * w; }
  const float re_il[4] = {1.f, 0.f, -1.f, 0.f};
  float sum = 0.f;
  for (int k = 0; k < a.nab; ++k) {
    const float cx = a.coeff[2 * k], cy = a.coeff[2 * k + 1];
    if (cx * cx + cy * cy > 1.0e-15f) {
      const int m = a.mn[2 * k], n = a.mn[2 * k + 1];
      float term = 0.f;
      for (int l = 0; l <= n; ++l) {
        float c = 0.f;
        c = c + re_il[l & 3] * cx;
        c = c + re_il[(l + 3) & 3] * cy;
        term = term + c * (float)a.binom[n + 9 * l] * px[n - l] * py[l];
      }
      sum = sum + term * pw[m - n] / (float)m;
    }
  }
  return sum * (6.28318530717958648f / a.wl);
}

__global__ void k_probe(float2* __restrict__ out, double* __restrict__ partial, ProbeArgsHost a) {
  __shared__ double red[256];
  const long long n = (long long)a.nx * a.ny;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double pw = 0.0;
  if (idx < n) {
    const int i = (int)(idx % a.nx), j = (int)(idx / a.nx);
    const float qx = (float)fidx(i, a.nx) / a.ax, qy = (float)fidx(j, a.ny) / a.ay;
    float amp = 0.f;
    if (a.qlim > 0.f && a.open) {
      const float dqx = (qx - a.btx) - a.qlcx, dqy = (qy - a.bty) - a.qlcy;
      const float ux = ((1.f - a.a1x) * dqx - a.a1y * dqy) * a.radet;
      const float uy = (-a.a1y * dqx + (1.f + a.a1x) * dqy) * a.radet;
      const float dqm = sqrtf(ux * ux + uy * uy);
      const float dqs = fabsf(a.qrs * a.qlim);
      if (dqm <= 0.f) amp = 1.f;
      else if (dqs > 0.f) amp = (1.f - tanhf(3.14159265358979324f * (dqm - a.qlim) / dqs)) * 0.5f;
      else if (dqm < a.qlim) amp = 1.f;
    }
    float2 v = make_float2(0.f, 0.f);
    if (amp > 1.0e-6f) {
      const float chi = a.nab > 0 ? probe_chi(qx, qy, a) : 0.f;
      float sn, cs;
      sincosf(chi, &sn, &cs);
      v = make_float2(amp * cs, -amp * sn);
      pw = (double)(amp * amp);
    }
    out[idx] = v;
  }
  red[threadIdx.x] = pw;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_probe_scale(float2* __restrict__ out, const double* __restrict__ partial, int nblk,
                              long long n) {
  __shared__ double red[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partial[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  const float fsca = 1.f / sqrtf((float)red[0]);
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) {
    float2 v = out[idx];
    v.x *= fsca;
    v.y *= fsca;
    out[idx] = v;
  }
}

// ---------------------------------------------------------------------------
// phase gratings of the Gaussian projected-potential model on the device
// (the host generator jms_gaussian_phase_grating, evaluated per pixel): each
// pixel adds amp * exp(-r^2 / (2 s^2)) of every atom image within the cut-off
// in the host's order (atoms ascending; per atom, pixel rows then columns of
// the unwrapped range), in double, then t = exp(i sig_int V).
// grid: (ceil(nx/16), ceil(ny/16), nvar), block: 16 x 16
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gauss_gratings(float2* __restrict__ out, int nx, int ny, double hx,
                                                        double hy, double rc2, double inv2s2, double sig_int,
                                                        const GaussAtom* __restrict__ atoms,
                                                        const int* __restrict__ off, const int* __restrict__ idx) {
  const int px = blockIdx.x * 16 + threadIdx.x, py = blockIdx.y * 16 + threadIdx.y;
  const int v = blockIdx.z;
  const int ntx = (nx + 15) / 16, nty = (ny + 15) / 16, ntiles = ntx * nty;
  const int t = blockIdx.y * ntx + blockIdx.x;
  const int* o = off + (long long)v * (ntiles + 1);
  const int a0 = o[t], a1 = o[t + 1];
  if (px >= nx || py >= ny) return;
  double pot = 0.0;
  for (int k = a0; k < a1; ++k) {
    const GaussAtom at = atoms[idx[k]];
    // unwrapped rows j = py + m ny and columns i = px + l nx inside the range
    int j = at.j0 + (((py - at.j0) % ny) + ny) % ny;
    for (; j <= at.j1; j += ny) {
      const double dy = j * hy - at.ya;
      const double dy2 = dy * dy;
      if (dy2 > rc2) continue;
      int i = at.i0 + (((px - at.i0) % nx) + nx) % nx;
      for (; i <= at.i1; i += nx) {
        const double dx = i * hx - at.xa;
        const double r2 = dx * dx + dy2;
        if (r2 > rc2) continue;
        pot += at.amp * exp(-r2 * inv2s2);
      }
    }
  }
  const double ph = sig_int * pot;
  out[(long long)v * nx * ny + (long long)py * nx + px] = make_float2((float)cos(ph), (float)sin(ph));
}

// ---------------------------------------------------------------------------
// phase gratings from projected atomic form factors (CalculatePhasegratingGPU,
// JEElaSca.cpp:978-1159; AddFormFactor2dKernel / PotToPgrKernel,
// scapot.cu:29-41,80-89).  Spatial frequencies in FFT order [1/nm],
// q = fidx(i) * (1 / a) (InitHelper, JEElaSca.cpp:529-541); every product
// and sum that decides an aperture edge is rounded as written, so that the
// edge pixels do not depend on contraction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float ff_freq(int i, int n, float itog) { return __fmul_rn((float)fidx(i, n), itog); }
__device__ __forceinline__ float ff_gmag(float qx, float qy) {
  return __fsqrt_rn(__fadd_rn(__fmul_rn(qx, qx), __fmul_rn(qy, qy)));
}

// Fourier-space projected potential of atom set v (atoms off[v] .. off[v+1],
// grouped by type, ascending atom order inside a type):
//   pot(q) = sum_t ff_t(q) ap(q) sum_{atoms of t} occ exp(-2 pi i (x qx + y qy))
// ap(q) = 1/2 - 1/2 tanh(30 (|q|/kmax - 0.9)) with the potential band limit,
// else 1 (JEElaSca.cpp:592-595).  The atoms pass through LDS 256 at a time;
// every thread walks the same list, so the type switches are uniform.
// grid: (ceil(nx*ny/256), nvar), block: 256
__global__ __launch_bounds__(256) void k_ff_sum(float2* __restrict__ out, const float2* __restrict__ ff,
                                                const FFAtom* __restrict__ atoms, const int* __restrict__ off,
                                                int nx, int ny, float itogx, float itogy, float kmax,
                                                int bwl_pot) {
  __shared__ FFAtom s_at[256];
  const long long n = (long long)nx * ny;
  const int v = blockIdx.y;
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  const bool valid = p < n;
  const int ix = valid ? (int)(p % nx) : 0, iy = valid ? (int)(p / nx) : 0;
  const float qx = ff_freq(ix, nx, itogx), qy = ff_freq(iy, ny, itogy);
  float ap = 1.f;
  if (bwl_pot) ap = 0.5f - 0.5f * tanhf((ff_gmag(qx, qy) / kmax - 0.9f) * 30.f);
  const int a0 = off[v], a1 = off[v + 1];
  float2 pot = make_float2(0.f, 0.f), s = make_float2(0.f, 0.f);
  int cur = -1;  // type of the running sum s
  for (int c0 = a0; c0 < a1; c0 += 256) {
    const int m = (a1 - c0 < 256) ? (a1 - c0) : 256;
    __syncthreads();  // the previous chunk's reads are done
    if ((int)threadIdx.x < m) s_at[threadIdx.x] = atoms[c0 + threadIdx.x];
    __syncthreads();
    for (int k = 0; k < m; ++k) {
      const FFAtom at = s_at[k];
      if (at.type != cur) {
        if (cur >= 0 && valid) {
          const float2 f = ff[(long long)cur * n + p];
          pot.x += ap * (f.x * s.x - f.y * s.y);
          pot.y += ap * (f.x * s.y + f.y * s.x);
        }
        s = make_float2(0.f, 0.f);
        cur = at.type;
      }
      // phase in turns, formed in double (x q reaches ~100 turns) and
      // reduced before the float sine and cosine
      double ud = (double)at.x * (double)qx + (double)at.y * (double)qy;
      ud -= rint(ud);
      float sn, cs;
      sincospif(2.f * (float)ud, &sn, &cs);
      s.x += at.occ * cs;
      s.y -= at.occ * sn;
    }
  }
  if (!valid) return;
  if (cur >= 0) {
    const float2 f = ff[(long long)cur * n + p];
    pot.x += ap * (f.x * s.x - f.y * s.y);
    pot.y += ap * (f.x * s.y + f.y * s.x);
  }
  out[(long long)v * n + p] = pot;
}

// a <- exp(i sg sc a): real-space potential (complex: the imaginary part
// absorbs) to transmission function, exp(-sg Im V) (cos sg Re V, sin sg Re V)
// with V = sc a (JEElaSca.cpp:1129-1141; PotToPgrKernel)
__global__ void k_pot_pgr(float2* __restrict__ a, long long n, float sc, float sg) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 p = a[i];
  const float vr = p.x * sc, vi = p.y * sc;
  const float m = expf(-sg * vi);
  float sn, cs;
  sincosf(sg * vr, &sn, &cs);
  a[i] = make_float2(m * cs, m * sn);
}

// a(q) <- a(q) * (|q| > kthr ? 0 : ftrn) on nvar Fourier-space gratings
// (phase-grating band limit, JEElaSca.cpp:597-604 with ftrn = 1/(nx ny))
__global__ void k_pgr_aperture(float2* __restrict__ a, int nx, int ny, long long total, float itogx,
                               float itogy, float kthr, float ftrn) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long p = i % ((long long)nx * ny);
  const int ix = (int)(p % nx), iy = (int)(p / nx);
  const float w = (ff_gmag(ff_freq(ix, nx, itogx), ff_freq(iy, ny, itogy)) > kthr) ? 0.f : ftrn;
  float2 z = a[i];
  z.x *= w;
  z.y *= w;
  a[i] = z;
}

// dst += w * src (complex), n items
__global__ void k_axpy_c(float2* __restrict__ dst, const float2* __restrict__ src, long long n, float w) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 s = src[i];
  float2 d = dst[i];
  d.x += w * s.x;
  d.y += w * s.y;
  dst[i] = d;
}

// ---------------------------------------------------------------------------
// dispatch over the compiled line lengths
// ---------------------------------------------------------------------------
bool size_supported(int n) {
  switch (n) {
#define JMS_CASE(N) case N: return true;
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return false;
  }
}

int row_tiles(int nx, int ny) {
  switch (nx) {
#define JMS_CASE(N) case N: return (ny + Cfg<N>::LPB - 1) / Cfg<N>::LPB;
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return 0;
  }
}

int col_tiles(int ny, int nx) {
  switch (ny) {
#define JMS_CASE(N) case N: return (nx + Cfg<N>::CC - 1) / Cfg<N>::CC;
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return 0;
  }
}

hipError_t launch_row_pass(int nx, int ny, int nwave, float2* wave, long long wave_stride,
                           const float2* pgr, long long slc_off, long long npg, const int* var_tab,
                           int var_ld, int jslc, const float2* twx, const RowChanHost& rc, int mode,
                           bool panel, hipStream_t st) {
  switch (nx) {
#define JMS_CASE(N)                                                                         \
  case N:                                                                                   \
    return launch_row_pass_n<N>(ny, nwave, wave, wave_stride, pgr, slc_off, npg, var_tab, \
                                var_ld, jslc, twx, rc, mode, panel, st);
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_col_pass(int ny, int nwave, bool init, bool exit_, const ColArgsHost& a,
                           hipStream_t st) {
  switch (ny) {
#define JMS_CASE(N) case N: return launch_col_pass_n<N>(nwave, init, exit_, a, st);
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return hipErrorInvalidValue;
  }
}

bool panel_ok(int nx, int ny) {
  bool rx = false, cy = false;
  switch (nx) {
#define JMS_CASE(N) case N: rx = Cfg<N>::PANEL; break;
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return false;
  }
  switch (ny) {
#define JMS_CASE(N) case N: cy = Cfg<N>::PANEL; break;
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return false;
  }
  return rx && cy;
}

bool grp_supported(int nx, int ny) {
  if (!size_supported(nx)) return false;
  switch (ny) {
#define JMS_CASE(N) case N: return Cfg<N>::WL || Cfg<N>::XL;
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return false;
  }
}

hipError_t launch_row_pass_grp(int nx, int ny, int nwave, long long wave_stride, long long npg,
                               const float2* twx, const GrpArgs& ga, bool panel, hipStream_t st) {
  switch (nx) {
#define JMS_CASE(N) case N: return launch_row_pass_grp_n<N>(ny, nwave, wave_stride, npg, twx, ga, panel, st);
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_col_pass_grp(int nx, int ny, int nwave, bool init, bool exit_, bool detect,
                               long long part_dstride, const float2* twy, const GrpArgs& ga,
                               bool panel, hipStream_t st) {
  switch (ny) {
#define JMS_CASE(N) \
  case N: return launch_col_pass_grp_n<N>(nx, nwave, init, exit_, detect, part_dstride, twy, ga, panel, st);
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_col_fft(int nx, int ny, int nbatch, float2* data, int dir, const float2* twy,
                          hipStream_t st) {
  const long long stride = (long long)nx * ny;
  switch (ny) {
#define JMS_CASE(N) case N: return launch_col_fft_n<N>(nx, nbatch, data, stride, dir, twy, st);
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fft2d(int nx, int ny, int nbatch, float2* data, int dir, const float2* twx,
                        const float2* twy, hipStream_t st) {
  const long long stride = (long long)nx * ny;
  hipError_t e = hipErrorInvalidValue;
  switch (nx) {
#define JMS_CASE(N) case N: e = launch_row_fft_n<N>(ny, nbatch, data, stride, dir, twx, st); break;
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  switch (ny) {
#define JMS_CASE(N) case N: return launch_col_fft_n<N>(nx, nbatch, data, stride, dir, twy, st);
    JMS_SIZES(JMS_CASE)
#undef JMS_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_combine(const float* partial, int ntiles, int ndet, int ndetslc, int npos,
                          int npass, const float* weights, float* out, int accumulate,
                          hipStream_t st) {
  const int total = npos * ndetslc * ndet;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_combine, dim3((total + 255) / 256), dim3(256), 0, st, partial, ntiles, ndet,
                     ndetslc, npos, npass, weights, out, accumulate);
  return hipGetLastError();
}

hipError_t launch_zero_partial(float* partial, long long part_w, long long off, int len, int nwave,
                               hipStream_t st) {
  if (len <= 0 || nwave <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_partial, dim3((len + 255) / 256, nwave), dim3(256), 0, st, partial, part_w, off, len);
  return hipGetLastError();
}

hipError_t launch_offset_wave(float2* out, const float2* in, int nx, int ny, float kscx,
                              float kscy, float dx, float dy, float fz, const float* cppl,
                              int ish0, int ish1, hipStream_t st) {
  const long long n = (long long)nx * ny;
  hipLaunchKernelGGL(k_offset_wave, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, in,
                     nx, ny, kscx, kscy, dx, dy, fz, cppl, ish0, ish1);
  return hipGetLastError();
}

hipError_t launch_gauss_gratings(float2* out, int nx, int ny, int nvar, double hx, double hy,
                                 double rc2, double inv2s2, double sig_int, const GaussAtom* atoms,
                                 const int* off, const int* idx, hipStream_t st) {
  dim3 g((nx + 15) / 16, (ny + 15) / 16, nvar), b(16, 16);
  hipLaunchKernelGGL(k_gauss_gratings, g, b, 0, st, out, nx, ny, hx, hy, rc2, inv2s2, sig_int, atoms, off,
                     idx);
  return hipGetLastError();
}

hipError_t launch_ff_gratings(float2* out, int nx, int ny, int nvar, const float2* ff, const FFAtom* atoms,
                              const int* off, float itogx, float itogy, float kmax, int bwl_pot, float sc,
                              float sg, int bwl_pgr, float kthr, float ftrn, const float2* twx,
                              const float2* twy, hipStream_t st) {
  const long long n = (long long)nx * ny, total = n * nvar;
  hipLaunchKernelGGL(k_ff_sum, dim3((unsigned)((n + 255) / 256), nvar), dim3(256), 0, st, out, ff, atoms, off, nx,
                     ny, itogx, itogy, kmax, bwl_pot);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = launch_fft2d(nx, ny, nvar, out, +1, twx, twy, st);
  if (e != hipSuccess) return e;
  const dim3 g((unsigned)((total + 255) / 256)), b(256);
  hipLaunchKernelGGL(k_pot_pgr, g, b, 0, st, out, total, sc, sg);
  e = hipGetLastError();
  if (e != hipSuccess || !bwl_pgr) return e;
  e = launch_fft2d(nx, ny, nvar, out, -1, twx, twy, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_pgr_aperture, g, b, 0, st, out, nx, ny, total, itogx, itogy, kthr, ftrn);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_fft2d(nx, ny, nvar, out, +1, twx, twy, st);
}

hipError_t launch_probe(float2* out, double* partial, const ProbeArgsHost& a, hipStream_t st) {
  const long long n = (long long)a.nx * a.ny;
  const unsigned nblk = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_probe, dim3(nblk), dim3(256), 0, st, out, partial, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_probe_scale, dim3(nblk), dim3(256), 0, st, out, (const double*)partial, (int)nblk, n);
  return hipGetLastError();
}

hipError_t launch_axpy_c(float2* dst, const float2* src, long long n, float w, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy_c, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, src, n, w);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free batch entry (jms_hip_run_batch_noalloc):
// the readout partial sums of one batch.  A file-scope device array, so each
// GPU the library runs on holds its own copy until the process ends; host
// code must use the address scratch_address() resolves for the current
// device, never the array's name.
// ---------------------------------------------------------------------------
static __device__ __align__(256) float scratch[kScratchFloats];

hipError_t scratch_address(int device, float** out) {
  static std::mutex mu;
  static std::vector<float*> addr;  // per device, resolved once
  std::lock_guard<std::mutex> lock(mu);
  if (device < 0) return hipErrorInvalidDevice;
  if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
  if (!addr[device]) {
    int cur = -1;
    hipError_t err = hipGetDevice(&cur);
    if (err == hipSuccess && cur != device) err = hipSetDevice(device);
    float* p = nullptr;
    if (err == hipSuccess) err = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
    if (err != hipSuccess) return err;
    addr[device] = p;
  }
  *out = addr[device];
  return hipSuccess;
}

}  // namespace jms
