// jms_kernels_impl.hpp -- HIP kernels of the batched multislice slice loop (gfx950).
//
// The wave functions of a batch of probe passes are stored one after the
// other in HBM, each as ny rows of nx complex64 values (index x + y*nx, x
// fastest; Fourier-space arrays in FFT order), exactly the layout of the
// reference's device wave m_d_wav (JMultiSlice.cpp:1772-1797).
//
// One slice of the reference loop (JMultiSlice.cpp:4590-4651) is
//     [readout] -> IFT -> * grating -> FT -> * propagator
// and is executed here as two fused passes over the wave:
//   row pass    : per row y:     IFFT_x, * t(x,y), FFT_x           (k_row_pass)
//   column pass : per column kx: FFT_y, * P(kx,ky), [readout], IFFT_y (k_col_pass)
// The column pass of slice j also performs the readout of plane j+1 and the
// first half (IFFT_y) of slice j+1's inverse transform, so each slice costs two
// read+write sweeps of the wave.  The first column pass of a run (INIT) builds
// the incident wave from the probe and the per-position phase plate
// (OffsetIncomingWave, JMultiSlice.cpp:1961-2084; ArrayOps.cu:390-422); the
// last one (EXIT) stops after the exit-plane readout and leaves psi(k).
//
// The kernels are templates on the line length; jms_kernels_size.hip
// instantiates them once per supported length (one translation unit each).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "jms_fft_device.hpp"
#include "jms_kernels.h"
#include "jms_cfg.hpp"

namespace jms {

constexpr int kMaxDet = 8;  // detectors handled per readout sweep

struct SyncBlock {
  __device__ __forceinline__ void operator()() const { __syncthreads(); }
};

// Barrier for LDS regions private to one wavefront: the wave's earlier LDS
// operations complete (they are processed in issue order) before later ones
// issue; the memory clobber keeps the compiler from moving LDS accesses across.
struct SyncWave {
  __device__ __forceinline__ void operator()() const { __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }
};

// line-major LDS layout for column tiles: column c at c*STRIDE, element e at
// e + (e >> PADSH); STRIDE = 2 (mod 32) so that the 16 columns of a staging
// access (c fastest) fall on distinct bank pairs.
template <int N, int PADSH, int SW = 0>
struct ColLayout : LdsPos<SW, PADSH> {
  using B = LdsPos<SW, PADSH>;
  static constexpr int base = B::span(N);
  static constexpr int STRIDE = base + ((34 - (base % 32)) % 32);
  __device__ static __forceinline__ int idx(int line, int e) { return line * STRIDE + B::pos(e); }
  static constexpr int words(int lines) { return lines * STRIDE; }
};

// index of i - d on an n-point periodic axis (0 <= i < n, any d)
__device__ __forceinline__ int wrap_sub(int i, int d, int n) {
  int r = (i - d) % n;
  return r < 0 ? r + n : r;
}

// frequency index of position i on an n-point grid in FFT order
// (JMultiSlice.cpp:1699-1712)
__device__ __forceinline__ int fidx(int i, int n) {
  const int h = (n - (n & 1)) >> 1;
  return ((i + h) % n) - h;
}

// Block coordinates (bx, by) of a launch.  The hardware hands block
// b = blockIdx.x + gridDim.x * blockIdx.y to XCD b % 8, and each of the 8
// XCDs has its own L2.  With XCD, the blocks one XCD receives take
// consecutive logical ids and the logical id runs over by fastest, so an XCD
// works through all block rows by (waves) of a block column bx back to back:
// the grating row or propagator tile of bx enters that XCD's L2 once and
// serves every wave.  A bijection of the grid for any size; identity
// without XCD.
__device__ __forceinline__ unsigned xcd_block_id() {
  const unsigned nb = gridDim.x * gridDim.y;
  const unsigned b = blockIdx.x + gridDim.x * blockIdx.y;
  const unsigned x8 = b & 7u, q = nb >> 3, r = nb & 7u;
  return x8 * q + (x8 < r ? x8 : r) + (b >> 3);
}
template <bool XCD>
__device__ __forceinline__ void block_xy(int& bx, int& by) {
  if constexpr (XCD) {
    const unsigned id = xcd_block_id(), nby = gridDim.y;
    bx = (int)(id / nby);
    by = (int)(id % nby);
  } else {
    bx = (int)blockIdx.x;
    by = (int)blockIdx.y;
  }
}
// The same consecutive ids per XCD over groups of G block rows: within a
// group the id runs over by fastest, then bx, so the blocks in flight on one
// XCD are G block rows (waves) of a few neighbouring block columns (tiles):
// with G = 1 tiles that share 128-byte lines meet in one L2, with G > 1 a
// tile's propagator columns serve G waves from it.  Group gi starts at block
// column (gi * ROT) mod gridDim.x (a bijection of each group), so that the
// XCDs, each on its own groups, work at different column offsets.
template <int G, int ROT>
__device__ __forceinline__ void block_xy_groups(int& bx, int& by) {
  const unsigned id = xcd_block_id(), nbx = gridDim.x, nby = gridDim.y;
  const unsigned gi = id / (G * nbx), rem = id % (G * nbx);
  const unsigned g0 = gi * G, ge = (nby - g0 < (unsigned)G) ? nby - g0 : (unsigned)G;
  by = (int)(g0 + rem % ge);
  bx = (int)((rem / ge + gi * ROT) % nbx);
}

// Panel layout of a wave (the staged batches where the line length allows
// it): panels of kPanel columns, each NY rows x kPanel contiguous, so that a
// column tile is one contiguous span and a row is NX / kPanel segments of
// 128 bytes.  Element (y, x) sits at panel_off<NY>(x) + kPanel * y.
constexpr int kPanel = 16;
template <int NY>
__device__ __forceinline__ int panel_off(int x) {
  return (x & ~(kPanel - 1)) * NY + (x & (kPanel - 1));
}

// real-space readout channels of one detection plane (per wave: weight
// cwgt[w], channel stride chan_wstride items, plane offset chan_off items)
struct RowChan {
  float* img;
  float* img_avg;
  float2* wfr;
  float2* wfr_avg;
  const float* cwgt;
  long long chan_wstride;
  long long chan_off;
};

// ---------------------------------------------------------------------------
// row pass: wave[w] rows <- FFT_x( t[jslc, var] * IFFT_x(rows) )
//   MODE 0: the slice;
//   MODE 1: the slice, with the real-space readout of the plane before it
//           (ReadoutImgDet_d, JMultiSlice.cpp:4369-4465) between IFFT_x and
//           the grating multiply, where the wave is IFFT2 of the plane's
//           Fourier-space wave (the column pass did IFFT_y);
//   MODE 2: IFFT_x and the readout only (exit plane after a column IFFT);
//           the wave is not written back.
//   GRP:    grouped launch (MODE 0): block row y is wave ga.w0 + y of the
//           wave map; wave buffer, variant table, gratings and the slice of
//           step jslc come from its job's tables.
// grid: (ceil(ny/LPB), nwave); block: T*LPB
// ---------------------------------------------------------------------------
template <int N, int MODE = 0, class C_ = Cfg<N>, bool GRP = false, bool PANEL = false>
__global__ __launch_bounds__(C_::T * C_::LPB) __attribute__((amdgpu_waves_per_eu(C_::WPE_ROW))) void k_row_pass(
    float2* __restrict__ wave, int ny, long long wave_stride, const float2* __restrict__ pgr,
    long long slc_off, long long npg, const int* __restrict__ var_tab, int var_ld, int jslc,
    const float2* __restrict__ twg, RowChan rc, GrpArgs ga) {
  static_assert(!GRP || MODE == 0, "grouped row passes carry no real-space readout");
  static_assert(!PANEL || (MODE == 0 && N % kPanel == 0), "panel waves: whole panels, staged batches only");
  using P = typename C_::P;
  constexpr int T = C_::T, LPB = C_::LPB;
  using L = LineLayout<N, C_::PADSH_ROW, C_::SW>;
  constexpr int NREG = regs_needed<N, T, P>();
  // twiddles from an LDS copy of the table, or (GTW_ROW) read through the
  // caches from the global table
  __shared__ float2 s_tw[C_::GTW_ROW ? 1 : N];
  __shared__ float2 s_buf[L::words(LPB)];
  const int tid = threadIdx.x;
  const int t = tid % T, line = tid / T;
  if constexpr (!C_::GTW_ROW)
    for (int e = tid; e < N; e += T * LPB) s_tw[e] = twg[e];
  const float2* tw = C_::GTW_ROW ? twg : s_tw;
  int bx, by;
  block_xy<C_::XCD_ROW>(bx, by);
  long long w = by;
  if constexpr (GRP) {  // this wave's job: its buffers and the slice of step jslc
    const GrpWave gw = ga.wav[ga.w0 + by];
    const GrpJob& jb = ga.job[gw.job];
    w = gw.w;
    wave = (float2*)jb.wave;
    var_tab = jb.var;
    var_ld = jb.nslc;
    pgr = (const float2*)jb.pgr;
    slc_off = ga.pl[(long long)gw.job * ga.pl_ld + jslc].slc_off;
  }
  const int row = bx * LPB + line;
  const bool valid = row < ny;
  // PANEL: element x of the row at (x / 16) * 16 ny + x % 16 from the row start
  float2* wrow = wave + w * wave_stride + (long long)row * (PANEL ? kPanel : N);
  auto wx = [&](int x) { return PANEL ? (x & ~(kPanel - 1)) * ny + (x & (kPanel - 1)) : x; };
  const int var = MODE == 2 ? 0 : var_tab[w * var_ld + jslc];
  const float2* grow = pgr + slc_off + (long long)var * npg + (long long)row * N;
  using M0 = IOMap<N, T, P, false>;
  using M1 = IOMap<N, T, P, true>;
  float2 v[NREG];
  if (valid) {
    static_for<0, M0::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M0::FULL || b < M0::NBF) {
        static_for<0, M0::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          v[i * M0::R0 + q] = ldg<C_::NT_ROW>(wrow + wx(b + q * M0::NBF));
        });
      }
    });
  }
  if constexpr (!C_::GTW_ROW) __syncthreads();  // twiddle table complete
  // rows lie inside one wavefront when T divides 64: wave-level barriers suffice
  using RowSync = std::conditional_t<(64 % T == 0), SyncWave, SyncBlock>;
  LineFFT<N, T, +1, P, false, L, NREG, C_::TWR_ROW>::run(v, s_buf, line, tw, t, RowSync{});
  if constexpr (MODE >= 1) {
    if (valid) {  // weight * |psi(r)|^2 and weight * psi(r) into the channels of this plane
      const float wg = rc.cwgt[w];
      const long long cb = w * rc.chan_wstride + rc.chan_off + (long long)row * N;
      static_for<0, M1::NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const int b = t + i * T;
        if (M1::FULL || b < M1::NBF) {
          static_for<0, M1::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            const long long idx = cb + b + q * M1::NBF;
            const float2 z = v[i * M1::R0 + q];
            const float pz = z.x * z.x + z.y * z.y;
            if (rc.img) rc.img[idx] += wg * pz;
            if (rc.img_avg) rc.img_avg[idx] += wg * pz;
            if (rc.wfr) { float2 o = rc.wfr[idx]; o.x += z.x * wg; o.y += z.y * wg; rc.wfr[idx] = o; }
            if (rc.wfr_avg) { float2 o = rc.wfr_avg[idx]; o.x += z.x * wg; o.y += z.y * wg; rc.wfr_avg[idx] = o; }
          });
        }
      });
    }
  }
  if constexpr (MODE == 2) return;
  if (valid) {
    static_for<0, M1::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M1::FULL || b < M1::NBF) {
        static_for<0, M1::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          v[i * M1::R0 + q] = cmul(v[i * M1::R0 + q], grow[b + q * M1::NBF]);
        });
      }
    });
  }
  LineFFT<N, T, -1, P, true, L, NREG, C_::TWR_ROW>::run(v, s_buf, line, tw, t, RowSync{});
  if (valid) {
    static_for<0, M0::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M0::FULL || b < M0::NBF) {
        static_for<0, M0::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          stg<C_::NT_ROW>(wrow + wx(b + q * M0::NBF), v[i * M0::R0 + q]);
        });
      }
    });
  }
}

// ---------------------------------------------------------------------------
// block reduction of kMaxDet per-thread partial sums; thread 0 writes
// partial[d * dstride] for d < ndet.  red: LDS scratch of >= 64*kMaxDet floats.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_reduce_det(float (&acc)[kMaxDet], int ndet, float* red,
                                                 float* out, long long dstride) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int d = 0; d < kMaxDet; ++d) {
    float x = acc[d];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    acc[d] = x;
  }
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < kMaxDet; ++d) red[wv * kMaxDet + d] = acc[d];
  }
  __syncthreads();
  if (threadIdx.x < ndet && threadIdx.x < kMaxDet) {
    const int d = threadIdx.x;
    float s = 0.f;
    for (int i = 0; i < nw; ++i) s += red[i * kMaxDet + d];
    out[d * dstride] = s;
  }
}

// ---------------------------------------------------------------------------
// column pass.
//   INIT : psi <- src (shared probe, or one wave per batch entry) times the
//          phase plate exp(-i[2 pi (kx dx + ky dy + k^2 fz) + cppl]) when ppar
//          is given (MulPhasePlateC01Kernel, ArrayOps.cu:406-422)
//   else : psi <- P * FFT_y(wave)
//   detect: partial[(w, idet, d, tile)] = sum_{column tile} |psi|^2 * detw[d]
//   EXIT : store psi(k); else store IFFT_y(psi)
// grid: (ceil(nx/C), nwave); block: TC*C
// ---------------------------------------------------------------------------
struct ColArgs {
  int dsx, dsy;             // diffraction de-scan [pixel]: readouts see psi(i + d)
  float* dif_avg;           // averaged diffraction channel (or nullptr)
  float2* wff_avg;          // averaged Fourier-space wave channel (or nullptr)
  float2* wave;
  int nx;
  long long wave_stride;
  const float2* pro;        // propagator of this slice (not INIT)
  const float2* src;        // INIT source
  long long src_stride;     // 0: shared probe; n: per-wave input
  const float4* ppar;       // INIT: per-wave (dx, dy, fz, 0) or nullptr
  float kscx, kscy;         // 1/a0x, 1/a0y [1/nm]
  const float* cppl;        // constant phase plate (or nullptr)
  int detect;               // 1: readout at this plane
  int ndet;
  const float* detw;        // [ndet][n] effective detector weights
  float* partial;           // partial sums
  long long part_wstride;   // per-wave stride in partial
  long long part_dstride;   // per-detector stride (= ntiles)
  long long part_off;       // plane offset (idetslc * ndet * ntiles)
  float* dif;               // diffraction channel (or nullptr)
  float2* wff;              // Fourier-space wave channel (or nullptr)
  const float* cwgt;        // per-wave channel weight
  long long chan_wstride;   // per-wave stride of the channels (items)
  long long chan_off;       // plane offset in the channels (items)
  const float2* twg;
  // column-major copies used by the wave-local column kernel: [x][y]
  const float2* pro_t;      // propagator of this slice
  const float* detw_t;      // [ndet][nx][ny] detector weights
  const float* cppl_t;      // constant phase plate
  GrpArgs grp;              // grouped launch: wave map and job / plane tables
};

template <int N, bool INIT, bool EXIT, class C_ = Cfg<N>>
__global__ __launch_bounds__(C_::TC * C_::C) void k_col_pass(ColArgs a) {
  using P = typename C_::PC;
  constexpr int T = C_::TC, C = C_::C;
  using L = TileLayout<N, C, C_::PADSH>;
  constexpr int NREG = regs_needed<N, T, P>();
  __shared__ float2 s_tw[N];
  __shared__ float2 s_buf[L::words(1)];
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  for (int e = tid; e < N; e += T * C) s_tw[e] = a.twg[e];
  const long long w = blockIdx.y;
  const int nx = a.nx;
  const int x = blockIdx.x * C + c;
  const bool valid = x < nx;
  float2* wcol = a.wave + w * a.wave_stride + x;
  using M0 = IOMap<N, T, P, false>;
  using M1 = IOMap<N, T, P, true>;
  float2 v[NREG];
  if constexpr (INIT) {
    if (valid) {
      const float2* s = a.src + w * a.src_stride + x;
      float dx = 0.f, dy = 0.f, fz = 0.f;
      const bool pp = a.ppar != nullptr;
      if (pp) {
        const float4 q4 = a.ppar[w];
        dx = q4.x; dy = q4.y; fz = q4.z;
      }
      const float kx = (float)fidx(x, nx) * a.kscx;
      static_for<0, M1::NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const int b = t + i * T;
        if (M1::FULL || b < M1::NBF) {
          static_for<0, M1::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            const int e = b + q * M1::NBF;
            float2 val = s[(long long)e * nx];
            if (pp) {
              const float ky = (float)fidx(e, N) * a.kscy;
              const float ksqr = kx * kx + ky * ky;
              float chi = -6.28318531f * (dx * kx + dy * ky + fz * ksqr);
              if (a.cppl) chi -= a.cppl[(long long)e * nx + x];
              float sn, cs;
              sincosf(chi, &sn, &cs);
              val = cmul(val, make_float2(cs, sn));
            }
            v[i * M1::R0 + q] = val;
          });
        }
      });
    }
    __syncthreads();
  } else {
    if (valid) {
      static_for<0, M0::NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const int b = t + i * T;
        if (M0::FULL || b < M0::NBF) {
          static_for<0, M0::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            v[i * M0::R0 + q] = wcol[(long long)(b + q * M0::NBF) * nx];
          });
        }
      });
    }
    __syncthreads();
    LineFFT<N, T, -1, P, false, L, NREG, C_::TWR>::run(v, s_buf, c, s_tw, t, SyncBlock{});
    if (valid) {
      const float2* pcol = a.pro + x;
      static_for<0, M1::NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const int b = t + i * T;
        if (M1::FULL || b < M1::NBF) {
          static_for<0, M1::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            v[i * M1::R0 + q] = cmul(v[i * M1::R0 + q], pcol[(long long)(b + q * M1::NBF) * nx]);
          });
        }
      });
    }
  }
  if (a.detect) {
    float acc[kMaxDet];
#pragma unroll
    for (int d = 0; d < kMaxDet; ++d) acc[d] = 0.f;
    if (valid) {
      const long long n = (long long)N * nx;
      // de-scan (DescanDifN, JMultiSlice.cpp:2184-2275): the detectors integrate
      // psi(i + d), i.e. pixel i is weighted with det(i - d)
      const int xs = wrap_sub(x, a.dsx, nx);
      for (int d = 0; d < a.ndet && d < kMaxDet; ++d) {  // uniform loop over detectors
        const float* dw = a.detw + d * n + xs;
        float wt[NREG];
        static_for<0, M1::NB>([&](auto I) {  // all weight loads of this detector first
          constexpr int i = decltype(I)::value;
          const int b = t + i * T;
          static_for<0, M1::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            wt[i * M1::R0 + q] = (M1::FULL || b < M1::NBF)
                                     ? dw[(long long)wrap_sub(b + q * M1::NBF, a.dsy, N) * nx] : 0.f;
          });
        });
        float sacc = 0.f;
        static_for<0, M1::NB>([&](auto I) {
          constexpr int i = decltype(I)::value;
          const int b = t + i * T;
          if (M1::FULL || b < M1::NBF) {
            static_for<0, M1::R0>([&](auto Q) {
              constexpr int q = decltype(Q)::value;
              const float2 z = v[i * M1::R0 + q];
              sacc += (z.x * z.x + z.y * z.y) * wt[i * M1::R0 + q];
            });
          }
        });
#pragma unroll
        for (int dd = 0; dd < kMaxDet; ++dd)
          if (dd == d) acc[dd] += sacc;
      }
    }
    block_reduce_det(acc, a.ndet, (float*)s_buf,
                     a.partial + w * a.part_wstride + a.part_off + blockIdx.x, a.part_dstride);
    __syncthreads();
    if (valid && (a.dif || a.wff || a.dif_avg || a.wff_avg)) {  // ReadoutDifDet_d channels (JMultiSlice.cpp:4310-4356)
      const float wg = a.cwgt[w];
      const long long cb = w * a.chan_wstride + a.chan_off;
      static_for<0, M1::NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const int b = t + i * T;
        if (M1::FULL || b < M1::NBF) {
          static_for<0, M1::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            const long long idx = cb + (long long)wrap_sub(b + q * M1::NBF, a.dsy, N) * nx + wrap_sub(x, a.dsx, nx);
            const float2 z = v[i * M1::R0 + q];
            const float pz = z.x * z.x + z.y * z.y;
            if (a.dif) a.dif[idx] += wg * pz;
            if (a.dif_avg) a.dif_avg[idx] += wg * pz;
            if (a.wff) { float2 o = a.wff[idx]; o.x += z.x * wg; o.y += z.y * wg; a.wff[idx] = o; }
            if (a.wff_avg) { float2 o = a.wff_avg[idx]; o.x += z.x * wg; o.y += z.y * wg; a.wff_avg[idx] = o; }
          });
        }
      });
    }
  }
  if constexpr (EXIT) {
    if (valid) {
      static_for<0, M1::NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const int b = t + i * T;
        if (M1::FULL || b < M1::NBF) {
          static_for<0, M1::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            wcol[(long long)(b + q * M1::NBF) * nx] = v[i * M1::R0 + q];
          });
        }
      });
    }
  } else {
    LineFFT<N, T, +1, P, true, L, NREG, C_::TWR>::run(v, s_buf, c, s_tw, t, SyncBlock{});
    if (valid) {
      static_for<0, M0::NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const int b = t + i * T;
        if (M0::FULL || b < M0::NBF) {
          static_for<0, M0::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            wcol[(long long)(b + q * M0::NBF) * nx] = v[i * M0::R0 + q];
          });
        }
      });
    }
  }
}


// Tile of C adjacent columns x N rows staged through LDS (column c at
// L::idx(c, e)); thread tid moves elements i = tid + k*BS, i.e. row i / C,
// column i % C, so a wavefront covers 64/C whole 8*C-byte row segments.
template <int N, class C_>
struct ColTile {
  static constexpr int T = C_::TW, C = C_::CW, BS = T * C;
  static constexpr int KST = (N * C + BS - 1) / BS;
  static constexpr bool STFULL = (N * C) % BS == 0;
  using L = ColLayout<N, C_::PADSH_COL, C_::SW>;
  // global -> registers (columns past the grid edge read column 0 of the
  // tile and are discarded)
  static __device__ __forceinline__ void load(float2 (&st)[KST], const float2* src, int nx, int ncol,
                                              int tid) {
    static_for<0, KST>([&](auto K) {
      constexpr int k = decltype(K)::value;
      const int i = tid + k * BS;
      if (STFULL || i < N * C) {
        const int e = i / C, cc = i % C;
        st[k] = ldg<C_::NT_COL>(src + e * nx + (cc < ncol ? cc : 0));
      }
    });
  }
  static __device__ __forceinline__ void to_lds(const float2 (&st)[KST], float2* s_buf, int tid) {
    static_for<0, KST>([&](auto K) {
      constexpr int k = decltype(K)::value;
      const int i = tid + k * BS;
      if (STFULL || i < N * C) s_buf[L::idx(i % C, i / C)] = st[k];
    });
  }
  static __device__ __forceinline__ void from_lds(float2 (&st)[KST], const float2* s_buf, int tid) {
    static_for<0, KST>([&](auto K) {
      constexpr int k = decltype(K)::value;
      const int i = tid + k * BS;
      if (STFULL || i < N * C) st[k] = s_buf[L::idx(i % C, i / C)];
    });
  }
  static __device__ __forceinline__ void store(const float2 (&st)[KST], float2* dst, int nx, int ncol,
                                               int tid) {
    static_for<0, KST>([&](auto K) {
      constexpr int k = decltype(K)::value;
      const int i = tid + k * BS;
      if ((STFULL || i < N * C) && i % C < ncol) stg<C_::NT_COL>(dst + (i / C) * nx + i % C, st[k]);
    });
  }
};

// The per-column work of the column pass on a tile already in LDS, done by
// the wavefront lanes of column c (thread t of T): forward FFT along y (or,
// for INIT, the phase plate on the probe), * propagator, readout, inverse
// FFT along y (unless EXIT); the result is left in the column's LDS region.
// Only wave-level barriers inside, except for the readout's block reduction.
// FIN: the forward transform's first stage was done by the caller (its
// output is in LDS), so the body resumes at stage 1.  FOUT: the inverse
// transform stops after writing its second-to-last stage to LDS; the caller
// does the last stage.
template <int N, bool INIT, bool EXIT, bool DET, class C_, bool FIN = false, bool FOUT = false>
__device__ __forceinline__ void col_body(const ColArgs& a, long long w, float2* s_buf, const float2* s_tw,
                                         float* s_red, int c, int t, int x, bool valid) {
  using P = typename C_::PC;
  constexpr int T = C_::TW;
  using L = ColLayout<N, C_::PADSH_COL, C_::SW>;
  constexpr int NREG = regs_needed<N, T, P>();
  using M0 = IOMap<N, T, P, false>;
  using M1 = IOMap<N, T, P, true>;
  // a column inside one wavefront needs wave barriers only; a column over
  // several wavefronts (T = 128) synchronises the block
  using CS = std::conditional_t<(T <= 64), SyncWave, SyncBlock>;
  const int nx = a.nx;
  const long long n = (long long)N * nx;
  const long long colb = (long long)(valid ? x : x - c) * N;  // column-major offset
  float2 v[NREG];
  if constexpr (INIT) {
    float dx = 0.f, dy = 0.f, fz = 0.f;
    const bool pp = a.ppar != nullptr;
    if (pp) {
      const float4 q4 = a.ppar[w];
      dx = q4.x; dy = q4.y; fz = q4.z;
    }
    const float kx = (float)fidx(valid ? x : 0, nx) * a.kscx;
    static_for<0, M1::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M1::FULL || b < M1::NBF) {
        static_for<0, M1::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          const int e = b + q * M1::NBF;
          float2 val = s_buf[lds_slot<L, M1::NBF, q>(c, b, L::idx(c, b))];
          if (pp && valid) {
            const float ky = (float)fidx(e, N) * a.kscy;
            const float ksqr = kx * kx + ky * ky;
            float chi = -6.28318531f * (dx * kx + dy * ky + fz * ksqr);
            if (a.cppl_t) chi -= a.cppl_t[colb + e];
            float sn, cs;
            sincosf(chi, &sn, &cs);
            val = cmul(val, make_float2(cs, sn));
          }
          v[i * M1::R0 + q] = val;
        });
      }
    });
  } else {
    if constexpr (FIN) {
      LineFFT<N, T, -1, P, false, L, NREG, C_::TWR, -1, 1>::read_input(v, s_buf, c, t);
      CS{}();
      LineFFT<N, T, -1, P, false, L, NREG, C_::TWR, -1, 1>::run(v, s_buf, c, s_tw, t, CS{});
    } else {
      static_for<0, M0::NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const int b = t + i * T;
        if (M0::FULL || b < M0::NBF) {
          static_for<0, M0::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            v[i * M0::R0 + q] = s_buf[lds_slot<L, M0::NBF, q>(c, b, L::idx(c, b))];
          });
        }
      });
      CS{}();
      LineFFT<N, T, -1, P, false, L, NREG, C_::TWR>::run(v, s_buf, c, s_tw, t, CS{});
    }
    const float2* pcol = a.pro_t + colb;
    static_for<0, M1::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M1::FULL || b < M1::NBF) {
        static_for<0, M1::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          v[i * M1::R0 + q] = cmul(v[i * M1::R0 + q], pcol[b + q * M1::NBF]);
        });
      }
    });
  }
  // a.detect is uniform over the block (one wave per block row): in a grouped
  // launch, jobs that do not read this plane out skip the readout
  if (DET && a.detect) {
    float acc[kMaxDet];
#pragma unroll
    for (int d = 0; d < kMaxDet; ++d) acc[d] = 0.f;
    if (valid) {
      float pw[NREG];
      static_for<0, M1::NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        static_for<0, M1::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          const float2 z = v[i * M1::R0 + q];
          pw[i * M1::R0 + q] = z.x * z.x + z.y * z.y;
        });
      });
      // de-scan (DescanDifN, JMultiSlice.cpp:2184-2275): pixel i is weighted
      // with det(i - d)
      const long long colbs = (long long)wrap_sub(x, a.dsx, nx) * N;
      for (int d = 0; d < a.ndet && d < kMaxDet; ++d) {  // uniform loop over detectors
        const float* dw = a.detw_t + d * n + colbs;
        float wt[NREG];
        static_for<0, M1::NB>([&](auto I) {  // all weight loads of this detector first
          constexpr int i = decltype(I)::value;
          const int b = t + i * T;
          static_for<0, M1::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            wt[i * M1::R0 + q] = (M1::FULL || b < M1::NBF) ? dw[wrap_sub(b + q * M1::NBF, a.dsy, N)] : 0.f;
          });
        });
        float sacc = 0.f;
        static_for<0, M1::NB>([&](auto I) {
          constexpr int i = decltype(I)::value;
          const int b = t + i * T;
          if (M1::FULL || b < M1::NBF) {
            static_for<0, M1::R0>([&](auto Q) {
              constexpr int q = decltype(Q)::value;
              sacc += pw[i * M1::R0 + q] * wt[i * M1::R0 + q];
            });
          }
        });
#pragma unroll
        for (int dd = 0; dd < kMaxDet; ++dd)
          if (dd == d) acc[dd] += sacc;
      }
    }
    block_reduce_det(acc, a.ndet, s_red, a.partial + w * a.part_wstride + a.part_off + blockIdx.x,
                     a.part_dstride);
    if (valid && (a.dif || a.wff || a.dif_avg || a.wff_avg)) {  // ReadoutDifDet_d channels (JMultiSlice.cpp:4310-4356)
      const float wg = a.cwgt[w];
      const long long cb = w * a.chan_wstride + a.chan_off;
      static_for<0, M1::NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const int b = t + i * T;
        if (M1::FULL || b < M1::NBF) {
          static_for<0, M1::R0>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            const long long idx = cb + (long long)wrap_sub(b + q * M1::NBF, a.dsy, N) * nx + wrap_sub(x, a.dsx, nx);
            const float2 z = v[i * M1::R0 + q];
            const float pz = z.x * z.x + z.y * z.y;
            if (a.dif) a.dif[idx] += wg * pz;
            if (a.dif_avg) a.dif_avg[idx] += wg * pz;
            if (a.wff) { float2 o = a.wff[idx]; o.x += z.x * wg; o.y += z.y * wg; a.wff[idx] = o; }
            if (a.wff_avg) { float2 o = a.wff_avg[idx]; o.x += z.x * wg; o.y += z.y * wg; a.wff_avg[idx] = o; }
          });
        }
      });
    }
  }
  if constexpr (EXIT) {
    CS{}();
    static_for<0, M1::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M1::FULL || b < M1::NBF) {
        static_for<0, M1::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          s_buf[lds_slot<L, M1::NBF, q>(c, b, L::idx(c, b))] = v[i * M1::R0 + q];
        });
      }
    });
  } else if constexpr (FOUT) {
    CS{}();
    LineFFT<N, T, +1, P, true, L, NREG, C_::TWR, P::S - 2>::run(v, s_buf, c, s_tw, t, CS{});
  } else {
    CS{}();
    LineFFT<N, T, +1, P, true, L, NREG, C_::TWR>::run(v, s_buf, c, s_tw, t, CS{});
    CS{}();
    static_for<0, M0::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M0::FULL || b < M0::NBF) {
        static_for<0, M0::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          s_buf[lds_slot<L, M0::NBF, q>(c, b, L::idx(c, b))] = v[i * M0::R0 + q];
        });
      }
    });
  }
}

// ---------------------------------------------------------------------------
// column pass, wave-local form (N <= 768): the block stages a tile of CW
// adjacent columns through LDS with coalesced row-segment loads and stores;
// each column is transformed inside one wavefront (TW threads, TW divides 64),
// so the FFT exchanges need only wave-level barriers.  Propagator, detector
// weights and phase plate are read from column-major copies (pro_t, detw_t,
// cppl_t) so that the lanes of a column access them contiguously.
// grid: (ceil(nx/CW), nwave); block: TW*CW
// ---------------------------------------------------------------------------
template <int N, bool INIT, bool EXIT, bool DET, class C_ = Cfg<N>, bool GRP = false, bool PANEL = false>
__global__ __launch_bounds__(C_::TW * C_::CW) __attribute__((amdgpu_waves_per_eu(C_::WPE_COL))) void k_col_pass_w(
    ColArgs a0) {
  constexpr int T = C_::TW, C = C_::CW, BS = T * C;
  static_assert(64 % T == 0 || T % 64 == 0, "a column is part of one wavefront or whole wavefronts");
  static_assert(BS % 64 == 0, "block must be whole wavefronts");
  using L = ColLayout<N, C_::PADSH_COL, C_::SW>;
  using TL = ColTile<N, C_>;
  __shared__ float2 s_tw[N];
  __shared__ float2 s_buf[L::words(C)];
  __shared__ float s_red[(BS / 64) * kMaxDet];
  const int tid = threadIdx.x;
  const int c = tid / T;  // column within the tile
  const int t = tid % T;
  for (int e = tid; e < N; e += BS) s_tw[e] = a0.twg[e];
  // GRP: block row y is wave grp.w0 + y of the wave map; its job supplies the
  // buffers, and the plane table the propagator and the readout of plane
  // grp.step (each job reads out on its own planes)
  ColArgs a = a0;
  long long w = blockIdx.y;
  if constexpr (GRP) {
    const GrpWave gw = a0.grp.wav[a0.grp.w0 + blockIdx.y];
    const GrpJob& jb = a0.grp.job[gw.job];
    const GrpPlane* pl = a0.grp.pl + (long long)gw.job * a0.grp.pl_ld;
    w = gw.w;
    a.wave = (float2*)jb.wave;
    a.src = (const float2*)jb.src;
    a.src_stride = 0;
    a.ppar = (const float4*)jb.ppar;
    a.kscx = jb.kscx;
    a.kscy = jb.kscy;
    a.cppl_t = jb.cppl_t;
    a.ndet = jb.ndet;
    a.detw_t = jb.detw_t;
    a.partial = jb.partial;
    a.part_wstride = jb.part_wstride;
    const long long po = pl[a0.grp.step].part_off;
    a.detect = po >= 0 ? 1 : 0;
    a.part_off = po >= 0 ? po : 0;
    if constexpr (!INIT) a.pro_t = (const float2*)pl[a0.grp.step - 1].pro_t;
  }
  const int nx = a.nx;
  const int x0 = blockIdx.x * C;
  const int ncol = (nx - x0 < C) ? (nx - x0) : C;
  // PANEL: the waves are stored in panels of 16 columns (element (y, x) at
  // (x / 16) * 16 N + 16 y + x % 16), so the rows of a tile are contiguous
  static_assert(!PANEL || C <= kPanel, "a panel holds whole tiles");
  const int wrs = PANEL ? kPanel : nx;  // row stride of the wave
  float2* wv_base = a.wave + w * a.wave_stride + (PANEL ? panel_off<N>(x0) : x0);
  // Fused tile I/O: when a column's first-stage butterflies are exactly the
  // T threads per column, the staging sweep already leaves in thread tid the
  // inputs of butterfly tid / C of column tid % C (rows b + k*N/R0), so the
  // first forward stage runs on the staged registers and writes straight into
  // the exchange layout, and the last inverse stage (the same radix, reversed
  // plan) produces the rows that the store sweep writes.  This removes the
  // LDS round trips of staging in and out.
  using P = typename C_::PC;
  constexpr int NREG = regs_needed<N, T, P>();
  constexpr bool FIO = C_::FIO && P::S >= 2 && N / radix_at<P, false>(0) == T && TL::KST == NREG;
  constexpr bool FIN = FIO && !INIT, FOUT = FIO && !EXIT;
  const int cs = tid % C, bs = tid / C;  // sweep mapping: column, first-stage butterfly
  {
    float2 st[TL::KST];
    TL::load(st, INIT ? (a.src + w * a.src_stride + x0) : wv_base, INIT ? nx : wrs, ncol, tid);
    if constexpr (FIN) {
      // twiddles from the global table: s_tw is complete only after the barrier
      LineFFT<N, T, -1, P, false, L, NREG, C_::TWR, 0>::run(st, s_buf, cs, a.twg, bs, SyncWave{});
    } else {
      TL::to_lds(st, s_buf, tid);
    }
  }
  __syncthreads();
  col_body<N, INIT, EXIT, DET, C_, FIN, FOUT>(a, w, s_buf, s_tw, s_red, c, t, x0 + c, c < ncol);
  __syncthreads();
  {
    float2 st[TL::KST];
    if constexpr (FOUT) {
      LineFFT<N, T, +1, P, true, L, NREG, C_::TWR, -1, P::S - 1>::read_input(st, s_buf, cs, bs);
      LineFFT<N, T, +1, P, true, L, NREG, C_::TWR, -1, P::S - 1>::run(st, s_buf, cs, s_tw, bs, SyncWave{});
    } else {
      TL::from_lds(st, s_buf, tid);
    }
    TL::store(st, wv_base, wrs, ncol, tid);
  }
}

// ---------------------------------------------------------------------------
// column pass, register-resident form for long columns, N = 64*M with M = 16
// or 32 (N = 1024, 2048).  A block of 64*CT threads owns CT = 16 adjacent
// columns (every row segment it reads or writes is 128 bytes, one block per
// CU) or CT = 8 (64-byte segments; blocks of 512 threads, two per CU, so
// that one block's transforms overlap the other's memory phase, in an
// XCD-aware order that runs both halves of a 128-byte line in one L2); the
// tile (256 KiB at N = 2048, CT = 16: more than the LDS holds) stays in
// registers, M elements per thread, and the column transform is split as
// N = 64 * M (four-step):
//   A layout (global loads and stores): thread tid holds column tid % CT,
//     rows n1 + 64*n2 (n1 = tid / CT, n2 = 0..M-1); an M-point DFT over n2 in
//     registers, then the twiddle W_N^(n1*k2);
//   exchange through LDS in rounds of 16 values of k2 (8 KiB per column);
//   B layout: k2 = 16r + kk in round r, kk = wavefront wv (CT = 16) or
//     2 wv + lane / 32 (CT = 8); the lanes of a quad (j = lane % 4) hold
//     column (lane / 4) % CT, slot m <-> n1 = j + 4m; a 16-point DFT over m
//     in registers, the twiddle W_64^(j*c), and a 4-point DFT across the
//     quad by two DPP butterflies.  Slot c of lane j = j0 + 2*j1 then holds
//     frequency ky = k2 + M*(c + 16*(j1 + 2*j0)).
// The phase plate (INIT), propagator, readouts and the exit store act in the
// B layout on row-major tables (a wavefront touches four 128-byte or eight
// 64-byte row segments per access); the inverse transform runs the steps
// backwards.  GRP: block row by is wave w0 + by of a grouped launch's wave
// map, whose job and plane tables supply the buffers and tables.
// grid: (ceil(nx/CT), nwave); block: 64*CT
// ---------------------------------------------------------------------------
// the value of x, hidden from the optimiser: table reads indexed with it are
// repeated where they are needed instead of being kept live in registers
// from an earlier identical read
__device__ __forceinline__ int opaque(int x) {
  __asm__ volatile("" : "+v"(x));
  return x;
}

// base[idx] addressed by a 32-bit byte offset (idx * sizeof(T) < 2^32), so
// that a wavefront-uniform base stays in scalar registers
template <class T>
__device__ __forceinline__ T* at32(T* base, int idx) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (unsigned)(idx * (int)sizeof(T)));
}
template <class T>
__device__ __forceinline__ const T* at32(const T* base, int idx) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (unsigned)(idx * (int)sizeof(T)));
}

// v[k] *= W_N^(DIR * n1 * k), k = 1..M-1, N = 64*M, from two exact table
// reads per factor: W^(n1*(8a + b)) = W^(8a*n1) * W^(b*n1) (n1*k < N); tw
// holds the first half of the table, W^(e + N/2) = -W^e
template <int M, int DIR>
__device__ __forceinline__ void twiddle_a(float2 (&v)[M], const float2* tw, int n1) {
  constexpr int H = 32 * M;
  static_assert(7 * 63 < H, "W^(b*n1) in the half table");
  float2 wb[8];
  static_for<1, 8>([&](auto B) {
    constexpr int b = decltype(B)::value;
    wb[b] = tw[b * n1];
  });
  static_for<0, M / 8>([&](auto A) {
    constexpr int a8 = decltype(A)::value;
    const int e = 8 * a8 * n1;
    const bool hi = e >= H;
    const float2 th = tw[hi ? e - H : e];
    const float2 wa = hi ? make_float2(-th.x, -th.y) : th;
    static_for<0, 8>([&](auto B) {
      constexpr int b = decltype(B)::value;
      constexpr int k = 8 * a8 + b;
      if constexpr (k > 0) {
        const float2 wk = (a8 == 0) ? wb[b] : ((b == 0) ? wa : cmul(wa, wb[b]));
        v[k] = (DIR < 0) ? cmul(v[k], wk) : cmulconj(v[k], wk);
      }
    });
  });
}

template <int N, bool INIT, bool EXIT, bool DET, class C_ = Cfg<N>, bool PANEL = false, bool GRP = false>
__global__ __launch_bounds__(64 * C_::XCT) __attribute__((amdgpu_waves_per_eu(4))) void k_col_pass_x(ColArgs a0) {
  constexpr int M = N / 64;
  static_assert(M == 16 || M == 32, "long-column kernel: N = 1024 or 2048");
  constexpr int R = M / 16;   // exchange rounds
  constexpr int CT = C_::XCT;  // columns per block
  static_assert(CT == 16 || CT == 8, "long-column kernel: 8 or 16 columns per block");
  constexpr int NT = 64 * CT;  // threads
  __shared__ float2 s_x[16 * 64 * CT];  // one round: [k2 % 16][n1][column]
  __shared__ float2 s_tw[N / 2];  // W_N^e, e < N/2
  // W_64^(j*c) at [c][j] (j = 0..3, c = 0..15): the four lanes of a quad
  // read four adjacent entries (in s_tw they would share one bank)
  __shared__ float2 s_w64[64];
  __shared__ float s_red[16 * kMaxDet];
  const int tid = threadIdx.x;
  for (int e = tid; e < N / 2; e += NT) s_tw[e] = a0.twg[e];
  if (tid < 64) s_w64[tid] = a0.twg[((tid & 3) * (tid >> 2) * M) % N];
  int bx, by;
  if constexpr (C_::XG_X > 0) block_xy_groups<C_::XG_X, C_::XROT_X>(bx, by);
  else block_xy<C_::XCD_COL>(bx, by);
  // GRP: block row by is wave grp.w0 + by of the wave map; its job supplies
  // the buffers, and the plane table the propagator and the readout of plane
  // grp.step (each job reads out on its own planes)
  ColArgs a = a0;
  long long w = by;
  if constexpr (GRP) {
    const GrpWave gw = a0.grp.wav[a0.grp.w0 + by];
    const GrpJob& jb = a0.grp.job[gw.job];
    const GrpPlane* pl = a0.grp.pl + (long long)gw.job * a0.grp.pl_ld;
    w = gw.w;
    a.wave = (float2*)jb.wave;
    a.src = (const float2*)jb.src;
    a.src_stride = 0;
    a.ppar = (const float4*)jb.ppar;
    a.kscx = jb.kscx;
    a.kscy = jb.kscy;
    a.cppl = jb.cppl;
    a.ndet = jb.ndet;
    a.detw = jb.detw;
    a.partial = jb.partial;
    a.part_wstride = jb.part_wstride;
    const long long po = pl[a0.grp.step].part_off;
    a.detect = po >= 0 ? 1 : 0;
    a.part_off = po >= 0 ? po : 0;
    if constexpr (!INIT) a.pro = (const float2*)pl[a0.grp.step - 1].pro;
  }
  const int nx = a.nx;
  const int x0 = bx * CT;
  const int ncol = (nx - x0 < CT) ? (nx - x0) : CT;
  // PANEL: the tile lies in one panel (element (y, x) at p N + 16 y + x - p,
  // p = x0 rounded down to a multiple of 16)
  static_assert(!PANEL || kPanel % CT == 0, "a tile lies in one panel");
  const int wrs = PANEL ? kPanel : nx;  // row stride of the wave
  float2* wbase =
      a.wave + w * a.wave_stride + (PANEL ? (long long)(x0 & ~(kPanel - 1)) * N + (x0 & (kPanel - 1)) : x0);
  // A layout (columns past the grid edge work on column 0 of the tile and
  // are not stored)
  const int ca = tid % CT, n1 = tid / CT;
  const int xa = (ca < ncol) ? ca : 0;
  // B layout
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int kk = CT == 16 ? wv : 2 * wv + (lane >> 5);  // k2 % 16
  const int j = lane & 3, cb = (lane >> 2) & (CT - 1);
  const int j0 = j & 1, j1 = j >> 1;
  const int dq = j1 + 2 * j0;
  const float sg1 = j0 ? -1.f : 1.f, sg2 = j1 ? -1.f : 1.f;
  const bool validb = cb < ncol;
  const int xb = x0 + (validb ? cb : 0);
  // exchange buffer [k2 % 16][n1][column].  CT = 16: columns XOR-swizzled,
  // rows with bit 1 set hold column c at c ^ 8, so that the B-layout
  // accesses (rows j + 4m of a quad's four lanes, 8 columns per
  // half-wavefront) fall on 64 distinct banks instead of two lanes per bank;
  // the A-layout accesses stay a permutation inside one row.  CT = 8: the
  // four rows of a half-wavefront's quads are 64 bytes apart, distinct banks
  // as they lie
  const int swa = CT == 16 ? ca ^ (((n1 >> 1) & 1) << 3) : ca;
  const int swb = CT == 16 ? cb ^ (j1 << 3) : cb;
  // slot (r, c) of the B layout <-> row ky = 16r + M*c (uniform part) + kyb
  const int kyb = kk + 16 * M * dq;
  auto kyof = [&](int r, int c) { return 16 * r + M * c + kyb; };
  // rows of the B layout are addressed as (row 16r + M*c of a table) +
  // (per-lane offset kyb*nx + x): the first part is wavefront-uniform
  const int ob = kyb * nx + xb;
  const long long rstep = (long long)M * nx;  // elements from slot c to c + 1
  // A layout: row n1 + 64*n2 = (uniform row 64*n2) + per-thread offset
  const long long astep = 64LL * wrs;
  float2 v[M];  // A: v[n2] (rows n1 + 64*n2); B: v[16r + c]
  if constexpr (!INIT) {
    static_for<0, M>([&](auto I) {
      constexpr int n2 = decltype(I)::value;
      v[n2] = ldg<C_::NT_COL>(at32(wbase + n2 * astep, n1 * wrs + xa));
    });
    DFT<M, -1>::run(v);
  }
  __builtin_amdgcn_sched_barrier(0);
  __syncthreads();  // twiddle table complete
  if constexpr (!INIT) twiddle_a<M, -1>(v, s_tw, opaque(n1));
  float acc[kMaxDet];
#pragma unroll
  for (int d = 0; d < kMaxDet; ++d) acc[d] = 0.f;
  float dx = 0.f, dy = 0.f, fz = 0.f;
  if constexpr (INIT) {
    if (a.ppar) {
      const float4 q4 = a.ppar[w];
      dx = q4.x; dy = q4.y; fz = q4.z;
    }
  }
  // B phase, one round (16 values of k2) at a time
  static_for<0, R>([&](auto RR) {
    constexpr int r = decltype(RR)::value;
    float2* u = &v[16 * r];
    if constexpr (INIT) {
      const float2* s = a.src + w * a.src_stride + 16 * r * nx;
      const float kx = (float)fidx(xb, nx) * a.kscx;
      static_for<0, 16>([&](auto K) {
        constexpr int c = decltype(K)::value;
        const int ky = kyof(r, c);
        float2 val = *at32(s + c * rstep, ob);
        if (a.ppar) {
          const float kyf = (float)fidx(ky, N) * a.kscy;
          const float ksqr = kx * kx + kyf * kyf;
          float chi = -6.28318531f * (dx * kx + dy * kyf + fz * ksqr);
          if (a.cppl) chi -= *at32(a.cppl + 16 * r * nx + c * rstep, ob);
          float sn, cs;
          sincosf(chi, &sn, &cs);
          val = cmul(val, make_float2(cs, sn));
        }
        u[c] = val;
      });
    } else {
      if constexpr (r > 0) __syncthreads();  // the previous round's reads are done
      static_for<0, 16>([&](auto K) {
        constexpr int k = decltype(K)::value;
        s_x[(k * 64 + n1) * CT + swa] = u[k];
      });
      __syncthreads();
      static_for<0, 16>([&](auto K) {
        constexpr int m = decltype(K)::value;
        u[m] = s_x[(kk * 64 + j + 4 * m) * CT + swb];
      });
      DFT<16, -1>::run(u);
      __builtin_amdgcn_sched_barrier(0);
      const float2* w64 = s_w64 + opaque(j);
      static_for<1, 16>([&](auto K) {
        constexpr int c = decltype(K)::value;
        u[c] = cmul(u[c], w64[4 * c]);  // W_64^(j*c)
      });
      static_for<0, 16>([&](auto K) {
        constexpr int c = decltype(K)::value;
        float2 z = quad_bfly<2>(u[c], sg2);
        if (j0 & j1) z = mul_si<-1>(z);  // W_4^(j0*e0)
        u[c] = cmul(quad_bfly<1>(z, sg1), *at32(a.pro + 16 * r * nx + c * rstep, ob));
      });
    }
    if constexpr (DET) {
      if (validb && a.detect) {
        // de-scan (DescanDifN, JMultiSlice.cpp:2184-2275): pixel i is
        // weighted with det(i - d)
        const int xs = wrap_sub(xb, a.dsx, nx);
        const int qs = (kyb - a.dsy) & (N - 1);  // row of slot (r, c): (16r + M*c + qs) mod N
        for (int d = 0; d < a.ndet && d < kMaxDet; ++d) {  // uniform loop over detectors
          const float* dw = a.detw + d * (long long)N * nx;
          const int q = opaque(qs);  // row offsets recomputed per detector
          float wt[16];
          static_for<0, 16>([&](auto K) {
            constexpr int c = decltype(K)::value;
            wt[c] = *at32(dw, ((16 * r + M * c + q) & (N - 1)) * nx + xs);
          });
          float sacc = 0.f;
          static_for<0, 16>([&](auto K) {
            constexpr int c = decltype(K)::value;
            sacc += (u[c].x * u[c].x + u[c].y * u[c].y) * wt[c];
          });
#pragma unroll
          for (int dd = 0; dd < kMaxDet; ++dd)
            if (dd == d) acc[dd] += sacc;
        }
        if (a.dif || a.wff || a.dif_avg || a.wff_avg) {  // ReadoutDifDet_d channels (JMultiSlice.cpp:4310-4356)
          const float wg = a.cwgt[w];
          const long long cbase = w * a.chan_wstride + a.chan_off;
          const int q = opaque(qs);
          static_for<0, 16>([&](auto K) {
            constexpr int c = decltype(K)::value;
            const long long idx = cbase + ((16 * r + M * c + q) & (N - 1)) * nx + xs;
            const float2 z = u[c];
            const float pz = z.x * z.x + z.y * z.y;
            if (a.dif) a.dif[idx] += wg * pz;
            if (a.dif_avg) a.dif_avg[idx] += wg * pz;
            if (a.wff) { float2 o = a.wff[idx]; o.x += z.x * wg; o.y += z.y * wg; a.wff[idx] = o; }
            if (a.wff_avg) { float2 o = a.wff_avg[idx]; o.x += z.x * wg; o.y += z.y * wg; a.wff_avg[idx] = o; }
          });
        }
      }
    }
    if constexpr (EXIT) {
      if (validb) {
        float2* wv0 = PANEL ? wbase + 16 * r * kPanel : a.wave + w * a.wave_stride + 16 * r * nx;
        const long long rs_w = PANEL ? (long long)M * kPanel : rstep;
        const int ob_w = PANEL ? kyb * kPanel + cb : ob;
        static_for<0, 16>([&](auto K) {
          constexpr int c = decltype(K)::value;
          stg<C_::NT_COL>(at32(wv0 + c * rs_w, ob_w), u[c]);
        });
      }
    }
  });
  if constexpr (DET) {
    if (a.detect)  // uniform over the block (one wave)
      block_reduce_det(acc, a.ndet, s_red, a.partial + w * a.part_wstride + a.part_off + bx, a.part_dstride);
  }
  if constexpr (!EXIT) {
    static_for<0, R>([&](auto RR) {
      constexpr int r = decltype(RR)::value;
      float2* u = &v[16 * r];
      static_for<0, 16>([&](auto K) {
        constexpr int c = decltype(K)::value;
        float2 z = quad_bfly<1>(u[c], sg1);
        if (j0 & j1) z = mul_si<+1>(z);
        u[c] = quad_bfly<2>(z, sg2);
      });
      const float2* w64 = s_w64 + opaque(j);
      static_for<1, 16>([&](auto K) {
        constexpr int c = decltype(K)::value;
        u[c] = cmulconj(u[c], w64[4 * c]);
      });
      DFT<16, +1>::run(u);
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();  // earlier users of the exchange buffer are done
      static_for<0, 16>([&](auto K) {
        constexpr int m = decltype(K)::value;
        s_x[(kk * 64 + j + 4 * m) * CT + swb] = u[m];
      });
      __syncthreads();
      static_for<0, 16>([&](auto K) {
        constexpr int k = decltype(K)::value;
        u[k] = s_x[(k * 64 + n1) * CT + swa];
      });
    });
    __builtin_amdgcn_sched_barrier(0);
    twiddle_a<M, +1>(v, s_tw, opaque(n1));
    DFT<M, +1>::run(v);
    if (ca < ncol) {
      static_for<0, M>([&](auto I) {
        constexpr int n2 = decltype(I)::value;
        stg<C_::NT_COL>(at32(wbase + n2 * astep, n1 * wrs + ca), v[n2]);
      });
    }
  }
}

// ---------------------------------------------------------------------------
// standalone line transforms (2-D FFT = row transform + column transform)
// ---------------------------------------------------------------------------
template <int N, int DIR, class C_ = Cfg<N>>
__global__ __launch_bounds__(C_::T * C_::LPB) void k_row_fft(float2* __restrict__ data,
                                                                      int ny, long long stride,
                                                                      const float2* __restrict__ twg) {
  using P = typename C_::P;
  constexpr int T = C_::T, LPB = C_::LPB;
  using L = LineLayout<N, C_::PADSH, C_::SW>;
  constexpr int NREG = regs_needed<N, T, P>();
  __shared__ float2 s_tw[N];
  __shared__ float2 s_buf[L::words(LPB)];
  const int tid = threadIdx.x, t = tid % T, line = tid / T;
  for (int e = tid; e < N; e += T * LPB) s_tw[e] = twg[e];
  const int row = blockIdx.x * LPB + line;
  const bool valid = row < ny;
  float2* d = data + blockIdx.y * stride + (long long)row * N;
  using M0 = IOMap<N, T, P, false>;
  using M1 = IOMap<N, T, P, true>;
  float2 v[NREG];
  if (valid) {
    static_for<0, M0::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M0::FULL || b < M0::NBF)
        static_for<0, M0::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          v[i * M0::R0 + q] = d[b + q * M0::NBF];
        });
    });
  }
  __syncthreads();
  LineFFT<N, T, DIR, P, false, L, NREG, C_::TWR>::run(v, s_buf, line, s_tw, t, SyncBlock{});
  if (valid) {
    static_for<0, M1::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M1::FULL || b < M1::NBF)
        static_for<0, M1::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          d[b + q * M1::NBF] = v[i * M1::R0 + q];
        });
    });
  }
}

template <int N, int DIR, class C_ = Cfg<N>>
__global__ __launch_bounds__(C_::TC * C_::C) void k_col_fft(float2* __restrict__ data,
                                                                     int nx, long long stride,
                                                                     const float2* __restrict__ twg) {
  using P = typename C_::PC;
  constexpr int T = C_::TC, C = C_::C;
  using L = TileLayout<N, C, C_::PADSH>;
  constexpr int NREG = regs_needed<N, T, P>();
  __shared__ float2 s_tw[N];
  __shared__ float2 s_buf[L::words(1)];
  const int tid = threadIdx.x, c = tid % C, t = tid / C;
  for (int e = tid; e < N; e += T * C) s_tw[e] = twg[e];
  const int x = blockIdx.x * C + c;
  const bool valid = x < nx;
  float2* d = data + blockIdx.y * stride + x;
  using M0 = IOMap<N, T, P, false>;
  using M1 = IOMap<N, T, P, true>;
  float2 v[NREG];
  if (valid) {
    static_for<0, M0::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M0::FULL || b < M0::NBF)
        static_for<0, M0::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          v[i * M0::R0 + q] = d[(long long)(b + q * M0::NBF) * nx];
        });
    });
  }
  __syncthreads();
  LineFFT<N, T, DIR, P, false, L, NREG, C_::TWR>::run(v, s_buf, c, s_tw, t, SyncBlock{});
  if (valid) {
    static_for<0, M1::NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if (M1::FULL || b < M1::NBF)
        static_for<0, M1::R0>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          d[(long long)(b + q * M1::NBF) * nx] = v[i * M1::R0 + q];
        });
    });
  }
}


// ---------------------------------------------------------------------------
// per-size launchers (instantiated in jms_kernels_size.hip)
// ---------------------------------------------------------------------------
template <int N>
hipError_t launch_row_pass_n(int ny, int nwave, float2* wave, long long wave_stride,
                             const float2* pgr, long long slc_off, long long npg,
                             const int* var_tab, int var_ld, int jslc, const float2* twx,
                             const RowChanHost& h, int mode, bool panel, hipStream_t st) {
  dim3 g((ny + Cfg<N>::LPB - 1) / Cfg<N>::LPB, nwave), b(Cfg<N>::T * Cfg<N>::LPB);
  RowChan rc;
  rc.img = h.img; rc.img_avg = h.img_avg;
  rc.wfr = (float2*)h.wfr; rc.wfr_avg = (float2*)h.wfr_avg;
  rc.cwgt = h.cwgt; rc.chan_wstride = h.chan_wstride; rc.chan_off = h.chan_off;
  if (panel && (mode != 0 || !Cfg<N>::PANEL)) return hipErrorInvalidValue;
  if (mode == 1)
    hipLaunchKernelGGL((k_row_pass<N, 1>), g, b, 0, st, wave, ny, wave_stride, pgr, slc_off, npg,
                       var_tab, var_ld, jslc, twx, rc, GrpArgs{});
  else if (mode == 2)
    hipLaunchKernelGGL((k_row_pass<N, 2>), g, b, 0, st, wave, ny, wave_stride, pgr, slc_off, npg,
                       var_tab, var_ld, jslc, twx, rc, GrpArgs{});
  else if (!panel)
    hipLaunchKernelGGL((k_row_pass<N, 0>), g, b, 0, st, wave, ny, wave_stride, pgr, slc_off, npg,
                       var_tab, var_ld, jslc, twx, rc, GrpArgs{});
  else if constexpr (Cfg<N>::PANEL)
    hipLaunchKernelGGL((k_row_pass<N, 0, Cfg<N>, false, true>), g, b, 0, st, wave, ny, wave_stride, pgr, slc_off,
                       npg, var_tab, var_ld, jslc, twx, rc, GrpArgs{});
  return hipGetLastError();
}

// grouped row pass (waves of several jobs, one slice step)
template <int N>
hipError_t launch_row_pass_grp_n(int ny, int nwave, long long wave_stride, long long npg, const float2* twx,
                                 const GrpArgs& ga, bool panel, hipStream_t st) {
  dim3 g((ny + Cfg<N>::LPB - 1) / Cfg<N>::LPB, nwave), b(Cfg<N>::T * Cfg<N>::LPB);
  if (panel && !Cfg<N>::PANEL) return hipErrorInvalidValue;
  if (!panel)
    hipLaunchKernelGGL((k_row_pass<N, 0, Cfg<N>, true>), g, b, 0, st, nullptr, ny, wave_stride, nullptr, 0LL, npg,
                       nullptr, 0, ga.step, twx, RowChan{}, ga);
  else if constexpr (Cfg<N>::PANEL)
    hipLaunchKernelGGL((k_row_pass<N, 0, Cfg<N>, true, true>), g, b, 0, st, nullptr, ny, wave_stride, nullptr, 0LL,
                       npg, nullptr, 0, ga.step, twx, RowChan{}, ga);
  return hipGetLastError();
}

// the column kernels' instantiations by pass kind (INIT, EXIT, readout)
template <int N, class CF, bool GRP, bool PANEL>
void launch_col_w(dim3 g, dim3 b, bool init, bool exit_, bool detect, const ColArgs& a, hipStream_t st) {
  // detection planes get their own instantiation: the readout's registers
  // would otherwise halve the occupancy of every other slice
  if (detect) {
    if (init && exit_) hipLaunchKernelGGL((k_col_pass_w<N, true, true, true, CF, GRP, PANEL>), g, b, 0, st, a);
    else if (init) hipLaunchKernelGGL((k_col_pass_w<N, true, false, true, CF, GRP, PANEL>), g, b, 0, st, a);
    else if (exit_) hipLaunchKernelGGL((k_col_pass_w<N, false, true, true, CF, GRP, PANEL>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_col_pass_w<N, false, false, true, CF, GRP, PANEL>), g, b, 0, st, a);
  } else {
    if (init && exit_) hipLaunchKernelGGL((k_col_pass_w<N, true, true, false, CF, GRP, PANEL>), g, b, 0, st, a);
    else if (init) hipLaunchKernelGGL((k_col_pass_w<N, true, false, false, CF, GRP, PANEL>), g, b, 0, st, a);
    else if (exit_) hipLaunchKernelGGL((k_col_pass_w<N, false, true, false, CF, GRP, PANEL>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_col_pass_w<N, false, false, false, CF, GRP, PANEL>), g, b, 0, st, a);
  }
}
template <int N, class CF, bool PANEL, bool GRP = false>
void launch_col_x(dim3 g, dim3 b, bool init, bool exit_, bool detect, const ColArgs& a, hipStream_t st) {
  if (detect) {
    if (init && exit_) hipLaunchKernelGGL((k_col_pass_x<N, true, true, true, CF, PANEL, GRP>), g, b, 0, st, a);
    else if (init) hipLaunchKernelGGL((k_col_pass_x<N, true, false, true, CF, PANEL, GRP>), g, b, 0, st, a);
    else if (exit_) hipLaunchKernelGGL((k_col_pass_x<N, false, true, true, CF, PANEL, GRP>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_col_pass_x<N, false, false, true, CF, PANEL, GRP>), g, b, 0, st, a);
  } else {
    if (init && exit_) hipLaunchKernelGGL((k_col_pass_x<N, true, true, false, CF, PANEL, GRP>), g, b, 0, st, a);
    else if (init) hipLaunchKernelGGL((k_col_pass_x<N, true, false, false, CF, PANEL, GRP>), g, b, 0, st, a);
    else if (exit_) hipLaunchKernelGGL((k_col_pass_x<N, false, true, false, CF, PANEL, GRP>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_col_pass_x<N, false, false, false, CF, PANEL, GRP>), g, b, 0, st, a);
  }
}

template <int N>
hipError_t launch_col_pass_n(int nwave, bool init, bool exit_, const ColArgsHost& h,
                             hipStream_t st) {
  ColArgs a;
  a.wave = (float2*)h.wave; a.nx = h.nx; a.wave_stride = h.wave_stride;
  a.pro = (const float2*)h.pro; a.src = (const float2*)h.src; a.src_stride = h.src_stride;
  a.ppar = (const float4*)h.ppar; a.kscx = h.kscx; a.kscy = h.kscy; a.cppl = h.cppl;
  a.detect = h.detect; a.ndet = h.ndet; a.detw = h.detw; a.partial = h.partial;
  a.part_wstride = h.part_wstride; a.part_dstride = h.part_dstride; a.part_off = h.part_off;
  a.dif = h.dif; a.wff = (float2*)h.wff; a.cwgt = h.cwgt; a.chan_wstride = h.chan_wstride;
  a.dif_avg = h.dif_avg; a.wff_avg = (float2*)h.wff_avg;
  a.dsx = h.dsx; a.dsy = h.dsy;
  a.chan_off = h.chan_off;
  a.twg = (const float2*)h.twy;
  a.pro_t = (const float2*)h.pro_t;
  a.detw_t = h.detw_t;
  a.cppl_t = h.cppl_t;
  using CF = Cfg<N>;
  if (h.panel && !CF::PANEL) return hipErrorInvalidValue;
  if constexpr (CF::XL) {
    dim3 g((h.nx + CF::XCT - 1) / CF::XCT, nwave), b(64 * CF::XCT);
    if (!h.panel) launch_col_x<N, CF, false>(g, b, init, exit_, h.detect, a, st);
    else if constexpr (CF::PANEL) launch_col_x<N, CF, true>(g, b, init, exit_, h.detect, a, st);
  } else if constexpr (CF::WL) {
    dim3 g((h.nx + CF::CW - 1) / CF::CW, nwave), b(CF::TW * CF::CW);
    if (!h.panel) launch_col_w<N, CF, false, false>(g, b, init, exit_, h.detect, a, st);
    else if constexpr (CF::PANEL) launch_col_w<N, CF, false, true>(g, b, init, exit_, h.detect, a, st);
  } else {
    if (h.panel) return hipErrorInvalidValue;
    dim3 g((h.nx + CF::C - 1) / CF::C, nwave), b(CF::TC * CF::C);
    if (init && exit_) hipLaunchKernelGGL((k_col_pass<N, true, true>), g, b, 0, st, a);
    else if (init) hipLaunchKernelGGL((k_col_pass<N, true, false>), g, b, 0, st, a);
    else if (exit_) hipLaunchKernelGGL((k_col_pass<N, false, true>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_col_pass<N, false, false>), g, b, 0, st, a);
  }
  return hipGetLastError();
}

// grouped column pass (waves of several jobs; the wave-local and the
// long-column kernel)
template <int N>
hipError_t launch_col_pass_grp_n(int nx, int nwave, bool init, bool exit_, bool detect, long long part_dstride,
                                 const float2* twy, const GrpArgs& ga, bool panel, hipStream_t st) {
  using CF = Cfg<N>;
  if constexpr (CF::XL) {
    if (panel) return hipErrorInvalidValue;
    ColArgs a{};
    a.nx = nx;
    a.wave_stride = (long long)N * nx;
    a.part_dstride = part_dstride;
    a.twg = twy;
    a.grp = ga;
    dim3 g((nx + CF::XCT - 1) / CF::XCT, nwave), b(64 * CF::XCT);
    launch_col_x<N, CF, false, true>(g, b, init, exit_, detect, a, st);
    return hipGetLastError();
  } else if constexpr (CF::WL) {
    if (panel && !CF::PANEL) return hipErrorInvalidValue;
    ColArgs a{};
    a.nx = nx;
    a.wave_stride = (long long)N * nx;
    a.part_dstride = part_dstride;
    a.twg = twy;
    a.grp = ga;
    dim3 g((nx + CF::CW - 1) / CF::CW, nwave), b(CF::TW * CF::CW);
    if (!panel) launch_col_w<N, CF, true, false>(g, b, init, exit_, detect, a, st);
    else if constexpr (CF::PANEL) launch_col_w<N, CF, true, true>(g, b, init, exit_, detect, a, st);
    return hipGetLastError();
  } else {
    return hipErrorInvalidValue;
  }
}

template <int N>
hipError_t launch_row_fft_n(int ny, int nbatch, float2* data, long long stride, int dir,
                            const float2* twx, hipStream_t st) {
  dim3 g((ny + Cfg<N>::LPB - 1) / Cfg<N>::LPB, nbatch), b(Cfg<N>::T * Cfg<N>::LPB);
  if (dir < 0) hipLaunchKernelGGL((k_row_fft<N, -1>), g, b, 0, st, data, ny, stride, twx);
  else hipLaunchKernelGGL((k_row_fft<N, +1>), g, b, 0, st, data, ny, stride, twx);
  return hipGetLastError();
}

template <int N>
hipError_t launch_col_fft_n(int nx, int nbatch, float2* data, long long stride, int dir,
                            const float2* twy, hipStream_t st) {
  dim3 g((nx + Cfg<N>::C - 1) / Cfg<N>::C, nbatch), b(Cfg<N>::TC * Cfg<N>::C);
  if (dir < 0) hipLaunchKernelGGL((k_col_fft<N, -1>), g, b, 0, st, data, nx, stride, twy);
  else hipLaunchKernelGGL((k_col_fft<N, +1>), g, b, 0, st, data, nx, stride, twy);
  return hipGetLastError();
}

template <int N> constexpr int row_lines_per_block() { return Cfg<N>::LPB; }
template <int N> constexpr int col_lines_per_block() { return Cfg<N>::C; }

}  // namespace jms
