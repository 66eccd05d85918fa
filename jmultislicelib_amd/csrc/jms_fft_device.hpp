// jms_fft_device.hpp -- device-side building blocks of the line FFTs used by
// the multislice kernels (gfx950 / CDNA4, wave64).
//
// A "line" is one row or one column of the wave function grid.  A line of N
// complex points is transformed by T cooperating threads with a mixed-radix,
// self-sorting Stockham FFT (decimation in frequency; the method of Stockham
// as written in Van Loan, "Computational Frameworks for the FFT", 1992).  One
// stage of radix R over butterflies b in [0, N/R), with s = product of the
// radices already applied, computes
//     y[(b % s) + (b / s) * s * R + s * k] =
//         W_N^{(b / s) * s * k} * sum_q x[b + q * N/R] * w_R^{q k}
// so every stage reads its inputs at the same stride N/R.  Butterfly b is
// owned by thread b % T (butterflies t, t+T, t+2T, ...).  Stage inputs live in
// registers; between stages the outputs go through LDS once.  The last stage
// leaves X[b + k * N/R_last] in the registers of the thread that owns b, which
// is the layout the first stage of a plan with the radices in reverse order
// reads -- so an inverse transform followed by a pointwise multiply and a
// forward transform with the reversed plan needs no LDS round trip between the
// two transforms.
//
// Sign conventions follow the reference: forward = exp(-2 pi i nk/N),
// backward = exp(+2 pi i nk/N), no scaling (JFFTCUDAcore.cpp:299-336).
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

namespace jms {

// ---------------------------------------------------------------------------
// compile-time helpers
// ---------------------------------------------------------------------------
template <int I>
using ic = std::integral_constant<int, I>;

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(ic<B>{});
    static_for<B + 1, E>(static_cast<F&&>(f));
  }
}

// sin/cos of 2*pi*m/R evaluated at compile time (Taylor series after range
// reduction to [-pi/4, pi/4]); used for the codelet constants.
constexpr double kPiD = 3.14159265358979323846264338327950288;
constexpr double cx_sin_small(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 14; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos_small(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
// cos/sin of 2*pi*m/R, exact octant reduction on the integer m.
constexpr double cx_cos2pi(long m, long R) {
  m %= R;
  if (m < 0) m += R;
  // angle = 2 pi m / R; fold using integer arithmetic on 8*m vs R
  long m8 = 8 * m;  // angle in units of pi/4 is m8 / R
  if (m8 <= R) return cx_cos_small(2.0 * kPiD * (double)m / (double)R);
  if (m8 <= 3 * R) return -cx_sin_small(2.0 * kPiD * (double)m / (double)R - kPiD / 2);
  if (m8 <= 5 * R) return -cx_cos_small(2.0 * kPiD * (double)m / (double)R - kPiD);
  if (m8 <= 7 * R) return cx_sin_small(2.0 * kPiD * (double)m / (double)R - 3 * kPiD / 2);
  return cx_cos_small(2.0 * kPiD * (double)m / (double)R - 2 * kPiD);
}
constexpr double cx_sin2pi(long m, long R) {
  m %= R;
  if (m < 0) m += R;
  long m8 = 8 * m;
  if (m8 <= R) return cx_sin_small(2.0 * kPiD * (double)m / (double)R);
  if (m8 <= 3 * R) return cx_cos_small(2.0 * kPiD * (double)m / (double)R - kPiD / 2);
  if (m8 <= 5 * R) return -cx_sin_small(2.0 * kPiD * (double)m / (double)R - kPiD);
  if (m8 <= 7 * R) return -cx_cos_small(2.0 * kPiD * (double)m / (double)R - 3 * kPiD / 2);
  return cx_sin_small(2.0 * kPiD * (double)m / (double)R - 2 * kPiD);
}
template <int R, int M>
struct Root {  // exp(+2 pi i M / R) = (c, s)
  static constexpr float c = (float)cx_cos2pi(M, R);
  static constexpr float s = (float)cx_sin2pi(M, R);
};

// ---------------------------------------------------------------------------
// complex helpers (float2 = interleaved complex64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulconj(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// global loads/stores with the nontemporal hint: streamed data then does not
// displace reused arrays (phase gratings, propagators) from L2
typedef float f2v __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ float2 ldg(const float2* p) {
  if constexpr (NT) {
    const f2v r = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
    return make_float2(r.x, r.y);
  } else {
    return *p;
  }
}
template <bool NT>
__device__ __forceinline__ void stg(float2* p, float2 v) {
  if constexpr (NT) {
    f2v r;
    r.x = v.x;
    r.y = v.y;
    __builtin_nontemporal_store(r, reinterpret_cast<f2v*>(p));
  } else {
    *p = v;
  }
}
// multiply by sigma*i, sigma = +1 (DIR > 0) or -1 (DIR < 0)
template <int DIR>
__device__ __forceinline__ float2 mul_si(float2 a) {
  if constexpr (DIR > 0) return make_float2(-a.y, a.x);
  else return make_float2(a.y, -a.x);
}
// multiply by exp(DIR * 2 pi i M / R) with compile-time constants
template <int R, int M, int DIR>
__device__ __forceinline__ float2 mul_root(float2 a) {
  constexpr int m = ((M % R) + R) % R;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (4 * m == R) {
    return mul_si<DIR>(a);
  } else if constexpr (2 * m == R) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (4 * m == 3 * R) {
    return mul_si<-DIR>(a);
  } else {
    constexpr float c = Root<R, m>::c;
    constexpr float s = (DIR > 0) ? Root<R, m>::s : -Root<R, m>::s;
    return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// ---------------------------------------------------------------------------
// in-register R-point DFT codelets: x[k] <- sum_j x[j] exp(DIR 2 pi i jk/R)
// ---------------------------------------------------------------------------
template <int R, int DIR>
struct DFT;

template <int DIR>
struct DFT<2, DIR> {
  static __device__ __forceinline__ void run(float2* x) {
    float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};

template <int DIR>
struct DFT<4, DIR> {
  static __device__ __forceinline__ void run(float2* x) {
    float2 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
    float2 t2 = cadd(x[1], x[3]), t3 = mul_si<DIR>(csub(x[1], x[3]));
    x[0] = cadd(t0, t2);
    x[2] = csub(t0, t2);
    x[1] = cadd(t1, t3);
    x[3] = csub(t1, t3);
  }
};

template <int DIR>
struct DFT<8, DIR> {
  static __device__ __forceinline__ void run(float2* x) {
    // radix-2 split into even/odd 4-point DFTs
    float2 e[4] = {x[0], x[2], x[4], x[6]};
    float2 o[4] = {x[1], x[3], x[5], x[7]};
    DFT<4, DIR>::run(e);
    DFT<4, DIR>::run(o);
    constexpr float h = 0.70710678118654752440f;
    float2 o1 = (DIR > 0) ? make_float2((o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h)
                          : make_float2((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
    float2 o2 = mul_si<DIR>(o[2]);
    float2 o3 = (DIR > 0) ? make_float2(-(o[3].x + o[3].y) * h, (o[3].x - o[3].y) * h)
                          : make_float2((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
    x[0] = cadd(e[0], o[0]);
    x[4] = csub(e[0], o[0]);
    x[1] = cadd(e[1], o1);
    x[5] = csub(e[1], o1);
    x[2] = cadd(e[2], o2);
    x[6] = csub(e[2], o2);
    x[3] = cadd(e[3], o3);
    x[7] = csub(e[3], o3);
  }
};

// odd prime R: symmetric-pair form
template <int R, int DIR>
struct DFTPrime {
  static __device__ __forceinline__ void run(float2* x) {
    constexpr int H = (R - 1) / 2;
    float2 a[H], d[H];
    float2 x0 = x[0], s0 = x[0];
    static_for<1, H + 1>([&](auto J) {
      constexpr int j = decltype(J)::value;
      a[j - 1] = cadd(x[j], x[R - j]);
      d[j - 1] = csub(x[j], x[R - j]);
      s0 = cadd(s0, a[j - 1]);
    });
    static_for<1, H + 1>([&](auto K) {
      constexpr int k = decltype(K)::value;
      float2 re = x0, im = make_float2(0.f, 0.f);
      static_for<1, H + 1>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr int m = (j * k) % R;
        constexpr float c = Root<R, m>::c;
        constexpr float s = Root<R, m>::s;
        re.x += a[j - 1].x * c;
        re.y += a[j - 1].y * c;
        im.x += d[j - 1].x * s;
        im.y += d[j - 1].y * s;
      });
      float2 sim = mul_si<DIR>(im);
      x[k] = cadd(re, sim);
      x[R - k] = csub(re, sim);
    });
    x[0] = s0;
  }
};
template <int DIR> struct DFT<3, DIR> : DFTPrime<3, DIR> {};
template <int DIR> struct DFT<5, DIR> : DFTPrime<5, DIR> {};
template <int DIR> struct DFT<7, DIR> : DFTPrime<7, DIR> {};
template <int DIR> struct DFT<11, DIR> : DFTPrime<11, DIR> {};
template <int DIR> struct DFT<13, DIR> : DFTPrime<13, DIR> {};

// composite R = A*B: inputs x[jb + B*ja]; A-point DFTs over ja, twiddle
// w_R^{jb*ka}, B-point DFTs over jb; output index A*kb + ka.
template <int A, int B, int DIR>
struct DFTComp {
  static __device__ __forceinline__ void run(float2* x) {
    constexpr int R = A * B;
    float2 y[R];
    static_for<0, B>([&](auto JB) {
      constexpr int jb = decltype(JB)::value;
      float2 t[A];
      static_for<0, A>([&](auto JA) {
        constexpr int ja = decltype(JA)::value;
        t[ja] = x[jb + B * ja];
      });
      DFT<A, DIR>::run(t);
      static_for<0, A>([&](auto KA) {
        constexpr int ka = decltype(KA)::value;
        y[jb * A + ka] = mul_root<R, jb * ka, DIR>(t[ka]);
      });
    });
    static_for<0, A>([&](auto KA) {
      constexpr int ka = decltype(KA)::value;
      float2 u[B];
      static_for<0, B>([&](auto JB) {
        constexpr int jb = decltype(JB)::value;
        u[jb] = y[jb * A + ka];
      });
      DFT<B, DIR>::run(u);
      static_for<0, B>([&](auto KB) {
        constexpr int kb = decltype(KB)::value;
        x[A * kb + ka] = u[kb];
      });
    });
  }
};
template <int DIR> struct DFT<6, DIR> : DFTComp<2, 3, DIR> {};
template <int DIR> struct DFT<9, DIR> : DFTComp<3, 3, DIR> {};
template <int DIR> struct DFT<10, DIR> : DFTComp<2, 5, DIR> {};
template <int DIR> struct DFT<12, DIR> : DFTComp<4, 3, DIR> {};
template <int DIR> struct DFT<14, DIR> : DFTComp<2, 7, DIR> {};
template <int DIR> struct DFT<15, DIR> : DFTComp<3, 5, DIR> {};
template <int DIR> struct DFT<16, DIR> : DFTComp<4, 4, DIR> {};
template <int DIR> struct DFT<32, DIR> : DFTComp<8, 4, DIR> {};

// ---------------------------------------------------------------------------
// quad (4-lane) exchanges by DPP: lane l reads lane l ^ 1 or l ^ 2 of its quad
// ---------------------------------------------------------------------------
template <int X>
__device__ __forceinline__ float quad_xor(float v) {
  static_assert(X == 1 || X == 2, "quad partner");
  // quad_perm [1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E
  constexpr int CTRL = (X == 1) ? 0xB1 : 0x4E;
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
// radix-2 butterfly between lanes l and l ^ X: the lane with bit X clear
// (hi = false) gets a + p, the other p - a (a: own value, p: partner's)
template <int X>
__device__ __forceinline__ float2 quad_bfly(float2 a, float sgn) {
  const float px = quad_xor<X>(a.x), py = quad_xor<X>(a.y);
  return make_float2(fmaf(sgn, a.x, px), fmaf(sgn, a.y, py));
}

// ---------------------------------------------------------------------------
// FFT plans: the radix sequence for a line length N (product = N)
// ---------------------------------------------------------------------------
template <int... Rs>
struct Radices {
  static constexpr int S = sizeof...(Rs);
  static constexpr int R[S] = {Rs...};
};

template <int N> struct PlanFor;  // specialised per supported size
template <> struct PlanFor<16> { using type = Radices<4, 4>; };
template <> struct PlanFor<32> { using type = Radices<8, 4>; };
template <> struct PlanFor<48> { using type = Radices<8, 6>; };
template <> struct PlanFor<60> { using type = Radices<6, 10>; };
template <> struct PlanFor<64> { using type = Radices<8, 8>; };
template <> struct PlanFor<72> { using type = Radices<8, 9>; };
template <> struct PlanFor<80> { using type = Radices<8, 10>; };
template <> struct PlanFor<96> { using type = Radices<8, 12>; };
template <> struct PlanFor<100> { using type = Radices<10, 10>; };
template <> struct PlanFor<112> { using type = Radices<8, 14>; };
template <> struct PlanFor<120> { using type = Radices<8, 15>; };
template <> struct PlanFor<128> { using type = Radices<8, 16>; };
template <> struct PlanFor<144> { using type = Radices<12, 12>; };
template <> struct PlanFor<160> { using type = Radices<10, 16>; };
template <> struct PlanFor<192> { using type = Radices<12, 16>; };
template <> struct PlanFor<200> { using type = Radices<8, 5, 5>; };
template <> struct PlanFor<240> { using type = Radices<16, 15>; };
template <> struct PlanFor<256> { using type = Radices<16, 16>; };
template <> struct PlanFor<288> { using type = Radices<8, 4, 9>; };
template <> struct PlanFor<320> { using type = Radices<8, 8, 5>; };
template <> struct PlanFor<360> { using type = Radices<8, 9, 5>; };
template <> struct PlanFor<384> { using type = Radices<8, 8, 6>; };
template <> struct PlanFor<400> { using type = Radices<8, 10, 5>; };
template <> struct PlanFor<448> { using type = Radices<8, 8, 7>; };
template <> struct PlanFor<480> { using type = Radices<8, 6, 10>; };
template <> struct PlanFor<500> { using type = Radices<10, 10, 5>; };
template <> struct PlanFor<512> { using type = Radices<8, 8, 8>; };
template <> struct PlanFor<576> { using type = Radices<8, 8, 9>; };
template <> struct PlanFor<600> { using type = Radices<8, 15, 5>; };
template <> struct PlanFor<625> { using type = Radices<5, 5, 5, 5>; };
template <> struct PlanFor<640> { using type = Radices<8, 8, 10>; };
template <> struct PlanFor<720> { using type = Radices<8, 9, 10>; };
template <> struct PlanFor<768> { using type = Radices<8, 12, 8>; };
template <> struct PlanFor<800> { using type = Radices<8, 10, 10>; };
template <> struct PlanFor<960> { using type = Radices<8, 12, 10>; };
template <> struct PlanFor<1000> { using type = Radices<10, 10, 10>; };
template <> struct PlanFor<1024> { using type = Radices<8, 16, 8>; };
template <> struct PlanFor<1152> { using type = Radices<8, 16, 9>; };
template <> struct PlanFor<1280> { using type = Radices<8, 16, 10>; };
template <> struct PlanFor<1440> { using type = Radices<8, 12, 15>; };
template <> struct PlanFor<1536> { using type = Radices<8, 16, 12>; };
template <> struct PlanFor<2048> { using type = Radices<8, 16, 16>; };

template <class P, bool REV>
constexpr int radix_at(int s) { return REV ? P::R[P::S - 1 - s] : P::R[s]; }
template <class P, bool REV>
constexpr int stride_at(int s) {
  int p = 1;
  for (int i = 0; i < s; ++i) p *= radix_at<P, REV>(i);
  return p;
}
constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }
// registers per thread: max over stages of (butterflies per thread * radix)
template <int N, int T, class P>
constexpr int regs_needed() {
  int m = 0;
  for (int s = 0; s < P::S; ++s) {
    int r = P::R[s];
    int v = cdiv(N / r, T) * r;
    if (v > m) m = v;
  }
  return m;
}
template <class P>
constexpr int plan_product() {
  int p = 1;
  for (int s = 0; s < P::S; ++s) p *= P::R[s];
  return p;
}

// ---------------------------------------------------------------------------
// LDS access policy for a line.  Element e of line `l` lives at
//   ROWMAJOR: lds[l * STRIDE + pos(e)]
//   interleaved (columns): lds[pos(e) * LINES + l]
// with pos(e) one of
//   padded   (SW = 0): e + (e >> PADSH), one spare slot per 2^PADSH elements;
//   swizzled (SW = 1): e ^ h(e >> 5), an XOR of the low five bits by
//            h(r) = (r & 7) | ((r >> 1) & 3) << 3.
// gfx950 serves 8-byte LDS accesses 32 lanes per cycle over 64 four-byte
// banks, so an access is conflict free when the 32 lanes' 8-byte word
// addresses are distinct mod 32.  For the Stockham exchanges of a 64-thread
// line (reads of consecutive elements, writes at stride R*s) the swizzle is
// conflict free for power-of-two radix-8/16 plans, the padding is better for
// mixed radices; the per-size choice is made in Cfg.
//
// step(d) (padded policy only) is the index distance from element e0 to
// element e0 + d of the same line when the number of padding slots between
// them is fixed at compile time, i.e. when d is a multiple of the padding
// period 2^PADSH or e0 lies on a period boundary; then idx(l, e0 + d) =
// idx(l, e0) + step(d) and a thread's accesses share one base address with
// immediate offsets.
// ---------------------------------------------------------------------------
template <int SW, int PADSH>
struct LdsPos {
  static constexpr bool AFFINE = (SW == 0);
  // PADSH >= 20: no padding at all (every element at its own index, so any
  // stride gives constant offsets from one base)
  static constexpr bool FLAT = (SW == 0) && (PADSH >= 20);
  static constexpr int PAD = PADSH;
  __device__ static __forceinline__ int pos(int e) {
    if constexpr (FLAT) {
      return e;
    } else if constexpr (SW == 0) {
      return e + (e >> PADSH);
    } else {
      const int r = e >> 5;
      return e ^ ((r & 7) | (((r >> 1) & 3) << 3));
    }
  }
  // slots a line of n elements occupies
  static constexpr int span(int n) { return FLAT ? n : (SW == 0 ? n + (n >> PADSH) + 1 : ((n + 31) / 32) * 32); }
  static constexpr int step(int d) { return FLAT ? d : d + (d >> PADSH); }
};

template <int N, int PADSH, int SW = 0>
struct LineLayout : LdsPos<SW, PADSH> {  // one line per group of threads, contiguous
  using B = LdsPos<SW, PADSH>;
  static constexpr int STRIDE = B::span(N);
  __device__ static __forceinline__ int idx(int line, int e) { return line * STRIDE + B::pos(e); }
  static constexpr int words(int lines) { return lines * STRIDE; }
};
template <int N, int LINES, int PADSH>
struct TileLayout : LdsPos<0, PADSH> {  // LINES lines interleaved (line fastest)
  static constexpr int ROWS = N + (N >> PADSH) + 1;
  __device__ static __forceinline__ int idx(int line, int e) { return (e + (e >> PADSH)) * LINES + line; }
  static constexpr int step(int d) { return (d + (d >> PADSH)) * LINES; }
  static constexpr int words(int) { return ROWS * LINES; }
};

// element b + q*NBF of line `line` for q = 0..R-1 (the register slots of one
// butterfly), by a shared base when the layout allows it
template <class Layout, int NBF, int Q>
__device__ __forceinline__ int lds_slot(int line, int b, int a0) {
  if constexpr (Layout::AFFINE && (Layout::FLAT || NBF % (1 << Layout::PAD) == 0)) return a0 + Layout::step(Q * NBF);
  else return Layout::idx(line, b + Q * NBF);
}

// ---------------------------------------------------------------------------
// Line FFT driver.
//   v   : per-thread registers, stage-0 input layout on entry
//         (v[i*R0 + q] = x[b + q*N/R0], b = t + i*T), last-stage output layout
//         on exit (v[i*RL + k] = X[b + k*N/RL]).
//   lds : LDS scratch for this line group (through Layout::idx(line, e))
//   tw  : W_N^e = exp(-2 pi i e/N), e in [0, N) (forward twiddles)
//   SYNC: callable issuing the barrier appropriate for the line group
//   STOP: when >= 0, the transform returns right after writing the output of
//         stage STOP to LDS (no barrier), so that the caller can finish the
//         remaining stages with another thread mapping; -1 runs all stages.
// LineFFT<..., S>::read_input() loads stage S's input layout from LDS (the
// exchange read), for a caller that resumes a transform at stage S.
// ---------------------------------------------------------------------------
template <int N, int T, int DIR, class P, bool REV, class Layout, int NREG, bool TWR = false, int STOP = -1,
          int S = 0>
struct LineFFT {
  static __device__ __forceinline__ void read_input(float2 (&v)[NREG], const float2* lds, int line, int t) {
    constexpr int R = radix_at<P, REV>(S);
    constexpr int NBF = N / R;
    constexpr int NB = cdiv(NBF, T);
    static_for<0, NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if ((NB * T == NBF) || b < NBF) {
        const int a0 = Layout::idx(line, b);
        static_for<0, R>([&](auto Q) {
          constexpr int q = decltype(Q)::value;
          v[i * R + q] = lds[lds_slot<Layout, NBF, q>(line, b, a0)];
        });
      }
    });
  }

  template <class Sync>
  static __device__ __forceinline__ void run(float2 (&v)[NREG], float2* lds, int line,
                                             const float2* tw, int t, Sync&& sync) {
    constexpr int R = radix_at<P, REV>(S);
    constexpr int s = stride_at<P, REV>(S);
    constexpr int NBF = N / R;
    constexpr int NB = cdiv(NBF, T);
    constexpr bool LAST = (S == P::S - 1);
    static_assert(STOP < P::S - 1, "STOP must leave a stage to the caller");
    static_for<0, NB>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const int b = t + i * T;
      if ((NB * T == NBF) || b < NBF) {
        DFT<R, DIR>::run(&v[i * R]);
        if constexpr (!LAST) {
          const int base = (b / s) * s;
          if constexpr (TWR) {
            // one table read, W^(base*k) for k >= 2 by a product tree of
            // depth <= log2(k): w^k = w^(k/2) * w^(k - k/2)
            float2 pw[R];
            pw[1] = tw[base];
            static_for<2, R>([&](auto K) {
              constexpr int k = decltype(K)::value;
              pw[k] = cmul(pw[k / 2], pw[k - k / 2]);
            });
            static_for<1, R>([&](auto K) {
              constexpr int k = decltype(K)::value;
              v[i * R + k] = (DIR < 0) ? cmul(v[i * R + k], pw[k]) : cmulconj(v[i * R + k], pw[k]);
            });
          } else {
            static_for<1, R>([&](auto K) {
              constexpr int k = decltype(K)::value;
              const float2 w = tw[base * k];
              v[i * R + k] = (DIR < 0) ? cmul(v[i * R + k], w) : cmulconj(v[i * R + k], w);
            });
          }
        }
      }
    });
    if constexpr (!LAST) {
      static_for<0, NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const int b = t + i * T;
        if ((NB * T == NBF) || b < NBF) {
          const int o0 = (b % s) + (b / s) * s * R;
          // o0 is on a padding-period boundary when s == 1 and R is a multiple
          // of the period; s*k is a multiple of it when s is
          constexpr bool CONST_OFF = Layout::AFFINE && (Layout::FLAT || (s % (1 << Layout::PAD) == 0) ||
                                                        (s == 1 && R % (1 << Layout::PAD) == 0));
          const int a0 = Layout::idx(line, o0);
          static_for<0, R>([&](auto K) {
            constexpr int k = decltype(K)::value;
            if constexpr (CONST_OFF) lds[a0 + Layout::step(s * k)] = v[i * R + k];
            else lds[Layout::idx(line, o0 + s * k)] = v[i * R + k];
          });
        }
      });
      if constexpr (S != STOP) {
        sync();
        LineFFT<N, T, DIR, P, REV, Layout, NREG, TWR, STOP, S + 1>::read_input(v, lds, line, t);
        sync();
        LineFFT<N, T, DIR, P, REV, Layout, NREG, TWR, STOP, S + 1>::run(v, lds, line, tw, t, sync);
      }
    }
  }
};

// element index held in register slot (i, q) for the first stage of a plan
// (REV selects the reversed radix order); the same mapping describes the last
// stage of the opposite plan.
template <int N, int T, class P, bool REV>
struct IOMap {
  static constexpr int R0 = radix_at<P, REV>(0);
  static constexpr int NBF = N / R0;
  static constexpr int NB = cdiv(NBF, T);
  static constexpr bool FULL = (NB * T == NBF);
};

}  // namespace jms
