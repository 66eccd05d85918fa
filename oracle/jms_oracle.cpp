// jms_oracle.cpp -- CPU oracle for the multislice slice loop.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the host, the algorithm of
// the reference's CPU multislice path (CJMultiSlice::CPUMultislice and the
// functions it calls) so that tests, the smoke check and the CPU-baseline leg of
// the benchmark can check the HIP product against it.  Nothing in the product
// library links, loads or calls this code.
//
// Every function cites the reference file:line it follows (paths relative to
// the reference repository root).  The two vendor libraries the reference path
// relies on for its FFT (Intel MKL DFTI, JFFTMKLcore.cpp:83-221; cuFFT,
// JFFTCUDAcore.cpp:96-336) are absent; the transform here is this project's own
// mixed-radix Stockham FFT (self-sorting, decimation in frequency), written from
// the published method (Stockham 1966; formulation as in Van Loan, "Computational
// Frameworks for the FFT", 1992), with the reference's conventions: forward =
// exp(-2 pi i nk/N), backward = exp(+2 pi i nk/N), no scaling either way.
//
// Both a float32 path (the reference arithmetic type) and a float64 twin (for the
// "error against FP64" statement) are provided through templates.
//
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math).

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <complex>
#include <thread>
#include <vector>

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

// --------------------------------------------------------------------------
// Physical constants, CODATA 2018 (NatureConstants.h:36-65)
// --------------------------------------------------------------------------
const double kPI = 3.1415926535897932384626433832795;   // NatureConstants.h:36
const double kTPI = 6.283185307179586476925286766559;   // NatureConstants.h:37
const double kD2R = kPI / 180.;                          // NatureConstants.h:42
const double kC0 = 299792458.;                           // NatureConstants.h:48
const double kHPL = 6.62607015E-34;                      // NatureConstants.h:51
const double kQEL = 1.602176634E-19;                     // NatureConstants.h:52
const double kME0 = 9.1093837015E-31;                    // NatureConstants.h:60
const double kEEL0KEV = kME0 * kC0 * kC0 / kQEL / 1000.; // NatureConstants.h:64
const double kWLELKEV = kC0 * kHPL / kQEL * 1.0E+6;      // NatureConstants.h:66
const double kRELAPERTURE = 0.666666666666667;          // JMultiSlice.h:157
const size_t kSUMBUF = 0x1000;                           // JMultiSlice.h:153 (_JMS_SUMMATION_BUFFER)
const size_t kSUMBTF = 32;                               // JMultiSlice.h:154 (_JMS_SUMMATION_BTF_THR)

// --------------------------------------------------------------------------
// Mixed-radix Stockham FFT (self-sorting, decimation in frequency).
//
// One stage with radix r on a length-n sub-problem with stride s (s = product of
// the radices already applied) maps, for p in [0, n/r) and q in [0, s):
//   y[q + s*(r*p + k)] = w_n^{p k} * sum_j x[q + s*(p + j*n/r)] * w_r^{j k}
// and recurses on length n/r with stride r*s.  After the last stage the output
// is in natural order.  The q index is the innermost loop, which also lets one
// call transform many interleaved sequences (all columns of a 2-D array at once)
// by starting with s = number of sequences.
// --------------------------------------------------------------------------
template <typename T>
struct FFTPlan {
  int n = 0;
  std::vector<int> radix;                          // stage radices, product = n
  std::vector<std::vector<std::complex<T>>> tw;    // per stage: w_n^{p k}, [p*r + k]
  std::vector<std::vector<std::complex<T>>> wr;    // per stage: w_r^{m}, m in [0,r)
};

// Radix order: 4s first, then 2, 3, 5, 7, then any remaining prime factor.
inline std::vector<int> factorize(int n) {
  std::vector<int> f;
  while (n % 4 == 0) { f.push_back(4); n /= 4; }
  while (n % 2 == 0) { f.push_back(2); n /= 2; }
  for (int p = 3; p <= n; p += 2) {
    while (n % p == 0) { f.push_back(p); n /= p; }
  }
  return f;
}

template <typename T>
FFTPlan<T> make_plan(int n) {
  FFTPlan<T> pl;
  pl.n = n;
  if (n <= 1) return pl;
  pl.radix = factorize(n);
  int len = n;
  for (int r : pl.radix) {
    int m = len / r;
    std::vector<std::complex<T>> t((size_t)m * r), w((size_t)r);
    for (int p = 0; p < m; ++p)
      for (int k = 0; k < r; ++k) {
        double a = -kTPI * (double)((long long)p * k % len) / (double)len;
        t[(size_t)p * r + k] = std::complex<T>((T)std::cos(a), (T)std::sin(a));
      }
    for (int k = 0; k < r; ++k) {
      double a = -kTPI * (double)k / (double)r;
      w[k] = std::complex<T>((T)std::cos(a), (T)std::sin(a));
    }
    pl.tw.push_back(std::move(t));
    pl.wr.push_back(std::move(w));
    len = m;
  }
  return pl;
}

// complex multiply written out (std::complex operator* may go through
// __mulsc3); for finite values this is the formula the reference's
// fcmplx '*=' evaluates (JFFTMKLcore.cpp:397-399).
template <typename T>
inline std::complex<T> cmul(std::complex<T> a, std::complex<T> b) {
  return std::complex<T>(a.real() * b.real() - a.imag() * b.imag(),
                         a.real() * b.imag() + a.imag() * b.real());
}
template <typename T>
inline std::complex<T> conj_if(std::complex<T> a, bool c) {
  return c ? std::complex<T>(a.real(), -a.imag()) : a;
}

// Transform `howmany` interleaved sequences of length n: element i of sequence
// q sits at x[q + howmany*i].  dir < 0: forward exp(-i..), dir > 0: backward.
// `work` must hold n*howmany complex values.  Result is left in x.
template <typename T>
void stockham(const FFTPlan<T>& pl, std::complex<T>* x, std::complex<T>* work,
              int howmany, int dir) {
  const int n = pl.n;
  if (n <= 1) return;
  const bool inv = dir > 0;
  std::complex<T>* src = x;
  std::complex<T>* dst = work;
  int len = n;
  size_t s = (size_t)howmany;
  for (size_t st = 0; st < pl.radix.size(); ++st) {
    const int r = pl.radix[st];
    const int m = len / r;
    const std::complex<T>* tw = pl.tw[st].data();
    const std::complex<T>* wr = pl.wr[st].data();
    for (int p = 0; p < m; ++p) {
      const std::complex<T>* twp = tw + (size_t)p * r;
      if (r == 2) {
        const std::complex<T> w1 = conj_if(twp[1], inv);
        for (size_t q = 0; q < s; ++q) {
          std::complex<T> a = src[q + s * p], b = src[q + s * (p + m)];
          dst[q + s * (2 * p)] = a + b;
          dst[q + s * (2 * p + 1)] = cmul(a - b, w1);
        }
      } else if (r == 4) {
        const std::complex<T> w1 = conj_if(twp[1], inv), w2 = conj_if(twp[2], inv),
                              w3 = conj_if(twp[3], inv);
        for (size_t q = 0; q < s; ++q) {
          std::complex<T> a0 = src[q + s * p], a1 = src[q + s * (p + m)],
                          a2 = src[q + s * (p + 2 * m)], a3 = src[q + s * (p + 3 * m)];
          std::complex<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = a1 - a3;
          // t3 * (-i) for forward, t3 * (+i) for backward
          std::complex<T> t3r = inv ? std::complex<T>(-t3.imag(), t3.real())
                                    : std::complex<T>(t3.imag(), -t3.real());
          dst[q + s * (4 * p)] = t0 + t2;
          dst[q + s * (4 * p + 1)] = cmul(t1 + t3r, w1);
          dst[q + s * (4 * p + 2)] = cmul(t0 - t2, w2);
          dst[q + s * (4 * p + 3)] = cmul(t1 - t3r, w3);
        }
      } else {
        std::complex<T> a[64];
        for (size_t q = 0; q < s; ++q) {
          for (int j = 0; j < r; ++j) a[j] = src[q + s * (p + (size_t)j * m)];
          for (int k = 0; k < r; ++k) {
            std::complex<T> acc = a[0];
            for (int j = 1; j < r; ++j)
              acc += cmul(a[j], conj_if(wr[(long long)j * k % r], inv));
            dst[q + s * ((size_t)r * p + k)] =
                (k == 0) ? acc : cmul(acc, conj_if(twp[k], inv));
          }
        }
      }
    }
    std::swap(src, dst);
    s *= (size_t)r;
    len = m;
  }
  if (src != x) std::memcpy(x, src, sizeof(std::complex<T>) * (size_t)n * howmany);
}

// 2-D transform of an ny x nx array, x fastest (index i + j*nx), in place.
// Rows (x, length nx) are transformed first, then columns (y, length ny); the
// reference's DFTI descriptor over {ny, nx} computes the same separable DFT
// (JFFTMKLcore.cpp:117-142).
template <typename T>
struct FFT2D {
  int nx = 0, ny = 0;
  FFTPlan<T> px, py;
  std::vector<std::complex<T>> work;
  void init(int nx_, int ny_) {
    nx = nx_; ny = ny_;
    px = make_plan<T>(nx);
    py = make_plan<T>(ny);
    work.assign((size_t)nx * ny, std::complex<T>(0, 0));
  }
  void run(std::complex<T>* d, int dir) {
    for (int j = 0; j < ny; ++j) stockham(px, d + (size_t)j * nx, work.data(), 1, dir);
    stockham(py, d, work.data(), nx, dir);
  }
};

// --------------------------------------------------------------------------
// Mersenne twister MT19937 (Matsumoto & Nishimura 1998), the engine behind the
// reference's CRng (rng.h:33-37, std::mt19937).  CRng::rand() draws
// uniform_int_distribution<int>(0, 0x7FFFFFFF) from it; in libstdc++ (GCC >= 11)
// that is exactly the engine output shifted right by one bit (Lemire's method
// with a power-of-two range), which is what rand() returns here.
// --------------------------------------------------------------------------
struct MT19937 {
  uint32_t mt[624];
  int idx = 624;
  void seed(uint32_t s) {
    mt[0] = s;
    for (int i = 1; i < 624; ++i)
      mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    idx = 624;
  }
  uint32_t next() {
    if (idx >= 624) {
      for (int i = 0; i < 624; ++i) {
        uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
        mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      }
      idx = 0;
    }
    uint32_t y = mt[idx++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
  }
};

struct OrcRng {
  MT19937 e;
  // Marsaglia polar method state (libstdc++ std::normal_distribution caches
  // the second deviate of each accepted pair).
  bool saved_ok = false;
  double saved = 0.;
  // rng.cpp:50-59 (seed 0 means "from time"; the oracle requires explicit seeds)
  void seed(int s) { e.seed((uint32_t)s); saved_ok = false; }
  // rng.cpp:66-69
  int rand() { return (int)(e.next() >> 1); }
  // rng.cpp:71-74
  double unirand(double rmin, double rmax) {
    return rmin + ((double)rand() / (double)((unsigned long)0x7FFFFFFF)) * (rmax - rmin);
  }
  // generate_canonical<double,53> over a 32-bit engine: two draws,
  // (x0 + x1 * 2^32) / 2^64, clamped below 1.
  double canonical() {
    const double r = 4294967296.0;
    double sum = (double)e.next();
    sum += (double)e.next() * r;
    double res = sum / (r * r);
    if (res >= 1.0) res = std::nextafter(1.0, 0.0);
    return res;
  }
  // rng.cpp:81-84: mu + sdev * N(0,1); N(0,1) by the Marsaglia polar method as
  // std::normal_distribution<double> evaluates it (Marsaglia & Bray 1964).
  double normrand(double mu, double sdev) {
    double ret;
    if (saved_ok) {
      saved_ok = false;
      ret = saved;
    } else {
      double x, y, r2;
      do {
        x = 2.0 * canonical() - 1.0;
        y = 2.0 * canonical() - 1.0;
        r2 = x * x + y * y;
      } while (r2 > 1.0 || r2 == 0.0);
      const double mult = std::sqrt(-2.0 * std::log(r2) / r2);
      saved = x * mult;
      saved_ok = true;
      ret = y * mult;
    }
    return mu + sdev * ret;
  }
};

inline int imod(int a, int b) { int r = a % b; return r >= 0 ? r : r + b; }  // JMultiSlice.cpp:35

// --------------------------------------------------------------------------
// Summation helpers, JMultiSlice.cpp:38-242
// --------------------------------------------------------------------------
void ffstrsum(const float* a, size_t n, float* s) {  // JMultiSlice.cpp:38-49
  float t = 0.f;
  for (size_t i = 0; i < n; i++) t += a[i];
  *s = t;
}

// strided 2-fold butterfly summation, modifies a (JMultiSlice.cpp:80-157)
void fdncs2m(float* a, size_t n, float* s) {
  if (n == 0) { *s = 0.f; return; }
  if (n == 1) { *s = a[0]; return; }
  if (n == 2) { *s = a[0] + a[1]; return; }
  if (n < kSUMBTF) { ffstrsum(a, n, s); return; }
  *s = 0.f;
  size_t n0 = 0, n1 = 1, n2 = 2, nc = 0, idx = 0, itmp = 0;
  size_t ntmp = (size_t)std::ceil((double)n / (double)kSUMBUF);
  float r = 0.0f;
  float* dst = s;
  if (ntmp > 1) {
    dst = (float*)calloc(ntmp, sizeof(float));
  } else {
    r = 0.0f; n2 = 2; n1 = 1;
    while (n2 <= n) {
      for (idx = n2 - 1; idx < n; idx += n2) a[idx] += a[idx - n1];
      if (n1 <= n % n2) r += a[idx - n1];
      n1 = n2; n2 = n2 << 1;
    }
    *s += (a[n1 - 1] + r);
    return;
  }
  float* tmp = a;
  while (n0 < n) {
    nc = std::min(kSUMBUF, n - n0);
    if (nc == kSUMBUF) {
      tmp = &a[n0];
      n2 = 2; n1 = 1;
      while (n2 < kSUMBUF) {
        for (idx = n2 - 1; idx < kSUMBUF; idx += n2) tmp[idx] += tmp[idx - n1];
        n1 = n2; n2 = n2 << 1;
      }
      dst[itmp] += (tmp[n1 - 1] + tmp[n2 - 1]);
    } else {
      if (1 == nc) { dst[itmp] = a[n0]; }
      else if (2 == nc) { dst[itmp] = a[n0] + a[1 + n0]; }
      else if (kSUMBTF > nc) { ffstrsum(&a[n0], nc, &dst[itmp]); }
      else {
        tmp = &a[n0];
        n2 = 2; n1 = 1;
        while (n2 <= nc) {
          for (idx = n2 - 1; idx < nc; idx += n2) tmp[idx] += tmp[idx - n1];
          if (n1 <= nc % n2) r += tmp[idx - n1];
          n1 = n2; n2 = n2 << 1;
        }
        dst[itmp] += (tmp[n1 - 1] + r);
      }
    }
    n0 += kSUMBUF;
    itmp++;
  }
  if (ntmp > 1) {
    fdncs2m(dst, ntmp, s);
    free(dst);
  }
}

// CJMultiSlice::MaskedDotProduct_h (JMultiSlice.cpp:3303-3330)
float masked_dot(const int* mask, const float* in1, const float* in2, size_t lenmask) {
  float fsum = 0.0f;
  if (lenmask == 0) return fsum;
  size_t nodd = lenmask % 2, neve = lenmask - nodd, nlen2 = 1 + neve / 2, idx = 0;
  std::vector<float> fdat(nlen2, 0.f);
  for (size_t i = 0; i < neve; i += 2) {
    unsigned m0 = (unsigned)mask[i], m1 = (unsigned)mask[i + 1];
    fdat[idx] = in1[m0] * in2[m0] + in1[m1] * in2[m1];
    idx++;
  }
  if (nodd) {
    unsigned m0 = (unsigned)mask[lenmask - 1];
    fdat[idx] = in1[m0] * in2[m0];
  }
  fdncs2m(fdat.data(), nlen2, &fsum);
  return fsum;
}

// CJMultiSlice::DotProduct_h (JMultiSlice.cpp:3279-3301); fdncs2 is fdncs2m on a copy
float plain_dot(const float* in1, const float* in2, size_t len) {
  float fsum = 0.0f;
  if (len == 0) return fsum;
  size_t nodd = len % 2, nlen2 = (len + nodd) / 2, idx = 0;
  std::vector<float> fdat(nlen2, 0.f);
  for (size_t i = 0; i + 1 < len; i += 2) {
    fdat[idx] = in1[i] * in2[i] + in1[i + 1] * in2[i + 1];
    idx++;
  }
  if (nodd) fdat[idx] = in1[len - 1] * in2[len - 1];
  fdncs2m(fdat.data(), nlen2, &fsum);
  return fsum;
}

// Double-precision twins of the dot products (plain double accumulation), used
// by the FP64 re-run of the recipe.
double masked_dot_d(const int* mask, const double* in1, const float* in2, size_t lenmask) {
  double s = 0.;
  for (size_t i = 0; i < lenmask; ++i) s += in1[mask[i]] * (double)in2[mask[i]];
  return s;
}
double plain_dot_d(const double* in1, const float* in2, size_t len) {
  double s = 0.;
  for (size_t i = 0; i < len; ++i) s += in1[i] * (double)in2[i];
  return s;
}

}  // namespace

// ==========================================================================
// Exported oracle functions (C ABI for the test harness)
// ==========================================================================

// CJMultiSlice::SetHighTension (JMultiSlice.cpp:372-378): wavelength in nm
ORC_API float orc_wavelength(float ht) {
  return (float)(kWLELKEV / std::sqrt(ht * (ht + 2. * kEEL0KEV)));
}

// 1-D transform of `howmany` interleaved sequences (element i of sequence q at
// q + howmany*i), in place.  data: interleaved re/im.
ORC_API int orc_fft1d_f32(float* data, int n, int howmany, int dir) {
  if (n < 1 || howmany < 1 || !data) return 1;
  FFTPlan<float> pl = make_plan<float>(n);
  std::vector<std::complex<float>> w((size_t)n * howmany);
  stockham(pl, (std::complex<float>*)data, w.data(), howmany, dir);
  return 0;
}
ORC_API int orc_fft1d_f64(double* data, int n, int howmany, int dir) {
  if (n < 1 || howmany < 1 || !data) return 1;
  FFTPlan<double> pl = make_plan<double>(n);
  std::vector<std::complex<double>> w((size_t)n * howmany);
  stockham(pl, (std::complex<double>*)data, w.data(), howmany, dir);
  return 0;
}
// 2-D in-place transform of ny rows of nx (x fastest)
ORC_API int orc_fft2d_f32(float* data, int nx, int ny, int dir) {
  if (nx < 1 || ny < 1 || !data) return 1;
  FFT2D<float> f; f.init(nx, ny);
  f.run((std::complex<float>*)data, dir);
  return 0;
}
ORC_API int orc_fft2d_f64(double* data, int nx, int ny, int dir) {
  if (nx < 1 || ny < 1 || !data) return 1;
  FFT2D<double> f; f.init(nx, ny);
  f.run((std::complex<double>*)data, dir);
  return 0;
}

// CJMultiSlice::CalculatePropagator (JMultiSlice.cpp:998-1106).
// a0x, a0y [nm], wl [nm], fthick [nm], otx, oty [deg]; ntype 0 planar, 1 Fresnel.
// pro: 2*nx*ny floats (interleaved), scrambled (FFT) order, x fastest.
ORC_API int orc_propagator(int nx, int ny, float a0x, float a0y, float wlf, float fthick,
                           float otx, float oty, int ntype, float* pro) {
  if (!pro || ntype < 0 || ntype > 1) return 1;
  const int nyqx = (nx - (nx % 2)) >> 1, nyqy = (ny - (ny % 2)) >> 1;
  const double itogx = 1. / a0x, itogy = 1. / a0y;
  const double gmax = std::min(itogx * nyqx, itogy * nyqy);
  const double gthr = gmax * kRELAPERTURE, gthr2 = gthr * gthr;
  const double otxr = (double)otx * kD2R, otyr = (double)oty * kD2R;
  const double wl = (double)wlf;
  const double ot = std::sqrt(otxr * otxr + otyr * otyr);
  const float pronorm = 1.f / ((float)nx * ny);
  if (ntype == 0) {  // planar, :1021-1056
    const double pfac = 2. * kPI / wl, cot = std::cos(ot), sot = std::sin(ot);
    double od = 0.;
    if (ot > 0.0) od = std::atan2(otyr, otxr);
    for (int j = 0; j < ny; j++) {
      const int j1 = ((j + nyqy) % ny) - nyqy;
      const double gy = itogy * j1, gy2 = gy * gy;
      for (int i = 0; i < nx; i++) {
        const int i1 = ((i + nyqx) % nx) - nyqx;
        const double gx = itogx * i1, g2 = gx * gx + gy2;
        float* o = pro + 2 * ((size_t)i + (size_t)j * nx);
        if (g2 > gthr2) { o[0] = 0.f; o[1] = 0.f; continue; }
        const double dt = std::asin(wl * std::sqrt(g2));
        double dd = 0.;
        if (dt > 0.) dd = std::atan2(gy, gx);
        const double chi = pfac * (double)fthick *
                           (1. / (cot * std::cos(dt) + sot * std::sin(dt) * std::cos(od - dd)) - 1. / cot);
        o[0] = (float)std::cos(chi) * pronorm;
        o[1] = (float)(-std::sin(chi) * pronorm);
      }
    }
  } else {  // Fresnel, :1058-1103
    const double pfac = kPI / wl * std::cos(ot);
    const double pt0 = pfac * (otxr * otxr + otyr * otyr);
    const double itowx = itogx * wl, itowy = itogy * wl;
    const double wthr2 = gthr2 * wl * wl;
    for (int j = 0; j < ny; j++) {
      const int j1 = ((j + nyqy) % ny) - nyqy;
      const double wy = itowy * j1, wy2 = wy * wy, uy = wy - otyr, uy2 = uy * uy;
      for (int i = 0; i < nx; i++) {
        const int i1 = ((i + nyqx) % nx) - nyqx;
        const double wx = itowx * i1, w2 = wx * wx + wy2, ux = wx - otxr, u2 = ux * ux + uy2;
        float* o = pro + 2 * ((size_t)i + (size_t)j * nx);
        if (w2 > wthr2) { o[0] = 0.f; o[1] = 0.f; continue; }
        const double chi = (double)fthick * (pfac * u2 - pt0);
        o[0] = (float)std::cos(chi) * pronorm;
        o[1] = (float)(-std::sin(chi) * pronorm);
      }
    }
  }
  return 0;
}

// CJMultiSlice::GetRadialDetectorSensitivity (JMultiSlice.cpp:1247-1272)
static float radial_sens(float theta, const float* profile, int len, float refpix, float beta1) {
  float sens = 1.0f;
  float rpix = std::max(0.f, refpix - 1.f);
  if (profile && len > 0 && rpix >= 0.f) {
    float tpix = std::max(0.f, theta * rpix / beta1);
    int ipix = (int)tpix;
    float tfrac = tpix - ipix;
    if (ipix > (len - 2)) sens = profile[len - 1];
    else sens = (1.f - tfrac) * profile[ipix] + tfrac * profile[ipix + 1];
  }
  return sens;
}

// CJMultiSlice::CalculateRingDetector (JMultiSlice.cpp:1274-1438).  The
// sensitivity profile is passed as an array (profile, plen, refpix) instead of
// a file name; plen = 0 means no profile.  det: nx*ny floats; msk: nx*ny ints
// (may be NULL); *msklen receives the mask length.  nprotype selects the
// angle -> k mapping (1: Fresnel, small angle; otherwise planar).
ORC_API int orc_ring_detector(int nx, int ny, float a0x, float a0y, float wlf, int nprotype,
                              float beta0, float beta1, float phi0, float phi1,
                              float theta0x, float theta0y, const float* profile, int plen,
                              float refpix, float* det, int* msklen, int* msk, int kmom1) {
  const int nitems = nx * ny;
  *msklen = 0;
  std::memset(det, 0, sizeof(float) * nitems);
  if (msk) std::memset(msk, 0, sizeof(int) * nitems);
  const bool busesprof = (profile != nullptr && plen > 0);
  const int nyqx = (nx - (nx % 2)) >> 1, nyqy = (ny - (ny % 2)) >> 1;
  const double itogx = 1. / a0x, itogy = 1. / a0y;
  const double wl = (double)wlf, k0 = 1. / wl;
  double dr1 = (double)beta0 * 0.001, dr2 = (double)beta1 * 0.001;
  double da1 = (double)phi0 * kD2R, da2 = (double)phi1 * kD2R;
  const double dc1 = (double)theta0x * 0.001, dc2 = (double)theta0y * 0.001;
  double dgc1 = 0., dgc2 = 0.;
  if (nprotype == 1) {
    // :1331-1336 -- in Fresnel mode the detector centre stays at 0 (dgc1/2 are
    // scaled but never set from dc1/2); restated as found.
    dr1 *= k0; dr2 *= k0; dgc1 *= k0; dgc2 *= k0;
  } else {
    dr1 = k0 * std::sin(dr1);
    dr2 = k0 * std::sin(dr2);
    double phi = 0.;
    double gy2 = k0 * std::sin(std::sqrt(dc1 * dc1 + dc2 * dc2));
    if (gy2 > 0.) phi = std::atan2(dc2, dc1);
    dgc1 = gy2 * std::cos(phi);
    dgc2 = gy2 * std::sin(phi);
  }
  int nring = 0;
  if (std::abs(da1 - da2) < 0.0001) nring++;
  if (std::abs(std::abs(da1 - da2) - kTPI) < 0.0001) nring++;
  if (nring > 0) nring = 1;
  if (nring == 0) {
    while (da1 < 0.0 || da2 < 0.0) { da1 += kTPI; da2 += kTPI; }
    if (da1 >= da2) da2 += kTPI;
    while (da1 >= kTPI) { da1 -= kTPI; da2 -= kTPI; }
  }
  int ml = 0;
  for (int j = 0; j < ny; j++) {
    const int j1 = ((j + nyqy) % ny) - nyqy;
    const double gy = itogy * j1 - dgc2, gy2 = gy * gy;
    for (int i = 0; i < nx; i++) {
      const int i1 = ((i + nyqx) % nx) - nyqx;
      const double gx = itogx * i1 - dgc1;
      const int idx = i + j * nx;
      const double g = std::sqrt(gy2 + gx * gx);
      double phi = 0.;
      if (g > 0.) phi = std::atan2(gy, gx);
      if (g < dr1 || g >= dr2) continue;
      if (nring == 0) {
        while (phi < da1) phi += kTPI;
        if (phi < da1 || phi >= da2) continue;
      }
      float sens = 1.0f;
      if (busesprof) {
        if (dr1 > 0.) sens = radial_sens((float)g, profile, plen, refpix, (float)dr1);
        else if (dr2 > 0.) sens = radial_sens((float)g, profile, plen, refpix, (float)dr2);
      }
      if (kmom1 == 1) sens = sens * (float)gx;
      if (kmom1 == 2) sens = sens * (float)gy;
      det[idx] = sens;
      if (msk) { msk[ml] = idx; ml++; }
    }
  }
  *msklen = ml;
  return 0;
}

// Axial aberration table of CJProbeGen (JProbeGen.cpp:187-245): aberration k
// of the interface list has order m and rotational symmetry n, enumerated
// for m = 1..8, n = 0..m with m + n even: (1,1) image shift, (2,0) defocus,
// (2,2) 2-fold astigmatism, (3,1) coma, (3,3), (4,0) spherical, ...
static const int kAbOrderMax = 8;  // _JPG_ABERRATION_ORDER_MAX (JProbeGen.h:53)
struct AbTable {
  int num = 0;
  std::vector<int> widx;   // (m, n) pairs
  std::vector<int> binom;  // binom[n + k*nol2] = C(n, k), nol2 = 2*8+1
  AbTable() {
    num = (4 * kAbOrderMax - kAbOrderMax % 2 + kAbOrderMax * kAbOrderMax) / 4;
    widx.assign(2 * num, 0);
    int l = 0;
    for (int m = 1; m <= kAbOrderMax; m++)
      for (int n = 0; n <= m; n++)
        if ((m + n) % 2 == 0) { widx[2 * l] = m; widx[2 * l + 1] = n; l++; }
    const int nol = 1 + 2 * kAbOrderMax;
    binom.assign(nol * nol, 0);
    for (int k = 0; k < nol; k++)
      for (int n = 0; n < nol; n++) {
        int res = 0;
        if (n >= k) {
          res = 1;
          const int lk = (k > n - k) ? n - k : k;
          for (int i = 0; i < lk; ++i) { res *= (n - i); res /= (i + 1); }
        }
        binom[n + k * nol] = res;
      }
  }
};
static const AbTable& ab_table() { static const AbTable t; return t; }

// CJProbeGen::AberrationFunction (JProbeGen.cpp:508-571): the phase chi [rad]
// at spatial frequency (qx, qy) [1/nm] for nabrr coefficient pairs (x, y) [nm]
// of the interface list: chi = 2pi/wl * sum_k Re[(ax - i ay)(wx + i wy)^n]
// |w|^(m-n) / m with w = wl * q, evaluated term by term in float.
static float aberration_chi(float qx, float qy, float wl, int nabrr, const float* coeff) {
  if (nabrr == 0 || !coeff || wl <= 0.f) return 0.f;
  const AbTable& T = ab_table();
  const int nol = kAbOrderMax + 1, nol2 = kAbOrderMax * 2 + 1;
  const int nab = std::min(nabrr, T.num);
  const float sgnprm[4] = {1.f, 0.f, -1.f, 0.f};
  const float wx = qx * wl, wy = qy * wl;
  const float w = std::sqrt(wx * wx + wy * wy);
  const float prefac = (float)(2.0 * kPI) / wl;
  std::vector<float> wfield(2 * nol), wabs(nol);
  float pwx = 1.f, pwy = 1.f, pw = 1.f;
  wfield[0] = pwx; wfield[1] = pwy; wabs[0] = pw;
  for (int i = 1; i < nol; i++) {
    pwx *= wx; pwy *= wy; pw *= w;
    wfield[2 * i] = pwx; wfield[2 * i + 1] = pwy; wabs[i] = pw;
  }
  float rtmp = 0.f;
  for (int k = 0; k < nab; k++) {
    const float wax = coeff[2 * k], way = coeff[2 * k + 1];
    if ((wax * wax + way * way) > (float)(1.E-15)) {  // _JPG_ABERRATION_STRTHRESH (JProbeGen.h:54)
      const int m = T.widx[2 * k], n = T.widx[2 * k + 1];
      float ttmp = 0.f;
      for (int l = 0; l <= n; l++) {
        float ttmp1 = 0.f;
        ttmp1 = ttmp1 + sgnprm[l % 4] * wax;
        ttmp1 = ttmp1 + sgnprm[(l + 3) % 4] * way;
        const int j = n - l;
        ttmp = ttmp + ttmp1 * (float)T.binom[n + l * nol2] * wfield[2 * j] * wfield[2 * l + 1];
      }
      ttmp = ttmp * wabs[m - n] / ((float)m);
      rtmp = rtmp + ttmp;
    }
  }
  return rtmp * prefac;
}

ORC_API float orc_aberration_chi(float qx, float qy, float wl, int nabrr, const float* coeff) {
  return aberration_chi(qx, qy, wl, nabrr, coeff);
}

ORC_API int orc_aberration_table(int* num, int* widx) {
  const AbTable& T = ab_table();
  if (num) *num = T.num;
  if (widx) std::copy(T.widx.begin(), T.widx.end(), widx);
  return 0;
}

// CJProbeGen::CalculateProbeWaveFourier (JProbeGen.cpp:707-777) with
// ApertureFunctionA (JProbeGen.cpp:471-505) and the aberration phase chi of
// nabrr coefficient pairs (none: chi = 0).  Angles in mrad, adir in degrees.
// wav: 2*nx*ny floats.
ORC_API int orc_probe_wave_ab(int nx, int ny, float ax, float ay, float wl, float alpha,
                              float alpha_rs, float alpha_asym, float alpha_adir,
                              float alpha_x0, float alpha_y0, float btx_mrad, float bty_mrad,
                              int nabrr, const float* abrr, float* wav);

ORC_API int orc_probe_wave(int nx, int ny, float ax, float ay, float wl, float alpha,
                           float alpha_rs, float alpha_asym, float alpha_adir,
                           float alpha_x0, float alpha_y0, float btx_mrad, float bty_mrad,
                           float* wav) {
  return orc_probe_wave_ab(nx, ny, ax, ay, wl, alpha, alpha_rs, alpha_asym, alpha_adir, alpha_x0, alpha_y0,
                           btx_mrad, bty_mrad, 0, nullptr, wav);
}

ORC_API int orc_probe_wave_ab(int nx, int ny, float ax, float ay, float wl, float alpha,
                              float alpha_rs, float alpha_asym, float alpha_adir,
                              float alpha_x0, float alpha_y0, float btx_mrad, float bty_mrad,
                              int nabrr, const float* abrr, float* wav) {
  if (!wav || wl <= 0.f || ax * ay <= 0.f || nx * ny <= 0 || alpha <= 0.f) return 1;
  const int nx2 = (nx - (nx % 2)) >> 1, ny2 = (ny - (ny % 2)) >> 1;
  const float qlim = alpha * 0.001f / wl;
  const float qrs = alpha_rs;
  const float alim = alpha_asym;
  const float adir = alpha_adir * 0.0174533f;
  const float qlcx = alpha_x0 * 0.001f / wl, qlcy = alpha_y0 * 0.001f / wl;
  const float btx = btx_mrad * 0.001f / wl, bty = bty_mrad * 0.001f / wl;
  std::vector<float> qnx(nx), qny(ny);
  std::memset(wav, 0, sizeof(float) * 2 * (size_t)nx * ny);
  for (int i = 0; i < nx; i++) qnx[i] = (float)(((i + nx2) % nx) - nx2) / ax;
  for (int j = 0; j < ny; j++) qny[j] = (float)(((j + ny2) % ny) - ny2) / ay;
  float wpow = 0.f;
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      float chi = 0.f;
      // ApertureFunctionA(qnx[i]-btx, qny[j]-bty, qlcx, qlcy, qlim, alim, adir, qrs)
      float wap = 0.f;
      const float qx = qnx[i] - btx, qy = qny[j] - bty;
      if (qlim > 0.f) {
        const float a1x = alim * std::cos(2.f * adir);
        const float a1y = alim * std::sin(2 * adir);
        const float adet = 1.f - alim * alim;
        if (adet != 0.f) {
          const float radet = 1.f / adet;
          const float dqx = qx - qlcx, dqy = qy - qlcy;
          const float dqx1 = ((1.f - a1x) * dqx - a1y * dqy) * radet;
          const float dqy1 = (-a1y * dqx + (1.f + a1x) * dqy) * radet;
          const float dq2 = dqx1 * dqx1 + dqy1 * dqy1;
          const float dqm = std::sqrt(dq2);
          const float dqs = std::abs(qrs * qlim);
          if (dqm > 0.f) {
            if (dqs > 0.f) {
              const float darg = (float)kPI * (dqm - qlim) / dqs;
              wap = (1.f - std::tanh(darg)) * 0.5f;
            } else if (dqm < qlim) {
              wap = 1.f;
            }
          } else {
            wap = 1.f;
          }
        }
      }
      if (wap > (float)(1.E-6)) {  // _JPG_PROBE_APERTURE_THRESH, JProbeGen.h:52
        if (nabrr > 0) chi = aberration_chi(qnx[i], qny[j], wl, nabrr, abrr);  // axial, untilted q
        float* o = wav + 2 * ((size_t)i + (size_t)j * nx);
        o[0] = wap * std::cos(chi);
        o[1] = -wap * std::sin(chi);
        wpow += (wap * wap);
      }
    }
  }
  const float fsca = 1.f / std::sqrt(wpow);
  for (size_t k = 0; k < 2 * (size_t)nx * ny; ++k) wav[k] = wav[k] * fsca;
  return 0;
}

// CJMultiSlice::SetDetectionPlanes, overload 1 (JMultiSlice.cpp:579-627).
// det_objslc (nobjslc+1 ints, may be NULL) receives the plane -> detection
// index hash; returns the number of detection planes.
ORC_API int orc_detection_planes(int ndetper, int nobjslc, int* det_objslc) {
  int ndetslc = 0;
  const int noslc = std::max(0, nobjslc);
  if (det_objslc && nobjslc > 0)
    for (int i = 0; i <= nobjslc; i++) det_objslc[i] = -1;
  if (ndetper > 0) {
    if (noslc > 0) {
      for (int i = 0; i <= noslc; i++) {
        if (0 == i % ndetper) {
          if (det_objslc) det_objslc[i] = ndetslc;
          ndetslc++;
        }
      }
      if (0 != noslc % ndetper) {
        if (det_objslc) det_objslc[noslc] = ndetslc;
        ndetslc++;
      }
    } else {
      if (det_objslc) det_objslc[0] = 0;
      ndetslc = 1;
    }
  } else {
    if (det_objslc) det_objslc[noslc] = 0;
    ndetslc = 1;
  }
  return ndetslc;
}

// ---- RNG (rng.cpp:30-89) -------------------------------------------------
ORC_API void* orc_rng_create(int seed) {
  OrcRng* r = new OrcRng();
  r->seed(seed);
  return r;
}
ORC_API void orc_rng_destroy(void* r) { delete (OrcRng*)r; }
ORC_API void orc_rng_seed(void* r, int seed) { ((OrcRng*)r)->seed(seed); }
ORC_API int orc_rng_rand(void* r) { return ((OrcRng*)r)->rand(); }
ORC_API double orc_rng_unirand(void* r, double a, double b) { return ((OrcRng*)r)->unirand(a, b); }
ORC_API double orc_rng_normrand(void* r, double mu, double s) { return ((OrcRng*)r)->normrand(mu, s); }
// raw engine output (for pinning against std::mt19937)
ORC_API unsigned orc_rng_raw(void* r) { return ((OrcRng*)r)->e.next(); }

// CJMultiSlice::GetRandomVariantSequence (JMultiSlice.cpp:1571-1676)
ORC_API int orc_variant_sequence(void* rngp, int nscslc, const int* nvarslc, int nobjslc,
                                 const int* objslc, int* out) {
  if (!out) return 1;
  if (!nvarslc || !objslc || nscslc < 1 || nobjslc < 1) return 2;
  OrcRng* rng = (OrcRng*)rngp;
  std::vector<int> qlen(nscslc, 0), qlast(nscslc, -1);
  int qlenmax = 0;
  for (int i = 0; i < nscslc; i++) qlenmax = std::max(qlenmax, nvarslc[i]);
  std::vector<int> q((size_t)nscslc * std::max(1, qlenmax), 0);
  for (int i = 0; i < nobjslc; i++) {
    const int islc = objslc[i];
    const int nv = nvarslc[islc];
    if (nv < 2) { out[i] = 0; continue; }
    int* qcur = q.data() + (int64_t)islc * qlenmax;
    if (qlen[islc] <= 0) {
      for (int j = 0; j < nv; j++) qcur[j] = j;
      qlen[islc] = nv;
    }
    int k = (int)rng->unirand(0., (double)nv);
    k = k % nv;
    const int k0 = k;
    while (qcur[k] < 0 || qcur[k] == qlast[islc]) {
      k++;
      k = k % nv;
      if (k == k0) break;
    }
    int ivar = qcur[k];
    if (ivar < 0) {
      ivar = (int)rng->unirand(0.0, (double)nv);
      qlen[islc] = 0;
    } else {
      qcur[k] = -1;
      qlen[islc]--;
    }
    qlast[islc] = ivar;
    out[i] = ivar;
  }
  return 0;
}

// Fixed summation routine exported for pinning (JMultiSlice.cpp:80-157); a is modified.
ORC_API float orc_fdncs2m(float* a, long long n) {
  float s = 0.f;
  fdncs2m(a, (size_t)n, &s);
  return s;
}
ORC_API float orc_masked_dot(const int* mask, const float* in1, const float* in2, long long lenmask) {
  return masked_dot(mask, in1, in2, (size_t)lenmask);
}
ORC_API float orc_plain_dot(const float* in1, const float* in2, long long len) {
  return plain_dot(in1, in2, (size_t)len);
}

// CJMultiSlice::OffsetIncomingWave, CPU branch (JMultiSlice.cpp:1961-2034).
// wav0, wav_out: 2*nx*ny floats; cppl: nx*ny floats or NULL (= no constant
// phase plate); dx, dy, dz in nm; ibtx, ibty: beam tilt in Fourier pixels.
ORC_API int orc_offset_wave(int nx, int ny, float a0x, float a0y, float wl, const float* wav0,
                            const float* cppl, float dx, float dy, float dz, int ibtx, int ibty,
                            float* wav_out) {
  const int nitems = nx * ny;
  if (nitems <= 0) return 1;
  const int nx2 = (nx - (nx % 2)) >> 1, ny2 = (ny - (ny % 2)) >> 1;
  const float ca = 1.f / a0x, cb = 1.f / a0y;
  const float fz = dz * wl * 0.5f;
  std::vector<std::complex<float>> pp(nitems);
  for (int j = 0; j < ny; j++) {
    const float ky = (float)(((j + ny2) % ny) - ny2) * cb;  // m_h_fny (:1704-1711)
    const float ppy = ky * (dy + ky * fz);
    for (int i = 0; i < nx; i++) {
      const int idx = i + j * nx;
      const float kx = (float)(((i + nx2) % nx) - nx2) * ca;
      const float chic = cppl ? cppl[idx] : 0.f;
      const float chi = ((float)kTPI) * (ppy + kx * (dx + kx * fz)) + chic;
      pp[idx] = std::complex<float>(std::cos(chi), -std::sin(chi));
    }
  }
  const std::complex<float>* w0 = (const std::complex<float>*)wav0;
  std::complex<float>* wo = (std::complex<float>*)wav_out;
  const unsigned ish0 = (unsigned)imod(ibtx, nx), ish1 = (unsigned)imod(ibty, ny);
  if (ish0 > 0 || ish1 > 0) {
    for (int j = 0; j < ny; j++) {
      const int jj1 = imod(j - ibty, ny) * nx;
      for (int i = 0; i < nx; i++) {
        const int idx = i + j * nx, idx1 = imod(i - ibtx, nx) + jj1;
        wo[idx] = cmul(w0[idx1], pp[idx]);
      }
    }
  } else {
    for (int i = 0; i < nitems; i++) wo[i] = cmul(w0[i], pp[i]);
  }
  return 0;
}

// ==========================================================================
// The slice loop: CJMultiSlice::CPUMultislice (JMultiSlice.cpp:3589-3720) with
// ReadoutDifDet_h (:3369-3470), GetPhaseGrating_h (:3334-3352) and
// GetPropagator_h (:3354-3366), for the INTEGRATED and WAVEFOURIER channels.
// ==========================================================================
struct OrcSetup {
  int nx, ny;
  int nscslc;              // number of structure slices
  const int* nvarslc;      // variants per structure slice
  const long long* slcoffset;  // item offset of each structure slice in pgr
  const float* pgr;        // all phase gratings, complex interleaved
  int npro;
  const int* proidx;       // propagator index per structure slice
  const float* pro;        // propagators, complex interleaved, [npro][n]
  int nobjslc;
  const int* objslc;       // structure slice per object slice
  const int* det_objslc;   // plane -> detection index (nobjslc+1)
  int ndetslc;
  int ndet;
  const float* det;        // [ndet][n] detector functions
  const int* msk;          // [ndet][n] masks (may be NULL)
  const int* msklen;       // [ndet] mask lengths (may be NULL)
  int imagedet;            // flags: 1 integrated, 16 wave Fourier
  int dsx = 0, dsy = 0;    // diffraction de-scan [pixel] (0, 0: off)
};

// detection flags (JMultiSlice.h:141-150); the _AVG channels accumulate the
// same quantities as their plain counterparts into separate buffers, so the
// oracle computes only the plain ones
static const int kDetInt = 1, kDetImg = 2, kDetDif = 4, kDetWFR = 8, kDetWFF = 16;

// optional per-plane output channels (ndetslc planes of n items each):
// dif, img: intensities; wfr, wff: complex interleaved
template <typename T>
struct OrcChan {
  T* dif = nullptr;
  T* img = nullptr;
  T* wfr = nullptr;
};

// ReadoutImgDet_h (JMultiSlice.cpp:3473-3538): the real-space wave at a
// detection plane (the unscaled inverse transform of the Fourier-space wave)
template <typename T>
static void readout_img(const OrcSetup& S, const std::complex<T>* wav, int iplane, float weight,
                        const OrcChan<T>& ch) {
  const int n = S.nx * S.ny;
  const int idetslc = S.det_objslc[iplane];
  if (idetslc < 0) return;
  if ((S.imagedet & kDetImg) && ch.img) {
    T* o = ch.img + (size_t)n * idetslc;
    for (int i = 0; i < n; ++i) {  // GetDataPow then out += weight * img
      const T p = wav[i].real() * wav[i].real() + wav[i].imag() * wav[i].imag();
      o[i] += (T)weight * p;
    }
  }
  if ((S.imagedet & kDetWFR) && ch.wfr) {
    T* o = ch.wfr + (size_t)2 * n * idetslc;
    for (int i = 0; i < n; ++i) {
      o[2 * i] += wav[i].real() * (T)weight;
      o[2 * i + 1] += wav[i].imag() * (T)weight;
    }
  }
}

// DescanDifN / DescanDifWavN_h (JMultiSlice.cpp:2184-2320): the de-scanned
// copy of a Fourier-space array, out[(i - dnx) mod nx + ((j - dny) mod ny) nx]
// = in[i + j nx]
template <typename V>
static void descan_copy(const V* in, V* out, int nx, int ny, int dnx, int dny) {
  for (int j = 0; j < ny; j++) {
    const int j1 = ((j - dny) % ny + ny) % ny;
    for (int i = 0; i < nx; i++) out[((i - dnx) % nx + nx) % nx + (size_t)j1 * nx] = in[i + (size_t)j * nx];
  }
}

template <typename T>
static void readout_dif(const OrcSetup& S, const std::complex<T>* wav_in, int iplane, float weight,
                        float* out_int, T* out_wff, std::vector<float>& difbuf,
                        std::vector<T>& difbuf_d, const OrcChan<T>& ch = OrcChan<T>()) {
  const int n = S.nx * S.ny;
  const int idetslc = S.det_objslc[iplane];
  if (idetslc < 0) return;
  // with de-scan, every Fourier-space readout sees the shifted wave
  std::vector<std::complex<T>> shifted;
  const std::complex<T>* wav = wav_in;
  if (S.dsx != 0 || S.dsy != 0) {
    shifted.resize(n);
    descan_copy(wav_in, shifted.data(), S.nx, S.ny, S.dsx, S.dsy);
    wav = shifted.data();
  }
  if ((S.imagedet & kDetDif) && ch.dif) {  // ReadoutDifDet_h (:3369-3470): out += weight * |psi|^2
    T* o = ch.dif + (size_t)n * idetslc;
    for (int i = 0; i < n; ++i) {
      const T p = wav[i].real() * wav[i].real() + wav[i].imag() * wav[i].imag();
      o[i] += (T)weight * p;
    }
  }
  if (S.ndet > 0 && (S.imagedet & kDetInt)) {
    if (sizeof(T) == 4) {
      for (int i = 0; i < n; ++i) {  // JFFTMKLcore::GetDataPow (:690-706)
        const float re = (float)wav[i].real(), im = (float)wav[i].imag();
        difbuf[i] = re * re + im * im;
      }
    } else {
      for (int i = 0; i < n; ++i) difbuf_d[i] = wav[i].real() * wav[i].real() + wav[i].imag() * wav[i].imag();
    }
    for (int idet = 0; idet < S.ndet; ++idet) {
      const float* det = S.det + (size_t)idet * n;
      const int* msk = S.msk ? S.msk + (size_t)idet * n : nullptr;
      const int ml = S.msklen ? S.msklen[idet] : 0;
      float* out = out_int + (size_t)idetslc * S.ndet + idet;
      if (sizeof(T) == 4) {
        if (msk && ml > 0) *out += weight * masked_dot(msk, difbuf.data(), det, (size_t)ml);
        else *out += weight * plain_dot(difbuf.data(), det, (size_t)n);
      } else {
        double v = (msk && ml > 0) ? masked_dot_d(msk, (const double*)difbuf_d.data(), det, (size_t)ml)
                                   : plain_dot_d((const double*)difbuf_d.data(), det, (size_t)n);
        *out += (float)((double)weight * v);
      }
    }
  }
  if ((S.imagedet & kDetWFF) && out_wff) {
    T* o = out_wff + (size_t)2 * n * idetslc;
    for (int i = 0; i < n; ++i) {
      o[2 * i] += wav[i].real() * (T)weight;
      o[2 * i + 1] += wav[i].imag() * (T)weight;
    }
  }
}

template <typename T>
static int multislice_run(const OrcSetup& S, FFT2D<T>& fft, std::complex<T>* wav, int islc0,
                          const int* var, float weight, float* out_int, T* out_wff,
                          std::vector<float>& difbuf, std::vector<T>& difbuf_d,
                          const OrcChan<T>& ch = OrcChan<T>()) {
  const int n = S.nx * S.ny;
  const bool bimg = (S.imagedet & (kDetImg | kDetWFR)) != 0;
  int islc = islc0;
  if (islc < 0) islc = 0;
  if (islc >= S.nobjslc) islc = S.nobjslc;
  if (islc < S.nobjslc) {
    for (int jslc = islc; jslc < S.nobjslc; ++jslc) {
      const bool bdetect = S.det_objslc[jslc] >= 0;
      if (bdetect) readout_dif<T>(S, wav, jslc, weight, out_int, out_wff, difbuf, difbuf_d, ch);
      fft.run(wav, +1);  // IFT
      if (bdetect && bimg) readout_img<T>(S, wav, jslc, weight, ch);
      const int js = S.objslc[jslc] % S.nscslc;
      const std::complex<float>* pg =
          (const std::complex<float>*)(S.pgr) + S.slcoffset[js] + (int64_t)var[jslc] * n;
      for (int i = 0; i < n; ++i)
        wav[i] = cmul(wav[i], std::complex<T>((T)pg[i].real(), (T)pg[i].imag()));
      fft.run(wav, -1);  // FT
      const std::complex<float>* pr =
          (const std::complex<float>*)(S.pro) + (size_t)S.proidx[js] * n;
      for (int i = 0; i < n; ++i)
        wav[i] = cmul(wav[i], std::complex<T>((T)pr[i].real(), (T)pr[i].imag()));
    }
  }
  readout_dif<T>(S, wav, S.nobjslc, weight, out_int, out_wff, difbuf, difbuf_d, ch);
  if (bimg) {  // exit-plane real-space readout (JMultiSlice.cpp:3700-3712)
    fft.run(wav, +1);
    readout_img<T>(S, wav, S.nobjslc, weight, ch);
  }
  return 0;
}

static OrcSetup make_setup(int nx, int ny, int nscslc, const int* nvarslc, const long long* slcoffset,
                           const float* pgr, int npro, const int* proidx, const float* pro,
                           int nobjslc, const int* objslc, const int* det_objslc, int ndetslc,
                           int ndet, const float* det, const int* msk, const int* msklen,
                           int imagedet) {
  OrcSetup S;
  S.nx = nx; S.ny = ny; S.nscslc = nscslc; S.nvarslc = nvarslc; S.slcoffset = slcoffset;
  S.pgr = pgr; S.npro = npro; S.proidx = proidx; S.pro = pro; S.nobjslc = nobjslc;
  S.objslc = objslc; S.det_objslc = det_objslc; S.ndetslc = ndetslc; S.ndet = ndet;
  S.det = det; S.msk = msk; S.msklen = msklen; S.imagedet = imagedet;
  return S;
}

// One CPUMultislice call: wave_in is the (already offset) incident wave in
// Fourier space; var the variant sequence (nobjslc ints).  out_int
// (ndetslc*ndet floats) and out_wff (2*n*ndetslc floats, may be NULL) are
// accumulated with `weight` (accmode INTEGRATE); callers zero them for
// accmode NONE.  prec 32: float arithmetic; prec 64: the FP64 twin (wave and
// exit waves in double; out_wff then holds doubles).
ORC_API int orc_multislice(int prec, int nx, int ny, int nscslc, const int* nvarslc,
                           const long long* slcoffset, const float* pgr, int npro,
                           const int* proidx, const float* pro, int nobjslc, const int* objslc,
                           const int* det_objslc, int ndetslc, int ndet, const float* det,
                           const int* msk, const int* msklen, int imagedet, const float* wave_in,
                           const int* var, int islc0, float weight, float* out_int, void* out_wff) {
  if (nx < 1 || ny < 1 || !wave_in || !det_objslc) return 1;
  OrcSetup S = make_setup(nx, ny, nscslc, nvarslc, slcoffset, pgr, npro, proidx, pro, nobjslc,
                          objslc, det_objslc, ndetslc, ndet, det, msk, msklen, imagedet);
  const int n = nx * ny;
  std::vector<float> difbuf(n);
  if (prec == 64) {
    FFT2D<double> f; f.init(nx, ny);
    std::vector<std::complex<double>> w(n);
    for (int i = 0; i < n; ++i) w[i] = std::complex<double>(wave_in[2 * i], wave_in[2 * i + 1]);
    std::vector<double> dd(n);
    return multislice_run<double>(S, f, w.data(), islc0, var, weight, out_int, (double*)out_wff, difbuf, dd);
  }
  FFT2D<float> f; f.init(nx, ny);
  std::vector<std::complex<float>> w(n);
  std::memcpy((void*)w.data(), wave_in, sizeof(float) * 2 * n);
  std::vector<float> dd(1);
  return multislice_run<float>(S, f, w.data(), islc0, var, weight, out_int, (float*)out_wff, difbuf, dd);
}

// orc_multislice with the intensity and wave channels: out_dif and out_img
// (n*ndetslc items), out_wfr and out_wff (2*n*ndetslc items), each NULL or
// accumulated with `weight`; float for prec 32, double for prec 64.  dsx, dsy:
// diffraction de-scan [pixel] applied to the Fourier-space readouts.
ORC_API int orc_multislice_ch(int prec, int nx, int ny, int nscslc, const int* nvarslc,
                              const long long* slcoffset, const float* pgr, int npro,
                              const int* proidx, const float* pro, int nobjslc, const int* objslc,
                              const int* det_objslc, int ndetslc, int ndet, const float* det,
                              const int* msk, const int* msklen, int imagedet, const float* wave_in,
                              const int* var, int islc0, float weight, float* out_int, void* out_dif,
                              void* out_img, void* out_wfr, void* out_wff, int dsx, int dsy) {
  if (nx < 1 || ny < 1 || !wave_in || !det_objslc) return 1;
  OrcSetup S = make_setup(nx, ny, nscslc, nvarslc, slcoffset, pgr, npro, proidx, pro, nobjslc,
                          objslc, det_objslc, ndetslc, ndet, det, msk, msklen, imagedet);
  S.dsx = dsx;
  S.dsy = dsy;
  const int n = nx * ny;
  std::vector<float> difbuf(n);
  if (prec == 64) {
    FFT2D<double> f; f.init(nx, ny);
    std::vector<std::complex<double>> w(n);
    for (int i = 0; i < n; ++i) w[i] = std::complex<double>(wave_in[2 * i], wave_in[2 * i + 1]);
    std::vector<double> dd(n);
    OrcChan<double> ch;
    ch.dif = (double*)out_dif; ch.img = (double*)out_img; ch.wfr = (double*)out_wfr;
    return multislice_run<double>(S, f, w.data(), islc0, var, weight, out_int, (double*)out_wff, difbuf, dd, ch);
  }
  FFT2D<float> f; f.init(nx, ny);
  std::vector<std::complex<float>> w(n);
  std::memcpy((void*)w.data(), wave_in, sizeof(float) * 2 * n);
  std::vector<float> dd(1);
  OrcChan<float> ch;
  ch.dif = (float*)out_dif; ch.img = (float*)out_img; ch.wfr = (float*)out_wfr;
  return multislice_run<float>(S, f, w.data(), islc0, var, weight, out_int, (float*)out_wff, difbuf, dd, ch);
}

// A STEM scan the way the reference bench drives the library
// (test/JMSBench1/multislice.cpp:469-688 run_multislice, threaded as in
// multithread_stem :1220-1417): for every position p and pass r,
// OffsetIncomingWave(dx[p], dy[p], dz[p]) then CPUMultislice(0, acc, weight[r]),
// accmode NONE on the first pass and INTEGRATE after.  The variant sequences
// (npos*npass*nobjslc ints, position-major, pass-minor) are drawn by the caller
// from one seeded stream in that serial order.  Work is spread over `nthreads`
// std::threads, each with its own FFT plan and buffers (JMultiSlice.cpp:1723-1753).
// out_int: npos*ndetslc*ndet floats.  out_wff: npos*2*n*ndetslc (or NULL).
ORC_API int orc_stem_scan(int prec, int nthreads, int nx, int ny, float a0x, float a0y, float wl,
                          int nscslc, const int* nvarslc, const long long* slcoffset,
                          const float* pgr, int npro, const int* proidx, const float* pro,
                          int nobjslc, const int* objslc, const int* det_objslc, int ndetslc,
                          int ndet, const float* det, const int* msk, const int* msklen,
                          int imagedet, const float* wav0, const float* cppl, int npos,
                          const float* dx, const float* dy, const float* dz, int npass,
                          const float* weights, const int* var_all, float* out_int,
                          void* out_wff) {
  if (npos < 1 || npass < 1 || nthreads < 1) return 1;
  OrcSetup S = make_setup(nx, ny, nscslc, nvarslc, slcoffset, pgr, npro, proidx, pro, nobjslc,
                          objslc, det_objslc, ndetslc, ndet, det, msk, msklen, imagedet);
  const int n = nx * ny;
  const size_t nint = (size_t)ndetslc * ndet;
  std::vector<std::thread> th;
  std::vector<int> errs(nthreads, 0);
  auto worker = [&](int tid) {
    std::vector<float> wf(2 * (size_t)n), difbuf(n);
    if (prec == 64) {
      FFT2D<double> f; f.init(nx, ny);
      std::vector<std::complex<double>> w(n);
      std::vector<double> dd(n);
      for (int p = tid; p < npos; p += nthreads) {
        float* oi = out_int + nint * p;
        double* ow = out_wff ? (double*)out_wff + (size_t)2 * n * ndetslc * p : nullptr;
        std::fill(oi, oi + nint, 0.f);
        if (ow) std::fill(ow, ow + (size_t)2 * n * ndetslc, 0.);
        for (int r = 0; r < npass; ++r) {
          orc_offset_wave(nx, ny, a0x, a0y, wl, wav0, cppl, dx[p], dy[p], dz ? dz[p] : 0.f, 0, 0, wf.data());
          for (int i = 0; i < n; ++i) w[i] = std::complex<double>(wf[2 * i], wf[2 * i + 1]);
          multislice_run<double>(S, f, w.data(), 0, var_all + ((size_t)p * npass + r) * nobjslc,
                                 weights[r], oi, ow, difbuf, dd);
        }
      }
    } else {
      FFT2D<float> f; f.init(nx, ny);
      std::vector<float> dd(1);
      for (int p = tid; p < npos; p += nthreads) {
        float* oi = out_int + nint * p;
        float* ow = out_wff ? (float*)out_wff + (size_t)2 * n * ndetslc * p : nullptr;
        std::fill(oi, oi + nint, 0.f);
        if (ow) std::fill(ow, ow + (size_t)2 * n * ndetslc, 0.f);
        for (int r = 0; r < npass; ++r) {
          orc_offset_wave(nx, ny, a0x, a0y, wl, wav0, cppl, dx[p], dy[p], dz ? dz[p] : 0.f, 0, 0, wf.data());
          multislice_run<float>(S, f, (std::complex<float>*)wf.data(), 0,
                                var_all + ((size_t)p * npass + r) * nobjslc, weights[r], oi, ow,
                                difbuf, dd);
        }
      }
    }
  };
  for (int t = 0; t < nthreads; ++t) th.emplace_back(worker, t);
  for (auto& t : th) t.join();
  return 0;
}

// ==========================================================================
// Synthetic specimen recipe (the benchmark configurations' own recipe, not a
// reference function): cubic perovskite ABO3 cells, Gaussian projected
// potential per atom, frozen-phonon site displacements drawn the way the
// reference's QEP model draws them (JEElaSca.cpp:893-905: per site, x then y,
// ux * NormRand() each), phase grating t = exp(i * sigma * V).
//
// Slice type 0 (z = 0): A at (0,0), O at (1/2,1/2); slice type 1 (z = a/2):
// B at (1/2,1/2), O at (1/2,0), O at (0,1/2) -- fractional cell coordinates.
// Sites are visited cell row by cell row (cy outer, cx inner), basis order
// within a cell.  The displacement stream of (type, variant) is seeded with
// seed_base + 8*type + variant; variant 0 of an nvar = 1 setup is undisplaced.
// V(r) = S/(2 pi s^2) exp(-|r-ra|^2/(2 s^2)), S = 2.0 * Z^0.7 [V A^2], summed
// over all periodic images within rcut = 8 s; s, positions and sampling in A.
// ==========================================================================
ORC_API int orc_perovskite_sites(int ncx, int ncy, double a_ang, int slice_type, int zA, int zB,
                                 int zO, double rms_ang, int seed, int displace, double* xy,
                                 int* zlist) {
  static const double fr0[2][2] = {{0.0, 0.0}, {0.5, 0.5}};
  static const double fr1[3][2] = {{0.5, 0.5}, {0.5, 0.0}, {0.0, 0.5}};
  OrcRng rng;
  rng.seed(seed);
  int k = 0;
  for (int cy = 0; cy < ncy; ++cy)
    for (int cx = 0; cx < ncx; ++cx) {
      const int nb = slice_type == 0 ? 2 : 3;
      for (int b = 0; b < nb; ++b) {
        const double fx = slice_type == 0 ? fr0[b][0] : fr1[b][0];
        const double fy = slice_type == 0 ? fr0[b][1] : fr1[b][1];
        double x = ((double)cx + fx) * (double)a_ang;
        double y = ((double)cy + fy) * (double)a_ang;
        if (displace) {
          x += rms_ang * rng.normrand(0.0, 1.0);
          y += rms_ang * rng.normrand(0.0, 1.0);
        }
        xy[2 * k] = x;
        xy[2 * k + 1] = y;
        zlist[k] = slice_type == 0 ? (b == 0 ? zA : zO) : (b == 0 ? zB : zO);
        ++k;
      }
    }
  return k;
}

// Phase grating of a list of atoms (positions in A) on an nx x ny grid over a
// cell of ax x ay A.  out: 2*nx*ny floats.
ORC_API int orc_gaussian_grating(int nx, int ny, double ax, double ay, int nat, const double* xy,
                                 const int* zlist, double sigma_ang, double sig_int,
                                 float* out) {
  const size_t n = (size_t)nx * ny;
  std::vector<double> v(n, 0.0);
  const double dxp = ax / nx, dyp = ay / ny;
  const double rcut = 8.0 * sigma_ang, rcut2 = rcut * rcut;
  const double i2s2 = 1.0 / (2.0 * sigma_ang * sigma_ang);
  for (int a = 0; a < nat; ++a) {
    const double S = 2.0 * std::pow((double)zlist[a], 0.7);
    const double amp = S / (2.0 * kPI * sigma_ang * sigma_ang);
    // wrap the atom into the cell
    double xa = std::fmod(xy[2 * a], ax); if (xa < 0) xa += ax;
    double ya = std::fmod(xy[2 * a + 1], ay); if (ya < 0) ya += ay;
    const int i0 = (int)std::floor((xa - rcut) / dxp), i1 = (int)std::ceil((xa + rcut) / dxp);
    const int j0 = (int)std::floor((ya - rcut) / dyp), j1 = (int)std::ceil((ya + rcut) / dyp);
    for (int j = j0; j <= j1; ++j) {
      const double ry = j * dyp - ya;
      if (ry * ry > rcut2) continue;
      const int jw = imod(j, ny);
      for (int i = i0; i <= i1; ++i) {
        const double rx = i * dxp - xa;
        const double r2 = rx * rx + ry * ry;
        if (r2 > rcut2) continue;
        v[(size_t)imod(i, nx) + (size_t)jw * nx] += amp * std::exp(-r2 * i2s2);
      }
    }
  }
  for (size_t i = 0; i < n; ++i) {
    const double ph = sig_int * v[i];
    out[2 * i] = (float)std::cos(ph);
    out[2 * i + 1] = (float)std::sin(ph);
  }
  return 0;
}
