// jms_elasca.cpp -- electron scattering factors and structure slicing on the
// host (include/jms_elasca.h): the Weickenmeier-Kohl (WK) parameterisation
// of the elastic atomic form factor with relativistic correction,
// Debye-Waller damping and absorptive parts, radial tabulation, projection
// to the 2-D form factors of a grid, and z-slicing of a structure.  The
// formulae follow the reference's CWEKOScat (src/wekoscat.cpp), CAtomType
// (src/Atoms.cpp), CJEElaSca (src/JEElaSca.cpp) and CStructure
// (src/Structure.cpp) step for step, in the same precision, so that the
// tables equal the reference's; the parameter tables are the published data
// of Weickenmeier & Kohl, Acta Cryst. A 47 (1991) 590-597, as the reference
// lists them.
#include "jms_elasca.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <vector>

namespace {

const double kPi = 3.1415926535897932384626433832795;
const double kTwoPi = 6.283185307179586476925286766559;
const double kFourPi = 12.566370614359172953850573533118;
// NatureConstants.h:47-67 (exact SI values and the derived form-factor
// amplitude m0 e^2 / (2 h^2) / (4 pi eps0) * 1e-10 [1/A])
const double kC0 = 299792458.;
const double kEps0 = 8.8541878176203898505365630317108E-12;
const double kHpl = 6.62607015E-34;
const double kQel = 1.602176634E-19;
const double kMe0 = 9.1093837015E-31;
const double kE0kV = kMe0 * kC0 * kC0 / kQel / 1000.;
const double kCffa = kMe0 * kQel * kQel / (2.0 * kHpl * kHpl) / (kFourPi * kEps0) * 1.0E-10;
const double kEps = 1.0E-8;                            // wekoscat.h:59
const double kR8Pi2A = 100. / (8. * kPi * kPi);         // wekoscat.h:60
const double kK0pA = kTwoPi * kQel / (kHpl * kC0) * 1.0E-7;  // wekoscat.h:61

const int kNprm = 6;
const int kMaxZ = 98;

// --- Weickenmeier & Kohl (1991) parameters, Z = 0..98 ----------------------
// V per element (the a1 / a0 ratio)
const double kWkV[kMaxZ + 1] = {
    0.0, 0.5, 0.5, 0.5, 0.3, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5,
    0.5, 0.5, 0.4, 0.5, 0.5, 0.5, 0.5, 0.5, 0.2, 0.3, 0.5,
    0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5,
    0.5, 0.5, 0.5, 0.5, 0.2, 0.3, 0.5, 0.5, 0.5, 0.5, 0.5,
    0.4, 0.5, 0.5, 0.5, 0.3, 0.4, 0.6, 0.6, 0.6, 0.4, 0.4,
    0.1, 0.1, 0.3, 0.3, 0.2, 0.2, 0.2, 0.2, 0.1, 0.2, 0.1,
    0.2, 0.1, 0.2, 0.1, 0.1, 0.1, 0.1, 0.4, 0.2, 0.5, 0.4,
    0.5, 0.5, 0.4, 0.4, 0.4, 0.3, 0.4, 0.4, 0.4, 0.4, 0.1,
    0.2, 0.2, 0.3, 0.2, 0.2, 0.2, 0.2, 0.2, 0.3, 0.2, 0.3,
};
// B1..B6 [A^2] per element
const double kWkB[kMaxZ + 1][kNprm] = {
#include "jms_wk_table.inc"
};
// x exp(-x) Ei(x) at 1/x = 0.005 i, i = 0..20, to six decimals
const double kRih2[21] = {
    1.000000, 1.005051, 1.010206, 1.015472, 1.020852, 1.026355, 1.031985, 1.037751, 1.043662, 1.049726, 1.055956,
    1.062364, 1.068965, 1.075780, 1.082830, 1.090140, 1.097737, 1.105647, 1.113894, 1.122497, 1.131470};
// rational approximation of x exp(x) E1(x), x >= 1 (Abramowitz & Stegun
// 5.1.56), coefficients to five decimals as the reference carries them
const double kEiA[4] = {8.57332, 18.05901, 8.63476, 0.26777};
const double kEiB[4] = {9.57332, 25.63295, 21.09965, 3.95849};

int weko_params(int z, double* a, double* b) {
  if (z < 0 || z > kMaxZ) return 1;
  if (!a || !b) return 2;
  a[0] = kCffa * (double)z / (3.0 * (1.0 + kWkV[z]));
  a[1] = kWkV[z] * a[0];
  for (int i = 0; i < kNprm; ++i) b[i] = kWkB[z][i];
  return 0;
}

// f(s) = sum_i a_{i/3} (1 - exp(-b_i s^2)) / s^2 with the small- and
// large-argument branches of the reference
double weko_f(const double* a, const double* b, double s) {
  const double s2 = s > 0.0 ? 1. / (s * s) : 0.0;
  double ff = 0.0;
  for (int i = 0; i < kNprm; ++i) {
    const double aj = a[i / 3];
    const double argu = b[i] * s * s;
    if (argu < 0.1) ff += aj * b[i] * (1.0 - 0.5 * argu);
    else if (argu > 20.0) ff += aj * s2;
    else ff += aj * (1.0 - std::exp(-argu)) * s2;
  }
  return ff;
}

double ei(double x) {
  if (x > 60. || x < -60.) return 0.0;
  if (x < -1.) {
    const double t = std::fabs(x);
    const double num = kEiA[3] + t * (kEiA[2] + t * (kEiA[1] + t * (kEiA[0] + t)));
    const double den = kEiB[3] + t * (kEiB[2] + t * (kEiB[1] + t * (kEiB[0] + t)));
    return -num / den * std::exp(-t) / t;
  }
  // Ei(x) = gamma + ln|x| + sum_k x^k / (k k!), terms until |term / x| <= 1e-6
  double term = x, sum = x;
  int k = 0;
  while (std::fabs(term / x) > 1.e-6) {
    ++k;
    const double kk = (double)k, k1 = kk + 1.;
    term *= (x * kk / (k1 * k1));
    sum += term;
  }
  return .577216 + std::log(std::fabs(x)) + sum;
}

double rih2(double x) {
  const double r = 1. / x;
  const int i = (int)(200. * r);
  return kRih2[i] + 200. * (kRih2[i + 1] - kRih2[i]) * (r - 0.005 * (double)i);
}

// exp(-x1) (Ei(x2) - Ei(x3)), large arguments through rih2
double rih1(double x1, double x2, double x3) {
  if (x2 <= 20. && x3 <= 20.) return std::exp(-x1) * (ei(x2) - ei(x3));
  double r = (x2 > 20.) ? std::exp(x2 - x1) * rih2(x2) / x2 : std::exp(-x1) * ei(x2);
  r -= (x3 > 20.) ? std::exp(x3 - x1) * rih2(x3) / x3 : std::exp(-x1) * ei(x3);
  return r;
}

double ri1(double bi, double bj, double g) {
  if (g == 0.0) return kPi * (bi * std::log((bi + bj) / bi) + bj * std::log((bi + bj) / bj));
  const double c = 0.5772157;
  const double g2 = g * g, big2 = bi * g2, bjg2 = bj * g2;
  double r = 2. * c + std::log(big2) + std::log(bjg2) - 2. * ei(-bi * bj * g2 / (bi + bj));
  r += rih1(big2, big2 * bi / (bi + bj), big2);
  r += rih1(bjg2, bjg2 * bj / (bi + bj), bjg2);
  return r * (kPi / g2);
}

double ri2(double bi, double bj, double g, double u) {
  const double u2 = u * u;
  if (g == 0.) {
    double r = (bi + u2) * std::log((bi + bj + u2) / (bi + u2));
    r += bj * std::log((bi + bj + u2) / (bj + u2));
    r += u2 * std::log(u2 / (bj + u2));
    return r * kPi;
  }
  const double g2 = g * g;
  const double biuh = bi + 0.5 * u2, bjuh = bj + 0.5 * u2;
  const double biu = bi + u2, bju = bj + u2;
  double r = ei(-0.5 * u2 * g2 * biuh / biu) + ei(-0.5 * u2 * g2 * bjuh / bju);
  r -= (ei(-biuh * bjuh * g2 / (biuh + bjuh)) + ei(-.25 * u2 * g2));
  r *= 2.0;
  r += rih1(0.5 * u2 * g2, 0.25 * u2 * g2, 0.25 * u2 * u2 * g2 / biu);
  r += rih1(0.5 * u2 * g2, 0.25 * u2 * g2, 0.25 * u2 * u2 * g2 / bju);
  r += rih1(biuh * g2, biuh * biuh * g2 / (biuh + bjuh), biuh * biuh * g2 / biu);
  r += rih1(bjuh * g2, bjuh * bjuh * g2 / (biuh + bjuh), bjuh * bjuh * g2 / bju);
  return r * (kPi / g2);
}

// sum_ij a'_i a'_j [DW(g) ri1 - ri2], a' = 16 pi^2 a, b' = b / (16 pi^2).
// The reference takes DW(g) = exp(-ul g^2 / 2) with the rms displacement ul
// itself (wekoscat.cpp:366) where the thermal-diffuse absorption formula has
// exp(-ul^2 g^2 / 2), the Debye-Waller factor that ri2 carries; that makes
// the absorptive form factor negative beyond small g.  dw_sq = false keeps
// the reference's term, dw_sq = true uses ul^2.
double weko_imag(double g, double ul, const double* a, const double* b, bool dw_sq = false) {
  const double fp2 = kFourPi * kFourPi;
  double a1[2], b1[kNprm];
  for (int i = 0; i < 2; ++i) a1[i] = a[i] * fp2;
  for (int i = 0; i < kNprm; ++i) b1[i] = b[i] / fp2;
  const double g2 = g * g;
  const double dewa = std::exp(-0.5 * (dw_sq ? ul * ul : ul) * g2);
  double fi = 0.0;
  for (int j = 0; j < kNprm; ++j)
    for (int i = 0; i < kNprm; ++i)
      fi += (a1[j / 3] * a1[i / 3] * (dewa * ri1(b1[i], b1[j], g) - ri2(b1[i], b1[j], g, ul)));
  return fi;
}

// Ewald-sphere integrand of the absorptive part due to a band-width limiting
// aperture sgmax (wekomugap, wekoscat.cpp:421-456)
struct ApertureIntegrand {
  const double* a;
  const double* b;
  double g, ap, k0;
  double operator()(double theta, double phi) const {
    const double ct = std::cos(theta), st = std::sin(theta), cp = std::cos(phi);
    const double k = k0, twok = k + k;
    const double q = k * std::sqrt(2. - 2. * ct);
    const double qg = std::sqrt(g * g + twok * k - twok * (k * ct + g * cp * st));
    const double sgmax = ap * 0.5;
    const double sg = 0.5 * g, sq = 0.5 * q, sqg = 0.5 * qg;
    const double fq = weko_f(a, b, sq), fqg = weko_f(a, b, sqg);
    const double ag = (sg >= sgmax) ? 0. : 1.;
    const double aq = (sq >= sgmax) ? 0. : 1.;
    const double aqg = (sqg >= sgmax) ? 0. : 1.;
    return fq * fqg * (ag - aq * aqg) * st;
  }
};

// trapezoidal sum on a power-law grid x = lo + (hi - lo) (i / (n - 1))^p
// (dsgrid2d, integration.h:40-100), dimension 1 inner
template <class F>
double grid_trapezoid_2d(const F& f, double a1, double b1, double a2, double b2, double p1, double p2, unsigned n1,
                         unsigned n2) {
  std::vector<double> x1(n1), x2(n2), s1(n2, 0.0), y((size_t)n1 * n2);
  const double c1 = 1.0 / (double)(n1 - 1), c2 = 1.0 / (double)(n2 - 1);
  for (unsigned i = 0; i < n1; ++i) x1[i] = a1 + (b1 - a1) * std::pow(c1 * (double)i, p1);
  for (unsigned j = 0; j < n2; ++j) {
    x2[j] = a2 + (b2 - a2) * std::pow(c2 * (double)j, p2);
    for (unsigned i = 0; i < n1; ++i) y[i + (size_t)j * n1] = f(x1[i], x2[j]);
  }
  for (unsigned j = 0; j < n2; ++j) {
    const double* row = &y[(size_t)j * n1];
    for (unsigned i = 0; i + 1 < n1; ++i) s1[j] += (row[i + 1] + row[i]) * (x1[i + 1] - x1[i]);
    s1[j] *= 0.5;
  }
  double c = 0.;
  for (unsigned j = 0; j + 1 < n2; ++j) c += (s1[j + 1] + s1[j]) * (x2[j + 1] - x2[j]);
  return 0.5 * c;
}

double weko_abs_aperture(double g, double ap, const double* a, const double* b, double k0) {
  if (!(g < ap)) return 0.;
  ApertureIntegrand f{a, b, g, ap, k0};
  // theta 0..pi on 128 quadratically spaced samples, phi 0..pi (half side,
  // symmetric: times 2) on 64 linear samples
  return grid_trapezoid_2d(f, 0., kPi, 0., kPi, 2.0, 1.0, 128, 64) * 2.0;
}

// absflg: 0 elastic only, 1 absorptive part as the reference computes it,
// 2 the same with the squared displacement in the thermal Debye-Waller term
void weko_form_factor(double g, double dw, int z, double ekv, bool dwflg, int absflg, double* out) {
  out[0] = out[1] = 0.0;
  double a[2], b[kNprm];
  if (weko_params(z, a, b) != 0) return;
  const double rc = (kE0kV + ekv) / kE0kV;
  const double dwf = dwflg ? std::exp(-0.25 * dw * g * g) : 1.0;
  const double sa = 0.050 * g;  // s = g / 2 [1/A]
  const double ff = 0.1 * weko_f(a, b, sa) * rc * kFourPi * dwf;  // [nm], with 4 pi as the EMS convention
  out[0] = ff;
  if (absflg && dwflg && dw > kEps && ff > kEps) {  // thermal diffuse scattering
    const double k0 = kK0pA * std::sqrt((2. * kE0kV + ekv) * ekv);  // 2 pi k [1/A]
    const double ga = 0.1 * g * kTwoPi;
    const double ua = std::sqrt(dw * kR8Pi2A);
    out[1] = 0.1 * rc * rc * weko_imag(ga, ua, a, b, absflg == 2) / k0;
  }
  if (absflg && !dwflg && g < dw) {  // band-width limiting aperture dw [1/nm]
    const double k0 = kK0pA * std::sqrt((2. * kE0kV + ekv) * ekv);
    out[1] = 0.1 * rc * rc * weko_abs_aperture(0.1 * g, 0.1 * dw, a, b, k0 / kTwoPi) * k0;
  }
}

}  // namespace

extern "C" {

int jms_weko_params(int z, double* a, double* b) { return weko_params(z, a, b); }
double jms_weko_f(const double* a, const double* b, double s) { return (a && b) ? weko_f(a, b, s) : 0.0; }
double jms_weko_ei(double x) { return ei(x); }
double jms_weko_rih2(double x) { return (x > 10.) ? rih2(x) : 0.0; }
double jms_weko_ri1(double bi, double bj, double g) { return ri1(bi, bj, g); }
double jms_weko_ri2(double bi, double bj, double g, double u) { return ri2(bi, bj, g, u); }
double jms_weko_imag(double g, double ul, const double* a, const double* b) {
  return (a && b) ? weko_imag(g, ul, a, b) : 0.0;
}

int jms_weko_form_factor(double g, double dw, int z, double ekv, int dwflg, int absflg, double* out) {
  if (!out) return 2;
  if (z < 0 || z > kMaxZ) { out[0] = out[1] = 0.0; return 1; }
  weko_form_factor(g, dw, z, ekv, dwflg != 0, absflg < 0 ? 0 : std::min(absflg, 2), out);
  return 0;
}

int jms_radial_form_factors(int z, float biso, float ekv, int dwf, int absorb, float abf, long long nk,
                            float sk, float bwlk, float* out) {
  if (!out || nk < 1) return 2;
  if (z < 0 || z > kMaxZ) return 1;
  const double sq = (double)sk, ht = (double)ekv;
  const double bpar = dwf ? (double)biso : (double)bwlk;
  for (long long ik = 0; ik < nk; ++ik) {
    const double q = sq * (double)ik;
    double v[2] = {0.0, 0.0};
    if (absorb == 2 || absorb == 3) {
      weko_form_factor(q, bpar, z, ht, dwf != 0, absorb == 3 ? 2 : 1, v);
    } else {
      weko_form_factor(q, bpar, z, ht, dwf != 0, 0, v);
      v[0] = (double)(float)v[0];  // the elastic value is stored through float first
      v[1] = (absorb == 1) ? v[0] * (double)abf : 0.0;
    }
    out[2 * ik] = (float)v[0];
    out[2 * ik + 1] = (float)v[1];
  }
  return 0;
}

int jms_form_factor_sampling(int nx, int ny, int nz, float ax, float ay, float az, long long* nk, float* sk,
                             float* bwlk) {
  if (nx < 1 || ny < 1 || !(ax > 0.f) || !(ay > 0.f)) return 1;
  if (!nk || !sk || !bwlk) return 2;
  const int nzz = std::max(0, nz);
  const long long n = 12LL * std::max(std::max(nx, ny), nzz);
  const float zt = (nzz > 0 && az > 0.f) ? (float)nzz / az : 0.f;
  const float s = std::max(std::max((float)nx / ax, (float)ny / ay), zt) / (float)n;
  const unsigned nyqx = (unsigned)(nx - (nx % 2)) >> 1, nyqy = (unsigned)(ny - (ny % 2)) >> 1;
  const float kmax = std::min((float)nyqx / ax, (float)nyqy / ay);
  *nk = n;
  *sk = s;
  *bwlk = kmax * 0.6666667f;
  return 0;
}

int jms_form_factors_2d(int nx, int ny, float ax, float ay, int bwl_pot, int ntype, long long nk, float sk,
                        const float* radial, float* ff2d) {
  if (nx < 1 || ny < 1 || !(ax > 0.f) || !(ay > 0.f) || ntype < 1 || nk < 2 || !(sk > 0.f)) return 1;
  if (!radial || !ff2d) return 2;
  typedef std::complex<float> cf;
  const float itogx = 1.f / ax, itogy = 1.f / ay;
  const int nx2 = (nx - nx % 2) >> 1, ny2 = (ny - ny % 2) >> 1;
  std::vector<float> qx(nx), qy(ny);
  for (int i = 0; i < nx; ++i) qx[i] = (float)(((i + nx2) % nx) - nx2) * itogx;
  for (int j = 0; j < ny; ++j) qy[j] = (float)(((j + ny2) % ny) - ny2) * itogy;
  const float kmax = std::min((float)nx2 * itogx, (float)ny2 * itogy);
  const size_t npix = (size_t)nx * ny;
  for (int t = 0; t < ntype; ++t) {
    const cf* tab = reinterpret_cast<const cf*>(radial) + (size_t)t * nk;
    cf* dst = reinterpret_cast<cf*>(ff2d) + (size_t)t * npix;
    for (int j = 0; j < ny; ++j) {
      const float gy2 = qy[j] * qy[j];
      for (int i = 0; i < nx; ++i) {
        const float g2 = qx[i] * qx[i] + gy2;
        const float ap = bwl_pot ? 0.5f - 0.5f * std::tanh((std::sqrt(g2) / kmax - 0.9f) * 30.f) : 1.f;
        // linear interpolation of the radial table at |q| (GetFormFactor, Atoms.cpp:414-441)
        const float fp = std::fabs(std::sqrt(g2)) / sk;
        size_t i0 = (size_t)fp;
        const float fi = fp - (float)i0;
        i0 = std::min(i0, (size_t)nk - 2);  // the table spans the grid; the clamp only guards its end
        const cf ff = tab[i0] * (1.0f - fi) + tab[i0 + 1] * fi;
        dst[i + (size_t)j * nx] = ff * ap;
      }
    }
  }
  return 0;
}

int jms_zslice(int nat, const float* fz, float fz0, float fz1, int* idx, float* fz_new) {
  const float w = fz1 - fz0;
  if (nat < 0 || (nat > 0 && !fz)) return -1;
  if (w <= 0.f) return -2;
  int k = 0;
  for (int i = 0; i < nat; ++i) {
    float z = std::fmod(fz[i], 1.f);
    if (z < 0.f) z += 1.f;
    if (z >= fz0 && z < fz1) {
      if (idx) idx[k] = i;
      if (fz_new) fz_new[k] = (z - fz0) / w;
      ++k;
    }
  }
  return k;
}

int jms_suggest_zslices(int nat, const int* z, const float* fz, float cz, float dmin) {
  if (nat < 2) return 1;
  if (!z || !fz) return -1;
  if (cz <= 0.f) return 0;
  const float dstep = std::max(0.0001f, dmin);
  const float rfac = -1.f / (dstep * dstep);
  const int imin = 8 * (1 + (int)(cz / dstep));
  const int nh = std::max(64, (int)std::pow(2.f, std::ceil(std::log((float)imin) / std::log(2.0f))));
  const int nyq = nh >> 1;
  const float hstep = cz / (float)nh;
  // z density: a Gaussian of width dstep and height Z per atom
  std::vector<float> hist(nh, 0.f);
  for (int a = 0; a < nat; ++a) {
    const float zamp = (float)z[a];
    const float zpos = fz[a] * cz;
    for (int i = 0; i < nh; ++i) {
      const float dz = (float)i * hstep - zpos;
      hist[i] += zamp * std::exp(rfac * dz * dz);
    }
  }
  // |DFT| at frequencies 1 .. nyq-1; the strongest one from 2 on, if it
  // beats frequency 1, gives the slice count
  auto amp = [&](int k) {
    double re = 0., im = 0.;
    for (int i = 0; i < nh; ++i) {
      const double ph = -kTwoPi * (double)((long long)k * i % nh) / (double)nh;
      re += (double)hist[i] * std::cos(ph);
      im += (double)hist[i] * std::sin(ph);
    }
    return (float)std::sqrt(re * re + im * im);
  };
  int nslc = 1;
  float best = amp(1);
  for (int k = 2; k < nyq; ++k) {
    const float a = amp(k);
    if (a > best) { best = a; nslc = k; }
  }
  return nslc;
}

}  // extern "C"
