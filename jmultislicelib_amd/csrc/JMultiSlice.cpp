// JMultiSlice.cpp -- CJMultiSlice on MI355X: the reference class interface
// (src/JMultiSlice.h:187-937) over the HIP engine of include/jms_hip.h.
//
// Host-side calculators (propagator, ring detector, probe, detection planes,
// variant sequences) follow the reference algorithms cited per function; the
// slice loop, readouts and wave offsets run on the GPU through jms_hip_*.
#include "jms/JMultiSlice.h"
#include "jms_sequence.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <thread>

namespace {
const double kPI = 3.1415926535897932384626433832795;   // NatureConstants.h:36
const double kTPI = 6.283185307179586476925286766559;   // NatureConstants.h:37
const double kD2R = kPI / 180.;                          // NatureConstants.h:42
// CODATA 2018 (NatureConstants.h:48-66): wavelength [nm keV], rest energy [keV]
const double kC0 = 299792458.;
const double kHPL = 6.62607015E-34;
const double kQEL = 1.602176634E-19;
const double kME0 = 9.1093837015E-31;
const double kWLELKEV = kC0 * kHPL / kQEL * 1.0E+6;
const double kEEL0KEV = kME0 * kC0 * kC0 / kQEL / 1000.;

// frequency number of grid index i on an n-point FFT grid (JMultiSlice.cpp:1699-1712)
inline int freq_index(int i, int n) {
  const int h = (n - (n % 2)) >> 1;
  return ((i + h) % n) - h;
}

bool gpu_code(int whichcode) { return (whichcode & _JMS_CODE_GPU) != 0; }
bool cpu_code(int whichcode) { return (whichcode & _JMS_CODE_CPU) != 0; }

int reject_cpu(const char* fn) {
  std::cerr << "Error: (" << fn << "): the CPU code path is not part of this library." << std::endl;
  return 1;
}

// transposed copy: input has nx rows of ny, output ny rows of nx
// (JMultiSlice.cpp:1818-1833)
template <class T>
std::vector<T> transpose_in(const T* src, int nx, int ny) {
  std::vector<T> out((size_t)nx * ny);
  for (int j = 0; j < nx; ++j)
    for (int i = 0; i < ny; ++i) out[(size_t)j + (size_t)i * nx] = src[(size_t)i + (size_t)j * ny];
  return out;
}
}  // namespace

CJMultiSlice::CJMultiSlice() : m_lrng(), m_prng(&m_lrng) {}

CJMultiSlice::~CJMultiSlice() { Cleanup(); }

int CJMultiSlice::SetDebugLevel(int dbg) {
  const int old = m_dbg;
  m_dbg = dbg;
  return old;
}

int CJMultiSlice::Cleanup(void) {
  if (m_eng) {
    jms_hip_destroy(m_eng);
    m_eng = nullptr;
  }
  m_objslc.clear();
  m_det_objslc.clear();
  m_nvarslc.clear();
  m_proidx.clear();
  m_slcthick.clear();
  m_d_det_int.clear();
  m_nscslc = m_nslcvar = m_nobjslc = 0;
  m_npgx = m_npgy = 0;
  m_npro = 0;
  m_ndet = m_ndetper = m_ndetslc = 0;
  m_imagedet = 0;
  m_status_setup_GPU = 0;
  m_wave_ready = false;
  return 0;
}

void CJMultiSlice::FreeLibMem(void) {}

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------
float CJMultiSlice::SetHighTension(float ht) {  // JMultiSlice.cpp:372-378
  const float old = m_ht;
  m_ht = ht;
  m_wl = (float)(kWLELKEV / std::sqrt(ht * (ht + 2. * kEEL0KEV)));
  return old;
}
float CJMultiSlice::GetHighTension(void) { return m_ht; }
float CJMultiSlice::GetWaveLength(void) { return m_wl; }

void CJMultiSlice::SetSupercellSize(float* a0) {
  if (a0) { m_a0[0] = a0[0]; m_a0[1] = a0[1]; m_a0[2] = a0[2]; }
}
void CJMultiSlice::SetSupercellSize(float a, float b, float c) {
  m_a0[0] = a; m_a0[1] = b; m_a0[2] = c;
}
void CJMultiSlice::SetGridSize(int nx, int ny) { m_nscx = nx; m_nscy = ny; }
int CJMultiSlice::GetDetetionSliceNum(void) { return m_ndetslc; }
int CJMultiSlice::GetDetNum(void) { return m_ndet; }
void CJMultiSlice::GetGridSize(int& nx, int& ny) { nx = m_nscx; ny = m_nscy; }
void CJMultiSlice::GetSupercellSize(float* a0) {
  if (a0) { a0[0] = m_a0[0]; a0[1] = m_a0[1]; a0[2] = m_a0[2]; }
}
void CJMultiSlice::GetSupercellSize(float& a, float& b, float& c) {
  a = m_a0[0]; b = m_a0[1]; c = m_a0[2];
}
void CJMultiSlice::SetSliceThickness(int islc, float t) {
  if (islc >= 0 && islc < (int)m_slcthick.size()) m_slcthick[islc] = t;
}
float CJMultiSlice::GetSliceThickness(int islc) {
  return (islc >= 0 && islc < (int)m_slcthick.size()) ? m_slcthick[islc] : 0.f;
}
void CJMultiSlice::SetPlasmonMC(bool use, float, float, float, unsigned int nexmax) {
  m_plasmonmc_flg = (use && nexmax > 0) ? 1 : 0;
  if (m_plasmonmc_flg)
    std::cerr << "Warning: (SetPlasmonMC): plasmon Monte-Carlo is not implemented; multislice calls will fail."
              << std::endl;
}
void CJMultiSlice::DiffractionDescan(bool b) { m_dif_descan_flg = b ? 1 : 0; }

// SetDiffractionDescanN / SetDiffractionDescan / SetDiffractionDescanMRad
// (JMultiSlice.cpp:1118-1164), GPU branch: the pixel shift takes effect in the
// next multislice run while the de-scan flag is on
void CJMultiSlice::SetDiffractionDescanN(int whichcode, int ndescanx, int ndescany, int iThread) {
  (void)iThread;
  if (cpu_code(whichcode) || !gpu_code(whichcode)) return;
  if ((m_status_setup_GPU & _JMS_THRESHOLD_CALC) == 0) return;
  m_d_dif_ndescanx = ndescanx;
  m_d_dif_ndescany = ndescany;
}

void CJMultiSlice::SetDiffractionDescan(int whichcode, float descanx, float descany, int iThread) {
  SetDiffractionDescanN(whichcode, (int)std::round(descanx * m_a0[0]), (int)std::round(descany * m_a0[1]),
                        iThread);  // 1/nm -> pixel
}

void CJMultiSlice::SetDiffractionDescanMRad(int whichcode, float descanx, float descany, int iThread) {
  SetDiffractionDescan(whichcode, descanx * 0.001f / m_wl, descany * 0.001f / m_wl, iThread);  // mrad -> 1/nm
}

int CJMultiSlice::ApplyDescan() {
  const int dx = m_dif_descan_flg ? m_d_dif_ndescanx : 0;
  const int dy = m_dif_descan_flg ? m_d_dif_ndescany : 0;
  return jms_hip_set_descan(m_eng, dx, dy);
}

// JMultiSlice.cpp:579-627: plane -> detection index hash, returns #planes
int CJMultiSlice::SetDetectionPlanes(int ndetper, int nobjslc, int* det_objslc) {
  const int nos = std::max(0, nobjslc);
  if (det_objslc && nobjslc > 0) std::fill(det_objslc, det_objslc + nobjslc + 1, -1);
  int ndetslc = 0;
  if (ndetper <= 0) {  // exit plane only
    if (det_objslc) det_objslc[nos] = 0;
    return 1;
  }
  if (nos == 0) {  // entrance plane only
    if (det_objslc) det_objslc[0] = 0;
    return 1;
  }
  for (int i = 0; i <= nos; i += ndetper) {  // entrance plane and every ndetper-th plane
    if (det_objslc) det_objslc[i] = ndetslc;
    ++ndetslc;
  }
  if (nos % ndetper != 0) {  // exit plane off the period
    if (det_objslc) det_objslc[nos] = ndetslc;
    ++ndetslc;
  }
  return ndetslc;
}

// JMultiSlice.cpp:630-665 (detection planes listed per plane, -1 = none)
int CJMultiSlice::SetDetectionPlanes(std::vector<int> slc_det, int nobjslc, int* det_objslc) {
  const int nlist = (int)slc_det.size();
  const int nos = std::min(std::max(0, nobjslc), nlist);
  if (det_objslc && nobjslc > 0) std::fill(det_objslc, det_objslc + nobjslc + 1, -1);
  if (nlist <= 0) {
    if (det_objslc) det_objslc[nobjslc] = 0;
    return 1;
  }
  if (nos <= 0) {
    if (det_objslc) det_objslc[0] = 0;
    return 1;
  }
  int ndetslc = 0;
  for (int i = 0; i <= nos && i < nlist; ++i) {
    if (slc_det[i] >= 0) {
      if (det_objslc) det_objslc[i] = slc_det[i];
      ++ndetslc;
    }
  }
  return ndetslc;
}

// ---------------------------------------------------------------------------
// calculators
// ---------------------------------------------------------------------------
// Axial aberrations of CJProbeGen: the interface list enumerates (order m,
// rotational symmetry n) for m = 1..8, n = 0..m, m + n even
// (JProbeGen.cpp:187-216), and the phase is
//   chi(q) = 2 pi / wl * sum_k Re[(ax - i ay)(wx + i wy)^n] |w|^(m-n) / m,
// w = wl q, summed term by term in float as AberrationFunction does
// (JProbeGen.cpp:508-571).
namespace {
constexpr int kAbOrderMax = 8;  // _JPG_ABERRATION_ORDER_MAX (JProbeGen.h:53)
constexpr int kAbNum = (4 * kAbOrderMax - kAbOrderMax % 2 + kAbOrderMax * kAbOrderMax) / 4;

struct AberrationTable {
  int mn[2 * kAbNum];
  int binom[(2 * kAbOrderMax + 1) * (2 * kAbOrderMax + 1)];
  AberrationTable() {
    int l = 0;
    for (int m = 1; m <= kAbOrderMax; m++)
      for (int n = 0; n <= m; n++)
        if ((m + n) % 2 == 0) { mn[2 * l] = m; mn[2 * l + 1] = n; l++; }
    const int nol = 2 * kAbOrderMax + 1;
    for (int k = 0; k < nol; k++)
      for (int n = 0; n < nol; n++) {  // C(n, k) by the multiplicative formula
        int res = 0;
        if (n >= k) {
          res = 1;
          const int lk = std::min(k, n - k);
          for (int i = 0; i < lk; ++i) { res *= (n - i); res /= (i + 1); }
        }
        binom[n + k * nol] = res;
      }
  }
};

float aberration_phase(float qx, float qy, float wl, int nabrr, const float* coeff) {
  static const AberrationTable T;
  if (nabrr == 0 || !coeff || wl <= 0.f) return 0.f;
  constexpr int nol = kAbOrderMax + 1, nol2 = 2 * kAbOrderMax + 1;
  const int nab = std::min(nabrr, kAbNum);
  const float re_il[4] = {1.f, 0.f, -1.f, 0.f};  // Re(i^l); Im(i^l) = re_il[(l + 3) % 4]
  const float wx = qx * wl, wy = qy * wl;
  const float w = std::sqrt(wx * wx + wy * wy);
  const float prefac = (float)(2.0 * kPI) / wl;
  float pwr_x[nol], pwr_y[nol], pwr_w[nol];
  pwr_x[0] = pwr_y[0] = pwr_w[0] = 1.f;
  for (int i = 1; i < nol; i++) {
    pwr_x[i] = pwr_x[i - 1] * wx;
    pwr_y[i] = pwr_y[i - 1] * wy;
    pwr_w[i] = pwr_w[i - 1] * w;
  }
  float sum = 0.f;
  for (int k = 0; k < nab; k++) {
    const float cx = coeff[2 * k], cy = coeff[2 * k + 1];
    if ((cx * cx + cy * cy) > (float)(1.E-15)) {  // _JPG_ABERRATION_STRTHRESH (JProbeGen.h:54)
      const int m = T.mn[2 * k], n = T.mn[2 * k + 1];
      float term = 0.f;
      for (int l = 0; l <= n; l++) {  // binomial expansion of (wx + i wy)^n
        float c = 0.f;
        c = c + re_il[l % 4] * cx;
        c = c + re_il[(l + 3) % 4] * cy;
        term = term + c * (float)T.binom[n + l * nol2] * pwr_x[n - l] * pwr_y[l];
      }
      term = term * pwr_w[m - n] / ((float)m);
      sum = sum + term;
    }
  }
  return sum * prefac;
}
}  // namespace

// CJProbeGen::CalculateProbeWaveFourier (JProbeGen.cpp:707-777) with the
// aperture of ApertureFunctionA (JProbeGen.cpp:471-505) and the axial
// aberration phase of m_abrr_num coefficient pairs: psi = A exp(-i chi).
int CJMultiSlice::CalculateProbeWaveFourier(CJProbeParams* prm, fcmplx* wav) {
  if (!prm || !wav) return 1;
  if (m_wl <= 0.f || m_a0[0] * m_a0[1] <= 0.f || m_nscx * m_nscy <= 0 || prm->m_alpha <= 0.f) return 1;
  const int nab = (prm->m_abrr_coeff && prm->m_abrr_num > 0) ? prm->m_abrr_num : 0;
  const int nx = m_nscx, ny = m_nscy;
  const float wl = m_wl, ax = m_a0[0], ay = m_a0[1];
  const float qlim = prm->m_alpha * 0.001f / wl;
  const float qrs = prm->m_alpha_rs, alim = prm->m_alpha_asym;
  const float adir = prm->m_alpha_adir * 0.0174533f;
  const float qlcx = prm->m_alpha_x0 * 0.001f / wl, qlcy = prm->m_alpha_y0 * 0.001f / wl;
  const float btx = prm->m_btx * 0.001f / wl, bty = prm->m_bty * 0.001f / wl;
  // aperture distortion (asymmetry) matrix, inverted
  const float a1x = alim * std::cos(2.f * adir), a1y = alim * std::sin(2 * adir);
  const float adet = 1.f - alim * alim;
  std::vector<float> qx(nx), qy(ny);
  for (int i = 0; i < nx; ++i) qx[i] = (float)freq_index(i, nx) / ax;
  for (int j = 0; j < ny; ++j) qy[j] = (float)freq_index(j, ny) / ay;
  std::fill(wav, wav + (size_t)nx * ny, fcmplx(0.f, 0.f));
  float wpow = 0.f;
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      float amp = 0.f;
      if (qlim > 0.f && adet != 0.f) {
        const float radet = 1.f / adet;
        const float dqx = (qx[i] - btx) - qlcx, dqy = (qy[j] - bty) - qlcy;
        const float ux = ((1.f - a1x) * dqx - a1y * dqy) * radet;
        const float uy = (-a1y * dqx + (1.f + a1x) * dqy) * radet;
        const float dqm = std::sqrt(ux * ux + uy * uy);
        const float dqs = std::abs(qrs * qlim);
        if (dqm <= 0.f) amp = 1.f;
        else if (dqs > 0.f) amp = (1.f - std::tanh((float)kPI * (dqm - qlim) / dqs)) * 0.5f;
        else if (dqm < qlim) amp = 1.f;
      }
      if (amp > 1.0e-6f) {  // _JPG_PROBE_APERTURE_THRESH (JProbeGen.h:52)
        const float chi = nab > 0 ? aberration_phase(qx[i], qy[j], wl, nab, prm->m_abrr_coeff) : 0.f;
        wav[(size_t)i + (size_t)j * nx] = fcmplx(amp * std::cos(chi), -amp * std::sin(chi));
        wpow += amp * amp;
      }
    }
  }
  const float fsca = 1.f / std::sqrt(wpow);
  for (size_t k = 0; k < (size_t)nx * ny; ++k) wav[k] = wav[k] * fsca;
  return 0;
}

// CJMultiSlice::CalculatePropagator (JMultiSlice.cpp:998-1106)
int CJMultiSlice::CalculatePropagator(float fthick, float otx, float oty, fcmplx* pro, int ntype) {
  if (!pro || ntype < 0 || ntype > 1) return 1;
  const int nx = m_nscx, ny = m_nscy;
  const int nyqx = (nx - (nx % 2)) >> 1, nyqy = (ny - (ny % 2)) >> 1;
  const double itogx = 1. / m_a0[0], itogy = 1. / m_a0[1];
  const double gthr = std::min(itogx * nyqx, itogy * nyqy) * _JMS_RELAPERTURE;
  const double gthr2 = gthr * gthr;
  const double otxr = (double)otx * kD2R, otyr = (double)oty * kD2R;
  const double wl = (double)m_wl;
  const double ot = std::sqrt(otxr * otxr + otyr * otyr);
  const float pronorm = 1.f / ((float)nx * ny);  // restores the norm after IFT+FT
  m_nprotype = ntype;
  if (ntype == 0) {  // planar (large-angle) propagator
    const double pfac = 2. * kPI / wl, cot = std::cos(ot), sot = std::sin(ot);
    const double od = (ot > 0.0) ? std::atan2(otyr, otxr) : 0.;
    for (int j = 0; j < ny; ++j) {
      const double gy = itogy * freq_index(j, ny);
      for (int i = 0; i < nx; ++i) {
        const double gx = itogx * freq_index(i, nx);
        const double g2 = gx * gx + gy * gy;
        fcmplx& o = pro[(size_t)i + (size_t)j * nx];
        if (g2 > gthr2) { o = fcmplx(0.f, 0.f); continue; }
        const double dt = std::asin(wl * std::sqrt(g2));
        const double dd = (dt > 0.) ? std::atan2(gy, gx) : 0.;
        const double chi = pfac * (double)fthick *
                           (1. / (cot * std::cos(dt) + sot * std::sin(dt) * std::cos(od - dd)) - 1. / cot);
        o = fcmplx((float)std::cos(chi) * pronorm, (float)(-std::sin(chi) * pronorm));
      }
    }
  } else {  // Fresnel (paraxial) propagator with object tilt
    const double pfac = kPI / wl * std::cos(ot);
    const double pt0 = pfac * (otxr * otxr + otyr * otyr);
    const double wthr2 = gthr2 * wl * wl;
    for (int j = 0; j < ny; ++j) {
      const double wy = itogy * wl * freq_index(j, ny);
      const double uy = wy - otyr;
      for (int i = 0; i < nx; ++i) {
        const double wx = itogx * wl * freq_index(i, nx);
        const double ux = wx - otxr;
        fcmplx& o = pro[(size_t)i + (size_t)j * nx];
        if (wx * wx + wy * wy > wthr2) { o = fcmplx(0.f, 0.f); continue; }
        const double chi = (double)fthick * (pfac * (ux * ux + uy * uy) - pt0);
        o = fcmplx((float)std::cos(chi) * pronorm, (float)(-std::sin(chi) * pronorm));
      }
    }
  }
  return 0;
}

// CJMultiSlice::CalculateRingDetector (JMultiSlice.cpp:1274-1438), with the
// sensitivity profile file of LoadSTEMDetectorProfile (:1127-1245) and the
// radial lookup of GetRadialDetectorSensitivity (:1247-1272).
int CJMultiSlice::CalculateRingDetector(float beta0, float beta1, float phi0, float phi1, float theta0x,
                                        float theta0y, std::string sdsprofile, float* det, int& msklen,
                                        int* msk, int kmom1) {
  const int nx = m_nscx, ny = m_nscy, nitems = nx * ny;
  msklen = 0;
  if (!det || nitems <= 0) return 1;
  std::fill(det, det + nitems, 0.f);
  if (msk) std::fill(msk, msk + nitems, 0);
  // optional radial sensitivity profile: header "len, refpix", then len values
  std::vector<float> prof;
  float refpix = 0.f;
  if (!sdsprofile.empty()) {
    std::ifstream in(sdsprofile);
    bool ok = in.is_open();
    int len = 0;
    if (ok) {
      std::string line;
      ok = (bool)std::getline(in, line);
      if (ok) {
        for (char& c : line) if (c == '\'' || c == ',') c = ' ';
        ok = (std::sscanf(line.c_str(), "%d %f", &len, &refpix) == 2) && len > 0;
      }
      while (ok && (int)prof.size() < len && std::getline(in, line)) prof.push_back(std::stof(line));
      ok = ok && (int)prof.size() == len;
    }
    if (!ok) {
      std::cout << "(CJMultiSlice::CalculateRingDetector) Warning: Failed to read detector sensitivity profile."
                << std::endl;
      prof.clear();
    }
  }
  auto radial = [&](float theta, float beta) {
    float sens = 1.0f;
    const float rpix = std::max(0.f, refpix - 1.f);
    const int len = (int)prof.size();
    if (len > 0 && rpix >= 0.f) {
      const float tpix = std::max(0.f, theta * rpix / beta);
      const int ipix = (int)tpix;
      const float tfrac = tpix - ipix;
      sens = (ipix > len - 2) ? prof[len - 1] : (1.f - tfrac) * prof[ipix] + tfrac * prof[ipix + 1];
    }
    return sens;
  };
  const double itogx = 1. / m_a0[0], itogy = 1. / m_a0[1];
  const double k0 = 1. / (double)m_wl;
  double dr1 = (double)beta0 * 0.001, dr2 = (double)beta1 * 0.001;
  double da1 = (double)phi0 * kD2R, da2 = (double)phi1 * kD2R;
  const double dc1 = (double)theta0x * 0.001, dc2 = (double)theta0y * 0.001;
  double gcx = 0., gcy = 0.;  // detector centre [1/nm]
  if (m_nprotype == 1) {
    // small-angle mapping; the reference leaves the centre at zero in this
    // mode (JMultiSlice.cpp:1331-1336), kept for identical detector masks
    dr1 *= k0;
    dr2 *= k0;
  } else {
    dr1 = k0 * std::sin(dr1);
    dr2 = k0 * std::sin(dr2);
    const double gc = k0 * std::sin(std::sqrt(dc1 * dc1 + dc2 * dc2));
    const double ph = (gc > 0.) ? std::atan2(dc2, dc1) : 0.;
    gcx = gc * std::cos(ph);
    gcy = gc * std::sin(ph);
  }
  const bool full_ring = (std::abs(da1 - da2) < 0.0001) || (std::abs(std::abs(da1 - da2) - kTPI) < 0.0001);
  if (!full_ring) {
    while (da1 < 0.0 || da2 < 0.0) { da1 += kTPI; da2 += kTPI; }
    if (da1 >= da2) da2 += kTPI;
    while (da1 >= kTPI) { da1 -= kTPI; da2 -= kTPI; }
  }
  for (int j = 0; j < ny; ++j) {
    const double gy = itogy * freq_index(j, ny) - gcy;
    for (int i = 0; i < nx; ++i) {
      const double gx = itogx * freq_index(i, nx) - gcx;
      const double g = std::sqrt(gy * gy + gx * gx);
      if (g < dr1 || g >= dr2) continue;  // half-open ring [dr1, dr2)
      if (!full_ring) {
        double phi = (g > 0.) ? std::atan2(gy, gx) : 0.;
        while (phi < da1) phi += kTPI;
        if (phi >= da2) continue;
      }
      float sens = 1.0f;
      if (!prof.empty()) {
        if (dr1 > 0.) sens = radial((float)g, (float)dr1);
        else if (dr2 > 0.) sens = radial((float)g, (float)dr2);
      }
      if (kmom1 == 1) sens *= (float)gx;
      if (kmom1 == 2) sens *= (float)gy;
      const int idx = i + j * nx;
      det[idx] = sens;
      if (msk) msk[msklen++] = idx;
    }
  }
  return 0;
}

// JMultiSlice.cpp:2086-2110: a_unscrambled[idx] = a_scrambled[phash[idx]]
int CJMultiSlice::GetUnscrambleHash(unsigned int* phash) {
  if (!phash) return 1;
  const int nx = m_nscx, ny = m_nscy;
  const int nx2 = (nx - (nx % 2)) >> 1, ny2 = (ny - (ny % 2)) >> 1;
  for (int j = 0; j < ny; ++j) {
    const int j1 = (freq_index(j, ny) + ny2) % ny;
    for (int i = 0; i < nx; ++i) {
      const int i1 = (freq_index(i, nx) + nx2) % nx;
      phash[(size_t)i1 + (size_t)j1 * nx] = (unsigned int)(i + j * nx);
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// RNG and variant sequences
// ---------------------------------------------------------------------------
void CJMultiSlice::SetRng(CRng* prng) { m_prng = prng ? prng : &m_lrng; }
CRng* CJMultiSlice::GetRng(void) { return m_prng; }
void CJMultiSlice::SeedRngEx(int nseed) { (m_prng ? m_prng : &m_lrng)->seed(nseed); }

namespace {
// JMultiSlice.cpp:1571-1676: per object slice a variant index drawn without
// returning from a per-structure-slice queue; a new round never starts with
// the variant that ended the previous one.
int variant_sequence(CRng* rng, int nscslc, const int* nvarslc, int nobjslc, const int* objslc, int* out) {
  if (!out) return 1;
  if (!nvarslc || !objslc || nscslc < 1 || nobjslc < 1) return 2;
  int qmax = 0;
  for (int s = 0; s < nscslc; ++s) qmax = std::max(qmax, nvarslc[s]);
  std::vector<int> queue((size_t)nscslc * std::max(1, qmax));
  std::vector<int> left(nscslc, 0), last(nscslc, -1);
  for (int i = 0; i < nobjslc; ++i) {
    const int s = objslc[i];
    const int nv = nvarslc[s];
    if (nv < 2) { out[i] = 0; continue; }
    int* q = &queue[(size_t)s * qmax];
    if (left[s] <= 0) {  // refill
      for (int v = 0; v < nv; ++v) q[v] = v;
      left[s] = nv;
    }
    int k = ((int)rng->unirand(0., (double)nv)) % nv;
    const int k0 = k;
    while (q[k] < 0 || q[k] == last[s]) {
      k = (k + 1) % nv;
      if (k == k0) break;
    }
    int var = q[k];
    if (var < 0) {  // not reachable for nv >= 2; kept as the reference's fallback
      var = (int)rng->unirand(0.0, (double)nv);
      left[s] = 0;
    } else {
      q[k] = -1;
      --left[s];
    }
    last[s] = var;
    out[i] = var;
  }
  return 0;
}
}  // namespace

int CJMultiSlice::GetRandomVariantSequence(int* out) {
  return variant_sequence(m_prng, m_nscslc, m_nvarslc.data(), m_nobjslc, m_objslc.data(), out);
}

extern "C" int jms_variant_sequences(int seed, int nscslc, const int* nvarslc, int nobjslc,
                                     const int* objslc, int nseq, int* out) {
  if (nseq < 0 || !out) return 1;
  for (int s = 0; s < nscslc; ++s)
    if (!nvarslc || nvarslc[s] < 1) return 3;
  for (int j = 0; j < nobjslc; ++j)
    if (!objslc || objslc[j] < 0 || objslc[j] >= nscslc) return 4;
  CRng rng(seed);
  for (int k = 0; k < nseq; ++k) {
    const int rc = variant_sequence(&rng, nscslc, nvarslc, nobjslc, objslc, out + (size_t)k * nobjslc);
    if (rc) return rc;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// device setup
// ---------------------------------------------------------------------------
int CJMultiSlice::EnsureEngine(void) {
  if (!m_eng) {
    if (jms_hip_create(&m_eng) != 0) {
      std::cerr << "Error: failed to create the HIP engine (no usable GPU?)" << std::endl;
      m_eng = nullptr;
      return 101;
    }
  }
  const float ax = m_a0[0] > 0.f ? m_a0[0] : 1.f;
  const float ay = m_a0[1] > 0.f ? m_a0[1] : 1.f;
  int nx = 0, ny = 0;
  (void)nx; (void)ny;
  if (jms_hip_set_grid(m_eng, m_nscx, m_nscy, ax, ay, m_wl) != 0) return 102;
  return 0;
}

int CJMultiSlice::PhaseGratingSetup(int whichcode, int nx, int ny, int nslc, int nvarmax, int* nslcvar) {
  if (cpu_code(whichcode)) return reject_cpu("PhaseGratingSetup");
  m_npgx = m_npgy = m_nscslc = 0;
  m_nslcvar = 1;
  m_nvarslc.clear();
  m_slcthick.clear();
  if (nslc < 1 || !nslcvar) return 0;  // no phase gratings
  if (nx != m_nscx || ny != m_nscy) {
    std::cerr << "Error: (PhaseGratingSetup): phase gratings must have the wave grid size." << std::endl;
    return 1;
  }
  m_npgx = nx; m_npgy = ny; m_nscslc = nslc; m_nslcvar = nvarmax;
  m_nvarslc.assign(nslcvar, nslcvar + nslc);
  for (int& v : m_nvarslc) v = std::min(v, m_nslcvar);  // JMultiSlice.cpp:470
  m_slcthick.assign(nslc, 0.f);
  if (gpu_code(whichcode)) {
    int rc = EnsureEngine();
    if (rc) return rc;
    if (jms_hip_setup_gratings(m_eng, nx, ny, nslc, m_nvarslc.data()) != 0) return 104;
    m_status_setup_GPU |= _JMS_STATUS_PGR;
  }
  return 0;
}

int CJMultiSlice::ObjectSliceSetup(int nobjslc, int* objslc) {
  m_nobjslc = 0;
  m_objslc.clear();
  if (nobjslc >= 1 && objslc) {
    m_nobjslc = nobjslc;
    m_objslc.assign(objslc, objslc + nobjslc);
  }
  // no object slices is a valid set-up (JMultiSlice.cpp:528-537): the
  // multislice then reads the incident wave out at the entrance plane
  m_status_setup_GPU |= _JMS_STATUS_OBJ;
  return 0;
}

int CJMultiSlice::PropagatorSetup(int whichcode, int npro, int* proidx) {
  if (cpu_code(whichcode)) return reject_cpu("PropagatorSetup");
  m_npro = 0;
  if (npro < 1 || !proidx) return 0;
  m_npro = npro;
  m_proidx.assign(proidx, proidx + m_nscslc);
  if (gpu_code(whichcode)) {
    int rc = EnsureEngine();
    if (rc) return rc;
    if (jms_hip_setup_propagators(m_eng, npro, m_nscslc, m_proidx.data()) != 0) return 102;
    m_status_setup_GPU |= _JMS_STATUS_PRO;
  }
  return 0;
}

int CJMultiSlice::DetectorSetup(int whichcode, int ndetper, int ndetint, int imagedet, int nthreads_CPU) {
  (void)nthreads_CPU;
  if (cpu_code(whichcode)) return reject_cpu("DetectorSetup");
  if ((m_status_setup_GPU & _JMS_STATUS_OBJ) == 0) {
    std::cerr << "Error: (DetectorSetup): no previous object slice setup for GPU code." << std::endl;
    return 101;
  }
  m_ndetper = std::max(0, ndetper);
  m_det_objslc.assign((size_t)m_nobjslc + 1, -1);
  m_ndetslc = SetDetectionPlanes(m_ndetper, m_nobjslc, m_det_objslc.data());
  return DetectorChannels(whichcode, ndetint, imagedet);
}

int CJMultiSlice::DetectorSetup(int whichcode, std::vector<int> slc_det, int ndetint, int imagedet,
                                int nthreads_CPU) {
  (void)nthreads_CPU;
  if (cpu_code(whichcode)) return reject_cpu("DetectorSetup");
  if ((m_status_setup_GPU & _JMS_STATUS_OBJ) == 0) {
    std::cerr << "Error: (DetectorSetup): no previous object slice setup for GPU code." << std::endl;
    return 101;
  }
  m_ndetper = 0;
  m_det_objslc.assign((size_t)m_nobjslc + 1, -1);
  m_ndetslc = SetDetectionPlanes(slc_det, m_nobjslc, m_det_objslc.data());
  return DetectorChannels(whichcode, ndetint, imagedet);
}

// channel part of DetectorSetup (JMultiSlice.cpp:700-811, GPU branch)
int CJMultiSlice::DetectorChannels(int whichcode, int ndetint, int imagedet) {
  m_imagedet = imagedet;
  m_ndet = std::max(0, ndetint);
  if (imagedet & ~JMS_HIP_DETECT_ALL) {
    std::cerr << "Error: (DetectorSetup): unknown detection channel flags." << std::endl;
    return 102;
  }
  m_d_det_int.assign((size_t)m_ndet * m_ndetslc, 0.f);
  if (gpu_code(whichcode)) {
    int rc = EnsureEngine();
    if (rc) return rc;
    // the engine validates the plane hash against the object slices
    if (jms_hip_setup_objects(m_eng, m_nobjslc, m_objslc.data()) != 0) return 104;
    if (jms_hip_setup_detection(m_eng, m_ndetslc, m_det_objslc.data(), m_ndet, imagedet) != 0) return 103;
    m_status_setup_GPU |= _JMS_STATUS_DET;
  }
  return 0;
}

int CJMultiSlice::SetPhaseGratingData(int whichcode, int islc, int nvar, fcmplx* pgr) {
  if (cpu_code(whichcode)) return reject_cpu("SetPhaseGratingData");
  if (islc < 0 || islc >= m_nscslc || !pgr) return 0;  // ignored as in the reference
  const int nl = std::min(m_nvarslc[islc], nvar);
  if (nl < m_nvarslc[islc]) m_nvarslc[islc] = nl;  // JMultiSlice.cpp:1450-1453
  if (nl <= 0) return 1;
  if (gpu_code(whichcode)) {
    if (!m_eng) return 101;
    if (jms_hip_upload_gratings(m_eng, islc, nl, (const float*)pgr) != 0) return 102;
  }
  return 0;
}

int CJMultiSlice::SetPropagatorData(int whichcode, int ipro, fcmplx* pro) {
  if (cpu_code(whichcode)) return reject_cpu("SetPropagatorData");
  if (ipro < 0 || ipro >= m_npro || !pro) return 0;
  if (gpu_code(whichcode)) {
    if (!m_eng) return 101;
    if (jms_hip_upload_propagator(m_eng, ipro, (const float*)pro) != 0) return 102;
  }
  return 0;
}

int CJMultiSlice::SetDetectorData(int whichcode, int idet, float* det, int msklen, int* msk) {
  if (cpu_code(whichcode)) return reject_cpu("SetDetectorData");
  if (idet < 0 || idet >= m_ndet || !det) return 0;
  if (gpu_code(whichcode)) {
    if (!m_eng) return 101;
    if (jms_hip_upload_detector(m_eng, idet, det, (msk && msklen > 0) ? msklen : 0, msk) != 0) return 102;
  }
  return 0;
}

int CJMultiSlice::InitCore(int whichcode, int nCPUthreads) {
  (void)nCPUthreads;
  if (cpu_code(whichcode)) return reject_cpu("InitCore");
  if (!gpu_code(whichcode)) return 0;
  if ((m_status_setup_GPU & _JMS_THRESHOLD_CORE) != _JMS_THRESHOLD_CORE) {
    std::cerr << "Error: (InitCore): incomplete GPU setup chain." << std::endl;
    return 103;
  }
  int rc = EnsureEngine();  // refreshes cell size and wavelength
  if (rc) return rc;
  if (jms_hip_setup_objects(m_eng, m_nobjslc, m_objslc.data()) != 0) return 104;
  m_status_setup_GPU |= _JMS_STATUS_CORE;
  return 0;
}

int CJMultiSlice::SetIncidentWave(int whichcode, fcmplx* wav, bool bTranspose) {
  if (cpu_code(whichcode)) return reject_cpu("SetIncidentWave");
  if (!wav) return 0;
  if ((m_status_setup_GPU & _JMS_THRESHOLD_CALC) != _JMS_THRESHOLD_CALC || !m_eng) return 0;
  std::vector<fcmplx> t;
  const fcmplx* src = wav;
  if (bTranspose) { t = transpose_in(wav, m_nscx, m_nscy); src = t.data(); }
  if (jms_hip_set_probe(m_eng, (const float*)src) != 0) return 102;
  return 0;
}

int CJMultiSlice::SetIncidentWavePhaseplate(int whichcode, float* ppl, bool bTranspose) {
  if (cpu_code(whichcode)) return reject_cpu("SetIncidentWavePhaseplate");
  if (!ppl || !m_eng) return 0;
  std::vector<float> t;
  const float* src = ppl;
  if (bTranspose) { t = transpose_in(ppl, m_nscx, m_nscy); src = t.data(); }
  if (jms_hip_set_phaseplate(m_eng, src) != 0) return 102;
  return 0;
}

int CJMultiSlice::ZeroIncidentWavePhaseplate(int whichcode) {
  if (cpu_code(whichcode)) return reject_cpu("ZeroIncidentWavePhaseplate");
  if (!m_eng) return 0;
  return jms_hip_set_phaseplate(m_eng, nullptr) ? 101 : 0;
}

// ---------------------------------------------------------------------------
// device info
// ---------------------------------------------------------------------------
int CJMultiSlice::GetCPUNum(void) { return (int)std::thread::hardware_concurrency(); }
int CJMultiSlice::GetGPUNum(void) {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess ? n : 0;
}
int CJMultiSlice::GetGPUName(int idev, char* name) {
  if (!name) return 1;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, idev) != hipSuccess) return 2;
  std::strncpy(name, p.name, 255);
  name[255] = 0;
  return 0;
}
int CJMultiSlice::GetCurrentGPU(void) {
  int d = -1;
  return hipGetDevice(&d) == hipSuccess ? d : -1;
}
int CJMultiSlice::SetCurrentGPU(int idev) {
  if (m_eng) {
    std::cerr << "Error: (SetCurrentGPU): the device cannot change after the GPU setup started." << std::endl;
    return 2;
  }
  return hipSetDevice(idev) == hipSuccess ? 0 : 1;
}
int CJMultiSlice::GetGPUStats(int idev, int& iCMajor, int& iCMinor, int& iMaxThread, int64_t& memtotal,
                              int64_t& memfree) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, idev) != hipSuccess) return 1;
  iCMajor = p.major;
  iCMinor = p.minor;
  iMaxThread = p.maxThreadsPerBlock;
  size_t f = 0, t = 0;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(idev);
  const hipError_t e = hipMemGetInfo(&f, &t);
  (void)hipSetDevice(cur);
  if (e != hipSuccess) return 2;
  memtotal = (int64_t)t;
  memfree = (int64_t)f;
  return 0;
}
int CJMultiSlice::GetGPUMemInfo(size_t& memtotal, size_t& memfree) {
  return hipMemGetInfo(&memfree, &memtotal) == hipSuccess ? 0 : 1;
}
int CJMultiSlice::SetGPUPgrLoading(int npgrload) {
  // all phase gratings are resident in HBM (288 GB per MI355X); the
  // load-on-demand mode of the reference (JMultiSlice.cpp:3865) is not needed
  return npgrload == 0 ? 0 : 1;
}

// ---------------------------------------------------------------------------
// multislice
// ---------------------------------------------------------------------------
int CJMultiSlice::ClearDetMem_d(void) {
  std::fill(m_d_det_int.begin(), m_d_det_int.end(), 0.f);
  return 0;
}

int CJMultiSlice::ClearDetMem_h(int) { return reject_cpu("ClearDetMem_h"); }
int CJMultiSlice::SetIncomingWaveCPU(fcmplx*, bool, int) { return reject_cpu("SetIncomingWaveCPU"); }
int CJMultiSlice::CPUMultislice(int, int, float, int) { return reject_cpu("CPUMultislice"); }

int CJMultiSlice::OffsetIncomingWave(int whichcode, float dx, float dy, float dz, int iThread, int ibtx,
                                     int ibty, bool bcppl) {
  (void)iThread;
  if (cpu_code(whichcode)) return reject_cpu("OffsetIncomingWave");
  if ((m_status_setup_GPU & _JMS_THRESHOLD_CALC) != _JMS_THRESHOLD_CALC || !m_eng) return 0;
  if (m_nscx * m_nscy <= 0) return 1;
  if (jms_hip_wave_offset(m_eng, dx, dy, dz, ibtx, ibty, bcppl ? 1 : 0) != 0) return 106;
  m_wave_ready = true;
  return 0;
}

int CJMultiSlice::SetIncomingWaveGPU(fcmplx* wav, bool bTranspose) {
  if ((m_status_setup_GPU & _JMS_THRESHOLD_CALC) != _JMS_THRESHOLD_CALC || !m_eng || !wav) return 101;
  std::vector<fcmplx> t;
  const fcmplx* src = wav;
  if (bTranspose) { t = transpose_in(wav, m_nscx, m_nscy); src = t.data(); }
  if (jms_hip_wave_set(m_eng, (const float*)src) != 0) return 102;
  m_wave_ready = true;
  return 0;
}

int CJMultiSlice::GPUMultislice(int islc0, int accmode, float weight) {
  if (!m_eng || !m_wave_ready || m_nscx * m_nscy <= 0) return 1;
  if ((m_status_setup_GPU & _JMS_THRESHOLD_CALC) != _JMS_THRESHOLD_CALC) return 2;
  if (m_plasmonmc_flg) {
    std::cerr << "Error: (GPUMultislice): plasmon Monte-Carlo is not implemented." << std::endl;
    return 5;
  }
  if (ApplyDescan() != 0) return 6;
  std::vector<int> var((size_t)std::max(1, m_nobjslc), 0);
  const int i0 = std::max(0, std::min(islc0, m_nobjslc));
  if (i0 < m_nobjslc) GetRandomVariantSequence(var.data());  // drawn for the whole stack (:4582)
  if (jms_hip_multislice(m_eng, i0, var.data(), weight, accmode == 0 ? 0 : 1) != 0) return 200;
  if (m_ndet > 0 && (m_imagedet & _JMS_DETECT_INTEGRATED)) {
    if (jms_hip_get_result(m_eng, JMS_HIP_DETECT_INTEGRATED, m_d_det_int.data()) != 0) return 700;
  }
  return 0;
}

int CJMultiSlice::GetResult(int whichcode, int whichresult, float* dst, int iThread) {
  (void)iThread;
  if (!dst) return 0;
  if (cpu_code(whichcode)) return reject_cpu("GetResult");
  if (!gpu_code(whichcode) || m_ndetslc <= 0 || !m_eng) return 0;
  if (whichresult == _JMS_DETECT_INTEGRATED) {
    std::memcpy(dst, m_d_det_int.data(), sizeof(float) * m_d_det_int.size());
    return 0;
  }
  if (whichresult == _JMS_DETECT_DIFFRACTION || whichresult == _JMS_DETECT_WAVEFOURIER ||
      whichresult == _JMS_DETECT_IMAGE || whichresult == _JMS_DETECT_WAVEREAL) {
    return jms_hip_get_result(m_eng, whichresult, dst) ? 102 : 0;
  }
  return 1;
}

int CJMultiSlice::GetAvgResult(int whichcode, int whichresult, float* dst, float& weight, int iThread) {
  (void)iThread;
  weight = 0.f;
  if (!dst) return 0;
  if (cpu_code(whichcode)) return reject_cpu("GetAvgResult");
  if (!gpu_code(whichcode) || m_ndetslc <= 0 || !m_eng) return 0;
  return jms_hip_get_avg_result(m_eng, whichresult, dst, &weight) ? 102 : 0;
}

float CJMultiSlice::GetImageAveragingWeight(int whichcode, int iThread) {
  (void)iThread;
  float wi = 0.f;
  if (gpu_code(whichcode) && m_eng) jms_hip_get_averaging_weights(m_eng, &wi, nullptr);
  return wi;
}

float CJMultiSlice::GetWaveAveragingWeight(int whichcode, int iThread) {
  (void)iThread;
  float ww = 0.f;
  if (gpu_code(whichcode) && m_eng) jms_hip_get_averaging_weights(m_eng, nullptr, &ww);
  return ww;
}

int CJMultiSlice::ResetImageAveraging(int whichcode, int iThread) {
  (void)iThread;
  if (cpu_code(whichcode)) return reject_cpu("ResetImageAveraging");
  if (!gpu_code(whichcode) || !m_eng) return 0;
  return jms_hip_reset_averaging(m_eng, 0) ? 101 : 0;
}

int CJMultiSlice::ResetWaveAveraging(int whichcode, int iThread) {
  (void)iThread;
  if (cpu_code(whichcode)) return reject_cpu("ResetWaveAveraging");
  if (!gpu_code(whichcode) || !m_eng) return 0;
  return jms_hip_reset_averaging(m_eng, 1) ? 103 : 0;
}

int CJMultiSlice::ReadoutDetAvg(int whichcode, int whichresult, float* dst, float& weight, int iThread) {
  (void)iThread;
  weight = 0.f;
  if (!dst) return 0;
  if (cpu_code(whichcode)) return reject_cpu("ReadoutDetAvg");
  if (!gpu_code(whichcode) || m_ndetslc <= 0 || !m_eng) return 0;
  return jms_hip_readout_det_avg(m_eng, whichresult, dst, &weight) ? 121 : 0;
}

int CJMultiSlice::GPUMultisliceScanVar(int npos, const float* dx, const float* dy, const float* dz, int npass,
                                       const float* weights, const int* var, float* out) {
  if (!m_eng || npos < 1 || npass < 1 || !dx || !dy || !weights || !out) return 1;
  if ((m_status_setup_GPU & _JMS_THRESHOLD_CALC) != _JMS_THRESHOLD_CALC) return 2;
  if (m_plasmonmc_flg) return 5;
  if (ApplyDescan() != 0) return 6;
  if (jms_hip_stage_scan(m_eng, npos, dx, dy, dz, npass, weights, var) != 0) return 101;
  if (jms_hip_run_staged(m_eng, 0, npos, 0) != 0) return 102;
  if (jms_hip_get_scan_results(m_eng, out) != 0) return 103;
  return 0;
}

int CJMultiSlice::GPUMultisliceScan(int npos, const float* dx, const float* dy, const float* dz, int npass,
                                    const float* weights, float* out) {
  if (npos < 1 || npass < 1) return 1;
  const size_t S = (size_t)m_nobjslc;
  std::vector<int> var((size_t)npos * npass * std::max<size_t>(1, S), 0);
  if (S > 0)
    for (size_t k = 0; k < (size_t)npos * npass; ++k) GetRandomVariantSequence(&var[k * S]);
  return GPUMultisliceScanVar(npos, dx, dy, dz, npass, weights, var.data(), out);
}
