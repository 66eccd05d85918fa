// jms_specimen.cpp -- synthetic perovskite specimens for the benchmark
// configurations (see include/jms_specimen.h).
#include "jms_specimen.h"

#include <cmath>
#include <vector>

#include "jms/rng.h"

extern "C" int jms_perovskite_sites(int ncx, int ncy, double a, int slice_type, int zA, int zB, int zO,
                                    double rms, int seed, int displace, double* xy, int* z) {
  if (ncx < 1 || ncy < 1 || !xy || !z) return 0;
  struct Basis { double fx, fy; int which; };  // which: 0 A, 1 B, 2 O
  static const Basis type0[2] = {{0.0, 0.0, 0}, {0.5, 0.5, 2}};
  static const Basis type1[3] = {{0.5, 0.5, 1}, {0.5, 0.0, 2}, {0.0, 0.5, 2}};
  const Basis* basis = slice_type == 0 ? type0 : type1;
  const int nb = slice_type == 0 ? 2 : 3;
  const int zz[3] = {zA, zB, zO};
  CRng rng(seed);
  int k = 0;
  for (int cy = 0; cy < ncy; ++cy) {
    for (int cx = 0; cx < ncx; ++cx) {
      for (int b = 0; b < nb; ++b, ++k) {
        double x = (cx + basis[b].fx) * a;
        double y = (cy + basis[b].fy) * a;
        if (displace) {  // QEP site displacement, x then y (JEElaSca.cpp:896-899)
          x += rms * rng.normrand(0.0, 1.0);
          y += rms * rng.normrand(0.0, 1.0);
        }
        xy[2 * k] = x;
        xy[2 * k + 1] = y;
        z[k] = zz[basis[b].which];
      }
    }
  }
  return k;
}

extern "C" int jms_gaussian_phase_grating(int nx, int ny, double ax, double ay, int nat, const double* xy,
                                          const int* z, double sigma, double sig_int, float* out) {
  if (nx < 1 || ny < 1 || !out || (nat > 0 && (!xy || !z))) return 1;
  const double PI = 3.14159265358979323846264338327950288;
  const double hx = ax / nx, hy = ay / ny;
  const double rc = 8.0 * sigma, rc2 = rc * rc;
  const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
  std::vector<double> pot((size_t)nx * ny, 0.0);
  std::vector<double> gx, gy;
  for (int at = 0; at < nat; ++at) {
    const double amp = 2.0 * std::pow((double)z[at], 0.7) / (2.0 * PI * sigma * sigma);
    double xa = std::fmod(xy[2 * at], ax);
    if (xa < 0) xa += ax;
    double ya = std::fmod(xy[2 * at + 1], ay);
    if (ya < 0) ya += ay;
    const int i0 = (int)std::floor((xa - rc) / hx), i1 = (int)std::ceil((xa + rc) / hx);
    const int j0 = (int)std::floor((ya - rc) / hy), j1 = (int)std::ceil((ya + rc) / hy);
    for (int j = j0; j <= j1; ++j) {
      const double dy = j * hy - ya;
      const double dy2 = dy * dy;
      if (dy2 > rc2) continue;
      const int jw = ((j % ny) + ny) % ny;
      double* row = &pot[(size_t)jw * nx];
      for (int i = i0; i <= i1; ++i) {
        const double dx = i * hx - xa;
        const double r2 = dx * dx + dy2;
        if (r2 > rc2) continue;
        row[((i % nx) + nx) % nx] += amp * std::exp(-r2 * inv2s2);
      }
    }
  }
  for (size_t p = 0; p < pot.size(); ++p) {
    const double ph = sig_int * pot[p];
    out[2 * p] = (float)std::cos(ph);
    out[2 * p + 1] = (float)std::sin(ph);
  }
  return 0;
}
