/* jms_specimen.h -- synthetic specimens for the benchmark configurations:
 * frozen-phonon (QEP) atom sites and phase gratings of the Gaussian projected
 * potential model.  Host code; no counterpart in the reference library, which
 * takes phase gratings from its caller (SetPhaseGratingData) -- the
 * displacement draw follows the reference's QEP model (JEElaSca.cpp:893-905:
 * per site, x then y, each ux * NormRand()).
 */
#ifndef JMS_SPECIMEN_H
#define JMS_SPECIMEN_H

#ifdef __cplusplus
extern "C" {
#endif

/* Sites of one slice type of an ncx x ncy block of cubic ABO3 perovskite
 * cells (lattice constant a [A]).  Slice type 0 (z = 0): A at (0,0), O at
 * (1/2,1/2); type 1 (z = a/2): B at (1/2,1/2), O at (1/2,0), O at (0,1/2)
 * (fractional coordinates).  Sites are listed cell row by cell row (cy outer,
 * cx inner), basis order within the cell.  With displace != 0 each site moves
 * by rms * N(0,1) in x, then in y, from a generator seeded with `seed`.
 * xy: 2 doubles per site [A]; z: atomic number per site.  Returns the number
 * of sites (5 per cell pair of slices: 2 for type 0, 3 for type 1). */
int jms_perovskite_sites(int ncx, int ncy, double a, int slice_type, int zA, int zB, int zO,
                         double rms, int seed, int displace, double* xy, int* z);

/* Phase grating t = exp(i * sig_int * V) of atoms with projected potential
 * V(r) = S / (2 pi s^2) exp(-|r - ra|^2 / (2 s^2)), S = 2.0 * Z^0.7 [V A^2],
 * on an nx x ny grid over ax x ay [A] (periodic, images within 8 s).
 * out: 2*nx*ny floats (complex64, x fastest). */
int jms_gaussian_phase_grating(int nx, int ny, double ax, double ay, int nat, const double* xy,
                               const int* z, double sigma, double sig_int, float* out);

#ifdef __cplusplus
}
#endif
#endif
