/* jms_elasca.h -- host-side electron scattering factors and structure
 * slicing of the MI355X multislice library: what turns a structure into the
 * projected 2-D form factors that jms_hip_formfactor_gratings (jms_hip.h)
 * consumes.  Restated from the reference's CWEKOScat (src/wekoscat.cpp),
 * CAtomType (src/Atoms.cpp), CJEElaSca (src/JEElaSca.cpp) and CStructure
 * (src/Structure.cpp); each entry cites the function it replaces.
 *
 * Units as in the reference: spatial frequencies g, q [1/nm]; Biso [nm^2];
 * high tension [kV]; form factors [nm] (including 4 pi and the relativistic
 * correction); cell lengths [nm]; fractional coordinates.  Return codes:
 * 0 = success, >0 = error. */
#ifndef JMS_ELASCA_H
#define JMS_ELASCA_H

#ifdef __cplusplus
extern "C" {
#endif

/* Weickenmeier-Kohl parameters of atomic number z (0..98): a[2] [1/A^2 * A],
 * b[6] [A^2] -- CWEKOScat::getweko, wekoscat.cpp:175-185 */
int jms_weko_params(int z, double* a, double* b);
/* elastic WK form factor [A] at s = g/2 [1/A] -- CWEKOScat::weko :201-224 */
double jms_weko_f(const double* a, const double* b, double s);
/* exponential integral Ei(x), -60 <= x <= 60 (0 outside) -- CWEKOScat::ei
 * :227-256 (series; Abramowitz & Stegun 5.1.56 for x < -1) */
double jms_weko_ei(double x);
/* x exp(-x) Ei(x) for x > 20 from the table at 1/x = 0.005 i --
 * CWEKOScat::rih2 :282-290 */
double jms_weko_rih2(double x);
/* analytic absorptive integrals (2 pi k notation) -- CWEKOScat::ri1
 * :293-316, ri2 :318-356 */
double jms_weko_ri1(double bi, double bj, double g);
double jms_weko_ri2(double bi, double bj, double g, double u);
/* absorptive TDS integral of parameters (a, b) at g for rms displacement ul,
 * with the reference's Debye-Waller term exp(-ul g^2 / 2) --
 * CWEKOScat::wekoimag :358-379 */
double jms_weko_imag(double g, double ul, const double* a, const double* b);
/* atomic form factor (out[0] real, out[1] imaginary) [nm] of atomic number z
 * at g [1/nm], high tension ekv [kV]; dwflg: Debye-Waller damping with
 * dw = Biso [nm^2]; absflg: absorptive part (TDS with dwflg, else due to the
 * band-width limit with dw = aperture [1/nm]): 0 none, 1 as the reference
 * computes it, 2 with the thermal Debye-Waller term exp(-<u^2> g^2 / 2)
 * where the reference writes exp(-sqrt(<u^2>) g^2 / 2) (no counterpart in
 * the reference; the aperture case is the same as 1) -- CWEKOScat::wekosca
 * :529-559 (wekoscar :513-527 is its real part without absflg; the band-width
 * absorptive part integrates wekomugap :421-456 with wekoabsap :479-499) */
int jms_weko_form_factor(double g, double dw, int z, double ekv, int dwflg, int absflg, double* out);

/* radial form-factor table of one atom type: out[2*ik], out[2*ik+1] = form
 * factor at q = ik*sk, ik < nk.  dwf: damp with biso; otherwise bwlk is the
 * band-width limit [1/nm] used by the absorptive model 2.  absorb: 0 none,
 * 1 imaginary = abf * real (Hashimoto et al.), 2 computed (Hall & Hirsch)
 * as the reference computes it, 3 computed with the squared displacement
 * in the thermal Debye-Waller term (jms_weko_form_factor absflg 2; no
 * counterpart in the reference) --
 * CAtomType::CalculateRadialFormFactors, Atoms.cpp:260-320 */
int jms_radial_form_factors(int z, float biso, float ekv, int dwf, int absorb, float abf, long long nk,
                            float sk, float bwlk, float* out);
/* sampling of the radial tables for a grid: nk = 12 max(nx, ny, nz),
 * sk = max(n/a) / nk, bwlk = 2/3 of the smaller x/y Nyquist frequency --
 * CJEElaSca::CalculateFormFactors, JEElaSca.cpp:418-480 */
int jms_form_factor_sampling(int nx, int ny, int nz, float ax, float ay, float az, long long* nk, float* sk,
                             float* bwlk);
/* projected 2-D form factors ff2d[t][ny][nx] (complex64, FFT order) of ntype
 * radial tables (radial[t][2*nk]), linearly interpolated at |q|, times the
 * potential band-limit aperture when bwl_pot -- CJEElaSca::InitHelper
 * :529-607 (q^2, aperture) and PrepareScatteringFunctions :685-740 */
int jms_form_factors_2d(int nx, int ny, float ax, float ay, int bwl_pot, int ntype, long long nk, float sk,
                        const float* radial, float* ff2d);

/* atoms of one z-slice: the indices (idx, ascending) of the atoms whose
 * fractional z, wrapped to [0, 1), lies in [fz0, fz1), and their fractional
 * z in the slice, (fz - fz0) / (fz1 - fz0); returns the count (<0: error) --
 * CStructure::GetZSliceStructure, Structure.cpp:258-291 */
int jms_zslice(int nat, const float* fz, float fz0, float fz1, int* idx, float* fz_new);
/* suggested number of equidistant z-slices of a structure: the strongest
 * Fourier component above the first of a Gaussian-smeared (width dmin [nm])
 * z-density weighted by Z -- CStructure::SuggestEquidistantZSlices,
 * Structure.cpp:194-256 */
int jms_suggest_zslices(int nat, const int* z, const float* fz, float cz, float dmin);

#ifdef __cplusplus
}
#endif
#endif /* JMS_ELASCA_H */
