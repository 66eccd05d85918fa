/* jms_hip.h -- thin C-ABI between the host-side CJMultiSlice class and the
 * HIP kernels of the multislice slice loop (MI355X / gfx950).
 *
 * All functions return 0 on success and >0 on failure, with GPU-side failures
 * in the 1xx range (the reference's convention, JMultiSlice.cpp:499,566,762-800).
 * Plain pointers and sizes only; complex arrays are interleaved complex64
 * (re, im float pairs), row-major with x fastest (index x + y*nx), Fourier-space
 * arrays in FFT ("scrambled") order -- the reference's fcmplx layout
 * (fcomplex.h:27, JMultiSlice.cpp:1036-1041).
 *
 * An engine owns one HIP stream on the device that is current when it is
 * created.  Calls on one engine must not overlap; distinct engines are
 * independent.  Batch runs are enqueued on that stream and return at once;
 * the set-up and upload calls first wait for the work already on it, so a
 * new scan, probe or table never replaces data a queued launch still reads.
 */
#ifndef JMS_HIP_H
#define JMS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jms_hip_engine jms_hip_engine;

/* detection channel flags (JMultiSlice.h:141-150) */
#define JMS_HIP_DETECT_INTEGRATED 1
#define JMS_HIP_DETECT_IMAGE 2
#define JMS_HIP_DETECT_DIFFRACTION 4
#define JMS_HIP_DETECT_WAVEREAL 8
#define JMS_HIP_DETECT_WAVEFOURIER 16
#define JMS_HIP_DETECT_IMAGE_AVG 32
#define JMS_HIP_DETECT_DIFFR_AVG 64
#define JMS_HIP_DETECT_WAVER_AVG 128
#define JMS_HIP_DETECT_WAVEF_AVG 256
#define JMS_HIP_DETECT_ALL 511

/* engine lifetime ---------------------------------------------------------- */
int jms_hip_create(jms_hip_engine** out);
void jms_hip_destroy(jms_hip_engine* e);
/* 1 when nx (row length) and ny (column length) have compiled kernels:
 * 16 32 48 64 96 128 160 192 256 320 360 384 480 512 625 640 720 768 1024 2048 */
int jms_hip_grid_supported(int nx, int ny);

/* setup (replaces the device side of the reference setup chain,
 * JMultiSlice.h:76-90) --------------------------------------------------- */
/* grid size, super-cell size [nm] and electron wavelength [nm] */
int jms_hip_set_grid(jms_hip_engine* e, int nx, int ny, float a0x, float a0y, float wl);
/* phase-grating store: nslc structure slices with nvarslc[i] variants each,
 * gratings of npgx*npgy (must equal the grid) -- PhaseGratingSetup :447 */
int jms_hip_setup_gratings(jms_hip_engine* e, int npgx, int npgy, int nslc, const int* nvarslc);
/* upload nvar variants of structure slice islc -- SetPhaseGratingData :1440 */
int jms_hip_upload_gratings(jms_hip_engine* e, int islc, int nvar, const float* pgr);
/* generate nvar variants of structure slice islc on the device, in place of
 * an upload: phase gratings t = exp(i sig_int V) of the Gaussian
 * projected-potential model V(r) = S/(2 pi s^2) exp(-|r - ra|^2/(2 s^2)),
 * S = 2.0 Z^0.7 [V A^2], periodic over the grid cell, images within 8 s
 * (the host generator of jms_specimen.h, evaluated in the same order).
 * nat[v] atoms per variant, their positions xy (2 doubles per atom [A]) and
 * atomic numbers z concatenated over the variants; sigma [A]; sig_int
 * [rad / (V A)].  No counterpart in the reference library, which takes
 * gratings from its caller. */
int jms_hip_gaussian_gratings(jms_hip_engine* e, int islc, int nvar, const int* nat, const double* xy,
                              const int* z, double sigma, double sig_int);
/* generate nvar variants of structure slice islc on the device from
 * projected atomic form factors, in place of an upload -- the device path of
 * CJEElaSca::CalculatePhasegratingGPU (JEElaSca.cpp:978-1159, kernels of
 * scapot.cu:29-89) for neutral atoms:
 *   V(q) = sum_t ff_t(q) ap(q) sum_{atoms of t} occ exp(-2 pi i (x qx + y qy)),
 *   V(r) = IFT[V(q)] * POT_SCA / (a0x a0y cz),
 *   t(r) = exp(i PGR_SIG cz wl V(r))  (complex V: the imaginary part absorbs),
 * then with JMS_HIP_FF_BWL_PGR t = IFT[FT[t] * (|q| <= 2/3 kmax) / n].
 * ff2d: ntype 2-D form factors [nm] (2*nx*ny floats each, FFT order, the
 * reference's m_h_att_ff2d, including any Debye-Waller and absorptive
 * parts); per variant v: nat[v] atoms, positions xy (2 floats per atom [nm]
 * in the grid cell), type index and occupancy (occ NULL = 1), concatenated
 * over the variants; cz: slice thickness [nm]; flags: JMS_HIP_FF_BWL_POT
 * (potential band limit 1/2 - 1/2 tanh(30 (|q|/kmax - 0.9))),
 * JMS_HIP_FF_BWL_PGR (phase-grating band limit, the reference's default).
 * Atom displacements (QEP draws) are the caller's: positions are final. */
#define JMS_HIP_FF_BWL_POT 1
#define JMS_HIP_FF_BWL_PGR 2
int jms_hip_formfactor_gratings(jms_hip_engine* e, int islc, int nvar, int ntype, const float* ff2d,
                                const int* nat, const float* xy, const int* type, const float* occ, float cz,
                                int flags);
/* copy nvar variants of structure slice islc (2*nx*ny floats each) */
int jms_hip_get_gratings(jms_hip_engine* e, int islc, int nvar, float* out);
/* npro propagators, proidx[nslc] selects one per structure slice -- :542 */
int jms_hip_setup_propagators(jms_hip_engine* e, int npro, int nslc, const int* proidx);
int jms_hip_upload_propagator(jms_hip_engine* e, int ipro, const float* pro);
/* object slice sequence (structure slice index per object slice) -- :522 */
int jms_hip_setup_objects(jms_hip_engine* e, int nobjslc, const int* objslc);
/* detection: det_objslc[nobjslc+1] maps planes to detection indices
 * (-1 = none), ndet integrating detectors, flags = JMS_HIP_DETECT_* -- :668.
 * Codes: 5 an index >= ndetslc, 6 an index used by two planes (the
 * reference would add both planes' readouts into it). */
int jms_hip_setup_detection(jms_hip_engine* e, int ndetslc, const int* det_objslc, int ndet,
                            int flags);
/* detector function det[nx*ny]; with msklen>0 only the listed pixel indices
 * count (each listing once) -- SetDetectorData :1505 */
int jms_hip_upload_detector(jms_hip_engine* e, int idet, const float* det, int msklen,
                            const int* msk);
/* incident (probe) wave in Fourier space -- SetIncidentWave :1805 */
int jms_hip_set_probe(jms_hip_engine* e, const float* wav0);
/* probe in Fourier space computed on the device into the incident-wave slot:
 * aperture (ApertureFunctionA, JProbeGen.cpp:471-505) times exp(-i chi) of
 * the axial aberrations (AberrationFunction :508-571), unit intensity
 * (CalculateProbeWaveFourier :707-777).  prm: alpha [mrad], edge smoothness,
 * aperture centre x, y [mrad], asymmetry, asymmetry direction [deg], beam
 * tilt x, y [mrad]; abrr: nabrr (x, y) coefficient pairs [nm] in the
 * interface order (image shift, defocus, 2-fold astigmatism, coma, ...);
 * out: optional host copy (2*nx*ny floats) */
int jms_hip_probe(jms_hip_engine* e, const float* prm, int nabrr, const float* abrr, float* out);
/* constant phase plate [rad] (nx*ny floats), NULL = zero -- :1858-1932 */
int jms_hip_set_phaseplate(jms_hip_engine* e, const float* cppl);

/* batched STEM scan -------------------------------------------------------- */
/* Stage a scan on the device: npos positions (dx, dy, dz in nm; dz may be
 * NULL), npass passes per position with weights[npass], and the variant
 * sequences var[(p*npass + r)*nobjslc + j] (position-major, pass-minor).
 * Results are zeroed. */
int jms_hip_stage_scan(jms_hip_engine* e, int npos, const float* dx, const float* dy,
                       const float* dz, int npass, const float* weights, const int* var);
/* run positions [pos0, pos0+npos) of the staged scan from plane islc0 on
 * device-resident inputs (asynchronous; jms_hip_sync waits).  Each position
 * gets sum_r weights[r] * (readout of pass r) on its detection planes. */
int jms_hip_run_staged(jms_hip_engine* e, int pos0, int npos, int islc0);
int jms_hip_sync(jms_hip_engine* e);
/* Allocation-free batch (no counterpart in the reference; for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops).  Runs the slice loop from plane islc0 on nwave = npos * npass waves
 * of the set-up specimen, as one batch of jms_hip_run_staged does, with every
 * buffer in device memory that the caller owns: waves (nwave * nx * ny
 * complex64 of working storage, laid out as the staged batches keep their
 * waves), ppar (float4 per wave: dx, dy [nm],
 * dz * lambda / 2, 0), var (nobjslc ints per wave), weights (npass), res
 * (npos * ndetslc * ndet floats, overwritten).  Enqueued on the engine's
 * stream; jms_hip_sync waits.  The readout partial sums use a fixed 16 MiB
 * scratch array per GPU (nwave * ndetslc * ndet * column tiles floats at
 * most 4 Mi; larger batches are refused with code 5); calls must not overlap
 * on one GPU, since all share that GPU's scratch.  No hipMalloc, hipFree or
 * allocating library call. */
int jms_hip_run_batch_noalloc(jms_hip_engine* e, int islc0, int nwave, void* waves, const float* ppar,
                              const int* var, int npos, int npass, const float* weights, float* res);
/* Uneven batch in grouped launches (no counterpart in the reference, which
 * runs one probe at a time, JMultiSlice.cpp:4590): a group holds n engines
 * with the same grid and device, each set up for its own specimen, probe
 * and detectors and with its own staged scan.  jms_hip_group_run runs every
 * member's staged scan in full, as jms_hip_run_staged(e, 0, npos, 0) on each
 * would (the same results to float32 rounding: the grouped kernels are
 * instantiations of their own), but the waves of all members advance
 * together: one row-pass and one column-pass launch per slice step over the
 * waves of the members that still have slices (plus one column launch for
 * the members whose exit plane it is), driven by per-wave, per-member and
 * per-plane descriptor tables.  Enqueued on the stream of the member with
 * the most slices (the first such in the order given), after the work
 * already on every member's stream; every member's stream waits for the
 * group, so jms_hip_sync on any member waits for it.  Members with more
 * positions than one batch (jms_hip_set_batch) continue in further rounds.
 * Integrated detectors only, every supported grid.  create: 2 invalid list,
 * 3 an engine twice, 4 grids or devices differ (or no grid), 5 grid not
 * supported; run: 18 a member's grid or device changed since create, 19 a
 * member was destroyed (the group stays valid to destroy), 20 a member has a
 * diffraction de-scan.  Destroying a group waits for its work on every live
 * member's stream. */
typedef struct jms_hip_group jms_hip_group;
int jms_hip_group_create(jms_hip_group** out, jms_hip_engine* const* engines, int n);
int jms_hip_group_run(jms_hip_group* g);
const char* jms_hip_group_last_error(jms_hip_group* g);
void jms_hip_group_destroy(jms_hip_group* g);
/* integrated results of the staged scan: npos*ndetslc*ndet floats,
 * [pos][idetslc][idet] */
int jms_hip_get_scan_results(jms_hip_engine* e, float* out);
/* batch size in probe passes (waves resident at once, at most 65535); 0 =
 * automatic (about 2 GiB of waves, at most 8192) */
int jms_hip_set_batch(jms_hip_engine* e, int nwave_max);
/* 1: capture the per-batch launch sequence in a HIP graph and replay it */
int jms_hip_set_graph(jms_hip_engine* e, int use_graph);

/* single-wave path (CJMultiSlice::GPUMultislice) --------------------------- */
/* wave <- probe * exp(-i[2pi(kx dx + ky dy + k^2 dz wl/2) + cppl]), with a
 * cyclic beam-tilt shift (ibtx, ibty) pixels -- OffsetIncomingWave :1961 */
int jms_hip_wave_offset(jms_hip_engine* e, float dx, float dy, float dz, int ibtx, int ibty,
                        int use_cppl);
/* wave <- host data (Fourier space) -- SetIncomingWaveGPU :4468 */
int jms_hip_wave_set(jms_hip_engine* e, const float* wav);
/* one multislice run of the wave from plane islc0 with variant sequence
 * var[nobjslc]; readouts are accumulated with `weight`, after clearing the
 * channels when accumulate == 0 -- GPUMultislice :4513 */
int jms_hip_multislice(jms_hip_engine* e, int islc0, const int* var, float weight, int accumulate);
/* copy a result channel of the single-wave path: which = JMS_HIP_DETECT_*;
 * INTEGRATED: ndetslc*ndet floats; DIFFRACTION, IMAGE: ndetslc*nx*ny floats;
 * WAVEFOURIER, WAVEREAL: 2*ndetslc*nx*ny floats -- GetResult :2406.  IMAGE
 * and WAVEREAL hold |psi(r)|^2 and psi(r) of the unscaled inverse transform
 * of each detection plane (ReadoutImgDet_d :4369). */
int jms_hip_get_result(jms_hip_engine* e, int which, float* dst);
/* copy the current wave (2*nx*ny floats) */
int jms_hip_wave_get(jms_hip_engine* e, float* dst);
/* averaging channels of the single-wave path: they accumulate over runs
 * (jms_hip_multislice does not clear them) until reset; which =
 * JMS_HIP_DETECT_IMAGE_AVG / DIFFR_AVG (ndetslc*nx*ny floats, weight = summed
 * run weights of the image group) or WAVER_AVG / WAVEF_AVG (2*ndetslc*nx*ny
 * floats, weight of the wave group) -- GetAvgResult :2489 */
int jms_hip_get_avg_result(jms_hip_engine* e, int which, float* dst, float* weight);
/* diffraction de-scan [pixel] of all Fourier-space readouts (integrated
 * detectors, diffraction and Fourier-space wave channels): they see the wave
 * shifted cyclically, psi'(i - d) = psi(i) (DescanDifN, JMultiSlice.cpp:
 * 2184-2275); (0, 0) switches it off */
int jms_hip_set_descan(jms_hip_engine* e, int dnx, int dny);
/* summed run weights of the image and wave averaging groups -- :2567-2590 */
int jms_hip_get_averaging_weights(jms_hip_engine* e, float* img_weight, float* wave_weight);
/* group 0: image group (IMAGE_AVG, DIFFR_AVG), group 1: wave group (WAVER_AVG,
 * WAVEF_AVG); clears the channels and the weight -- ResetImageAveraging :2592,
 * ResetWaveAveraging :2629 */
int jms_hip_reset_averaging(jms_hip_engine* e, int group);
/* intensities of the averaged (elastic) waves: INTEGRATED (ndetslc*ndet
 * floats, from the averaged Fourier-space wave), DIFFRACTION or IMAGE
 * (ndetslc*nx*ny floats); weight = wave-group weight -- ReadoutDetAvg :2667 */
int jms_hip_readout_det_avg(jms_hip_engine* e, int which, float* dst, float* weight);

/* utilities ---------------------------------------------------------------- */
/* in-place 2-D FFT of nbatch host arrays (dir < 0: forward exp(-i), dir > 0:
 * backward exp(+i), unscaled) on the device -- JFFTCUDAcore::FT/IFT :299-336 */
int jms_hip_fft2d(int nx, int ny, int nbatch, float* data, int dir);
/* timing with HIP events on the engine stream: enable, then read the summed
 * duration [ms] and count per kind (0 = row pass, 1 = column pass,
 * 2 = combine: eager launches; 3 = one replay of a batch's HIP graph, which
 * runs the batch's whole launch sequence) */
int jms_hip_set_timing(jms_hip_engine* e, int enable);
int jms_hip_get_timing(jms_hip_engine* e, int kind, double* ms, long long* count);
/* last HIP error text of the engine (static storage) */
const char* jms_hip_last_error(jms_hip_engine* e);

#ifdef __cplusplus
}
#endif
#endif /* JMS_HIP_H */
