// jms/JMultiSlice.h -- CJMultiSlice, the host-side class of the MI355X
// multislice library.  It keeps the interface of the reference class
// (src/JMultiSlice.h:187-937) for the GPU multislice path: grid, phase-grating,
// propagator, detector and incident-wave setup, the GPU multislice entry and
// the detector readout.  Device work goes through the thin C-ABI of
// include/jms_hip.h (HIP kernels for gfx950).
//
// Differences from the reference, by design:
//  - The CPU code path (_JMS_CODE_CPU: CPUMultislice, SetIncomingWaveCPU,
//    ClearDetMem_h and the host multislice buffers) is not part of this
//    library; its entry points exist and return an error (1).
//  - Plasmon Monte-Carlo and sub-frame phase gratings are not implemented;
//    setups that need them fail with an error.
//  - The detection channels (integrated, image, diffraction, real- and
//    Fourier-space waves and their averages) are filled by the single-wave
//    path (GPUMultislice); the batched scan fills the integrated channel.
//  - CalculateProbeWaveFourier takes the axial aberrations of the
//    reference's probe generator (24 terms, m = 1..8, n <= m, m+n even);
//    partial spatial and temporal coherence envelopes are not applied.
//  - Additive batched entry GPUMultisliceScan runs many positions and passes
//    in one batched schedule; a batch of one gives the numbers of
//    OffsetIncomingWave + GPUMultislice.
#pragma once
#include <complex>
#include <string>
#include <vector>

#include "jms/rng.h"
#include "jms_hip.h"

typedef std::complex<float> fcmplx;  // fcomplex.h:27

// code IDs, status IDs and detection flags (JMultiSlice.h:126-157)
#define _JMS_CODE_CPU 1
#define _JMS_CODE_GPU 2
#define _JMS_STATUS_NONE 0
#define _JMS_STATUS_PGR 1
#define _JMS_STATUS_OBJ 2
#define _JMS_STATUS_PRO 4
#define _JMS_STATUS_DET 8
#define _JMS_STATUS_CORE 16
#define _JMS_THRESHOLD_CORE 15
#define _JMS_THRESHOLD_CALC 31
#define _JMS_ACCMODE_NONE 0
#define _JMS_ACCMODE_INTEGRATE 1
#define _JMS_DETECT_NONE 0
#define _JMS_DETECT_INTEGRATED 1
#define _JMS_DETECT_IMAGE 2
#define _JMS_DETECT_DIFFRACTION 4
#define _JMS_DETECT_WAVEREAL 8
#define _JMS_DETECT_WAVEFOURIER 16
#define _JMS_DETECT_IMAGE_AVG 32
#define _JMS_DETECT_DIFFR_AVG 64
#define _JMS_DETECT_WAVER_AVG 128
#define _JMS_DETECT_WAVEF_AVG 256
constexpr double _JMS_RELAPERTURE = 0.666666666666667;  // JMultiSlice.h:157

#include "JMultiSliceLib.h"  // CJProbeParams (standard-layout, shared with the C-ABI)

class CJMultiSlice {
 public:
  CJMultiSlice();
  ~CJMultiSlice();

  // library handling
  int SetDebugLevel(int dbg);
  int Cleanup(void);
  void FreeLibMem(void);

  // setup
  float SetHighTension(float ht);
  void SetSupercellSize(float* a0);
  void SetSupercellSize(float a, float b, float c);
  void SetGridSize(int nx, int ny);
  void SetSliceThickness(int islc, float fthickness);
  int SetDetectionPlanes(int ndetper, int nobjslc, int* det_objslc = nullptr);
  int SetDetectionPlanes(std::vector<int> slc_det, int nobjslc, int* det_objslc = nullptr);
  void SetPlasmonMC(bool use, float q_e, float q_c, float mfp, unsigned int nexmax);
  void DiffractionDescan(bool bActivate);
  // de-scan of the Fourier-space readouts (JMultiSlice.h:471-495): in pixels,
  // in 1/nm (rounded to pixels with the cell size) or in mrad
  void SetDiffractionDescanN(int whichcode, int ndescanx, int ndescany, int iThread = 0);
  void SetDiffractionDescan(int whichcode, float descanx, float descany, int iThread = 0);
  void SetDiffractionDescanMRad(int whichcode, float descanx, float descany, int iThread = 0);

  int CalculateProbeWaveFourier(CJProbeParams* prm, fcmplx* wav);
  int CalculatePropagator(float fthick, float otx, float oty, fcmplx* pro, int ntype = 0);
  int CalculateRingDetector(float beta0, float beta1, float phi0, float phi1, float theta0x,
                            float theta0y, std::string sdsprofile, float* det, int& msklen,
                            int* msk = nullptr, int kmom1 = 0);

  int PhaseGratingSetup(int whichcode, int nx, int ny, int nslc, int nvarmax, int* nslcvar);
  int ObjectSliceSetup(int nobjslc, int* objslc);
  int PropagatorSetup(int whichcode, int npro, int* proidx);
  int DetectorSetup(int whichcode, int ndetper, int ndetint, int imagedet, int nthreads_CPU = 1);
  int DetectorSetup(int whichcode, std::vector<int> slc_det, int ndetint, int imagedet,
                    int nthreads_CPU = 1);
  int SetPhaseGratingData(int whichcode, int islc, int nvar, fcmplx* pgr);
  int SetPropagatorData(int whichcode, int ipro, fcmplx* pro);
  int SetDetectorData(int whichcode, int idet, float* det, int msklen = 0, int* msk = nullptr);
  int InitCore(int whichcode, int nCPUthreads = 1);
  int SetIncidentWave(int whichcode, fcmplx* wav, bool bTranspose = false);
  int SetIncidentWavePhaseplate(int whichcode, float* ppl, bool bTranspose = false);
  int ZeroIncidentWavePhaseplate(int whichcode);
  int GetUnscrambleHash(unsigned int* phash);

  // getters
  int GetDetetionSliceNum(void);
  int GetDetNum(void);
  float GetHighTension(void);
  float GetWaveLength(void);
  void GetGridSize(int& nx, int& ny);
  void GetSupercellSize(float* a0);
  void GetSupercellSize(float& a, float& b, float& c);
  float GetSliceThickness(int islc);

  // device interface
  int GetCPUNum(void);
  int GetGPUNum(void);
  int GetGPUName(int idev, char* name);
  int GetCurrentGPU(void);
  int SetCurrentGPU(int idev);
  int GetGPUStats(int idev, int& iCMajor, int& iCMinor, int& iMaxThread, int64_t& memtotal,
                  int64_t& memfree);
  int GetGPUMemInfo(size_t& memtotal, size_t& memfree);
  int SetGPUPgrLoading(int npgrload = 0);

  // random numbers (variant sequences)
  void SetRng(CRng* prng = nullptr);
  CRng* GetRng(void);
  void SeedRngEx(int nseed = 0);

  // detector memory
  int ClearDetMem_d(void);
  int ClearDetMem_h(int iThread);  // CPU path: returns 1 (JMultiSlice.cpp:3128)

  // multislice
  int OffsetIncomingWave(int whichcode, float dx, float dy, float dz, int iThread = 0, int ibtx = 0,
                         int ibty = 0, bool bcppl = true);
  int SetIncomingWaveGPU(fcmplx* wav, bool bTranspose = false);
  // CPU path: returns 1 (JMultiSlice.cpp:3541, :3589)
  int SetIncomingWaveCPU(fcmplx* wav, bool bTranspose = false, int iThread = 0);
  int CPUMultislice(int islc0, int accmode, float weight = 1.0f, int iThread = 0);
  int GPUMultislice(int islc0, int accmode, float weight = 1.0f);
  int GetResult(int whichcode, int whichresult, float* dst, int iThread = 0);
  // averaging channels (JMultiSlice.h:903-925)
  int GetAvgResult(int whichcode, int whichresult, float* dst, float& weight, int iThread = 0);
  float GetImageAveragingWeight(int whichcode, int iThread = 0);
  float GetWaveAveragingWeight(int whichcode, int iThread = 0);
  int ResetImageAveraging(int whichcode, int iThread = 0);
  int ResetWaveAveraging(int whichcode, int iThread = 0);
  // intensities of the averaged (elastic) waves (JMultiSlice.h:927-936)
  int ReadoutDetAvg(int whichcode, int whichresult, float* dst, float& weight, int iThread = 0);

  // batched STEM scan (additive entry): for every position p and pass r,
  // the equivalent of OffsetIncomingWave(GPU, dx[p], dy[p], dz[p]) +
  // GPUMultislice(0, r ? INTEGRATE : NONE, weights[r]); variant sequences are
  // drawn from the current RNG in that serial order (position-major).
  // out: npos * ndetslc * ndet floats, [p][idetslc][idet].
  int GPUMultisliceScan(int npos, const float* dx, const float* dy, const float* dz, int npass,
                        const float* weights, float* out);
  // same with caller-provided variant sequences var[(p*npass + r)*nobjslc + j]
  int GPUMultisliceScanVar(int npos, const float* dx, const float* dy, const float* dz, int npass,
                           const float* weights, const int* var, float* out);
  // random variant sequence for the object slices (JMultiSlice.cpp:1571-1676)
  int GetRandomVariantSequence(int* out);
  // engine (for batch tuning, timing); nullptr before InitCore
  jms_hip_engine* Engine(void) { return m_eng; }

 protected:
  int m_dbg = 0;
  CRng m_lrng;
  CRng* m_prng = nullptr;
  float m_ht = 0.f, m_wl = 0.f;
  float m_a0[3] = {0.f, 0.f, 0.f};
  int m_nscslc = 0, m_nslcvar = 0, m_nobjslc = 0;
  int m_nscx = 0, m_nscy = 0, m_npgx = 0, m_npgy = 0;
  int m_npro = 0, m_nprotype = 0;
  int m_ndet = 0, m_ndetper = 0, m_imagedet = 0, m_ndetslc = 0;
  int m_status_setup_GPU = 0;
  int m_plasmonmc_flg = 0, m_dif_descan_flg = 0;
  int m_d_dif_ndescanx = 0, m_d_dif_ndescany = 0;  // de-scan [pixel]
  int ApplyDescan();  // hands the active de-scan to the engine
  std::vector<int> m_objslc, m_det_objslc, m_nvarslc, m_proidx;
  std::vector<float> m_slcthick;
  std::vector<float> m_d_det_int;  // integrated GPU results live on the host (JMultiSlice.h:349-352)
  bool m_wave_ready = false;
  jms_hip_engine* m_eng = nullptr;

  // creates the engine on first use and (re)applies grid, cell and wavelength
  int EnsureEngine(void);
  // detection channels and device detector setup shared by both DetectorSetup overloads
  int DetectorChannels(int whichcode, int ndetint, int imagedet);
};
