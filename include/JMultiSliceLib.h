/* JMultiSliceLib.h -- flat C-ABI of the MI355X multislice library
 * (libjmultislice_hip.so).  The entry points carry the names and argument
 * meaning of the reference's exported interface (src/JMultiSliceLib.h:111-380),
 * each cited below, and forward to one process-global CJMultiSlice object, as
 * the reference does (src/JMultiSliceLib.h:106).  Linux default visibility,
 * no __stdcall.  Return codes: 0 = success, >0 = error (1xx: GPU side).
 *
 * Deliberate differences from the reference exports:
 *  - C types only: const char* instead of std::string, pointers instead of
 *    C++ references, int for bool, no default arguments.
 *  - CalculateRingDetector passes theta0y through (the reference wrapper passes
 *    theta0x twice, JMultiSliceLib.cpp:254).
 *  - OffsetIncomingWave takes bcppl and SetIncidentWave takes bTranspose (the
 *    reference wrappers drop them, JMultiSliceLib.h:246,326).
 *  - whichcode _JMS_CODE_CPU (1) is rejected: this library has no CPU path.
 *    The CPU-only entries (SetIncomingWaveCPU, CPUMultislice, ClearDetMem_h)
 *    are exported and return 1.
 *  - SetPlasmonMC is exported; activating it makes GPUMultislice fail
 *    (plasmon Monte-Carlo is not part of this library).
 *  - GPUMultisliceScan / GPUMultisliceScanVar are additive batched entries.
 */
#ifndef JMULTISLICELIB_HIP_H
#define JMULTISLICELIB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JMS_API __attribute__((visibility("default")))

/* Probe parameters (fields of CJProbeParams, JProbeGen.h:84-122).  Angles in
 * mrad except m_alpha_adir [deg]; m_wl is set by the library.  Aberrations:
 * m_abrr_num coefficient pairs (x, y) [nm] in m_abrr_coeff, in the order of
 * the reference's aberration table (JProbeGen.cpp:508-571: order m = 1..8,
 * n <= m, m+n even; up to 24 terms). */
typedef struct CJProbeParams {
  float m_wl;
  float m_alpha;
  float m_alpha_rs;
  float m_alpha_x0;
  float m_alpha_y0;
  float m_alpha_asym;
  float m_alpha_adir;
  float m_btx;
  float m_bty;
  float m_source_width;
  int m_source_shape;
  float m_fspread_width;
  float m_fspread_kernel_width;
  int m_fspread_kernel_samples;
  float* m_abrr_coeff;
  int m_abrr_num;
} CJProbeParams;

/* opaque RNG object (class CRng in jms/rng.h) */
#ifdef __cplusplus
class CRng;
#else
typedef struct CRng CRng;
#endif

/* basic interface */
JMS_API float GetJMSVersion(void);                                  /* JMultiSliceLib.h:111 */
JMS_API int SetJMSDebugLevel(int debuglevel);                       /* :114 */
JMS_API void SetRng(CRng* prng);                                    /* :117 */
JMS_API CRng* GetRng(void);                                         /* :120 */
JMS_API void SeedRngEx(int rngseed);                                /* :123 */

/* physics interface */
JMS_API float SetHighTension(float ht);                             /* :129 */
JMS_API float GetHighTension(void);                                 /* :132 */
JMS_API float GetWaveLength(void);                                  /* :135 */
JMS_API void SetSupercellSize(float* a0);                           /* :138 */
JMS_API void SetSupercellSizeABC(float a, float b, float c);        /* :141 */
JMS_API void SetSliceThickness(int islc, float fthickness);         /* :146 */
JMS_API void DiffractionDescan(int bActivate);                      /* :149 */
JMS_API void SetPlasmonMC(int bActivate, float fq_e, float fq_c, float fmfp,
                          unsigned int nexmax);                       /* :152 */
JMS_API void SetDiffractionDescanN(int whichcode, int ndescanx, int ndescany,
                                   int iThread);                      /* :160 */
JMS_API void SetDiffractionDescan(int whichcode, float descanx, float descany,
                                  int iThread);                       /* :168 */
JMS_API void SetDiffractionDescanMRad(int whichcode, float descanx, float descany,
                                      int iThread);                   /* :176 */

/* numerics interface */
JMS_API void SetGridSize(int nx, int ny);                           /* :183 */
JMS_API void GetGridSize(int* nx, int* ny);                         /* :186 */
JMS_API int GetDetetionSliceNum(void);                              /* :189 */
JMS_API int GetDetNum(void);                                        /* :192 */
JMS_API int GetCPUNum(void);                                        /* :195 */
JMS_API int GetGPUNum(void);                                        /* :198 */
JMS_API int GetGPUName(int idev, char* name);                       /* :201 */
JMS_API int GetCurrentGPU(void);                                    /* :204 */
JMS_API int SetCurrentGPU(int idev);                                /* :207 */
JMS_API int GetGPUStats(int idev, int* iCMajor, int* iCMinor, int* iMaxThread,
                        int64_t* memtotal, int64_t* memfree);       /* :210 */
JMS_API int GetUnscrambleHash(unsigned int* phash);                 /* :220 */
JMS_API float GetImageAveragingWeight(int whichcode, int iThread);  /* :223 */
JMS_API float GetWaveAveragingWeight(int whichcode, int iThread);   /* :226 */

/* setup interface */
JMS_API int PhaseGratingSetup(int whichcode, int nx, int ny, int nslc, int nvarmax,
                              int* nslcvar);                        /* :231 */
JMS_API int ObjectSliceSetup(int nobjslc, int* objslc);             /* :234 */
JMS_API int PropagatorSetup(int whichcode, int npro, int* proidx);  /* :237 */
JMS_API int DetectorSetup(int whichcode, int ndetper, int ndetint, int imagedet,
                          int nthreads_CPU);                        /* :240 */
JMS_API int InitCore(int whichcode, int nCPUthreads);               /* :243 */
JMS_API int SetIncidentWave(int whichcode, float* wav, int bTranspose); /* :246 */
JMS_API int SetIncidentWavePhaseplate(int whichcode, float* ppl, int bTranspose); /* :249 */
JMS_API int ZeroIncidentWavePhaseplate(int whichcode);              /* :252 */
JMS_API int SetSlicePhaseGratings(int whichcode, int islc, int nvar, float* pgr); /* :255 */
JMS_API int SetPropagatorData(int whichcode, int ipro, float* pro); /* :258 */
JMS_API int SetDetectorData(int whichcode, int idet, float* det, int msklen, int* msk); /* :261 */
JMS_API int SetGPUPgrLoading(int npgrload);                         /* :266 */
JMS_API int ResetImageAveraging(int whichcode, int iThread);        /* :269 */
JMS_API int ResetWaveAveraging(int whichcode, int iThread);         /* :272 */

/* calculation interface */
JMS_API int CalculateProbeWaveFourier(CJProbeParams* prm, float* wav);   /* :280 */
JMS_API int SetIncomingWaveGPU(float* wav, int bTranspose);              /* :286 */
JMS_API int SetIncomingWaveCPU(float* wav, int bTranspose, int iThread); /* :293, returns 1 */
JMS_API int CalculatePropagator(float fthick, float otx, float oty, float* pro,
                                int ntype);                              /* :303 */
JMS_API int CalculateRingDetector(float beta0, float beta1, float phi0, float phi1,
                                  float theta0x, float theta0y, const char* sdsprofile,
                                  float* det, int* msklen, int* msk, int kmom1); /* :318 */
JMS_API int OffsetIncomingWave(int whichcode, float dx, float dy, float dz, int iThread,
                               int ibtx, int ibty, int bcppl);           /* :326 */
JMS_API int CPUMultislice(int islc0, int accmode, float weight, int iThread); /* :333, returns 1 */
JMS_API int GPUMultislice(int islc0, int accmode, float weight);         /* :339 */
JMS_API int GetResult(int whichcode, int whichresult, float* dst, int iThread); /* :346 */
JMS_API int GetAvgResult(int whichcode, int whichresult, float* dst, float* wgt,
                         int iThread);                                  /* :354 */
JMS_API int ClearDetMem_h(int iThread);                                  /* :357, returns 1 */
JMS_API int ClearDetMem_d(void);                                         /* :363 */

/* batched STEM scan (no counterpart in the reference; additive).  For each
 * position p and pass r: OffsetIncomingWave(GPU, dx[p], dy[p], dz ? dz[p] : 0)
 * then GPUMultislice(0, r ? INTEGRATE : NONE, weights[r]), all passes in
 * batched launches; out[(p*ndetslc + idetslc)*ndet + idet].  Variant
 * sequences come from the library RNG in that serial order (Scan) or from
 * var[(p*npass + r)*nobjslc + j] (ScanVar). */
JMS_API int GPUMultisliceScan(int npos, const float* dx, const float* dy, const float* dz,
                              int npass, const float* weights, float* out);
JMS_API int GPUMultisliceScanVar(int npos, const float* dx, const float* dy, const float* dz,
                                 int npass, const float* weights, const int* var, float* out);
/* random variant sequence of the current object-slice setup
 * (CJMultiSlice::GetRandomVariantSequence, JMultiSlice.cpp:1571-1676) */
JMS_API int GetRandomVariantSequence(int* out);

/* cleanup interface */
JMS_API int Cleanup(void);                                               /* :375 */
JMS_API void FreeLibMem(void);                                           /* :380 */

#ifdef __cplusplus
}
#endif
#endif /* JMULTISLICELIB_HIP_H */
