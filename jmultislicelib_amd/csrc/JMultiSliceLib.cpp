// JMultiSliceLib.cpp -- flat C-ABI (include/JMultiSliceLib.h) forwarding to one
// process-global CJMultiSlice, as the reference's JMultiSliceLib.cpp does.
#include "JMultiSliceLib.h"

#include <string>

#include "jms/JMultiSlice.h"

namespace {
CJMultiSlice& jms_obj() {
  static CJMultiSlice obj;
  return obj;
}
}  // namespace

extern "C" {

float GetJMSVersion(void) { return 0.46f; }
int SetJMSDebugLevel(int d) { return jms_obj().SetDebugLevel(d); }
void SetRng(CRng* prng) { jms_obj().SetRng(prng); }
CRng* GetRng(void) { return jms_obj().GetRng(); }
void SeedRngEx(int s) { jms_obj().SeedRngEx(s); }

float SetHighTension(float ht) { return jms_obj().SetHighTension(ht); }
float GetHighTension(void) { return jms_obj().GetHighTension(); }
float GetWaveLength(void) { return jms_obj().GetWaveLength(); }
void SetSupercellSize(float* a0) { jms_obj().SetSupercellSize(a0); }
void SetSupercellSizeABC(float a, float b, float c) { jms_obj().SetSupercellSize(a, b, c); }
void SetSliceThickness(int islc, float t) { jms_obj().SetSliceThickness(islc, t); }
void DiffractionDescan(int bActivate) { jms_obj().DiffractionDescan(bActivate != 0); }
void SetPlasmonMC(int bActivate, float fq_e, float fq_c, float fmfp, unsigned int nexmax) {
  jms_obj().SetPlasmonMC(bActivate != 0, fq_e, fq_c, fmfp, nexmax);
}
void SetDiffractionDescanN(int whichcode, int ndescanx, int ndescany, int iThread) {
  jms_obj().SetDiffractionDescanN(whichcode, ndescanx, ndescany, iThread);
}
void SetDiffractionDescan(int whichcode, float descanx, float descany, int iThread) {
  jms_obj().SetDiffractionDescan(whichcode, descanx, descany, iThread);
}
void SetDiffractionDescanMRad(int whichcode, float descanx, float descany, int iThread) {
  jms_obj().SetDiffractionDescanMRad(whichcode, descanx, descany, iThread);
}

void SetGridSize(int nx, int ny) { jms_obj().SetGridSize(nx, ny); }
void GetGridSize(int* nx, int* ny) {
  int a = 0, b = 0;
  jms_obj().GetGridSize(a, b);
  if (nx) *nx = a;
  if (ny) *ny = b;
}
int GetDetetionSliceNum(void) { return jms_obj().GetDetetionSliceNum(); }
int GetDetNum(void) { return jms_obj().GetDetNum(); }
int GetCPUNum(void) { return jms_obj().GetCPUNum(); }
int GetGPUNum(void) { return jms_obj().GetGPUNum(); }
int GetGPUName(int idev, char* name) { return jms_obj().GetGPUName(idev, name); }
int GetCurrentGPU(void) { return jms_obj().GetCurrentGPU(); }
int SetCurrentGPU(int idev) { return jms_obj().SetCurrentGPU(idev); }
int GetGPUStats(int idev, int* iCMajor, int* iCMinor, int* iMaxThread, int64_t* memtotal, int64_t* memfree) {
  int a = 0, b = 0, c = 0;
  int64_t t = 0, f = 0;
  int rc = jms_obj().GetGPUStats(idev, a, b, c, t, f);
  if (iCMajor) *iCMajor = a;
  if (iCMinor) *iCMinor = b;
  if (iMaxThread) *iMaxThread = c;
  if (memtotal) *memtotal = t;
  if (memfree) *memfree = f;
  return rc;
}
int GetUnscrambleHash(unsigned int* phash) { return jms_obj().GetUnscrambleHash(phash); }
float GetImageAveragingWeight(int whichcode, int iThread) {
  return jms_obj().GetImageAveragingWeight(whichcode, iThread);
}
float GetWaveAveragingWeight(int whichcode, int iThread) {
  return jms_obj().GetWaveAveragingWeight(whichcode, iThread);
}

int PhaseGratingSetup(int whichcode, int nx, int ny, int nslc, int nvarmax, int* nslcvar) {
  return jms_obj().PhaseGratingSetup(whichcode, nx, ny, nslc, nvarmax, nslcvar);
}
int ObjectSliceSetup(int nobjslc, int* objslc) { return jms_obj().ObjectSliceSetup(nobjslc, objslc); }
int PropagatorSetup(int whichcode, int npro, int* proidx) { return jms_obj().PropagatorSetup(whichcode, npro, proidx); }
int DetectorSetup(int whichcode, int ndetper, int ndetint, int imagedet, int nthreads_CPU) {
  return jms_obj().DetectorSetup(whichcode, ndetper, ndetint, imagedet, nthreads_CPU);
}
int InitCore(int whichcode, int nCPUthreads) { return jms_obj().InitCore(whichcode, nCPUthreads); }
int SetIncidentWave(int whichcode, float* wav, int bTranspose) {
  return jms_obj().SetIncidentWave(whichcode, (fcmplx*)wav, bTranspose != 0);
}
int SetIncidentWavePhaseplate(int whichcode, float* ppl, int bTranspose) {
  return jms_obj().SetIncidentWavePhaseplate(whichcode, ppl, bTranspose != 0);
}
int ZeroIncidentWavePhaseplate(int whichcode) { return jms_obj().ZeroIncidentWavePhaseplate(whichcode); }
int SetSlicePhaseGratings(int whichcode, int islc, int nvar, float* pgr) {
  return jms_obj().SetPhaseGratingData(whichcode, islc, nvar, (fcmplx*)pgr);
}
int SetPropagatorData(int whichcode, int ipro, float* pro) {
  return jms_obj().SetPropagatorData(whichcode, ipro, (fcmplx*)pro);
}
int SetDetectorData(int whichcode, int idet, float* det, int msklen, int* msk) {
  return jms_obj().SetDetectorData(whichcode, idet, det, msklen, msk);
}
int SetGPUPgrLoading(int npgrload) { return jms_obj().SetGPUPgrLoading(npgrload); }
int ResetImageAveraging(int whichcode, int iThread) { return jms_obj().ResetImageAveraging(whichcode, iThread); }
int ResetWaveAveraging(int whichcode, int iThread) { return jms_obj().ResetWaveAveraging(whichcode, iThread); }

int CalculateProbeWaveFourier(CJProbeParams* prm, float* wav) {
  return jms_obj().CalculateProbeWaveFourier(prm, (fcmplx*)wav);
}
int SetIncomingWaveCPU(float* wav, int bTranspose, int iThread) {
  return jms_obj().SetIncomingWaveCPU((fcmplx*)wav, bTranspose != 0, iThread);
}
int SetIncomingWaveGPU(float* wav, int bTranspose) {
  return jms_obj().SetIncomingWaveGPU((fcmplx*)wav, bTranspose != 0);
}
int CalculatePropagator(float fthick, float otx, float oty, float* pro, int ntype) {
  return jms_obj().CalculatePropagator(fthick, otx, oty, (fcmplx*)pro, ntype);
}
int CalculateRingDetector(float beta0, float beta1, float phi0, float phi1, float theta0x, float theta0y,
                          const char* sdsprofile, float* det, int* msklen, int* msk, int kmom1) {
  int ml = 0;
  int rc = jms_obj().CalculateRingDetector(beta0, beta1, phi0, phi1, theta0x, theta0y,
                                           std::string(sdsprofile ? sdsprofile : ""), det, ml, msk, kmom1);
  if (msklen) *msklen = ml;
  return rc;
}
int OffsetIncomingWave(int whichcode, float dx, float dy, float dz, int iThread, int ibtx, int ibty, int bcppl) {
  return jms_obj().OffsetIncomingWave(whichcode, dx, dy, dz, iThread, ibtx, ibty, bcppl != 0);
}
int CPUMultislice(int islc0, int accmode, float weight, int iThread) {
  return jms_obj().CPUMultislice(islc0, accmode, weight, iThread);
}
int GPUMultislice(int islc0, int accmode, float weight) { return jms_obj().GPUMultislice(islc0, accmode, weight); }
int GetAvgResult(int whichcode, int whichresult, float* dst, float* wgt, int iThread) {
  float w = 0.f;
  const int rc = jms_obj().GetAvgResult(whichcode, whichresult, dst, w, iThread);
  if (wgt) *wgt = w;
  return rc;
}

int GetResult(int whichcode, int whichresult, float* dst, int iThread) {
  return jms_obj().GetResult(whichcode, whichresult, dst, iThread);
}
int ClearDetMem_h(int iThread) { return jms_obj().ClearDetMem_h(iThread); }
int ClearDetMem_d(void) { return jms_obj().ClearDetMem_d(); }

int GPUMultisliceScan(int npos, const float* dx, const float* dy, const float* dz, int npass,
                      const float* weights, float* out) {
  return jms_obj().GPUMultisliceScan(npos, dx, dy, dz, npass, weights, out);
}
int GPUMultisliceScanVar(int npos, const float* dx, const float* dy, const float* dz, int npass,
                         const float* weights, const int* var, float* out) {
  return jms_obj().GPUMultisliceScanVar(npos, dx, dy, dz, npass, weights, var, out);
}
int GetRandomVariantSequence(int* out) { return jms_obj().GetRandomVariantSequence(out); }

int Cleanup(void) { return jms_obj().Cleanup(); }
void FreeLibMem(void) { jms_obj().FreeLibMem(); }

}  // extern "C"
