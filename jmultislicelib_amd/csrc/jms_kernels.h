// jms_kernels.h -- host-side launchers of the multislice kernels (jms_kernels.hip)
#pragma once
#include <hip/hip_runtime.h>

namespace jms {

// Arguments of one column pass (see k_col_pass in jms_kernels.hip).
struct ColArgsHost {
  void* wave = nullptr;          // batch of waves (float2), wave_stride items apart
  int nx = 0;
  long long wave_stride = 0;
  const void* pro = nullptr;     // propagator of this slice (float2, nx*ny)
  const void* src = nullptr;     // INIT source waves (float2)
  long long src_stride = 0;      // 0: one shared source; nx*ny: one per batch entry
  const void* ppar = nullptr;    // INIT: float4 (dx, dy, fz, 0) per batch entry, or nullptr
  float kscx = 0.f, kscy = 0.f;  // 1/a0x, 1/a0y [1/nm]
  const float* cppl = nullptr;   // constant phase plate [rad] or nullptr
  int detect = 0;                // readout at this plane
  int ndet = 0;
  const float* detw = nullptr;   // [ndet][nx*ny] effective detector weights
  float* partial = nullptr;      // partial readout sums
  long long part_wstride = 0;    // per batch entry
  long long part_dstride = 0;    // per detector (= number of column tiles)
  long long part_off = 0;        // offset of this detection plane
  float* dif = nullptr;          // diffraction channel |psi|^2 (or nullptr)
  void* wff = nullptr;           // Fourier-space wave channel (float2, or nullptr)
  float* dif_avg = nullptr;      // averaged diffraction channel (or nullptr)
  int dsx = 0, dsy = 0;          // diffraction de-scan [pixel]
  void* wff_avg = nullptr;       // averaged Fourier-space wave channel (or nullptr)
  const float* cwgt = nullptr;   // per batch entry weight of the channels
  long long chan_wstride = 0;    // per batch entry stride of the channels (items)
  long long chan_off = 0;        // offset of this detection plane in the channels
  const void* twy = nullptr;     // twiddles W_ny^e (float2, ny)
  // column-major ([x][y]) copies for the wave-local column kernel
  const void* pro_t = nullptr;   // propagator of this slice (float2)
  const float* detw_t = nullptr; // [ndet][nx][ny] detector weights
  const float* cppl_t = nullptr; // constant phase plate or nullptr
  bool panel = false;            // the waves are in 16-column panels (panel_ok)
};

// real-space readout channels of one detection plane for the row pass
// (float, or float2 for the wave channels; any may be nullptr)
struct RowChanHost {
  float* img = nullptr;
  float* img_avg = nullptr;
  void* wfr = nullptr;
  void* wfr_avg = nullptr;
  const float* cwgt = nullptr;   // per batch entry weight
  long long chan_wstride = 0;    // per batch entry stride (items)
  long long chan_off = 0;        // offset of this detection plane (items)
};

// parameters of the device probe generator: spatial frequencies [1/nm]
// already derived from the probe's angles as the host generator does
struct ProbeArgsHost {
  int nx = 0, ny = 0;
  float ax = 0.f, ay = 0.f, wl = 0.f;           // cell [nm], wavelength [nm]
  float qlim = 0.f, qrs = 0.f;                  // aperture radius, edge smoothness
  float a1x = 0.f, a1y = 0.f, radet = 1.f;      // aperture distortion
  int open = 1;                                 // 0: aperture closed by distortion
  float qlcx = 0.f, qlcy = 0.f, btx = 0.f, bty = 0.f;  // aperture centre, beam tilt
  int nab = 0;                                  // aberration pairs used
  float coeff[48] = {};                         // (x, y) per aberration [nm]
  int mn[48] = {};                              // (order, symmetry) per aberration
  int binom[81] = {};                           // C(n, l) at n + 9 l, n, l <= 8
};

// one atom of the Gaussian projected-potential model, prepared on the host:
// wrapped position [A], amplitude S/(2 pi s^2) [V A], and the unwrapped pixel
// range [i0, i1] x [j0, j1] its 8 s disc covers
struct GaussAtom {
  double xa, ya, amp;
  int i0, i1, j0, j1;
};

// one atom of the form-factor model: position [nm] in the slice cell,
// occupancy, atom type (index of its 2-D form factor)
struct FFAtom {
  float x, y, occ;
  int type;
};

// Grouped launches (the uneven batch): the waves of several jobs with one
// grid advance together, one launch per pass and step over all of them.  A
// launch's block row y takes wave w0 + y of the wave map, which names its job
// and its index in that job's batch; the job's constant buffers and the
// per-plane slice data come from the job and plane tables.
struct GrpJob {
  void* wave = nullptr;           // the job's batch of waves (float2, nx*ny apart)
  const void* src = nullptr;      // INIT source: the job's probe (float2)
  const void* ppar = nullptr;     // float4 (dx, dy, fz, 0) per wave
  const int* var = nullptr;       // variant indices [wave][nslc]
  float* partial = nullptr;       // readout partial sums [wave][plane][detector][tile]
  const void* pgr = nullptr;      // phase gratings (float2)
  const float* detw_t = nullptr;  // [ndet][nx][ny] detector weights
  const float* cppl_t = nullptr;  // constant phase plate [x][y] or nullptr
  const float* detw = nullptr;    // the same row-major [ndet][ny][nx] (long-column kernel)
  const float* cppl = nullptr;    // the same row-major [y][x] or nullptr
  float kscx = 0.f, kscy = 0.f;   // 1/a0x, 1/a0y [1/nm]
  int nslc = 0;                   // slices of the job (row length of var)
  int ndet = 0;
  long long part_wstride = 0;     // per wave in partial
};
// plane p of a job: slice data of step p (p < nslc) and the readout of
// plane p (made by the column pass of step p - 1, or the INIT pass for p = 0)
struct GrpPlane {
  long long slc_off = 0;          // grating offset of the slice of step p (variant 0)
  const void* pro_t = nullptr;    // propagator of step p, column-major (float2)
  const void* pro = nullptr;      // the same row-major (long-column kernel)
  long long part_off = -1;        // offset of plane p in partial; -1: no readout
};
struct GrpWave {
  int job = 0;                    // index in the job table
  int w = 0;                      // wave index in that job's batch
};
struct GrpArgs {
  const GrpWave* wav = nullptr;   // wave map
  const GrpJob* job = nullptr;    // job table
  const GrpPlane* pl = nullptr;   // plane table [job][pl_ld]
  int pl_ld = 0;
  int w0 = 0;                     // first wave of the launch in the wave map
  int step = 0;                   // row pass: slice step; column pass: plane read out
};
// grouped launches exist for grids whose column pass is the wave-local or
// the long-column kernel
bool grp_supported(int nx, int ny);
// row pass of step ga.step over nwave waves of the map (from ga.w0)
hipError_t launch_row_pass_grp(int nx, int ny, int nwave, long long wave_stride, long long npg,
                               const float2* twx, const GrpArgs& ga, bool panel, hipStream_t st);
// column pass producing plane ga.step (propagator of step ga.step - 1 unless
// init) over nwave waves of the map; detect: some job of the launch reads
// plane ga.step out (each wave checks its own job's plane table)
hipError_t launch_col_pass_grp(int nx, int ny, int nwave, bool init, bool exit_, bool detect,
                               long long part_dstride, const float2* twy, const GrpArgs& ga,
                               bool panel, hipStream_t st);

constexpr int kMaxDetPerSweep = 8;

bool size_supported(int n);
int row_tiles(int nx, int ny);
int col_tiles(int ny, int nx);

// mode 0: slice; 1: slice with the real-space readout of `rc` first;
// 2: inverse x-transform and readout only (exit plane, after a column IFFT).
// panel: the waves are in 16-column panels (mode 0 only; see panel_ok)
hipError_t launch_row_pass(int nx, int ny, int nwave, float2* wave, long long wave_stride,
                           const float2* pgr, long long slc_off, long long npg, const int* var_tab,
                           int var_ld, int jslc, const float2* twx, const RowChanHost& rc, int mode,
                           bool panel, hipStream_t st);
// whether the staged batches of an nx x ny grid keep their waves in panels of
// 16 columns: element (y, x) at (x / 16) * 16 ny + 16 y + x % 16 (column
// tiles contiguous); chosen per line length where the passes run faster so
bool panel_ok(int nx, int ny);
// column transforms only (all columns of nbatch grids)
hipError_t launch_col_fft(int nx, int ny, int nbatch, float2* data, int dir, const float2* twy,
                          hipStream_t st);
hipError_t launch_col_pass(int ny, int nwave, bool init, bool exit_, const ColArgsHost& a,
                           hipStream_t st);
hipError_t launch_fft2d(int nx, int ny, int nbatch, float2* data, int dir, const float2* twx,
                        const float2* twy, hipStream_t st);
// zero partial[w * part_w + off + k], k < len, for the nwave waves (the
// partial sums of a detection plane a batch does not read out)
hipError_t launch_zero_partial(float* partial, long long part_w, long long off, int len, int nwave,
                               hipStream_t st);
hipError_t launch_combine(const float* partial, int ntiles, int ndet, int ndetslc, int npos,
                          int npass, const float* weights, float* out, int accumulate,
                          hipStream_t st);
hipError_t launch_offset_wave(float2* out, const float2* in, int nx, int ny, float kscx,
                              float kscy, float dx, float dy, float fz, const float* cppl,
                              int ish0, int ish1, hipStream_t st);
hipError_t launch_axpy_c(float2* dst, const float2* src, long long n, float w, hipStream_t st);
// phase gratings exp(i sig_int V) of nvar atom sets into out + v * n: tiles of
// 16 x 16 pixels, atoms of tile t of set v at idx[off[v * (ntiles + 1) + t] ..
// off[v * (ntiles + 1) + t + 1]) in ascending atom order
hipError_t launch_gauss_gratings(float2* out, int nx, int ny, int nvar, double hx, double hy,
                                 double rc2, double inv2s2, double sig_int, const GaussAtom* atoms,
                                 const int* off, const int* idx, hipStream_t st);
// phase gratings of nvar atom sets from projected form factors into out +
// v * n: structure-factor sum (ff[t][n] complex per atom type, atoms of set v
// at off[v] .. off[v+1] grouped by type), inverse 2-D transform, scale sc and
// exp(i sg V), and with bwl_pgr the band limit |q| <= kthr (weight ftrn)
hipError_t launch_ff_gratings(float2* out, int nx, int ny, int nvar, const float2* ff, const FFAtom* atoms,
                              const int* off, float itogx, float itogy, float kmax, int bwl_pot, float sc,
                              float sg, int bwl_pgr, float kthr, float ftrn, const float2* twx,
                              const float2* twy, hipStream_t st);
// probe into out (n complex), partial: ceil(n/256) doubles of scratch
hipError_t launch_probe(float2* out, double* partial, const ProbeArgsHost& a, hipStream_t st);
// device address of the allocation-free entry's scratch (kScratchFloats
// floats) on `device`, resolved once per device with that device current
constexpr long long kScratchFloats = 4LL << 20;  // 16 MiB
hipError_t scratch_address(int device, float** out);

}  // namespace jms
