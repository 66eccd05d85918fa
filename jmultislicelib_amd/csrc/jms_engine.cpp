// jms_engine.cpp -- implementation of the thin C-ABI in include/jms_hip.h:
// device memory, twiddle tables, the batched slice-loop schedule (eager or
// captured in a HIP graph) and the single-wave path used by CJMultiSlice.
//
// Schedule of one batch of W probe passes (all resident in HBM at once):
//   col_pass INIT  (plane islc0 readout)          1 launch
//   for j in [islc0, nobjslc):
//     row_pass  (slice j grating, per-wave variant)  1 launch
//     col_pass  (slice j propagator, plane j+1)       1 launch (EXIT for the last)
//   combine partial readouts -> per-position results 1 launch
// This is the reference loop JMultiSlice.cpp:4590-4667 with readout order
// unchanged (plane j read before slice j is applied; exit plane last).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <array>
#include <map>
#include <string>
#include <vector>

#include "jms_hip.h"
#include "jms_kernels.h"

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // (re)allocate to hold at least nb bytes; contents undefined
  hipError_t ensure(size_t nb) {
    if (nb <= bytes && p) return hipSuccess;
    release();
    if (nb == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, nb);
    if (e == hipSuccess) bytes = nb;
    else p = nullptr;
    return e;
  }
  template <class T>
  T* as() const { return (T*)p; }
};

struct Timer {
  double ms = 0.0;
  long long count = 0;
};

}  // namespace

// a member of group g is being destroyed (defined with the grouped launches)
void group_member_gone(jms_hip_group* g, jms_hip_engine* e);

struct jms_hip_engine {
  int device = 0;
  hipStream_t st = nullptr;
  std::string err;
  std::vector<jms_hip_group*> groups;  // grouped launches this engine is a member of
  // grid
  int nx = 0, ny = 0;
  long long n = 0;
  float a0x = 0.f, a0y = 0.f, wl = 0.f;
  DevBuf twx, twy;
  // gratings
  int npgx = 0, npgy = 0, nscslc = 0;
  std::vector<int> nvarslc;
  std::vector<long long> slcoffset;
  DevBuf pgr;
  // propagators
  int npro = 0;
  std::vector<int> proidx;
  DevBuf pro, pro_t;  // row-major and column-major ([x][y]) copies
  // object slices
  int nobjslc = 0;
  std::vector<int> objslc;
  // detection
  int ndetslc = 0, ndet = 0, detflags = 0;
  std::vector<int> det_objslc;
  DevBuf detw, detw_t;
  // incident wave
  DevBuf wav0, cppl, cppl_t;
  bool have_cppl = false;
  // staged scan
  int s_npos = 0, s_npass = 0;
  DevBuf s_ppar, s_weights, s_var, s_res;
  // batch working buffers
  int bcap_req = 0;
  DevBuf waves, partial, cur_var, cur_ppar, cur_res;
  // single-wave path: wave, variant sequence, results and channels
  DevBuf wav, var1, res1, w1, dif1, wff1, img1, wfr1;
  // averaging channels (JMultiSlice.cpp:2489-2660): accumulated over runs until
  // reset, with the summed run weights of the image and wave groups
  DevBuf dif_avg, wff_avg, img_avg, wfr_avg;
  float w_imgavg = 0.f, w_wavavg = 0.f;
  // diffraction de-scan [pixel] of the Fourier-space readouts
  int dsx = 0, dsy = 0;
  // graph cache, key = (nwave, islc0, positions, passes).  A captured launch
  // sequence holds raw device addresses, so the cache is also tied to the
  // addresses of every buffer the batch reads or writes (graph_bufs): when
  // any of them moves, all graphs are dropped before the next capture.
  bool use_graph = false;
  std::map<std::array<int, 4>, hipGraphExec_t> graphs;
  std::array<const void*, 16> graph_bufs{};
  // this device's scratch of the allocation-free batch entry (resolved at create)
  float* scratch = nullptr;
  // staged batches keep their waves in 16-column panels (set with the grid)
  bool panel = false;
  // timing
  bool timing = false;
  Timer tm[4];  // row pass, column pass, combine, graph replay
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  size_t ev_next = 0;

  int fail(int code, const char* what, hipError_t e) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s: %s (%d)", what, e == hipSuccess ? "" : hipGetErrorString(e), (int)e);
    err = buf;
    fprintf(stderr, "jms_hip error: %s\n", buf);
    return code;
  }
  int fail(int code, const char* what) {
    err = what;
    fprintf(stderr, "jms_hip error: %s\n", what);
    return code;
  }
  void drop_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }
  // device addresses a captured batch depends on
  std::array<const void*, 16> batch_buffers() const {
    return {waves.p, partial.p, cur_ppar.p, cur_var.p, cur_res.p, s_weights.p, wav0.p, pgr.p,
            pro.p, pro_t.p, detw.p, detw_t.p, twx.p, twy.p, cppl.p, cppl_t.p};
  }
  int ntiles_col() const { return jms::col_tiles(ny, nx); }
};

namespace {

void make_twiddles(int n, std::vector<float>& tw) {
  tw.resize(2 * (size_t)n);
  for (int e = 0; e < n; ++e) {
    const double a = -2.0 * 3.14159265358979323846264338327950288 * (double)e / (double)n;
    tw[2 * e] = (float)std::cos(a);
    tw[2 * e + 1] = (float)std::sin(a);
  }
}

// column-major copy of a row-major ny x nx array (element (x, y) at x*ny + y)
template <class T>
std::vector<T> to_colmajor(const T* src, int nx, int ny) {
  std::vector<T> out((size_t)nx * ny);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) out[(size_t)x * ny + y] = src[(size_t)y * nx + x];
  return out;
}

// events for the optional per-launch timing
hipEvent_t take_event(jms_hip_engine* e) {
  if (e->ev_next >= e->ev_pool.size()) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    e->ev_pool.push_back(ev);
  }
  return e->ev_pool[e->ev_next++];
}

void flush_timing(jms_hip_engine* e) {
  if (e->ev_pending.empty()) return;
  (void)hipStreamSynchronize(e->st);
  for (auto& pe : e->ev_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
      e->tm[pe.first].ms += ms;
      e->tm[pe.first].count += 1;
    }
  }
  e->ev_pending.clear();
  e->ev_next = 0;
}

struct TimedLaunch {
  jms_hip_engine* e;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  bool on;
  TimedLaunch(jms_hip_engine* e_, int kind_, bool on_) : e(e_), kind(kind_), on(on_) {
    if (on) {
      a = take_event(e);
      b = take_event(e);
      if (a && b) (void)hipEventRecord(a, e->st);
      else on = false;
    }
  }
  ~TimedLaunch() {
    if (on) {
      (void)hipEventRecord(b, e->st);
      e->ev_pending.push_back({kind, {a, b}});
      if (e->ev_pending.size() > 4096) flush_timing(e);
    }
  }
};

// per-wave readout channels of a batch (any may be null; cwgt is the
// per-wave weight, the channels hold ndetslc planes of n items per wave)
struct BatchChans {
  float* dif = nullptr;
  float* dif_avg = nullptr;
  float* img = nullptr;
  float* img_avg = nullptr;
  void* wff = nullptr;
  void* wff_avg = nullptr;
  void* wfr = nullptr;
  void* wfr_avg = nullptr;
  const float* cwgt = nullptr;
  bool real_space() const { return img || img_avg || wfr || wfr_avg; }
};

// Launch sequence of one batch.  `waves` holds the nwave working waves
// (stride n); `ppar` (float4 per wave) and `var` (nobjslc ints per wave)
// describe them; src/src_stride the INIT source.  Readout partial sums go to
// `partial` ([wave][plane][detector][column tile]), combined into `res`
// (npos results) when npass > 0; `ch` are the single-wave channels.
int enqueue_batch(jms_hip_engine* e, void* waves, float* partial, int nwave, int islc0, const void* src,
                  long long src_stride, const void* ppar, const int* var, int npos, int npass,
                  const float* weights, float* res, int accumulate, const BatchChans& ch, bool timed,
                  bool panel = false) {
  // panel waves (staged batches): integrated readouts only, the waves are internal
  if (panel && (ch.real_space() || ch.dif || ch.wff || ch.dif_avg || ch.wff_avg))
    return e->fail(135, "panel waves carry no readout channels");
  const int S = e->nobjslc;
  const int ntc = e->ntiles_col();
  const long long part_w = (long long)e->ndetslc * e->ndet * ntc;
  jms::ColArgsHost ca;
  ca.wave = waves;
  ca.nx = e->nx;
  ca.wave_stride = e->n;
  ca.kscx = 1.f / e->a0x;
  ca.kscy = 1.f / e->a0y;
  ca.cppl = e->have_cppl ? e->cppl.as<float>() : nullptr;
  ca.cppl_t = e->have_cppl ? e->cppl_t.as<float>() : nullptr;
  ca.detw_t = e->detw_t.as<float>();
  ca.ndet = (e->detflags & JMS_HIP_DETECT_INTEGRATED) ? e->ndet : 0;
  ca.detw = e->detw.as<float>();
  ca.partial = partial;
  ca.part_wstride = part_w;
  ca.part_dstride = ntc;
  ca.dif = ch.dif;
  ca.wff = ch.wff;
  ca.dif_avg = ch.dif_avg;
  ca.wff_avg = ch.wff_avg;
  ca.dsx = e->dsx;
  ca.dsy = e->dsy;
  ca.cwgt = ch.cwgt;
  ca.chan_wstride = (long long)e->ndetslc * e->n;
  jms::RowChanHost rc;
  rc.img = ch.img;
  rc.img_avg = ch.img_avg;
  rc.wfr = ch.wfr;
  rc.wfr_avg = ch.wfr_avg;
  rc.cwgt = ch.cwgt;
  rc.chan_wstride = ca.chan_wstride;
  ca.twy = e->twy.p;
  ca.panel = panel;
  hipError_t err;
  auto plane_setup = [&](int plane) {
    const int id = (plane >= 0 && plane <= S) ? e->det_objslc[plane] : -1;
    ca.detect = id >= 0 ? 1 : 0;
    ca.part_off = id >= 0 ? (long long)id * e->ndet * ntc : 0;
    ca.chan_off = id >= 0 ? (long long)id * e->n : 0;
  };
  // INIT column pass: plane islc0
  {
    ca.src = src;
    ca.src_stride = src_stride;
    ca.ppar = ppar;
    ca.pro = nullptr;
    plane_setup(islc0);
    TimedLaunch tl(e, 1, timed);
    err = jms::launch_col_pass(e->ny, nwave, true, islc0 >= S, ca, e->st);
    if (err != hipSuccess) return e->fail(130, "column pass (init)", err);
  }
  ca.src = nullptr;
  ca.ppar = nullptr;
  for (int j = islc0; j < S; ++j) {
    const int js = e->objslc[j] % e->nscslc;
    {
      // real-space readout of plane j inside the row pass (ReadoutImgDet_d)
      const int id = e->det_objslc[j];
      const int mode = (id >= 0 && ch.real_space()) ? 1 : 0;
      rc.chan_off = id >= 0 ? (long long)id * e->n : 0;
      TimedLaunch tl(e, 0, timed);
      err = jms::launch_row_pass(e->nx, e->ny, nwave, (float2*)waves, e->n,
                                 e->pgr.as<float2>(), e->slcoffset[js], (long long)e->npgx * e->npgy,
                                 var, S, j, e->twx.as<float2>(), rc, mode, panel, e->st);
      if (err != hipSuccess) return e->fail(131, "row pass", err);
    }
    ca.pro = e->pro.as<float2>() + (long long)e->proidx[js] * e->n;
    ca.pro_t = e->pro_t.as<float2>() + (long long)e->proidx[js] * e->n;
    plane_setup(j + 1);
    {
      TimedLaunch tl(e, 1, timed);
      err = jms::launch_col_pass(e->ny, nwave, false, j == S - 1, ca, e->st);
      if (err != hipSuccess) return e->fail(132, "column pass", err);
    }
  }
  if (npass > 0 && res && ca.ndet > 0) {
    // detection planes before islc0 are not read out in this batch: their
    // partial sums hold nothing of it (GPUMultislice reads only the planes
    // from islc0 on, JMultiSlice.cpp:4568-4590)
    for (int p = 0; p < islc0 && p <= S; ++p) {
      const int id = e->det_objslc[p];
      bool later = false;
      for (int q = islc0; q <= S && !later; ++q) later = e->det_objslc[q] == id;
      if (id < 0 || later) continue;
      err = jms::launch_zero_partial(partial, part_w, (long long)id * e->ndet * ntc, e->ndet * ntc, nwave, e->st);
      if (err != hipSuccess) return e->fail(134, "clear unread planes", err);
    }
    TimedLaunch tl(e, 2, timed);
    err = jms::launch_combine(partial, ntc, e->ndet, e->ndetslc, npos, npass, weights,
                              res, accumulate, e->st);
    if (err != hipSuccess) return e->fail(133, "combine", err);
  }
  return 0;
}

// exit-plane real-space readout (GPUMultislice :4662-4666): the column IFFT
// in place, then the inverse row transforms with the readout; the batch
// slots are left in a mixed state and must be treated as scratch.
int enqueue_exit_real_readout(jms_hip_engine* e, int nwave, const BatchChans& ch) {
  const int id = e->det_objslc[e->nobjslc];
  if (id < 0 || !ch.real_space()) return 0;
  jms::RowChanHost rc;
  rc.img = ch.img;
  rc.img_avg = ch.img_avg;
  rc.wfr = ch.wfr;
  rc.wfr_avg = ch.wfr_avg;
  rc.cwgt = ch.cwgt;
  rc.chan_wstride = (long long)e->ndetslc * e->n;
  rc.chan_off = (long long)id * e->n;
  hipError_t err = jms::launch_col_fft(e->nx, e->ny, nwave, e->waves.as<float2>(), +1, e->twy.as<float2>(), e->st);
  if (err == hipSuccess)
    err = jms::launch_row_pass(e->nx, e->ny, nwave, e->waves.as<float2>(), e->n, e->pgr.as<float2>(), 0, 0,
                               nullptr, 0, 0, e->twx.as<float2>(), rc, 2, false, e->st);
  if (err != hipSuccess) return e->fail(134, "exit-plane real-space readout", err);
  return 0;
}

int check_ready(jms_hip_engine* e) {
  if (!e) return 1;
  if (e->n <= 0) return e->fail(2, "grid not set");
  if (!e->pgr.p && e->nobjslc > 0) return e->fail(3, "phase gratings not set up");
  if (!e->pro.p && e->nobjslc > 0) return e->fail(4, "propagators not set up");
  if (e->det_objslc.size() != (size_t)e->nobjslc + 1) return e->fail(5, "detection not set up");
  return 0;
}

int ensure_batch(jms_hip_engine* e, int nwave) {
  hipError_t err = e->waves.ensure(sizeof(float) * 2 * (size_t)e->n * nwave);
  if (err != hipSuccess) return e->fail(110, "allocate wave batch", err);
  const size_t np = sizeof(float) * (size_t)nwave * std::max(1, e->ndetslc) *
                    std::max(1, e->ndet) * (size_t)e->ntiles_col();
  err = e->partial.ensure(np);
  if (err != hipSuccess) return e->fail(111, "allocate readout partials", err);
  return 0;
}

// default batch: about 2 GiB of resident waves
int auto_batch(jms_hip_engine* e, int total_waves) {
  if (e->bcap_req > 0) return std::min(total_waves, e->bcap_req);
  const long long per = 8LL * e->n;
  long long cap = (2LL << 30) / std::max(1LL, per);
  cap = std::max(1LL, std::min(cap, 8192LL));
  return (int)std::min<long long>(total_waves, cap);
}

}  // namespace

// ===========================================================================
extern "C" {

int jms_hip_create(jms_hip_engine** out) {
  if (!out) return 1;
  *out = nullptr;
  jms_hip_engine* e = new jms_hip_engine();
  hipError_t err = hipGetDevice(&e->device);
  if (err != hipSuccess) { delete e; return 101; }
  err = hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking);
  if (err != hipSuccess) { delete e; return 102; }
  err = jms::scratch_address(e->device, &e->scratch);
  if (err != hipSuccess) {
    fprintf(stderr, "jms_hip error: scratch address: %s\n", hipGetErrorString(err));
    (void)hipStreamDestroy(e->st);
    delete e;
    return 103;
  }
  *out = e;
  return 0;
}

void jms_hip_destroy(jms_hip_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->st) (void)hipStreamSynchronize(e->st);
  for (jms_hip_group* g : e->groups) group_member_gone(g, e);  // those groups refuse to run from now on
  e->groups.clear();
  e->drop_graphs();
  for (auto ev : e->ev_pool) (void)hipEventDestroy(ev);
  DevBuf* bufs[] = {&e->twx, &e->twy, &e->pgr, &e->pro, &e->pro_t, &e->detw, &e->detw_t, &e->wav0,
                    &e->cppl, &e->cppl_t,
                    &e->s_ppar, &e->s_weights, &e->s_var, &e->s_res, &e->waves, &e->partial,
                    &e->cur_var, &e->cur_ppar, &e->cur_res, &e->wav, &e->var1, &e->res1,
                    &e->w1, &e->dif1, &e->wff1, &e->img1, &e->wfr1, &e->dif_avg, &e->wff_avg,
                    &e->img_avg, &e->wfr_avg};
  for (DevBuf* b : bufs) b->release();
  if (e->st) (void)hipStreamDestroy(e->st);
  delete e;
}

int jms_hip_grid_supported(int nx, int ny) {
  return (jms::size_supported(nx) && jms::size_supported(ny)) ? 1 : 0;
}

const char* jms_hip_last_error(jms_hip_engine* e) { return e ? e->err.c_str() : "no engine"; }

int jms_hip_set_grid(jms_hip_engine* e, int nx, int ny, float a0x, float a0y, float wl) {
  if (!e) return 1;
  if (!jms_hip_grid_supported(nx, ny)) return e->fail(2, "grid size has no compiled kernels");
  if (!(a0x > 0.f) || !(a0y > 0.f)) return e->fail(3, "invalid super-cell size");
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  e->drop_graphs();
  e->nx = nx; e->ny = ny; e->n = (long long)nx * ny;
  e->a0x = a0x; e->a0y = a0y; e->wl = wl;
  e->panel = jms::panel_ok(nx, ny);
  std::vector<float> tw;
  make_twiddles(nx, tw);
  hipError_t err = e->twx.ensure(tw.size() * sizeof(float));
  if (err == hipSuccess) err = hipMemcpy(e->twx.p, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice);
  if (err != hipSuccess) return e->fail(101, "twiddles x", err);
  make_twiddles(ny, tw);
  err = e->twy.ensure(tw.size() * sizeof(float));
  if (err == hipSuccess) err = hipMemcpy(e->twy.p, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice);
  if (err != hipSuccess) return e->fail(102, "twiddles y", err);
  return 0;
}

int jms_hip_setup_gratings(jms_hip_engine* e, int npgx, int npgy, int nslc, const int* nvarslc) {
  if (!e) return 1;
  if (nslc < 1 || !nvarslc) return e->fail(2, "no phase gratings");
  if (npgx != e->nx || npgy != e->ny)
    return e->fail(3, "phase-grating grid differs from the wave grid (sub-frame gratings are not supported)");
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  e->drop_graphs();
  e->npgx = npgx; e->npgy = npgy; e->nscslc = nslc;
  e->nvarslc.assign(nvarslc, nvarslc + nslc);
  e->slcoffset.assign(nslc, 0);
  long long off = 0;
  for (int i = 0; i < nslc; ++i) {
    if (e->nvarslc[i] < 1) return e->fail(4, "structure slice without variants");
    e->slcoffset[i] = off;
    off += (long long)e->nvarslc[i] * npgx * npgy;
  }
  hipError_t err = e->pgr.ensure(sizeof(float) * 2 * (size_t)off);
  if (err != hipSuccess) return e->fail(104, "allocate phase gratings", err);
  err = hipMemset(e->pgr.p, 0, sizeof(float) * 2 * (size_t)off);
  if (err != hipSuccess) return e->fail(105, "clear phase gratings", err);
  return 0;
}

int jms_hip_upload_gratings(jms_hip_engine* e, int islc, int nvar, const float* pgr) {
  if (!e) return 1;
  if (islc < 0 || islc >= e->nscslc || !pgr || nvar < 1) return e->fail(2, "invalid grating upload");
  const int nl = std::min(nvar, e->nvarslc[islc]);
  const size_t nb = sizeof(float) * 2 * (size_t)nl * e->npgx * e->npgy;
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  hipError_t err = hipMemcpy(e->pgr.as<float2>() + e->slcoffset[islc], pgr, nb, hipMemcpyHostToDevice);
  if (err != hipSuccess) return e->fail(102, "upload phase gratings", err);
  return 0;
}

int jms_hip_gaussian_gratings(jms_hip_engine* e, int islc, int nvar, const int* nat, const double* xy,
                              const int* z, double sigma, double sig_int) {
  if (!e) return 1;
  if (islc < 0 || islc >= e->nscslc || nvar < 1 || !nat || sigma <= 0.0) return e->fail(2, "invalid grating setup");
  if (e->npgx != e->nx || e->npgy != e->ny) return e->fail(3, "gratings must match the grid");
  const int nl = std::min(nvar, e->nvarslc[islc]);
  const int nx = e->nx, ny = e->ny;
  const double PI = 3.14159265358979323846264338327950288;
  const double ax = 10.0 * (double)e->a0x, ay = 10.0 * (double)e->a0y;  // nm -> A
  const double hx = ax / nx, hy = ay / ny;
  const double rc = 8.0 * sigma, rc2 = rc * rc;
  const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
  const int ntx = (nx + 15) / 16, nty = (ny + 15) / 16, ntiles = ntx * nty;
  // atom records as the host generator computes them, and per-set tile lists
  std::vector<jms::GaussAtom> atoms;
  std::vector<int> off((size_t)nl * (ntiles + 1), 0), idx;
  std::vector<std::vector<int>> tiles_of;
  size_t a_base = 0;
  for (int v = 0; v < nl; ++v) {
    if (nat[v] < 0 || (nat[v] > 0 && (!xy || !z))) return e->fail(4, "invalid atom list");
    std::vector<std::vector<int>> lists(ntiles);
    std::vector<char> colhit(ntx), rowhit(nty);
    for (int at = 0; at < nat[v]; ++at) {
      const size_t k = a_base + at;
      jms::GaussAtom g;
      g.amp = 2.0 * std::pow((double)z[k], 0.7) / (2.0 * PI * sigma * sigma);
      g.xa = std::fmod(xy[2 * k], ax);
      if (g.xa < 0) g.xa += ax;
      g.ya = std::fmod(xy[2 * k + 1], ay);
      if (g.ya < 0) g.ya += ay;
      g.i0 = (int)std::floor((g.xa - rc) / hx); g.i1 = (int)std::ceil((g.xa + rc) / hx);
      g.j0 = (int)std::floor((g.ya - rc) / hy); g.j1 = (int)std::ceil((g.ya + rc) / hy);
      const int id = (int)atoms.size();
      atoms.push_back(g);
      std::fill(colhit.begin(), colhit.end(), 0);
      std::fill(rowhit.begin(), rowhit.end(), 0);
      for (int i = g.i0; i <= g.i1 && i < g.i0 + nx; ++i) colhit[(((i % nx) + nx) % nx) / 16] = 1;
      for (int j = g.j0; j <= g.j1 && j < g.j0 + ny; ++j) rowhit[(((j % ny) + ny) % ny) / 16] = 1;
      for (int ty = 0; ty < nty; ++ty)
        if (rowhit[ty])
          for (int tx = 0; tx < ntx; ++tx)
            if (colhit[tx]) lists[ty * ntx + tx].push_back(id);
    }
    int* o = off.data() + (size_t)v * (ntiles + 1);
    for (int t = 0; t < ntiles; ++t) {
      o[t] = (int)idx.size();
      idx.insert(idx.end(), lists[t].begin(), lists[t].end());
    }
    o[ntiles] = (int)idx.size();
    a_base += nat[v];
  }
  if (idx.empty()) idx.push_back(0);
  if (atoms.empty()) atoms.push_back(jms::GaussAtom{});
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  DevBuf d_atoms, d_off, d_idx;
  hipError_t err = d_atoms.ensure(sizeof(jms::GaussAtom) * atoms.size());
  if (err == hipSuccess) err = d_off.ensure(sizeof(int) * off.size());
  if (err == hipSuccess) err = d_idx.ensure(sizeof(int) * idx.size());
  if (err == hipSuccess) err = hipMemcpy(d_atoms.p, atoms.data(), sizeof(jms::GaussAtom) * atoms.size(), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(d_off.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(d_idx.p, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = jms::launch_gauss_gratings(e->pgr.as<float2>() + e->slcoffset[islc], nx, ny, nl, hx, hy, rc2, inv2s2,
                                     sig_int, (const jms::GaussAtom*)d_atoms.p, d_off.as<int>(), d_idx.as<int>(),
                                     e->st);
  if (err == hipSuccess) err = hipStreamSynchronize(e->st);
  d_atoms.release(); d_off.release(); d_idx.release();
  if (err != hipSuccess) return e->fail(102, "phase gratings on the device", err);
  return 0;
}

int jms_hip_formfactor_gratings(jms_hip_engine* e, int islc, int nvar, int ntype, const float* ff2d,
                                const int* nat, const float* xy, const int* type, const float* occ, float cz,
                                int flags) {
  if (!e) return 1;
  if (islc < 0 || islc >= e->nscslc || nvar < 1 || ntype < 1 || !ff2d || !nat || !(cz > 0.f))
    return e->fail(2, "invalid grating setup");
  if (e->npgx != e->nx || e->npgy != e->ny) return e->fail(3, "gratings must match the grid");
  const int nl = std::min(nvar, e->nvarslc[islc]);
  const int nx = e->nx, ny = e->ny;
  const size_t n = (size_t)nx * ny;
  // atoms of each set grouped by type, ascending atom order inside a type
  // (the reference's loop order: types outer, atoms inner, JEElaSca.cpp:1088-1122)
  std::vector<jms::FFAtom> atoms;
  std::vector<int> off(nl + 1, 0);
  size_t a_base = 0;
  for (int v = 0; v < nl; ++v) {
    if (nat[v] < 0 || (nat[v] > 0 && (!xy || !type))) return e->fail(4, "invalid atom list");
    for (int at = 0; at < nat[v]; ++at) {
      const int t = type[a_base + at];
      if (t < 0 || t >= ntype) return e->fail(4, "atom type out of range");
    }
    for (int t = 0; t < ntype; ++t)
      for (int at = 0; at < nat[v]; ++at) {
        const size_t k = a_base + at;
        if (type[k] != t) continue;
        atoms.push_back(jms::FFAtom{xy[2 * k], xy[2 * k + 1], occ ? occ[k] : 1.f, t});
      }
    off[v + 1] = (int)atoms.size();
    a_base += nat[v];
  }
  if (atoms.empty()) atoms.push_back(jms::FFAtom{0.f, 0.f, 0.f, 0});
  // JEElaSca.h:88,91,94: potential scale [V nm^2 per form-factor nm] and
  // interaction constant (their product is 1 / (4 pi)), band-limit fraction
  constexpr float POT_SCA = 0.03809982f, PGR_SIG = 2.08865737f, PGR_BWL = 0.6666667f;
  const float itogx = 1.f / e->a0x, itogy = 1.f / e->a0y;
  const float kmax = std::min((float)(nx / 2) * itogx, (float)(ny / 2) * itogy);
  const float vol = e->a0x * e->a0y * cz;
  const float sc = POT_SCA / vol, sg = PGR_SIG * cz * e->wl;
  const float ftrn = 1.f / (float)nx / (float)ny;
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  DevBuf d_ff, d_atoms, d_off;
  hipError_t err = d_ff.ensure(sizeof(float) * 2 * n * ntype);
  if (err == hipSuccess) err = d_atoms.ensure(sizeof(jms::FFAtom) * atoms.size());
  if (err == hipSuccess) err = d_off.ensure(sizeof(int) * off.size());
  if (err == hipSuccess) err = hipMemcpy(d_ff.p, ff2d, sizeof(float) * 2 * n * ntype, hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = hipMemcpy(d_atoms.p, atoms.data(), sizeof(jms::FFAtom) * atoms.size(), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(d_off.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = jms::launch_ff_gratings(e->pgr.as<float2>() + e->slcoffset[islc], nx, ny, nl, d_ff.as<float2>(),
                                  (const jms::FFAtom*)d_atoms.p, d_off.as<int>(), itogx, itogy, kmax,
                                  (flags & JMS_HIP_FF_BWL_POT) ? 1 : 0, sc, sg, (flags & JMS_HIP_FF_BWL_PGR) ? 1 : 0,
                                  kmax * PGR_BWL, ftrn, e->twx.as<float2>(), e->twy.as<float2>(), e->st);
  if (err == hipSuccess) err = hipStreamSynchronize(e->st);
  d_ff.release(); d_atoms.release(); d_off.release();
  if (err != hipSuccess) return e->fail(102, "form-factor phase gratings on the device", err);
  return 0;
}

int jms_hip_get_gratings(jms_hip_engine* e, int islc, int nvar, float* out) {
  if (!e || !out) return 1;
  if (islc < 0 || islc >= e->nscslc || nvar < 1) return e->fail(2, "invalid grating read");
  const int nl = std::min(nvar, e->nvarslc[islc]);
  (void)hipSetDevice(e->device);
  hipError_t err = hipMemcpy(out, e->pgr.as<float2>() + e->slcoffset[islc],
                             sizeof(float) * 2 * (size_t)nl * e->npgx * e->npgy, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->fail(102, "read phase gratings", err);
  return 0;
}

int jms_hip_setup_propagators(jms_hip_engine* e, int npro, int nslc, const int* proidx) {
  if (!e) return 1;
  if (npro < 1 || !proidx || nslc < 1) return e->fail(2, "no propagators");
  for (int i = 0; i < nslc; ++i)
    if (proidx[i] < 0 || proidx[i] >= npro) return e->fail(3, "propagator index out of range");
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  e->drop_graphs();
  e->npro = npro;
  e->proidx.assign(proidx, proidx + nslc);
  hipError_t err = e->pro.ensure(sizeof(float) * 2 * (size_t)npro * e->n);
  if (err == hipSuccess) err = e->pro_t.ensure(sizeof(float) * 2 * (size_t)npro * e->n);
  if (err != hipSuccess) return e->fail(102, "allocate propagators", err);
  return 0;
}

int jms_hip_upload_propagator(jms_hip_engine* e, int ipro, const float* pro) {
  if (!e) return 1;
  if (ipro < 0 || ipro >= e->npro || !pro) return e->fail(2, "invalid propagator upload");
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  hipError_t err = hipMemcpy(e->pro.as<float2>() + (long long)ipro * e->n, pro,
                             sizeof(float) * 2 * (size_t)e->n, hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    std::vector<float2> t = to_colmajor((const float2*)pro, e->nx, e->ny);
    err = hipMemcpy(e->pro_t.as<float2>() + (long long)ipro * e->n, t.data(), sizeof(float2) * t.size(),
                    hipMemcpyHostToDevice);
  }
  if (err != hipSuccess) return e->fail(102, "upload propagator", err);
  return 0;
}

int jms_hip_setup_objects(jms_hip_engine* e, int nobjslc, const int* objslc) {
  if (!e) return 1;
  if (nobjslc < 0 || (nobjslc > 0 && !objslc)) return e->fail(2, "invalid object slices");
  for (int i = 0; i < nobjslc; ++i)
    if (objslc[i] < 0) return e->fail(3, "negative structure slice index");
  e->drop_graphs();
  e->nobjslc = nobjslc;
  e->objslc.assign(objslc, objslc + nobjslc);
  return 0;
}

int jms_hip_setup_detection(jms_hip_engine* e, int ndetslc, const int* det_objslc, int ndet, int flags) {
  if (!e) return 1;
  if (!det_objslc || ndetslc < 1 || ndet < 0) return e->fail(2, "invalid detection setup");
  if (ndet > jms::kMaxDetPerSweep) return e->fail(3, "too many integrating detectors (max 8)");
  if (flags & ~JMS_HIP_DETECT_ALL) return e->fail(4, "unknown detection channel flag");
  std::vector<char> used(ndetslc, 0);
  for (int i = 0; i <= e->nobjslc; ++i) {
    const int id = det_objslc[i];
    if (id >= ndetslc) return e->fail(5, "detection index out of range");
    // every plane's readout fills its index's partial sums: an index shared
    // by two planes would keep only the later one
    if (id >= 0 && used[id]++) return e->fail(6, "a detection index used by two planes");
  }
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  e->drop_graphs();
  e->ndetslc = ndetslc;
  e->ndet = ndet;
  e->detflags = flags;
  e->det_objslc.assign(det_objslc, det_objslc + e->nobjslc + 1);
  // averaging channels start empty (DetectorSetup, JMultiSlice.cpp:784-800)
  {
    const size_t nchan = (size_t)ndetslc * e->n;
    struct { DevBuf* b; int flag; size_t nb; } avg[] = {
        {&e->dif_avg, JMS_HIP_DETECT_DIFFR_AVG, sizeof(float) * nchan},
        {&e->img_avg, JMS_HIP_DETECT_IMAGE_AVG, sizeof(float) * nchan},
        {&e->wff_avg, JMS_HIP_DETECT_WAVEF_AVG, 2 * sizeof(float) * nchan},
        {&e->wfr_avg, JMS_HIP_DETECT_WAVER_AVG, 2 * sizeof(float) * nchan}};
    for (auto& a : avg) {
      if (!(flags & a.flag)) { a.b->release(); continue; }
      hipError_t err = a.b->ensure(a.nb);
      if (err == hipSuccess) err = hipMemset(a.b->p, 0, a.nb);
      if (err != hipSuccess) return e->fail(104, "allocate averaging channels", err);
    }
    e->w_imgavg = 0.f;
    e->w_wavavg = 0.f;
  }
  if (ndet > 0) {
    hipError_t err = e->detw.ensure(sizeof(float) * (size_t)ndet * e->n);
    if (err == hipSuccess) err = hipMemset(e->detw.p, 0, sizeof(float) * (size_t)ndet * e->n);
    if (err == hipSuccess) err = e->detw_t.ensure(sizeof(float) * (size_t)ndet * e->n);
    if (err == hipSuccess) err = hipMemset(e->detw_t.p, 0, sizeof(float) * (size_t)ndet * e->n);
    if (err != hipSuccess) return e->fail(103, "allocate detector functions", err);
  }
  return 0;
}

int jms_hip_upload_detector(jms_hip_engine* e, int idet, const float* det, int msklen, const int* msk) {
  if (!e) return 1;
  if (idet < 0 || idet >= e->ndet || !det) return e->fail(2, "invalid detector upload");
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  // effective weights: det at every listed mask pixel (a pixel listed k times
  // counts k times, as the masked sum does), or det everywhere without a mask
  std::vector<float> wgt((size_t)e->n, 0.f);
  if (msk && msklen > 0) {
    for (int i = 0; i < msklen; ++i) {
      if (msk[i] < 0 || msk[i] >= e->n) return e->fail(3, "detector mask index out of range");
      wgt[msk[i]] += det[msk[i]];
    }
  } else {
    std::memcpy(wgt.data(), det, sizeof(float) * (size_t)e->n);
  }
  std::vector<float> wt = to_colmajor(wgt.data(), e->nx, e->ny);
  const size_t nb = sizeof(float) * (size_t)e->n;
  hipError_t err = hipMemcpy(e->detw.as<float>() + (long long)idet * e->n, wgt.data(), nb, hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = hipMemcpy(e->detw_t.as<float>() + (long long)idet * e->n, wt.data(), nb, hipMemcpyHostToDevice);
  if (err != hipSuccess) return e->fail(104, "upload detector", err);
  return 0;
}

int jms_hip_set_probe(jms_hip_engine* e, const float* wav0) {
  if (!e) return 1;
  if (!wav0 || e->n <= 0) return e->fail(2, "invalid probe");
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  const size_t nb = sizeof(float) * 2 * (size_t)e->n;
  hipError_t err = e->wav0.ensure(nb);
  if (err == hipSuccess) err = hipMemcpy(e->wav0.p, wav0, nb, hipMemcpyHostToDevice);
  if (err != hipSuccess) return e->fail(102, "upload probe", err);
  return 0;
}

int jms_hip_probe(jms_hip_engine* e, const float* prm, int nabrr, const float* abrr, float* out) {
  if (!e) return 1;
  if (!prm || e->n <= 0 || e->wl <= 0.f) return e->fail(2, "invalid probe parameters or grid");
  if (nabrr < 0 || (nabrr > 0 && !abrr)) return e->fail(3, "invalid aberration list");
  const float wl = e->wl;
  const float alpha = prm[0], alpha_rs = prm[1], ax0 = prm[2], ay0 = prm[3], asym = prm[4];
  const float adir = prm[5] * 0.0174533f, btx = prm[6], bty = prm[7];
  if (alpha <= 0.f) return e->fail(4, "aperture must be open");
  jms::ProbeArgsHost a;
  a.nx = e->nx; a.ny = e->ny; a.ax = e->a0x; a.ay = e->a0y; a.wl = wl;
  a.qlim = alpha * 0.001f / wl;  // mrad -> 1/nm, as CalculateProbeWaveFourier
  a.qrs = alpha_rs;
  a.a1x = asym * std::cos(2.f * adir);
  a.a1y = asym * std::sin(2 * adir);
  const float adet = 1.f - asym * asym;
  a.open = adet != 0.f;
  a.radet = a.open ? 1.f / adet : 1.f;
  a.qlcx = ax0 * 0.001f / wl; a.qlcy = ay0 * 0.001f / wl;
  a.btx = btx * 0.001f / wl; a.bty = bty * 0.001f / wl;
  // aberration list (order m, symmetry n), m = 1..8, m + n even (JProbeGen.cpp:187-216)
  int l = 0;
  for (int m = 1; m <= 8; m++)
    for (int n = 0; n <= m; n++)
      if ((m + n) % 2 == 0) { a.mn[2 * l] = m; a.mn[2 * l + 1] = n; l++; }
  for (int n = 0; n <= 8; n++)
    for (int k = 0; k <= n; k++) {
      int c = 1;
      for (int i = 0; i < std::min(k, n - k); ++i) { c *= (n - i); c /= (i + 1); }
      a.binom[n + 9 * k] = c;
    }
  a.nab = std::min(nabrr, 24);
  for (int k = 0; k < 2 * a.nab; ++k) a.coeff[k] = abrr[k];
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  const size_t nb = sizeof(float) * 2 * (size_t)e->n;
  DevBuf partial;
  hipError_t err = e->wav0.ensure(nb);
  if (err == hipSuccess) err = partial.ensure(sizeof(double) * (size_t)((e->n + 255) / 256));
  if (err == hipSuccess) err = jms::launch_probe(e->wav0.as<float2>(), partial.as<double>(), a, e->st);
  if (err == hipSuccess) err = hipStreamSynchronize(e->st);
  if (err == hipSuccess && out) err = hipMemcpy(out, e->wav0.p, nb, hipMemcpyDeviceToHost);
  partial.release();
  if (err != hipSuccess) return e->fail(102, "probe on the device", err);
  return 0;
}

int jms_hip_set_phaseplate(jms_hip_engine* e, const float* cppl) {
  if (!e) return 1;
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  e->drop_graphs();
  if (!cppl) {
    e->have_cppl = false;
    return 0;
  }
  const size_t nb = sizeof(float) * (size_t)e->n;
  std::vector<float> t = to_colmajor(cppl, e->nx, e->ny);
  hipError_t err = e->cppl.ensure(nb);
  if (err == hipSuccess) err = hipMemcpy(e->cppl.p, cppl, nb, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = e->cppl_t.ensure(nb);
  if (err == hipSuccess) err = hipMemcpy(e->cppl_t.p, t.data(), nb, hipMemcpyHostToDevice);
  if (err != hipSuccess) return e->fail(102, "upload phase plate", err);
  e->have_cppl = true;
  return 0;
}

int jms_hip_set_batch(jms_hip_engine* e, int nwave_max) {
  if (!e) return 1;
  e->bcap_req = std::min(std::max(0, nwave_max), 65535);  // waves are the launch grid's y dimension
  e->drop_graphs();
  return 0;
}

int jms_hip_set_graph(jms_hip_engine* e, int use_graph) {
  if (!e) return 1;
  e->use_graph = use_graph != 0;
  return 0;
}

int jms_hip_set_timing(jms_hip_engine* e, int enable) {
  if (!e) return 1;
  flush_timing(e);
  e->timing = enable != 0;
  for (auto& t : e->tm) { t.ms = 0.0; t.count = 0; }
  return 0;
}

int jms_hip_get_timing(jms_hip_engine* e, int kind, double* ms, long long* count) {
  if (!e || kind < 0 || kind > 3) return 1;
  flush_timing(e);
  if (ms) *ms = e->tm[kind].ms;
  if (count) *count = e->tm[kind].count;
  return 0;
}

int jms_hip_stage_scan(jms_hip_engine* e, int npos, const float* dx, const float* dy, const float* dz,
                       int npass, const float* weights, const int* var) {
  if (!e) return 1;
  if (npos < 1 || npass < 1 || !dx || !dy || !weights) return e->fail(2, "invalid scan");
  if (e->nobjslc > 0 && !var) return e->fail(3, "missing variant sequences");
  int rc = check_ready(e);
  if (rc) return rc;
  // variant range check
  const long long nv = (long long)npos * npass * e->nobjslc;
  for (long long i = 0; i < nv; ++i) {
    const int j = (int)(i % e->nobjslc);
    const int js = e->objslc[j] % e->nscslc;
    if (var[i] < 0 || var[i] >= e->nvarslc[js]) return e->fail(4, "variant index out of range");
  }
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  std::vector<float> pp(4 * (size_t)npos * npass);
  for (int p = 0; p < npos; ++p)
    for (int r = 0; r < npass; ++r) {
      float* q = &pp[4 * ((size_t)p * npass + r)];
      q[0] = dx[p];
      q[1] = dy[p];
      q[2] = (dz ? dz[p] : 0.f) * e->wl * 0.5f;  // fz = dz * lambda / 2 (JMultiSlice.cpp:1973)
      q[3] = 0.f;
    }
  hipError_t err = e->s_ppar.ensure(pp.size() * sizeof(float));
  if (err == hipSuccess) err = hipMemcpy(e->s_ppar.p, pp.data(), pp.size() * sizeof(float), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = e->s_weights.ensure(sizeof(float) * npass);
  if (err == hipSuccess) err = hipMemcpy(e->s_weights.p, weights, sizeof(float) * npass, hipMemcpyHostToDevice);
  if (err == hipSuccess && nv > 0) err = e->s_var.ensure(sizeof(int) * (size_t)nv);
  if (err == hipSuccess && nv > 0) err = hipMemcpy(e->s_var.p, var, sizeof(int) * (size_t)nv, hipMemcpyHostToDevice);
  const size_t nres = sizeof(float) * (size_t)npos * e->ndetslc * std::max(1, e->ndet);
  if (err == hipSuccess) err = e->s_res.ensure(nres);
  if (err == hipSuccess) err = hipMemset(e->s_res.p, 0, nres);
  if (err != hipSuccess) return e->fail(102, "stage scan", err);
  e->s_npos = npos;
  e->s_npass = npass;
  return 0;
}

int jms_hip_run_staged(jms_hip_engine* e, int pos0, int npos, int islc0) {
  if (!e) return 1;
  if (pos0 < 0 || npos < 1 || pos0 + npos > e->s_npos) return e->fail(2, "position range outside the staged scan");
  int rc = check_ready(e);
  if (rc) return rc;
  if (!e->wav0.p) return e->fail(6, "probe not set");
  islc0 = std::max(0, std::min(islc0, e->nobjslc));
  (void)hipSetDevice(e->device);
  const int npass = e->s_npass;
  const int total = npos * npass;
  int bw = auto_batch(e, total);
  int bpos = std::max(1, bw / npass);  // whole positions per batch
  bw = bpos * npass;
  rc = ensure_batch(e, bw);
  if (rc) return rc;
  const int S = e->nobjslc;
  const int per = e->ndetslc * std::max(1, e->ndet);
  for (int p = pos0; p < pos0 + npos; p += bpos) {
    const int np = std::min(bpos, pos0 + npos - p);
    const int nw = np * npass;
    const float4* ppar = e->s_ppar.as<float4>() + (long long)p * npass;
    const int* var = e->s_var.as<int>() + (long long)p * npass * S;
    float* res = e->s_res.as<float>() + (long long)p * per;
    if (e->use_graph) {
      // the graph reads fixed "current batch" buffers, refreshed before replay
      hipError_t err = e->cur_ppar.ensure(sizeof(float4) * (size_t)bw);
      if (err == hipSuccess && S > 0) err = e->cur_var.ensure(sizeof(int) * (size_t)bw * S);
      if (err == hipSuccess) err = e->cur_res.ensure(sizeof(float) * (size_t)bpos * per);
      if (err == hipSuccess) err = hipMemcpyAsync(e->cur_ppar.p, ppar, sizeof(float4) * nw, hipMemcpyDeviceToDevice, e->st);
      if (err == hipSuccess && S > 0)
        err = hipMemcpyAsync(e->cur_var.p, var, sizeof(int) * (size_t)nw * S, hipMemcpyDeviceToDevice, e->st);
      if (err != hipSuccess) return e->fail(120, "stage batch", err);
      const auto bufs = e->batch_buffers();
      if (bufs != e->graph_bufs) {
        e->drop_graphs();
        e->graph_bufs = bufs;
      }
      const std::array<int, 4> key{nw, islc0, np, npass};
      auto it = e->graphs.find(key);
      if (it == e->graphs.end()) {
        hipGraph_t g = nullptr;
        err = hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal);
        if (err != hipSuccess) return e->fail(121, "begin capture", err);
        rc = enqueue_batch(e, e->waves.p, e->partial.as<float>(), nw, islc0, e->wav0.p, 0, e->cur_ppar.p,
                           e->cur_var.as<int>(), np,
                           npass, e->s_weights.as<float>(), e->cur_res.as<float>(), 0, BatchChans(), false,
                           e->panel);
        hipError_t err2 = hipStreamEndCapture(e->st, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (err2 != hipSuccess) return e->fail(122, "end capture", err2);
        hipGraphExec_t ge = nullptr;
        err = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err != hipSuccess) return e->fail(123, "instantiate graph", err);
        it = e->graphs.emplace(key, ge).first;
      }
      {
        TimedLaunch tl(e, 3, e->timing);  // one replay = the batch's whole launch sequence
        err = hipGraphLaunch(it->second, e->st);
      }
      if (err == hipSuccess)
        err = hipMemcpyAsync(res, e->cur_res.p, sizeof(float) * (size_t)np * per, hipMemcpyDeviceToDevice, e->st);
      if (err != hipSuccess) return e->fail(124, "graph launch", err);
    } else {
      rc = enqueue_batch(e, e->waves.p, e->partial.as<float>(), nw, islc0, e->wav0.p, 0, ppar, var, np, npass,
                         e->s_weights.as<float>(), res,
                         0, BatchChans(), e->timing, e->panel);
      if (rc) return rc;
    }
  }
  return 0;
}

int jms_hip_run_batch_noalloc(jms_hip_engine* e, int islc0, int nwave, void* waves, const float* ppar,
                              const int* var, int npos, int npass, const float* weights, float* res) {
  if (!e) return 1;
  int rc = check_ready(e);
  if (rc) return rc;
  if (!e->wav0.p) return e->fail(6, "probe not set");
  if (nwave < 1 || nwave > 65535 || npos < 1 || npass < 1 || (long long)npos * npass != nwave)
    return e->fail(2, "batch shape: need 1 <= nwave = npos * npass <= 65535");
  if (!waves || !ppar || !weights || !res || (e->nobjslc > 0 && !var)) return e->fail(3, "missing buffer");
  if (!(e->detflags & JMS_HIP_DETECT_INTEGRATED) || e->ndet < 1)
    return e->fail(4, "no integrating detectors set up");
  const long long need = (long long)nwave * std::max(1, e->ndetslc) * e->ndet * e->ntiles_col();
  if (need > jms::kScratchFloats) {
    char msg[200];
    snprintf(msg, sizeof(msg), "batch needs %lld floats of readout scratch, more than the %lld available "
             "(fewer waves per call)", need, (long long)jms::kScratchFloats);
    return e->fail(5, msg);
  }
  if (!e->scratch) return e->fail(7, "no scratch address");
  islc0 = std::max(0, std::min(islc0, e->nobjslc));
  (void)hipSetDevice(e->device);
  return enqueue_batch(e, waves, e->scratch, nwave, islc0, e->wav0.p, 0, ppar, var, npos, npass, weights, res, 0,
                       BatchChans(), false, e->panel);
}

// ---------------------------------------------------------------------------
// Grouped launches over several engines with one grid (the uneven batch).
// Every member's staged scan runs in full; the waves of all members advance
// together, one row and one column launch per slice step over the waves of
// the members that still have slices (members sorted by slice count, so
// those are a prefix of the wave map), the members that reach their exit
// plane at a step in a launch of their own.  Tables:
//   plane table [member][maxS + 1] (GrpPlane): slice offset and propagator
//     of each step, readout offset of each plane;
//   per round: job table [member] (GrpJob) and wave map (GrpWave).
// A round takes at most one batch of each member (auto_batch); members with
// more positions than one batch continue in later rounds.
// ---------------------------------------------------------------------------
}  // extern "C"

struct jms_hip_group {
  std::vector<jms_hip_engine*> eng;  // members, by slice count descending
  int device = 0, nx = 0, ny = 0, smax = 0;
  bool orphaned = false;             // a member was destroyed: the group cannot run
  std::string err;
  struct Round {
    size_t job_off = 0, wav_off = 0;  // byte offsets in tab
    int nwave = 0;
    std::vector<int> active;          // active[p]: waves whose member has > p slices (p = 0..smax)
    std::vector<int> member_w;        // waves of each member in this round
    std::vector<int> member_p0;       // first position of each member in this round
    std::vector<std::vector<char>> det;  // det[p][member]: member reads plane p out
  };
  std::vector<Round> rounds;
  std::vector<char> host;             // tables as last uploaded
  DevBuf tab;
  hipEvent_t ev_done = nullptr;
  std::vector<hipEvent_t> ev_in;
  int fail(int code, const char* what, hipError_t e = hipSuccess) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s%s%s", what, e == hipSuccess ? "" : ": ", e == hipSuccess ? "" : hipGetErrorString(e));
    err = buf;
    fprintf(stderr, "jms_hip error: group: %s\n", buf);
    return code;
  }
};

void group_member_gone(jms_hip_group* g, jms_hip_engine* e) {
  for (auto& m : g->eng)
    if (m == e) m = nullptr;
  g->orphaned = true;
}

namespace {

// host tables of the group for the members' current state; fills g->rounds
int group_build(jms_hip_group* g, std::vector<char>& host) {
  const int m = (int)g->eng.size();
  g->smax = 0;
  for (auto* e : g->eng) g->smax = std::max(g->smax, e->nobjslc);
  const int ld = g->smax + 1;
  std::vector<int> bpos(m), nround(m);
  int R = 0;
  for (int k = 0; k < m; ++k) {
    jms_hip_engine* e = g->eng[k];
    const int total = e->s_npos * e->s_npass;
    int bw = auto_batch(e, total);
    bpos[k] = std::max(1, bw / e->s_npass);
    int rc = ensure_batch(e, bpos[k] * e->s_npass);
    if (rc) return g->fail(110, "allocate member batch");
    nround[k] = (e->s_npos + bpos[k] - 1) / bpos[k];
    R = std::max(R, nround[k]);
  }
  // plane table
  std::vector<jms::GrpPlane> planes((size_t)m * ld);
  for (int k = 0; k < m; ++k) {
    jms_hip_engine* e = g->eng[k];
    const int ntc = e->ntiles_col();
    const int ndet = (e->detflags & JMS_HIP_DETECT_INTEGRATED) ? e->ndet : 0;
    for (int p = 0; p <= e->nobjslc; ++p) {
      jms::GrpPlane& q = planes[(size_t)k * ld + p];
      if (p < e->nobjslc) {
        const int js = e->objslc[p] % e->nscslc;
        q.slc_off = e->slcoffset[js];
        q.pro_t = e->pro_t.as<float2>() + (long long)e->proidx[js] * e->n;
        q.pro = e->pro.as<float2>() + (long long)e->proidx[js] * e->n;
      }
      const int id = e->det_objslc[p];
      q.part_off = (id >= 0 && ndet > 0) ? (long long)id * ndet * ntc : -1;
    }
  }
  size_t off = sizeof(jms::GrpPlane) * planes.size();
  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  off = align(off);
  g->rounds.assign(R, jms_hip_group::Round());
  std::vector<std::vector<jms::GrpJob>> jobs(R);
  std::vector<std::vector<jms::GrpWave>> wmap(R);
  for (int r = 0; r < R; ++r) {
    auto& rd = g->rounds[r];
    rd.member_w.assign(m, 0);
    rd.member_p0.assign(m, 0);
    rd.active.assign(ld, 0);
    rd.det.assign(ld, std::vector<char>(m, 0));
    jobs[r].resize(m);
    for (int k = 0; k < m; ++k) {
      jms_hip_engine* e = g->eng[k];
      const int p0 = r * bpos[k];
      const int np = std::max(0, std::min(bpos[k], e->s_npos - p0));
      const int nw = np * e->s_npass;
      const int S = e->nobjslc;
      jms::GrpJob& jb = jobs[r][k];
      jb.wave = e->waves.p;
      jb.src = e->wav0.p;
      jb.ppar = e->s_ppar.as<float4>() + (long long)p0 * e->s_npass;
      jb.var = e->s_var.as<int>() + (long long)p0 * e->s_npass * S;
      jb.partial = e->partial.as<float>();
      jb.pgr = e->pgr.p;
      jb.detw_t = e->detw_t.as<float>();
      jb.cppl_t = e->have_cppl ? e->cppl_t.as<float>() : nullptr;
      jb.detw = e->detw.as<float>();
      jb.cppl = e->have_cppl ? e->cppl.as<float>() : nullptr;
      jb.kscx = 1.f / e->a0x;
      jb.kscy = 1.f / e->a0y;
      jb.nslc = S;
      jb.ndet = (e->detflags & JMS_HIP_DETECT_INTEGRATED) ? e->ndet : 0;
      jb.part_wstride = (long long)e->ndetslc * e->ndet * e->ntiles_col();
      rd.member_w[k] = nw;
      rd.member_p0[k] = p0;
      for (int w = 0; w < nw; ++w) wmap[r].push_back({k, w});
      for (int p = 0; p < ld; ++p) {
        if (S > p) rd.active[p] += nw;
        if (nw > 0 && p <= S && planes[(size_t)k * ld + p].part_off >= 0) rd.det[p][k] = 1;
      }
    }
    rd.nwave = (int)wmap[r].size();
    rd.job_off = off;
    off = align(off + sizeof(jms::GrpJob) * m);
    rd.wav_off = off;
    off = align(off + sizeof(jms::GrpWave) * std::max<size_t>(1, wmap[r].size()));
  }
  host.assign(off, 0);
  memcpy(host.data(), planes.data(), sizeof(jms::GrpPlane) * planes.size());
  for (int r = 0; r < R; ++r) {
    memcpy(host.data() + g->rounds[r].job_off, jobs[r].data(), sizeof(jms::GrpJob) * m);
    if (!wmap[r].empty())
      memcpy(host.data() + g->rounds[r].wav_off, wmap[r].data(), sizeof(jms::GrpWave) * wmap[r].size());
  }
  return 0;
}

// any member among the waves [w0, w1) of round rd reads plane p out
bool group_detects(const jms_hip_group* g, const jms_hip_group::Round& rd, int p, int w0, int w1) {
  int base = 0;
  for (size_t k = 0; k < g->eng.size(); ++k) {
    const int nw = rd.member_w[k];
    if (nw > 0 && base < w1 && base + nw > w0 && rd.det[p][k]) return true;
    base += nw;
  }
  return false;
}

}  // namespace

extern "C" {

int jms_hip_group_create(jms_hip_group** out, jms_hip_engine* const* engines, int n) {
  if (!out) return 1;
  *out = nullptr;
  if (!engines || n < 1) return 2;
  for (int k = 0; k < n; ++k) {
    if (!engines[k]) return 2;
    for (int j = 0; j < k; ++j)
      if (engines[j] == engines[k]) return 3;  // a member twice
  }
  jms_hip_engine* e0 = engines[0];
  if (e0->n <= 0) return 4;
  for (int k = 0; k < n; ++k) {
    jms_hip_engine* e = engines[k];
    if (e->nx != e0->nx || e->ny != e0->ny || e->device != e0->device) return 4;
  }
  if (!jms::grp_supported(e0->nx, e0->ny)) return 5;
  jms_hip_group* g = new jms_hip_group();
  g->eng.assign(engines, engines + n);
  for (jms_hip_engine* e : g->eng) e->groups.push_back(g);
  std::stable_sort(g->eng.begin(), g->eng.end(),
                   [](const jms_hip_engine* a, const jms_hip_engine* b) { return a->nobjslc > b->nobjslc; });
  g->device = e0->device;
  g->nx = e0->nx;
  g->ny = e0->ny;
  (void)hipSetDevice(g->device);
  if (hipEventCreateWithFlags(&g->ev_done, hipEventDisableTiming) != hipSuccess) {
    jms_hip_group_destroy(g);
    return 101;
  }
  g->ev_in.resize(n, nullptr);
  for (int k = 0; k < n; ++k)
    if (hipEventCreateWithFlags(&g->ev_in[k], hipEventDisableTiming) != hipSuccess) {
      jms_hip_group_destroy(g);
      return 101;
    }
  *out = g;
  return 0;
}

void jms_hip_group_destroy(jms_hip_group* g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  for (jms_hip_engine* e : g->eng) {  // live members: wait for the group's work, forget the group
    if (!e) continue;
    (void)hipStreamSynchronize(e->st);
    e->groups.erase(std::remove(e->groups.begin(), e->groups.end(), g), e->groups.end());
  }
  g->tab.release();
  for (auto ev : g->ev_in) if (ev) (void)hipEventDestroy(ev);
  if (g->ev_done) (void)hipEventDestroy(g->ev_done);
  delete g;
}

const char* jms_hip_group_last_error(jms_hip_group* g) { return g ? g->err.c_str() : "null group"; }

int jms_hip_group_run(jms_hip_group* g) {
  if (!g) return 1;
  if (g->orphaned) return g->fail(19, "a member engine was destroyed");
  for (auto* e : g->eng) {
    if (e->nx != g->nx || e->ny != g->ny || e->device != g->device)
      return g->fail(18, "member grid or device changed since the group was made");
    if (e->dsx != 0 || e->dsy != 0) return g->fail(20, "member with a diffraction de-scan (not grouped)");
    int rc = check_ready(e);
    if (rc) return g->fail(10 + rc, "member not ready");
    if (!e->wav0.p) return g->fail(16, "member probe not set");
    if (e->s_npos < 1) return g->fail(17, "member has no staged scan");
  }
  (void)hipSetDevice(g->device);
  jms_hip_engine* lead = g->eng[0];
  hipStream_t st = lead->st;
  // the group's launches follow every member's earlier work
  for (size_t k = 1; k < g->eng.size(); ++k) {
    hipError_t err = hipEventRecord(g->ev_in[k], g->eng[k]->st);
    if (err == hipSuccess) err = hipStreamWaitEvent(st, g->ev_in[k], 0);
    if (err != hipSuccess) return g->fail(102, "order member streams", err);
  }
  std::vector<char> host;
  int rc = group_build(g, host);
  if (rc) return rc;
  if (host != g->host) {  // tables changed: upload once no launch reads the old ones
    hipError_t err = hipStreamSynchronize(st);
    if (err == hipSuccess) err = g->tab.ensure(host.size());
    if (err == hipSuccess) err = hipMemcpy(g->tab.p, host.data(), host.size(), hipMemcpyHostToDevice);
    if (err != hipSuccess) return g->fail(103, "upload group tables", err);
    g->host.swap(host);
  }
  const char* base = (const char*)g->tab.p;
  const long long n = (long long)g->nx * g->ny;
  const int ld = g->smax + 1;
  const int ntc = lead->ntiles_col();
  hipError_t err = hipSuccess;
  for (const auto& rd : g->rounds) {
    if (rd.nwave == 0) continue;
    jms::GrpArgs ga;
    ga.wav = (const jms::GrpWave*)(base + rd.wav_off);
    ga.job = (const jms::GrpJob*)(base + rd.job_off);
    ga.pl = (const jms::GrpPlane*)base;
    ga.pl_ld = ld;
    // INIT column pass: plane 0 for every wave; members without slices exit
    auto col = [&](int plane, bool init, int w0, int w1, bool exit_) -> hipError_t {
      if (w1 <= w0) return hipSuccess;
      ga.w0 = w0;
      ga.step = plane;
      TimedLaunch tl(lead, 1, lead->timing);
      return jms::launch_col_pass_grp(g->nx, g->ny, w1 - w0, init, exit_, group_detects(g, rd, plane, w0, w1), ntc,
                                      lead->twy.as<float2>(), ga, lead->panel, st);
    };
    err = col(0, true, 0, rd.active[0], false);
    if (err == hipSuccess) err = col(0, true, rd.active[0], rd.nwave, true);
    for (int s = 0; err == hipSuccess && s < g->smax && rd.active[s] > 0; ++s) {
      ga.w0 = 0;
      ga.step = s;
      {
        TimedLaunch tl(lead, 0, lead->timing);
        err = jms::launch_row_pass_grp(g->nx, g->ny, rd.active[s], n, n, lead->twx.as<float2>(), ga, lead->panel, st);
      }
      if (err == hipSuccess) err = col(s + 1, false, 0, rd.active[s + 1], false);
      if (err == hipSuccess) err = col(s + 1, false, rd.active[s + 1], rd.active[s], true);
    }
    if (err != hipSuccess) return g->fail(131, "grouped launch", err);
    for (size_t k = 0; k < g->eng.size(); ++k) {
      jms_hip_engine* e = g->eng[k];
      const int nw = rd.member_w[k];
      if (nw == 0 || !(e->detflags & JMS_HIP_DETECT_INTEGRATED) || e->ndet < 1) continue;
      const int per = e->ndetslc * e->ndet;
      TimedLaunch tl(lead, 2, lead->timing);
      err = jms::launch_combine(e->partial.as<float>(), ntc, e->ndet, e->ndetslc, nw / e->s_npass, e->s_npass,
                                e->s_weights.as<float>(), e->s_res.as<float>() + (long long)rd.member_p0[k] * per, 0,
                                st);
      if (err != hipSuccess) return g->fail(133, "combine", err);
    }
  }
  // members' later work (and jms_hip_sync on any member) follows the group
  err = hipEventRecord(g->ev_done, st);
  for (size_t k = 1; err == hipSuccess && k < g->eng.size(); ++k) err = hipStreamWaitEvent(g->eng[k]->st, g->ev_done, 0);
  if (err != hipSuccess) return g->fail(104, "order member streams", err);
  return 0;
}

int jms_hip_sync(jms_hip_engine* e) {
  if (!e) return 1;
  (void)hipSetDevice(e->device);
  hipError_t err = hipStreamSynchronize(e->st);
  if (err != hipSuccess) return e->fail(101, "synchronize", err);
  flush_timing(e);
  return 0;
}

int jms_hip_get_scan_results(jms_hip_engine* e, float* out) {
  if (!e || !out) return 1;
  (void)hipSetDevice(e->device);
  hipError_t err = hipStreamSynchronize(e->st);
  const size_t nb = sizeof(float) * (size_t)e->s_npos * e->ndetslc * e->ndet;
  if (err == hipSuccess && nb > 0) err = hipMemcpy(out, e->s_res.p, nb, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->fail(102, "read scan results", err);
  return 0;
}

// --------------------------------------------------------------------------
// single-wave path
// --------------------------------------------------------------------------
int jms_hip_wave_offset(jms_hip_engine* e, float dx, float dy, float dz, int ibtx, int ibty, int use_cppl) {
  if (!e) return 1;
  if (!e->wav0.p) return e->fail(2, "probe not set");
  (void)hipSetDevice(e->device);
  hipError_t err = e->wav.ensure(sizeof(float) * 2 * (size_t)e->n);
  if (err != hipSuccess) return e->fail(103, "allocate wave", err);
  int ish0 = ibtx % e->nx; if (ish0 < 0) ish0 += e->nx;  // JMultiSlice.cpp:1990-1991
  int ish1 = ibty % e->ny; if (ish1 < 0) ish1 += e->ny;
  err = jms::launch_offset_wave(e->wav.as<float2>(), e->wav0.as<float2>(), e->nx, e->ny, 1.f / e->a0x,
                                1.f / e->a0y, dx, dy, dz * e->wl * 0.5f,
                                (use_cppl && e->have_cppl) ? e->cppl.as<float>() : nullptr, ish0, ish1, e->st);
  if (err == hipSuccess) err = hipStreamSynchronize(e->st);
  if (err != hipSuccess) return e->fail(106, "offset incoming wave", err);
  return 0;
}

int jms_hip_wave_set(jms_hip_engine* e, const float* wav) {
  if (!e || !wav) return 1;
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  hipError_t err = e->wav.ensure(sizeof(float) * 2 * (size_t)e->n);
  if (err == hipSuccess) err = hipMemcpy(e->wav.p, wav, sizeof(float) * 2 * (size_t)e->n, hipMemcpyHostToDevice);
  if (err != hipSuccess) return e->fail(102, "set wave", err);
  return 0;
}

int jms_hip_wave_get(jms_hip_engine* e, float* dst) {
  if (!e || !dst) return 1;
  if (!e->wav.p) return e->fail(2, "no wave");
  (void)hipSetDevice(e->device);
  hipError_t err = hipMemcpy(dst, e->wav.p, sizeof(float) * 2 * (size_t)e->n, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->fail(102, "get wave", err);
  return 0;
}

int jms_hip_multislice(jms_hip_engine* e, int islc0, const int* var, float weight, int accumulate) {
  if (!e) return 1;
  int rc = check_ready(e);
  if (rc) return rc;
  if (!e->wav.p) return e->fail(6, "no incoming wave (call jms_hip_wave_offset or jms_hip_wave_set)");
  const int S = e->nobjslc;
  if (S > 0 && !var) return e->fail(7, "missing variant sequence");
  for (int j = 0; j < S; ++j) {
    const int js = e->objslc[j] % e->nscslc;
    if (var[j] < 0 || var[j] >= e->nvarslc[js]) return e->fail(8, "variant index out of range");
  }
  islc0 = std::max(0, std::min(islc0, S));
  (void)hipSetDevice(e->device);
  rc = ensure_batch(e, 1);
  if (rc) return rc;
  const int per = e->ndetslc * std::max(1, e->ndet);
  const size_t nchan = (size_t)e->ndetslc * e->n;
  hipError_t err = e->res1.ensure(sizeof(float) * per);
  if (err == hipSuccess) err = e->w1.ensure(sizeof(float));
  if (err == hipSuccess && S > 0) err = e->var1.ensure(sizeof(int) * S);
  const int fl = e->detflags;
  // per-run channels (cleared unless accumulating) and their buffers
  struct { DevBuf* b; int flag; size_t nb; } run[] = {
      {&e->dif1, JMS_HIP_DETECT_DIFFRACTION, sizeof(float) * nchan},
      {&e->img1, JMS_HIP_DETECT_IMAGE, sizeof(float) * nchan},
      {&e->wff1, JMS_HIP_DETECT_WAVEFOURIER, 2 * sizeof(float) * nchan},
      {&e->wfr1, JMS_HIP_DETECT_WAVEREAL, 2 * sizeof(float) * nchan}};
  for (auto& r : run)
    if (err == hipSuccess && (fl & r.flag)) err = r.b->ensure(r.nb);
  if (err != hipSuccess) return e->fail(110, "allocate single-wave buffers", err);
  if (!accumulate) {  // ClearDetMem_d (JMultiSlice.cpp:4054-4108); the averaging channels stay
    err = hipMemsetAsync(e->res1.p, 0, sizeof(float) * per, e->st);
    for (auto& r : run)
      if (err == hipSuccess && (fl & r.flag)) err = hipMemsetAsync(r.b->p, 0, r.nb, e->st);
    if (err != hipSuccess) return e->fail(111, "clear detector channels", err);
  }
  err = hipMemcpyAsync(e->w1.p, &weight, sizeof(float), hipMemcpyHostToDevice, e->st);
  if (err == hipSuccess && S > 0) err = hipMemcpyAsync(e->var1.p, var, sizeof(int) * S, hipMemcpyHostToDevice, e->st);
  if (err != hipSuccess) return e->fail(112, "stage single wave", err);
  BatchChans ch;
  ch.cwgt = e->w1.as<float>();
  if (fl & JMS_HIP_DETECT_DIFFRACTION) ch.dif = e->dif1.as<float>();
  if (fl & JMS_HIP_DETECT_IMAGE) ch.img = e->img1.as<float>();
  if (fl & JMS_HIP_DETECT_WAVEFOURIER) ch.wff = e->wff1.p;
  if (fl & JMS_HIP_DETECT_WAVEREAL) ch.wfr = e->wfr1.p;
  if (fl & JMS_HIP_DETECT_DIFFR_AVG) ch.dif_avg = e->dif_avg.as<float>();
  if (fl & JMS_HIP_DETECT_IMAGE_AVG) ch.img_avg = e->img_avg.as<float>();
  if (fl & JMS_HIP_DETECT_WAVEF_AVG) ch.wff_avg = e->wff_avg.p;
  if (fl & JMS_HIP_DETECT_WAVER_AVG) ch.wfr_avg = e->wfr_avg.p;
  rc = enqueue_batch(e, e->waves.p, e->partial.as<float>(), 1, islc0, e->wav.p, e->n, nullptr,
                     e->var1.as<int>(), 1, 1, e->w1.as<float>(),
                     e->res1.as<float>(), 1, ch, e->timing);
  if (rc) return rc;
  // the run leaves psi(k) of the exit plane in the batch slot; keep it as the wave
  err = hipMemcpyAsync(e->wav.p, e->waves.p, sizeof(float) * 2 * (size_t)e->n, hipMemcpyDeviceToDevice, e->st);
  if (err != hipSuccess) return e->fail(113, "multislice run", err);
  rc = enqueue_exit_real_readout(e, 1, ch);
  if (rc) return rc;
  err = hipStreamSynchronize(e->st);
  if (err != hipSuccess) return e->fail(113, "multislice run", err);
  // averaging weights count the runs that fed the averaging channels (:4668-4673)
  if (fl & (JMS_HIP_DETECT_DIFFR_AVG | JMS_HIP_DETECT_IMAGE_AVG)) e->w_imgavg += weight;
  if (fl & (JMS_HIP_DETECT_WAVEF_AVG | JMS_HIP_DETECT_WAVER_AVG)) e->w_wavavg += weight;
  return 0;
}

int jms_hip_get_result(jms_hip_engine* e, int which, float* dst) {
  if (!e || !dst) return 1;
  (void)hipSetDevice(e->device);
  hipError_t err = hipStreamSynchronize(e->st);
  const size_t nchan = (size_t)e->ndetslc * e->n;
  if (err != hipSuccess) return e->fail(101, "synchronize", err);
  if (which == JMS_HIP_DETECT_INTEGRATED) {
    if (!e->res1.p) return e->fail(2, "no result");
    err = hipMemcpy(dst, e->res1.p, sizeof(float) * (size_t)e->ndetslc * e->ndet, hipMemcpyDeviceToHost);
  } else if (which == JMS_HIP_DETECT_DIFFRACTION) {
    if (!e->dif1.p) return e->fail(3, "diffraction channel not active");
    err = hipMemcpy(dst, e->dif1.p, sizeof(float) * nchan, hipMemcpyDeviceToHost);
  } else if (which == JMS_HIP_DETECT_WAVEFOURIER) {
    if (!e->wff1.p) return e->fail(4, "wave channel not active");
    err = hipMemcpy(dst, e->wff1.p, sizeof(float) * 2 * nchan, hipMemcpyDeviceToHost);
  } else if (which == JMS_HIP_DETECT_IMAGE) {
    if (!e->img1.p) return e->fail(6, "image channel not active");
    err = hipMemcpy(dst, e->img1.p, sizeof(float) * nchan, hipMemcpyDeviceToHost);
  } else if (which == JMS_HIP_DETECT_WAVEREAL) {
    if (!e->wfr1.p) return e->fail(7, "real-space wave channel not active");
    err = hipMemcpy(dst, e->wfr1.p, sizeof(float) * 2 * nchan, hipMemcpyDeviceToHost);
  } else {
    return e->fail(5, "unsupported result channel");
  }
  if (err != hipSuccess) return e->fail(102, "read result", err);
  return 0;
}

int jms_hip_get_avg_result(jms_hip_engine* e, int which, float* dst, float* weight) {
  if (!e || !dst) return 1;
  (void)hipSetDevice(e->device);
  hipError_t err = hipStreamSynchronize(e->st);
  if (err != hipSuccess) return e->fail(101, "synchronize", err);
  const size_t nchan = (size_t)e->ndetslc * e->n;
  const DevBuf* b = nullptr;
  size_t nb = 0;
  float w = 0.f;
  switch (which) {  // GetAvgResult GPU branch (JMultiSlice.cpp:2534-2560)
    case JMS_HIP_DETECT_IMAGE_AVG: b = &e->img_avg; nb = sizeof(float) * nchan; w = e->w_imgavg; break;
    case JMS_HIP_DETECT_DIFFR_AVG: b = &e->dif_avg; nb = sizeof(float) * nchan; w = e->w_imgavg; break;
    case JMS_HIP_DETECT_WAVER_AVG: b = &e->wfr_avg; nb = 2 * sizeof(float) * nchan; w = e->w_wavavg; break;
    case JMS_HIP_DETECT_WAVEF_AVG: b = &e->wff_avg; nb = 2 * sizeof(float) * nchan; w = e->w_wavavg; break;
    default: return e->fail(5, "unsupported averaging channel");
  }
  if (!b->p) return e->fail(6, "averaging channel not active");
  err = hipMemcpy(dst, b->p, nb, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->fail(102, "read averaging channel", err);
  if (weight) *weight = w;
  return 0;
}

int jms_hip_set_descan(jms_hip_engine* e, int dnx, int dny) {
  if (!e) return 1;
  if (e->dsx != dnx || e->dsy != dny) e->drop_graphs();
  e->dsx = dnx;
  e->dsy = dny;
  return 0;
}

int jms_hip_get_averaging_weights(jms_hip_engine* e, float* img_weight, float* wave_weight) {
  if (!e) return 1;
  if (img_weight) *img_weight = e->w_imgavg;
  if (wave_weight) *wave_weight = e->w_wavavg;
  return 0;
}

int jms_hip_reset_averaging(jms_hip_engine* e, int group) {
  if (!e) return 1;
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);  // launches already enqueued may still read what this replaces
  hipError_t err = hipStreamSynchronize(e->st);
  const size_t nchan = (size_t)e->ndetslc * e->n;
  if (group == 0) {  // ResetImageAveraging (JMultiSlice.cpp:2592-2627)
    e->w_imgavg = 0.f;
    if (err == hipSuccess && e->img_avg.p) err = hipMemset(e->img_avg.p, 0, sizeof(float) * nchan);
    if (err == hipSuccess && e->dif_avg.p) err = hipMemset(e->dif_avg.p, 0, sizeof(float) * nchan);
  } else if (group == 1) {  // ResetWaveAveraging (:2629-2664)
    e->w_wavavg = 0.f;
    if (err == hipSuccess && e->wfr_avg.p) err = hipMemset(e->wfr_avg.p, 0, 2 * sizeof(float) * nchan);
    if (err == hipSuccess && e->wff_avg.p) err = hipMemset(e->wff_avg.p, 0, 2 * sizeof(float) * nchan);
  } else {
    return e->fail(2, "unknown averaging group");
  }
  if (err != hipSuccess) return e->fail(101, "reset averaging channels", err);
  return 0;
}

int jms_hip_readout_det_avg(jms_hip_engine* e, int which, float* dst, float* weight) {
  if (!e || !dst) return 1;
  (void)hipSetDevice(e->device);
  hipError_t err = hipStreamSynchronize(e->st);
  if (err != hipSuccess) return e->fail(101, "synchronize", err);
  const size_t n = (size_t)e->n, nchan = (size_t)e->ndetslc * n;
  if (weight) *weight = e->w_wavavg;
  // intensities of the averaged waves (ReadoutDetAvg GPU branch, JMultiSlice.cpp:2806-2890)
  const DevBuf* src = nullptr;
  if (which == JMS_HIP_DETECT_IMAGE) src = &e->wfr_avg;
  else if (which == JMS_HIP_DETECT_INTEGRATED || which == JMS_HIP_DETECT_DIFFRACTION) src = &e->wff_avg;
  else return e->fail(5, "unsupported result channel");
  if (!src->p) return e->fail(6, "the averaged wave of that space is not collected");
  std::vector<float> wv(2 * nchan);
  err = hipMemcpy(wv.data(), src->p, sizeof(float) * 2 * nchan, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->fail(102, "read averaged wave", err);
  if (which != JMS_HIP_DETECT_INTEGRATED) {
    for (size_t i = 0; i < nchan; ++i) dst[i] = wv[2 * i] * wv[2 * i] + wv[2 * i + 1] * wv[2 * i + 1];
    return 0;
  }
  if (e->ndet <= 0) return 0;
  std::vector<float> det((size_t)e->ndet * n);
  err = hipMemcpy(det.data(), e->detw.p, sizeof(float) * det.size(), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return e->fail(103, "read detector functions", err);
  for (int p = 0; p < e->ndetslc; ++p) {
    const float* w = wv.data() + 2 * n * p;
    for (int d = 0; d < e->ndet; ++d) {
      const float* dw = det.data() + n * d;
      double acc = 0.0;
      for (size_t i = 0; i < n; ++i) acc += (double)(w[2 * i] * w[2 * i] + w[2 * i + 1] * w[2 * i + 1]) * dw[i];
      dst[(size_t)p * e->ndet + d] = (float)acc;
    }
  }
  return 0;
}

int jms_hip_fft2d(int nx, int ny, int nbatch, float* data, int dir) {
  if (!data || nbatch < 1 || dir == 0) return 1;
  if (!jms_hip_grid_supported(nx, ny)) return 2;
  std::vector<float> tx, ty;
  make_twiddles(nx, tx);
  make_twiddles(ny, ty);
  const size_t nb = sizeof(float) * 2 * (size_t)nx * ny * nbatch;
  DevBuf d, dtx, dty;
  hipError_t err = d.ensure(nb);
  if (err == hipSuccess) err = dtx.ensure(tx.size() * sizeof(float));
  if (err == hipSuccess) err = dty.ensure(ty.size() * sizeof(float));
  if (err == hipSuccess) err = hipMemcpy(d.p, data, nb, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(dtx.p, tx.data(), tx.size() * sizeof(float), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(dty.p, ty.data(), ty.size() * sizeof(float), hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = jms::launch_fft2d(nx, ny, nbatch, d.as<float2>(), dir, dtx.as<float2>(), dty.as<float2>(), nullptr);
  if (err == hipSuccess) err = hipDeviceSynchronize();
  if (err == hipSuccess) err = hipMemcpy(data, d.p, nb, hipMemcpyDeviceToHost);
  d.release(); dtx.release(); dty.release();
  if (err != hipSuccess) {
    fprintf(stderr, "jms_hip_fft2d: %s\n", hipGetErrorString(err));
    return 101;
  }
  return 0;
}

}  // extern "C"
