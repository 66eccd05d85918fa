// tune_size.hip -- times the launch variants of the row and column passes
// (middle slices, about 1 GiB of waves) for one line length, -DJMS_N=<n>,
// and prints one JSON object per variant.  Used by tune.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>
#include "jms_kernels_impl.hpp"

using namespace jms;
#ifndef JMS_N
#error "compile with -DJMS_N=<line length>"
#endif
constexpr int N = JMS_N;
using B = Cfg<N>;

// row variants: rows per block (default, x2, /2) x global twiddles
template <int LPBV, bool G>
struct RowV : B {
  static constexpr int LPB = LPBV;
  static constexpr bool GTW_ROW = G;
  static constexpr bool TWR_ROW = G;
};
// the tuned row shape with the other load/store cache policy
struct RowNT : B {
  static constexpr bool NT_ROW = !B::NT_ROW;
};
// rows on TRV threads (several butterflies per thread where a stage has more)
template <int TRV, int LPBV>
struct RowT : B {
  static constexpr int T = TRV;
  static constexpr int LPB = LPBV;
};
// the tuned row shape with a given LDS padding shift (3: one slot per 8
// elements; >= 20: none)
template <int PADV>
struct RowP : B {
  static constexpr int PADSH_ROW = PADV;
};
// column variants (wave-local kernel): columns per block x waves per SIMD x
// threads per column (0: default)
template <int CWV, int WPE, int TWV = 0, bool TWRV = B::TWR, int PADV = B::PADSH_COL, bool NTV = B::NT_COL>
struct ColV : B {
  static constexpr bool NT_COL = NTV;
  static constexpr int CW = CWV;
  static constexpr int CC = CWV;
  static constexpr int WPE_COL = WPE;
  static constexpr int TW = TWV > 0 ? TWV : B::TW;
  static constexpr bool TWR = TWRV;
  static constexpr int PADSH_COL = PADV;
};

#define CK(x)                                                                  \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::printf("{\"error\": \"%s\", \"line\": %d}\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                                \
    }                                                                          \
  } while (0)

// long-column kernel: CT columns per block, block order (groups of G waves
// rotated by ROT tiles; G = 0 the hardware order)
// (default cache policy: the one jms_cfg.hpp gives a tile of CT columns)
template <int CT, int G, int ROT, bool NT = (CT == 16 && N >= 256 && Tuned<N>::NTC)>
struct XcolV : B {
  static constexpr int XCT = CT;
  static constexpr int CC = CT;
  static constexpr int XG_X = G;
  static constexpr int XROT_X = ROT;
  static constexpr bool NT_COL = NT;
};

struct Bufs {
  float2 *w, *pro, *tw, *pg;
  int* var;
  int nw;
};

template <class CF, bool PANEL = false>
float time_row(const Bufs& b, int reps) {
  dim3 g((N + CF::LPB - 1) / CF::LPB, b.nw), bl(CF::T * CF::LPB);
  RowChan rc{};
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0);
  for (int i = 0; i < reps; ++i)
    hipLaunchKernelGGL((k_row_pass<N, 0, CF, false, PANEL>), g, bl, 0, 0, b.w, N, (long long)N * N, b.pg, 0LL,
                       (long long)N * N, b.var, 1, 0, b.tw, rc, GrpArgs{});
  (void)hipEventRecord(e1);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  return ms / reps;
}

// column pass of a middle slice with the kernel the launcher would pick for CF
template <class CF, bool PANEL = false>
float time_col(const Bufs& b, int reps) {
  ColArgs a{};
  a.wave = b.w; a.nx = N; a.wave_stride = (long long)N * N; a.pro = b.pro; a.pro_t = b.pro; a.twg = b.tw;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0);
  for (int i = 0; i < reps; ++i) {
    if constexpr (CF::XL) {
      hipLaunchKernelGGL((k_col_pass_x<N, false, false, false, CF, PANEL>), dim3((N + CF::XCT - 1) / CF::XCT, b.nw),
                         dim3(64 * CF::XCT), 0, 0, a);
    } else if constexpr (CF::WL) {
      hipLaunchKernelGGL((k_col_pass_w<N, false, false, false, CF, false, PANEL>), dim3((N + CF::CW - 1) / CF::CW, b.nw),
                         dim3(CF::TW * CF::CW), 0, 0, a);
    } else {
      hipLaunchKernelGGL((k_col_pass<N, false, false, CF>), dim3((N + CF::C - 1) / CF::C, b.nw),
                         dim3(CF::TC * CF::C), 0, 0, a);
    }
  }
  (void)hipEventRecord(e1);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  return ms / reps;
}

template <class CF>
void row_variant(const Bufs& b, const char* name, int lpb, bool gtw, double bytes) {
  float best = 1e30f;
  for (int r = 0; r < 3; ++r) best = std::fmin(best, time_row<CF>(b, 10));
  std::printf("{\"n\": %d, \"pass\": \"row\", \"variant\": \"%s\", \"lpb\": %d, \"gtw\": %d, \"ms\": %.5f, \"gbs\": %.1f}\n",
              N, name, lpb, (int)gtw, best, bytes / (best * 1e-3) / 1e9);
}

template <class CF>
void col_variant(const Bufs& b, int cw, int wpe, double bytes, int tw = 0) {
  float best = 1e30f;
  for (int r = 0; r < 3; ++r) best = std::fmin(best, time_col<CF>(b, 10));
  std::printf("{\"n\": %d, \"pass\": \"col\", \"cw\": %d, \"wpe\": %d, \"tw\": %d, \"twr\": %d, \"pad\": %d, "
              "\"ms\": %.5f, \"gbs\": %.1f}\n", N, cw, wpe, tw, (int)CF::TWR, CF::PADSH_COL, best,
              bytes / (best * 1e-3) / 1e9);
}

// FFT plan candidates (a file written by tune.py: JMS_TUNE_PLANS(X) lists
// X("r0,r1,...", Radices<...>)), each timed with the tuned launch knobs
#ifdef JMS_PLANS_FILE
#include JMS_PLANS_FILE
#endif

template <class PL>
void plan_variant(const Bufs& b, const char* name, double bytes) {
  using C = CfgP<N, PL, PL>;
  float br = 1e30f, bc = 1e30f;
  for (int r = 0; r < 3; ++r) {
    br = std::fmin(br, time_row<C>(b, 10));
    bc = std::fmin(bc, time_col<C>(b, 10));
  }
  std::printf("{\"n\": %d, \"pass\": \"plan\", \"plan\": \"%s\", \"row_ms\": %.5f, \"col_ms\": %.5f, "
              "\"gbs\": %.1f}\n", N, name, br, bc, 2.0 * bytes / ((br + bc) * 1e-3) / 1e9);
}

int main() {
  Bufs b{};
  b.nw = (int)std::min<size_t>(65535, ((size_t)1 << 30) / ((size_t)N * N * 8));  // grid y limit
  CK(hipMalloc(&b.w, (size_t)N * N * b.nw * 8));
  CK(hipMalloc(&b.pro, (size_t)N * N * 8));
  CK(hipMalloc(&b.pg, (size_t)N * N * 8));
  CK(hipMalloc(&b.tw, N * 8));
  CK(hipMalloc(&b.var, b.nw * 4));
  CK(hipMemset(b.w, 0, (size_t)N * N * b.nw * 8));
  CK(hipMemset(b.var, 0, b.nw * 4));
  std::vector<float2> h(N);
  for (int e = 0; e < N; ++e) {
    const double an = -2.0 * 3.14159265358979323846 * e / N;
    h[e] = make_float2((float)std::cos(an), (float)std::sin(an));
  }
  CK(hipMemcpy(b.tw, h.data(), N * 8, hipMemcpyHostToDevice));
  std::vector<float2> one((size_t)N * N, make_float2(1.f, 0.f));
  CK(hipMemcpy(b.pro, one.data(), one.size() * 8, hipMemcpyHostToDevice));
  CK(hipMemcpy(b.pg, one.data(), one.size() * 8, hipMemcpyHostToDevice));
  const double bytes = 16.0 * N * N * b.nw;
  constexpr int L0 = (256 / B::T) > 0 ? (256 / B::T) : 1;  // default rows per block
  constexpr int L2 = (B::T * 2 * L0 <= 1024) ? 2 * L0 : L0;
  constexpr int LH = L0 > 1 ? L0 / 2 : 1;
  if constexpr (N % 16 == 0) {  // waves in 16-column panels: the tuned row and column kernels
    float br = 1e30f, bp = 1e30f, cr = 1e30f, cp = 1e30f;
    for (int r = 0; r < 3; ++r) {
      br = std::fmin(br, time_row<B, false>(b, 10));
      bp = std::fmin(bp, time_row<B, true>(b, 10));
      cr = std::fmin(cr, time_col<B, false>(b, 10));
      cp = std::fmin(cp, time_col<B, true>(b, 10));
    }
    std::printf("{\"n\": %d, \"pass\": \"layout\", \"row_ms\": %.5f, \"row_panel_ms\": %.5f, \"col_ms\": %.5f, "
                "\"col_panel_ms\": %.5f}\n", N, br, bp, cr, cp);
    // panel rows with twice and half the rows per block (longer / shorter contiguous runs per panel)
    constexpr int LX = 2 * B::LPB, LH2 = B::LPB > 1 ? B::LPB / 2 : 1;
    float b2 = 1e30f, bh = 1e30f;
    for (int r = 0; r < 3; ++r) {
      if constexpr (B::T * LX <= 1024) b2 = std::fmin(b2, time_row<RowV<LX, B::GTW_ROW>, true>(b, 10));
      bh = std::fmin(bh, time_row<RowV<LH2, B::GTW_ROW>, true>(b, 10));
    }
    std::printf("{\"n\": %d, \"pass\": \"panel_rows\", \"lpb\": %d, \"ms\": %.5f, \"lpb_x2\": %d, \"ms_x2\": %.5f, "
                "\"lpb_half\": %d, \"ms_half\": %.5f}\n", N, B::LPB, bp, LX, b2, LH2, bh);
  }
  if constexpr (B::T > 64) {  // one wavefront per row
    row_variant<RowT<64, 4>>(b, "tr64", 4, B::GTW_ROW, bytes);
    row_variant<RowT<64, 2>>(b, "tr64", 2, B::GTW_ROW, bytes);
    row_variant<RowT<64, 8>>(b, "tr64", 8, B::GTW_ROW, bytes);
    if constexpr (N % 16 == 0) {
      float bp = 1e30f;
      for (int r = 0; r < 3; ++r) bp = std::fmin(bp, time_row<RowT<64, 4>, true>(b, 10));
      std::printf("{\"n\": %d, \"pass\": \"tr64_panel\", \"lpb\": 4, \"ms\": %.5f}\n", N, bp);
    }
  }
  row_variant<RowP<3>>(b, "tuned", B::LPB, B::GTW_ROW, bytes);
  row_variant<RowP<24>>(b, "tuned_flat", B::LPB, B::GTW_ROW, bytes);
  row_variant<RowP<2>>(b, "tuned_pad2", B::LPB, B::GTW_ROW, bytes);
  row_variant<RowP<4>>(b, "tuned_pad4", B::LPB, B::GTW_ROW, bytes);
  row_variant<RowP<5>>(b, "tuned_pad5", B::LPB, B::GTW_ROW, bytes);
  row_variant<RowV<L0, false>>(b, "default", L0, false, bytes);
  row_variant<RowV<L0, true>>(b, "default", L0, true, bytes);
  if (L2 != L0) {
    row_variant<RowV<L2, false>>(b, "x2", L2, false, bytes);
    row_variant<RowV<L2, true>>(b, "x2", L2, true, bytes);
  }
  if (LH != L0) {
    row_variant<RowV<LH, false>>(b, "half", LH, false, bytes);
    row_variant<RowV<LH, true>>(b, "half", LH, true, bytes);
  }
  {  // the row pass with the other cache policy, in both layouts
    float bn = 1e30f, bp = 1e30f;
    for (int r = 0; r < 3; ++r) {
      bn = std::fmin(bn, time_row<RowNT, false>(b, 10));
      if constexpr (N % 16 == 0) bp = std::fmin(bp, time_row<RowNT, true>(b, 10));
    }
    std::printf("{\"n\": %d, \"pass\": \"row_nt\", \"nt\": %d, \"ms\": %.5f, \"panel_ms\": %.5f}\n", N,
                (int)RowNT::NT_ROW, bn, bp);
  }
  if constexpr (B::WL && !B::XL) {
    col_variant<ColV<16, 1>>(b, 16, 1, bytes);
    col_variant<ColV<16, 8>>(b, 16, 8, bytes);
    if constexpr (B::TW % 8 == 0) {
      col_variant<ColV<8, 1>>(b, 8, 1, bytes);
      col_variant<ColV<8, 8>>(b, 8, 8, bytes);
    }
    // 32 threads per column (two columns per wavefront): half the block, so
    // two 16-column blocks can share a CU where the LDS tile allows it
    if constexpr (B::TW == 64) {
      col_variant<ColV<16, 1, 32>>(b, 16, 1, bytes, 32);
      col_variant<ColV<16, 4, 32>>(b, 16, 4, bytes, 32);
      col_variant<ColV<12, 1, 32>>(b, 12, 1, bytes, 32);
      col_variant<ColV<12, 1>>(b, 12, 1, bytes);
    }
    // one table read per twiddle leg instead of the product tree (fewer
    // VALU instructions, more LDS reads)
    col_variant<ColV<16, 1, 0, false>>(b, 16, 1, bytes);
    if constexpr (B::TW % 8 == 0) col_variant<ColV<8, 1, 0, false>>(b, 8, 1, bytes);
    // no LDS padding: every exchange address a constant offset from one base
    col_variant<ColV<16, 1, 0, B::TWR, 24>>(b, 16, 1, bytes);
    if constexpr (B::TW % 8 == 0) col_variant<ColV<8, 1, 0, B::TWR, 24>>(b, 8, 1, bytes);
    // the tuned shape with the other load/store cache policy, in both layouts
    {
      using CN = ColV<B::CW, B::WPE_COL, B::TW, B::TWR, B::PADSH_COL, !B::NT_COL>;
      float bn = 1e30f, bp = 1e30f;
      for (int r = 0; r < 3; ++r) {
        bn = std::fmin(bn, time_col<CN, false>(b, 10));
        if constexpr (N % 16 == 0) bp = std::fmin(bp, time_col<CN, true>(b, 10));
      }
      std::printf("{\"n\": %d, \"pass\": \"col_nt\", \"cw\": %d, \"nt\": %d, \"ms\": %.5f, \"panel_ms\": %.5f}\n",
                  N, B::CW, (int)CN::NT_COL, bn, bp);
    }
    // 16 and 8 columns per block, each with both cache policies, in both layouts
    {
      auto cn = [&](auto cf) {
        using CF = decltype(cf);
        float bn = 1e30f, bp = 1e30f;
        if constexpr ((CF::TW * CF::CW) % 64 == 0) {
          for (int r = 0; r < 3; ++r) {
            bn = std::fmin(bn, time_col<CF, false>(b, 10));
            if constexpr (N % 16 == 0) bp = std::fmin(bp, time_col<CF, true>(b, 10));
          }
        }
        std::printf("{\"n\": %d, \"pass\": \"col_cw_nt\", \"cw\": %d, \"nt\": %d, \"ms\": %.5f, "
                    "\"panel_ms\": %.5f}\n", N, CF::CW, (int)CF::NT_COL, bn, bp);
      };
      if constexpr ((B::TW * 16) % 64 == 0) {
        cn(ColV<16, B::WPE_COL, B::TW, B::TWR, B::PADSH_COL, true>{});
        cn(ColV<16, B::WPE_COL, B::TW, B::TWR, B::PADSH_COL, false>{});
      }
      if constexpr ((B::TW * 8) % 64 == 0) {
        cn(ColV<8, B::WPE_COL, B::TW, B::TWR, B::PADSH_COL, true>{});
        cn(ColV<8, B::WPE_COL, B::TW, B::TWR, B::PADSH_COL, false>{});
      }
    }
    // one padding slot per 4 / 16 / 32 elements (per 4: exchange offsets stay
    // constants where the strides are multiples of 4)
    col_variant<ColV<16, 1, 0, B::TWR, 2>>(b, 16, 1, bytes);
    if constexpr (B::TW % 8 == 0) col_variant<ColV<8, 1, 0, B::TWR, 2>>(b, 8, 1, bytes);
    col_variant<ColV<16, 1, 0, B::TWR, 4>>(b, 16, 1, bytes);
    if constexpr (B::TW % 8 == 0) col_variant<ColV<8, 1, 0, B::TWR, 4>>(b, 8, 1, bytes);
    col_variant<ColV<16, 1, 0, B::TWR, 5>>(b, 16, 1, bytes);
    if constexpr (B::TW % 8 == 0) col_variant<ColV<8, 1, 0, B::TWR, 5>>(b, 8, 1, bytes);
    // 128 threads per column (two wavefronts, block barriers in the
    // transform): every radix stage in one round, far fewer registers
    if constexpr (B::T > 64 && B::T <= 128) {
      col_variant<ColV<8, 1, 128>>(b, 8, 1, bytes, 128);
      col_variant<ColV<8, 8, 128>>(b, 8, 8, bytes, 128);
    }
  }
  if constexpr (B::XL) {
    // 16 or 8 columns per block (8: two blocks per CU); hardware order,
    // XCD-contiguous with block columns fastest, and that with each block
    // row rotated by 8 block columns
    auto xcol = [&](auto cf) {
      using CF = decltype(cf);
      float best = 1e30f;
      for (int r = 0; r < 3; ++r) best = std::fmin(best, time_col<CF>(b, 10));
      std::printf("{\"n\": %d, \"pass\": \"xcol\", \"ct\": %d, \"g\": %d, \"rot\": %d, \"nt\": %d, "
                  "\"ms\": %.5f, \"gbs\": %.1f}\n", N, CF::XCT, CF::XG_X, CF::XROT_X, (int)CF::NT_COL, best,
                  bytes / (best * 1e-3) / 1e9);
    };
    xcol(XcolV<16, 0, 0>{});
    // the other load/store cache policy for each tile width
    xcol(XcolV<16, 0, 0, false>{});
    xcol(XcolV<16, 2, 16, false>{});
    xcol(XcolV<8, 2, 16, true>{});
    xcol(XcolV<8, 1, 32, true>{});
    xcol(XcolV<16, 2, 16>{});
    xcol(XcolV<16, 2, 8>{});
    xcol(XcolV<16, 2, 32>{});
    xcol(XcolV<8, 0, 0>{});
    xcol(XcolV<8, 1, 8>{});
    xcol(XcolV<8, 1, 32>{});
    xcol(XcolV<8, 2, 8>{});
    xcol(XcolV<8, 2, 16>{});
    xcol(XcolV<8, 2, 32>{});
    xcol(XcolV<8, 4, 8>{});
    xcol(XcolV<8, 4, 16>{});
    xcol(XcolV<8, 4, 32>{});
    xcol(XcolV<8, 8, 16>{});
  }
#ifdef JMS_PLANS_FILE
#define JMS_PLAN_CASE(NAME, ...) plan_variant<__VA_ARGS__>(b, NAME, bytes);
  JMS_TUNE_PLANS(JMS_PLAN_CASE)
#undef JMS_PLAN_CASE
#endif
  CK(hipGetLastError());
  return 0;
}
