// jms_cfg.hpp -- launch configuration of the line kernels per line length N
//   T   : threads per row (one radix butterfly per thread per stage at most)
//   LPB : rows per block of the row pass
//   C   : adjacent columns per block of the column pass (C*8 bytes per row
//         segment: 128 B for N <= 640)
//   TC  : threads per column in the column pass (block = TC*C <= 1024)
//   PADSH: LDS padding, one complex slot per 2^PADSH elements
#pragma once
#include "jms_fft_device.hpp"

namespace jms {

// per-size launch configuration
template <int N>
struct Cfg {
  using P = typename PlanFor<N>::type;
  static_assert(plan_product<P>() == N, "plan does not multiply to N");
  static constexpr int max_bfly() {
    int m = 0;
    for (int s = 0; s < P::S; ++s) m = (N / P::R[s] > m) ? N / P::R[s] : m;
    return m;
  }
  static constexpr int T = max_bfly();                   // row pass: threads per line
  static constexpr int LPB = (256 / T) > 0 ? (256 / T) : 1;  // rows per block
  static constexpr int C = (N <= 640) ? 16 : ((N <= 1024) ? 8 : 4);  // columns per block
  static constexpr int TC0 = 1024 / C;
  static constexpr int TC1 = (T < TC0) ? T : TC0;
  // column pass: threads per line, rounded down so that a block is a whole
  // number of 64-lane wavefronts (the readout reduction shuffles over 64 lanes)
  static constexpr int TC = (C >= 64) ? TC1 : (TC1 / (64 / C)) * (64 / C);
  static_assert(TC >= 1 && (TC * C) % 64 == 0, "column block must be whole wavefronts");
  // wave-local column kernel: TW threads per column (largest power of two
  // <= min(T, 64), so a column stays inside one wavefront), CW columns per
  // block; used up to N = 768 (the staged tile must fit LDS)
  static constexpr int pow2_floor(int v) { int p = 1; while (2 * p <= v) p *= 2; return p; }
  static constexpr int TW = pow2_floor(T < 64 ? T : 64);
  static constexpr int CW = 16;
  static constexpr bool WL = (N <= 768) && ((TW * CW) % 64 == 0);
  // columns per block of the column kernel in use (partial-sum tiles)
  static constexpr int CC = WL ? CW : C;
  static constexpr int PADSH = 3;
};


}  // namespace jms
