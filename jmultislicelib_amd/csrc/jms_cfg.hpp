// jms_cfg.hpp -- launch configuration of the line kernels per line length N
//   T   : threads per row (one radix butterfly per thread per stage at most)
//   LPB : rows per block of the row pass
//   C   : adjacent columns per block of the column pass (C*8 bytes per row
//         segment: 128 B for N <= 640)
//   TC  : threads per column in the column pass (block = TC*C <= 1024)
//   PADSH: LDS padding, one complex slot per 2^PADSH elements
#pragma once
#include "jms_fft_device.hpp"
#include "jms_tuned.h"

namespace jms {

// launch configuration of length N with FFT plans PL for the row transforms
// and PLC for the column transforms (radix orders; the tuned plans of
// jms_tuned.h by default: the two passes are independent 1-D transforms, so
// each may take the plan that suits its kernel)
template <int N, class PL = typename TunedPlan<N>::type, class PLC = typename TunedColPlan<N>::type>
struct CfgP {
  using P = PL;
  using PC = PLC;
  static_assert(plan_product<P>() == N, "row plan does not multiply to N");
  static_assert(plan_product<PC>() == N, "column plan does not multiply to N");
  template <class Q>
  static constexpr int max_bfly_of() {
    int m = 0;
    for (int s = 0; s < Q::S; ++s) m = (N / Q::R[s] > m) ? N / Q::R[s] : m;
    return m;
  }
  // row pass: threads per line, one per butterfly of the widest stage, or
  // (TR, jms_tuned.h) fewer with several butterflies each, e.g. 64 so that a
  // row is one wavefront and its exchanges need wave barriers only
  static constexpr int T = (Tuned<N>::TR > 0 && Tuned<N>::TR < max_bfly_of<P>()) ? Tuned<N>::TR : max_bfly_of<P>();
  static constexpr int TCOL = max_bfly_of<PC>();         // butterflies per column stage (at most)
  // rows per block: 256 threads' worth, or the tuned value (jms_tuned.h)
  // when it fits a block with this plan's threads per row
  static constexpr int LPB = (Tuned<N>::LPB > 0 && T * Tuned<N>::LPB <= 1024) ? Tuned<N>::LPB
                                                                               : ((256 / T) > 0 ? (256 / T) : 1);
  static constexpr int C = (N <= 640) ? 16 : ((N <= 1024) ? 8 : 4);  // columns per block
  static constexpr int TC0 = 1024 / C;
  static constexpr int TC1 = (TCOL < TC0) ? TCOL : TC0;
  // column pass: threads per line, rounded down so that a block is a whole
  // number of 64-lane wavefronts (the readout reduction shuffles over 64 lanes)
  static constexpr int TC = (C >= 64) ? TC1 : (TC1 / (64 / C)) * (64 / C);
  static_assert(TC >= 1 && (TC * C) % 64 == 0, "column block must be whole wavefronts");
  // wave-local column kernel: TW threads per column (largest power of two
  // <= min(T, 64), so a column stays inside one wavefront, unless the tuner
  // chose fewer), CW columns per
  // block; used up to N = 768 (the staged tile must fit LDS)
  static constexpr int pow2_floor(int v) { int p = 1; while (2 * p <= v) p *= 2; return p; }
  static constexpr int TW = Tuned<N>::TW > 0 ? Tuned<N>::TW : pow2_floor(TCOL < 64 ? TCOL : 64);
  // 16 columns (128-byte segments) unless the tuner found 8 faster (where a
  // 16-column block cannot share its CU)
  static constexpr int CW = Tuned<N>::CW;
  static constexpr bool WL = (N <= 768) && ((TW * CW) % 64 == 0);
  // long columns (N = 1024, 2048): register-resident column kernel with 16
  // columns per block (128-byte row segments; a 2048^2 column-tile copy runs
  // at 0.93 / 2.6 / 5.1 TB/s with 4 / 8 / 16 columns per block on MI355X in
  // the hardware block order), or 8 (CW = 8, jms_tuned.h) in blocks of 512
  // threads, two per CU, ordered so that the neighbouring tiles sharing a
  // 128-byte line run on one XCD at the same time
  static constexpr bool XL = (N == 1024 || N == 2048);
  static constexpr int XCT = (XL && CW == 8) ? 8 : 16;
  // block order of the long-column kernel: XG_X > 0 takes block_xy_groups
  // with groups of XG_X waves rotated by XROT_X tiles, 0 the hardware order.
  // Per 1 GiB of waves on MI355X (tuner, nontemporal loads and stores):
  // N = 1024 16 columns 0.706 ms, 8 columns in the hardware order 0.793 ms,
  // in groups of 1 / 2 / 4 waves 0.707 / 0.597 / 0.618 ms; N = 2048 16
  // columns 0.706 ms, 8 columns 0.70-1.05 ms.  With ordinary loads and
  // stores (NT_COL below) 8 columns in groups of 2 waves take 0.582 ms at
  // 1024 and 0.636 ms at 2048, which both lengths use
  static constexpr int XG_X = XCT == 8 ? 2 : 0;
  static constexpr int XROT_X = XCT == 8 ? 16 : 0;
  // columns per block of the column kernel in use (partial-sum tiles)
  static constexpr int CC = XL ? XCT : (WL ? CW : C);
  static constexpr int PADSH = 3;
  // LDS padding of the wave-local column kernel's tile (a shift >= 20: no
  // padding, every exchange address a constant offset from one base)
  static constexpr int PADSH_COL = Tuned<N>::PADC > 0 ? Tuned<N>::PADC : PADSH;
  // the same for the row pass's line exchanges
  static constexpr int PADSH_ROW = Tuned<N>::PADR > 0 ? Tuned<N>::PADR : PADSH;
  // waves of staged batches stored in panels of 16 columns (see panel_off):
  // contiguous column tiles, rows in 128-byte segments; where the tuner found
  // the pair of passes faster that way
  static constexpr bool PANEL = Tuned<N>::PANEL && (N % 16 == 0);
  // minimum waves per SIMD requested from the register allocator for the
  // wave-local column kernel (1 = no constraint; jms_tuned.h).  Eight waves
  // let two 1024-thread blocks share a CU, so one block's FFTs overlap the
  // other's memory phase (N = 512: 1.35 -> 1.11 ms per 1024 waves on MI355X).
  static constexpr int WPE_COL = Tuned<N>::WPE;
  static constexpr int WPE_ROW = 1;

  // nontemporal wave loads/stores in the row / column pass: the streamed
  // waves then leave the reused phase gratings and propagators in L2
  // (row pass 0.84 -> 0.78 ms per 1024 waves at N = 512)
  static constexpr bool NT_ROW = N >= 256;
  // except the long-column kernel with 8-column tiles: the neighbouring
  // tile reads the other half of each 128-byte line a little later, and a
  // nontemporal line is gone by then (2048 columns: 0.746 -> 0.636 ms per
  // GiB of waves on MI355X; 1024: 0.597 -> 0.582)
  // and where the tuner found ordinary ones faster (NTC, jms_tuned.h: 480
  // columns in 8-column half-panels 0.543 -> 0.482 ms per GiB of waves)
  static constexpr bool NT_COL = N >= 256 && Tuned<N>::NTC && !(XL && XCT == 8);
  // LDS position policy of line layouts (0 padded, 1 swizzled) and twiddle
  // powers by product tree (one table read per butterfly) instead of one
  // table read per leg; measured at N = 512 (2048 waves, interleaved runs):
  // column pass 2.54 -> 2.21 ms with it, row pass 1.60 -> 1.63 ms, so the
  // row pass keeps the table reads
  static constexpr int SW = 0;
  static constexpr bool TWR = true;
  // Where the tuner chose it (jms_tuned.h; long rows): the row pass reads its
  // twiddles through the caches from the global table (GTW_ROW) instead of
  // an LDS copy that every block fills (16 KiB per 16 KiB row at N = 2048),
  // with one read per butterfly and the product tree (TWR_ROW); per 1 GiB of
  // waves on MI355X: N = 2048 0.535 -> 0.392 ms, N = 1024 0.417 -> 0.365 ms
  static constexpr bool TWR_ROW = Tuned<N>::GTW;
  static constexpr bool GTW_ROW = Tuned<N>::GTW;
  // fused tile I/O in the wave-local column kernel where the plan allows it
  static constexpr bool FIO = true;
  // XCD-aware block order (block_xy) in the row pass: where a slice's
  // grating (8 n bytes, 32 MiB at N = 2048) is far larger than one XCD's
  // 4 MiB L2, the waves sharing a grating row then run on one XCD back to
  // back (2048^2, 32 waves: 0.453 -> 0.427 ms per launch on MI355X).  Off in
  // the long-column kernel: there the blocks in flight on an XCD would all
  // read the same column tile of different waves, addresses a whole wave
  // (32 MiB) apart, and the launch slowed from 0.755 to 0.819 ms.
  static constexpr bool XCD_ROW = N >= 1024;
  static constexpr bool XCD_COL = false;
};

// the configuration the library is built with
template <int N>
struct Cfg : CfgP<N> {};


}  // namespace jms
