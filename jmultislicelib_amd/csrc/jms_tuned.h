// jms_tuned.h -- launch shapes per line length chosen by the autotuner
// (csrc/tune/tune.py) from measured kernel times on MI355X.  Generated:
// edit by re-running the tuner.  Fields not given keep the defaults of
// jms_cfg.hpp: LPB (rows per row-pass block; 0 = default), GTW (row-pass
// twiddles from the global table with the product tree), CW (columns per
// wave-local column block), WPE (waves per SIMD asked of the register
// allocator for that kernel), TW (threads per column of that kernel; 0 =
// default), PADC (LDS padding shift of that kernel's tile; 0 = default, >= 20
// none), PADR (the same for the row pass), PANEL (waves of staged batches in
// 16-column panels), TR (row-pass threads per line; 0 = one per butterfly),
// NTC (nontemporal wave loads and stores in the column pass);
// at N = 1024, 2048 CW is the long-column kernel's columns per block (16 or
// 8).  Included by jms_cfg.hpp after the plans.
#pragma once

namespace jms {

// FFT plan (radix order) per length: PlanFor<N> unless the tuner chose another
template <int N>
struct TunedPlan {
  using type = typename PlanFor<N>::type;
};

// FFT plan of the column transforms: the row plan unless the tuner chose another
template <int N>
struct TunedColPlan {
  using type = typename TunedPlan<N>::type;
};

template <int N>
struct Tuned {
  static constexpr int LPB = 0;
  static constexpr bool GTW = false;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};

template <> struct TunedColPlan<360> { using type = Radices<9, 5, 8>; };
template <> struct TunedColPlan<480> { using type = Radices<12, 8, 5>; };
template <> struct TunedColPlan<720> { using type = Radices<12, 12, 5>; };
template <> struct TunedPlan<32> { using type = Radices<4, 8>; };
template <> struct TunedPlan<48> { using type = Radices<6, 8>; };
template <> struct TunedPlan<128> { using type = Radices<4, 8, 4>; };
template <> struct TunedPlan<320> { using type = Radices<5, 8, 8>; };
template <> struct Tuned<16> {
  static constexpr int LPB = 32;
  static constexpr bool GTW = true;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<48> {
  static constexpr int LPB = 64;
  static constexpr bool GTW = false;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<64> {
  static constexpr int LPB = 64;
  static constexpr bool GTW = false;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<96> {
  static constexpr int LPB = 42;
  static constexpr bool GTW = false;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<128> {
  static constexpr int LPB = 16;
  static constexpr bool GTW = true;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = true;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<160> {
  static constexpr int LPB = 16;
  static constexpr bool GTW = true;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<192> {
  static constexpr int LPB = 8;
  static constexpr bool GTW = true;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = true;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<256> {
  static constexpr int LPB = 8;
  static constexpr bool GTW = true;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = true;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<320> {
  static constexpr int LPB = 8;
  static constexpr bool GTW = false;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<360> {
  static constexpr int LPB = 0;
  static constexpr bool GTW = false;
  static constexpr int CW = 8;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 24;
  static constexpr int PADR = 24;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = false;
};
template <> struct Tuned<384> {
  static constexpr int LPB = 8;
  static constexpr bool GTW = false;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = true;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<480> {
  static constexpr int LPB = 4;
  static constexpr bool GTW = false;
  static constexpr int CW = 8;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 24;
  static constexpr int PADR = 24;
  static constexpr bool PANEL = true;
  static constexpr int TR = 64;
  static constexpr bool NTC = false;
};
template <> struct Tuned<512> {
  static constexpr int LPB = 4;
  static constexpr bool GTW = true;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 24;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = true;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<625> {
  static constexpr int LPB = 4;
  static constexpr bool GTW = false;
  static constexpr int CW = 16;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = false;
};
template <> struct Tuned<640> {
  static constexpr int LPB = 3;
  static constexpr bool GTW = true;
  static constexpr int CW = 8;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 24;
  static constexpr bool PANEL = true;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<720> {
  static constexpr int LPB = 2;
  static constexpr bool GTW = true;
  static constexpr int CW = 8;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 24;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<768> {
  static constexpr int LPB = 2;
  static constexpr bool GTW = true;
  static constexpr int CW = 8;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = true;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<1024> {
  static constexpr int LPB = 2;
  static constexpr bool GTW = true;
  static constexpr int CW = 8;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};
template <> struct Tuned<2048> {
  static constexpr int LPB = 1;
  static constexpr bool GTW = true;
  static constexpr int CW = 8;
  static constexpr int WPE = 1;
  static constexpr int TW = 0;
  static constexpr int PADC = 0;
  static constexpr int PADR = 0;
  static constexpr bool PANEL = false;
  static constexpr int TR = 0;
  static constexpr bool NTC = true;
};

}  // namespace jms
