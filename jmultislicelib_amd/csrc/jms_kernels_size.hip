// jms_kernels_size.hip -- instantiates the line-length templated kernels for
// one length, selected at compile time with -DJMS_N=<length>.
#include "jms_kernels_impl.hpp"

#ifndef JMS_N
#error "compile with -DJMS_N=<line length>"
#endif

namespace jms {
template hipError_t launch_row_pass_n<JMS_N>(int, int, float2*, long long, const float2*, long long,
                                             long long, const int*, int, int, const float2*,
                                             const RowChanHost&, int, bool, hipStream_t);
template hipError_t launch_col_pass_n<JMS_N>(int, bool, bool, const ColArgsHost&, hipStream_t);
template hipError_t launch_row_pass_grp_n<JMS_N>(int, int, long long, long long, const float2*, const GrpArgs&,
                                                 bool, hipStream_t);
template hipError_t launch_col_pass_grp_n<JMS_N>(int, int, bool, bool, bool, long long, const float2*,
                                                 const GrpArgs&, bool, hipStream_t);
template hipError_t launch_row_fft_n<JMS_N>(int, int, float2*, long long, int, const float2*,
                                            hipStream_t);
template hipError_t launch_col_fft_n<JMS_N>(int, int, float2*, long long, int, const float2*,
                                            hipStream_t);
}  // namespace jms
