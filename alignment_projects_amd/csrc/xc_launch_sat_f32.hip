// xc_launch_sat_f32.hip -- table sweeps: FP32 values, FP64 (S, S^2) entries (the batched form).
#include "xc_launch_sat.hpp"

namespace xc {

template int launch_sats<float, double>(xc_ctx *, hipStream_t, const TileDev *, const std::vector<SatGroup> &);

}  // namespace xc
