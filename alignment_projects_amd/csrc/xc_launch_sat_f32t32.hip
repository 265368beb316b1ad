// xc_launch_sat_f32t32.hip -- table sweeps: FP32 values, FP32 (S, S^2) entries (tiles that ask for FP32 tables).
#include "xc_launch_sat.hpp"

namespace xc {

template int launch_sats<float, float>(xc_ctx *, hipStream_t, const TileDev *, const std::vector<SatGroup> &);

}  // namespace xc
