// xc_launch_sat_x.hip -- table sweeps: exact integer tables of unmasked u8 tiles (FP32 transforms).
#include "xc_launch_sat.hpp"

namespace xc {

template int launch_sats<float, TabX>(xc_ctx *, hipStream_t, const TileDev *, const std::vector<SatGroup> &);

}  // namespace xc
