// xc_launch_sat_f64.hip -- table sweeps: FP64 values, FP64 (S, S^2) entries (the plain form).
#include "xc_launch_sat.hpp"

namespace xc {

template int launch_sats<double, double>(xc_ctx *, hipStream_t, const TileDev *, const std::vector<SatGroup> &);

}  // namespace xc
