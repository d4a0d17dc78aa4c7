// ls_fast.hip -- fast-mode kernels (placeholder until implemented).
#include "ls_fast.hpp"
#include "ls_internal.hpp"

namespace lsf {
template <int DIM>
int launch_stage_fast(const lsx::Geom&, int, const double*, const double*, double*, const double*,
                      const double*, const double*, const double*, double, hipStream_t) {
    return lsi::set_error(LS_ERR_INVALID, "fast mode is not available");
}
template int launch_stage_fast<2>(const lsx::Geom&, int, const double*, const double*, double*,
                                  const double*, const double*, const double*, const double*, double,
                                  hipStream_t);
template int launch_stage_fast<3>(const lsx::Geom&, int, const double*, const double*, double*,
                                  const double*, const double*, const double*, const double*, double,
                                  hipStream_t);
int launch_hcr2_fast(const lsx::Geom&, const double*, const double*, double*, hipStream_t) {
    return lsi::set_error(LS_ERR_INVALID, "fast mode is not available");
}
}  // namespace lsf
