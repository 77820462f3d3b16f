// Level operator instantiations for the "weights" cell-matrix kind (see apply_impl.h).
#include "apply_impl.h"

namespace lp {

void launch_apply_weights(const Grid& G, int level, int eb, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                     hipStream_t s) {
    if (eb == 8)
        apply_detail::launch_kind<double, LK_WEIGHTS>(G, level, static_cast<const double*>(u), vstride, nrhs, E, s);
    else
        apply_detail::launch_kind<float, LK_WEIGHTS>(G, level, static_cast<const float*>(u), vstride, nrhs, E, s);
}

}  // namespace lp
