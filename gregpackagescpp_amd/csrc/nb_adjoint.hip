// nb_adjoint.hip -- adjoint (placeholder until the adjoint kernels land).
#include "nb_system.hpp"

namespace nbe {
void adjoint_run(nb_system *, uint64_t, uint64_t, const double *, const double *, double *, double *, double *,
                 int64_t *, int64_t *) {
    throw ApiError(NB_ERR_NOT_SUPPORTED, "adjoint not built yet");
}
}  // namespace nbe
