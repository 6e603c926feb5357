// nb_bh.hip -- Barnes-Hut octree (placeholder until the tree kernels land).
#include "nb_system.hpp"

namespace nbe {
struct BHWork {};
void bh_work_free(BHWork *w) { delete w; }
void bh_forces(nb_system *, int64_t *) { throw ApiError(NB_ERR_NOT_SUPPORTED, "Barnes-Hut not built yet"); }
void bh_steps(nb_system *, uint64_t) { throw ApiError(NB_ERR_NOT_SUPPORTED, "Barnes-Hut not built yet"); }
int bh_topology(nb_system *, int64_t, int32_t *, int32_t *, int64_t *, int64_t *, double *, double *, double *,
                int64_t *) {
    throw ApiError(NB_ERR_NOT_SUPPORTED, "Barnes-Hut not built yet");
}
}  // namespace nbe
