// nb_coll.hip -- collision pass (placeholder until the collision kernels land).
#include "nb_system.hpp"

namespace nbe {
struct CollWork {};
void coll_work_free(CollWork *w) { delete w; }
void collision_pass(nb_system *, uint64_t) { throw ApiError(NB_ERR_NOT_SUPPORTED, "collisions not built yet"); }
}  // namespace nbe

struct nb_batch {
    std::string err;
};
extern "C" {
int nb_batch_create(int, int64_t, const int64_t *, const double *, const double *, const double *, const double *,
                    const double *, nb_batch **out) {
    if (out) *out = nullptr;
    return NB_ERR_NOT_SUPPORTED;
}
int nb_batch_destroy(nb_batch *b) { delete b; return NB_OK; }
const char *nb_batch_last_error(const nb_batch *) { return "batched ensembles not built yet"; }
int nb_batch_set_params(nb_batch *, double, double, double, int, int, double, int) { return NB_ERR_NOT_SUPPORTED; }
int nb_batch_evolve(nb_batch *, uint64_t, int64_t *) { return NB_ERR_NOT_SUPPORTED; }
int nb_batch_get_state(const nb_batch *, int64_t, int64_t *, double *, double *, double *, double *, int64_t *) {
    return NB_ERR_NOT_SUPPORTED;
}
int nb_batch_merge_log(const nb_batch *, int64_t, int64_t *, int64_t *) { return NB_ERR_NOT_SUPPORTED; }
}
