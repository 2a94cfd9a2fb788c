// Entry points of include/hmsnn.h whose implementation is not built yet.
// Each fails with HMSNN_EUNSUPPORTED and a message naming the entry point.
#include "../../include/hmsnn.h"
#include "hm_api_util.h"

namespace hm {
struct MsnnState;
void msnn_release(MsnnState*) {}
}  // namespace hm

#define HM_UNSUPPORTED(name) return hm::fail(HMSNN_EUNSUPPORTED, name " is not implemented yet")

extern "C" {

int hmsnn_msnn_begin(hmsnn_ctx*, int) { HM_UNSUPPORTED("hmsnn_msnn_begin"); }
int hmsnn_msnn_frames(hmsnn_ctx*, int, hmsnn_msnn_stats*) { HM_UNSUPPORTED("hmsnn_msnn_frames"); }
int hmsnn_msnn_get_image(hmsnn_ctx*, int, float*) { HM_UNSUPPORTED("hmsnn_msnn_get_image"); }
int hmsnn_render_hair_msnn(hmsnn_ctx*, int, int, float*, hmsnn_msnn_stats*) { HM_UNSUPPORTED("hmsnn_render_hair_msnn"); }
int hmsnn_mlp_create(int, uint64_t, hmsnn_mlp**) { HM_UNSUPPORTED("hmsnn_mlp_create"); }
int hmsnn_mlp_destroy(hmsnn_mlp*) { return HMSNN_OK; }
int hmsnn_mlp_get_weights(const hmsnn_mlp*, float*, int) { HM_UNSUPPORTED("hmsnn_mlp_get_weights"); }
int hmsnn_mlp_set_weights(hmsnn_mlp*, const float*, int) { HM_UNSUPPORTED("hmsnn_mlp_set_weights"); }
int hmsnn_mlp_encode_dev(const float*, int, float*, void*) { HM_UNSUPPORTED("hmsnn_mlp_encode_dev"); }
int hmsnn_mlp_forward_dev(hmsnn_mlp*, const float*, int, float*, void*) { HM_UNSUPPORTED("hmsnn_mlp_forward_dev"); }
int hmsnn_mlp_backward_dev(hmsnn_mlp*, const float*, const float*, int, int, float*, float*, void*) {
  HM_UNSUPPORTED("hmsnn_mlp_backward_dev");
}
int hmsnn_mlp_train_step_dev(hmsnn_mlp*, const float*, const float*, int, int, float, float*, void*) {
  HM_UNSUPPORTED("hmsnn_mlp_train_step_dev");
}
int hmsnn_adjoint_render(hmsnn_ctx*, int, const float*, double*, double[3]) {
  HM_UNSUPPORTED("hmsnn_adjoint_render");
}

}  // extern "C"
