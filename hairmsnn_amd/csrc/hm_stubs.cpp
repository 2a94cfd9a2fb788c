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
int hmsnn_adjoint_render(hmsnn_ctx*, int, const float*, double*, double[3]) {
  HM_UNSUPPORTED("hmsnn_adjoint_render");
}

}  // extern "C"
