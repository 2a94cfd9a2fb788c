// Entry points of include/hmsnn.h whose implementation is not built yet.
// Each fails with HMSNN_EUNSUPPORTED and a message naming the entry point.
#include "../../include/hmsnn.h"
#include "hm_api_util.h"

#define HM_UNSUPPORTED(name) return hm::fail(HMSNN_EUNSUPPORTED, name " is not implemented yet")

extern "C" {


}  // extern "C"
