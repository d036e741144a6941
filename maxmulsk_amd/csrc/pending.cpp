// Temporary: entry points whose subsystems are not built yet.
#include "maxmulsk.h"
#include "plan.hpp"
#include "runtime.hpp"
using namespace mmsk;
namespace mmsk {
bool selector_pick(size_t, size_t, size_t, int, maxmulsk_plan*) { return false; }
}
#define NB(...) return fail(MAXMULSK_ERR_UNSUPPORTED, "not built yet")
extern "C" {

size_t maxmulsk_mlp_num_params(void) { return 0; }
maxmulsk_status maxmulsk_mlp_forward(const float*, const float*, int, float*, float*, float*, void*) { NB(); }
maxmulsk_status maxmulsk_mlp_backward(const float*, const float*, int, const float*, const float*, const float*, const float*, const float*, float*, float*, void*) { NB(); }
maxmulsk_status maxmulsk_mlp_adam(float*, const float*, float*, float*, int, float, float, float, float, void*) { NB(); }
maxmulsk_status maxmulsk_selector_set(const float*) { NB(); }
maxmulsk_status maxmulsk_selector_predict(size_t, size_t, size_t, maxmulsk_mode, float*, maxmulsk_plan*) { NB(); }
}
