#include "renderer.h"
namespace pt {
namespace host {
int render_streaming(pt_renderer &, const pt_sample_spec &, const gpu::DeviceScene &,
                     uint32_t, const pt_settings &, const gpu::RngParams &, uint32_t,
                     uint32_t, const float **) {
  return set_error(PT_ERR_UNSUPPORTED, "Streaming integrator not built yet");
}
} // namespace host
} // namespace pt
