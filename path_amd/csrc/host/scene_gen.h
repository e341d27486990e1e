#pragma once

#include "scene.h"

namespace pt {
namespace host {

struct GeneratedScene {
  Scene scene;
  Camera camera;
  uint32_t resolution = 64;
};

// Deterministic random scene for training-data generation (see scene_gen.cpp).
GeneratedScene generate_scene(uint64_t seed);

} // namespace host
} // namespace pt
