// Host-side SAH BVH builder and scene digestion for the HIP renderer.
//
// raydar has no acceleration structure: trace_ray scans every object
// (src/renderer/cpu.rs:344-352).  The BVH here must return exactly the scan's
// answer, including phantom hits that the reference's ill-conditioned FP32
// sphere test produces near silhouettes, so primitive boxes are padded by a
// bound on that test's rounding error (see pad_for in bvh.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/raydar_hip.h"
#include "layout.hpp"

namespace rdh {

struct BuiltScene {
  std::vector<DevNode> nodes;
  std::vector<DevNode4> nodes4;    // the same tree, collapsed to 4-wide nodes
  uint32_t depth4 = 0;             // levels of nodes4
  uint32_t stack4 = 0;             // traversal stack entries nodes4 can need
  uint32_t max_leaf = 0;           // primitives in the largest leaf
  std::vector<DevPrim> prims;      // BVH leaf order
  std::vector<DevObject> objects;  // original order
  uint32_t depth = 0;              // max node depth (root = 1)
  float r0 = 0.f;                  // |o|_1 bound covered by the static padding
  double sah_cost = 0.0;
};

// Validates the scene (finite numbers, known kinds) and builds the device
// image.  Returns false with a message in *err on invalid input.
bool build_scene(const rd_scene &scene, int leaf_size, BuiltScene *out, std::string *err);

// Structural invariants of a built BVH (every prim referenced once, child
// boxes contain their prims' padded boxes, depth within the stack bound).
bool check_bvh(const BuiltScene &b, std::string *err);

// Host reference traversal of the device BVH format, used by the CPU tests of
// the builder: same culling rule as the extend kernel, exact hit tests.
int32_t host_trace(const BuiltScene &b, const float o[3], const float d[3], float *t_out);

// The same over the 4-wide nodes (children visited nearest first).
int32_t host_trace4(const BuiltScene &b, const float o[3], const float d[3], float *t_out);

// Structural invariants of nodes4 (as check_bvh) and its stack bound.
bool check_bvh4(const BuiltScene &b, std::string *err);

}  // namespace rdh
