// Host helpers shared by the on-disk scene readers (assets.cpp, gltf.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace aph {

// Foundation::calculateNormals (Geometry.cpp:29-101): angle-weighted vertex normals of
// an indexed triangle mesh (3 floats per vertex, 3 indices per triangle).
void angleWeightedNormals(const std::vector<float>& pos, const std::vector<uint32_t>& idx, std::vector<float>& out);

}  // namespace aph
