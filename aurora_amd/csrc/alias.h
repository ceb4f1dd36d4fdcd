// Environment alias map (host build).
#pragma once

#include <stdint.h>

#include "scene.h"

namespace aph {

// AliasMap::build (Source/AliasMap.cpp:31-166): luminance x row solid angle,
// luminance integral, Vose-style small/large two-cursor pairing.  Reads RGB
// floats with stride 3 (AliasMap.cpp:86).
void aliasBuild(const float* rgb, uint32_t w, uint32_t h, AliasEntry* out, float* integral);

}  // namespace aph
