// Environment alias map, built on the host when the environment is uploaded.
#include "alias.h"

#include <cmath>
#include <vector>

namespace aph {

void aliasBuild(const float* rgb, uint32_t w, uint32_t h, AliasEntry* out, float* integralOut)
{
    const uint32_t n = w * h;
    std::vector<float> lum(n), lumArea(n);
    std::vector<uint32_t> order(n);
    // Row solid angle (sin(top) - sin(bottom)) * 2pi/w, latitude stepped down from pi/2.
    const float lonStep = static_cast<float>(2.0f * M_PI / w);
    const float latStep = static_cast<float>(M_PI / h);
    float lat = static_cast<float>(M_PI_2);
    float integral = 0.0f;
    for (uint32_t y = 0, i = 0; y < h; y++) {
        const float solidAngle = (std::sin(lat) - std::sin(lat - latStep)) * lonStep;
        lat -= latStep;
        for (uint32_t x = 0; x < w; x++, i++) {
            const float* p = rgb + 3 * (size_t)i;
            const float l = p[0] * 0.2125f + p[1] * 0.7154f + p[2] * 0.0721f;
            lum[i] = l;
            lumArea[i] = solidAngle * l;
            integral += lumArea[i];
        }
    }
    const float average = integral / (float)n;
    uint32_t small = 0, large = n;
    for (uint32_t i = 0; i < n; i++) {
        out[i].prob = lumArea[i] / average;
        out[i].pdf = lum[i] / integral;
        out[i].pad = 0.0f;
        out[i].alias = i;
        order[out[i].prob < 1.0f ? small++ : --large] = i;
    }
    // Pair each below-average entry with the current above-average one.
    for (small = 0; small < large && large < n; small++) {
        const uint32_t s = order[small], l = order[large];
        out[s].alias = l;
        out[l].prob -= 1.0f - out[s].prob;
        if (out[l].prob < 1.0f) large++;
    }
    *integralOut = integral;
}

}  // namespace aph
