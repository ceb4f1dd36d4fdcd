// Drives aurora_amd/apps/sample_counter.h with a simulated clock: each frame renders
// for (per-sample time) x (its sample count) simulated milliseconds.
// usage: sample_counter_test active_ms idle_ms max_samples idle_after_ms restart_at ms_per_sample...
//   one ms_per_sample value per phase of 30 frames; prints "frame start count" lines.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../aurora_amd/apps/sample_counter.h"

namespace {
double g_now = 0.0;  // simulated milliseconds
struct FakeTimer {
    double start = 0.0;
    void reset() { start = g_now; }
    float elapsed() const { return (float)(g_now - start); }
};
}  // namespace

int main(int argc, char** argv)
{
    if (argc < 7) return 2;
    const float active = (float)std::atof(argv[1]), idle = (float)std::atof(argv[2]);
    const uint32_t maxSamples = (uint32_t)std::atoi(argv[3]);
    const double idleAfter = std::atof(argv[4]);  // simulated pause before the next phase
    const int restartAt = std::atoi(argv[5]);     // frame that restarts (-1: none)
    std::vector<double> perSample;
    for (int i = 6; i < argc; i++) perSample.push_back(std::atof(argv[i]));
    aph_app::SampleCounter<FakeTimer> c(active, idle, maxSamples);
    int frame = 0;
    for (size_t ph = 0; ph < perSample.size(); ph++) {
        for (int k = 0; k < 30; k++, frame++) {
            uint32_t start = 0;
            const uint32_t n = c.update(start, frame == restartAt);
            std::printf("%d %u %u\n", frame, start, n);
            g_now += perSample[ph] * n;
        }
        g_now += idleAfter;
    }
    return 0;
}
