// Adaptive samples-per-frame for progressive rendering: the SampleCounter of
// Libraries/Foundation/API/Aurora/Foundation/Timer.h:97-281, which Plasma's update loop
// calls each frame (Plasma.cpp:1061) to accumulate as many samples as fit a target frame
// time.  The clock is a template parameter (elapsed() in milliseconds, reset()) so the
// bookkeeping can be driven by a simulated clock in tests.
#pragma once

#include <algorithm>
#include <cassert>
#include <chrono>
#include <cstdint>
#include <vector>

namespace aph_app {

// Milliseconds on the steady clock since the last reset.
class SteadyTimer {
public:
    SteadyTimer() { reset(); }
    void reset() { start_ = std::chrono::steady_clock::now(); }
    float elapsed() const
    {
        return std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - start_).count();
    }

private:
    std::chrono::steady_clock::time_point start_;
};

template <class Timer = SteadyTimer>
class SampleCounter {
public:
    // activeFrameTime: target milliseconds per frame while the view keeps changing;
    // idleFrameTime: target once it has not been restarted for a second; maxSamples: the
    // number of samples at which rendering is complete.
    SampleCounter(float activeFrameTime = 33.3f, float idleFrameTime = 200.0f, uint32_t maxSamples = 1000,
                  Timer timer = Timer(), Timer idleTimer = Timer())
        : timer_(timer), idleTimer_(idleTimer), activeFrameTime_(activeFrameTime), idleFrameTime_(idleFrameTime),
          maxSamples_(maxSamples)
    {
        assert(activeFrameTime > 0.0f && idleFrameTime > 0.0f && maxSamples > 0);
        records_.resize(kTimeCount);
        reset();
    }

    void reset()
    {
        timer_.reset();
        idleTimer_.reset();
        records_[0].samples = 1;
        isFull_ = false;
        index_ = 0;
        sampleStart_ = 0;
        storedSamples_ = 0;
        storedTime_ = 0.0f;
    }

    uint32_t currentSamples() const { return sampleStart_; }
    bool isComplete() const { return sampleStart_ == maxSamples_; }
    void setMaxSamples(uint32_t n) { maxSamples_ = n; }

    // The number of samples for the next frame (0 once complete), starting at sampleStart.
    // The time since the previous call is taken as the previous frame's render time; until
    // the record ring (20 frames) is full, frames get one sample each.
    uint32_t update(uint32_t& sampleStart, bool restart = false)
    {
        uint32_t samples = 1;
        const bool wasDone = isComplete();
        if (restart) {
            sampleStart_ = 0;
            idleTimer_.reset();
            if (wasDone) {
                // after a completed render the elapsed time is idle time, not render time
                timer_.reset();
                samples = sampleCount(activeFrameTime_);
                sampleStart = 0;
                sampleStart_ += samples;
                return samples;
            }
        }
        if (wasDone) {
            sampleStart = maxSamples_;
            return 0;
        }
        const float elapsed = timer_.elapsed();
        timer_.reset();
        Record& prev = records_[index_];
        if (isFull_) storedTime_ -= prev.time;
        prev.time = elapsed;
        storedTime_ += elapsed;
        if (isFull_) {
            const float kIdleDuration = 1000.0f;
            samples = sampleCount(idleTimer_.elapsed() > kIdleDuration ? idleFrameTime_ : activeFrameTime_);
        }
        if (++index_ == kTimeCount) {
            isFull_ = true;
            index_ = 0;
        }
        Record& next = records_[index_];
        if (isFull_) storedSamples_ -= next.samples;
        next.samples = samples;
        storedSamples_ += samples;
        sampleStart = sampleStart_;
        sampleStart_ += samples;
        return samples;
    }

private:
    static constexpr uint32_t kTimeCount = 20;
    struct Record {
        float time = 0.0f;
        uint32_t samples = 0;
    };

    // Samples that fit the target frame time at the recorded average time per sample, at
    // least one, at most what remains (all that remains while no time has been measured,
    // where the reference's float-to-integer conversion would be undefined).
    uint32_t sampleCount(float target) const
    {
        const float average = storedTime_ / (float)storedSamples_;
        if (!(average > 0.0f)) return maxSamples_ - sampleStart_;  // no measurable time yet
        uint32_t n = std::max((uint32_t)(target / average), 1u);
        return std::min(maxSamples_ - sampleStart_, n);
    }

    Timer timer_, idleTimer_;
    std::vector<Record> records_;
    float activeFrameTime_, idleFrameTime_;
    uint32_t maxSamples_;
    bool isFull_ = false;
    uint32_t index_ = 0;
    uint32_t sampleStart_ = 0;
    uint32_t storedSamples_ = 0;
    float storedTime_ = 0.0f;
};

}  // namespace aph_app
