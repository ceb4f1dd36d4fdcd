// HIP-event timing of kernel launches on a stream (host only).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace aph {

struct PendingEvent {
    std::string name;
    hipEvent_t a, b;
};

// Records an event pair around the launches made during its lifetime when
// `sink` is non-null; the pair is resolved later by the context.
struct EventScope {
    std::vector<PendingEvent>* sink;
    const char* name;
    hipStream_t stream;
    hipEvent_t a = nullptr, b = nullptr;
    EventScope(std::vector<PendingEvent>* s, const char* n, hipStream_t st) : sink(s), name(n), stream(st)
    {
        if (sink) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, stream);
        }
    }
    ~EventScope()
    {
        if (sink) {
            hipEventRecord(b, stream);
            sink->push_back(PendingEvent{name, a, b});
        }
    }
};

}  // namespace aph
