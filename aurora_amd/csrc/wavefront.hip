// Wavefront path-tracing pipeline for gfx950.
#include "wavefront.h"

namespace aph {

hipError_t wavefrontRender(WavefrontState& st, const WavefrontRequest* reqs, uint32_t nreq, DevCounters* ctr,
                           bool stats, hipStream_t stream, std::vector<PendingEvent>* timing)
{
    (void)st; (void)reqs; (void)nreq; (void)ctr; (void)stats; (void)stream; (void)timing;
    return hipErrorNotSupported;
}

void wavefrontRelease(WavefrontState& st)
{
    if (st.buf) hipFree(st.buf);
    st.buf = nullptr;
    st.bytes = 0;
    st.capacity = 0;
}

}  // namespace aph
