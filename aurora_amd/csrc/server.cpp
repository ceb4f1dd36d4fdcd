// Serving loop: continuous batching of render jobs over device-resident scenes.
//
// No reference counterpart: Aurora renders one request at a time, sample by
// sample (HGIRenderer::render, HGIRenderer.cpp:477-589; progressive
// render(sampleStart, sampleCount) driven by SampleCounter, Timer.h:97-286).
// Here jobs of uneven resolution and spp arrive against TLASes already built
// in the context.  Each step
//   1. admits waiting jobs in arrival order while a slot is free and the HBM
//      budget allows, allocating the job's RGBA32F accumulation buffer.  The
//      budget covers the resident scenes (BVHs, meshes, instance records, their
//      environments), the images bound to the material table, every admitted
//      job's accumulation buffer and the wavefront working set (path state,
//      queues and tables of the paths in flight), whose chunk is lowered from
//      the context's tuned one until it fits;
//   2. gives every admitted job its next run of sample indices - an equal
//      number per job, sized so the whole batch is about step_samples
//      pixel-samples - and renders them all in one shared wavefront;
//   3. completes jobs whose last sample is in; without keep_results their
//      buffers go back to the budget at once.
// Jobs arriving during a step join at the next step boundary, so new work
// overlaps the tail of older jobs on the device.  A job's image depends only
// on its own (scene, frame, sample range): the wavefront accumulates each
// pixel in sample order whatever the batch, so a job renders the same alone
// and in any mix.  APH_MODE_MEGAKERNEL serves the plain way instead: one job
// at a time, all of its samples in one megakernel launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <thread>

#include "context.h"

using namespace aph;

namespace {

struct Job {
    uint64_t ticket = 0;
    aph_job desc{};
    Tlas* tlas = nullptr;
    DevBuf accum;
    size_t bytes = 0;       // accumulation buffer charged to the budget
    uint32_t done = 0;      // samples accumulated
    int state = APH_JOB_PENDING;
    double tArrival = 0.0, tAdmit = 0.0, tDone = 0.0;
};

}  // namespace

struct aph_server {
    aph_ctx* ctx = nullptr;
    aph_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    uint64_t nextTicket = 1;
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<Job*> pending;
    std::vector<Job*> active;
    std::set<const Tlas*> resident;  // scenes charged to the budget
    size_t sceneBytes = 0, jobBytes = 0;
    size_t residentBytes = 0;        // scenes + material images, refreshed at each admission
    size_t wfBytes = 0;              // wavefront working set of the last batched step
    uint32_t wfPaths = 0;            // its paths in flight
    uint64_t samplesRendered = 0;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

constexpr uint32_t kDefaultStepSamples = 8u << 20;
constexpr uint32_t kMaxQuantum = 1024;  // samples of one job per step (bounds the step's tables)

size_t accumBytes(const aph_job& j) { return (size_t)j.width * j.height * sizeof(float4); }

uint32_t pixels(const Job* j) { return j->desc.width * j->desc.height; }

// Device bytes of the images bound to the context's material table, each counted once.
size_t textureBytes(const aph_ctx* c)
{
    std::set<uint64_t> seen;
    size_t bytes = 0;
    for (const aph_texture_binding& b : c->texBindings) {
        if (!b.image || !seen.insert(b.image).second) continue;
        auto it = c->images.find(b.image);
        if (it != c->images.end()) bytes += it->second->texels.bytes;
    }
    return bytes;
}

// Upper bound of a step's call table: one request record and up to kMaxQuantum
// sample segments (and chunk records) per active job.
size_t tableReserve(size_t jobs)
{
    return ((size_t)1 << 20) + jobs * (sizeof(WavefrontRequest) + (size_t)kMaxQuantum * 48);
}

// Wavefront working set for `paths` paths in flight and `jobs` requests.
size_t workingSet(uint32_t paths, size_t jobs) { return wavefrontBufferBytes(paths) + tableReserve(jobs); }

// Paths in flight of a batched step: the context's tuned chunk, halved (down to
// the largest active frame, which one chunk must hold) until the working set
// fits what the budget leaves after scenes and job buffers.
uint32_t stepPaths(const aph_server* s, uint32_t maxPix)
{
    uint32_t cap = std::max(s->ctx->wf.tuning.chunkPaths, maxPix);
    const size_t budget = s->cfg.hbm_budget;
    if (!budget) return cap;
    const size_t used = s->residentBytes + s->jobBytes;
    const size_t avail = budget > used ? budget - used : 0;
    while (cap > maxPix && workingSet(cap, s->active.size()) > avail) cap = std::max(maxPix, cap / 2);
    return cap;
}

// Admits waiting jobs in arrival order.  Returns APH_OK or an error for a job
// that can never fit.
int admit(aph_server* s)
{
    const bool plain = s->cfg.mode == APH_MODE_MEGAKERNEL;
    const size_t slots = plain ? 1 : (s->cfg.max_active ? s->cfg.max_active : (size_t)-1);
    s->residentBytes = s->sceneBytes + textureBytes(s->ctx);
    uint32_t maxPix = 0;
    for (const Job* j : s->active) maxPix = std::max(maxPix, pixels(j));
    while (!s->pending.empty() && s->active.size() < slots) {
        Job* j = s->pending.front();
        size_t need = accumBytes(j->desc);
        // batched jobs also need the smallest working set that holds one frame of every active job
        const uint32_t pix = std::max(maxPix, pixels(j));
        const size_t wf = plain ? 0 : workingSet(pix, s->active.size() + 1);
        size_t budget = s->cfg.hbm_budget;
        if (budget && s->residentBytes + s->jobBytes + need + wf > budget) {
            if (s->active.empty() && s->jobBytes == 0) {
                s->pending.pop_front();
                j->state = APH_JOB_FAILED;
                return fail(s->ctx, APH_E_MEMORY,
                            "job %llu needs %zu bytes (image) + %zu (working set) beyond the resident %zu; budget %zu",
                            (unsigned long long)j->ticket, need, wf, s->residentBytes, budget);
            }
            break;  // wait for running jobs to return memory
        }
        hipError_t e = j->accum.alloc(need);
        if (e == hipErrorOutOfMemory && !s->active.empty()) break;
        if (e != hipSuccess) return hipFail(s->ctx, e, "job accumulation buffer");
        j->bytes = need;
        s->jobBytes += need;
        j->state = APH_JOB_ACTIVE;
        j->tAdmit = s->now();
        s->pending.pop_front();
        s->active.push_back(j);
        maxPix = pix;
    }
    return APH_OK;
}

void finish(aph_server* s, Job* j)
{
    j->state = APH_JOB_DONE;
    j->tDone = s->now();
    if (!s->cfg.keep_results) {
        j->accum.reset();
        s->jobBytes -= j->bytes;
        j->bytes = 0;
    }
}

int stepOnce(aph_server* s)
{
    aph_ctx* c = s->ctx;
    int rc = admit(s);
    if (rc) return rc;
    if (s->active.empty()) return APH_OK;
    if (s->cfg.mode == APH_MODE_MEGAKERNEL) {
        Job* j = s->active.front();
        RenderScene rs = renderScene(c, j->tlas);
        aph_frame fp = clampFrame(j->desc.frame);
        HIPCALL(c, launchRenderMega(rs, fp, j->desc.width, j->desc.height, 0, j->desc.spp, j->desc.seed_offset,
                                    j->accum.as<float4>(), c->counters.as<DevCounters>(), c->statsLevel > 0,
                                    c->stream),
                "megakernel launch");
        HIPCALL(c, hipStreamSynchronize(c->stream), "serving step");
        j->done = j->desc.spp;
        s->samplesRendered += (uint64_t)j->desc.width * j->desc.height * j->desc.spp;
        finish(s, j);
        s->active.erase(s->active.begin());
        return APH_OK;
    }
    uint64_t pix = 0;
    uint32_t maxPix = 0;
    for (Job* j : s->active) {
        pix += pixels(j);
        maxPix = std::max(maxPix, pixels(j));
    }
    uint64_t stepSamples = s->cfg.step_samples ? s->cfg.step_samples : kDefaultStepSamples;
    uint32_t quantum = (uint32_t)std::min<uint64_t>(
        kMaxQuantum, std::max<uint64_t>(1, stepSamples / std::max<uint64_t>(pix, 1)));
    std::vector<WavefrontRequest> batch;
    std::vector<uint32_t> counts;
    batch.reserve(s->active.size());
    for (Job* j : s->active) {
        if ((rc = ensureTransmit(c, j->tlas))) return rc;
        uint32_t n = std::min(quantum, j->desc.spp - j->done);
        WavefrontRequest w;
        w.rs = renderScene(c, j->tlas);
        w.frame = clampFrame(j->desc.frame);
        w.width = j->desc.width;
        w.height = j->desc.height;
        w.sampleStart = j->done;
        w.sampleCount = n;
        w.seedOffset = j->desc.seed_offset;
        w.accum = j->accum.as<float4>();
        batch.push_back(w);
        counts.push_back(n);
    }
    const uint32_t paths = stepPaths(s, maxPix);
    const size_t wf = wavefrontBufferBytes(paths) +
                      std::max(wavefrontCallTableBytes(batch.data(), (uint32_t)batch.size(), paths), (size_t)1 << 20);
    hipError_t e = wavefrontRender(c->wf, batch.data(), (uint32_t)batch.size(), c->counters.as<DevCounters>(),
                                   c->statsLevel > 0, c->stream, c->timing ? &c->pending : nullptr, paths);
    if (e != hipSuccess) return hipFail(c, e, "serving step");
    s->wfPaths = paths;
    s->wfBytes = std::max(wf, c->wf.bytes + c->wf.tblBytes);
    HIPCALL(c, hipStreamSynchronize(c->stream), "serving step");
    std::vector<Job*> still;
    for (size_t k = 0; k < s->active.size(); k++) {
        Job* j = s->active[k];
        j->done += counts[k];
        s->samplesRendered += (uint64_t)j->desc.width * j->desc.height * counts[k];
        if (j->done >= j->desc.spp) finish(s, j);
        else still.push_back(j);
    }
    s->active.swap(still);
    return APH_OK;
}

Job* findJob(aph_server* s, uint64_t ticket)
{
    auto it = s->jobs.find(ticket);
    return it == s->jobs.end() ? nullptr : it->second.get();
}

}  // namespace

extern "C" {

int aph_server_create(aph_ctx_t c, const aph_server_config* cfg, aph_server_t* out)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!cfg || !out) return fail(c, APH_E_INVALID, "null config or output");
    if (cfg->mode != APH_MODE_WAVEFRONT && cfg->mode != APH_MODE_MEGAKERNEL)
        return fail(c, APH_E_INVALID, "unknown mode %d", cfg->mode);
    std::unique_ptr<aph_server> s(new aph_server());
    s->ctx = c;
    s->cfg = *cfg;
    s->t0 = std::chrono::steady_clock::now();
    *out = s.release();
    return APH_OK;
}

void aph_server_destroy(aph_server_t s)
{
    if (!s) return;
    if (enter(s->ctx) == APH_OK) hipStreamSynchronize(s->ctx->stream);
    delete s;
}

double aph_server_clock(aph_server_t s) { return s ? s->now() : 0.0; }

int aph_server_submit(aph_server_t s, const aph_job* job, double tArrival, uint64_t* ticket)
{
    if (!s) return APH_E_INVALID;
    aph_ctx* c = s->ctx;
    int rc = enter(c);
    if (rc) return rc;
    if (!job || !ticket) return fail(c, APH_E_INVALID, "null job or ticket");
    Tlas* tl = lookup(c->tlases, job->tlas);
    if (!tl) return fail(c, APH_E_INVALID, "job: unknown tlas %llu", (unsigned long long)job->tlas);
    if (job->width == 0 || job->height == 0 || job->spp == 0)
        return fail(c, APH_E_INVALID, "job: width, height and spp must be positive");
    if ((uint64_t)job->width * job->height > (1ull << 28)) return fail(c, APH_E_INVALID, "job: image too large");
    if (job->frame.light_count < 0 || job->frame.light_count > 4)
        return fail(c, APH_E_INVALID, "job: light_count must be in [0,4]");
    if (!s->resident.count(tl)) {
        // BVHs, instance and shading records, and the scene's own environment images
        const size_t bytes = tl->bytes + (tl->environment ? tl->environment->bytes() : 0);
        if (s->cfg.hbm_budget && s->sceneBytes + textureBytes(c) + bytes > s->cfg.hbm_budget)
            return fail(c, APH_E_MEMORY, "scene of %zu bytes does not fit the budget", bytes);
        s->resident.insert(tl);
        s->sceneBytes += bytes;
        s->residentBytes = s->sceneBytes + textureBytes(c);
    }
    std::unique_ptr<Job> j(new Job());
    j->ticket = s->nextTicket++;
    j->desc = *job;
    j->tlas = tl;
    j->tArrival = tArrival >= 0.0 ? tArrival : s->now();
    s->pending.push_back(j.get());
    *ticket = j->ticket;
    s->jobs[j->ticket] = std::move(j);
    return APH_OK;
}

int aph_server_step(aph_server_t s, uint32_t* active, uint32_t* pending)
{
    if (!s) return APH_E_INVALID;
    int rc = enter(s->ctx);
    if (rc) return rc;
    if ((rc = checkTextures(s->ctx))) return rc;
    rc = stepOnce(s);
    if (active) *active = (uint32_t)s->active.size();
    if (pending) *pending = (uint32_t)s->pending.size();
    return rc;
}

int aph_server_status(aph_server_t s, uint64_t ticket, aph_job_status* st)
{
    if (!s || !st) return APH_E_INVALID;
    Job* j = findJob(s, ticket);
    if (!j) return fail(s->ctx, APH_E_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
    st->state = j->state;
    st->samples_done = j->done;
    st->t_arrival = j->tArrival;
    st->t_admit = j->tAdmit;
    st->t_done = j->tDone;
    return APH_OK;
}

int aph_server_result(aph_server_t s, uint64_t ticket, float* dst, size_t bytes)
{
    if (!s) return APH_E_INVALID;
    aph_ctx* c = s->ctx;
    int rc = enter(c);
    if (rc) return rc;
    Job* j = findJob(s, ticket);
    if (!j || !dst) return fail(c, APH_E_INVALID, "unknown ticket or null destination");
    if (j->state != APH_JOB_DONE) return fail(c, APH_E_STATE, "job %llu is not finished", (unsigned long long)ticket);
    if (!j->accum.p) return fail(c, APH_E_STATE, "job %llu kept no result (keep_results is 0)",
                                 (unsigned long long)ticket);
    size_t need = accumBytes(j->desc);
    if (bytes < need) return fail(c, APH_E_INVALID, "destination too small (%zu < %zu)", bytes, need);
    HIPCALL(c, hipMemcpy(dst, j->accum.p, need, hipMemcpyDeviceToHost), "job result");
    return APH_OK;
}

int aph_server_usage(aph_server_t s, uint64_t* residentBytes, uint64_t* jobBytes, uint32_t* active,
                     uint32_t* pending)
{
    if (!s) return APH_E_INVALID;
    if (residentBytes) *residentBytes = s->residentBytes;
    if (jobBytes) *jobBytes = s->jobBytes;
    if (active) *active = (uint32_t)s->active.size();
    if (pending) *pending = (uint32_t)s->pending.size();
    return APH_OK;
}

int aph_server_working_set(aph_server_t s, uint64_t* wavefront_bytes, uint32_t* paths_in_flight)
{
    if (!s) return APH_E_INVALID;
    if (wavefront_bytes) *wavefront_bytes = s->wfBytes;
    if (paths_in_flight) *paths_in_flight = s->wfPaths;
    return APH_OK;
}

int aph_server_release(aph_server_t s, uint64_t ticket)
{
    if (!s) return APH_E_INVALID;
    int rc = enter(s->ctx);
    if (rc) return rc;
    Job* j = findJob(s, ticket);
    if (!j) return fail(s->ctx, APH_E_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
    if (j->state == APH_JOB_ACTIVE) return fail(s->ctx, APH_E_STATE, "job %llu is running", (unsigned long long)ticket);
    if (j->state == APH_JOB_PENDING) s->pending.erase(std::find(s->pending.begin(), s->pending.end(), j));
    s->jobBytes -= j->bytes;
    s->jobs.erase(ticket);
    return APH_OK;
}

int aph_server_replay(aph_server_t s, const aph_job* jobs, const double* arrival, uint32_t n, uint64_t* tickets)
{
    if (!s) return APH_E_INVALID;
    aph_ctx* c = s->ctx;
    int rc = enter(c);
    if (rc) return rc;
    if (n && (!jobs || !arrival || !tickets)) return fail(c, APH_E_INVALID, "null jobs, arrivals or tickets");
    for (uint32_t i = 1; i < n; i++)
        if (arrival[i] < arrival[i - 1]) return fail(c, APH_E_INVALID, "arrival times must not decrease");
    const double start = s->now();
    uint32_t next = 0;
    while (true) {
        double t = s->now() - start;
        while (next < n && arrival[next] <= t) {
            rc = aph_server_submit(s, &jobs[next], start + arrival[next], &tickets[next]);
            if (rc) return rc;
            next++;
        }
        if (s->active.empty() && s->pending.empty()) {
            if (next == n) break;
            double wait = arrival[next] - (s->now() - start);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            continue;
        }
        rc = stepOnce(s);
        if (rc) return rc;
    }
    return APH_OK;
}

}  // extern "C"
