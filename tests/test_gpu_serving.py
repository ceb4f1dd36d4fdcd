"""Serving loop (aph_server_*): continuous batching with HBM-budget admission.

Bar: a job's image is bit-identical whether it is served alone or in a mix of
other jobs (other scenes, resolutions and spp, arriving while it runs), and
identical to a plain aph_render of the same sample range; admission never
exceeds the slot count or the memory budget; the plain (megakernel) server
runs one job at a time and agrees with the batched server to rounding.
"""
import numpy as np
import pytest

from aurora_amd import capi, scenes
from aurora_amd.session import AphError, Server

pytestmark = pytest.mark.gpu


@pytest.fixture
def two_scenes(renderer, teapot_path):
    a = scenes.teapot_scene(teapot_path, 40, 32)
    b = scenes.config1(width=48, height=40, trace_depth=3)
    return (a, renderer.load_scene(a)), (b, renderer.load_scene(b, capi.APH_BUILDER_SAH))


def plain_render(renderer, tlas, frame, w, h, spp):
    tgt = renderer.target(w, h)
    renderer.render([renderer.request(tgt, tlas, frame, 0, spp)], capi.APH_MODE_WAVEFRONT)
    renderer.wait()
    return renderer.readback(tgt, w, h)


def test_job_alone_equals_job_in_mix(renderer, two_scenes):
    (a, ta), (b, tb) = two_scenes
    ref = plain_render(renderer, ta, a.frame, 40, 32, 7)
    alone = Server(renderer, step_samples=40 * 32 * 2)  # two samples per step: several steps
    t = alone.submit(Server.job(ta, a.frame, 40, 32, 7))
    alone.run()
    np.testing.assert_array_equal(alone.result(t, 40, 32), ref)
    mix = Server(renderer, step_samples=5000)
    tickets = [mix.submit(Server.job(tb, b.frame, 48, 40, 5))]
    mix.step()
    tickets.append(mix.submit(Server.job(ta, a.frame, 40, 32, 7)))  # joins a running batch
    mix.step()
    tickets.append(mix.submit(Server.job(tb, b.frame, 24, 20, 3)))
    tickets.append(mix.submit(Server.job(ta, a.frame, 40, 32, 9, seed_offset=3)))
    mix.run()
    for tk in tickets:
        assert mix.status(tk).state == capi.APH_JOB_DONE
    np.testing.assert_array_equal(mix.result(tickets[1], 40, 32), ref)
    np.testing.assert_array_equal(mix.result(tickets[0], 48, 40), plain_render(renderer, tb, b.frame, 48, 40, 5))


def test_admission_respects_slots_and_budget(renderer, two_scenes):
    (a, ta), _ = two_scenes
    roomy = Server(renderer, hbm_budget=10 ** 12, keep_results=False)
    tks = [roomy.submit(Server.job(ta, a.frame, 40, 32, 4)) for _ in range(3)]
    roomy.run()
    assert all(roomy.status(t).state == capi.APH_JOB_DONE for t in tks)
    # two slots: the third job is admitted only when one of the first two is done
    slots = Server(renderer, max_active=2, step_samples=40 * 32, keep_results=False)
    tks = [slots.submit(Server.job(ta, a.frame, 40, 32, 3 + k)) for k in range(3)]
    slots.run()
    st = [slots.status(t) for t in tks]
    assert st[2].t_admit >= min(st[0].t_done, st[1].t_done)
    assert all(s.t_arrival <= s.t_admit <= s.t_done for s in st)
    # a scene that cannot fit the budget is refused at submission
    tiny = Server(renderer, hbm_budget=1)
    with pytest.raises(AphError) as e:
        tiny.submit(Server.job(ta, a.frame, 40, 32, 1))
    assert e.value.code == capi.APH_E_MEMORY


def test_budget_queues_jobs_until_memory_returns(renderer, two_scenes):
    """With room for one job buffer beside the scene, jobs run strictly one after another."""
    (a, ta), _ = two_scenes
    # the smallest budget that admits one job: resident scene + one buffer
    lo, hi = 1, 1 << 34
    while hi - lo > 1:
        mid = (lo + hi) // 2
        probe = Server(renderer, hbm_budget=mid, keep_results=False)
        try:
            probe.submit(Server.job(ta, a.frame, 40, 32, 1))
            probe.run()
            hi = mid
        except AphError:
            lo = mid
    need_one = hi
    assert need_one > 40 * 32 * 16
    one = Server(renderer, hbm_budget=need_one, step_samples=40 * 32, keep_results=False)
    tks = [one.submit(Server.job(ta, a.frame, 40, 32, 2)) for _ in range(3)]
    one.run()
    st = [one.status(t) for t in tks]
    assert st[1].t_admit >= st[0].t_done and st[2].t_admit >= st[1].t_done


def test_plain_server_runs_one_job_at_a_time(renderer, two_scenes):
    (a, ta), (b, tb) = two_scenes
    plain = Server(renderer, mode=capi.APH_MODE_MEGAKERNEL)
    batched = Server(renderer)
    jobs = [Server.job(ta, a.frame, 40, 32, 4), Server.job(tb, b.frame, 48, 40, 3)]
    tp = [plain.submit(j) for j in jobs]
    tb_ = [batched.submit(j) for j in jobs]
    plain.run()
    batched.run()
    s0, s1 = plain.status(tp[0]), plain.status(tp[1])
    assert s1.t_admit >= s0.t_done
    for k, (w, h) in enumerate([(40, 32), (48, 40)]):
        g, m = batched.result(tb_[k], w, h), plain.result(tp[k], w, h)
        same = np.isclose(g, m, rtol=1e-5, atol=1e-6).all(axis=-1).mean()
        assert same >= 0.99, same


def test_replay_of_an_arrival_trace(renderer, two_scenes):
    (a, ta), (b, tb) = two_scenes
    rng = np.random.default_rng(4)
    jobs, arr, t = [], [], 0.0
    for k in range(12):
        t += rng.exponential(0.01)
        arr.append(t)
        if k % 2:
            jobs.append(Server.job(ta, a.frame, 40, 32, int(rng.integers(1, 6))))
        else:
            jobs.append(Server.job(tb, b.frame, 48, 40, int(rng.integers(1, 6))))
    srv = Server(renderer, max_active=4, keep_results=True)
    tickets = srv.replay(jobs, arr)
    for k, tk in enumerate(tickets):
        st = srv.status(tk)
        assert st.state == capi.APH_JOB_DONE and st.samples_done == jobs[k].spp
        assert st.t_arrival <= st.t_admit <= st.t_done
    # the replayed images equal plain renders of the same jobs
    for k in (0, 1):
        j = jobs[k]
        ref = plain_render(renderer, j.tlas, j.frame, j.width, j.height, j.spp)
        np.testing.assert_array_equal(srv.result(tickets[k], j.width, j.height), ref)


def _device_free():
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_budget_covers_the_wavefront_working_set(renderer, two_scenes):
    """The context is tuned for 64M-path chunks (about 35 GB of path state and queues).
    A server whose budget is a few GB beyond the resident scene serves its jobs with a
    smaller chunk, charges that working set, and the device memory it takes - measured
    with hipMemGetInfo around every step - stays within the budget.  Images are the
    same as plain renders (chunking never changes results)."""
    (a, ta), (b, tb) = two_scenes
    r = renderer
    r.set_tuning(1 << 16, 16, 8)
    plain_render(r, ta, a.frame, 40, 32, 1)  # the context's buffers shrink to a 64K-path chunk
    r.set_tuning(64 << 20, 16, 8)
    budget = 3 << 30
    srv = Server(r, hbm_budget=budget, step_samples=1 << 22, keep_results=True)
    jobs = [Server.job(ta, a.frame, 40, 32, 6), Server.job(tb, b.frame, 48, 40, 5), Server.job(ta, a.frame, 40, 32, 3)]
    tickets = [srv.submit(j) for j in jobs]
    free0 = _device_free()
    resident = srv.usage()[0]
    assert 0 < resident < budget
    peak = 0
    while True:
        act, pend = srv.step()
        used = free0 - _device_free()
        peak = max(peak, used)
        wf, paths = srv.working_set()
        res, jb, _, _ = srv.usage()
        assert res + jb + wf <= budget
        if act == 0 and pend == 0:
            break
    assert 0 < paths < (64 << 20) and wf <= budget - resident
    assert peak <= budget - resident, (peak, budget, resident)
    r.set_tuning(4 << 20, 16, 8)
    for k, j in enumerate(jobs):
        ref = plain_render(r, j.tlas, j.frame, j.width, j.height, j.spp)
        np.testing.assert_array_equal(srv.result(tickets[k], j.width, j.height), ref)
    srv.close()
