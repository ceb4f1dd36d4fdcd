"""The progressive renderer's adaptive sample count (aurora_amd/apps/sample_counter.h),
the SampleCounter of Libraries/Foundation/API/Aurora/Foundation/Timer.h:97-281 that
Plasma's update loop calls each frame.  Driven by a simulated clock on the CPU: the
expected counts follow from the behaviour the reference documents (one sample per frame
until 20 frame times are recorded, then the samples that fit the target frame time at
the recorded average time per sample, the idle target after a second without restarts,
completion at the maximum), not from its code."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("sc") / "sample_counter_test")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-o", exe, os.path.join(ROOT, "tests", "cpp",
                                                                                  "sample_counter_test.cpp")],
                   check=True)

    def run(active, idle, max_samples, idle_after, restart_at, *per_sample):
        out = subprocess.run([exe, str(active), str(idle), str(max_samples), str(idle_after), str(restart_at)]
                             + [str(p) for p in per_sample], check=True, capture_output=True, text=True).stdout
        return [tuple(int(x) for x in line.split()) for line in out.splitlines()]
    return run


def test_priming_then_frame_time_over_sample_time(driver):
    rows = driver(33.3, 200.0, 100000, 0.0, -1, 2.0)  # 2 ms per sample
    counts = [n for _, _, n in rows]
    assert counts[:20] == [1] * 20           # the ring of 20 frame times fills first
    # then the samples that fit 33.3 ms at 2 ms each; the first full window still holds
    # the zero time recorded before the first frame, so it starts one lower
    assert counts[20] in (15, 16) and counts[21:] == [16] * 9
    starts = [s for _, s, _ in rows]
    assert starts == [sum(counts[:k]) for k in range(len(counts))]  # consecutive sample ranges


def test_average_follows_a_change_of_speed_then_goes_idle(driver):
    # no restart: after a second of simulated time the idle target (200 ms) applies
    rows = driver(33.3, 200.0, 100000, 0.0, -1, 2.0, 4.0)
    counts = [n for _, _, n in rows]
    t, switched = 0.0, None
    for k, n in enumerate(counts):
        if t > 1000.0 and switched is None:
            switched = k
        t += (2.0 if k < 30 else 4.0) * n
    assert switched is not None and switched > 30
    # slowing down: the counts fall toward floor(33.3 / 4) while the window takes in 4 ms frames
    mid = counts[30:switched]
    assert all(a >= b for a, b in zip(mid, mid[1:])) and 8 <= min(mid) <= 12
    # idle: close to 200 / 4 once the window has settled
    assert all(abs(n - 50) <= 2 for n in counts[-8:])


def test_idle_frame_time_after_a_pause(driver):
    # a 1.5 s pause after the first phase; the pause is recorded as one frame's time and
    # leaves the window over 20 frames, the count rising to the idle target 200 / 2
    rows = driver(33.3, 200.0, 100000, 1500.0, -1, 2.0, 2.0)
    counts = [n for _, _, n in rows]
    assert counts[21:30] == [16] * 9
    assert all(a <= b for a, b in zip(counts[30:], counts[31:]))
    assert all(abs(n - 100) <= 2 for n in counts[-5:])


def test_completion_and_restart(driver):
    rows = driver(33.3, 200.0, 100, 0.0, -1, 1.0)
    counts = [n for _, _, n in rows]
    total = 0
    for _, s, n in rows:
        assert s == min(total, 100)
        total += n
    assert total == 100 and counts[-1] == 0            # complete: no more samples
    last = max(k for k, n in enumerate(counts) if n)
    assert counts[last] == 100 - sum(counts[:last]) <= 33  # the last frame takes what remains
    # a restart after completion renders from 0 at once, sized for the active frame time
    # (about 33 at 1 ms; the window's last record paired the final frame's samples with a
    # time from the ring's previous lap, so the estimate runs a little high)
    rows = driver(33.3, 200.0, 100, 0.0, 40, 1.0, 1.0)
    assert rows[39][2] == 0
    assert rows[40][1] == 0 and 30 <= rows[40][2] <= 45
    assert rows[41][1] == rows[40][2]
