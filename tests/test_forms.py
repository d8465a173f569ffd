"""Host logic of the batched and served forms (job generators, queueing arithmetic)."""
import importlib

import numpy as np

forms = importlib.import_module("wavefront-renderer_amd.forms")


def test_batched_jobs_follow_the_config():
    jobs = forms.batched_jobs(64)
    assert len(jobs) == 64
    assert all(128 <= j.width <= 1024 and j.width == j.height for j in jobs)
    assert {j.spp for j in jobs} <= {4, 16, 64, 256}
    assert all(8 <= j.max_depth <= 12 for j in jobs)
    assert {j.scene for j in jobs} == {0, 1}
    assert len({j.tag for j in jobs}) == 64
    assert forms.batched_jobs(64)[5] == jobs[5]          # seeded


def test_served_jobs_arrivals_and_sizes():
    jobs = forms.served_jobs(400, rate=4.0)
    t = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 0.25) < 0.05                # Poisson at 4 jobs/s
    assert all(64 <= j.width <= 1920 and 64 <= j.height <= 1080 for j in jobs)
    side = np.median([j.width for j in jobs])
    assert 400 < side < 650                              # lognormal around 512
    assert all(j.adaptive and abs(j.threshold ** 2 - 1e-3) < 1e-9 for j in jobs)


def test_fcfs_latency():
    lat = forms.fcfs_latency([0.0, 0.5, 0.6, 5.0], [1.0, 1.0, 1.0, 0.5])
    assert np.allclose(lat, [1.0, 1.5, 2.4, 0.5])
