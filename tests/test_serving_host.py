"""Host logic of the serving loop (CPU): the seeded request stream and the
latency summary."""
import numpy as np

from particleclem_amd import serving as sv


def test_request_stream_is_seeded_and_within_bounds():
    a = sv.request_stream(500, rate=100.0, seed=5)
    b = sv.request_stream(500, rate=100.0, seed=5)
    c = sv.request_stream(500, rate=100.0, seed=6)
    key = lambda rs: [(r.kind, r.n, r.nsteps, r.ratio, r.phi, r.t_arrive, r.seed) for r in rs]
    assert key(a) == key(b) and key(a) != key(c)
    n = np.array([r.n for r in a])
    assert n.min() >= 32 and n.max() <= 128 and 56 <= np.median(n) <= 72
    st = np.array([r.nsteps for r in a])
    assert st.min() >= 10 and st.max() <= 200
    assert all(2.0 <= r.ratio <= 20.0 and 0.5 <= r.phi <= 1.2 for r in a)
    kinds = np.array([r.kind for r in a])
    assert 0.4 < kinds.mean() < 0.6
    t = np.array([r.t_arrive for r in a])
    assert np.all(np.diff(t) >= 0)
    # Poisson at 100/s: mean gap 10 ms
    assert 0.008 < np.diff(t).mean() < 0.012
    assert len({r.seed for r in a}) == len(a)


def test_summary_percentiles():
    reqs = sv.request_stream(100, rate=50.0, seed=1)
    for i, r in enumerate(reqs):
        r.t_admit = r.t_arrive + 0.001
        r.t_done = r.t_arrive + 0.01 * (i + 1)
    s = sv.summarize(reqs, wall=10.0)
    assert abs(s["latency_ms"]["p50"] - 505.0) < 1e-6
    assert abs(s["latency_ms"]["p99"] - 990.1) < 1e-6
    assert s["requests"] == 100
    cells = sum(r.n ** 3 * r.nsteps for r in reqs)
    assert abs(s["mcell_updates_per_s"] - cells / 10.0 / 1e6) < 1e-9
