"""Multi-stream pool and serving loop on the GPU, through the C ABI.

The bar: every stream's output in a mixed tick (and replayed from a captured hipGraph) is
byte-identical to the same stream processed alone (no computation crosses streams, and
the threshold sums are order-independent integers); admission never exceeds the budget.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import wema_amd  # noqa: E402

STREAMS = [
    dict(w=64, h=48, fps=30.0, kw=dict(amp=120.0, f_low=0.5, f_high=3.0, levels=3), preset=4),
    dict(w=96, h=54, fps=30.0, kw=dict(amp=60.0, f_low=0.4, f_high=4.0, color=True), preset=2),
    dict(w=80, h=60, fps=60.0, kw=dict(amp=90.0, use_iir=False, window=8), preset=4),
    dict(w=50, h=33, fps=30.0, kw=dict(amp=30.0, f_low=1.0, f_high=5.0, edge_aware=False), preset=4),
    dict(w=120, h=68, fps=60.0, kw=dict(amp=150.0, f_low=0.3, f_high=5.0, use_iir=False,
                                       window=16, bilateral=True, color=True), preset=2),
]


def clips(n):
    out = []
    for i, s in enumerate(STREAMS):
        syn = wema_amd.Synth(s["preset"], s["w"], s["h"], s["fps"], seed_offset=10 + i)
        out.append(torch.from_numpy(np.stack([syn.frame(t) for t in range(n)])).cuda())
    return out


def alone(n, frames):
    res = []
    for s, fr in zip(STREAMS, frames):
        sess = wema_amd.Session(s["w"], s["h"], s["fps"], **s["kw"])
        out = torch.empty_like(fr)
        for t in range(n):
            sess.process_device(fr[t:t + 1], d_out_u8=out[t:t + 1], n=1)
        sess.sync()
        res.append(out.cpu().numpy())
        sess.close()
    return res


@pytest.mark.parametrize("graphs", [False, True])
def test_mixed_ticks_equal_alone(graphs):
    n = 20
    frames = clips(n)
    ref = alone(n, frames)
    pool = wema_amd.Pool(4e9, graphs=graphs)
    ids = [pool.add(s["w"], s["h"], s["fps"], **s["kw"]) for s in STREAMS]
    outs = [torch.empty_like(f) for f in frames]
    rng = np.random.default_rng(0)
    done = [0] * len(STREAMS)
    # uneven ticks: each tick takes a random subset of the streams that still have frames
    while min(done) < n:
        sel = [i for i in range(len(STREAMS)) if done[i] < n and rng.random() < 0.7]
        if not sel:
            continue
        pool.tick([ids[i] for i in sel], [frames[i][done[i]] for i in sel],
                  [outs[i][done[i]] for i in sel])
        for i in sel:
            done[i] += 1
    pool.sync()
    for i in range(len(STREAMS)):
        assert np.array_equal(outs[i].cpu().numpy(), ref[i]), STREAMS[i]
    pool.close()


def test_admission_respects_budget():
    one = wema_amd.Session.query_bytes(1920, 1080, 30.0, f_high=3.0)
    pool = wema_amd.Pool(int(2.5 * one))
    a = pool.add(1920, 1080, 30.0, f_high=3.0)
    b = pool.add(1920, 1080, 30.0, f_high=3.0)
    with pytest.raises(wema_amd.WemaError, match="budget"):
        pool.add(1920, 1080, 30.0, f_high=3.0)
    assert pool.bytes_in_use <= 2.5 * one
    pool.remove(a)
    c = pool.add(1920, 1080, 30.0, f_high=3.0)
    assert c == a and pool.bytes_in_use == 2 * one
    pool.remove(b)
    pool.remove(c)
    assert pool.bytes_in_use == 0
    pool.close()


def test_serving_loop_runs():
    st = wema_amd.serve(duration_s=2.0, drain_s=0.5, arrival_rate=6.0, life=(0.5, 1.5),
                        hbm_budget=40e9, seed=3)
    assert st["sessions_admitted"] > 0 and st["frames"] > 0
    assert st["peak_bytes"] <= 40e9
    assert 0 < st["lat_p50_ms"] <= st["lat_p99_ms"] <= st["lat_max_ms"]
