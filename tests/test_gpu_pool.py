"""Multi-stream pool and serving loop on the GPU, through the C ABI.

The bar: every stream's output in a mixed tick (and replayed from a captured hipGraph) is
byte-identical to the same stream processed alone (no computation crosses streams, and
the threshold sums are order-independent integers), and equal to the oracle's (within
1 LSB with bilateral weights); admission never exceeds the budget.
"""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import wema_amd  # noqa: E402
from oracle import oracle as O  # noqa: E402

DIGESTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_digests.json")

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
        s = STREAMS[i]
        orc = O.OraclePipe(s["w"], s["h"], fps=s["fps"], **s["kw"])
        fr = frames[i].cpu().numpy()
        want = np.stack([orc.frame(fr[t])[0] for t in range(n)])
        tol = 1 if s["kw"].get("bilateral") else 0
        assert int(np.abs(want.astype(int) - ref[i].astype(int)).max()) <= tol, s
    pool.close()


@pytest.mark.parametrize("graphs", [False, True])
def test_tick_runs_of_frames_equal_alone(graphs):
    """A session listed k times in a row in one tick (k <= its max_batch) has its next k
    frames processed as one call of k frames: with random runs of 1-3 frames per session
    per tick, every stream's bytes equal the stream alone; a session in two runs of one
    tick, or a run longer than max_batch, is refused."""
    n = 18
    frames = clips(n)
    ref = alone(n, frames)
    pool = wema_amd.Pool(4e9, graphs=graphs)
    ids = [pool.add(s["w"], s["h"], s["fps"], max_batch=3, **s["kw"]) for s in STREAMS]
    outs = [torch.empty_like(f) for f in frames]
    rng = np.random.default_rng(7)
    done = [0] * len(STREAMS)
    while min(done) < n:
        t_ids, t_in, t_out, took = [], [], [], {}
        for i in range(len(STREAMS)):
            k = min(int(rng.integers(0, 4)), n - done[i])
            for j in range(k):
                t_ids.append(ids[i])
                t_in.append(frames[i][done[i] + j])
                t_out.append(outs[i][done[i] + j])
            if k:
                took[i] = k
        if not t_ids:
            continue
        pool.tick(t_ids, t_in, t_out)
        for i, k in took.items():
            done[i] += k
    pool.sync()
    for i in range(len(STREAMS)):
        assert np.array_equal(outs[i].cpu().numpy(), ref[i]), STREAMS[i]
    with pytest.raises(wema_amd.WemaError, match="two runs"):
        pool.tick([ids[0], ids[1], ids[0]], [frames[0][0]] * 3, [outs[0][0]] * 3)
    with pytest.raises(wema_amd.WemaError, match="max_batch"):
        pool.tick([ids[0]] * 4, [frames[0][0]] * 4, [outs[0][0]] * 4)
    pool.close()


def test_c4_mix_full_size_vs_alone_and_oracle():
    """The 24-stream mix at its full sizes (8 x 640x480@30, 8 x 1280x720@30, 6 x 1920x1080@60,
    2 x 3840x2160@30; per-stream band, amplification and window; automatic levels), 20 ticks
    of all 24 streams, once with direct launches and once with hipGraph replay (captured at
    the second tick): every stream's bytes equal the stream run alone, and the 640x480 and
    1280x720 streams' bytes and thresholds equal the oracle's for every frame (digests
    written by tests/golden/make_digests.py; frames 15-19 are past the fade-in)."""
    with open(DIGESTS) as f:
        doc = json.load(f)
    streams, cases = doc["c4_streams"], doc["cases"]
    n = 20
    inputs = [wema_amd.Synth(s["preset"], s["w"], s["h"], s["fps"], s["seed_offset"]).frames_device(0, n)
              for s in streams]
    alone = []
    for s, x in zip(streams, inputs):
        sess = wema_amd.Session(s["w"], s["h"], s["fps"], **s["kw"])
        out = torch.empty_like(x)
        for t in range(n):
            sess.process_device(x[t:t + 1], d_out_u8=out[t:t + 1], n=1)
        sess.sync()
        sess.close()
        alone.append(out)
    for graphs in (False, True):
        print("c4 mix, graphs=%s" % graphs, flush=True)
        pool = wema_amd.Pool(40e9, graphs=graphs)
        ids = [pool.add(s["w"], s["h"], s["fps"], **s["kw"]) for s in streams]
        outs = [torch.empty_like(x) for x in inputs]
        thr = {i: [] for i in range(len(streams))}
        for t in range(n):
            pool.tick(ids, [x[t] for x in inputs], [o[t] for o in outs])
            for i in range(len(streams)):
                if "c4_%02d" % i in cases:
                    thr[i].append(pool.thresholds(ids[i]))
        pool.sync()
        for i, s in enumerate(streams):
            assert torch.equal(outs[i], alone[i]), (graphs, i, s)
            name = "c4_%02d" % i
            if name not in cases:
                continue
            got = outs[i].cpu().numpy()
            for t in range(n):
                assert list(thr[i][t][:1].view(np.uint32)) == cases[name]["thr_bits"][t][:1], (graphs, name, t)
                assert hashlib.sha256(got[t].tobytes()).hexdigest() == cases[name]["out_sha256"][t], (graphs, name, t)
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


def test_serving_c5_outputs_equal_alone_and_oracle():
    """The serving loop at the c5 load (Poisson 4 sessions/s, lifetime 2-20 s, 30/60 fps,
    resolutions from the 8:8:6:2 mix, 200 GB budget), from steady state for 3 s, with frames
    uploaded from pinned host memory and downloaded back (copies overlapping the ticks):
    the outputs of two sessions of each resolution class, recorded as they are served (up to
    24 frames each, in whichever ticks they shared with the others), equal the same streams
    run alone, frame by frame; one 640x480 session also equals the oracle."""
    cfg = dict(duration_s=3.0, arrival_rate=4.0, life=(2.0, 20.0), hbm_budget=200e9,
               warm_start=True, seed=21)
    plan = wema_amd.serve_plan(**{k: v for k, v in cfg.items() if k != "hbm_budget"})
    pick, per_cls = [], {}
    for i, (p, arrive, c) in enumerate(plan):
        if arrive == 0.0 and per_cls.get(c, 0) < 2:
            per_cls[c] = per_cls.get(c, 0) + 1
            pick.append(i)
    assert len(per_cls) == 4, per_cls
    want_frames = 24
    got = {i: {} for i in pick}

    def on_frame(session, frame, arr):
        if session in got and frame < want_frames:
            got[session][frame] = hashlib.sha256(arr.tobytes()).hexdigest()

    st = wema_amd.serve(batched=True, graphs=True, host_io=True, on_frame=on_frame, **cfg)
    assert st["frames"] > 0 and st["peak_bytes"] <= 200e9
    assert 0 < st["lat_p50_ms"] <= st["lat_p99_ms"] and st["link_gbs"] > 0
    inputs = {}
    for c, (w, h, fps) in enumerate(wema_amd.SERVE_MIX):
        syn = wema_amd.Synth(4, w, h, fps, c)
        inputs[c] = np.stack([syn.frame(k) for k in range(wema_amd.SERVE_CLASS_FRAMES)])
    checked_oracle = False
    for i in pick:
        p, _, c = plan[i]
        n = len(got[i])
        assert n >= 10 and sorted(got[i]) == list(range(n)), (i, n)
        sess = wema_amd.Session(p.width, p.height, p.fps, amp=p.amp, f_low=p.f_low, f_high=p.f_high,
                                window=p.window)
        seq = np.stack([inputs[c][t % wema_amd.SERVE_CLASS_FRAMES] for t in range(n)])
        d_in = torch.from_numpy(seq).cuda()
        d_out = torch.empty_like(d_in)
        for t in range(n):
            sess.process_device(d_in[t:t + 1], d_out_u8=d_out[t:t + 1], n=1)
        sess.sync()
        sess.close()
        alone = d_out.cpu().numpy()
        for t in range(n):
            assert hashlib.sha256(alone[t].tobytes()).hexdigest() == got[i][t], (i, c, t)
        if c == 0 and not checked_oracle:
            orc = O.OraclePipe(p.width, p.height, fps=p.fps, amp=p.amp, f_low=p.f_low, f_high=p.f_high,
                               window=p.window)
            for t in range(n):
                assert np.array_equal(orc.frame(seq[t])[0], alone[t]), (i, t)
            checked_oracle = True
    assert checked_oracle
