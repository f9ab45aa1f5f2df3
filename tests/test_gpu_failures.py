"""Failure handling through the C ABI: HIP failures injected from the host at the points a
real one can occur (wema_dev_debug_inject_error; no kernel is made to fault). The bar: the
failing call returns -1 with the error's text; a failure before anything of the call is
enqueued leaves the session usable and its later output equal to the oracle's, as if the
failed call had not been made; a failure after part of the call ran marks the session
failed, so later calls return -1 instead of using state of unknown content, and the session
still closes (or leaves its pool) cleanly. The reference returns -1 on its errors
(src/wema.c:437-439) and propagates its I/O threads' errors (src/async_io.c:60-66, 291-294)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import wema_amd  # noqa: E402
from oracle import oracle as O  # noqa: E402

W, H, FPS = 64, 48, 30.0
KW = dict(amp=100.0, f_low=0.5, f_high=3.0, levels=3)


def clip(n, seed=4):
    syn = wema_amd.Synth(4, W, H, FPS, seed)
    return np.stack([syn.frame(t) for t in range(n)])


def oracle_out(frames):
    orc = O.OraclePipe(W, H, fps=FPS, **KW)
    return np.stack([orc.frame(f)[0] for f in frames])


@pytest.fixture(autouse=True)
def disarm():
    yield
    wema_amd.inject_error(0)


def test_upload_failure_leaves_the_session_usable():
    n = 10
    frames = clip(n)
    sess = wema_amd.Session(W, H, FPS, max_batch=2, **KW)
    d_in = torch.from_numpy(frames).cuda()
    d_out = torch.empty_like(d_in)
    for t0 in range(0, n, 2):
        if t0 == 4:
            wema_amd.inject_error(1)
            with pytest.raises(wema_amd.WemaError, match="injected"):
                sess.process_device(d_in[t0:t0 + 2], d_out_u8=d_out[t0:t0 + 2], n=2)
            assert sess.frames == 4
        sess.process_device(d_in[t0:t0 + 2], d_out_u8=d_out[t0:t0 + 2], n=2)
    sess.sync()
    assert sess.frames == n
    assert np.array_equal(d_out.cpu().numpy(), oracle_out(frames))
    sess.close()


@pytest.mark.parametrize("after", [1, 3, 7])
def test_launch_failure_marks_the_session_failed(after):
    frames = clip(6)
    sess = wema_amd.Session(W, H, FPS, **KW)
    d_in = torch.from_numpy(frames).cuda()
    d_out = torch.empty_like(d_in)
    sess.process_device(d_in[0:1], d_out_u8=d_out[0:1], n=1)
    wema_amd.inject_error(2, after)
    with pytest.raises(wema_amd.WemaError, match="unknown"):
        sess.process_device(d_in[1:2], d_out_u8=d_out[1:2], n=1)
    with pytest.raises(wema_amd.WemaError, match="failed in an earlier call"):
        sess.process_device(d_in[2:3], d_out_u8=d_out[2:3], n=1)
    sess.sync()        # what was enqueued before the failure still completes
    sess.close()
    # a new session on the same device is unaffected
    s2 = wema_amd.Session(W, H, FPS, **KW)
    out = torch.empty_like(d_in)
    for t in range(6):
        s2.process_device(d_in[t:t + 1], d_out_u8=out[t:t + 1], n=1)
    s2.sync()
    assert np.array_equal(out.cpu().numpy(), oracle_out(frames))
    s2.close()


def test_frame_upload_failure_in_the_asynchronous_pipeline():
    n, batch = 9, 3
    frames = clip(n, seed=6)
    sess = wema_amd.Session(W, H, FPS, max_batch=batch, **KW)
    pin_in = wema_amd.PinnedArray((n, H, W, 3))
    pin_out = wema_amd.PinnedArray((n, H, W, 3))
    pin_in.array[:] = frames
    tickets = []
    for t0 in range(0, n, batch):
        if t0 == 3:
            wema_amd.inject_error(3)
            with pytest.raises(wema_amd.WemaError, match="injected"):
                sess.submit_host(pin_in.array[t0:t0 + batch], pin_out.array[t0:t0 + batch], batch)
        tickets.append(sess.submit_host(pin_in.array[t0:t0 + batch], pin_out.array[t0:t0 + batch], batch))
    for tk in tickets:
        sess.wait(tk)
    assert np.array_equal(pin_out.array, oracle_out(frames))
    sess.close()
    pin_in.close()
    pin_out.close()


def test_launch_failure_in_a_pool_tick():
    frames = torch.from_numpy(clip(4, seed=8)).cuda()
    pool = wema_amd.Pool(2e9)
    a = pool.add(W, H, FPS, **KW)
    b = pool.add(W, H, FPS, **KW)
    out = torch.empty_like(frames)
    pool.tick([a, b], [frames[0], frames[0]], [out[0], out[1]])
    wema_amd.inject_error(2, 2)
    with pytest.raises(wema_amd.WemaError, match="unknown"):
        pool.tick([a, b], [frames[1], frames[1]], [out[0], out[1]])
    with pytest.raises(wema_amd.WemaError, match="earlier tick"):
        pool.tick([a], [frames[2]], [out[2]])
    pool.remove(a)
    pool.remove(b)
    c = pool.add(W, H, FPS, **KW)
    res = torch.empty_like(frames)
    for t in range(4):
        pool.tick([c], [frames[t]], [res[t]])
    pool.sync()
    assert np.array_equal(res.cpu().numpy(), oracle_out(frames.cpu().numpy()))
    pool.close()
