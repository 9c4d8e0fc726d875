"""Batched uneven systems, hipGraph capture/replay and the walk-split tuner.

Each frame of a batch must equal its own single-system oracle call
(src/gpu/knearest.cpp:705-816 runs frames one after another; the batch
runs them in one pipeline), bit for bit, whatever the walk split."""
import numpy as np
import pytest

from gpu_util import cell12, same_bits
from oracle import lcoracle as orc
from linkcell_amd import workloads as wl

pytestmark = pytest.mark.gpu

# the first frames of the 256-frame uneven batch: n from 1e3 to 2e5,
# orthorhombic and triclinic, uniform and clustered, k in {4, 12, 32}
N_FRAMES = 24


@pytest.fixture(scope="module")
def ws(gpu):
    import linkcell_amd as lc
    w = lc.Workspace()
    yield w
    w.close()


@pytest.fixture(scope="module")
def c4():
    return wl.batched_c4(N_FRAMES)


@pytest.fixture(scope="module")
def c4_oracle(c4):
    out = []
    for fr in c4:
        nn, d2, reach, _ = orc.knearest_many(fr["xyz"], fr["cell"], fr["k"],
                                             cell_hint=fr["hint"], stats=True)
        out.append((nn, d2, reach))
    return out


class DeviceBatch:
    """Device buffers for a list of frames plus the frame records that
    point into them."""

    def __init__(self, frames, masks=None):
        import torch
        dev = torch.device("cuda:0")
        self.frames = frames
        self.x, self.nn, self.d2, self.reach, self.mask = [], [], [], [], []
        self.recs = []
        for i, fr in enumerate(frames):
            n, k = len(fr["xyz"]), fr["k"]
            x = torch.from_numpy(np.ascontiguousarray(fr["xyz"], dtype=np.float64)).to(dev)
            nn = torch.full((n, k), -7, dtype=torch.int32, device=dev)
            d2 = torch.full((n, k), -7.0, dtype=torch.float64, device=dev)
            rr = torch.full((n,), -1, dtype=torch.int32, device=dev)
            m = None
            if masks is not None and masks[i] is not None:
                m = torch.from_numpy(np.ascontiguousarray(masks[i], dtype=np.int32)).to(dev)
            self.x.append(x)
            self.nn.append(nn)
            self.d2.append(d2)
            self.reach.append(rr)
            self.mask.append(m)
            self.recs.append(dict(xyz=x.data_ptr(), n=n, cell=cell12(fr["cell"]), k=k,
                                  mask=m.data_ptr() if m is not None else 0,
                                  cell_hint=fr.get("hint") or 0.0, out_nn=nn.data_ptr(),
                                  out_d2=d2.data_ptr(), out_reach=rr.data_ptr()))
        torch.cuda.synchronize()

    def clear(self):
        import torch
        for nn, d2, rr in zip(self.nn, self.d2, self.reach):
            nn.fill_(-7)
            d2.fill_(-7.0)
            rr.fill_(-1)
        torch.cuda.synchronize()

    def results(self):
        return [(nn.cpu().numpy(), d2.cpu().numpy(), rr.cpu().numpy())
                for nn, d2, rr in zip(self.nn, self.d2, self.reach)]


def assert_frames_equal(got, want, what):
    for f, ((gn, gd, gr), (on, od, orr)) in enumerate(zip(got, want)):
        assert np.array_equal(gn, on), f"{what}: frame {f} indices"
        assert same_bits(gd, od), f"{what}: frame {f} d2 bits"
        assert np.array_equal(gr, orr), f"{what}: frame {f} reach"


def test_uneven_batch_matches_per_frame_oracle(ws, c4, c4_oracle):
    b = DeviceBatch(c4)
    ws.knearest_batch(b.recs)
    assert_frames_equal(b.results(), c4_oracle, "batch")


@pytest.mark.parametrize("dense_min", [-1, -2, 1, 16, 64, 1 << 30])
def test_batch_independent_of_walk_split(ws, c4, c4_oracle, dense_min):
    b = DeviceBatch(c4[:10])
    ws.set_dense_min(dense_min)
    try:
        ws.knearest_batch(b.recs)
    finally:
        ws.set_dense_min(0)
    assert_frames_equal(b.results(), c4_oracle[:10], f"dense_min={dense_min}")


def test_batch_with_masks_and_exact_mode_frames(ws):
    rng = np.random.default_rng(77)
    frames, masks = [], []
    for f in range(6):
        n = int(rng.integers(200, 3000))
        L = rng.uniform(6.0, 14.0, size=3)
        frames.append(dict(xyz=rng.random((n, 3)) * L, cell=L, hint=2.0,
                           k=[1, 5, 16, 33, 40, 100][f]))
        masks.append((rng.random(n) < 0.7).astype(np.int32) if f % 2 else None)
    b = DeviceBatch(frames, masks)
    ws.knearest_batch(b.recs)
    want = []
    for fr, m in zip(frames, masks):
        nn, d2, reach, _ = orc.knearest_many(fr["xyz"], fr["cell"], fr["k"], mask=m,
                                             cell_hint=fr["hint"], stats=True)
        want.append((nn, d2, reach))
    got = b.results()
    for f, ((gn, gd, _), (on, od, _)) in enumerate(zip(got, want)):
        assert np.array_equal(gn, on), f"frame {f}"
        assert same_bits(gd, od), f"frame {f}"


def test_graph_replay_follows_refilled_inputs(ws, c4):
    """Capture once; every replay reads whatever the captured buffers hold,
    with eager work and host waits between replays."""
    import torch
    base = c4[:8]
    b = DeviceBatch(base)
    ws.graph_capture(b.recs)
    try:
        for rep in range(3):
            new = []
            for f, fr in enumerate(base):
                rng = np.random.default_rng(9000 + 31 * rep + f)
                H = np.asarray(fr["cell"], dtype=np.float64)
                H = np.diag(H) if H.shape == (3,) else H
                xyz = rng.random((len(fr["xyz"]), 3)) @ H
                new.append(dict(fr, xyz=xyz))
                b.x[f].copy_(torch.from_numpy(xyz))
            b.clear()
            ws.graph_launch(wait=True)
            want = []
            for fr in new:
                nn, d2, reach, _ = orc.knearest_many(fr["xyz"], fr["cell"], fr["k"],
                                                     cell_hint=fr["hint"], stats=True)
                want.append((nn, d2, reach))
            assert_frames_equal(b.results(), want, f"replay {rep}")
            # eager work on the same workspace between replays
            ws.knearest_batch(b.recs[:2])
    finally:
        ws.graph_reset()


def test_learned_schedule_in_captured_graph(ws, c4, c4_oracle):
    """The learned schedule inside a captured batch: train the predictor on
    other uneven batches (labels from their plain walk), capture this batch
    under the learned schedule, replay it with eager work between replays;
    every frame stays bit-identical to its own oracle call."""
    train = wl.batched_c4(12, seed0=4700)
    ws.set_schedule(0)
    ws.sched_clear()
    tb = DeviceBatch(train)
    ws.knearest_batch(tb.recs)
    assert ws.sched_collect() == ws.tile_count()
    loss, margin = ws.sched_train(epochs=20, batch=1024, lr=2e-3, seed=3, quantile=0.98)
    assert np.isfinite(loss) and np.isfinite(margin)
    b = DeviceBatch(c4[:10])
    ws.set_schedule(1)
    try:
        ws.graph_capture(b.recs)
        for rep in range(3):
            b.clear()
            ws.graph_launch(wait=True)
            assert_frames_equal(b.results(), c4_oracle[:10], f"learned replay {rep}")
            ws.knearest_batch(b.recs[:2])
    finally:
        ws.graph_reset()
        ws.set_schedule(0)


def test_persistent_walk_in_captured_graph(ws, c4, c4_oracle):
    """The persistent row walk (waves taking tile groups from a device
    counter that a fill kernel clears) inside a captured batch: every
    replay, with eager work between replays, equals the oracle."""
    b = DeviceBatch(c4[:10])
    ws.set_dense_min(-2)
    try:
        ws.graph_capture(b.recs)
        for rep in range(3):
            b.clear()
            ws.graph_launch(wait=True)
            assert_frames_equal(b.results(), c4_oracle[:10], f"persistent replay {rep}")
            ws.knearest_batch(b.recs[:2])
    finally:
        ws.graph_reset()
        ws.set_dense_min(0)


def test_graph_launch_without_capture_raises(ws):
    import linkcell_amd as lc
    ws.graph_reset()
    with pytest.raises(lc.Error, match="no captured graph"):
        ws.graph_launch()


def test_autotune_keeps_results(ws, c4, c4_oracle):
    b = DeviceBatch(c4[:10])
    best, ms = ws.autotune(b.recs)
    assert best in (-1, -2, 1, 16, 32, 64, 128, 1 << 30)
    assert ms > 0.0
    b.clear()
    ws.knearest_batch(b.recs)
    assert_frames_equal(b.results(), c4_oracle[:10], f"tuned dense_min={best}")


def test_batch_errors_name_the_frame(ws, c4):
    import linkcell_amd as lc
    b = DeviceBatch(c4[:3])
    recs = [dict(r) for r in b.recs]
    recs[1]["k"] = 0
    with pytest.raises(lc.Error, match="frame 1: k must be at least 1"):
        ws.knearest_batch(recs)
    recs = [dict(r) for r in b.recs]
    recs[2]["cell"] = [0.0] * 12
    with pytest.raises(lc.Error, match="frame 2: bad cell"):
        ws.knearest_batch(recs)
    with pytest.raises(lc.Error, match="no points"):
        ws.knearest_batch([])
