"""Every benchmark configuration on the device at its own size.

* C5(i): all 128 small-box frames (N 64-512, edges 4-8, k = 100) in one
  batched call, each frame bit for bit against the oracle.
* C4: the full 256-frame uneven batch captured into a hipGraph and replayed
  three times with the inputs refilled and eager work in between, every
  replay bitwise equal to the eager batch, which is sampled against the
  oracle frame by frame.
* C3 (4,194,304 clustered points, k = 16) and C5(ii) (8,388,608 jittered FCC
  points, k = 32) with the learned scheduler on: trained on other frames of
  the same generator, the learned walk's outputs bitwise equal to the plain
  walk's, and sampled sources bitwise equal to the oracle.
"""
import numpy as np
import pytest

from gpu_util import cell12, run_device, same_bits
from linkcell_amd import workloads as wl
from oracle import lcoracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws(gpu):
    import linkcell_amd as lc
    w = lc.Workspace()
    yield w
    w.close()


def _device_batch(frames):
    import torch
    dev = torch.device("cuda:0")
    keep, recs = [], []
    for fr in frames:
        n, k = len(fr["xyz"]), fr["k"]
        x = torch.from_numpy(np.ascontiguousarray(fr["xyz"])).to(dev)
        nn = torch.full((n, k), -7, dtype=torch.int32, device=dev)
        d2 = torch.full((n, k), -7.0, dtype=torch.float64, device=dev)
        keep.append((x, nn, d2))
        recs.append(dict(xyz=x.data_ptr(), n=n, cell=cell12(fr["cell"]), k=k,
                         cell_hint=fr["hint"], out_nn=nn.data_ptr(), out_d2=d2.data_ptr()))
    torch.cuda.synchronize()
    return keep, recs


def test_c5i_all_128_frames_in_one_batch(ws):
    frames = wl.small_c5i(128)
    keep, recs = _device_batch(frames)
    ws.knearest_batch(recs)
    for f, (fr, (_, nn, d2)) in enumerate(zip(frames, keep)):
        on, od = orc.knearest_many(fr["xyz"], fr["cell"], 100, cell_hint=fr["hint"])
        assert np.array_equal(nn.cpu().numpy(), on), f
        assert same_bits(d2.cpu().numpy(), od), f


def test_c4_full_batch_graph_replays(ws):
    import torch
    frames = wl.batched_c4(256)
    keep, recs = _device_batch(frames)
    ws.knearest_batch(recs)
    eager = [(nn.clone(), d2.clone()) for _, nn, d2 in keep]
    rng = np.random.default_rng(34)
    for f, (fr, (nn, d2)) in enumerate(zip(frames, eager)):
        n = len(fr["xyz"])
        src = np.sort(rng.choice(n, min(n, 48), replace=False))
        on, od, _, _, _, _ = orc.knearest_sample(fr["xyz"], fr["cell"], fr["k"], src,
                                                 cell_hint=fr["hint"])
        assert np.array_equal(nn.cpu().numpy()[src], on[src]), f
        assert same_bits(d2.cpu().numpy()[src], od[src]), f
    ws.graph_capture(recs)
    try:
        for rep in range(3):
            for x, nn, d2 in keep:
                nn.fill_(-9)
                d2.fill_(-9.0)
            torch.cuda.synchronize()
            ws.graph_launch(wait=True)
            for f, ((_, nn, d2), (en, ed)) in enumerate(zip(keep, eager)):
                assert torch.equal(nn, en), (rep, f)
                assert torch.equal(d2.view(torch.int64), ed.view(torch.int64)), (rep, f)
            # eager work on torch's stream between replays
            torch.cuda.synchronize()
    finally:
        ws.graph_reset()


def _learned_vs_plain(ws, cfg, train_frames, samples, seed):
    """Plain search, training on the given device frames, learned search:
    returns nothing, asserts learned == plain bitwise and a sample == oracle."""
    import torch
    dev = torch.device("cuda:0")
    n, k = len(cfg["xyz"]), cfg["k"]
    x = torch.from_numpy(cfg["xyz"]).to(dev)
    nn = torch.empty((n, k), dtype=torch.int32, device=dev)
    d2 = torch.empty((n, k), dtype=torch.float64, device=dev)
    c12 = cell12(cfg["cell"])
    torch.cuda.synchronize()

    def search(xp, onn, od2):
        ws.knearest_ptr(xp, n, 1, c12, k, onn.data_ptr(), od2.data_ptr(), 0, cfg["hint"],
                        0, True)

    ws.set_schedule(0)
    search(x.data_ptr(), nn, d2)
    ref_nn, ref_d2 = nn.clone(), d2.clone()
    tn = torch.empty_like(nn)
    td = torch.empty_like(d2)
    ws.sched_clear()
    for make in train_frames:
        tx = make()
        torch.cuda.synchronize()
        search(tx.data_ptr(), tn, td)
        assert ws.sched_collect() > 0
        del tx
    loss, margin = ws.sched_train(epochs=6, batch=4096, lr=4e-3, seed=1, quantile=0.999)
    assert np.isfinite(loss) and np.isfinite(margin)
    ws.set_schedule(1)
    try:
        nn.fill_(-9)
        d2.fill_(-9.0)
        torch.cuda.synchronize()
        search(x.data_ptr(), nn, d2)
    finally:
        ws.set_schedule(0)
    assert torch.equal(nn, ref_nn)
    assert torch.equal(d2.view(torch.int64), ref_d2.view(torch.int64))
    rng = np.random.default_rng(seed)
    src = np.sort(rng.choice(n, samples, replace=False))
    on, od, _, _, _, _ = orc.knearest_sample(cfg["xyz"], cfg["cell"], k, src,
                                             cell_hint=cfg["hint"])
    gn, gd = nn.cpu().numpy(), d2.cpu().numpy()
    assert np.array_equal(gn[src], on[src])
    assert same_bits(gd[src], od[src])


def test_c3_full_size_learned(ws):
    import torch
    dev = torch.device("cuda:0")
    cfg = wl.clustered_c3(0)
    frames = [lambda f=f: wl.clustered_c3_device(f, dev) for f in range(1, 5)]
    _learned_vs_plain(ws, cfg, frames, 3000, 41)


def test_c5ii_full_size_learned(ws):
    import torch
    dev = torch.device("cuda:0")
    cfg = wl.fcc_c5ii()
    frames = [lambda s=s: wl.fcc_c5ii_device(s, dev) for s in (5501, 5502, 5503)]
    _learned_vs_plain(ws, cfg, frames, 1500, 42)
