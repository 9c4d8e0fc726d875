"""Learned scheduler on the device, through the C-ABI.

* The network's forward pass, loss gradient and Adam steps (HIP kernels on
  bf16 MFMA) against the CPU oracle oracle/mlp.py, which emulates the same
  BF16 rounding points; tolerances are those of FP32 against FP64
  accumulation.
* Training on the plain walk's own labels, then the learned walk: bitwise
  identical to the plain walk (and to the CPU oracle on a sample), also
  where the caps are wrong and rows are walked again.
* What the trained network gains: fewer tiles walked than the plain walk,
  and than an untrained network with the same margin.
"""
import numpy as np
import pytest

from gpu_util import cell12, run_device, same_bits
from linkcell_amd import workloads as wl
from oracle import lcoracle as orc
from oracle import mlp

pytestmark = pytest.mark.gpu

LEARNED, GIVEN, PLAIN, DEPTH = 1, 2, 0, 3


@pytest.fixture(scope="module")
def ws(gpu):
    import linkcell_amd as lc
    w = lc.Workspace()
    yield w
    w.close()


def _samples(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0.0, 1.0, (n, mlp.FEAT)).astype(np.float32)
    t = (X[:, :125].mean(axis=1) * 2.0 - 1.0 + 0.5 * X[:, 125]).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    return X, t, w


def _f64(*a):
    return [np.asarray(x, dtype=np.float64) for x in a]


def test_forward_matches_oracle(ws):
    p = mlp.init_params(1).astype(np.float32)
    ws.sched_set_params(p, 0.0)
    X, _, _ = _samples(2, 300)  # not a multiple of the 64-row block
    y = ws.sched_forward(X)
    p64, X64 = _f64(p, X)
    want = mlp.forward(p64, X64, emulate=True)["y"]
    assert np.max(np.abs(y - want)) <= 1e-4 * (1.0 + np.max(np.abs(want)))


def test_gradient_matches_oracle(ws):
    p = mlp.init_params(3).astype(np.float32)
    ws.sched_set_params(p, 0.0)
    X, t, w = _samples(4, 200)
    g, loss = ws.sched_gradient(X, t, w)
    lo, go = mlp.loss_grad(*_f64(p, X, t, w), emulate=True)
    assert abs(loss - lo) <= 1e-4 * (1.0 + lo)
    blocks = [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
              (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.NPARAMS)]
    for a, b in blocks:
        err = np.linalg.norm(g[a:b] - go[a:b])
        assert err <= 5e-3 * np.linalg.norm(go[a:b]) + 1e-7, (a, b, err)


def test_adam_steps_match_oracle(ws):
    p = mlp.init_params(5).astype(np.float32)
    ws.sched_set_params(p, 0.0)
    X, t, w = _samples(6, 128)
    lr = 1e-3
    loss = ws.sched_adam(X, t, w, 3, lr)
    pg, _ = ws.sched_get_params()
    po, lo = mlp.adam(*_f64(p, X, t, w), 3, lr, emulate=True)
    assert abs(loss - lo) <= 1e-4 * (1.0 + lo)
    d = np.abs(pg.astype(np.float64) - po)
    # every step moves a parameter by about lr; device and oracle agree to a
    # small fraction of that except where a gradient is within its own
    # rounding of zero
    assert np.quantile(d, 0.99) <= 1e-2 * lr
    assert d.max() <= 6.0 * lr
    moved = np.abs(pg.astype(np.float64) - p) > 0.5 * lr
    assert moved.mean() > 0.5


class _Search:
    """One search with device outputs kept alive (collect reads them)."""

    def __init__(self, ws, cfg):
        import torch
        dev = torch.device("cuda:0")
        self.ws = ws
        self.cfg = cfg
        x = cfg["xyz"]
        self.n, self.k = len(x), cfg["k"]
        self.x = torch.from_numpy(x).to(dev)
        self.m = (torch.from_numpy(cfg["mask"]).to(dev) if cfg.get("mask") is not None
                  else None)
        self.nn = torch.empty((self.n, self.k), dtype=torch.int32, device=dev)
        self.d2 = torch.empty((self.n, self.k), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()

    def run(self):
        self.ws.knearest_ptr(self.x.data_ptr(), self.n, 1, cell12(self.cfg["cell"]), self.k,
                             self.nn.data_ptr(), self.d2.data_ptr(),
                             self.m.data_ptr() if self.m is not None else 0,
                             self.cfg["hint"], 0, True)
        return self.nn.cpu().numpy(), self.d2.cpu().numpy()

    def work(self):
        self.ws.set_stats(True)
        self.run()
        st = self.ws.last_stats()
        self.ws.set_stats(False)
        self.ws.set_timing(True)
        self.run()
        rt = self.ws.last_retried()
        self.ws.set_timing(False)
        return st, rt


@pytest.fixture(scope="module")
def trained(ws):
    """Predictor trained on four frames of the clustered mixture (1/16
    scale, same densities), labelled by the plain walk."""
    import torch
    ws.set_schedule(PLAIN)
    ws.sched_clear()
    targets = []
    for f in range(1, 5):
        cfg = wl.clustered_c3_small(f)
        s = _Search(ws, cfg)
        s.run()
        nt = ws.tile_count()
        assert ws.sched_collect() == nt
        lab = torch.empty(nt, dtype=torch.float32, device="cuda:0")
        ws.tile_labels(lab.data_ptr(), nt)
        L = lab.cpu().numpy().astype(np.float64)
        targets.append(np.log2(L[np.isfinite(L)] / 3.0))  # cell_min = 3 (mesh 20x20x30)
    n = ws.sched_samples()
    loss, margin = ws.sched_train(epochs=30, batch=4096, lr=2e-3, seed=1, quantile=0.98)
    params, m2 = ws.sched_get_params()
    assert m2 == margin and np.isfinite(params).all()
    t = np.concatenate(targets)
    return dict(samples=n, loss=loss, margin=margin, params=params, target_var=t.var())


def test_training_fits_the_labels(ws, trained):
    # the loss is half the mean squared error in log2 units: a constant
    # prediction scores half the labels' variance; the network explains
    # most of it
    assert trained["samples"] > 4 * 15000
    assert trained["loss"] < 0.5 * (0.5 * trained["target_var"])
    assert 0.0 < trained["margin"] < 1.0


def test_learned_walk_is_bitwise_the_plain_walk(ws, trained):
    s = _Search(ws, wl.clustered_c3_small(0))
    ws.set_schedule(PLAIN)
    pn, pd = s.run()
    ws.set_schedule(LEARNED)
    try:
        ln, ld = s.run()
    finally:
        ws.set_schedule(PLAIN)
    assert np.array_equal(ln, pn) and same_bits(ld, pd)
    rng = np.random.default_rng(0)
    src = np.sort(rng.choice(s.n, 2000, replace=False))
    on, od, _, _, _, _ = orc.knearest_sample(s.cfg["xyz"], s.cfg["cell"], 16, src,
                                             cell_hint=3.0)
    assert np.array_equal(ln[src], on[src]) and same_bits(ld[src], od[src])


def test_trained_network_saves_work(ws, trained):
    s = _Search(ws, wl.clustered_c3_small(0))
    ws.set_schedule(PLAIN)
    plain, _ = s.work()
    ws.set_schedule(LEARNED)
    try:
        learned, retried = s.work()
        ntiles = ws.tile_count()
        # the same network untrained: random weights, the trained margin
        ws.sched_set_params(mlp.init_params(7).astype(np.float32), trained["margin"])
        untrained, retried_u = s.work()
    finally:
        ws.sched_set_params(trained["params"], trained["margin"])
        ws.set_schedule(PLAIN)
    print(f"\nrow tiles swept: plain {plain['row_useful']}, learned {learned['row_useful']} "
          f"({retried} of {ntiles} rows again), untrained {untrained['row_useful']} "
          f"({retried_u} rows again)")
    # the trained caps hold for nearly every row, and for fewer rows fail
    # than caps from the untrained network. The walk drains each row's tiles
    # nearest first and drops those beyond the row's threshold, so a cap
    # (correct or not) changes the swept tiles little: the trained caps
    # sweep about as many as the plain walk
    assert retried < 0.05 * ntiles
    assert retried < retried_u
    assert learned["row_useful"] < 1.03 * plain["row_useful"]


@pytest.mark.parametrize("margin", [None, 3.0])
def test_depth_schedule_is_bitwise_the_plain_walk(ws, trained, margin):
    """Rows list shells 1 .. predicted depth in one pass and are packed by
    that depth; with the trained margin most rows keep depth 1, with a
    margin of 3 (caps 8x) every row walks more shells than it needs, up to
    the wrap-free reach: the results stay the plain walk's."""
    s = _Search(ws, wl.clustered_c3_small(0))
    ws.set_schedule(PLAIN)
    pn, pd = s.run()
    if margin is not None:
        ws.sched_set_params(trained["params"], margin)
    ws.set_schedule(DEPTH)
    try:
        dn, dd = s.run()
        ws.set_timing(True)
        s.run()
        retried = ws.last_retried()
        ws.set_timing(False)
    finally:
        ws.set_schedule(PLAIN)
        ws.sched_set_params(trained["params"], trained["margin"])
    assert np.array_equal(dn, pn) and same_bits(dd, pd)
    if margin is not None:
        assert retried == 0  # caps 8x above the prediction never cut a row short
    rng = np.random.default_rng(1)
    src = np.sort(rng.choice(s.n, 2000, replace=False))
    on, od, _, _, _, _ = orc.knearest_sample(s.cfg["xyz"], s.cfg["cell"], 16, src,
                                             cell_hint=3.0)
    assert np.array_equal(dn[src], on[src]) and same_bits(dd[src], od[src])


@pytest.mark.parametrize("name", ["c2", "c5i", "c4"])
def test_depth_schedule_on_other_workloads_keeps_results(ws, trained, name):
    # another density (and, for the batch, frames of other k and cells): the
    # predicted depths are wrong in both directions
    if name == "c2":
        cfgs = [wl.triclinic_c2(n=65536)]
    elif name == "c5i":
        cfgs = [dict(fr, name="c5i") for fr in wl.small_c5i(n_frames=3)]
    else:
        cfgs = wl.batched_c4(12, seed0=7000)
    ws.sched_set_params(trained["params"], 1.5)
    try:
        for cfg in cfgs:
            ws.set_schedule(PLAIN)
            pn, pd = run_device(ws, cfg["xyz"], cfg["cell"], cfg["k"], cfg.get("mask"), 3.0)
            ws.set_schedule(DEPTH)
            dn, dd = run_device(ws, cfg["xyz"], cfg["cell"], cfg["k"], cfg.get("mask"), 3.0)
            ws.set_schedule(PLAIN)
            assert np.array_equal(dn, pn) and same_bits(dd, pd)
    finally:
        ws.set_schedule(PLAIN)
        ws.sched_set_params(trained["params"], trained["margin"])


def test_depth_schedule_in_a_batch_keeps_results(ws, trained):
    import torch
    from test_gpu_configs import _device_batch
    frames = wl.batched_c4(24, seed0=7100)
    keep, recs = _device_batch(frames)
    ws.set_schedule(PLAIN)
    ws.knearest_batch(recs)
    ref = [(nn.clone(), d2.clone()) for _, nn, d2 in keep]
    ws.sched_set_params(trained["params"], 2.0)
    ws.set_schedule(DEPTH)
    try:
        for _, nn, d2 in keep:
            nn.fill_(-5)
            d2.fill_(-5.0)
        torch.cuda.synchronize()
        ws.knearest_batch(recs)
    finally:
        ws.set_schedule(PLAIN)
        ws.sched_set_params(trained["params"], trained["margin"])
    for f, ((_, nn, d2), (rn, rd)) in enumerate(zip(keep, ref)):
        assert torch.equal(nn, rn), f
        assert torch.equal(d2.view(torch.int64), rd.view(torch.int64)), f


@pytest.mark.parametrize("name", ["c1", "c2", "c5i"])
def test_learned_schedule_on_other_workloads_keeps_results(ws, trained, name):
    # the network was trained on another density: many caps are wrong, and
    # the rows they cut short are walked again without them
    if name == "c1":
        cfgs = [wl.uniform_c1()]
    elif name == "c2":
        cfgs = [wl.triclinic_c2(n=65536)]
    else:
        cfgs = [dict(fr, name="c5i") for fr in wl.small_c5i(n_frames=3)]
    for cfg in cfgs:
        ws.set_schedule(PLAIN)
        pn, pd = run_device(ws, cfg["xyz"], cfg["cell"], cfg["k"], cfg.get("mask"), 3.0)
        ws.set_schedule(LEARNED)
        try:
            ln, ld = run_device(ws, cfg["xyz"], cfg["cell"], cfg["k"], cfg.get("mask"), 3.0)
        finally:
            ws.set_schedule(PLAIN)
        assert np.array_equal(ln, pn) and same_bits(ld, pd)


@pytest.mark.parametrize("scale", [0.0, 0.5, 1.0, 1e30])
def test_given_tile_caps_never_change_results(ws, scale):
    import torch
    s = _Search(ws, wl.clustered_c3_small(2))
    ws.set_schedule(PLAIN)
    pn, pd = s.run()
    nt = ws.tile_count()
    lab = torch.empty(nt, dtype=torch.float32, device="cuda:0")
    ws.tile_labels(lab.data_ptr(), nt)
    caps = (lab * scale).contiguous()
    ws.sched_set_tile_prediction(caps.data_ptr(), nt)
    ws.set_schedule(GIVEN)
    try:
        ws.set_timing(True)
        gn, gd = s.run()
        retried = ws.last_retried()
    finally:
        ws.set_timing(False)
        ws.set_schedule(PLAIN)
        ws.sched_set_tile_prediction(0, 0)
    assert np.array_equal(gn, pn) and same_bits(gd, pd)
    if scale < 1.0:
        assert retried > 0
    if scale == 0.0:
        assert retried == nt
    if scale >= 1.0:
        # the labels are the rows' own final k-th distances: no row retried
        # beyond float rounding at the cap
        assert retried <= nt // 1000


def test_labels_are_the_rows_largest_kth_distance(ws):
    import torch
    cfg = wl.clustered_c3_small(3)
    s = _Search(ws, cfg)
    ws.set_schedule(PLAIN)
    _, d2 = s.run()
    nt = ws.tile_count()
    lab = torch.empty(nt, dtype=torch.float32, device="cuda:0")
    ws.tile_labels(lab.data_ptr(), nt)
    L = lab.cpu().numpy().astype(np.float64)
    kth = np.sqrt(d2[:, -1])

    def f32_up(v):
        f = np.float32(v)
        return np.float64(np.nextafter(f, np.float32(np.inf)) if np.float64(f) < v else f)

    # every label is some source's k-th distance rounded up to float, the
    # largest label that of the largest k-th distance
    assert L.max() == f32_up(kth.max())
    assert np.all(L >= f32_up(kth.min()))
    ups = np.unique(np.array([f32_up(v) for v in np.unique(kth)]))
    assert np.isin(L, ups).all()
    assert len(L) >= s.n // 16


def _bf16_bits(x64):
    """BF16 bit pattern (round to nearest even through FP32) of float64 values."""
    return (mlp.bf16(x64).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def _bf16_ambiguous(x64, ulps=8):
    """True where an FP32 evaluation within `ulps` FP32 ulps of x could round
    to either BF16 neighbour (x sits that close to a BF16 rounding midpoint)."""
    x = np.asarray(x64, dtype=np.float64)
    lo = mlp.bf16(x * (1.0 - ulps * 2.0 ** -23))
    hi = mlp.bf16(x * (1.0 + ulps * 2.0 ** -23))
    return lo != hi


def test_feature_maps_match_between_training_inference_and_oracle(ws):
    """The 160 inputs of every tile of a clustered search, as training
    collects them and as inference reads them, are the same BF16 values;
    both equal mlp.tile_features on the tile layout (exactly, except where
    an FP32 log2 lies within rounding of a BF16 midpoint); the cell
    occupancies equal a numpy bin count; inference's caps follow
    mlp.forward on the dumped inputs."""
    cfg = wl.clustered_c3_small(5)
    s = _Search(ws, cfg)
    s.run()
    ws.sched_clear()
    ws.sched_collect()
    ws.sched_train(epochs=2, batch=4096, lr=4e-3, seed=1, quantile=0.999)
    p, margin = ws.sched_get_params()
    s.run()
    L = np.asarray(cfg["cell"], dtype=np.float64)
    oc = orc.Cell.ortho(*L)
    plan = oc.plan(cfg["hint"])
    nx, ny, nz = plan["nx"], plan["ny"], plan["nz"]
    cm = plan["cell_min"]
    d = ws.sched_dump(nx * ny * nz)
    xc, xp = d["x_collect"], d["x_predict"]
    # 1. training and inference read identical inputs
    assert np.array_equal(xc, xp)
    # 2. occupancies: the fold and bin of src/cell.rs:176-194 and
    #    src/knearest.rs:405-412 in numpy
    sf = cfg["xyz"] / np.array(oc.widths)
    sf = sf - np.floor(sf)
    sf = np.where(sf >= 1.0, 0.0, sf)
    nvec = np.array([nx, ny, nz])
    b = np.clip((sf * nvec).astype(np.int64), 0, nvec - 1)
    cid = (b[:, 2] * ny + b[:, 1]) * nx + b[:, 0]
    count = np.bincount(cid, minlength=nx * ny * nz)
    assert np.array_equal(d["count"], count)
    # 3. the CPU feature map on the device's tile layout
    cells, box = d["tile_cell"], d["tile_box"]
    nt = len(cells)
    assert nt == np.sum((count + 15) // 16)
    first = np.concatenate([[0], np.cumsum((count + 15) // 16)])
    tl = np.arange(nt) - first[cells]
    fill = np.minimum(16, count[cells] - 16 * tl)
    assert (fill >= 1).all()
    c3 = count.reshape(nz, ny, nx)
    hx, hy, hz = cells % nx, (cells // nx) % ny, cells // (nx * ny)
    ext = np.stack([(box[:, 1] - box[:, 0]) / cm, (box[:, 3] - box[:, 2]) / cm,
                    (box[:, 5] - box[:, 4]) / cm], axis=1).astype(np.float32)
    hinv = np.array(oc.hinv)
    mid = np.stack([0.5 * (box[:, 0] + box[:, 1]), 0.5 * (box[:, 2] + box[:, 3]),
                    0.5 * (box[:, 4] + box[:, 5])], axis=1)
    sfrac = np.stack([(hinv[0] * mid[:, 0] + hinv[3] * mid[:, 1]) + hinv[6] * mid[:, 2],
                      (hinv[1] * mid[:, 0] + hinv[4] * mid[:, 1]) + hinv[7] * mid[:, 2],
                      (hinv[2] * mid[:, 0] + hinv[5] * mid[:, 1]) + hinv[8] * mid[:, 2]],
                     axis=1)
    pos = (sfrac * nvec - np.stack([hx, hy, hz], axis=1)).astype(np.float32)
    want = np.empty((nt, mlp.FEAT))
    for t in range(nt):
        want[t] = mlp.tile_features(c3, (nx, ny, nz), (hx[t], hy[t], hz[t]),
                                    ext[t].astype(np.float64), pos[t].astype(np.float64),
                                    fill[t], cfg["k"])
    wb = _bf16_bits(want)
    amb = _bf16_ambiguous(want)
    # position, fill and k inputs are exact casts, never ambiguous
    amb[:, 128:] = False
    same = wb == xc
    assert same[~amb].all(), np.argwhere(~same & ~amb)[:5]
    near = (np.abs(wb.astype(np.int64) - xc.astype(np.int64)) <= 1)
    assert near[amb].all()
    assert amb.mean() < 2e-3
    # 4. inference's caps: cell_min * 2^(y + margin), y = mlp.forward on the
    #    dumped inputs. All but a few tiles within the tolerance of
    #    test_forward_matches_oracle; the rest differ by one BF16 rounding
    #    of a hidden activation, which an FP32 sum one ulp away from the
    #    oracle's FP64 sum can flip (a 2^-8 relative step in one input of
    #    the second layer)
    Xd = (xc.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    y = mlp.forward(p.astype(np.float64), Xd, emulate=True)["y"]
    y_dev = np.log2(d["caps"].astype(np.float64) / np.float32(cm)) - margin
    err = np.abs(y_dev - y)
    assert np.quantile(err, 0.999) <= 1e-4 * (1.0 + np.max(np.abs(y))) + 1e-5
    assert err.max() <= 2.0 ** -6
