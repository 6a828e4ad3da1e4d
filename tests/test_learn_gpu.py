"""GPU: the learned coarse solver.

* forward output, energy loss and the full parameter gradient of the device
  network (BF16 MFMA kernels) against the CPU reference network with the same
  rounding points;
* training on residuals harvested from the simulator lowers the loss;
* flexible CG with the trained two-grid preconditioner reaches the same FP64
  residual as the plain solvers, in fewer iterations than with the untrained
  network and than Jacobi-PCG, and agrees with their solution.
"""
import importlib.util
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

_spec = importlib.util.spec_from_file_location(
    "unet_ref", os.path.join(os.path.dirname(__file__), "..", "oracle", "unet_ref.py"))
U = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(U)


def _batch(rng, B, H, W):
    in5 = rng.normal(size=(B, H, W, 5)).astype(np.float32)
    in5[..., 1:] = rng.uniform(0.0, 0.3, (B, H, W, 4))
    d = rng.uniform(1.0, 2.0, (B, H, W))
    cx = rng.uniform(0.0, 0.4, (B, H, W))
    cy = rng.uniform(0.0, 0.4, (B, H, W))
    cx[:, :, -1] = 0.0
    cy[:, -1, :] = 0.0
    op = np.stack([d, cx, cy], axis=1).astype(np.float32)
    return in5, rng.normal(size=(B, H, W)).astype(np.float32), op


def test_device_network_matches_cpu_reference():
    from fluidsim_amd.learn import LearnedPreconditioner

    rng = np.random.default_rng(11)
    B, H, W = 2, 16, 24
    net = LearnedPreconditioner(0, 8, 8, seed=5)
    w = net.weights()
    assert w.size == U.offsets()[1]
    in5, rhat, op = _batch(rng, B, H, W)
    loss, grad, out = net.loss_grad(in5, rhat, op)
    ref = U.UNetRef(w.astype(np.float64), use_bf16=True)
    e, acts = ref.forward(in5.astype(np.float64))
    L, dout = ref.loss(e, rhat.astype(np.float64), op.astype(np.float64))
    g = ref.backward(dout, acts)
    # measured on the MI355X: output 1.4e-6, loss 7e-7, layer gradients
    # 2e-9 .. 3.7e-7 relative; the bounds sit just above.  FP32 accumulation
    # order is the only difference left (both sides round to BF16 at the same
    # points); on the CPU network, leaving out every BF16 rounding moves the
    # output by 6e-3 and transposing one layer's 3x3 taps by 0.1 .. 0.9.
    # Not every input stays this close: one 16x24 sample drawn from another
    # seed differed by 1.2e-3 (in both launch shapes; 32x32 and 16x16 samples
    # by 5e-9 and 1.4e-8).  The likely cause, not examined further: an FP32
    # sum that lies within its rounding of a BF16 rounding boundary rounds to
    # the neighbouring BF16 value on one side (one BF16 ulp, 0.4 %), and a deep
    # layer spreads that element over the output.
    print("out rel L2", rel_l2(out, e), "loss", loss, L)
    assert rel_l2(out, e) <= 5e-6
    assert loss == pytest.approx(L, rel=5e-6, abs=1e-9)
    off, n = U.offsets()
    for l in range(15):
        hi = off[l + 1] if l < 14 else n
        err = rel_l2(grad[off[l]:hi], g[off[l]:hi])
        print(f"layer {l} grad rel L2 {err:.2e}")
        assert err <= 1.5e-6, (l, err)
    assert rel_l2(grad, g) <= 1e-6


@pytest.mark.parametrize("H,W", [(16, 24), (64, 64)])
def test_application_launch_shape_equals_training_shape(H, W):
    """A single sample runs the forward in the network application's launch
    shape (one output tile per wave, the tiles along the grid's y), a batch in
    the training shape (every tile in each wave): each output is the same
    MFMA chain in the same order, so the two agree bitwise."""
    from fluidsim_amd.learn import LearnedPreconditioner

    rng = np.random.default_rng(3)
    net = LearnedPreconditioner(0, 8, 8, seed=7)
    in5, rhat, op = _batch(rng, 1, H, W)
    _, _, out1 = net.loss_grad(in5, rhat, op)
    dup = lambda a: np.concatenate([a, a])
    _, _, out2 = net.loss_grad(dup(in5), dup(rhat), dup(op))
    assert np.array_equal(out1[0], out2[0]) and np.array_equal(out2[0], out2[1])
    assert np.abs(out1).max() > 0


def _harvested(nx=64, steps=24, seed=0):
    from fluidsim_amd import Discharge, configs
    from fluidsim_amd.learn import LearnedPreconditioner

    c = configs.config0(rtol=1e-10, solver=1)
    c["nx"] = c["ny"] = nx
    d = Discharge(c)
    d.init_gaussian(seed=seed)
    net = LearnedPreconditioner(4000, nx // 2, nx // 2, seed=1)
    for k in range(steps):
        d.run(1)
        net.harvest(d)
    return d, net


def test_training_lowers_the_energy_loss():
    d, net = _harvested()
    assert net.samples > 100
    loss = net.train(300, batch=8, lr=2e-3, seed=3)
    first, last = loss[:20].mean(), loss[-20:].mean()
    print("samples", net.samples, "loss first/last", first, last)
    assert np.all(np.isfinite(loss))
    assert last < first - 0.05 * abs(first)


def test_learned_flexible_cg_reaches_the_plain_solution():
    from fluidsim_amd.learn import LearnedPreconditioner

    d, net = _harvested()
    net.train(400, batch=8, lr=2e-3, seed=3)
    fresh = LearnedPreconditioner(0, net.H, net.W, seed=1)  # same initial weights
    ph0 = d.get("ph")
    res = {}
    for name, solver, nn in (("jacobi", 0, None), ("mg", 1, None), ("no network", 2, None),
                             ("untrained", 2, fresh), ("learned", 2, net)):
        d.set("ph", 1, ph0)
        if nn is not None:
            nn.attach(d)
        elif solver == 2:
            LearnedPreconditioner.detach(d)
        d.set_solver(solver)
        lerr, st = d.lapl_solve()
        res[name] = (lerr, st.iters, st.resid, d.get("ph") - ph0)
        print(name, "lerr", lerr, "iters", st.iters, "true rel residual", st.resid)
    d.set("ph", 1, ph0)
    d.set_solver(1)
    for name in res:
        assert res[name][0] == 0 and res[name][2] <= 1e-10, (name, res[name][:3])
    assert res["learned"][1] < res["untrained"][1]
    assert res["learned"][1] < res["no network"][1]
    assert res["learned"][1] < res["jacobi"][1] / 3
    ref = res["mg"][3][1:-1, 1:-1]
    assert rel_l2(res["learned"][3][1:-1, 1:-1], ref) <= 1e-6


def test_harvest_from_ensemble_members():
    """Residuals harvested from the solves of parameter-sweep ensemble members
    (the training set mixes them with the 512^2 pulse's): every member whose
    learned level fits the network's grid contributes samples, a member of a
    wider grid is skipped, and a network trained on the mixed set keeps the
    learned solve's FP64 residual on a fresh discharge."""
    from fluidsim_amd import Discharge, configs
    from fluidsim_amd.learn import LearnedPreconditioner, harvest_ensemble

    net = LearnedPreconditioner(3000, 64, 64, seed=1)
    info = harvest_ensemble(net, 6, seeds=(0,), samples=400)
    print("ensemble harvest", info, "samples", net.samples)
    assert net.samples >= 400
    # members 198x107, 254x139, 108x193, 205x192, 155x198 have learned levels
    # of at most 64 x 64 (50x27 ... 52x48); 201x79's is 101x40
    assert info["members_used"] == 5 and info["members_skipped"] == 1
    loss = net.train(200, batch=8, lr=2e-3, seed=2)
    assert np.all(np.isfinite(loss)) and loss[-20:].mean() < loss[:20].mean()
    d = Discharge(configs.config0(rtol=1e-10))
    d.init_gaussian(seed=9)
    d.run(3)
    net.attach(d)
    d.set_solver(2)
    lerr, st = d.lapl_solve()
    assert lerr == 0 and st.resid <= 1e-10
    d.set_solver(1)


def test_learned_graph_replay_equals_eager_launches():
    """The learned iterations replayed from a captured HIP graph give bitwise
    the same correction as the same kernels launched one by one."""
    from fluidsim_amd import Discharge, configs

    d, net = _harvested(steps=6)
    net.train(50, batch=8, lr=2e-3, seed=1)
    out = []
    for graphs in (0, 1):
        c = configs.config0(rtol=1e-10, solver=2, graphs=graphs)
        e = Discharge(c)
        e.init_gaussian(seed=0)
        net.attach(e)
        e.run(8)
        out.append((e.get("ph"), e.get("ne", 1), e.log()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def test_learned_graphs_follow_reallocated_activations():
    """One network shared by two discharges whose learned levels differ in size
    (32x32 and 24x24 cells), alternating passes: every switch makes the network
    reallocate its inference activations, so graphs captured over one
    allocation must be captured again, never replayed over freed buffers.
    Graph replay (graphs = 1) equals eager launches (graphs = 0) bitwise."""
    from fluidsim_amd import Discharge, configs

    d, net = _harvested(steps=6)
    net.train(50, batch=8, lr=2e-3, seed=1)
    out = {}
    for graphs in (0, 1):
        hs = []
        for nx in (64, 48):
            c = configs.config0(rtol=1e-10, solver=2, graphs=graphs)
            c["nx"] = c["ny"] = nx
            e = Discharge(c)
            e.init_gaussian(seed=0)
            net.attach(e)
            hs.append(e)
        for _ in range(3):
            for e in hs:
                e.run(2)
        out[graphs] = [(e.get("ph"), e.get("ne", 1), e.log()) for e in hs]
    for a, b in zip(out[0], out[1]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
