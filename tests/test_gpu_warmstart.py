"""Learned warm start (cmg_mlp_*): the network kernels against the CPU FP64
restatement (oracle/mlp.py), and the warm-started Newmark run against the plain one.

Tolerances:
  forward (FP32 kernels)        |Y - Y_orc| <= 1e-5 max|Y_orc|
  loss / gradient (FP32)        rel. 1e-4 / ||g - g_orc|| <= 1e-4 ||g_orc||
  bfloat16 inference (MFMA)     |Y_bf16 - Y_orc| <= 3e-2 max|Y_orc|; against the same
                                roundings in numpy <= 2e-3 max|Y|
  Adam step                     <= 1e-6 relative to the oracle update
  warm-started Newmark steps    converge to the same tolerance; nodes within 1e-8 of
                                the plain run (relative to the displacement)
"""
import numpy as np
import pytest

from oracle import mlp as OM
from cardiomechanics_amd import materials as MT
from cardiomechanics_amd import mesh as M
from cardiomechanics_amd import solver as SV
from cardiomechanics_amd.context import MechanicsContext
from cardiomechanics_amd.hip import DeviceArray
from cardiomechanics_amd.warmstart import WarmStartNet

pytestmark = pytest.mark.gpu


def bf16(a):
    """round float32 values to bfloat16 (nearest even), returned as float64"""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_forward(w, X):
    """8 -> 128 -> 128 -> 3 with bfloat16 inputs, weights and hidden activations"""
    W1, b1 = w[0:1024].reshape(8, 128), w[1024:1152]
    W2, b2 = w[1152:17536].reshape(128, 128), w[17536:17664]
    W3, b3 = w[17664:18048].reshape(128, 3), w[18048:18051]
    h1 = bf16(np.tanh(bf16(X) @ bf16(W1) + b1))
    h2 = bf16(np.tanh(h1 @ bf16(W2) + b2))
    return h2 @ bf16(W3) + b3


@pytest.fixture(scope="module")
def net_ctx():
    ctx = MechanicsContext(M.cube_mesh(2), [MT.cube_guccione()])
    net = WarmStartNet(ctx, seed=3)
    yield ctx, net
    net.close()
    ctx.close()


def test_forward_and_gradient_match_oracle(net_ctx):
    ctx, net = net_ctx
    rng = np.random.default_rng(0)
    rows = 1000                                   # not a multiple of the 64-row tile
    X = rng.normal(size=(rows, 8)).astype(np.float32)
    T = rng.normal(size=(rows, 3)).astype(np.float32)
    w = net.weights().astype(np.float64)
    Xd, Td, Yd = DeviceArray(X), DeviceArray(T), DeviceArray(np.zeros((rows, 3), np.float32))
    net.forward(Xd, Yd, rows)
    net.sync()
    Yo, _, _ = OM.forward(w, X.astype(np.float64))
    assert np.abs(Yd.download() - Yo).max() <= 1e-5 * np.abs(Yo).max()
    net.forward(Xd, Yd, rows, bf16=True)
    net.sync()
    Yb = Yd.download()
    assert np.abs(Yb - Yo).max() <= 3e-2 * np.abs(Yo).max()
    # the matrix-core path against the same roundings in numpy (bfloat16 inputs, weights
    # and hidden activations, wide accumulation): pins the operand layout of the MFMA
    Ye = bf16_forward(w, X)
    err = np.abs(Yb - Ye).max() / np.abs(Ye).max()
    print(f"bf16 inference vs numpy emulation {err:.2e}, vs FP64 {np.abs(Yb - Yo).max() / np.abs(Yo).max():.2e}")
    assert err <= 2e-3
    loss, g = net.loss_grad(Xd, Td, rows, want_grad=True)
    lo, go = OM.loss_grad(w, X.astype(np.float64), T.astype(np.float64))
    print(f"loss {loss:.6e} oracle {lo:.6e}; grad rel err {np.linalg.norm(g - go) / np.linalg.norm(go):.2e}")
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go)
    # every block of the gradient is exercised
    for a, b in [(0, 1024), (1024, 1152), (1152, 17536), (17536, 17664), (17664, 18048), (18048, 18051)]:
        assert np.linalg.norm(g[a:b] - go[a:b]) <= 1e-3 * np.linalg.norm(go[a:b]) + 1e-12


def test_adam_step_matches_oracle(net_ctx):
    ctx, net = net_ctx
    rng = np.random.default_rng(1)
    rows = 300
    X = rng.normal(size=(rows, 8)).astype(np.float32)
    T = rng.normal(size=(rows, 3)).astype(np.float32)
    w0 = rng.normal(0, 0.1, net.n_weights).astype(np.float32)
    net.set_weights(w0)
    _, g = net.loss_grad(DeviceArray(X), DeviceArray(T), rows, want_grad=True)
    net.adam_step(1e-3)
    net.sync()
    w1 = net.weights()
    wo, _, _ = OM.adam(w0.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    assert np.abs(w1 - wo).max() <= 1e-6 * max(1.0, np.abs(wo).max())


def run_window(mesh, mat, steps, dt, weights=None, record=False):
    """Newmark steps of the beat's first `steps` * dt seconds; with `weights` the
    network proposes every step's guess from the third step on; with `record` the
    (features, increment) pairs are returned for training."""
    plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
    out, samples = [], []
    with MechanicsContext(mesh, [mat], newton_atol=1e-12, newton_rtol=1e-12, newton_stol=1e-15) as ctx:
        wn = WarmStartNet(ctx, seed=5)
        if weights is not None:
            wn.set_weights(weights)
        p_prev = 0.0
        for k in range(1, steps + 1):
            t = k * dt
            plug.apply(t, dt)
            p = plug.values(ctx.n_groups)
            ctx.set_pressures(p)
            rate = (p[0] - p_prev) / dt
            p_prev = p[0]
            if record:
                wn.record_before_step(t, rate)
            if weights is not None and k > 2:
                wn.propose(t, rate)
            rc, st = ctx.newmark_step(t, dt)
            assert rc == 2, st.as_dict()
            if record:
                wn.record_after_step()
            out.append((ctx.get_state()["nodes"], st.newton_iterations, st.krylov_iterations, st.guess_source,
                        st.guess_residual_ratio))
        samples = wn.samples
        wn.close()
    return out, samples


def test_warm_start_keeps_the_converged_solution_and_reports_its_gain():
    """Trained on three jittered windows, the network's guess is kept where its residual
    is below the predictor's, Newton converges to the same solution (1e-8), and the
    guess lowers the initial Newton residual (median ratio to the predictor's < 1 over
    the proposals).  Newton and Krylov totals of both runs are reported: a median ratio
    near 0.9 saves no Newton iteration (quadratic convergence needs orders of magnitude),
    so no iteration gain is asserted."""
    mesh = M.lv_mesh(dims=(2, 4, 8))
    base = MT.lv_holzapfel()
    dt, steps = 2e-3, 60
    # training trajectories: Tmax jittered by +-15 %, activation shifted by +-10 ms
    rng = np.random.default_rng(0)
    samples = []
    for b in range(3):
        m = MT.lv_holzapfel(tension_max=60e3 * rng.uniform(0.85, 1.15))
        mb = M.Mesh(coords=mesh.coords, elements=mesh.elements, bases=mesh.bases, bc=mesh.bc,
                    material=mesh.material, activation=np.maximum(mesh.activation + rng.uniform(-0.01, 0.01), 0.0),
                    surfaces=mesh.surfaces, surface_id=mesh.surface_id)
        samples += run_window(mb, m, steps, dt, record=True)[1]
    with MechanicsContext(mesh, [base]) as tctx:
        net = WarmStartNet(tctx, seed=5)
        net.samples = samples
        hist = net.train(epochs=40, lr=2e-3)
        weights = net.weights()
        net.close()
    print(f"training loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    assert hist[-1] < 0.5 * hist[0]
    plain, _ = run_window(mesh, base, steps, dt)
    warm, _ = run_window(mesh, base, steps, dt, weights=weights)
    n_plain = sum(s[1] for s in plain)
    n_warm = sum(s[1] for s in warm)
    k_plain = sum(s[2] for s in plain)
    k_warm = sum(s[2] for s in warm)
    kept = sum(1 for s in warm if s[3] == 2)
    ratio = np.median([s[4] for s in warm if s[3] in (2, 3)])
    print(f"Newton iterations plain {n_plain} warm {n_warm}; Krylov plain {k_plain} warm {k_warm}; "
          f"learned guess kept in {kept}/{steps - 2} steps, median residual ratio {ratio:.3f}")
    for (xp, *_), (xw, *_) in zip(plain, warm):
        u = xp - mesh.coords.ravel()
        assert np.linalg.norm(xw - xp) <= 1e-8 * max(np.linalg.norm(u), 1e-300)
    assert all(s[3] in (2, 3) for s in warm[2:])
    assert kept > 0 and ratio < 1.0
