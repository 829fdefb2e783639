"""Pins of the CPU oracle against mathematics it must satisfy (no reference test
holds golden numbers for this path).  Each check is chosen so that a plausible
restatement mistake -- a dropped term, a wrong sign, a transposed operand, a wrong
midside-node order -- fails it."""
import math
import os

import numpy as np
import pytest

import oracle as O
from cardiomechanics_amd import materials as MT
from cardiomechanics_amd import mesh as M

RNG = np.random.default_rng(7)


def rand_rotation(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return q


def rand_F(rng, amp=0.15):
    F = np.eye(3) + amp * rng.normal(size=(3, 3))
    assert np.linalg.det(F) > 0
    return F


def straight_t10(rng, curved=0.0):
    X4 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 1e-3
    X4 = X4 + 2e-4 * rng.normal(size=(4, 3))
    if np.linalg.det(np.stack([X4[1] - X4[0], X4[2] - X4[0], X4[3] - X4[0]])) < 0:
        X4[[1, 2]] = X4[[2, 1]]
    mids = np.array([(X4[a] + X4[b]) / 2 for a, b in M.T10_EDGES])
    mids += curved * 1e-3 * rng.normal(size=mids.shape)
    return np.concatenate([X4, mids])


GUC = MT.cube_guccione()
HO = MT.lv_holzapfel()
HO0 = MT.lv_holzapfel()
HO0.params[9] = 0.0           # step Heaviside, H' = 0 away from I4 = 1


# ---------------------------------------------------------------- analytic functions
def test_analytic_functions_closed_forms():
    w = 2 * math.pi / 0.3
    fn = [0.0, 0.3, 1, w, 0.0] + [0] * 7
    for t in np.linspace(0, 0.3, 13):
        assert O.analytic(O.FN_SINUS, fn, t) == pytest.approx(math.sin(w * t / 2) ** 2, abs=1e-15)
    assert O.analytic(O.FN_SINUS, fn, 0.31) == 0.0
    assert O.analytic(O.FN_SINUS, fn, -0.01) == 0.0
    lin = [0.2, 0.6, 1] + [0] * 9
    assert O.analytic(O.FN_LINEAR, lin, 0.1) == 0.0
    assert O.analytic(O.FN_LINEAR, lin, 0.4) == pytest.approx(0.5)
    assert O.analytic(O.FN_LINEAR, lin, 0.7) == 1.0
    lin[2] = 0
    assert O.analytic(O.FN_LINEAR, lin, 0.7) == 0.0
    kind, fn = MT.function_params("DoubleHill", start=0.0, stop=10.0)
    ts = np.linspace(0, 0.8, 1001)
    vals = [O.analytic(O.FN_DOUBLEHILL, fn, t) for t in ts]
    assert max(vals) == pytest.approx(1.0, abs=1e-12)     # normalised by its own maximum
    # host and oracle evaluations agree
    for t in (0.05, 0.2, 0.41):
        assert MT.analytic(kind, fn, t) == pytest.approx(O.analytic(kind, fn, t), rel=1e-14)


def test_sinus_driver_closed_form():
    """SinusDriver (CBAnalyticFunction.h:105-119): sin(3.14159 (t - t0) / (t1 - t0)) on the
    closed window [t0, t1], 0 outside.  The truncated pi leaves sin(3.14159) = 2.65e-6 at t1."""
    fn = [0.1, 0.5] + [0] * 10
    f = lambda t: O.analytic(O.FN_SINUSDRIVER, fn, t)     # noqa: E731
    assert f(0.1) == 0.0 and f(0.0999) == 0.0 and f(0.5001) == 0.0
    assert f(0.5) == pytest.approx(2.6535897933527304e-06, rel=1e-9)
    assert f(0.3) == pytest.approx(0.99999999999912, abs=1e-14)        # the top: sin(3.14159 / 2)
    for t in (0.15, 0.2):   # symmetric about the middle only to the truncation of pi
        assert 1e-7 < abs(f(t) - f(0.6 - t)) < 3e-6
    assert f(0.1 + 0.4 / 6) == pytest.approx(math.sin(3.14159 / 6), rel=1e-14)


def test_double_hill_shape_at_its_hill_offsets():
    """DoubleHill (CBAnalyticFunction.h:123-176, Stergiopulos 1996): at t - onset = the
    contraction offset the contraction Hill term is exactly 1/2, at the relaxation offset
    the relaxation term is exactly 1/2 -- swapping the offsets or the rate constants fails
    these; before onset and outside the window it is 0, and it repeats with the period."""
    c_off, r_off, a, b, onset, period = 0.215, 0.362, 1.32, 21.9, 0.05, 0.8
    kind, fn = MT.function_params("DoubleHill", start=0.0, stop=10.0, period=period, contr_offset=c_off,
                                  relax_offset=r_off, contr_rate=a, relax_rate=b, onset=onset)
    mx = fn[11]
    f = lambda t: O.analytic(kind, fn, t)     # noqa: E731
    assert f(onset + c_off) == pytest.approx(0.5 / (1.0 + (c_off / r_off) ** b) / mx, rel=1e-13)
    hr = (r_off / c_off) ** a
    assert f(onset + r_off) == pytest.approx(hr / (1.0 + hr) * 0.5 / mx, rel=1e-13)
    assert f(onset / 2) == 0.0 and f(onset) == 0.0
    for t in (0.1, 0.3, 0.5):
        assert f(t + period) == pytest.approx(f(t), rel=1e-12)
    fn2 = list(fn)
    fn2[1] = 0.6
    assert O.analytic(kind, fn2, 0.61) == 0.0 and O.analytic(kind, fn2, 0.6) > 0.0


def test_adam_restatement_matches_torch():
    """The warm-start trainer's Adam step (oracle/mlp.py) against torch.optim.Adam on the
    same gradients for five steps."""
    torch = pytest.importorskip("torch")
    from oracle import mlp as OM

    rng = np.random.default_rng(3)
    w = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(5)]
    tw = torch.tensor(w.copy(), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tw], lr=2e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros(50)
    v = np.zeros(50)
    for t, g in enumerate(grads, start=1):
        opt.zero_grad()
        tw.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        w, m, v = OM.adam(w, g, m, v, t, 2e-3)
        np.testing.assert_allclose(w, tw.detach().numpy(), rtol=0, atol=1e-15)


# ---------------------------------------------------------------- laws
@pytest.mark.parametrize("mat", [GUC, HO, HO0])
def test_stress_free_reference(mat):
    rc, S = O.pk2(mat, np.eye(3))
    assert rc == O.SUCCESS
    assert np.abs(S).max() < 1e-9


def test_guccione_stress_is_energy_derivative():
    for _ in range(5):
        F = rand_F(RNG)
        rc, S = O.pk2(GUC, F)
        rc2, P = O.energy_dF_cs(GUC, F)
        assert rc == rc2 == O.SUCCESS
        np.testing.assert_allclose(F @ S, P, rtol=1e-12, atol=1e-9 * np.abs(P).max())
        np.testing.assert_allclose(S, S.T, rtol=0, atol=1e-12 * np.abs(S).max())


def test_guccione_energy_closed_form_diagonal():
    lam = np.array([1.1, 0.95, 0.97])
    E = 0.5 * (lam ** 2 - 1)
    C, bf, bt, bfs, K = GUC.params[:5]
    Q = bf * E[0] ** 2 + bt * (E[1] ** 2 + E[2] ** 2)
    J = lam.prod()
    W = C / 2 * (math.exp(Q) - 1) + K / 2 * (J - 1) ** 2
    rc, Wo = O.energy(GUC, np.diag(lam))
    assert Wo == pytest.approx(W, rel=1e-13)


@pytest.mark.parametrize("mat", [HO, HO0])
def test_holzapfel_stress_vs_energy_with_coded_quirks(mat):
    """S_code = 2 dW/dC - a e^{b(I1-3)} I1/3 C^-1 - 2 H'(I4f) W_f f(x)f - 2 H'(I4s) W_s s(x)s:
    the Jm = pow(J,-2/3) = 1 quirk leaves the I1/3 C^-1 term without its J^-2/3, and
    the stress omits the derivative of the logistic switch."""
    a, b, af, bf, as_, bs, afs, bfs, kappa, k = mat.params[:10]
    for _ in range(6):
        F = rand_F(RNG, 0.12)
        rc, S = O.pk2(mat, F)
        rc2, P = O.energy_dF_cs(mat, F)
        assert rc == rc2 == O.SUCCESS
        S_energy = np.linalg.solve(F, P)
        C = F.T @ F
        I1 = np.trace(C)
        expected = -(a * math.exp(b * (I1 - 3)) * I1 / 3) * np.linalg.inv(C)
        for idx, (aa, bb) in ((0, (af, bf)), (1, (as_, bs))):
            I4 = C[idx, idx]
            if k != 0:
                H = 1 / (1 + math.exp(-k * (I4 - 1)))
                dH = k * H * (1 - H)
                Wq = aa / (2 * bb) * (math.exp(bb * (I4 - 1) ** 2) - 1)
                expected[idx, idx] -= 2 * dH * Wq
        np.testing.assert_allclose(S - S_energy, expected, rtol=1e-9, atol=1e-9 * np.abs(S).max())


def test_active_stress_is_derivative_of_T_sqrt_I4f():
    T = 6e4
    for _ in range(4):
        F = rand_F(RNG)
        Sa = O.active_stress(F, T)
        dW = np.zeros((3, 3))
        dW[:, 0] = T * F[:, 0] / np.linalg.norm(F[:, 0])     # d(T |F e_f|)/dF
        np.testing.assert_allclose(Sa, np.linalg.solve(F, dW), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(O.active_stress(np.eye(3), T), np.diag([T, 0, 0]))


def _tanh_mat(**kw):
    return MT.with_tanh(MT.cube_guccione(), **kw)


def test_tanh_tension_limits_and_window():
    """TanH tension (CBTensionModelTanH.cpp:71-97): zero before activation + EMD and at
    the end of the duration, zero for fibre stretch at or below l0, periodic in the
    cycle length, and in the saturated limit (t_s >> tau_c, t_dur - t_s >> tau_r) equal
    to Tmax S_peak tanh(ld (lambda - l0))."""
    m = _tanh_mat()
    emd, dur = 0.015, 0.3
    assert O.tanh_tension(m, 1.1, 0.010, 0.0) == 0.0              # t - t_act < EMD
    assert O.tanh_tension(m, 1.1, 0.2 + emd - 1e-9, 0.2) == 0.0    # just before onset of an element at 0.2 s
    assert O.tanh_tension(m, 1.1, emd + dur, 0.0) == 0.0           # t_s = t_dur
    assert O.tanh_tension(m, 1.1, 0.2, 0.0) > 0.0
    assert O.tanh_tension(m, 0.7, 0.2, 0.0) == 0.0                 # phi = 0
    assert O.tanh_tension(m, 0.65, 0.2, 0.0) == 0.0                # phi < 0 clipped
    for t in (0.05, 0.13, 0.27):
        assert O.tanh_tension(m, 1.05, t + 1.0, 0.0) == pytest.approx(O.tanh_tension(m, 1.05, t, 0.0), rel=1e-12)
    assert O.tanh_tension(m, 1.05, 0.13, 0.0) == O.tanh_tension(_tanh_mat(start_time=0.02), 1.05, 0.15, 0.0)
    sat = _tanh_mat(Duration=20.0, ContractionTime=0.01, RelaxationTime=0.01, LengthDependenceUpstrokeTime=0.0)
    for lam in (0.8, 1.0, 1.2):
        got = O.tanh_tension(sat, lam, 10.0, 0.0)
        assert got == pytest.approx(1000.0 * 120.0 * math.tanh(5.0 * (lam - 0.7)), rel=1e-14)
    # monotone in stretch (length dependence) inside the window
    lams = np.linspace(0.75, 1.3, 12)
    vals = [O.tanh_tension(m, l, 0.1, 0.0) for l in lams]
    assert np.all(np.diff(vals) > 0)


def test_tanh_active_stress_and_element_tangent():
    """The TanH active PK2 is kff T(lambda)/lambda on ff only; the element's complex-step
    tangent through the stretch-dependent tension matches the reference's central-
    difference tangent (transposed, CBElementSolidT10.cpp:511-567)."""
    m = _tanh_mat()
    rng = np.random.default_rng(5)
    F = np.eye(3) + 0.1 * rng.standard_normal((3, 3))
    S0 = O.active_stress_model(m, F, 0.1)
    lam = np.linalg.norm(F[:, 0])
    assert S0[0, 0] == pytest.approx(O.tanh_tension(m, lam, 0.1, 0.0) / lam, rel=1e-14)
    assert np.count_nonzero(S0) == 1
    X, x, dN, bases, w = element_case(np.random.default_rng(11))   # own stream: the module RNG feeds later tests
    bc = np.zeros(30, np.uint8)
    rc, Kcs = O.t10_cs_tangent(x.ravel(), dN, bases, w, bc, m, time=0.1)
    rc2, Kfd = O.t10_fd_tangent(x.ravel(), dN, bases, w, bc, m, time=0.1, eps=1e-8)
    assert rc == rc2 == O.SUCCESS
    np.testing.assert_allclose(Kfd.T, Kcs, rtol=0, atol=1e-5 * np.abs(Kcs).max())
    # the active tension contributes to the tangent (phi ~ 0 switches it off)
    flat = _tanh_mat(DegreeOfLengthDependence=1e-9)
    rc, Kflat = O.t10_cs_tangent(x.ravel(), dN, bases, w, bc, flat, time=0.1)
    assert np.abs(Kflat - Kcs).max() > 1e-3 * np.abs(Kcs).max()


def test_pk2_tangent_complex_step_matches_central_difference():
    for mat in (GUC, HO):
        F = rand_F(RNG, 0.1)
        rc, D = O.pk2_dF_cs(mat, F)
        h = 1e-6
        Dfd = np.zeros((9, 9))
        for b in range(9):
            Fp, Fm = F.copy().ravel(), F.copy().ravel()
            Fp[b] += h
            Fm[b] -= h
            Dfd[:, b] = (O.pk2(mat, Fp.reshape(3, 3))[1] - O.pk2(mat, Fm.reshape(3, 3))[1]).ravel() / (2 * h)
        np.testing.assert_allclose(D, Dfd, rtol=1e-6, atol=1e-6 * np.abs(D).max())


def test_corrupt_element_detection():
    F = np.diag([1.0, 1.0, -1.0])
    for mat in (GUC, HO):
        rc, _ = O.pk2(mat, F)
        assert rc == O.CORRUPT_ELEMENT
    m2 = MT.cube_guccione()
    m2.ignore_corrupt = 1
    rc, _ = O.pk2(m2, np.diag([1.0, 1.0, 0.5]))
    assert rc == O.SUCCESS


# ---------------------------------------------------------------- shape functions
@pytest.mark.parametrize("curved", [0.0, 0.05])
def test_t10_shape_identities(curved):
    X = straight_t10(RNG, curved)
    dN, detJ = O.t10_shape(X.ravel())
    dN = dN.reshape(5, 10, 3)
    for q in range(5):
        np.testing.assert_allclose(dN[q].sum(axis=0), 0, atol=1e-9)               # partition of unity
        np.testing.assert_allclose(X.T @ dN[q], np.eye(3), atol=1e-12)            # isoparametric identity
        A = RNG.normal(size=(3, 3))
        np.testing.assert_allclose((X @ A.T).T @ dN[q], A, atol=1e-11)            # linear patch
    if curved == 0.0:
        v6 = np.dot(X[1] - X[0], np.cross(X[2] - X[0], X[3] - X[0]))
        assert detJ == pytest.approx(v6, rel=1e-12)                               # detJ = 6 V
        # straight-edged T10 gradients at the QPs from the T4 gradients: (4 l_i - 1) dL_i
        dL, d4 = O.t4_shape(X[:4].ravel())
        dL = dL.reshape(4, 3)
        for q, l in enumerate(M.T10_QP_BARY):
            for i in range(4):
                np.testing.assert_allclose(dN[q, i], (4 * l[i] - 1) * dL[i], atol=1e-9 * np.abs(dL).max())
            for m_, (i, j) in enumerate(M.T10_EDGES):
                np.testing.assert_allclose(dN[q, 4 + m_], 4 * (l[i] * dL[j] + l[j] * dL[i]),
                                           atol=1e-9 * np.abs(dL).max())


def test_vectorised_shape_matches_oracle():
    mesh = M.cube_mesh(2)
    dn, dj = M.t10_shape_derivatives(mesh.coords, mesh.elements)
    ed = O.ElementData(mesh)
    np.testing.assert_array_equal(dn.reshape(-1, 150), ed.dNdX)
    np.testing.assert_array_equal(dj, ed.w)


# ---------------------------------------------------------------- element residual
def element_case(rng, npe=10, amp=0.02):
    if npe == 10:
        X = straight_t10(rng, 0.02)
        dN, w = O.t10_shape(X.ravel())
        bases = np.concatenate([rand_rotation(rng).ravel() for _ in range(5)])
    else:
        X = straight_t10(rng)[:4]
        dN, d = O.t4_shape(X.ravel())
        w = d / 6
        bases = rand_rotation(rng).ravel()
    x = X + amp * 1e-3 * rng.normal(size=X.shape)
    return X, x, dN, bases, w


@pytest.mark.parametrize("npe", [10, 4])
def test_residual_is_energy_gradient_and_rigid_invariant(npe):
    mat = MT.cube_guccione()
    for _ in range(3):
        X, x, dN, bases, w = element_case(RNG, npe)
        bc = np.zeros(3 * npe, np.uint8)
        if npe == 10:
            rc, f = O.t10_residual(x.ravel(), dN, bases, w, bc, mat)
            rc2, g = O.t10_energy_grad_cs(x.ravel(), dN, bases, w, mat)
            np.testing.assert_allclose(f, g, rtol=1e-11, atol=1e-11 * np.abs(f).max())
        else:
            rc, f = O.t4_residual(x.ravel(), dN, bases, w, bc, mat)
        np.testing.assert_allclose(f.reshape(-1, 3).sum(axis=0), 0, atol=1e-10 * np.abs(f).max())
        # rigid motion of the reference configuration: no force
        R = rand_rotation(RNG)
        xr = X @ R.T + np.array([1e-3, -2e-3, 5e-4])
        fr = (O.t10_residual if npe == 10 else O.t4_residual)(xr.ravel(), dN, bases, w, bc, mat)[1]
        assert np.abs(fr).max() < 1e-9 * np.abs(f).max()


def test_residual_bc_zeroing_and_active_tension():
    mat = MT.lv_holzapfel()
    X, x, dN, bases, w = element_case(RNG)
    bc = np.zeros(30, np.uint8)
    bc[[0, 4, 8]] = 1
    rc, f = O.t10_residual(x.ravel(), dN, bases, w, bc, mat, time=0.1, tact=0.0)
    assert rc == O.SUCCESS and f[0] == f[4] == f[8] == 0.0
    # active tension at the reference configuration: f_a = V/4 sum_q B (T e_f (x) e_f) B^-1 g_a
    T = mat.tension_max * math.sin(2 * math.pi / 0.3 * 0.1 / 2) ** 2
    rc, f0 = O.t10_residual(X.ravel(), dN, bases, w, np.zeros(30, np.uint8), mat, time=0.1)
    exp = np.zeros(30)
    for q in range(4):
        B = bases[9 * (q + 1):9 * (q + 2)].reshape(3, 3)
        P = B @ np.diag([T, 0, 0]) @ np.linalg.inv(B)
        g = dN[30 * q:30 * q + 30].reshape(10, 3)
        exp += w / 24 * (g @ P.T).ravel()
    np.testing.assert_allclose(f0, exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max())


# ---------------------------------------------------------------- tangents
@pytest.mark.parametrize("mat", [GUC, HO])
def test_t10_fd_tangent_is_transposed_complex_step(mat):
    X, x, dN, bases, w = element_case(RNG)
    bc = np.zeros(30, np.uint8)
    bc[[3, 4, 5]] = 1
    rc, Kcs = O.t10_cs_tangent(x.ravel(), dN, bases, w, bc, mat, time=0.1)
    rc2, Kfd = O.t10_fd_tangent(x.ravel(), dN, bases, w, bc, mat, time=0.1, eps=1e-8)
    assert rc == rc2 == O.SUCCESS
    np.testing.assert_allclose(Kfd.T, Kcs, rtol=0, atol=1e-5 * np.abs(Kcs).max())
    assert np.abs(Kcs[:, 3:6]).max() == 0 and np.abs(Kcs[3:6, :]).max() == 0
    if mat is GUC:
        np.testing.assert_allclose(Kcs, Kcs.T, atol=1e-11 * np.abs(Kcs).max())


def test_t4_fd_tangent_is_plain_complex_step():
    X, x, dN, basis, vol = element_case(RNG, 4)
    bc = np.zeros(12, np.uint8)
    rc, Kcs = O.t4_cs_tangent(x.ravel(), dN, basis, vol, bc, HO, time=0.1)
    rc, Kfd = O.t4_fd_tangent(x.ravel(), dN, basis, vol, bc, HO, time=0.1, eps=1e-8)
    np.testing.assert_allclose(Kfd, Kcs, rtol=0, atol=1e-5 * np.abs(Kcs).max())


# ---------------------------------------------------------------- mass
def test_mass_matrices_sum_to_element_mass():
    X = straight_t10(RNG)
    dN, detJ = O.t10_shape(X.ravel())
    rho = 1082.0
    V = detJ / 6
    bc = np.zeros(30, np.uint8)
    Mc = O.t10_mass_consistent(detJ, rho, bc)
    for k in range(3):
        assert Mc[k::3, k::3].sum() == pytest.approx(rho * V, rel=1e-13)
        assert np.abs(Mc[k::3, (k + 1) % 3::3]).max() == 0
    np.testing.assert_array_equal(Mc, Mc.T)
    Ml = O.t10_mass_lumped(V, rho, bc)
    np.testing.assert_allclose(np.diag(Ml), rho * V / 10)
    bc[7] = 1
    assert np.abs(O.t10_mass_consistent(detJ, rho, bc)[7]).max() == 0
    M4 = O.t4_mass_consistent(V, rho, np.zeros(12, np.uint8))
    assert M4[0::3, 0::3].sum() == pytest.approx(rho * V)
    # T4 lumped: the row sums of the consistent mass on the diagonal (row-sum lumping)
    L4 = O.t4_mass_lumped(V, rho, np.zeros(12, np.uint8))
    np.testing.assert_allclose(L4, np.diag(M4.sum(axis=1)), rtol=1e-15)
    assert L4[0::3, 0::3].sum() == pytest.approx(rho * V)
    bc4 = np.zeros(12, np.uint8)
    bc4[4] = 1
    assert O.t4_mass_lumped(V, rho, bc4)[4, 4] == 0


# ---------------------------------------------------------------- surfaces
def test_t3_pressure_force_and_volume():
    tri = np.array([[0, 0, 0], [2e-3, 0, 0], [0, 1e-3, 0]], float)
    p = 1500.0
    f = O.surface_forces(tri.ravel(), np.zeros(9, np.uint8), p).reshape(3, 3)
    n = np.cross(tri[0] - tri[1], tri[0] - tri[2])
    area = 0.5 * np.linalg.norm(n)
    np.testing.assert_allclose(f.sum(axis=0), -p * area * n / np.linalg.norm(n), rtol=1e-14)
    np.testing.assert_allclose(f[0], f[1])
    K = O.surface_cs_tangent(tri.ravel(), np.zeros(9, np.uint8), p)
    h = 1e-9
    for c in range(9):
        xp, xm = tri.ravel().copy(), tri.ravel().copy()
        xp[c] += h
        xm[c] -= h
        fd = (O.surface_forces(xp, np.zeros(9, np.uint8), p) - O.surface_forces(xm, np.zeros(9, np.uint8), p)) / (2 * h)
        np.testing.assert_allclose(K[:, c], fd, rtol=1e-6, atol=1e-6 * np.abs(K).max())


def closed_cube_triangles(L=1e-2):
    mesh = M.cube_mesh(1, length=L, order=4)
    tets = mesh.elements
    faces = M.orient_faces_outward(mesh.coords, M.boundary_faces(tets), tets)
    return mesh.coords, faces


def test_closed_surface_volume_independent_of_reference_point():
    coords, faces = closed_cube_triangles()
    for ref in ((0, 0, 0), (1, 1, 1), (-0.3, 0.2, 5.0)):
        vol = sum(O.surface_volume(coords[f].ravel(), ref) for f in faces)
        assert vol == pytest.approx(1e-6, rel=1e-12)


def test_t6_flat_triangle_total_force_and_fd_consistency():
    c = np.array([[0, 0, 0], [1e-3, 0, 0], [0, 1e-3, 0]], float)
    mids = np.array([(c[0] + c[1]) / 2, (c[1] + c[2]) / 2, (c[2] + c[0]) / 2])
    x = np.concatenate([c, mids]).ravel()
    p = 800.0
    f = O.surface_forces(x, np.zeros(18, np.uint8), p).reshape(6, 3)
    area = 0.5e-6
    np.testing.assert_allclose(f.sum(axis=0), [0, 0, -p * area], rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(f[:3], 0, atol=1e-12 * np.abs(f).max())   # corner loads vanish at these QPs
    v = O.surface_volume(x, (0, 0, -1e-3))
    assert v == pytest.approx(area * 1e-3 / 3, rel=1e-12)
    K = O.surface_cs_tangent(x, np.zeros(18, np.uint8), p)
    xd = x + 1e-5 * RNG.normal(size=18)
    Kfd = O.t6_fd_tangent(xd, np.zeros(18, np.uint8), p, eps=1e-9)
    Kc = O.surface_cs_tangent(xd, np.zeros(18, np.uint8), p)
    np.testing.assert_allclose(Kfd, Kc, rtol=0, atol=1e-5 * np.abs(Kc).max())
    assert np.isfinite(K).all()


def test_mlp_oracle_against_torch_autograd_and_finite_differences():
    """The warm-start network's CPU restatement: forward and gradient against torch
    (an independent implementation, FP64 autograd) and central differences."""
    import torch

    from oracle import mlp

    H = 16
    rng = np.random.default_rng(11)
    w = rng.normal(0, 0.3, mlp.n_weights(H))
    X = rng.normal(size=(37, 8))
    T = rng.normal(size=(37, 3))
    W1, b1, W2, b2, W3, b3 = mlp.unpack(w, H)
    tw = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in (W1, b1, W2, b2, W3, b3)]
    tx, tt = torch.tensor(X), torch.tensor(T)
    l1 = torch.nn.functional.linear(tx, tw[0].T, tw[1]).tanh()
    l2 = torch.nn.functional.linear(l1, tw[2].T, tw[3]).tanh()
    y = torch.nn.functional.linear(l2, tw[4].T, tw[5])
    lt = torch.nn.functional.mse_loss(y, tt)
    lt.backward()
    gt = np.concatenate([t.grad.numpy().ravel() for t in tw])
    lo, go = mlp.loss_grad(w, X, T, H)
    assert abs(lo - lt.item()) <= 1e-14 * abs(lo)
    assert np.abs(go - gt).max() <= 1e-12 * np.abs(gt).max()
    # central differences on a few coordinates of every block
    h = 1e-6
    for k in [0, 5, 8 * H + 3, 9 * H + 17, 9 * H + H * H + 2, mlp.n_weights(H) - 1]:
        e = np.zeros_like(w)
        e[k] = h
        fd = (mlp.loss(w + e, X, T, H) - mlp.loss(w - e, X, T, H)) / (2 * h)
        assert abs(fd - go[k]) <= 1e-7 * max(1.0, abs(go[k]))


def test_lv_generator_orients_the_endocardium_like_the_reference_mesh():
    """Endocardial normals n = (b-a) x (c-a) point into the cavity (towards the long
    axis) in the reference's ellipsoid.sur and in the generator, so a negative
    pressure amplitude inflates both (the oracle: the cavity volume grows)."""
    from cardiomechanics_amd import materials as MT
    from cardiomechanics_amd import mesh as M
    from oracle.drivers import CpuMechanics

    def inward_fraction(X, S, axis_xy):
        a, b, c = X[S[:, 0]], X[S[:, 1]], X[S[:, 2]]
        n = np.cross(b - a, c - a)
        rad = (a + b + c) / 3 - np.array([axis_xy[0], axis_xy[1], 0.0])
        rad[:, 2] = 0.0
        return np.mean(np.einsum("ij,ij->i", n, rad) < 0)

    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    z = np.load(os.path.join(golden, "ellipsoid_subset.npz"))
    assert inward_fraction(z["coords"], z["surfaces"], (0.0, 0.0)) == 1.0
    mesh = M.lv_mesh(dims=(2, 4, 8))
    assert inward_fraction(mesh.coords, mesh.surfaces, (0.0, 0.0)) == 1.0
    cpu = CpuMechanics(mesh, [MT.cube_guccione()])
    v0 = cpu.cavity_volumes()[0]
    cpu.pressures[:] = -500.0
    cpu.static_step(0.0, 1.0)
    cpu.static_step(1.0, 1.0)
    assert v0 < 0 and cpu.cavity_volumes()[0] < v0      # |V| grows


@pytest.mark.parametrize("alpha_over_w,beta_r_w", [(0.2, 0.0), (0.0, 0.2), (0.1, 0.1)])
def test_rayleigh_damping_reduces_to_the_damped_oscillator(alpha_over_w, beta_r_w):
    """Small vibrations of the clamped cube from rest in its lowest mode phi (K0 phi =
    w^2 M phi): C = alpha M + beta_R K is diagonal in the modes, so the modal coordinate
    q = phi^T M u of the oracle's Newmark run equals Newmark on q'' + (alpha + beta_R
    w^2) q' + w^2 q = 0 step for step, and follows the exact damped oscillation to the
    trapezoidal rule's O((w dt)^2) phase error."""
    import scipy.linalg as sla
    from oracle.drivers import CpuMechanics

    mesh, mats = M.cube_mesh(n=2), [MT.cube_guccione()]
    cpu = CpuMechanics(mesh, mats)
    X = mesh.coords.ravel().copy()
    free = ~cpu.bcdof
    K0 = cpu.tangent(X, 0.0, include_pressure=False).toarray()[np.ix_(free, free)]
    Mf = cpu.M.toarray()[np.ix_(free, free)]
    w2, V = sla.eigh(0.5 * (K0 + K0.T), Mf)
    w = math.sqrt(w2[0])
    phi = np.zeros(cpu.n)
    phi[free] = V[:, 0]                                   # phi^T M phi = 1
    alpha, beta_r = alpha_over_w * w, beta_r_w / w
    cpu = CpuMechanics(mesh, mats, rayleigh=(alpha, beta_r))
    eps = 1e-8 / np.abs(phi).max()
    cpu.nodes = X + eps * phi
    cpu.acc = -w2[0] * eps * phi
    dt = (2 * math.pi / w) / 40
    beta, gamma = 0.25, 0.5
    c, k = alpha + beta_r * w2[0], w2[0]
    q, v, a = eps, 0.0, -k * eps
    zw = 0.5 * c
    wd = math.sqrt(k - zw * zw)
    e_scale = 0.5 * k * eps * eps          # the mode's initial potential energy
    dissipated = 0.0
    for n in range(1, 41):
        cpu.newmark_step(n * dt, dt, tol=1e-13)
        # the same Newmark step on the modal equation
        dp = dt * v + dt * dt * 0.5 * (1 - 2 * beta) * a
        vp = v + (1 - gamma) * dt * a
        m1, c1 = 1 / (beta * dt * dt), gamma / (beta * dt)
        inc = (m1 * dp + c * c1 * dp - c * vp - k * q) / (m1 + c * c1 + k)
        a = m1 * (inc - dp)
        v = vp + gamma * dt * a
        q = q + inc
        qf = phi @ (cpu.M @ (cpu.nodes - X))
        assert abs(qf - q) <= 1e-4 * eps, (n, qf / eps, q / eps)
        # corrector energies in the modal limit: v^T M v / 2 = q'^2 / 2 and
        # (C v) . d = c q' dq (phi^T C phi = alpha + beta_R w^2)
        dissipated += c * v * inc
        assert abs(cpu.kinetic_energy - 0.5 * v * v) <= 1e-4 * e_scale, n
        assert abs(cpu.damping_energy - dissipated) <= 1e-4 * e_scale, n
        t = n * dt
        exact = eps * math.exp(-zw * t) * (math.cos(wd * t) + zw / wd * math.sin(wd * t))
        assert abs(q - exact) <= 2e-2 * eps, (n, q / eps, exact / eps)
    # damping ratio 0.1: after one period the amplitude is ~0.53 of the undamped one
    assert abs(q - eps * math.cos(w * 40 * dt)) > 0.3 * eps


def test_element_data_does_not_alias_the_mesh_constraints():
    """Editing the oracle's per-element constraint mask (element checks without
    constraints) must leave the mesh's boundary conditions intact."""
    mesh = M.cube_mesh(2)
    bc = mesh.bc.copy()
    ed = O.ElementData(mesh)
    ed.bcmask[:] = 0
    assert np.array_equal(mesh.bc, bc)


def test_damping_matrix_uses_the_reference_layout_of_the_tangent():
    """The reference builds C = alpha M + beta_R K from its finite-difference Jacobian
    (CBSolverNewmarkBeta.cpp:657-668), which T10 assembles with row = perturbed DOF
    (CBElementSolidT10.cpp:559-565), i.e. K^T.  The oracle's C applied to w must equal
    alpha M w + beta_R (assembled reference-layout FD Jacobian) w, to finite-difference
    noise, on Holzapfel elements with active tension, whose tangent as coded is
    unsymmetric: the untransposed K misses by far more than that noise."""
    import scipy.sparse as sp
    from oracle.drivers import CpuMechanics

    mesh, mats = M.lv_mesh(dims=(2, 4, 8)), [MT.lv_holzapfel()]
    alpha, beta_r = 500.0, 0.01
    cpu = CpuMechanics(mesh, mats, rayleigh=(alpha, beta_r))
    L = np.ptp(mesh.coords, axis=0).max()
    cpu.nodes = (mesh.coords + M.smooth_displacement(mesh.coords, L, 0.02, 1234)).ravel()
    t = 0.12
    C = cpu.damping_matrix(t)
    ed = O.ElementData(mesh)
    rc, Jfd, _ = O.tangent_batch(ed, cpu.nodes.reshape(-1, 3), mats, t, kind="fd", eps=1e-9)
    assert rc == O.SUCCESS
    nd = Jfd.shape[1]
    rows = np.repeat(cpu.dofs, nd, axis=1).ravel()
    cols = np.tile(cpu.dofs, (1, nd)).ravel()
    A_ref = sp.csr_matrix((Jfd.ravel(), (rows, cols)), shape=(cpu.n, cpu.n))
    rc, Kcs, _ = O.tangent_batch(ed, cpu.nodes.reshape(-1, 3), mats, t, kind="cs")
    K = sp.csr_matrix((Kcs.ravel(), (rows, cols)), shape=(cpu.n, cpu.n))
    D = sp.diags((~cpu.bcdof).astype(float))
    w = np.random.default_rng(11).normal(size=cpu.n)
    want = D @ (alpha * (cpu.M @ (D @ w)) + beta_r * (A_ref @ (D @ w)))
    got = C @ w
    wrong = D @ (alpha * (cpu.M @ (D @ w)) + beta_r * (K @ (D @ w)))
    scale = np.linalg.norm(beta_r * (A_ref @ (D @ w)))
    err = np.linalg.norm(got - want) / scale
    miss = np.linalg.norm(wrong - want) / scale
    print(f"C w vs reference layout: {err:.2e}; untransposed K: {miss:.2e}")
    assert err <= 1e-5
    assert miss >= 1e-3
