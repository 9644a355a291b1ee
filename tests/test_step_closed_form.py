"""The RK4 step's orchestration pinned by closed forms.

A spatially uniform state has vanishing derivatives, so away from the ghost
zones the BSSN equations of bssn_rhs.f90 (GAUGE 0 and 2, vacuum) reduce to
ordinary differential equations of the point values: every Christoffel
symbol, Ricci component and derivative is zero and what is left is

    chi'     = 2/3 chi alpha K                               (:160)
    K'       = alpha (K^2/3 + A_kl A^kl)                     (:783-787)
    g_ij'    = -2 alpha A_ij                                 (:170-180)
    A_ij'    = alpha/3 (2/3 K^2 - A_kl A^kl) g_ij            (the trace of
               + alpha (K A_ij - 2 A_ik A^k_j)                R_ij replaced by
                                                              the Hamiltonian
                                                              constraint, the
                                                              "bam" form of
                                                              :665-760)
    Gamma^i' = 0
    alpha'   = -2 alpha K                                    (:791)
    beta'    = 3/4 B                                         (:793-795)
    B'       = -eta B       (eta = 2 for GAUGE 0, :797-799; the GAUGE 2
                             damping is proportional to |d chi| = 0)

(indices raised with the conformal metric g).  One step of
bssn_class::Step (bssn_class.C:3042-3163, 3401-3526) is then textbook
classical RK4 on these ODEs with the reference's projections:

* enforce_ga (unit determinant of g, trace-free A) on U0 in place before the
  predictor (:3051-3057) -- the enforced U0 is the base of every stage --
  and on every stage input (:3415-3428), but not on the result;
* lowerboundset on every stage output and on the result (:3156, 3525):
  chi - 1 is raised to chitiny - 1 there, while the predictor's RHS sees the
  unbounded U0;
* the stage inputs are the previous stage's outputs (the pre/cor swaps,
  :3765-3781), the final combination uses k1 + 2 k2 + 2 k3 + k4.

The Hamiltonian residual formed in the predictor from the enforced U0
(co = 0, bssn_rhs.f90:1044-1075 with R = 0) is H = 2/3 K^2 - A_kl A^kl at
every interior point, so its L2 norm over the owned points is |H| sqrt(N dV).

Ghost zones are frozen (the benchmark configurations' copy boundary); their
influence spreads 3 points per RHS evaluation, so only points at least 12
points from the ghosts are compared.
"""
import numpy as np
import pytest

NF = 24
N_INT = 32          # interior points per axis
GW = 3
H = 0.25
MARGIN = GW + 12    # compared points: [MARGIN, n - MARGIN) on every axis
CHITINY = 1e-5
EPS = 0.02


def geometry():
    n = N_INT + 2 * GW
    lo = -(n / 2) * H
    X = lo + (np.arange(n) + 0.5) * H
    bbox = (lo, lo, lo, lo + n * H, lo + n * H, lo + n * H)
    return n, X, bbox


# stored-offset values: chi_s = chi - 1, K, metric offsets (dxx, gxy, gxz, dyy,
# gyz, dzz), A (Axx, Axy, Axz, Ayy, Ayz, Azz), Lap_s = alpha - 1, beta, B; dt
CASES = {
    # chi below the floor with no evolution: only the final bound changes it;
    # det(g) = 1.1 so enforce_ga changes the metric; the shift subsystem is
    # linear: (beta, B)' = M (beta, B), M = [[0, 3/4], [0, -eta]]
    "floor_enforce_shift": dict(chi=-0.999995, K=0.0, g=(0.1, 0.0, 0.0, 0.0, 0.0, 0.0), A=(0.0,) * 6, lap=0.0,
                                beta=(0.0, 0.0, 0.0), B=(0.3, -0.2, 0.5), dt=0.0625),
    # nonlinear lapse / trK / chi / A dynamics: non-diagonal metric with
    # det != 1, a trace-free A_ij (with respect to g), uniform shift
    "lapse_trK_chi": dict(chi=-0.5, K=0.3, g=(0.1, 0.05, -0.02, -0.04, 0.03, 0.07),
                          A=(0.04, -0.03, 0.02, 0.05, 0.01, None), lap=-0.2,
                          beta=(0.1, -0.05, 0.02), B=(0.3, -0.2, 0.5), dt=0.0625),
    # chi starts below the floor and grows above it within the step: the
    # predictor's output is bounded (so k2 sees chi = chitiny) but the result
    # is not -- pins the bound on the stage outputs, not only on the result
    "floor_stage_output": dict(chi=6e-6 - 1.0, K=1.0, g=(0.0,) * 6, A=(0.0,) * 6, lap=2.0,
                               beta=(0.0, 0.0, 0.0), B=(0.1, 0.2, -0.3), dt=0.5),
}


def _sym(v, offset):
    xx, xy, xz, yy, yz, zz = v
    return np.array([[xx + offset, xy, xz], [xy, yy + offset, yz], [xz, yz, zz + offset]])


def _unsym(m, offset):
    return [m[0, 0] - offset, m[0, 1], m[0, 2], m[1, 1] - offset, m[1, 2], m[2, 2] - offset]


def initial_values(case):
    """The 24 stored values of a case (A_zz solved for a trace-free A_ij
    where the case leaves it open)."""
    c = CASES[case]
    A = list(c["A"])
    if A[5] is None:
        gi = np.linalg.inv(_sym(c["g"], 1.0))
        A[5] = 0.0
        A[5] = -np.trace(gi @ _sym(A, 0.0)) / gi[2, 2]
    vals = np.zeros(NF)
    vals[0], vals[1] = c["chi"], c["K"]
    vals[2:8] = c["g"]
    vals[8:14] = A
    vals[17] = c["lap"]
    vals[18:21] = c["beta"]
    vals[21:24] = c["B"]
    return vals


def uniform_state(case, n):
    return [np.full((n, n, n), v) for v in initial_values(case)]


def enforce(y):
    """Unit determinant of the conformal metric, trace-free A_ij."""
    y = y.copy()
    g = _sym(y[2:8], 1.0)
    g = g * np.linalg.det(g) ** (-1.0 / 3.0)
    A = _sym(y[8:14], 0.0)
    A = A - g * np.trace(np.linalg.inv(g) @ A) / 3.0
    y[2:8] = _unsym(g, 1.0)
    y[8:14] = _unsym(A, 0.0)
    return y


def rhs(y, eta):
    chi, K, a = y[0] + 1.0, y[1], y[17] + 1.0
    g = _sym(y[2:8], 1.0)
    A = _sym(y[8:14], 0.0)
    gi = np.linalg.inv(g)
    Aup = gi @ A @ gi
    AA = float(np.sum(A * Aup))
    d = np.zeros(NF)
    d[0] = 2.0 / 3.0 * chi * a * K
    d[1] = a * (K * K / 3.0 + AA)
    d[2:8] = _unsym(-2.0 * a * A, 0.0)
    d[8:14] = _unsym(a / 3.0 * (2.0 / 3.0 * K * K - AA) * g + a * (K * A - 2.0 * A @ gi @ A), 0.0)
    d[17] = -2.0 * a * K
    d[18:21] = 0.75 * y[21:24]
    d[21:24] = -eta * y[21:24]
    return d


def expected_step(case, gauge, stage_bound=True):
    """Textbook RK4 on the uniform-state ODEs with the reference's
    projections (stored offsets in and out); also the predictor's H."""
    dt = CASES[case]["dt"]
    eta = 2.0 if gauge == 0 else 0.0
    tinny = CHITINY - 1.0

    def bound(y, on=True):
        y = y.copy()
        if on and y[0] < tinny:
            y[0] = tinny
        return y

    y0 = enforce(initial_values(case))
    k1 = rhs(y0, eta)
    k2 = rhs(enforce(bound(y0 + 0.5 * dt * k1, stage_bound)), eta)
    k3 = rhs(enforce(bound(y0 + 0.5 * dt * k2, stage_bound)), eta)
    k4 = rhs(enforce(bound(y0 + dt * k3, stage_bound)), eta)
    y1 = bound(y0 + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4))
    g = _sym(y0[2:8], 1.0)
    A = _sym(y0[8:14], 0.0)
    gi = np.linalg.inv(g)
    ham = 2.0 / 3.0 * y0[1] ** 2 - float(np.sum(A * (gi @ A @ gi)))
    return y1, ham


def check_centre(U, exp, what):
    n = U[0].shape[0]
    s = np.s_[MARGIN:n - MARGIN, MARGIN:n - MARGIN, MARGIN:n - MARGIN]
    for v in range(NF):
        got = U[v][s]
        tol = 1e-13 * max(1.0, abs(exp[v]))
        err = float(np.max(np.abs(got - exp[v])))
        assert err <= tol, f"{what}: field {v}: |got - closed form| = {err:.3e} (expected {exp[v]!r})"


def test_cases_are_discriminating():
    """Each case moves what it is meant to pin, by far more than the tolerance."""
    for gauge in (0, 2):
        e, _ = expected_step("floor_enforce_shift", gauge)
        assert e[0] == CHITINY - 1.0 and abs(e[2] - 0.1) > 1e-3
    e0, _ = expected_step("floor_enforce_shift", 0)
    e2, _ = expected_step("floor_enforce_shift", 2)
    assert np.max(np.abs(e0[18:24] - e2[18:24])) > 1e-3       # eta enters
    a, _ = expected_step("floor_stage_output", 0)
    b, _ = expected_step("floor_stage_output", 0, stage_bound=False)
    assert a[0] > CHITINY - 1.0 and abs(a[0] - b[0]) > 1e-7      # stage bound changes the result
    # the nonlinear case moves K, the metric and A by far more than rounding
    e, _ = expected_step("lapse_trK_chi", 0)
    y0 = enforce(initial_values("lapse_trK_chi"))
    assert abs(e[1] - y0[1]) > 1e-3 and np.max(np.abs(e[2:14] - y0[2:14])) > 1e-4


def test_linear_shift_subsystem_is_the_rk4_polynomial():
    """For the linear shift subsystem textbook RK4 equals P(dt M) with
    P(z) = 1 + z + z^2/2 + z^3/6 + z^4/24 (the closed form of one step)."""
    c = CASES["floor_enforce_shift"]
    for gauge, eta in ((0, 2.0), (2, 0.0)):
        M = np.array([[0.0, 0.75], [0.0, -eta]])
        z = c["dt"] * M
        P = np.eye(2) + z + z @ z / 2 + z @ z @ z / 6 + z @ z @ z @ z / 24
        e, _ = expected_step("floor_enforce_shift", gauge)
        for d in range(3):
            want = P @ np.array([c["beta"][d], c["B"][d]])
            assert np.allclose([e[18 + d], e[21 + d]], want, rtol=1e-15, atol=1e-16)


def _ham_l2_expected(ham, n):
    return abs(ham) * np.sqrt(float(n - 2 * GW) ** 3 * H ** 3)


@pytest.mark.parametrize("gauge", [0, 2])
@pytest.mark.parametrize("case", list(CASES))
def test_oracle_step_closed_form(orc, case, gauge):
    n, X, bbox = geometry()
    U = uniform_state(case, n)
    gont, hamf = orc.step(U, X, X, X, bbox, CASES[case]["dt"], EPS, CHITINY, gauge=gauge, bc=0, gw=GW,
                          want_ham=True)
    assert gont == 0
    exp, ham = expected_step(case, gauge)
    check_centre(U, exp, f"orc.step {case} gauge {gauge}")
    s = np.s_[GW:n - GW, GW:n - GW, GW:n - GW]
    assert np.allclose(hamf[s], ham, rtol=1e-13, atol=1e-20)


@pytest.mark.parametrize("case", list(CASES))
def test_step_driver_over_reference_closed_form(ref, case):
    """The Python statement of Step over the reference's own routines
    (GAUGE 2, the reference build's gauge)."""
    import threading
    import oracle
    n, X, bbox = geometry()
    U = uniform_state(case, n)
    out = {}

    def run():  # the reference keeps its temporaries on the stack
        out["r"] = oracle.step_driver(ref, U, X, X, X, bbox, CASES[case]["dt"], EPS, CHITINY, 0, 0, GW, True)

    old = threading.stack_size(256 * 1024 * 1024)
    try:
        t = threading.Thread(target=run)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    gont, hamf = out["r"]
    assert gont == 0
    exp, ham = expected_step(case, 2)
    check_centre(U, exp, f"step_driver(ref) {case}")
    s = np.s_[GW:n - GW, GW:n - GW, GW:n - GW]
    assert np.allclose(hamf[s], ham, rtol=1e-13, atol=1e-20)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("gauge", [0, 2])
@pytest.mark.parametrize("case", list(CASES))
def test_gpu_level_step_closed_form(case, gauge, mode):
    import amssbssn as A
    from amssbssn.grid import PatchGeom
    n, X, bbox = geometry()
    lv = A.Level([PatchGeom((n, n, n), bbox, GW)], gauge=gauge, eps=EPS, chitiny=CHITINY)
    try:
        lv.set_state(0, uniform_state(case, n))
        assert lv.step(CASES[case]["dt"], 1, mode, want_ham=A.RES_LAST_STEP) == 0
        U = lv.get_state(0)
        exp, ham = expected_step(case, gauge)
        check_centre(U, exp, f"GPU level {case} gauge {gauge} mode {mode}")
        got = lv.ham_l2()
        want = _ham_l2_expected(ham, n)
        # the kernels' fused multiply-adds leave the second-derivative
        # stencils of a constant field at rounding size rather than exactly
        # zero (the Ham residual of the vacuum case reads 3.6e-15 per point,
        # 8e-14 as the norm); a wrong term would be of the order of K^2
        assert abs(got - want) <= 1e-13 * want + 1e-12, (got, want)
    finally:
        lv.close()
