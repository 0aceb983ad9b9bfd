"""Pins the oracle's search functions: BFGS with the Armijo line search
(bfgs.h:30-147), mutation (mutate.cpp:35-81, mutate_conf_cuda
common.cuh:207-258), the Metropolis rule (monte_carlo.cu:92-97) and the Monte
Carlo chain (kernel<32, Config>, monte_carlo.cu:236-356).

BFGS runs on quadratics, where the iterates follow from the algorithm by hand
(1-D cases) or obey identities that any correct inverse-Hessian update must
meet (the secant equation, symmetry, the initial scaling on vectors
orthogonal to s and y).  Mutation and acceptance are checked against the
distributions the reference draws from.
"""
import math

import numpy as np
import pytest
from scipy import stats

from helpers import oracle_scene, random_poses
from oracle.oracle import bfgs_quadratic, metropolis


# ---------------------------------------------------------------- BFGS on quadratics
def test_bfgs_1d_unit_curvature_hand_iterates():
    """f = x^2 / 2 from x = 1: the first step (H = I, alpha = 1) lands on the
    minimum; y.p / y.y = 1 keeps H = 1; the second direction is 0, so its line
    search fails all 10 trials (alpha = 2^-10); the convergence test then sees
    g = 0 and stops.  12 evaluations."""
    r = bfgs_quadratic(np.array([[1.0]]), np.array([0.0]), np.array([1.0]))
    assert r["x"][0] == 0.0 and r["f"] == 0.0
    assert r["evals"] == 12 and r["steps"] == 2
    np.testing.assert_array_equal(r["alpha"], [1.0, 2.0 ** -10])
    np.testing.assert_array_equal(r["trials"], [1, 10])
    np.testing.assert_array_equal(r["H"][:, 0, 0], [1.0, 1.0])


def test_bfgs_1d_backtracking_hand_iterates():
    """f = 2 x^2 from x = 1: alpha = 1 gives x = -3 (f = 18 > 2), alpha = 1/2
    gives x = -1 (f = 2, not below the Armijo line), alpha = 1/4 gives x = 0.
    The initial scaling alpha y.p / y.y = 1/4 is the exact inverse curvature,
    which the rank-2 update keeps."""
    r = bfgs_quadratic(np.array([[4.0]]), np.array([0.0]), np.array([1.0]))
    assert r["x"][0] == 0.0
    np.testing.assert_array_equal(r["trials"], [3, 10])
    assert r["alpha"][0] == 0.25
    np.testing.assert_allclose(r["H"][:, 0, 0], [0.25, 0.25], rtol=0, atol=1e-15)
    assert r["evals"] == 1 + 3 + 10


@pytest.mark.parametrize("prec", [64, 32])
def test_bfgs_armijo_constant(prec):
    """For f = a x^2 / 2 the full step satisfies f1 - f0 < c0 alpha p.g exactly
    when a < 2 - 2 c0: with c0 = 1e-4 the boundary is a = 1.9998."""
    ok = bfgs_quadratic(np.array([[1.99975]]), np.array([0.0]), np.array([1.0]), max_steps=1, prec=prec)
    no = bfgs_quadratic(np.array([[1.99990]]), np.array([0.0]), np.array([1.0]), max_steps=1, prec=prec)
    assert ok["trials"][0] == 1 and ok["alpha"][0] == 1.0
    assert no["trials"][0] == 2 and no["alpha"][0] == 0.5


def test_bfgs_failed_line_search_takes_the_last_trial():
    """A direction along which f cannot drop below the Armijo line (here the
    first line search against f0 = 0 of the reference GPU rule, with f > 0
    everywhere): all 10 trials fail, alpha ends at 2^-10 and x still moves by
    the last trial's 2^-9 p (x = x_new after the search, bfgs.h:125-129)."""
    x0 = 1.0
    r = bfgs_quadratic(np.array([[1.0]]), np.array([0.0]), np.array([x0]), c=100.0, max_steps=1, f0_zero=True)
    assert r["trials"][0] == 10 and r["alpha"][0] == 2.0 ** -10
    assert r["x"][0] == pytest.approx(x0 - 2.0 ** -9 * x0, rel=0, abs=1e-15)
    # the CPU rule compares with f(x0) and takes the full step
    c = bfgs_quadratic(np.array([[1.0]]), np.array([0.0]), np.array([x0]), c=100.0, max_steps=1)
    assert c["trials"][0] == 1 and c["x"][0] == 0.0
    # the GPU rule only changes the first line search
    g = bfgs_quadratic(np.array([[1.0]]), np.array([0.0]), np.array([x0]), c=100.0, max_steps=30, f0_zero=True)
    assert g["trials"][1] == 1 and g["x"][0] == pytest.approx(0.0, abs=1e-12)


def _spd(n, seed, cond=20.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    d = np.geomspace(1.0, cond, n)
    return (Q * d) @ Q.T, rng.normal(size=n), rng.normal(size=n) * 3


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bfgs_secant_symmetry_and_scaling(seed):
    """Every rank-2 update (bfgs.h:46-63) must give H_new y = s (s = alpha p)
    and a symmetric H; positive definite for a convex quadratic (y.s > 0).  On
    a vector v orthogonal to s_0 and y_0 the first update leaves H_0 = d I
    unchanged, so H_1 v = d v with d = y.s / y.y (bfgs.h:135-138)."""
    n = 7
    A, b, x0 = _spd(n, seed)
    r = bfgs_quadratic(A, b, x0, max_steps=40)
    assert r["steps"] >= 3
    for k in range(r["steps"]):
        s, y, H = r["s"][k], r["y"][k], r["H"][k]
        np.testing.assert_allclose(y, A @ s, rtol=1e-9, atol=1e-12)  # quadratic: y = A s
        if k == r["steps"] - 1 or y @ s < 1.2e-7:
            continue  # no update: the stopping step, or alpha y.p < epsilon_fl (bfgs.h:48)
        np.testing.assert_allclose(H @ y, s, rtol=1e-8, atol=1e-10 * np.abs(s).max())
        np.testing.assert_allclose(H, H.T, rtol=0, atol=1e-12)
        assert np.linalg.eigvalsh(H).min() > 0
    s0, y0 = r["s"][0], r["y"][0]
    d = (y0 @ s0) / (y0 @ y0)
    B = np.linalg.svd(np.stack([s0, y0]))[2][2:]  # orthonormal basis of {s0, y0}^perp
    for v in B:
        np.testing.assert_allclose(r["H"][0] @ v, d * v, rtol=0, atol=1e-10)
    # the minimiser of the quadratic
    np.testing.assert_allclose(r["x"], np.linalg.solve(A, b), rtol=0, atol=1e-5)


def test_bfgs_update_sign_error_breaks_the_secant_identity():
    """The identity H_new y = s of the previous test fixes both terms of the
    update: flipping the sign of either one violates it for a generic (H, s, y)."""
    rng = np.random.default_rng(3)
    n = 5
    M = rng.normal(size=(n, n))
    H = M @ M.T + n * np.eye(n)
    p = rng.normal(size=n)
    y = rng.normal(size=n)
    y = y if y @ p > 0 else -y
    alpha = 0.5
    mhy = -H @ y
    yhy = -(y @ mhy)
    r = 1.0 / (alpha * (y @ p))
    for s1, s2 in ((1, 1), (-1, 1), (1, -1)):
        Hn = H + s1 * alpha * r * (np.outer(mhy, p) + np.outer(p, mhy)) + \
            s2 * alpha * alpha * (r * r * yhy + r) * np.outer(p, p)
        ok = np.allclose(Hn @ y, alpha * p, rtol=1e-10, atol=1e-12)
        assert ok == (s1 == 1 and s2 == 1)


# ---------------------------------------------------------------- mutation
def _mutations(o, li, pose, rgyr, n, seed=77):
    out = []
    for k in range(n):
        out.append(o.mutate(li, pose, rgyr, seed, k % 97, k // 97))
    return np.array(out)


def _scene():
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    li = max(range(len(ligs)), key=lambda i: ligs[i].n_tors)
    pose = random_poses(np.random.default_rng(8), ligs[li], cav, 2.0, 1)[0]
    return o, ligs, li, pose


def test_mutation_picks_one_entity_uniformly():
    """mutate_conf: one of the 2 + T entities (position, orientation, each
    torsion) is chosen uniformly and only that entity changes."""
    o, ligs, li, pose = _scene()
    T = ligs[li].n_tors
    M = _mutations(o, li, pose, 3.0, 4000)
    which = []
    for m in M:
        d = np.abs(m[:7 + T] - pose[:7 + T]) > 0
        changed = [bool(d[0:3].any()), bool(d[3:7].any())] + [bool(d[7 + k]) for k in range(T)]
        assert sum(changed) == 1
        which.append(changed.index(True))
    counts = np.bincount(which, minlength=2 + T)
    assert stats.chisquare(counts).pvalue > 1e-3, counts
    assert np.all(M[:, 7 + T:] == 0)


def test_mutation_position_step_uniform_in_ball():
    """Position entity: pose.pos += amplitude u, u uniform in the unit ball:
    |u| has CDF r^3 and the direction is uniform on the sphere (its z
    component uniform in [-1, 1], its azimuth uniform)."""
    o, ligs, li, pose = _scene()
    M = _mutations(o, li, pose, 3.0, 6000)
    d = M[:, :3] - pose[:3]
    mv = np.linalg.norm(d, axis=1) > 0
    u = d[mv] / 2.0  # amplitude 2 (mutate.cpp:42, vina.h:210)
    r = np.linalg.norm(u, axis=1)
    assert r.max() <= 1.0 + 1e-12
    assert stats.kstest(r, lambda t: np.clip(t, 0, 1) ** 3).pvalue > 1e-3
    z = u[:, 2] / r
    assert stats.kstest(z, "uniform", args=(-1, 2)).pvalue > 1e-3
    phi = np.arctan2(u[:, 1], u[:, 0])
    assert stats.kstest(phi, "uniform", args=(-math.pi, 2 * math.pi)).pvalue > 1e-3


def _qmul(p, q):
    a1, b1, c1, d1 = p
    a2, b2, c2, d2 = q
    return np.array([a1 * a2 - b1 * b2 - c1 * c2 - d1 * d2, a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2,
                     a1 * c2 - b1 * d2 + c1 * a2 + d1 * b2, a1 * d2 + b1 * c2 - c1 * b2 + d1 * a2])


def test_mutation_rotation_scaled_by_gyration_radius():
    """Orientation entity: q <- exp((amplitude / r_gyr) u / 2) q with u uniform
    in the unit ball (quaternion_increment), i.e. a rotation by angle
    (2 / r_gyr) |u| about u: the angle over 2 / r_gyr has CDF t^3.  A zero
    gyration radius leaves the orientation unchanged."""
    o, ligs, li, pose = _scene()
    rgyr = 4.0
    M = _mutations(o, li, pose, rgyr, 6000)
    q0 = pose[3:7] / np.linalg.norm(pose[3:7])
    conj = q0 * np.array([1, -1, -1, -1])
    ang = []
    for m in M:
        if np.any(m[3:7] != pose[3:7]):
            dq = _qmul(m[3:7], conj)
            ang.append(2 * math.atan2(np.linalg.norm(dq[1:]), abs(dq[0])))
    t = np.array(ang) / (2.0 / rgyr)
    assert len(t) > 300 and t.max() <= 1.0 + 1e-9
    assert stats.kstest(t, lambda s: np.clip(s, 0, 1) ** 3).pvalue > 1e-3
    Z = _mutations(o, li, pose, 0.0, 400)
    assert np.all(Z[:, 3:7] == pose[3:7])


def test_mutation_torsion_uniform():
    """Torsion entity: the torsion is replaced by a value uniform in [-pi, pi)."""
    o, ligs, li, pose = _scene()
    T = ligs[li].n_tors
    M = _mutations(o, li, pose, 3.0, 8000)
    vals = []
    for m in M:
        d = np.nonzero(m[7:7 + T] != pose[7:7 + T])[0]
        if len(d):
            vals.append(m[7 + d[0]])
    vals = np.array(vals)
    assert np.all(vals >= -math.pi) and np.all(vals < math.pi)
    assert stats.kstest(vals, "uniform", args=(-math.pi, 2 * math.pi)).pvalue > 1e-3


# ---------------------------------------------------------------- Metropolis
def test_metropolis_rule():
    """Accept a lower energy always; a higher one when u < exp((old - new) / T)."""
    T = 1.2
    assert metropolis(-5.0, -6.0, T, 0.999999)
    for de in (0.1, 0.5, 1.0, 3.0):
        p = math.exp(-de / T)
        assert metropolis(-5.0, -5.0 + de, T, p * 0.999)
        assert not metropolis(-5.0, -5.0 + de, T, p * 1.001)
    # acceptance frequency over uniform u equals exp(-de / T)
    u = np.random.default_rng(0).random(20000)
    for de in (0.3, 1.2):
        f = np.mean([metropolis(0.0, de, T, x) for x in u])
        assert f == pytest.approx(math.exp(-de / T), abs=0.015)


# ---------------------------------------------------------------- Monte Carlo chain
def test_mc_chain_best_is_scored_with_authentic_caps():
    """The chain's best pose comes out of the second BFGS with authentic_v =
    1000 (monte_carlo.cu:308-330): its energy re-evaluated with caps 1000
    equals the reported best energy, and the best-energy trace never rises."""
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    for li in range(3):
        e, best, ev, tr = o.mc_chain(li, 1234, 5 + li, 6)
        assert o.eval_deriv(li, best, (1000.0, 1000.0, 1000.0))[0] == pytest.approx(e, rel=1e-9, abs=1e-9)
        assert np.all(np.diff(tr) <= 1e-12)
        assert tr[-1] == e


def test_mc_chain_first_line_search_rule_changes_the_chain():
    """Both first-line-search rules are available to the chain: the GPU rule
    (f0 = 0) gives a different, still self-consistent, chain."""
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    a = o.mc_chain(2, 99, 1, 4)
    o.set_ls_f0_zero(True)
    b = o.mc_chain(2, 99, 1, 4)
    o.set_ls_f0_zero(False)
    c = o.mc_chain(2, 99, 1, 4)
    assert not np.array_equal(a[1], b[1])
    assert np.array_equal(a[1], c[1])
    assert o.eval_deriv(2, b[1], (1000.0, 1000.0, 1000.0))[0] == pytest.approx(b[0], rel=1e-9, abs=1e-9)
