"""Pins the CPU oracle against known answers, closed forms and finite differences.

The reference ships no numeric golden vectors for this path (its tests only
check score ranges), so each oracle function is pinned by what the mathematics
fixes: published Philox known-answer vectors, hand-computed potential values,
affine fields that trilinear interpolation must reproduce exactly, geometric
invariants of forward kinematics, and central finite differences of the
energy.  (tests/golden/ holds the reference's own 1iep / 1a30 input fixtures,
used by the ingestion and redock tests; the search functions are pinned in
test_oracle_search.py.)
"""
import math

import numpy as np
import pytest

from helpers import LIGAND_TYPES, oracle_scene, random_poses, small_ligands
from oracle.oracle import Oracle, philox

W = dict(g1=-0.035579, g2=-0.005156, rep=0.840245, hyd=-0.035069, hb=-0.587439)


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_vina_potential_hand_values():
    o = Oracle(64, 0)
    # C_H - C_H: optimal distance 1.9 + 1.9 = 3.8, both hydrophobic, no h-bond
    e = o.sf_eval(0, 0, 3.8)
    expect = W["g1"] * 1.0 + W["g2"] * math.exp(-(3.0 / 2.0) ** 2) + W["hyd"] * 1.0
    assert e == pytest.approx(expect, abs=1e-14)
    # d = -1: repulsion 1, gauss1 exp(-4), gauss2 exp(-4), hydrophobic 1
    e = o.sf_eval(0, 0, 2.8)
    expect = W["g1"] * math.exp(-4) + W["g2"] * math.exp(-4) + W["rep"] * 1.0 + W["hyd"]
    assert e == pytest.approx(expect, abs=1e-14)
    # hydrophobic slope_step(1.5, 0.5, d): half way at d = 1.0
    e = o.sf_eval(0, 0, 4.8)
    expect = W["g1"] * math.exp(-4) + W["g2"] * math.exp(-(2.0 / 2.0) ** 2) + W["hyd"] * 0.5
    assert e == pytest.approx(expect, abs=1e-14)
    # N_D (donor) - O_A (acceptor): od 1.8 + 1.7 = 3.5; h-bond slope_step(0, -0.7, -0.35) = 0.5
    e = o.sf_eval(3, 8, 3.15)
    expect = W["g1"] * math.exp(-(0.35 / 0.5) ** 2) + W["g2"] * math.exp(-(3.35 / 2.0) ** 2) + \
        W["rep"] * 0.35 ** 2 + W["hb"] * 0.5
    assert e == pytest.approx(expect, abs=1e-14)
    # C_P - N_P: no hydrophobic, no h-bond
    e = o.sf_eval(1, 2, 3.7)
    expect = W["g1"] * 1.0 + W["g2"] * math.exp(-(3.0 / 2.0) ** 2)
    assert e == pytest.approx(expect, abs=1e-14)
    # cutoff 8 A: all Vina terms vanish at and beyond it
    assert o.sf_eval(0, 0, 8.0) == 0.0
    assert o.sf_eval(3, 8, 9.5) == 0.0


def test_vinardo_potential_hand_values():
    o = Oracle(64, 1)
    # C_H - C_H vinardo radii 2.0 + 2.0; gauss(0, 0.8), rep, hydrophobic slope_step(2.5, 0, d)
    e = o.sf_eval(0, 0, 4.0)
    assert e == pytest.approx(-0.045 * 1.0 + -0.035 * 1.0, abs=1e-14)
    e = o.sf_eval(0, 0, 5.25)
    assert e == pytest.approx(-0.045 * math.exp(-(1.25 / 0.8) ** 2) + -0.035 * 0.5, abs=1e-14)
    e = o.sf_eval(0, 0, 3.0)
    assert e == pytest.approx(-0.045 * math.exp(-(1.0 / 0.8) ** 2) + 0.8 * 1.0 + -0.035 * 1.0, abs=1e-14)


def test_pair_table_samples_and_derivative():
    o = Oracle(64, 0)
    e, dor, fast = o.pair_table(1, 2)  # C_P - N_P: only the two gaussians and repulsion
    rs = np.sqrt(np.arange(2051) / 32.0)
    # samples of the closed form
    for k in (0, 1, 100, 555, 1024, 2047):
        assert e[k] == pytest.approx(o.sf_eval(1, 2, rs[k]), abs=1e-15)
    assert e[2048] == 0.0 and e[2050] == 0.0
    # dor approximates (1/r) dE/dr of the closed form (smooth region)
    od = 3.7

    def dE(r):
        a = (r - od) / 0.5
        b = (r - od - 3.0) / 2.0
        d = r - od
        return W["g1"] * math.exp(-a * a) * (-2 * a / 0.5) + W["g2"] * math.exp(-b * b) * (-2 * b / 2.0) + \
            (W["rep"] * 2 * d if d < 0 else 0.0)
    for k in (300, 500, 700, 900, 1500):
        r = rs[k]
        assert dor[k] == pytest.approx(dE(r) / r, rel=2e-3, abs=1e-6)
    assert dor[0] == 0.0 and dor[2050] == 0.0
    # fast is the midpoint of neighbouring samples
    assert fast[10] == pytest.approx(0.5 * (e[10] + e[11]), abs=1e-15)


def _affine_scene(prec):
    o = Oracle(prec, 0)
    n, geo = o.set_grid(np.array([1.0, -2.0, 3.0]), np.array([6.0, 7.5, 4.5]), 0.375)
    init, factor_inv = geo[:3], geo[6:9]
    xs = init[0] + factor_inv[0] * np.arange(n[0])
    ys = init[1] + factor_inv[1] * np.arange(n[1])
    zs = init[2] + factor_inv[2] * np.arange(n[2])
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    a, b, c, d = -5.0, 0.25, -0.5, 0.125
    vals = a + b * X + c * Y + d * Z  # negative everywhere: no curl
    assert vals.max() < 0
    o.set_map(0, vals.reshape(-1))
    return o, (a, b, c, d), init, init + factor_inv * (np.array(n) - 1), geo[3:6]


def test_grid_trilinear_reproduces_affine_fields():
    o, (a, b, c, d), lo, hi, factor = _affine_scene(64)
    rng = np.random.default_rng(0)
    for _ in range(50):
        p = lo + rng.random(3) * (hi - lo)
        e, g = o.grid_eval(0, p, v=1000.0)
        assert e == pytest.approx(a + b * p[0] + c * p[1] + d * p[2], abs=1e-12)
        np.testing.assert_allclose(g, [b, c, d], atol=1e-12)


def test_grid_out_of_box_penalty():
    o, (a, b, c, d), lo, hi, factor = _affine_scene(64)
    slope = 1e6
    p = np.array([lo[0] - 0.5, 0.5 * (lo[1] + hi[1]), hi[2] + 0.25])
    e, g = o.grid_eval(0, p, v=1000.0)
    inside = np.array([lo[0], p[1], hi[2]])
    f_edge = a + b * inside[0] + c * inside[1] + d * inside[2]
    assert e == pytest.approx(f_edge + slope * (0.5 + 0.25), rel=1e-12)
    assert g[0] == pytest.approx(-slope)
    assert g[1] == pytest.approx(c, abs=1e-12)
    assert g[2] == pytest.approx(slope)


def test_grid_curl_caps_positive_energies():
    o = Oracle(64, 0)
    n, geo = o.set_grid(np.zeros(3), np.array([3.0, 3.0, 3.0]), 0.375)
    o.set_map(0, np.full(int(np.prod(n)), 30.0))
    e, _ = o.grid_eval(0, np.zeros(3), v=10.0)
    assert e == pytest.approx(30.0 * 10.0 / 40.0)


def test_grid_build_matches_direct_sum():
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    rng = np.random.default_rng(1)
    n = o.n
    init, factor_inv = o.geo[:3], o.geo[6:9]
    m = o.get_map(0).reshape(n[2], n[1], n[0])
    tab = {t: o.pair_table(t, 0)[2] for t in set(int(t) for t in xs)}
    for _ in range(20):
        ix, iy, iz = (int(rng.integers(0, n[i])) for i in range(3))
        p = init + factor_inv * np.array([ix, iy, iz])
        r2 = ((xyz - p) ** 2).sum(1)
        s = sum(tab[int(xs[j])][int(32 * r2[j])] for j in np.nonzero(r2 <= 64.0)[0])
        assert m[iz, iy, ix] == pytest.approx(s, rel=1e-12, abs=1e-12)


def test_forward_kinematics_invariants():
    o, ligs, _ = oracle_scene(64, with_maps=False)
    rng = np.random.default_rng(2)
    for li, lig in enumerate(ligs):
        zero = np.zeros(7 + 48)  # UD_POSE_FLOATS
        zero[0:3] = lig.origin[0]
        zero[3] = 1.0
        _, _, coords, _ = o.eval_deriv(li, zero)
        np.testing.assert_allclose(coords, lig.coords, atol=1e-12)
        pose = random_poses(rng, lig, np.zeros(3), 5.0, 1)[0]
        _, _, c2, _ = o.eval_deriv(li, pose)
        for a, b in lig.bonds:
            assert np.linalg.norm(c2[a] - c2[b]) == pytest.approx(1.5, abs=1e-9)
        for k in range(len(lig.parent)):  # distances inside a rigid node are preserved
            sl = slice(lig.abegin[k], lig.aend[k])
            d0 = np.linalg.norm(lig.coords[sl][:, None] - lig.coords[sl][None], axis=-1)
            d1 = np.linalg.norm(c2[sl][:, None] - c2[sl][None], axis=-1)
            np.testing.assert_allclose(d1, d0, atol=1e-9)


def _rotq(axis, ang):
    axis = np.asarray(axis, float)
    return np.array([math.cos(ang / 2), *(math.sin(ang / 2) * axis)])


def _qmul(p, q):
    a1, b1, c1, d1 = p
    a2, b2, c2, d2 = q
    return np.array([a1 * a2 - b1 * b2 - c1 * c2 - d1 * d2, a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2,
                     a1 * c2 - b1 * d2 + c1 * a2 + d1 * b2, a1 * d2 + b1 * c2 - c1 * b2 + d1 * a2])


def _fd_grad(o, li, lig, pose, h):
    fd = np.zeros(6 + lig.n_tors)
    for i in range(3):
        p1, p2 = pose.copy(), pose.copy()
        p1[i] += h
        p2[i] -= h
        fd[i] = (o.eval_deriv(li, p1)[0] - o.eval_deriv(li, p2)[0]) / (2 * h)
    for i in range(3):
        ax = np.eye(3)[i]
        p1, p2 = pose.copy(), pose.copy()
        p1[3:7] = _qmul(_rotq(ax, h), pose[3:7])
        p2[3:7] = _qmul(_rotq(ax, -h), pose[3:7])
        fd[3 + i] = (o.eval_deriv(li, p1)[0] - o.eval_deriv(li, p2)[0]) / (2 * h)
    for k in range(lig.n_tors):
        p1, p2 = pose.copy(), pose.copy()
        p1[7 + k] += h
        p2[7 + k] -= h
        fd[6 + k] = (o.eval_deriv(li, p1)[0] - o.eval_deriv(li, p2)[0]) / (2 * h)
    return fd


def test_eval_deriv_matches_central_differences():
    """Grid term + tree derivative: exact to FD accuracy (pair-free copies of
    the ligands).  With intramolecular pairs the reference interpolates e and
    dor = (1/r) dE/dr from separate samples (precalculate.h:48-65), so the
    gradient is the derivative of the closed form rather than of the
    piecewise-linear energy; the bound there is the table's consistency."""
    import copy
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    bare = []
    for lig in ligs:
        b = copy.deepcopy(lig)
        b.pairs = np.zeros((0, 2), dtype=np.int32)
        bare.append(o.add_ligand(b))
    rng = np.random.default_rng(3)
    for li, lig in enumerate(ligs):
        for pose in random_poses(rng, lig, cav, 2.0, 3):
            _, pose, _ = o.bfgs(li, pose, max_steps=3)  # off the steepest walls
            g = o.eval_deriv(li, pose)[1]
            fd = _fd_grad(o, li, lig, pose, 1e-5)
            scale = max(1.0, np.abs(g).max())
            np.testing.assert_allclose(g, fd, atol=5e-3 * scale, rtol=0)
            gb = o.eval_deriv(bare[li], pose)[1]
            fdb = _fd_grad(o, bare[li], lig, pose, 1e-5)
            np.testing.assert_allclose(gb, fdb, atol=1e-6 * max(1.0, np.abs(gb).max()), rtol=1e-6)


def test_fp32_restatement_close_to_fp64():
    o64, ligs, (xyz, xs, cav) = oracle_scene(64)
    o32, _, _ = oracle_scene(32)
    rng = np.random.default_rng(4)
    for li, lig in enumerate(ligs):
        for pose in random_poses(rng, lig, cav, 2.0, 4):
            e64, g64, _, _ = o64.eval_deriv(li, pose)
            e32, g32, _, _ = o32.eval_deriv(li, pose)
            assert e32 == pytest.approx(e64, rel=1e-4, abs=1e-4)
            np.testing.assert_allclose(g32, g64, rtol=1e-3, atol=1e-3 * max(1.0, np.abs(g64).max()))


def test_bfgs_decreases_energy_and_reverts_on_failure():
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    rng = np.random.default_rng(5)
    for li, lig in enumerate(ligs[:3]):
        for pose in random_poses(rng, lig, cav, 2.0, 3):
            e0 = o.eval_deriv(li, pose)[0]
            e1, p1, ev = o.bfgs(li, pose)
            assert e1 <= e0 + 1e-12
            assert ev >= 2
            # the returned energy is the energy of the returned pose
            assert o.eval_deriv(li, p1)[0] == pytest.approx(e1, rel=1e-9, abs=1e-9)


def test_mc_chain_is_seeded():
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    a = o.mc_chain(0, 42, 3, 4)
    b = o.mc_chain(0, 42, 3, 4)
    c = o.mc_chain(0, 43, 3, 4)
    assert a[0] == b[0] and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[1], c[1])
    # best energy is monotone along the chain
    assert np.all(np.diff(a[3]) <= 1e-12)


def test_random_start_inside_box():
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    lo = o.geo[:3]
    hi = lo + o.geo[6:9] * (np.array(o.n) - 1)
    for ch in range(20):
        p = o.random_start(1, 7, ch)
        assert np.all(p[:3] >= lo) and np.all(p[:3] <= hi)
        assert np.linalg.norm(p[3:7]) == pytest.approx(1.0, abs=1e-12)
        assert np.all(np.abs(p[7:7 + ligs[1].n_tors]) <= math.pi)


def test_final_score_closed_form():
    o = Oracle(64, 0)
    # num_tors_div: x / (1 + 0.1 * (w_int + 1) * N / 5) with w_int = 5 w_rot / 0.1 - 1
    w_int = 5 * 0.05846 / 0.1 - 1
    for n in (0, 1, 5, 10):
        assert o.conf_independent(n, -7.5) == pytest.approx(-7.5 / (1 + 0.1 * (w_int + 1) * n / 5))


# ---------------------------------------------------------------- explicit receptor (refinement)
def test_explicit_receptor_equals_maps_at_grid_points():
    """non_cache::eval (fast table, energy only) at a grid point inside the box
    is the map value there: both sum eval_fast over the receptor atoms within
    the cutoff, then curl."""
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    rng = np.random.default_rng(4)
    n = o.n
    init, factor_inv = o.geo[:3], o.geo[6:9]
    for _ in range(40):
        idx = np.array([int(rng.integers(0, n[i])) for i in range(3)])
        p = init + factor_inv * idx
        t = int(rng.choice(LIGAND_TYPES))
        for v in (1000.0, 10.0):
            e_nc, _ = o.nc_atom(t, p, v=v, deriv=False)
            e_map, _ = o.grid_eval(t, p, v=v, deriv=False)
            assert e_nc == pytest.approx(e_map, rel=1e-9, abs=1e-9), (idx, t, v)


def test_explicit_receptor_out_of_box():
    """Outside the box the coordinate is clamped: penalty slope * |miss|, and
    the gradient is the clamped point's pair gradient plus slope * sign (the
    reference does not zero the pair part in clamped directions)."""
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    lo = o.geo[:3]
    hi = lo + o.geo[6:9] * (np.array(o.n) - 1)
    face = np.array([lo[0], 0.5 * (lo[1] + hi[1]), hi[2]])
    out = face + np.array([-0.7, 0.0, 0.4])
    for slope in (100.0, 1e6):
        e_f, d_f = o.nc_atom(0, face, slope=slope)
        e_o, d_o = o.nc_atom(0, out, slope=slope)
        assert e_o - e_f == pytest.approx(slope * 1.1, rel=1e-9)
        np.testing.assert_allclose(d_o, d_f + slope * np.array([-1.0, 0.0, 1.0]), rtol=1e-12, atol=1e-9)
        e_fast_f, _ = o.nc_atom(0, face, slope=slope, deriv=False)
        e_fast_o, _ = o.nc_atom(0, out, slope=slope, deriv=False)
        assert e_fast_o - e_fast_f == pytest.approx(slope * 1.1, rel=1e-9)


def test_explicit_receptor_gradient_central_differences():
    """Inside the box the gradient is that of the interpolated pair energy, to
    the table's own consistency: e and dor are interpolated from separate
    samples (dor a central difference of e), so a sum of ~100 receptor pairs
    differs from the finite difference of the piecewise-linear e by up to ~1 %
    of the gradient; a wrong sign or factor would be off by 100 %.  Points
    with a receptor atom in the table's last bin (2047/32 <= r^2 < 64) are
    skipped: the terms vanish at r = 8 A, so e falls to 0 within that bin while
    dor, a central difference across it, carries about half that slope.  The
    same happens, more weakly, in the bin of each kink of the piecewise-linear
    terms (hydrophobic and hydrogen-bond slope steps), so most points, not
    all, must agree to 1 %."""
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    rng = np.random.default_rng(5)
    h = 1e-5
    n_checked = n_close = 0
    for _ in range(60):
        p = cav + rng.uniform(-5.0, 5.0, 3)
        t = int(rng.choice(LIGAND_TYPES))
        e, d = o.nc_atom(t, p)
        r2 = ((xyz - p) ** 2).sum(1)
        if e > 5.0 or np.any((r2 >= 2047 / 32) & (r2 < 64.0)):  # off the steep walls and the last bin
            continue
        fd = np.zeros(3)
        for k in range(3):
            a, b = p.copy(), p.copy()
            a[k] += h
            b[k] -= h
            fd[k] = (o.nc_atom(t, a)[0] - o.nc_atom(t, b)[0]) / (2 * h)
        sc = max(1.0, np.abs(d).max())
        n_close += bool(np.all(np.abs(d - fd) <= 1e-2 * sc))
        assert np.all(np.abs(d - fd) <= 0.5 * sc), (p, t, d, fd)
        n_checked += 1
    assert n_checked >= 20
    assert n_close >= 0.8 * n_checked, (n_close, n_checked)


def test_refine_descends_and_reports_within():
    o, ligs, (xyz, xs, cav) = oracle_scene(64)
    lo = o.geo[:3]
    hi = lo + o.geo[6:9] * (np.array(o.n) - 1)
    rng = np.random.default_rng(6)
    n_in = 0
    for li, lig in enumerate(ligs):
        for pose in random_poses(rng, lig, cav, 1.5, 2):
            _, start, _ = o.bfgs(li, pose, max_steps=30)  # a docked-like start: a grid-map minimum
            e0 = o.eval_deriv_nc(li, start, slope=100.0)[0]
            e, P, within = o.refine(li, start, 3)
            c = o.eval_deriv(li, P)[2]
            heavy_in = bool(np.all((c[lig.heavy == 1] >= lo - 1e-4) & (c[lig.heavy == 1] <= hi + 1e-4)))
            assert within == heavy_in
            if within:
                n_in += 1
                # the energy is that of the returned pose for one of the pass slopes
                es = [o.eval_deriv_nc(li, P, slope=s)[0] for s in (100.0, 1e4, 1e6)]
                assert min(abs(e - x) for x in es) <= 1e-9 * max(1.0, abs(e))
                if abs(e - es[0]) <= 1e-9 * max(1.0, abs(e)):  # ended in the first pass
                    assert e <= e0 + 1e-9
            else:
                assert e == pytest.approx(3.4028234663852886e38)
    assert n_in >= 6
