"""AutoDock 4.2 scoring in the oracle (CPU): potentials, per-pair tables,
maps and the pose energy, pinned by closed forms, published parameters,
brute force and finite differences.

  potentials   12-6 vdW minimum -sqrt(eps_i eps_j) at R_i + R_j on the 0.5 A
               smoothing plateau; 12-10 H-bond minimum equal to the product
               of the H-bond depths at the H-bond radius sum; vdW is zero for
               H-bond pairs and H-bond zero otherwise; 8 A / 20.48 A cutoffs;
               the Mehler-Solmajer sigmoidal dielectric of the AutoDock 4
               papers (A = -8.5525, eps0 = 78.4, lambda = 0.003627,
               k = 7.7839) and the 332 q_i q_j / (r eps(r)) Coulomb term with its
               cap; the desolvation Gaussian (sigma 3.6 A) and its linearity
               in |q|
  parameters   the AutoDock 4.2 parameter file's rows for C, OA, HD, and
               volume = 4/3 pi R^3 for every heavy atom type
  tables       samples at rs[k] = sqrt(k / 32), 13424 of them (20.48 A),
               central-difference dor, midpoint fast table; the table of a pair
               is its uncharged table plus q_i q_j times the Coulomb samples plus
               the desolvation coefficient's charge part times the Gaussian
               samples (the decomposition the GPU kernels use)
  maps         map_t + q elec + |q| dsolv at grid points equals the brute-force
               sum of the pair potential over the receptor atoms
  energy       central differences of the FP64 energy (gradient 6 + T)
"""
import math

import numpy as np
import pytest

from helpers import oracle_scene_ad4, random_poses, small_receptor_ad4, small_ligands_ad4
from oracle.oracle import AD4_MAP_DSOLV, AD4_MAP_ELEC, AD4_TABLE_N, AD4_WEIGHTS, Oracle, ad4_kind

C, A, N, O, P, S, H, F, I, NA, OA, SA, HD = range(13)
W = np.array(AD4_WEIGHTS)


@pytest.fixture(scope="module")
def o():
    return Oracle(prec=64, sf=2)


def test_parameter_rows():
    # AD4.1_bound.dat rows (Rii / 2, epsii, hbond eps, Rij_hb, solpar, vol)
    assert np.allclose(ad4_kind(C), [2.0, 0.150, 0.0, 0.0, -0.00143, 33.5103])
    assert np.allclose(ad4_kind(OA), [1.6, 0.200, -5.0, 1.9, -0.00251, 17.1573])
    assert np.allclose(ad4_kind(HD), [1.0, 0.020, 1.0, 0.0, 0.00051, 0.0])
    for t in (C, A, N, O, P, S, F, I, NA, OA, SA, 18, 19, 20, 21):  # heavy non-metal types
        r, v = ad4_kind(t)[0], ad4_kind(t)[5]
        # the published Br volume (42.5661) is 0.14 % above the sphere's
        assert v == pytest.approx(4.0 / 3.0 * math.pi * r ** 3, rel=2e-3 if t == 19 else 2e-4), t


def _lj(rij, r, depth, m, n):
    return depth * (n * (rij / r) ** m - m * (rij / r) ** n) / (m - n)


def test_vdw_and_hbond(o):
    rij = 4.0  # C-C
    for r in (3.76, 4.0, 4.24):  # the smoothing plateau
        assert o.ad4_terms(C, 0, C, 0, r)[0] == pytest.approx(-0.15, abs=1e-12)
    for r in (2.5, 3.3, 5.0, 6.7):
        rs = r + 0.25 if r < rij else r - 0.25
        assert o.ad4_terms(C, 0, C, 0, r)[0] == pytest.approx(_lj(rij, rs, 0.15, 12, 6), rel=1e-10)
    assert o.ad4_terms(C, 0, C, 0, 8.0)[0] == 0.0
    assert o.ad4_terms(C, 0, C, 0, 0.3)[0] == 100000.0  # cap
    # OA...HD: 12-10 well of depth 5 * 1 at 1.9 A
    assert o.ad4_terms(OA, 0, HD, 0, 1.9)[1] == pytest.approx(-5.0, abs=1e-12)
    for r in (1.5, 2.6, 3.5):
        rs = r + 0.25 if r < 1.9 else r - 0.25
        assert o.ad4_terms(OA, 0, HD, 0, r)[1] == pytest.approx(_lj(1.9, rs, 5.0, 12, 10), rel=1e-10)
    assert o.ad4_terms(OA, 0, HD, 0, 2.0)[0] == 0.0  # an H-bond pair has no vdW term
    assert o.ad4_terms(C, 0, OA, 0, 3.0)[1] == 0.0   # no H-bond without a donor
    assert o.ad4_terms(OA, 0, OA, 0, 3.0)[1] == 0.0  # acceptor-acceptor: vdW
    assert o.ad4_terms(OA, 0, OA, 0, 3.0)[0] != 0.0


def _eps(r):
    a, e0, lam, k = -8.5525, 78.4, 0.003627, 7.7839
    b = e0 - a
    return a + b / (1.0 + k * math.exp(-lam * b * r))


def test_electrostatics(o):
    assert _eps(1e6) == pytest.approx(78.4)
    for r in (1.5, 3.0, 7.0, 15.0, 20.4):
        assert o.ad4_terms(C, 0.3, OA, -0.5, r)[2] == pytest.approx(332.0 * 0.3 * -0.5 / (r * _eps(r)), rel=1e-12)
    assert o.ad4_terms(C, 0.3, OA, -0.5, 0.005)[2] == pytest.approx(332.0 * 0.3 * -0.5 * 100.0, rel=1e-12)  # cap
    assert o.ad4_terms(C, 0.3, OA, -0.5, 20.48)[2] == 0.0


def test_desolvation(o):
    s = lambda q1, q2, r: o.ad4_terms(C, q1, OA, q2, r)[3]
    r1, r2 = 2.0, 5.0
    assert s(0.2, -0.4, r1) / s(0.2, -0.4, r2) == pytest.approx(math.exp(-(r1 ** 2 - r2 ** 2) / (2 * 3.6 ** 2)))
    # linear in |q1| and |q2|: the charge part is 0.01097 (|q1| V2 + |q2| V1) G
    g = math.exp(-0.5 * (r1 / 3.6) ** 2)
    d = s(0.2, -0.4, r1) - s(0.0, 0.0, r1)
    assert d == pytest.approx(0.01097 * (0.2 * 17.1573 + 0.4 * 33.5103) * g, rel=1e-10)
    assert s(-0.2, 0.4, r1) == pytest.approx(s(0.2, -0.4, r1))
    assert o.ad4_terms(C, 0.1, OA, -0.2, 3.3)[3] == pytest.approx(o.ad4_terms(OA, -0.2, C, 0.1, 3.3)[3])
    assert o.ad4_terms(HD, 0.2, HD, 0.3, 2.0)[3] == 0.0  # volume 0
    assert s(0.2, -0.4, 20.48) == 0.0


def test_weighted_sum(o):
    t = o.ad4_terms(N, -0.3, OA, -0.4, 3.1)
    assert t[4] == pytest.approx(float(np.dot(W[:4], t[:4])), rel=1e-12)


def test_pair_table_sampling(o):
    e, dor, fast = o.pair_table_ad4(NA, -0.35, HD, 0.2)
    assert len(e) == AD4_TABLE_N == int(32 * 20.48 ** 2) + 3
    rs = np.sqrt(np.arange(AD4_TABLE_N) / 32.0)
    for k in (1, 77, 150, 2047, 2048, 5000, 13421, 13422):
        assert e[k] == pytest.approx(o.ad4_terms(NA, -0.35, HD, 0.2, rs[k])[4], rel=1e-12, abs=1e-15)
    k = np.arange(1, AD4_TABLE_N - 1)
    np.testing.assert_allclose(dor[k], (e[k + 1] - e[k - 1]) / ((rs[k + 1] - rs[k - 1]) * rs[k]), rtol=1e-12)
    assert dor[0] == 0 and dor[-1] == 0
    np.testing.assert_allclose(fast[:-1], (e[:-1] + e[1:]) / 2)


def test_pair_table_charge_decomposition(o):
    """table(q1, q2) = table(0, 0) + q1 q2 w_elec C(r) + 0.01097 (|q1| V2 + |q2| V1) w_dsolv G(r)"""
    rs = np.sqrt(np.arange(AD4_TABLE_N) / 32.0)
    cr = np.array([o.ad4_terms(C, 1.0, C, 1.0, r)[2] for r in rs])  # 332 / (r eps), capped
    gr = np.where(rs < 20.48, np.exp(-0.5 * (rs / 3.6) ** 2), 0.0)  # the Gaussian of test_desolvation
    for (t1, q1, t2, q2) in ((OA, -0.41, HD, 0.21), (A, 0.02, NA, -0.33), (SA, -0.15, C, 0.07)):
        e, dor, fast = o.pair_table_ad4(t1, q1, t2, q2)
        e0, dor0, fast0 = o.pair_table_ad4(t1, 0.0, t2, 0.0)
        coef = 0.01097 * (abs(q1) * ad4_kind(t2)[5] + abs(q2) * ad4_kind(t1)[5])
        np.testing.assert_allclose(e, e0 + q1 * q2 * W[2] * cr + coef * W[3] * gr, rtol=1e-9, atol=1e-12)
        # beyond 8 A only the charge terms remain
        assert np.all(e0[2049:] == pytest.approx(W[3] * (ad4_kind(t1)[4] * ad4_kind(t2)[5] +
                                                         ad4_kind(t2)[4] * ad4_kind(t1)[5]) * gr[2049:], abs=1e-12))


def test_maps_equal_pair_potential_sums():
    xyz, ad, q, cav = small_receptor_ad4()
    sel = np.argsort(np.linalg.norm(xyz - cav, axis=1))[:40]  # the pocket lining
    o = Oracle(prec=64, sf=2)
    c = np.asarray(cav, dtype=np.float64)
    n, geo = o.set_grid(c, np.array([3.0, 3.0, 3.0]), 0.375)
    types = [C, A, OA, HD]
    o.build_maps_ad4(xyz[sel], ad[sel], q[sel], types)
    maps = {t: o.get_map(t).reshape(n[2], n[1], n[0]) for t in types + [AD4_MAP_ELEC, AD4_MAP_DSOLV]}
    rng = np.random.default_rng(3)
    for _ in range(12):
        ix, iy, iz = (int(rng.integers(0, n[d])) for d in range(3))
        p = geo[:3] + geo[6:9] * np.array([ix, iy, iz])
        for t in types:
            qq = float(rng.normal(0, 0.3))
            want = 0.0
            for j in sel:
                want += o.ad4_terms(t, qq, int(ad[j]), float(q[j]), float(np.linalg.norm(xyz[j] - p)))[4]
            got = maps[t][iz, iy, ix] + qq * maps[AD4_MAP_ELEC][iz, iy, ix] + abs(qq) * maps[AD4_MAP_DSOLV][iz, iy, ix]
            assert got == pytest.approx(want, rel=1e-10, abs=1e-10), (t, qq)


@pytest.fixture(scope="module")
def scene():
    return oracle_scene_ad4(64, 10.0)


def _fd(o, li, p, nt, h=1e-6):
    """Central differences of the energy: position, rotation about each world
    axis (the torque components) and torsions."""
    fd = np.zeros(6 + nt)
    for j in list(range(3)) + list(range(7, 7 + nt)):
        a, b = p.copy(), p.copy()
        a[j] += h
        b[j] -= h
        fd[j if j < 3 else j - 1] = (o.eval_deriv(li, a)[0] - o.eval_deriv(li, b)[0]) / (2 * h)
    for ax in range(3):
        fd[3 + ax] = (o.eval_deriv(li, _rotate(p, ax, h))[0] - o.eval_deriv(li, _rotate(p, ax, -h))[0]) / (2 * h)
    return fd


def test_energy_central_differences(scene):
    """Pair-free copies: the three-map grid term and the tree derivative are
    exact to FD accuracy.  With pairs the gradient uses the table's dor,
    interpolated separately from e (precalculate.h:48-65), so the bound is the
    table's consistency, as for Vina; AD4.2's smoothing plateaus put kinks
    inside table intervals, where the central-difference dor averages the two
    slopes, hence 2 % of the largest component."""
    o, ligs, (xyz, ad, q, cav) = scene
    bare = [o.add_ligand(l, pairs=np.zeros((0, 2), dtype=np.int32)) for l in ligs]
    rng = np.random.default_rng(7)
    for li, lig in enumerate(ligs):
        p = random_poses(rng, lig, cav, 1.5, 1)[0]
        p = o.bfgs(li, p, max_steps=15)[1][:7 + lig.n_tors]  # a relaxed pose: no capped clashes
        g = o.eval_deriv(li, p)[1]
        sc = max(1.0, np.abs(g).max())
        np.testing.assert_allclose(g, _fd(o, li, p, lig.n_tors), rtol=0, atol=2e-2 * sc, err_msg=f"ligand {li}")
        gb = o.eval_deriv(bare[li], p)[1]
        np.testing.assert_allclose(gb, _fd(o, bare[li], p, lig.n_tors), rtol=1e-6,
                                   atol=1e-6 * max(1.0, np.abs(gb).max()), err_msg=f"bare ligand {li}")


def _rotate(p, ax, h):
    """Pose p with its orientation rotated by angle h about axis ax (world frame)."""
    q = p[3:7]
    rq = np.array([math.cos(h / 2), 0.0, 0.0, 0.0])
    rq[1 + ax] = math.sin(h / 2)
    w1, x1, y1, z1 = rq
    w2, x2, y2, z2 = q
    out = p.copy()
    out[3:7] = [w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2]
    return out


def test_score_terms_and_final_score(scene):
    o, ligs, (xyz, ad, q, cav) = scene
    rng = np.random.default_rng(9)
    lig = ligs[1]
    p = random_poses(rng, lig, cav, 1.0, 1)[0]
    inter, intra = o.score_terms(1, p)
    e, _, _, terms = o.eval_deriv(1, p)
    # same grid part with the same caps (v = 1000 both) up to the fast-table intra
    assert inter == pytest.approx(terms[0], rel=1e-9)
    assert abs(intra - terms[1]) < 0.05 * max(1.0, abs(terms[1]))
    # ad4_tors_add: total = inter + w_rot TORSDOF
    assert o.conf_independent(lig.torsdof, inter) == pytest.approx(inter + W[5] * lig.torsdof)


def test_fp32_restatement_agrees(scene):
    o64, ligs, (xyz, ad, q, cav) = scene
    o32, _, _ = oracle_scene_ad4(32, 10.0)
    rng = np.random.default_rng(4)
    for li, lig in enumerate(ligs):
        p = random_poses(rng, lig, cav, 1.0, 1)[0]
        p = o64.bfgs(li, p, max_steps=15)[1][:7 + lig.n_tors]
        e64 = o64.eval_deriv(li, p)[0]
        e32 = o32.eval_deriv(li, p)[0]
        assert e32 == pytest.approx(e64, rel=1e-4, abs=2e-3)


def test_ligand_view():
    for lig in small_ligands_ad4():
        assert lig.ad is not None and len(lig.ad) == lig.n_atoms
        hd = lig.ad == HD
        assert np.all(lig.heavy[hd] == 0) and np.all(lig.xs[hd] == -1)
        assert lig.torsdof == lig.n_tors


def test_eval_batch_with_hydrogens(scene):
    """The threaded batch evaluation (the bench's FP64 error leg) on ligands
    whose polar hydrogens have no X-Score type."""
    o, ligs, (xyz, ad, q, cav) = scene
    assert any(np.any(l.xs < 0) for l in ligs)
    rng = np.random.default_rng(12)
    pl, P = [], []
    for li, lig in enumerate(ligs):
        for p in random_poses(rng, lig, cav, 1.0, 2):
            pl.append(li)
            P.append(p)
    e = o.eval_batch(np.array(pl), np.array(P), n_threads=4)
    for i, li in enumerate(pl):
        assert e[i] == o.eval_deriv(li, P[i])[0]
