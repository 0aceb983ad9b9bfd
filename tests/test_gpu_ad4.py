"""AutoDock 4.2 scoring on the GPU (through the C ABI) against the oracle.

Tolerances
  maps          FP32 sums over the receptor vs the FP64 brute force: 1e-4
                relative + 1e-4 absolute where |map| < 1000 on all but 0.1 %
                of the points (cutoff flips, each under 0.05); the capped
                12-6 / 12-10 cores at 1e-4 relative
  energy        vs the FP64 oracle on the GPU's own FP32 maps: 1e-4 relative
                (+ 2e-4 absolute); gradient 1e-3 * max(1, |g|_inf).  The GPU
                splits each pair's table into the type-pair part and the two
                charge terms (the oracle samples one table per pair, as the
                reference does); the sum is the same interpolation
  rescoring     FP64 terms kernel vs the oracle's fast tables: 1e-5 relative
  BFGS minima   0.01 kcal/mol from identical near-minimum starts as often as
                the FP32 oracle reaches the FP64 minimum, less 10 %; every run
                ends at or below the GPU's energy of its start (1e-6
                relative, for the lane order of the sums), which is the FP64
                one to the energy tolerance (starts in deep clashes, E ~ 1e5,
                sit 1e-5 relative from FP64 in FP32)
  post-process  final energy = inter + w_rot TORSDOF, unbound = -intra, no
                refinement, components equal to the rescoring of the pose
  weight grad   dE/d(w_vdw, w_hb, w_elec, w_dsolv) vs central differences of
                the FP64 oracle with maps and tables rebuilt per weight set
                (h = 1e-3): 2e-3 * max(1, |dE/dw|_inf); with curl off and the
                atoms in the box, sum_k w_k dE/dw_k = E (1e-4 relative)
"""
import numpy as np
import pytest

from helpers import (ad4_map_types, box_of, oracle_scene_ad4, random_poses, small_ligands_ad4,
                     small_receptor_ad4)
from oracle.oracle import AD4_MAP_DSOLV, AD4_MAP_ELEC, AD4_WEIGHTS, Oracle
from unidock_tools import _lib, synth

pytestmark = pytest.mark.gpu
EDGE = 10.0


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def sc():
    xyz, ad, q, cav = small_receptor_ad4()
    ligs = small_ligands_ad4()
    types = ad4_map_types(ligs)
    _say("[gpu] ad4: creating context")
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_AD4)
    c, s = box_of(cav, EDGE)
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps_ad4(xyz, ad, q, types, _lib.UD_GRID_F32)
    ctx.set_ligands(synth.LigandBatch(list(ligs)).arrays())
    o64, _, _ = oracle_scene_ad4(64, EDGE)
    # FP64 oracle on the GPU's FP32 maps: isolates the evaluation
    o64m, _, _ = oracle_scene_ad4(64, EDGE)
    o32m, _, _ = oracle_scene_ad4(32, EDGE)
    for t in types + [AD4_MAP_ELEC, AD4_MAP_DSOLV]:
        o64m.set_map(t, ctx.get_map(t).astype(np.float64))
        o32m.set_map(t, ctx.get_map(t).astype(np.float64))
    yield dict(ctx=ctx, o64=o64, o64m=o64m, o32m=o32m, ligs=ligs, cav=cav, types=types)
    ctx.close()


def _poses(sc, n_per, seed, spread):
    rng = np.random.default_rng(seed)
    pl, P = [], []
    for li, lig in enumerate(sc["ligs"]):
        for p in random_poses(rng, lig, sc["cav"], spread, n_per):
            pl.append(li)
            P.append(p)
    return np.array(pl, dtype=np.int32), np.array(P)


def test_maps_match_oracle(sc):
    _say("[gpu] ad4 maps vs FP64 oracle")
    ctx, o64 = sc["ctx"], sc["o64"]
    for t in sc["types"] + [AD4_MAP_ELEC, AD4_MAP_DSOLV]:
        g = ctx.get_map(t).astype(np.float64)
        r = o64.get_map(t)
        soft = np.abs(r) < 1000
        # a receptor atom whose r^2 lies on the 20.48 A cutoff in FP32 but not
        # in FP64 (or the reverse) moves a point of the Coulomb map by at most
        # w_elec 332 |q| / (20.48 eps(20.48)) ~ 0.02 per atom
        bad = np.abs(g[soft] - r[soft]) > 1e-4 * (1 + np.abs(r[soft]))
        assert bad.mean() <= 1e-3, (t, bad.mean())
        assert np.abs(g[soft] - r[soft]).max() < 0.05, (t, np.abs(g[soft] - r[soft]).max())
        np.testing.assert_allclose(g[~soft], r[~soft], rtol=1e-4, err_msg=f"map {t}")


def test_score_grad_matches_oracle(sc):
    _say("[gpu] ad4 score_grad vs oracle")
    ctx, o = sc["ctx"], sc["o64m"]
    pl, P = _poses(sc, 4, 1, 1.5)
    R = np.array([o.bfgs(int(l), p, max_steps=15)[1] for l, p in zip(pl, P)])  # off the capped walls
    for v in ((1000.0, 1000.0, 1000.0), (10.0, 10.0, 10.0)):
        e, g = ctx.score_grad(pl, R, v)
        for i in range(len(pl)):
            li = int(pl[i])
            nt = sc["ligs"][li].n_tors
            e64, g64, _, _ = o.eval_deriv(li, R[i], v)
            assert e[i] == pytest.approx(e64, rel=1e-4, abs=2e-4), (i, e[i], e64)
            s = max(1.0, np.abs(g64).max())
            np.testing.assert_allclose(g[i, :6 + nt], g64, rtol=0, atol=1e-3 * s, err_msg=f"pose {i}")
            assert np.all(g[i, 6 + nt:] == 0)


def test_score_terms_match_oracle(sc):
    _say("[gpu] ad4 score_terms vs oracle")
    ctx, o = sc["ctx"], sc["o64m"]
    pl, P = _poses(sc, 4, 2, 2.0)
    inter, intra = ctx.score_terms(pl, P)
    lo = o.geo[:3]
    hi = lo + o.geo[6:9] * (np.array(o.n) - 1)
    n_in = 0
    for i in range(len(pl)):
        li = int(pl[i])
        a, b = o.score_terms(li, P[i])
        c = o.eval_deriv(li, P[i])[2]
        lig = sc["ligs"][li]
        if np.all(c >= lo) and np.all(c <= hi):
            n_in += 1
            assert inter[i] == pytest.approx(a, rel=1e-5, abs=1e-4), (i, inter[i], a)
        else:  # the 1e6 slope penalty of each of the three maps, from FP32 coordinates
            allow = 1e6 * 4e-6 * lig.n_atoms * (1.0 + 2.0 * np.abs(lig.charge).max())
            assert abs(inter[i] - a) <= 1e-5 * abs(a) + allow, (i, inter[i], a)
        assert intra[i] == pytest.approx(b, rel=1e-5, abs=1e-4), (i, intra[i], b)
    assert n_in >= 6


def test_bfgs_reaches_same_minima(sc):
    _say("[gpu] ad4 optimize vs oracle bfgs")
    ctx, o = sc["ctx"], sc["o64m"]
    rng = np.random.default_rng(3)
    pl, P = _poses(sc, 4, 3, 1.5)
    S = []
    for l, p in zip(pl, P):
        m = o.bfgs(int(l), p, max_steps=300)[1]
        nt = sc["ligs"][int(l)].n_tors
        m[0:3] += rng.normal(scale=0.05, size=3)
        qq = m[3:7] + rng.normal(scale=0.01, size=4)
        m[3:7] = qq / np.linalg.norm(qq)
        m[7:7 + nt] += rng.normal(scale=0.02, size=nt)
        S.append(m)
    S = np.array(S)
    Pg, eg, _ = ctx.optimize(pl, S, (1000.0, 1000.0, 1000.0))
    es, _ = ctx.score_grad(pl, S, (1000.0, 1000.0, 1000.0))  # the GPU's energy of each start
    agree = base = 0
    for i in range(len(pl)):
        li = int(pl[i])
        e64 = o.bfgs(li, S[i])[0]
        e32 = sc["o32m"].bfgs(li, S[i])[0]
        agree += abs(eg[i] - e64) <= 0.01
        base += abs(e32 - e64) <= 0.01
        _say(f"  start {i}: gpu {eg[i]:.4f} fp32 {e32:.4f} fp64 {e64:.4f}")
        assert eg[i] == pytest.approx(o.eval_deriv(li, Pg[i].astype(np.float64))[0], rel=1e-4, abs=1e-3)
        # descends from its own start energy (score and search may sum in
        # different lane orders), which is the FP64 one to the energy tolerance
        e_start = o.eval_deriv(li, S[i])[0]
        assert es[i] == pytest.approx(e_start, rel=1e-4, abs=2e-4)
        assert eg[i] <= es[i] + 1e-6 * abs(es[i]) + 1e-4
    # starts in clashes or outside the box give chaotic FP32 trajectories: the
    # GPU must agree with FP64 as often as the FP32 restatement does, less 10 %
    assert agree >= base - 0.1 * len(pl), (agree, base, len(pl))


def test_dock_and_postprocess(sc):
    _say("[gpu] ad4 dock + postprocess")
    ctx, o = sc["ctx"], sc["o64m"]
    ligs = sc["ligs"]
    exh = 16
    P, E, _ = ctx.dock(exhaustiveness=exh, global_steps=6, seed=5)
    # every chain's reported best energy is the authentic-cap (v = 1000) energy of its pose
    for li in range(len(ligs)):
        for c in range(0, exh, 5):
            k = li * exh + c
            assert E[k] == pytest.approx(o.eval_deriv(li, P[k].astype(np.float64))[0], rel=1e-4, abs=2e-3)
    pl_all = np.repeat(np.arange(len(ligs)), exh).astype(np.int32)
    eg, _ = ctx.score_grad(pl_all, P)
    np.testing.assert_allclose(eg, E, rtol=1e-4, atol=1e-3)
    modes = ctx.postprocess(max_modes=9, min_rmsd=1.0, refine_step=3)  # AD4.2 modes are not refined
    w_rot = _lib.AD4_WEIGHTS[5]
    for li, lig in enumerate(ligs):
        got = modes[li]
        assert got
        energies = [m["energy"] for m in got]
        assert energies == sorted(energies)
        pl = np.full(len(got), li, dtype=np.int32)
        Pm = np.array([m["pose"] for m in got])
        inter, intra = ctx.score_terms(pl, Pm)
        # atoms on or beyond the box faces pay the 1e6 / Å penalty, where the
        # FP32 and FP64 coordinates of the same pose differ by ~1e-6 Å
        C = ctx.pose_coords(pl, Pm)[:, :lig.n_atoms]
        c0, s0 = box_of(sc["cav"], EDGE)
        span = 0.375 * np.ceil(s0 / 0.375)
        lo, hi = c0 - span / 2, c0 + span / 2
        near_face = ((C < lo + 1e-3) | (C > hi - 1e-3)).any(axis=(1, 2))
        for k, m in enumerate(got):
            np.testing.assert_array_equal(m["pose"], P[m["chain"]])  # not refined
            assert m["inter"] == inter[k] and m["intra"] == intra[k]
            assert m["energy"] == pytest.approx(m["inter"] + w_rot * lig.torsdof, rel=1e-6, abs=1e-5)
            assert m["unbound"] == pytest.approx(-m["intra"], abs=0)
            a, b = o.score_terms(li, m["pose"].astype(np.float64))
            # poses of ligands that do not fit the small box carry the box penalty
            boxed = abs(a) >= 100 or near_face[k]
            allow = 1e6 * 4e-6 * lig.n_atoms * (1.0 + 2.0 * np.abs(lig.charge).max()) if boxed else 0.0
            assert abs(m["inter"] - a) <= 1e-5 * abs(a) + 1e-4 + allow, (li, m["inter"], a)


def _unit(p):
    q = p.copy()
    q[3:7] /= np.linalg.norm(q[3:7])
    return q


def _oracle_w(weights):
    xyz, ad, q, cav = small_receptor_ad4()
    ligs = small_ligands_ad4()
    o = Oracle(prec=64, sf=2, weights=weights)
    c, s = box_of(cav, EDGE)
    o.set_grid(c, s, 0.375)
    o.build_maps_ad4(xyz, ad, q, ad4_map_types(ligs))
    for l in ligs:
        o.add_ligand(l)
    return o


def test_weight_gradient_central_differences(sc):
    _say("[gpu] ad4 energy_grad weight gradient vs FP64 central differences")
    ctx = sc["ctx"]
    pl, P = _poses(sc, 3, 21, 1.0)
    P = np.array([_unit(p) for p in P])
    e, _, gw = ctx.energy_grad(pl, P)
    # the same evaluation as the search's
    es, _ = ctx.score_grad(pl, P)
    np.testing.assert_allclose(e, es, rtol=1e-6, atol=1e-5)
    w0 = np.array(AD4_WEIGHTS, dtype=np.float64)
    h = 1e-3
    fd = np.zeros((len(pl), 4))
    for k in range(4):
        wp, wm = w0.copy(), w0.copy()
        wp[k] += h
        wm[k] -= h
        op, om = _oracle_w(wp), _oracle_w(wm)
        for i, li in enumerate(pl):
            fd[i, k] = (op.eval_deriv(li, P[i])[0] - om.eval_deriv(li, P[i])[0]) / (2 * h)
    _say(f"[gpu] ad4 |dE/dw|_inf per term: {np.abs(fd).max(axis=0).round(3).tolist()}")
    for i in range(len(pl)):
        s_ = max(1.0, np.abs(fd[i]).max())
        np.testing.assert_allclose(gw[i, :4], fd[i], rtol=0, atol=2e-3 * s_, err_msg=f"pose {i}")
    # glue joins no ligand pairs; the torsional weight is conformation independent
    assert np.all(gw[:, 4:] == 0)
    # vdW, electrostatic and desolvation are exercised on these poses
    assert np.all(np.abs(fd[:, [0, 2, 3]]).max(axis=0) > 0.1)


def test_weight_gradient_euler_identity(sc):
    _say("[gpu] ad4 energy_grad Euler identity (curl off)")
    ctx, o64 = sc["ctx"], sc["o64"]
    lo = o64.geo[:3]
    hi = lo + o64.geo[6:9] * (np.array(o64.n) - 1)
    pl, P = _poses(sc, 4, 22, 1.0)
    P = np.array([_unit(p) for p in P])
    v = (3.4028234663852886e38,) * 3
    e, _, gw = ctx.energy_grad(pl, P, v)
    coords = ctx.pose_coords(pl, P)
    w = np.array(AD4_WEIGHTS[:4], dtype=np.float64)
    n_in = 0
    for i, li in enumerate(pl):
        c = coords[i, :sc["ligs"][li].n_atoms]
        if np.all(c >= lo) and np.all(c <= hi):
            n_in += 1
            assert float(np.dot(w, gw[i, :4])) == pytest.approx(float(e[i]), rel=1e-4, abs=1e-4)
    assert n_in >= 6


def test_ad4_fp16_maps_energy_grad(sc):
    """AD4.2 over FP16 maps: energy_grad equals the search's evaluation on the
    same maps, and stays within the FP16 maps' error of the FP32-map results
    (weight gradients too: the per-term maps are FP32 in both)."""
    _say("[gpu] ad4 energy_grad FP16 maps")
    xyz, ad, q, cav = small_receptor_ad4()
    pl, P = _poses(sc, 3, 24, 1.0)
    P = np.array([_unit(p) for p in P])
    e32, g32, w32 = sc["ctx"].energy_grad(pl, P)
    ctx = _lib.Context(0)
    try:
        ctx.set_scoring(_lib.UD_SF_AD4)
        c, s = box_of(cav, EDGE)
        ctx.set_box(c, s, 0.375)
        ctx.compute_maps_ad4(xyz, ad, q, sc["types"], _lib.UD_GRID_F16)
        ctx.set_ligands(synth.LigandBatch(list(sc["ligs"])).arrays())
        e16, g16, w16 = ctx.energy_grad(pl, P)
        es16, gs16 = ctx.score_grad(pl, P)
    finally:
        ctx.close()
    # the search's evaluation on the same maps, up to the renormalised
    # quaternion (as test_matches_score_grad_at_unit_quaternions bounds it)
    for i in range(len(pl)):
        assert abs(e16[i] - es16[i]) <= 1e-6 * abs(es16[i]) + 2e-6 * max(1.0, np.abs(gs16[i]).max()), i
    for i in range(len(pl)):
        assert abs(e16[i] - e32[i]) <= 2e-3 * max(1.0, abs(float(e32[i]))), (i, e16[i], e32[i])
        np.testing.assert_allclose(g16[i], g32[i], rtol=0, atol=5e-3 * max(1.0, np.abs(g32[i]).max()))
        np.testing.assert_allclose(w16[i], w32[i], rtol=0, atol=2e-3 * max(1.0, np.abs(w32[i]).max()))


def test_torch_autograd_ad4(sc):
    """DockingEnergy under AD4.2: forward and backward equal ud_energy_grad on
    the same maps; new weights rebuild the maps along the weight gradient."""
    _say("[gpu] ad4 torch autograd energy")
    import torch
    from unidock_tools.torch_energy import DockingEnergy
    xyz, ad, q, cav = small_receptor_ad4()
    c, s = box_of(cav, EDGE)
    model = DockingEnergy(xyz, ad, list(sc["ligs"]), c, s, sf=_lib.UD_SF_AD4, receptor_charge=q)
    pl, P = _poses(sc, 3, 23, 1.0)
    P = np.array([_unit(p) for p in P])
    e_ref, gp_ref, gw_ref = sc["ctx"].energy_grad(pl, P)
    k = 7 + max(l.n_tors for l in sc["ligs"])
    poses = torch.tensor(P[:, :k], dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.tensor(AD4_WEIGHTS, dtype=torch.float64, requires_grad=True)
    e = model(poses, pl, w)
    np.testing.assert_allclose(e.detach().cpu().numpy(), e_ref, rtol=1e-6, atol=1e-6)
    r = torch.linspace(-1.0, 2.0, len(pl), dtype=torch.float64, device="cuda")
    (e * r).sum().backward()
    rn = r.cpu().numpy()[:, None]
    np.testing.assert_allclose(poses.grad.cpu().numpy(), rn * gp_ref[:, :k], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(w.grad.numpy(), (rn * gw_ref[:, :len(AD4_WEIGHTS)]).sum(0), rtol=1e-4, atol=1e-4)
    w2 = w.detach().clone()
    w2[2] *= 1.01
    e2 = model(poses.detach(), pl, w2).cpu().numpy()
    pred = e_ref + gw_ref[:, 2] * float(w2[2] - w.detach()[2])
    np.testing.assert_allclose(e2, pred, rtol=0, atol=1e-3 * max(1.0, np.abs(e_ref).max()))


def test_state_errors(sc):
    ctx = sc["ctx"]
    # term-weight gradients need the receptor the maps were built from
    c3 = _lib.Context(0)
    try:
        c3.set_scoring(_lib.UD_SF_AD4)
        c, s = box_of(sc["cav"], EDGE)
        n = c3.set_box(c, s, 0.375)
        types = sc["types"] + [AD4_MAP_ELEC, AD4_MAP_DSOLV]
        c3.set_maps(types, np.zeros(len(types) * int(np.prod(n)), dtype=np.float32))
        c3.set_ligands(synth.LigandBatch(list(sc["ligs"])).arrays())
        with pytest.raises(_lib.UdError) as ei:
            c3.energy_grad(np.zeros(1, np.int32), np.zeros((1, _lib.UD_POSE_FLOATS), np.float32))
        assert ei.value.code == _lib.UD_ERR_STATE
    finally:
        c3.close()
    with pytest.raises(_lib.UdError) as ei:
        ctx.refine(np.zeros(1, np.int32), np.zeros((1, _lib.UD_POSE_FLOATS), np.float32))
    assert ei.value.code == _lib.UD_ERR_STATE
    # ligands without AutoDock types are refused under AD4.2
    c2 = _lib.Context(0)
    try:
        c2.set_scoring(_lib.UD_SF_AD4)
        a = synth.LigandBatch(list(sc["ligs"])).arrays()
        for k in ("atom_ad", "atom_charge", "torsdof"):
            a.pop(k)
        with pytest.raises(_lib.UdError) as ei:
            c2.set_ligands(a)
        assert ei.value.code == _lib.UD_ERR_INVALID
    finally:
        c2.close()


def test_autogrid_files_round_trip(sc, tmp_path):
    """Maps written in AutoGrid's format and read back (ad4cache::read) score
    like the maps they came from (to the files' 4 significant digits).
    AutoGrid boxes have an even number of voxels: 26 x 0.375 A."""
    _say("[gpu] ad4 AutoGrid map files")
    xyz, ad, q, cav = small_receptor_ad4()
    ligs = list(sc["ligs"])
    center = np.round(np.asarray(cav, dtype=np.float64), 3)
    c1, c2 = _lib.Context(0), _lib.Context(0)
    try:
        c1.set_scoring(_lib.UD_SF_AD4)
        n = np.array(c1.set_box(center, (9.75, 9.75, 9.75), 0.375))
        assert np.all(n == 27)
        c1.compute_maps_ad4(xyz, ad, q, sc["types"], _lib.UD_GRID_F32)
        c1.set_ligands(synth.LigandBatch(ligs).arrays())
        names = {t: _lib.AD_TYPE_NAMES[t] for t in sc["types"]}
        names[AD4_MAP_ELEC], names[AD4_MAP_DSOLV] = "e", "d"
        for t, nm in names.items():
            _lib.write_ad4_map(str(tmp_path / f"rec.{nm}.map"), center, n - 1, 0.375, c1.get_map(t))
        c2.set_scoring(_lib.UD_SF_AD4)
        c2.load_ad4_maps(str(tmp_path / "rec"), [int(t) for l in ligs for t in l.ad])
        assert tuple(c2.npts) == tuple(c1.npts)
        c2.set_ligands(synth.LigandBatch(ligs).arrays())
        pl, P = _poses(sc, 2, 4, 1.0)
        e1, _ = c1.score_grad(pl, P)
        e2, _ = c2.score_grad(pl, P)
        np.testing.assert_allclose(e2, e1, rtol=2e-3, atol=0.05)
    finally:
        c1.close()
        c2.close()


def test_redock_1iep_ad4():
    """The reference's 1iep fixture (test_unidock.py:21 box) with AD4.2 maps
    built from the receptor: docking reaches a bound pose (negative final
    energy)."""
    import os
    _say("[gpu] ad4 redock 1iep")
    here = os.path.join(os.path.dirname(__file__), "golden")
    ls = _lib.LigandSet()
    ls.add_pdbqt(open(os.path.join(here, "1iep_ligand.pdbqt")).read())
    xyz, ad, q = _lib.receptor_ad4_from_pdbqt(open(os.path.join(here, "1iep_receptor.pdbqt")).read())
    ctx = _lib.Context(0)
    try:
        ctx.set_scoring(_lib.UD_SF_AD4)
        ctx.set_box((15.19, 53.903, 16.917), (20, 20, 20), 0.375)
        arr = ls.arrays(_lib.UD_SF_AD4)
        types = sorted({_lib.ad4_grid_type(int(t)) for t in arr["atom_ad"]} - {-1})
        ctx.compute_maps_ad4(xyz, ad, q, types, _lib.UD_GRID_F32)
        ctx.set_ligands(arr)
        ctx.dock(exhaustiveness=64, global_steps=20, seed=181129, fetch=False)
        modes = ctx.postprocess(max_modes=9, min_rmsd=1.0)
        best = modes[0][0]
        _say(f"  1iep AD4.2 best {best['energy']:.3f} inter {best['inter']:.3f} intra {best['intra']:.3f}")
        assert best["energy"] < -5.0
        # crystal pose: the input coordinates
        xyz0 = ls.ligands(_lib.UD_SF_AD4)[0].coords
        heavy = arr["atom_heavy"].astype(bool)
        c = ctx.pose_coords(np.zeros(1, np.int32), best["pose"][None, :])[0, :len(heavy)]
        rmsd = float(np.sqrt(((c[heavy] - xyz0[heavy]) ** 2).sum(1).mean()))
        _say(f"  1iep AD4.2 best-mode RMSD to the crystal pose {rmsd:.2f} A")
    finally:
        ctx.close()
