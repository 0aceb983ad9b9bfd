"""Ligands at the reference's largest size class (MaxConfig: up to 300 atoms
and 48 torsions, unidock/src/cuda/kernel.h:242-277) against the oracle.

The test ligand has 100 heavy atoms in a 40-torsion tree (synth.make_ligand)
and one hydrogen on each heavy atom, placed 1.0 A away in the heavy atom's
rigid node: 200 atoms, 41 nodes, 46 degrees of freedom and 4,590
intramolecular pairs (hydrogens carry no X-Score type, so they score neither
against the maps nor in pairs, as in the reference).  It lands in the
300-atom size bucket, whose workgroup keeps its 46 x 46 inverse Hessian and
per-pair buffer in LDS.  Scoring (energy and gradient), BFGS from identical
starts and the Monte Carlo search match the FP32 restatement; a 48-torsion
ligand and a 300-atom one are accepted, a 49th torsion or a 301st atom is a
capacity error.
"""
import numpy as np
import pytest
from scipy import stats

from helpers import LIGAND_TYPES, box_of, small_receptor
from oracle.oracle import Oracle
from unidock_tools import _lib, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def with_hydrogens(lig, seed, n_h=None):
    """lig with n_h (default: one per atom) untyped, non-heavy atoms bonded to
    its atoms in turn, each in its atom's node (atoms stay grouped by node)."""
    rng = np.random.default_rng(seed)
    n_h = lig.n_atoms if n_h is None else n_h
    per = [n_h // lig.n_atoms + (1 if i < n_h % lig.n_atoms else 0) for i in range(lig.n_atoms)]
    coords, xs, heavy, ab, ae, remap = [], [], [], [], [], np.zeros(lig.n_atoms, dtype=np.int64)
    for k in range(len(lig.parent)):
        ab.append(len(coords))
        for i in range(lig.abegin[k], lig.aend[k]):
            remap[i] = len(coords)
            coords.append(lig.coords[i])
            xs.append(lig.xs[i])
            heavy.append(1)
        for i in range(lig.abegin[k], lig.aend[k]):
            for _ in range(per[i]):
                d = rng.normal(size=3)
                coords.append(lig.coords[i] + d / np.linalg.norm(d))
                xs.append(-1)
                heavy.append(0)
        ae.append(len(coords))
    bonds = np.array([[remap[a], remap[b]] for a, b in lig.bonds], dtype=np.int32)
    return synth.Ligand(coords=np.array(coords), xs=np.array(xs, dtype=np.int32),
                        heavy=np.array(heavy, dtype=np.int32), parent=lig.parent.copy(),
                        abegin=np.array(ab, dtype=np.int32), aend=np.array(ae, dtype=np.int32),
                        origin=lig.origin.copy(), axis=lig.axis.copy(),
                        pairs=remap[lig.pairs].astype(np.int32), bonds=bonds, rotatable=lig.rotatable.copy(),
                        num_tors=lig.num_tors)


def _scene():
    xyz, xs, cav = small_receptor()
    big = with_hydrogens(synth.make_ligand(100, 40, seed=5), seed=6)
    assert big.n_atoms == 200 and big.n_tors == 40 and len(big.pairs) > 4096
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINA)
    # a box the ligand (about 20 A across) is not pressed against: an atom on
    # a face carries the 10^6 / A penalty, which turns coordinate rounding
    # into kcal/mol
    c, s = box_of(cav, 36.0)
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps(xyz, xs, LIGAND_TYPES, _lib.UD_GRID_F32)
    ctx.set_ligands(synth.LigandBatch([big]).arrays())
    os_ = {}
    for prec in (32, 64):
        o = Oracle(prec, 0)
        o.set_grid(c, s, 0.375)
        for t in LIGAND_TYPES:
            o.set_map(t, ctx.get_map(t).astype(np.float64))
        o.add_ligand(big)
        os_[prec] = o
    lo = os_[64].geo[:3]
    hi = lo + os_[64].geo[6:9] * (np.array(os_[64].n) - 1)
    return ctx, os_[32], os_[64], big, cav, (lo, hi)


def _poses(rng, lig, cav, n, spread):
    P = np.zeros((n, _lib.UD_POSE_FLOATS))
    P[:, :3] = cav + rng.uniform(-spread, spread, (n, 3))
    q = rng.normal(size=(n, 4))
    P[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
    P[:, 7:7 + lig.n_tors] = rng.uniform(-np.pi, np.pi, (n, lig.n_tors))
    return P


def test_200_atom_40_torsion_ligand_score_optimize_dock():
    _say("[gpu] 200-atom, 40-torsion ligand")
    ctx, o32, o64, big, cav, (lo, hi) = _scene()
    rng = np.random.default_rng(7)
    n = 6 + big.n_tors
    # energy and gradient
    P = _poses(rng, big, cav, 24, 2.0)
    pl = np.zeros(len(P), dtype=np.int32)
    for lanes in (64, 32):
        ctx.set_lanes(lanes)
        e, g = ctx.score_grad(pl, P, (10.0, 10.0, 10.0))
        for i in range(len(P)):
            e32, g32, _, _ = o32.eval_deriv(0, P[i], (10.0, 10.0, 10.0))
            e64, g64, _, _ = o64.eval_deriv(0, P[i], (10.0, 10.0, 10.0))
            # thousands of FP32 pair terms summed in different orders
            assert e[i] == pytest.approx(e32, rel=1e-4, abs=1e-3), (lanes, i)
            # gradient error against FP64 within a few times the FP32
            # restatement's (40 chained quaternion products round in FP32)
            sc = max(1.0, np.abs(g64).max())
            err_gpu = np.abs(g[i, :n] - g64).max()
            err_32 = np.abs(g32 - g64).max()
            assert err_gpu <= 10.0 * err_32 + 1e-4 * sc, (lanes, i, err_gpu, err_32)
    ctx.set_lanes(0)
    # BFGS from identical starts a small kick away from FP64 minima
    starts = []
    for p in P[:12]:
        m = o64.bfgs(0, p, (1000.0, 1000.0, 1000.0), max_steps=300)[1]
        m[:3] += rng.normal(scale=0.05, size=3)
        m[7:7 + big.n_tors] += rng.normal(scale=0.02, size=big.n_tors)
        starts.append(m)
    starts = np.array(starts)
    Pg, eg, evg = ctx.optimize(pl[:12], starts, (1000.0, 1000.0, 1000.0))
    e64s = np.array([o64.bfgs(0, x)[0] for x in starts])
    e32s = np.array([o32.bfgs(0, x)[0] for x in starts])
    agree_gpu = int(np.sum(np.abs(eg - e64s) <= 0.01))
    agree_32 = int(np.sum(np.abs(e32s - e64s) <= 0.01))
    _say(f"[gpu] 200-atom BFGS minima within 0.01 of FP64: GPU {agree_gpu}/12, FP32 restatement {agree_32}/12")
    assert agree_gpu >= agree_32 - 2
    # the reported energy is the returned pose's (GPU re-evaluation), and the
    # FP32 restatement's within the FP32 noise of ~10^4 kcal/mol clash sums
    ec_gpu, _ = ctx.score_grad(pl[:12], Pg, (1000.0, 1000.0, 1000.0))
    np.testing.assert_allclose(eg, ec_gpu, rtol=1e-5, atol=1e-3)
    checked = 0
    for i in range(12):
        e_o, _, c_o, _ = o32.eval_deriv(0, Pg[i].astype(np.float64))
        if np.any(c_o < lo + 1e-3) or np.any(c_o > hi - 1e-3):
            continue  # an atom on a box face
        checked += 1
        assert eg[i] == pytest.approx(e_o, rel=1e-3, abs=1e-2)
    assert checked >= 6
    # the search: 48 chains, 6 MC steps, against the oracle's chains
    Pd, ed, evd = ctx.dock(exhaustiveness=48, global_steps=6, seed=11)
    li = ctx.last_launches()
    assert [a["atom_cap"] for a in li] == [300]
    ec, _ = o32.dock_cpu([0], 48, 11, 6, [(25 + big.n_atoms) // 3], 16)
    ks = stats.ks_2samp(ed, ec).pvalue
    _say(f"[gpu] 200-atom search: GPU median {np.median(ed):.3f}, oracle {np.median(ec):.3f}, KS p = {ks:.3g}, "
         f"launch lanes {li[0]['lanes']}")
    assert np.all(np.isfinite(ed)) and ks > 1e-3
    # re-scored by the dock's own force path (64 lanes, pair buffer): the
    # same numbers; other paths sum the pairs in another order, which for
    # this ligand moves FP32 energies by up to ~1e-3 relative (the FP32
    # restatement differs from FP64 by more)
    ctx.set_lanes(li[0]["lanes"])
    es, _ = ctx.score_grad(np.zeros(48, dtype=np.int32), Pd, (1000.0, 1000.0, 1000.0))
    ctx.set_lanes(0)
    np.testing.assert_allclose(es, ed, rtol=1e-6, atol=1e-4)
    ctx.close()


def test_capacity_limits():
    """48 torsions and 300 atoms are accepted (and score as the oracle does);
    one more torsion or atom is UD_ERR_CAPACITY."""
    _say("[gpu] capacity limits")
    xyz, xs, cav = small_receptor()
    l48 = synth.make_ligand(60, 48, seed=9)
    l300 = with_hydrogens(synth.make_ligand(120, 20, seed=10), seed=11, n_h=180)
    l49 = synth.make_ligand(60, 49, seed=12)
    assert l48.n_tors == 48 and l300.n_atoms == 300 and l49.n_tors == 49
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINA)
    c, s = box_of(cav, 24.0)
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps(xyz, xs, LIGAND_TYPES, _lib.UD_GRID_F32)
    for bad in (l49, ):
        with pytest.raises(_lib.UdError) as ei:
            ctx.set_ligands(synth.LigandBatch([bad]).arrays())
        assert ei.value.code == _lib.UD_ERR_CAPACITY
    arr = synth.LigandBatch([l300]).arrays()
    arr["atom_offset"] = np.array([0, 301], dtype=np.int32)
    with pytest.raises(_lib.UdError) as ei:
        ctx.set_ligands(arr)
    assert ei.value.code == _lib.UD_ERR_CAPACITY
    assert len(l300.pairs) <= _lib.UD_MAX_PAIRS
    ctx.set_ligands(synth.LigandBatch([l48, l300]).arrays())
    o32, o64 = Oracle(32, 0), Oracle(64, 0)
    for o in (o32, o64):
        o.set_grid(c, s, 0.375)
        for t in LIGAND_TYPES:
            o.set_map(t, ctx.get_map(t).astype(np.float64))
        o.add_ligand(l48)
        o.add_ligand(l300)
    rng = np.random.default_rng(13)
    pl = np.array([0] * 8 + [1] * 8, dtype=np.int32)
    P = np.concatenate([_poses(rng, l48, cav, 8, 2.0), _poses(rng, l300, cav, 8, 2.0)])
    e, g = ctx.score_grad(pl, P, (10.0, 10.0, 10.0))
    for i in range(len(pl)):
        lig = (l48, l300)[pl[i]]
        e32, g32, _, _ = o32.eval_deriv(int(pl[i]), P[i], (10.0, 10.0, 10.0))
        e64, g64, _, _ = o64.eval_deriv(int(pl[i]), P[i], (10.0, 10.0, 10.0))
        assert e[i] == pytest.approx(e32, rel=1e-4, abs=1e-3), i
        # 48 chained quaternion products carry FP32 rounding into the
        # coordinates: the GPU's gradient error against FP64 stays within a
        # few times the FP32 restatement's
        sc = max(1.0, np.abs(g64).max())
        err_gpu = np.abs(g[i, :6 + lig.n_tors] - g64).max()
        err_32 = np.abs(g32 - g64).max()
        assert err_gpu <= 10.0 * err_32 + 1e-4 * sc, (i, err_gpu, err_32)
    Pd, ed, _ = ctx.dock(exhaustiveness=8, global_steps=2, seed=3)
    assert np.all(np.isfinite(ed))
    ctx.close()
