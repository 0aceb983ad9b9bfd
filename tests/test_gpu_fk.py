"""Forward kinematics of ligands with 9 and 10 torsions, the most the
register-Hessian search kernel takes: each lane quad of a 32-lane group
handles nodes j and j + 8, so these trees use both node slots (and 8- and
16-lane groups more).  At every lane width: energies and gradients against
the FP32 restatement, atom coordinates against the oracle's, and the BFGS
energies of the search kernel against the scoring kernel's energies of the
poses it returns."""
import numpy as np
import pytest

from helpers import LIGAND_TYPES, box_of, random_poses, small_receptor
from oracle.oracle import Oracle
from unidock_tools import _lib, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_nine_and_ten_torsion_trees_at_every_lane_width():
    _say("[gpu] 9- and 10-torsion ligands: score, coordinates, optimize at lanes 8 / 16 / 32 / 64")
    xyz, xs, cav = small_receptor()
    ligs = [synth.make_ligand(26, 10, seed=701), synth.make_ligand(22, 9, seed=702),
            synth.make_ligand(30, 10, seed=703)]
    assert [l.n_tors for l in ligs] == [10, 9, 10]
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINA)
    c, s = box_of(cav)
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps(xyz, xs, LIGAND_TYPES, _lib.UD_GRID_F32)
    ctx.set_ligands(synth.LigandBatch(ligs).arrays())
    o32 = Oracle(prec=32, sf=0)
    o32.set_grid(c, s, 0.375)
    for t in LIGAND_TYPES:
        o32.set_map(t, ctx.get_map(t).astype(np.float64))
    for l in ligs:
        o32.add_ligand(l)
    rng = np.random.default_rng(704)
    pl = np.repeat(np.arange(len(ligs), dtype=np.int32), 8)
    P = np.concatenate([random_poses(rng, l, cav, 2.0, 8) for l in ligs])
    caps = (10.0, 10.0, 10.0)  # curl keeps clash energies small, so 1e-5 relative is FP32 rounding
    try:
        for lanes in (8, 16, 32, 64):
            ctx.set_lanes(lanes)
            e, g = ctx.score_grad(pl, P, caps)
            for i in range(len(pl)):
                e32, g32, c32, _ = o32.eval_deriv(int(pl[i]), P[i], caps)
                nt = ligs[int(pl[i])].n_tors
                assert e[i] == pytest.approx(e32, rel=1e-5, abs=1e-4), (lanes, i)
                np.testing.assert_allclose(g[i, :6 + nt], g32, rtol=0, atol=1e-4 * max(1.0, np.abs(g32).max()))
            C = ctx.pose_coords(pl, P)
            for i in range(len(pl)):
                rc = o32.eval_deriv(int(pl[i]), P[i], caps)[2]
                np.testing.assert_allclose(C[i, :len(rc)], rc, atol=1e-4)
            # the search kernel (register inverse Hessian at these torsion
            # counts) reports the energy the scoring kernel gives its pose
            Pg, eg, _ = ctx.optimize(pl, P, (1000.0, 1000.0, 1000.0))
            es, _ = ctx.score_grad(pl, Pg, (1000.0, 1000.0, 1000.0))
            np.testing.assert_allclose(eg, es, rtol=1e-5, atol=1e-4, err_msg=f"lanes {lanes}")
            _say(f"[gpu] lanes={lanes}: {len(pl)} poses, BFGS energies {np.min(eg):.2f}..{np.max(eg):.2f}")
    finally:
        ctx.set_lanes(0)
        ctx.close()
