"""Seeded global search on the GPU against the oracle's CPU search.

The device and the oracle draw the same Philox counters (counter = (step,
chain, purpose, 0), key = seed), so a chain's random start and its first
mutation are the same numbers on both; trajectories then differ only by FP32
rounding.  Checks:
  * one Monte Carlo step per chain whose BFGS runs take one step each
    (random start, mutation, one line search with hunt caps, Metropolis, one
    with authentic caps): every chain's result equals the FP32 restatement's
    up to rounding, under both first-line-search rules; with full BFGS runs
    the chain-best energies follow one distribution;
  * the reference's own fixture (unidock/test/ut/test_unidock.py: 1iep, box
    centre (15.19, 53.903, 16.917), 20 A), fast preset (exhaustiveness 128,
    max_step 20) over 8 seeds and detail preset (512, 40) over 4 seeds: the
    pooled chain-best energies of GPU and oracle follow one distribution (KS),
    the per-seed best energies too (Mann-Whitney), the number of seeds whose
    best pose lies within 2 A of the crystal pose agrees, and at the detail
    preset every seed finds the crystal mode, as the oracle does.
"""
import os

import numpy as np
import pytest
from scipy import stats

from helpers import LIGAND_TYPES, box_of, oracle_scene, small_ligands, small_receptor
from unidock_tools import _lib, synth

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


def _search_scene(f0_zero):
    xyz, xs, cav = small_receptor()
    ligs = small_ligands()
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINA)
    c, s = box_of(cav)
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps(xyz, xs, LIGAND_TYPES, _lib.UD_GRID_F32)
    ctx.set_ligands(synth.LigandBatch(list(ligs)).arrays())
    o32, _, _ = oracle_scene(32)
    # the oracle on the GPU's own maps: isolates the search from map rounding
    for t in LIGAND_TYPES:
        o32.set_map(t, ctx.get_map(t).astype(np.float64))
    o32.set_ls_f0_zero(f0_zero)
    return ctx, o32, ligs


@pytest.mark.parametrize("f0_zero", [False, True])
def test_first_mutation_and_bfgs_step_match_oracle(f0_zero):
    """One MC step whose two BFGS runs take one quasi-Newton step each
    (local_steps = 1): random start, mutation (same Philox numbers), the
    steepest-descent line search with hunt caps, Metropolis, the one with
    authentic caps.  Every chain's best energy and pose equal the FP32
    restatement's up to rounding."""
    _say(f"[gpu] first mutation + BFGS step vs oracle, f0_zero={f0_zero}")
    ctx, o32, ligs = _search_scene(f0_zero)
    o64, _, _ = oracle_scene(64)
    for t in LIGAND_TYPES:
        o64.set_map(t, ctx.get_map(t).astype(np.float64))
    o64.set_ls_f0_zero(f0_zero)
    exh, seed = 24, 4242
    P, e, ev = ctx.dock(exhaustiveness=exh, global_steps=1, local_steps=1, seed=seed, first_ls_f0_zero=f0_zero)
    close, total, worst, flips_ref = 0, 0, 0.0, 0
    for li, lig in enumerate(ligs):
        for k in range(exh):
            ch = li * exh + k
            eo, po, evo, _ = o32.mc_chain(li, seed, ch, 1, local_steps=1)
            e64 = o64.mc_chain(li, seed, ch, 1, local_steps=1)[0]
            flips_ref += abs(eo - e64) / max(1.0, abs(e64)) > 1e-4
            total += 1
            d = abs(e[ch] - eo) / max(1.0, abs(eo))
            worst = max(worst, d)
            if d <= 1e-4:
                close += 1
                np.testing.assert_allclose(P[ch, :7 + lig.n_tors], po[:7 + lig.n_tors], atol=2e-3)
    _say(f"[gpu] first step: {close}/{total} chains within 1e-4 relative, worst {worst:.2e}; "
         f"FP32 and FP64 restatements differ on {flips_ref}")
    # with f0 = f(x) the first Armijo test compares f(x + a p) - f(x) at
    # clash starts, where rounding flips the branch: the GPU may part from
    # the FP32 restatement on at most 5 % of the chains, or twice as many as
    # the FP32 and FP64 restatements part on
    assert total - close <= max(0.05 * total, 2 * flips_ref)
    ctx.close()


@pytest.mark.parametrize("f0_zero", [False, True])
def test_first_mc_step_matches_oracle(f0_zero):
    """One full MC step per chain ((25 + atoms) / 3 BFGS steps, both cap
    sets).  From random starts, which are mostly clashes, FP32 BFGS
    trajectories diverge with rounding, so many chains but not all end at the
    same minimum; the chain-best energies of both follow one distribution."""
    _say(f"[gpu] first MC step vs oracle, f0_zero={f0_zero}")
    ctx, o32, ligs = _search_scene(f0_zero)
    exh, seed = 24, 4242
    P, e, ev = ctx.dock(exhaustiveness=exh, global_steps=1, seed=seed, first_ls_f0_zero=f0_zero)
    eo = np.array([o32.mc_chain(li, seed, li * exh + k, 1)[0] for li in range(len(ligs)) for k in range(exh)])
    close = int(np.sum(np.abs(e - eo) <= 1e-3 * np.maximum(1.0, np.abs(eo))))
    ks = stats.ks_2samp(e, eo).pvalue
    _say(f"[gpu] first MC step: {close}/{len(eo)} chains within 1e-3 relative of the FP32 restatement, KS p = {ks:.3g}")
    assert close >= 0.4 * len(eo)
    assert ks > 1e-3
    ctx.close()


def _fixture():
    rec_xyz, rec_xs = _lib.receptor_from_pdbqt(open(os.path.join(GOLDEN, "1iep_receptor.pdbqt")).read())
    ls = _lib.LigandSet()
    ls.add_pdbqt(open(os.path.join(GOLDEN, "1iep_ligand.pdbqt")).read())
    return rec_xyz, rec_xs, ls, ls.ligands()[0]


def _rmsd(a, b, heavy):
    return float(np.sqrt(np.mean(np.sum((a[heavy] - b[heavy]) ** 2, axis=1))))


@pytest.mark.parametrize("preset", ["fast", "detail"])
def test_1iep_search_distribution_matches_oracle(preset):
    from oracle.oracle import Oracle
    exh, steps, seeds = {"fast": (128, 20, list(range(1, 9))), "detail": (512, 40, list(range(1, 5)))}[preset]
    _say(f"[gpu] 1iep {preset} search: GPU vs oracle over seeds {seeds}")
    rec_xyz, rec_xs, ls, lig = _fixture()
    heavy = np.nonzero(lig.heavy)[0]
    center, size = (15.19, 53.903, 16.917), (20.0, 20.0, 20.0)
    types = sorted({_lib.grid_type(int(t)) for t in lig.xs} - {-1})
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINA)
    ctx.set_box(center, size, 0.375)
    ctx.compute_maps(rec_xyz, rec_xs, types, _lib.UD_GRID_F32)
    ctx.set_ligands(ls.arrays())
    o = Oracle(32, 0)
    o.set_grid(np.array(center), np.array(size), 0.375)
    for t in types:
        o.set_map(t, ctx.get_map(t).astype(np.float64))
    li = o.add_ligand(lig)
    lsteps = [(25 + lig.n_atoms) // 3]
    nth = min(16, len(os.sched_getaffinity(0)))
    g_best, c_best, g_pool, c_pool, g_near, c_near = [], [], [], [], 0, 0
    pl1 = np.zeros(1, dtype=np.int32)
    for seed in seeds:
        P, e, _ = ctx.dock(exhaustiveness=exh, global_steps=steps, seed=seed)
        k = int(np.argmin(e))
        rg = _rmsd(ctx.pose_coords(pl1, P[k][None])[0, :lig.n_atoms], lig.coords, heavy)
        be, _, Pc = o.dock_cpu([li], exh, seed, steps, lsteps, nth, poses=True)
        kc = int(np.argmin(be))
        rc = _rmsd(o.eval_deriv(li, Pc[kc])[2], lig.coords, heavy)
        g_best.append(float(e[k]))
        c_best.append(float(be[kc]))
        g_pool.extend(e.tolist())
        c_pool.extend(be.tolist())
        g_near += rg < 2.0
        c_near += rc < 2.0
        _say(f"[gpu] 1iep {preset} seed {seed}: GPU best {e[k]:.3f} ({rg:.2f} A), oracle best {be[kc]:.3f} ({rc:.2f} A)")
    ks = stats.ks_2samp(g_pool, c_pool).pvalue
    mw = stats.mannwhitneyu(g_best, c_best).pvalue
    _say(f"[gpu] 1iep {preset}: pooled chain-best KS p = {ks:.3g}, per-seed best Mann-Whitney p = {mw:.3g}, "
         f"seeds within 2 A: GPU {g_near}/{len(seeds)}, oracle {c_near}/{len(seeds)}")
    assert ks > 1e-3
    assert mw > 1e-2
    assert abs(g_near - c_near) <= max(2, len(seeds) // 3)
    if preset == "detail":
        assert g_near == len(seeds) and c_near == len(seeds)
    ctx.close()
