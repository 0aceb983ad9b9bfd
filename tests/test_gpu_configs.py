"""The five benchmark configurations on the GPU, each checked against the
oracle where the oracle finishes in seconds and by size-independent
properties at full size.

  config 0  one 24-atom, 5-torsion ligand, 2,000-atom receptor, 22.5 A box
            (61^3), exhaustiveness 8, max_step 20, FP32 maps: in full against
            the oracle (maps, energies and gradients, BFGS minima, the search
            distribution over 16 seeds)
  config 1  1,024 ligands (10-40 atoms, 0-10 torsions), exhaustiveness 128,
            max_step 20, FP32 maps: the whole batch (every chain's reported
            energy is its pose's energy, bound poses lie in the box) and 32
            of its ligands against the oracle's search
  config 2  screening distribution (atoms N(28, 7) in [8, 60], torsions
            Poisson(5) in [0, 15]), 5,000-atom receptor, exhaustiveness 512,
            max_step 40, autotuned lanes, Vina and Vinardo: 256 ligands that
            include every ligand of more than 10 torsions (the 38-entry search
            vectors) and of more than 48 atoms among the first 2,000 of the
            set, 8 of them against the oracle's search
  config 3  AD4.2 at 81^3 (30 A box, 8,000-atom receptor with a cleared
            pocket), FP16 and FP32 maps, 4,096 ligands: maps against the FP64
            brute force on an aligned 21^3 sub-box, the whole batch docked with
            both precisions, and per-pose energies on random poses separated
            into kernel error (GPU FP32 vs the FP32 restatement on the same
            maps), FP32 rounding (FP32 vs FP64 restatement) and FP16 storage
            (FP16 vs FP32 maps)
  config 4  65,536 poses over 2,048 ligands through the torch.autograd
            binding, forward and backward: finite, the pose gradient equal to
            the change-space gradient of ud_score_grad (the reference's
            analytic derivative), and central differences (h = 1e-3) of the
            FP64 oracle on a sample
Seeds follow the benchmark's (receptor seed = 1000 x configuration, ligands
+1), so these are the benchmark's own inputs or subsets of them.
"""
import numpy as np
import pytest
from scipy import stats

from oracle.oracle import AD4_MAP_DSOLV, AD4_MAP_ELEC, Oracle
from unidock_tools import _lib, synth

pytestmark = pytest.mark.gpu
BOX = np.array([22.5, 22.5, 22.5])
LIGAND_TYPES = [0, 1, 2, 3, 8, 9, 10]


def _say(msg):
    print(msg, flush=True)


def _threads():
    import os
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _global_steps(ligs, max_step):
    heur = max(l.n_atoms + 10 * (6 + l.n_tors) for l in ligs)  # vina.h:192-204
    return min(int(70 * 3 * (50 + heur) / 2), max_step)


def _ctx(rec, ligs, precision, sf=_lib.UD_SF_VINA, box=BOX, types=None):
    ctx = _lib.Context(0)
    ctx.set_scoring(sf)
    npts = ctx.set_box(rec.cavity, box, 0.375)
    ctx.compute_maps(rec.xyz, rec.xs, types or LIGAND_TYPES, precision)
    ctx.set_ligands(synth.LigandBatch(list(ligs)).arrays())
    return ctx, npts


def _oracle_on_gpu_maps(ctx, rec, ligs, prec, sf=0, box=BOX, types=None):
    o = Oracle(prec, sf)
    o.set_grid(rec.cavity, box, 0.375)
    for t in types or LIGAND_TYPES:
        o.set_map(t, ctx.get_map(t).astype(np.float64))
    for l in ligs:
        o.add_ligand(l)
    return o


def _random_poses(rng, ligs, pl, lo, hi):
    P = np.zeros((len(pl), _lib.UD_POSE_FLOATS))
    P[:, :3] = lo + rng.random((len(pl), 3)) * (hi - lo)
    q = rng.normal(size=(len(pl), 4))
    P[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
    for i, li in enumerate(pl):
        nt = ligs[li].n_tors
        P[i, 7:7 + nt] = rng.uniform(-np.pi, np.pi, nt)
    return P


def _box_bounds(o):
    lo = o.geo[:3]
    return lo, lo + o.geo[6:9] * (np.array(o.n) - 1)


def _batch_properties(ctx, ligs, exh, P, e, ev, sample, rng, tag):
    """Every chain's result is finite, evaluated, and its reported energy is
    the energy of its pose with authentic caps.  Returns a sample of the
    chains below 1000 kcal/mol with their coordinates: an atom outside the box
    would carry the 10^6 kcal/mol/A penalty, so their atoms lie inside within
    10^-3 A.  (The synthetic receptors pack atoms at 3 A, so most best poses
    still clash.)"""
    n = len(ligs)
    assert P.shape == (n * exh, _lib.UD_POSE_FLOATS)
    assert np.all(np.isfinite(e)) and np.all(ev > 0), tag
    pl_all = np.repeat(np.arange(n, dtype=np.int32), exh)
    idx = rng.choice(len(e), min(sample, len(e)), replace=False)
    es, _ = ctx.score_grad(pl_all[idx], P[idx], (1000.0, 1000.0, 1000.0))
    np.testing.assert_allclose(es, e[idx], rtol=1e-4, atol=1e-3, err_msg=tag)
    bound = idx[e[idx] < 1000.0]
    xyz = ctx.pose_coords(pl_all[bound], P[bound])
    return pl_all, bound, xyz


def _search_parity(g_e, c_e, tag, ks_min=1e-3):
    ks = stats.ks_2samp(g_e, c_e).pvalue
    _say(f"[gpu] {tag}: GPU chain-best median {np.median(g_e):.3f}, oracle {np.median(c_e):.3f}, KS p = {ks:.3g}")
    assert ks > ks_min, tag


# ---------------------------------------------------------------- config 0
def test_config0_single_ligand_in_full():
    _say("[gpu] config 0: single ligand, 2,000-atom receptor, 61^3, exhaustiveness 8")
    rec = synth.make_receptor(2000, seed=1000)
    lig = synth.make_ligand(24, 5, seed=0)
    types = sorted({int(t) for t in lig.xs if t >= 0})
    ctx, npts = _ctx(rec, [lig], _lib.UD_GRID_F32, types=types)
    assert tuple(npts) == (61, 61, 61)
    # maps over the whole box against the FP64 and FP32 brute force
    o64 = Oracle(64, 0)
    o64.set_grid(rec.cavity, BOX, 0.375)
    o64.build_maps(rec.xyz, rec.xs, types)
    o32 = Oracle(32, 0)
    o32.set_grid(rec.cavity, BOX, 0.375)
    o32.build_maps(rec.xyz, rec.xs, types)
    for t in types:
        g = ctx.get_map(t).astype(np.float64)
        assert g.size == 61 ** 3
        for o, frac in ((o32, 2e-3), (o64, 5e-3)):
            r = o.get_map(t)
            bad = np.abs(g - r) > 1e-5 * (1 + np.abs(r))
            assert bad.mean() <= frac, (t, bad.mean())
    # energies and gradients on the GPU's maps
    o64m = _oracle_on_gpu_maps(ctx, rec, [lig], 64, types=types)
    o32m = _oracle_on_gpu_maps(ctx, rec, [lig], 32, types=types)
    lo, hi = _box_bounds(o64m)
    rng = np.random.default_rng(0)
    pl = np.zeros(64, dtype=np.int32)
    P = _random_poses(rng, [lig], pl, rec.cavity - 3.0, rec.cavity + 3.0)
    e, g = ctx.score_grad(pl, P)
    for i in range(len(pl)):
        e32, g32, _, _ = o32m.eval_deriv(0, P[i])
        e64, g64, c64, _ = o64m.eval_deriv(0, P[i])
        # 1e-5 relative to the FP32 restatement; at clash poses, where
        # |dE/dx| reaches 10^6 kcal/mol/A and one ulp of a coordinate moves
        # the energy by more than that, no farther from it than twice the
        # restatement's own distance from FP64
        assert abs(e[i] - e32) <= max(1e-5 * abs(e32) + 1e-4, 2.0 * abs(e32 - e64)), (i, e[i], e32, e64)
        if np.all(c64 >= lo) and np.all(c64 <= hi):
            assert e[i] == pytest.approx(e64, rel=1e-4, abs=1e-4)
        np.testing.assert_allclose(g[i, :11], g32, rtol=0, atol=1e-4 * max(1.0, np.abs(g32).max()))
    # BFGS minima from identical starts a small kick away from the FP64
    # oracle's minima: the GPU reaches the FP64 minimum as often as the FP32
    # restatement does, less 10 % (the receptor packs atoms at 3 A, so the
    # minima are clash-shaped and FP32 line-search branches flip)
    starts = []
    for p in P[:32]:
        m = o64m.bfgs(0, p, max_steps=300)[1]
        m[0:3] += rng.normal(scale=0.05, size=3)
        q = m[3:7] + rng.normal(scale=0.01, size=4)
        m[3:7] = q / np.linalg.norm(q)
        m[7:12] += rng.normal(scale=0.02, size=5)
        starts.append(m)
    starts = np.array(starts)
    Pg, eg, _ = ctx.optimize(pl[:32], starts, (1000.0, 1000.0, 1000.0))
    e64s = np.array([o64m.bfgs(0, x)[0] for x in starts])
    e32s = np.array([o32m.bfgs(0, x)[0] for x in starts])
    agree_gpu = int(np.sum(np.abs(eg - e64s) <= 0.01))
    agree_32 = int(np.sum(np.abs(e32s - e64s) <= 0.01))
    _say(f"[gpu] config 0 BFGS minima within 0.01 kcal/mol of FP64: GPU {agree_gpu}/32, FP32 restatement {agree_32}/32")
    assert agree_gpu >= agree_32 - 3
    # the search: exhaustiveness 8, max_step 20, 16 seeds
    gs = _global_steps([lig], 20)
    g_e, c_e, g_best, c_best = [], [], [], []
    for seed in range(1, 17):
        _, eg_, _ = ctx.dock(exhaustiveness=8, global_steps=gs, seed=seed)
        ec, _ = o32m.dock_cpu([0], 8, seed, gs, [(25 + lig.n_atoms) // 3], _threads())
        g_e.extend(eg_)
        c_e.extend(ec)
        g_best.append(eg_.min())
        c_best.append(ec.min())
    _search_parity(np.array(g_e), np.array(c_e), "config 0 search (16 seeds x 8 chains)")
    assert stats.mannwhitneyu(g_best, c_best).pvalue > 1e-2
    ctx.close()


# ---------------------------------------------------------------- config 1
def test_config1_batch_of_1024():
    _say("[gpu] config 1: 1,024 ligands, exhaustiveness 128, max_step 20, FP32 maps")
    rec = synth.make_receptor(2000, seed=1000)
    ligs = synth.make_ligand_set(1024, seed=1001, atoms=(10, 40), tors=(0, 10))
    ctx, npts = _ctx(rec, ligs, _lib.UD_GRID_F32)
    assert tuple(npts) == (61, 61, 61)
    gs = _global_steps(ligs, 20)
    exh = 128
    P, e, ev = ctx.dock(exhaustiveness=exh, global_steps=gs, seed=181129)
    rng = np.random.default_rng(1)
    pl_all, bound, xyz = _batch_properties(ctx, ligs, exh, P, e, ev, 4096, rng, "config 1")
    o32 = _oracle_on_gpu_maps(ctx, rec, ligs[:0], 32)
    lo, hi = _box_bounds(o32)
    for k, c in enumerate(bound):
        na = ligs[pl_all[c]].n_atoms
        assert np.all(xyz[k, :na] >= lo - 1e-3) and np.all(xyz[k, :na] <= hi + 1e-3), c
    _say(f"[gpu] config 1: {len(e)} chains, {np.mean(e < 0):.3f} below 0 and {len(bound)} of a 4,096-chain "
         f"sample below 1000 kcal/mol (atoms in the box), evaluations {ev.sum():.3e}")
    assert len(bound) > 1000
    # 32 ligands against the oracle's search on the same maps
    sel = np.linspace(0, len(ligs) - 1, 32).astype(int)
    for li in sel:
        o32.add_ligand(ligs[li])
    ec, _ = o32.dock_cpu(list(range(32)), exh, 181129, gs, [(25 + ligs[li].n_atoms) // 3 for li in sel], _threads())
    # chain ids differ between the subset and the batch, so compare distributions
    g_e = e.reshape(len(ligs), exh)[sel]
    _search_parity(g_e.ravel(), ec, "config 1 search (32 ligands x 128 chains)")
    # per-ligand medians, paired: the oracle against itself with six seeds
    # (181129, 7..11) gives |median d| up to 0.76 over the 15 seed pairs
    # (spread about 0.4), so 1.0 bounds a shift beyond the search's own
    # noise; the signed-rank test checks the pairs' symmetry about 0
    d = np.median(g_e, axis=1) - np.median(ec.reshape(32, exh), axis=1)
    assert abs(np.median(d)) < 1.0, d
    assert stats.wilcoxon(d).pvalue > 1e-3, d
    ctx.close()


# ---------------------------------------------------------------- config 2
@pytest.mark.parametrize("sf", [_lib.UD_SF_VINA, _lib.UD_SF_VINARDO])
def test_config2_screening_subset(sf):
    name = {_lib.UD_SF_VINA: "vina", _lib.UD_SF_VINARDO: "vinardo"}[sf]
    _say(f"[gpu] config 2 ({name}): screening distribution, exhaustiveness 512, max_step 40")
    rec = synth.make_receptor(5000, seed=2000)
    pool = synth.make_ligand_set(2000, seed=2001, dist="screening")  # the first 2,000 of the 10,000
    big = [i for i, l in enumerate(pool) if l.n_tors > 10 or l.n_atoms > 48]
    rest = [i for i in range(len(pool)) if i not in set(big)]
    rng = np.random.default_rng(2)
    pick = sorted(big + list(rng.choice(rest, 256 - len(big), replace=False)))
    ligs = [pool[i] for i in pick]
    assert max(l.n_tors for l in ligs) > 10 and max(l.n_atoms for l in ligs) > 48
    types = sorted({int(t) for l in ligs for t in l.xs if t >= 0})
    ctx, npts = _ctx(rec, ligs, _lib.UD_GRID_F16, sf=sf, types=types)
    exh, gs = 512, _global_steps(ligs, 40)
    tuned = ctx.autotune(exhaustiveness=exh, global_steps=2, sample_ligands=256)
    P, e, ev = ctx.dock(exhaustiveness=exh, global_steps=gs, seed=181129)
    launches = ctx.last_launches()
    _say(f"[gpu] config 2 ({name}): lanes {list(tuned)}, launches {[(a['atom_cap'], a['lanes']) for a in launches]}")
    assert any(a["atom_cap"] == 64 for a in launches)
    _batch_properties(ctx, ligs, exh, P, e, ev, 4096, rng, f"config 2 {name}")
    # 6 ligands against the oracle: the two with most torsions, the two
    # largest, and the two smallest
    order = sorted(range(len(ligs)), key=lambda i: (ligs[i].n_tors, ligs[i].n_atoms))
    sel = sorted(set(order[-2:] + sorted(range(len(ligs)), key=lambda i: ligs[i].n_atoms)[-2:] + order[:2]))
    o = _oracle_on_gpu_maps(ctx, rec, [ligs[i] for i in sel], 32, sf=sf, types=types)
    ec, _ = o.dock_cpu(list(range(len(sel))), exh, 181129, gs, [(25 + ligs[i].n_atoms) // 3 for i in sel],
                       _threads())
    g_e = e.reshape(len(ligs), exh)[sel]
    _say(f"[gpu] config 2 ({name}) oracle subset: atoms {[ligs[i].n_atoms for i in sel]}, "
         f"torsions {[ligs[i].n_tors for i in sel]}")
    _search_parity(g_e.ravel(), ec, f"config 2 {name} search ({len(sel)} ligands x 512 chains)")
    ctx.close()


# ---------------------------------------------------------------- config 3
def test_config3_ad4_maps_81_cubed():
    _say("[gpu] config 3: AD4.2, 81^3 maps, FP16 and FP32, 4,096 ligands")
    rec = synth.make_receptor(8000, seed=3000)
    keep = np.linalg.norm(rec.xyz - rec.cavity, axis=1) > 7.0
    rec = synth.Receptor(xyz=rec.xyz[keep], xs=rec.xs[keep], edge=rec.edge, cavity=rec.cavity)
    rec_ad, rec_q = synth.ad4_receptor(rec, seed=3003)
    ligs = [synth.ad4_ligand(l, seed=3004 + i)
            for i, l in enumerate(synth.make_ligand_set(4096, seed=3001, dist="screening"))]
    types = sorted({_lib.ad4_grid_type(int(t)) for l in ligs for t in l.ad} - {-1})
    box = np.array([30.0, 30.0, 30.0])
    all_maps = types + [AD4_MAP_ELEC, AD4_MAP_DSOLV]
    ctxs = {}
    for name, prec in (("fp32", _lib.UD_GRID_F32), ("fp16", _lib.UD_GRID_F16)):
        ctx = _lib.Context(0)
        ctx.set_scoring(_lib.UD_SF_AD4)
        npts = ctx.set_box(rec.cavity, box, 0.375)
        assert tuple(npts) == (81, 81, 81)
        ctx.compute_maps_ad4(rec.xyz, rec_ad, rec_q, types, prec)
        ctx.set_ligands(synth.LigandBatch(ligs).arrays())
        ctxs[name] = ctx
    # maps against the FP64 brute force on the 21^3 sub-box at the cavity,
    # whose points are full-box points 30..50 along each axis
    osub = Oracle(64, 2)
    n_sub, _ = osub.set_grid(rec.cavity, np.array([7.5, 7.5, 7.5]), 0.375)
    assert tuple(n_sub) == (21, 21, 21)
    osub.build_maps_ad4(rec.xyz, rec_ad, rec_q, types)
    for t in all_maps:
        full = ctxs["fp32"].get_map(t).reshape(81, 81, 81)[30:51, 30:51, 30:51].ravel().astype(np.float64)
        r = osub.get_map(t)
        soft = np.abs(r) < 1000
        bad = np.abs(full[soft] - r[soft]) > 1e-4 * (1 + np.abs(r[soft]))
        assert bad.mean() <= 1e-3, (t, bad.mean())
        np.testing.assert_allclose(full[~soft], r[~soft], rtol=1e-4, err_msg=f"map {t}")
        h = ctxs["fp16"].get_map(t).reshape(81, 81, 81)[30:51, 30:51, 30:51].ravel().astype(np.float64)
        scale = np.abs(ctxs["fp32"].get_map(t)).max()
        assert np.abs(h - full).max() <= scale * 2.0 ** -10, t  # half precision with a power-of-two map scale
    # the whole batch with both map precisions
    exh, gs = 128, _global_steps(ligs, 20)
    rng = np.random.default_rng(3)
    res = {}
    for name, ctx in ctxs.items():
        P, e, ev = ctx.dock(exhaustiveness=exh, global_steps=gs, seed=181129)
        _batch_properties(ctx, ligs, exh, P, e, ev, 4096, rng, f"config 3 {name}")
        res[name] = (P, e)
        _say(f"[gpu] config 3 {name}: {len(e)} chains, {np.mean(e < 0):.3f} bound")
    # per-pose energies on 20,000 random poses over 32 ligands, on the GPU's FP32 maps
    sub = rng.choice(len(ligs), 32, replace=False)
    o32 = Oracle(32, 2)
    o64 = Oracle(64, 2)
    for o in (o32, o64):
        o.set_grid(rec.cavity, box, 0.375)
        for t in all_maps:
            o.set_map(t, ctxs["fp32"].get_map(t).astype(np.float64))
        for k in sub:
            o.add_ligand(ligs[int(k)])
    lo, hi = _box_bounds(o64)
    pls = rng.integers(0, len(sub), 20000).astype(np.int32)
    pl = sub[pls].astype(np.int32)
    P = _random_poses(rng, ligs, pl, lo, hi)
    e32 = ctxs["fp32"].score_grad(pl, P)[0].astype(np.float64)
    e16 = ctxs["fp16"].score_grad(pl, P)[0].astype(np.float64)
    r32 = o32.eval_batch(pls, P, (1000.0, 1000.0, 1000.0), _threads())
    r64 = o64.eval_batch(pls, P, (1000.0, 1000.0, 1000.0), _threads())
    inside = np.abs(r64) < 1e4  # no out-of-box penalty
    rel = lambda a, b: np.abs(a - b) / np.maximum(1.0, np.abs(b))
    kern = rel(e32, r32)[inside]
    rnd = rel(r32, r64)[inside]
    fp16 = rel(e16, e32)[inside]
    _say(f"[gpu] config 3 random poses ({int(inside.sum())} in the box): kernel (GPU FP32 vs FP32 restatement) "
         f"max {kern.max():.2e}, FP32 rounding (FP32 vs FP64 restatement) max {rnd.max():.2e}, "
         f"FP16 maps vs FP32 maps max {fp16.max():.2e}")
    assert np.percentile(kern, 99.9) <= 1e-5 and kern.max() <= 1e-3
    # the FP16 search's bound chain-best poses of the 32 ligands: FP16 vs FP64
    # on the same (FP32) maps
    Pd, ed = res["fp16"]
    ch = np.concatenate([np.arange(int(k) * exh, (int(k) + 1) * exh) for k in sub])
    ch = ch[ed[ch] < 0]
    pos = {int(k): j for j, k in enumerate(sub)}  # oracle index of each subset ligand
    pl_d = np.array([pos[int(l)] for l in ch // exh], dtype=np.int32)
    d64 = o64.eval_batch(pl_d, Pd[ch].astype(np.float64), (1000.0, 1000.0, 1000.0), _threads())
    d16 = ctxs["fp16"].score_grad((ch // exh).astype(np.int32), Pd[ch])[0]
    err = np.abs(d16 - d64)
    _say(f"[gpu] config 3 docked FP16 poses vs FP64: {len(err)} bound chains, max {err.max():.4f}, "
         f"mean {err.mean():.5f} kcal/mol")
    assert len(err) > 0 and err.max() <= 0.1
    for ctx in ctxs.values():
        ctx.close()


# ---------------------------------------------------------------- config 4
def test_config4_energy_grad_65536_poses():
    import torch
    from unidock_tools.torch_energy import DockingEnergy
    _say("[gpu] config 4: 65,536 poses over 2,048 ligands, forward + backward")
    rec = synth.make_receptor(5000, seed=4000)
    ligs = synth.make_ligand_set(2048, seed=4001, dist="screening")
    model = DockingEnergy(rec.xyz, rec.xs, ligs, rec.cavity, BOX)
    o = Oracle(64, 0)
    o.set_grid(rec.cavity, BOX, 0.375)
    lo, hi = _box_bounds(o)
    rng = np.random.default_rng(4002)
    pl = np.repeat(np.arange(len(ligs), dtype=np.int32), 32)
    P = _random_poses(rng, ligs, pl, lo, hi)
    poses = torch.tensor(P, dtype=torch.float32, device="cuda", requires_grad=True)
    w = torch.tensor(_lib.VINA_WEIGHTS, dtype=torch.float64, requires_grad=True)
    e = model(poses, pl, w)
    e.sum().backward()
    torch.cuda.synchronize()
    E = e.detach().cpu().numpy()
    G = poses.grad.detach().cpu().numpy()
    W = w.grad.detach().cpu().numpy()
    assert E.shape == (65536,) and np.all(np.isfinite(E)) and np.all(np.isfinite(G)) and np.all(np.isfinite(W))
    # the change-space gradient of ud_score_grad at the normalised pose: the
    # position part is the force, the torsion part the torsion derivatives
    Pn = P.copy()
    Pn[:, 3:7] /= np.linalg.norm(Pn[:, 3:7], axis=1, keepdims=True)
    idx = rng.choice(len(pl), 4096, replace=False)
    es, gs = model.ctx.score_grad(pl[idx], Pn[idx], (1000.0, 1000.0, 1000.0))
    # out-of-box poses carry the 10^6 / A penalty, which amplifies the
    # rounding of the two normalisations of q (kernel vs host)
    inb = np.abs(es) < 1e4
    np.testing.assert_allclose(E[idx][inb], es[inb], rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(E[idx][~inb], es[~inb], rtol=1e-3)
    # gradients: the position and torsion parts of the backward pass equal
    # the kernel's change-space gradient.  Compared on poses whose
    # orientation is one of the 24 unit quaternions with components in
    # {0, +-1/2, +-1}, which both paths normalise exactly, so that both
    # kernels place the atoms identically (an ulp of q moves coordinates,
    # and the curvature of clash poses turns that into gradient differences
    # of 10^-4 relative)
    units = [np.roll([s0, 0.0, 0.0, 0.0], j) for j in range(4) for s0 in (1.0, -1.0)]
    units += [0.5 * np.array([1.0, a, b, c]) * s0 for a in (1, -1) for b in (1, -1) for c in (1, -1)
              for s0 in (1.0, -1.0)]
    Pu = P[idx].copy()
    Pu[:, 3:7] = np.array(units)[rng.integers(0, len(units), len(idx))]
    eu, gu = model.ctx.score_grad(pl[idx], Pu, (1000.0, 1000.0, 1000.0))
    pu = torch.tensor(Pu, dtype=torch.float32, device="cuda", requires_grad=True)
    Eu = model(pu, pl[idx], w.detach())
    Eu.sum().backward()
    torch.cuda.synchronize()
    Gu = pu.grad.detach().cpu().numpy()
    np.testing.assert_allclose(Eu.detach().cpu().numpy(), eu, rtol=1e-6, atol=1e-5)
    for k, i in enumerate(idx):
        nt = ligs[pl[i]].n_tors
        sc = max(1.0, np.abs(gu[k, :6 + nt]).max())
        np.testing.assert_allclose(Gu[k, :3], gu[k, :3], atol=1e-6 * sc)
        np.testing.assert_allclose(Gu[k, 7:7 + nt], gu[k, 6:6 + nt], atol=1e-6 * sc)
    # central differences of the FP64 oracle (h = 1e-3) on poses inside the box
    o.build_maps(rec.xyz, rec.xs, sorted({int(t) for l in ligs for t in l.xs if t >= 0}))
    for l in ligs:
        o.add_ligand(l)
    rel = []
    for i in rng.choice(len(pl), 48, replace=False):
        li = int(pl[i])
        nt = ligs[li].n_tors
        x = P[i].copy()
        fd = np.zeros(7 + nt)
        for j in range(7 + nt):
            a, b = x.copy(), x.copy()
            a[j] += 1e-3
            b[j] -= 1e-3
            a[3:7] /= np.linalg.norm(a[3:7])
            b[3:7] /= np.linalg.norm(b[3:7])
            fd[j] = (o.eval_deriv(li, a)[0] - o.eval_deriv(li, b)[0]) / 2e-3
        rel.append(np.abs(G[i, :7 + nt] - fd).max() / max(1.0, np.abs(fd).max()))
    rel = np.array(rel)
    _say(f"[gpu] config 4 finite differences: median {np.median(rel):.2e}, 75th percentile {np.percentile(rel, 75):.2e}")
    # stencils that cross a map-cell face or the box edge differ by the kink
    assert np.median(rel) <= 1e-3 and np.percentile(rel, 75) <= 2e-2
