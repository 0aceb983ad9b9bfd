"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Primary target: the FP32 restatement of the reference's CPU path (the
reference computes in FP32 throughout, common.h:45).  Secondary: the FP64
oracle, with bounds that account for what FP32 itself cannot resolve.

Tolerances
  maps         |d| <= 1e-5 (1 + |v|) except rare table-bin flips (a receptor
               atom with r^2 on a 1/32 A^2 bin edge), each at most one step
  energy       1e-5 relative vs FP32 restatement; vs FP64 1e-4 relative for
               poses inside the box; out-of-box poses carry the 1e6 slope
               penalty, whose FP32 rounding is ~slope * ulp(coordinate)
  gradient     1e-4 * max(1, |g|_inf) per component vs FP32 restatement,
               3e-3 * max(1, |g|_inf) vs FP64
  BFGS minima  0.01 kcal/mol from identical near-minimum starts
  MC           identical random starts; per-ligand best energies and the
               pooled best-score distribution (KS test) vs the oracle
"""
import ctypes

import numpy as np
import pytest

from helpers import LIGAND_TYPES, box_of, oracle_scene, random_poses, small_ligands, small_receptor
from unidock_tools import _lib, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def scene():
    xyz, xs, cav = small_receptor()
    ligs = small_ligands()
    _say("[gpu] scene: creating context")
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINA)
    c, s = box_of(cav)
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps(xyz, xs, LIGAND_TYPES, _lib.UD_GRID_F32)
    ctx.set_ligands(synth.LigandBatch(list(ligs)).arrays())
    o64, _, _ = oracle_scene(64)
    o32, _, _ = oracle_scene(32)
    lo = o64.geo[:3]
    hi = lo + o64.geo[6:9] * (np.array(o64.n) - 1)
    # FP64 oracle on the GPU's own (FP32-built) maps: isolates evaluation from
    # the table-bin flips of map construction
    o64m, _, _ = oracle_scene(64, with_maps=False)
    for t in LIGAND_TYPES:
        o64m.set_map(t, ctx.get_map(t).astype(np.float64))
    yield dict(ctx=ctx, o64=o64, o32=o32, o64m=o64m, ligs=ligs, cav=cav, xyz=xyz, xs=xs, lo=lo, hi=hi)
    ctx.close()


def _poses(scene, n_per=6, seed=0, spread=3.0):
    rng = np.random.default_rng(seed)
    pl, P = [], []
    for li, lig in enumerate(scene["ligs"]):
        for p in random_poses(rng, lig, scene["cav"], spread, n_per):
            pl.append(li)
            P.append(p)
    return np.array(pl, dtype=np.int32), np.array(P)


def _in_box(scene, coords):
    return bool(np.all(coords >= scene["lo"]) and np.all(coords <= scene["hi"]))


def test_maps_match_oracle(scene):
    _say("[gpu] maps vs oracle")
    ctx, o64, o32 = scene["ctx"], scene["o64"], scene["o32"]
    for t in LIGAND_TYPES:
        g = ctx.get_map(t).astype(np.float64)
        for o, frac in ((o32, 2e-3), (o64, 5e-3)):
            r = o.get_map(t)
            bad = np.abs(g - r) > 1e-5 * (1 + np.abs(r))
            assert bad.mean() <= frac, (t, bad.mean())
            assert np.abs(g - r).max() < 0.5, (t, np.abs(g - r).max())


def test_score_grad_matches_oracle(scene):
    _say("[gpu] score_grad vs oracle")
    ctx, o64, o32 = scene["ctx"], scene["o64"], scene["o32"]
    pl, P = _poses(scene)
    n_in = 0
    for v in ((1000.0, 1000.0, 1000.0), (10.0, 10.0, 10.0)):
        e, g = ctx.score_grad(pl, P, v)
        for i in range(len(pl)):
            li = int(pl[i])
            nt = scene["ligs"][li].n_tors
            e64, g64, c64, _ = o64.eval_deriv(li, P[i], v)
            e32, g32, _, _ = o32.eval_deriv(li, P[i], v)
            assert e[i] == pytest.approx(e32, rel=1e-5, abs=1e-4), (i, e[i], e32)
            if _in_box(scene, c64):
                n_in += 1
                assert e[i] == pytest.approx(e64, rel=1e-4, abs=1e-4), (i, e[i], e64)
            else:
                assert abs(e[i] - e64) <= 1e-4 * abs(e64) + 1e6 * 4e-6 * len(c64), (i, e[i], e64)
            sc = max(1.0, np.abs(g32).max())
            np.testing.assert_allclose(g[i, :6 + nt], g32, rtol=0, atol=1e-4 * sc)
            np.testing.assert_allclose(g[i, :6 + nt], g64, rtol=0, atol=3e-3 * sc)
            assert np.all(g[i, 6 + nt:] == 0)
    assert n_in >= 10


def test_score_grad_relaxed_poses(scene):
    """Parity away from clashes: oracle-relaxed poses inside the box."""
    _say("[gpu] score_grad on relaxed poses")
    ctx, o64, o32 = scene["ctx"], scene["o64"], scene["o32"]
    pl, P = _poses(scene, n_per=4, seed=1, spread=1.5)
    R = np.array([o64.bfgs(int(l), p)[1] for l, p in zip(pl, P)])
    e, g = ctx.score_grad(pl, R)
    for i in range(len(pl)):
        e64, g64, c64, _ = o64.eval_deriv(int(pl[i]), R[i])
        e32, g32, _, _ = o32.eval_deriv(int(pl[i]), R[i])
        nt = scene["ligs"][int(pl[i])].n_tors
        assert e[i] == pytest.approx(e32, rel=1e-5, abs=1e-4)
        if _in_box(scene, c64):
            assert e[i] == pytest.approx(e64, rel=1e-4, abs=2e-4)
        sc = max(1.0, np.abs(g32).max())
        np.testing.assert_allclose(g[i, :6 + nt], g32, rtol=0, atol=1e-4 * sc)
        # FP64 on the same maps: evaluation parity without map-construction bin flips
        e64m, g64m, _, _ = scene["o64m"].eval_deriv(int(pl[i]), R[i])
        assert e[i] == pytest.approx(e64m, rel=1e-4, abs=2e-4)
        np.testing.assert_allclose(g[i, :6 + nt], g64m, rtol=0, atol=1e-3 * sc)
        # FP64 on FP64-built maps: a map point can differ by one table step
        np.testing.assert_allclose(g[i, :6 + nt], g64, rtol=1e-6, atol=0.5)


def _near_minimum_starts(scene, n_per, seed):
    """Starts a small kick away from oracle minima: both BFGS runs descend
    into the same basin unless a line-search branch flips."""
    o64 = scene["o64"]
    rng = np.random.default_rng(seed)
    pl, P = _poses(scene, n_per=n_per, seed=seed, spread=1.5)
    S = []
    for l, p in zip(pl, P):
        m = o64.bfgs(int(l), p, max_steps=300)[1]
        nt = scene["ligs"][int(l)].n_tors
        m[0:3] += rng.normal(scale=0.05, size=3)
        q = m[3:7] + rng.normal(scale=0.01, size=4)
        m[3:7] = q / np.linalg.norm(q)
        m[7:7 + nt] += rng.normal(scale=0.02, size=nt)
        S.append(m)
    return pl, np.array(S)


def test_bfgs_reaches_same_minima(scene):
    """From identical starts near a minimum, with the reference's iteration
    count ((25 + N) / 3), GPU and oracle BFGS end in the same minimum."""
    _say("[gpu] optimize vs oracle bfgs")
    ctx, o64, o32 = scene["ctx"], scene["o64m"], scene["o32"]
    pl, P = _near_minimum_starts(scene, 5, 2)
    Pg, eg, evg = ctx.optimize(pl, P, (1000.0, 1000.0, 1000.0))
    agree32 = agree64 = 0
    for i in range(len(pl)):
        li = int(pl[i])
        e32 = o32.bfgs(li, P[i])[0]
        e64 = o64.bfgs(li, P[i])[0]
        agree32 += abs(eg[i] - e32) <= 0.01
        agree64 += abs(eg[i] - e64) <= 0.01
        e_chk = o32.eval_deriv(li, Pg[i].astype(np.float64))[0]
        assert eg[i] == pytest.approx(e_chk, rel=1e-4, abs=1e-3)
        e_start = o32.eval_deriv(li, P[i])[0]
        assert eg[i] <= e_start + 1e-3 + 1e-5 * abs(e_start)
        assert evg[i] >= 2
    assert agree32 >= 0.9 * len(pl), (agree32, len(pl))
    assert agree64 >= 0.9 * len(pl), (agree64, len(pl))


def test_bfgs_from_random_starts_descends(scene):
    _say("[gpu] optimize from random starts")
    ctx, o32 = scene["ctx"], scene["o32"]
    pl, P = _poses(scene, n_per=5, seed=6)
    Pg, eg, evg = ctx.optimize(pl, P, (10.0, 10.0, 10.0))
    for i in range(len(pl)):
        e_start = o32.eval_deriv(int(pl[i]), P[i], (10.0, 10.0, 10.0))[0]
        assert eg[i] <= e_start + 1e-3 + 1e-5 * abs(e_start)
        e_chk = o32.eval_deriv(int(pl[i]), Pg[i].astype(np.float64), (10.0, 10.0, 10.0))[0]
        assert eg[i] == pytest.approx(e_chk, rel=1e-4, abs=1e-3)


def test_random_starts_match_oracle(scene):
    _say("[gpu] dock with zero steps = random starts")
    ctx, o64 = scene["ctx"], scene["o64"]
    exh = 8
    P, e, ev = ctx.dock(exhaustiveness=exh, global_steps=0, seed=77)
    for li, lig in enumerate(scene["ligs"]):
        for c in range(exh):
            ref = o64.random_start(li, 77, li * exh + c)
            got = P[li * exh + c]
            np.testing.assert_allclose(got[:7 + lig.n_tors], ref[:7 + lig.n_tors], atol=2e-5)
    assert np.all(np.isinf(e))


def test_dock_is_deterministic_and_consistent(scene):
    _say("[gpu] dock determinism")
    ctx, o64 = scene["ctx"], scene["o64"]
    for lanes in (64, 16):
        P1, e1, ev1 = ctx.dock(exhaustiveness=8, global_steps=3, seed=5, lanes=lanes)
        P2, e2, ev2 = ctx.dock(exhaustiveness=8, global_steps=3, seed=5, lanes=lanes)
        assert np.array_equal(P1, P2) and np.array_equal(e1, e2) and np.array_equal(ev1, ev2)
        assert np.all(ev1 > 0) and np.all(np.isfinite(e1))
        pl = np.repeat(np.arange(len(scene["ligs"])), 8).astype(np.int32)
        # the reported best energy is the authentic-cap energy of the reported pose
        eg, _ = ctx.score_grad(pl, P1)
        np.testing.assert_allclose(eg, e1, rtol=1e-4, atol=1e-3)


def test_chain_refill_gives_identical_chains(scene):
    """A chain's trajectory depends only on its index and the seed, not on
    which lane group runs it: workgroups whose groups run several chains in
    turn (refilled as chains end) report exactly the one-chain-per-group
    results."""
    _say("[gpu] dock chain refill")
    ctx = scene["ctx"]
    for lanes in (32, 16):
        P1, e1, ev1 = ctx.dock(exhaustiveness=12, global_steps=3, seed=7, lanes=lanes)
        for cpg in (3, 8):
            P2, e2, ev2 = ctx.dock(exhaustiveness=12, global_steps=3, seed=7, lanes=lanes, chains_per_group=cpg)
            assert np.array_equal(P1, P2) and np.array_equal(e1, e2) and np.array_equal(ev1, ev2), (lanes, cpg)
    p = ctx._params(8, 3, 7, 0, 0, 65)
    assert ctx.L.ud_dock(ctx.h, ctypes.byref(p), None, None, None) == _lib.UD_ERR_INVALID


def test_dock_lane_widths_agree(scene):
    """Every lane width runs the same chains from the same random stream; the
    FP32 sums differ in order, so trajectories may part after a line-search
    branch flips.  Per-ligand best energies agree."""
    from scipy.stats import ks_2samp
    _say("[gpu] dock lane widths")
    ctx = scene["ctx"]
    exh = 32
    res = {}
    for lanes in (8, 16, 32, 64):
        P, e, ev = ctx.dock(exhaustiveness=exh, global_steps=6, seed=9, lanes=lanes)
        res[lanes] = e.astype(np.float64)
    base = res[64]
    for lanes in (8, 16, 32):
        zg, zb = [], []
        for li in range(len(scene["ligs"])):
            a, b = res[lanes][li * exh:(li + 1) * exh], base[li * exh:(li + 1) * exh]
            mu, sd = np.median(b), max(np.std(b), 0.1)
            zg += list((a - mu) / sd)
            zb += list((b - mu) / sd)
        assert ks_2samp(zg, zb).pvalue > 0.01, lanes


@pytest.mark.parametrize("lanes", [8, 16, 32, 64])
def test_every_lane_width_matches_oracle(scene, lanes):
    """Energy/gradient and converged BFGS minima at each lane width."""
    _say(f"[gpu] lanes={lanes} score + optimize")
    ctx, o32 = scene["ctx"], scene["o32"]
    ctx.set_lanes(lanes)
    try:
        pl, P = _poses(scene, n_per=3, seed=40 + lanes)
        e, g = ctx.score_grad(pl, P, (10.0, 10.0, 10.0))
        for i in range(len(pl)):
            e32, g32, _, _ = o32.eval_deriv(int(pl[i]), P[i], (10.0, 10.0, 10.0))
            nt = scene["ligs"][int(pl[i])].n_tors
            assert e[i] == pytest.approx(e32, rel=1e-5, abs=1e-4)
            np.testing.assert_allclose(g[i, :6 + nt], g32, rtol=0, atol=1e-4 * max(1.0, np.abs(g32).max()))
        pl, S = _near_minimum_starts(scene, 2, 50 + lanes)
        Pg, eg, _ = ctx.optimize(pl, S, (1000.0, 1000.0, 1000.0))
        agree = sum(abs(eg[i] - o32.bfgs(int(pl[i]), S[i])[0]) <= 0.01 for i in range(len(pl)))
        assert agree >= 0.9 * len(pl), (agree, len(pl))
        a, b = ctx.score_terms(pl, S)
        C = ctx.pose_coords(pl, S)
        for i in range(len(pl)):
            ra, rb = o32.score_terms(int(pl[i]), S[i])
            _, _, rc, _ = o32.eval_deriv(int(pl[i]), S[i])
            slack = 0.0 if _in_box(scene, rc) else 1e6 * 4e-6 * len(rc)  # out-of-box slope penalty
            assert abs(a[i] - ra) <= 1e-5 * abs(ra) + 1e-4 + slack
            np.testing.assert_allclose(C[i, :len(rc)], rc, atol=1e-4)
    finally:
        ctx.set_lanes(0)


def test_dock_matches_oracle_chains(scene):
    """Seeded global search vs the oracle's MC chains on the same random
    stream (FP64).  Chains from random starts are chaotic in FP32, so the
    comparison is distributional: per-ligand best energies and a two-sample
    KS test on per-ligand standardised best-of-chain energies."""
    from scipy.stats import ks_2samp
    _say("[gpu] dock vs oracle chains")
    ctx, o64 = scene["ctx"], scene["o64"]
    exh, steps = 16, 8
    P, e, ev = ctx.dock(exhaustiveness=exh, global_steps=steps, seed=22)
    ref = np.array([o64.mc_chain(li, 22, li * exh + c, steps)[0] for li in range(len(scene["ligs"]))
                    for c in range(exh)])
    zg, zr = [], []
    for li in range(len(scene["ligs"])):
        g = e[li * exh:(li + 1) * exh].astype(np.float64)
        r = ref[li * exh:(li + 1) * exh]
        mu, sd = np.median(r), max(np.std(r), 0.1)
        zg += list((g - mu) / sd)
        zr += list((r - mu) / sd)
    # pooled standardised best energies: same location and same distribution
    assert abs(np.mean(zg) - np.mean(zr)) < 0.5
    assert ks_2samp(zg, zr).pvalue > 0.01


def test_fp16_maps_error_bound(scene):
    _say("[gpu] fp16 maps")
    xyz, xs, cav = scene["xyz"], scene["xs"], scene["cav"]
    ctx16 = _lib.Context(0)
    ctx16.set_scoring(_lib.UD_SF_VINA)
    c, s = box_of(cav)
    ctx16.set_box(c, s, 0.375)
    ctx16.compute_maps(xyz, xs, LIGAND_TYPES, _lib.UD_GRID_F16)
    ctx16.set_ligands(synth.LigandBatch(list(scene["ligs"])).arrays())
    o64 = scene["o64"]
    # stored maps: half-precision rounding of the FP32 maps (same table bins)
    for t in LIGAND_TYPES:
        m32 = scene["ctx"].get_map(t)
        np.testing.assert_allclose(ctx16.get_map(t), m32, rtol=2 ** -11, atol=1e-6 * np.abs(m32).max())
    pl, P = _poses(scene, n_per=4, seed=3, spread=1.5)
    R = np.array([o64.bfgs(int(l), p)[1] for l, p in zip(pl, P)])
    e16, _ = ctx16.score_grad(pl, R)
    for i in range(len(pl)):
        e64, _, c64, _ = o64.eval_deriv(int(pl[i]), R[i])
        if _in_box(scene, c64):
            assert abs(e16[i] - e64) < 0.1, (i, e16[i], e64)
    ctx16.close()


def test_score_terms_and_coords(scene):
    """The rescoring kernel interpolates and sums in FP64: on the GPU's own
    maps its inter term matches the FP64 oracle to the FP32 coordinates'
    resolution."""
    _say("[gpu] score terms + coords")
    ctx, o64, o32, o64m = scene["ctx"], scene["o64"], scene["o32"], scene["o64m"]
    pl, P = _poses(scene, n_per=3, seed=4)
    a, b = ctx.score_terms(pl, P)
    C = ctx.pose_coords(pl, P)
    for i in range(len(pl)):
        ra, rb = o32.score_terms(int(pl[i]), P[i])
        assert a[i] == pytest.approx(ra, rel=1e-5, abs=1e-4)
        if _in_box(scene, o64.eval_deriv(int(pl[i]), P[i])[2]):
            assert a[i] == pytest.approx(o64m.score_terms(int(pl[i]), P[i])[0], rel=2e-5, abs=2e-5)
        # intra uses the piecewise-constant table: a bin flip is one table step
        assert b[i] == pytest.approx(rb, rel=1e-4, abs=1e-4) or abs(b[i] - rb) < 0.05
        _, _, rc, _ = o64.eval_deriv(int(pl[i]), P[i])
        np.testing.assert_allclose(C[i, :len(rc)], rc, atol=1e-4)


def test_capacity_and_state_errors(scene):
    _say("[gpu] error paths")
    ctx = _lib.Context(0)
    with pytest.raises(_lib.UdError) as ei:
        ctx.dock(exhaustiveness=4, global_steps=1)
    assert ei.value.code == _lib.UD_ERR_STATE
    big = synth.make_ligand(24, 5, seed=1)
    arr = synth.LigandBatch([big]).arrays()
    arr["atom_offset"] = np.array([0, _lib.UD_MAX_ATOMS + 1], dtype=np.int32)
    with pytest.raises(_lib.UdError) as ei:
        ctx.set_ligands(arr)
    assert ei.value.code == _lib.UD_ERR_CAPACITY
    ctx.close()


def test_autotune_and_bucket_lanes(scene):
    """The autotuner stores one supported lane count and LDS layout (per-pair
    force buffer or not) per occupied size bucket; docking with the tuned
    setting equals docking with the same lanes set by hand, and the
    per-launch timings cover the whole dock."""
    _say("[gpu] autotune")
    ctx = scene["ctx"]
    nat = [l.n_atoms for l in scene["ligs"]]
    occupied = {min(b for b in _lib.BUCKET_ATOMS if n <= b) for n in nat}
    tuned = ctx.autotune(exhaustiveness=8, global_steps=1, sample_ligands=4)
    lanes_l, pb_l = ctx.bucket_layout()
    assert list(lanes_l) == list(tuned)
    for cap, lanes, pb in zip(_lib.BUCKET_ATOMS, tuned, pb_l):
        assert (lanes in (8, 16, 32, 64) and pb in (0, 1)) if cap in occupied else (lanes == 0 and pb == -1)
    with pytest.raises(_lib.UdError) as ei:
        ctx.fetch_dock()
    assert ei.value.code == _lib.UD_ERR_STATE
    P1, e1, ev1 = ctx.dock(exhaustiveness=8, global_steps=3, seed=4)
    total_ms, n = ctx.last_kernel_time()
    li = ctx.last_launches()
    assert n == len(li) == len(occupied)
    assert [a["atom_cap"] for a in li] == sorted(occupied, reverse=True)
    for a in li:
        b = _lib.BUCKET_ATOMS.index(a["atom_cap"])
        assert a["lanes"] == tuned[b] and a["pair_buffer"] == pb_l[b] and a["ms"] > 0 and a["lds_bytes"] > 0
    # bucket launches overlap on their own streams: each lies inside the dock's span
    assert max(a["ms"] for a in li) <= total_ms * (1 + 1e-3) + 1e-2
    assert total_ms <= sum(a["ms"] for a in li) * (1 + 1e-3) + 0.5
    ctx.set_bucket_lanes([0] * _lib.UD_N_BUCKETS)
    ctx.set_bucket_lanes(list(tuned))
    P2, e2, ev2 = ctx.dock(exhaustiveness=8, global_steps=3, seed=4)
    assert np.array_equal(P1, P2) and np.array_equal(e1, e2) and np.array_equal(ev1, ev2)
    with pytest.raises(_lib.UdError) as ei:
        ctx.set_bucket_lanes([12, 0, 0, 0, 0, 0])
    assert ei.value.code == _lib.UD_ERR_INVALID
    ctx.set_bucket_lanes([0] * _lib.UD_N_BUCKETS)
    # both LDS layouts on demand: the per-atom pair loops use less LDS and
    # search alike (pooled chain-best energies, KS test)
    res = {}
    for pb in (1, 0):
        ctx.set_bucket_pair_buffer([pb] * _lib.UD_N_BUCKETS)
        _, e, _ = ctx.dock(exhaustiveness=32, global_steps=4, seed=8)
        li = ctx.last_launches()
        assert all(a["pair_buffer"] == pb for a in li)
        res[pb] = (e, sum(a["lds_bytes"] for a in li))
    assert res[0][1] < res[1][1]
    from scipy.stats import ks_2samp
    assert ks_2samp(res[0][0], res[1][0]).pvalue > 1e-3
    with pytest.raises(_lib.UdError) as ei:
        ctx.set_bucket_pair_buffer([2] * _lib.UD_N_BUCKETS)
    assert ei.value.code == _lib.UD_ERR_INVALID
    ctx.set_bucket_pair_buffer([-1] * _lib.UD_N_BUCKETS)


def test_large_ligand_without_pair_buffer(scene):
    """A 120-atom ligand has too many pairs for the per-pair force buffer to
    fit a workgroup's LDS, so its forces take the two-sided partner path;
    a 30-atom one takes the pair-buffer path.  Both match the FP32 oracle."""
    from helpers import box_of as _box
    from oracle.oracle import Oracle
    _say("[gpu] large ligand (two-sided pair path)")
    big = synth.make_ligand(120, 12, seed=77)
    mid = synth.make_ligand(30, 6, seed=78)
    assert len(big.pairs) * 16 > 65536  # the buffer alone would exceed 64 KiB
    ligs = [big, mid]
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINA)
    c, s = _box(scene["cav"])
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps(scene["xyz"], scene["xs"], LIGAND_TYPES, _lib.UD_GRID_F32)
    ctx.set_ligands(synth.LigandBatch(ligs).arrays())
    o32 = Oracle(prec=32, sf=0)
    o32.set_grid(c, s, 0.375)
    for t in LIGAND_TYPES:
        o32.set_map(t, ctx.get_map(t).astype(np.float64))
    for l in ligs:
        o32.add_ligand(l)
    rng = np.random.default_rng(3)
    pl, P = [], []
    for li, lig in enumerate(ligs):
        for p in random_poses(rng, lig, scene["cav"], 2.0, 6):
            pl.append(li)
            P.append(p)
    pl = np.array(pl, dtype=np.int32)
    P = np.array(P)
    for lanes in (64, 16):
        ctx.set_lanes(lanes)
        e, g = ctx.score_grad(pl, P, (10.0, 10.0, 10.0))
        for i in range(len(pl)):
            li = int(pl[i])
            nt = ligs[li].n_tors
            e32, g32, _, _ = o32.eval_deriv(li, P[i], (10.0, 10.0, 10.0))
            assert e[i] == pytest.approx(e32, rel=1e-5, abs=1e-4), (lanes, i, e[i], e32)
            sc = max(1.0, np.abs(g32).max())
            np.testing.assert_allclose(g[i, :6 + nt], g32, rtol=0, atol=1e-4 * sc)
    ctx.set_lanes(0)
    ctx.close()


def _xs_element(t):
    if t in (0, 1, 19, 20, 22, 23, 25, 26, 28, 29):
        return 6
    if t in (2, 3, 4, 5, 33, 35):
        return 7
    if t in (6, 7, 8, 9, 32, 34):
        return 8
    return 100 + t


def test_postprocess_matches_oracle(scene):
    """ud_postprocess (clustering, rescoring, RMSD bounds) against the
    restatement in oracle/post.py fed with the same chains' coordinates and
    energy terms."""
    _say("[gpu] postprocess")
    _check_postprocess(scene["ctx"], scene["ligs"], exh=24, max_modes=9, min_rmsd=1.0, seed=12)


def test_postprocess_mode_capacity(scene):
    """The same check with the container at UD_MAX_MODES (many chains, a
    small clustering radius), and max_modes above it rejected."""
    _say("[gpu] postprocess at UD_MAX_MODES")
    ctx, ligs = scene["ctx"], scene["ligs"]
    n = _check_postprocess(ctx, ligs, exh=160, max_modes=_lib.UD_MAX_MODES, min_rmsd=0.25, seed=5)
    assert n >= 8 * len(ligs)
    with pytest.raises(_lib.UdError):
        ctx.postprocess(max_modes=_lib.UD_MAX_MODES + 1, min_rmsd=1.0, refine_step=0)


def _check_postprocess(ctx, ligs, exh, max_modes, min_rmsd, seed):
    from oracle import post
    P, E, _ = ctx.dock(exhaustiveness=exh, global_steps=4, seed=seed)
    modes = ctx.postprocess(max_modes=max_modes, min_rmsd=min_rmsd, refine_step=0)
    arr = synth.LigandBatch(list(ligs)).arrays()
    rot_internal = 5 * _lib.VINA_WEIGHTS[6] / 0.1 - 1
    n_checked = 0
    for li, lig in enumerate(ligs):
        a0, a1 = arr["atom_offset"][li], arr["atom_offset"][li + 1]
        heavy = np.nonzero(arr["atom_heavy"][a0:a1])[0]
        element = np.array([_xs_element(int(t)) for t in arr["atom_xs"][a0:a1][heavy]])
        pl = np.full(exh, li, dtype=np.int32)
        Pl = P[li * exh:(li + 1) * exh]
        C = ctx.pose_coords(pl, Pl)[:, heavy, :]
        inter, intra = ctx.score_terms(pl, Pl)
        ref = post.postprocess(list(C), list(E[li * exh:(li + 1) * exh]), lambda j: float(inter[j]),
                               lambda j: float(intra[j]), float(arr["num_tors"][li]), rot_internal, element,
                               min_rmsd, max_modes)
        got = modes[li]
        assert [m["chain"] - li * exh for m in got] == [m["chain"] for m in ref], li
        for g, r in zip(got, ref):
            assert g["energy"] == pytest.approx(r["energy"], rel=1e-5, abs=1e-5)
            assert g["inter"] == pytest.approx(r["inter"], rel=1e-6, abs=1e-6)
            assert g["intra"] == pytest.approx(r["intra"], rel=1e-6, abs=1e-6)
            assert g["rmsd_ub"] == pytest.approx(r["rmsd_ub"], rel=1e-5, abs=1e-5)
            assert g["rmsd_lb"] == pytest.approx(r["rmsd_lb"], rel=1e-5, abs=1e-5)
            np.testing.assert_array_equal(g["pose"], P[g["chain"]])
        energies = [m["energy"] for m in got]
        assert energies == sorted(energies) and got[0]["rmsd_ub"] == 0.0
        for i in range(len(got)):
            for j in range(i + 1, len(got)):
                ci, cj = C[got[i]["chain"] - li * exh], C[got[j]["chain"] - li * exh]
                assert post.rmsd_upper_bound(ci, cj) >= min_rmsd
        n_checked += len(got)
    assert n_checked >= 2 * len(ligs)
    return n_checked


def test_vinardo_maps_energy_and_gradient():
    """Vinardo scoring (scoring_function.h:60-70): maps vs the FP32
    restatement, energies and gradients vs FP32 and FP64 oracles."""
    _say("[gpu] vinardo")
    xyz, xs, cav = small_receptor()
    ligs = small_ligands()
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINARDO)
    c, s = box_of(cav)
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps(xyz, xs, LIGAND_TYPES, _lib.UD_GRID_F32)
    ctx.set_ligands(synth.LigandBatch(list(ligs)).arrays())
    o32, _, _ = oracle_scene(32, sf=1)
    o64m, _, _ = oracle_scene(64, sf=1, with_maps=False)
    for t in LIGAND_TYPES:
        m = ctx.get_map(t)
        r = o32.get_map(t)
        bad = ~np.isclose(m, r, rtol=1e-5, atol=1e-5)
        assert bad.mean() < 2e-3, (t, bad.mean())
        o64m.set_map(t, m.astype(np.float64))
    rng = np.random.default_rng(21)
    pl, P = [], []
    for li, lig in enumerate(ligs):
        for p in random_poses(rng, lig, cav, 2.5, 5):
            pl.append(li)
            P.append(p)
    pl = np.array(pl, dtype=np.int32)
    P = np.array(P)
    e, g = ctx.score_grad(pl, P, (10.0, 10.0, 10.0))
    lo = o64m.geo[:3]
    hi = lo + o64m.geo[6:9] * (np.array(o64m.n) - 1)
    for i in range(len(pl)):
        li = int(pl[i])
        nt = ligs[li].n_tors
        e32, g32, _, _ = o32.eval_deriv(li, P[i], (10.0, 10.0, 10.0))
        e64, g64, c64, _ = o64m.eval_deriv(li, P[i], (10.0, 10.0, 10.0))
        assert e[i] == pytest.approx(e32, rel=1e-5, abs=1e-4), (i, e[i], e32)
        if bool(np.all(c64 >= lo) and np.all(c64 <= hi)):
            assert e[i] == pytest.approx(e64, rel=1e-4, abs=1e-4), (i, e[i], e64)
        sc = max(1.0, np.abs(g32).max())
        np.testing.assert_allclose(g[i, :6 + nt], g32, rtol=0, atol=1e-4 * sc)
    ctx.close()


def test_redock_1iep_from_reference_fixture():
    """The reference's own test (unidock/test/ut/test_unidock.py): the 1iep
    ligand docked into its receptor, box centre (15.19, 53.903, 16.917), size
    20 A, fast search (exhaustiveness 128, max_step 20), seed 181129, Vina
    and Vinardo; the best mode's score lies in [-20, -1] kcal/mol.  The best
    mode is also compared with the crystal pose the ligand file holds."""
    import os
    golden = os.path.join(os.path.dirname(__file__), "golden")
    rec_xyz, rec_xs = _lib.receptor_from_pdbqt(open(os.path.join(golden, "1iep_receptor.pdbqt")).read())
    ls = _lib.LigandSet()
    ls.add_pdbqt(open(os.path.join(golden, "1iep_ligand.pdbqt")).read())
    lig = ls.ligands()[0]
    heavy = np.nonzero(lig.heavy)[0]
    for sf in (_lib.UD_SF_VINA, _lib.UD_SF_VINARDO):
        _say(f"[gpu] 1iep redock sf={sf}")
        ctx = _lib.Context(0)
        ctx.set_scoring(sf)
        ctx.set_box((15.19, 53.903, 16.917), (20.0, 20.0, 20.0), 0.375)
        types = sorted({int(t) for t in lig.xs if t >= 0})
        ctx.compute_maps(rec_xyz, rec_xs, types, _lib.UD_GRID_F32)
        ctx.set_ligands(ls.arrays())
        ctx.dock(exhaustiveness=128, global_steps=20, seed=181129)
        modes = ctx.postprocess(max_modes=9, min_rmsd=1.0)[0]
        best = modes[0]
        c = ctx.pose_coords(np.array([0], dtype=np.int32), best["pose"][None])[0, :lig.n_atoms]
        rmsd = float(np.sqrt(np.mean(np.sum((c[heavy] - lig.coords[heavy]) ** 2, axis=1))))
        _say(f"[gpu] 1iep sf={sf}: best {best['energy']:.3f} kcal/mol, {len(modes)} modes, "
             f"RMSD to crystal {rmsd:.2f} A")
        assert -20.0 <= best["energy"] <= -1.0
        # modes written as PDBQT read back with the docked coordinates
        pl_m = np.zeros(len(modes), dtype=np.int32)
        cm = ctx.pose_coords(pl_m, np.array([m["pose"] for m in modes]))[:, :lig.n_atoms]
        text = ls.write_pdbqt(0, modes, cm, how_many=9, energy_range=3.0)
        assert text.startswith("MODEL 1\nREMARK VINA RESULT:")
        body = "\n".join(ln for ln in text.split("ENDMDL\n")[0].splitlines()
                         if not ln.startswith(("MODEL", "REMARK VINA", "REMARK I", "REMARK U")))
        back = _lib.LigandSet()
        back.add_pdbqt(body)
        np.testing.assert_allclose(back.ligands()[0].coords, cm[0], atol=6e-4)
        # the crystal pose is a minimum of the scoring function: a local
        # optimisation from it stays within 1 A and lowers the energy
        P0 = np.zeros((1, _lib.UD_POSE_FLOATS), dtype=np.float32)
        P0[0, :3] = lig.origin[0]
        P0[0, 3] = 1.0
        pl = np.zeros(1, dtype=np.int32)
        e0, _ = ctx.score_grad(pl, P0)
        Po, eo, _ = ctx.optimize(pl, P0, max_steps=200)
        co = ctx.pose_coords(pl, Po)[0, :lig.n_atoms]
        r0 = float(np.sqrt(np.mean(np.sum((co[heavy] - lig.coords[heavy]) ** 2, axis=1))))
        _say(f"[gpu] 1iep sf={sf}: crystal pose {e0[0]:.3f} -> optimised {eo[0]:.3f} kcal/mol, moved {r0:.2f} A")
        assert r0 < 1.0 and eo[0] < e0[0]
        ctx.close()


def test_redock_1a30_sdf_from_reference_fixture():
    """The reference's SDF test (unidock/test/ut/test_unidock_sdf.py): the 1a30
    ligand (SDF with fragInfo / torsionInfo / atomInfo) docked into
    1a30_protein.pdbqt, box centre (8.729, 25.62, 4.682), size (19.12, 16.56,
    18.65), fast search, seed 181129, one mode written as SDF; the ENERGY of
    the "> <Uni-Dock RESULT>" field lies in [-20, -1] kcal/mol, for Vina and
    Vinardo."""
    import os
    golden = os.path.join(os.path.dirname(__file__), "golden")
    rec_xyz, rec_xs = _lib.receptor_from_pdbqt(open(os.path.join(golden, "1a30_protein.pdbqt")).read())
    ls = _lib.LigandSet()
    ls.add_sdf(open(os.path.join(golden, "1a30_ligand.sdf")).read())
    lig = ls.ligands()[0]
    for sf in (_lib.UD_SF_VINA, _lib.UD_SF_VINARDO):
        _say(f"[gpu] 1a30 SDF redock sf={sf}")
        ctx = _lib.Context(0)
        ctx.set_scoring(sf)
        ctx.set_box((8.729, 25.62, 4.682), (19.12, 16.56, 18.65), 0.375)
        types = sorted({_lib.grid_type(int(t)) for t in lig.xs} - {-1})
        ctx.compute_maps(rec_xyz, rec_xs, types, _lib.UD_GRID_F32)
        ctx.set_ligands(ls.arrays())
        ctx.dock(exhaustiveness=128, global_steps=20, seed=181129)
        modes = ctx.postprocess(max_modes=1, min_rmsd=1.0)[0]
        xyz = ctx.pose_coords(np.zeros(len(modes), np.int32), np.array([m["pose"] for m in modes]))
        text = ls.write_sdf(0, modes, xyz[:, :lig.n_atoms], how_many=1)
        lines = text.splitlines()
        score_line = lines[lines.index("> <Uni-Dock RESULT> ") + 1]
        score = float(score_line[len("ENERGY="):].split()[0])
        _say(f"[gpu] 1a30 sf={sf}: {score_line}")
        assert -20.0 <= score <= -1.0
        ctx.close()
