"""Benchmark: ligands/s of batched Monte Carlo docking on one MI355X.

Workload (the metric's configuration): synthetic receptor of 2,000 heavy
atoms (protein-like X-Score mix, >= 3.0 A spacing in a 40 A cube), 1,024
synthetic ligands (10-40 heavy atoms, 0-10 torsions), search box 22.5 A at
0.375 A (61^3 points per map), Vina scoring, exhaustiveness 128, max_step 20.

A step is one ud_dock call over the whole batch with every input already
resident on the device.  Three forms are timed:
  batched  FP16 grid maps (per-map power-of-two scale, FP32 arithmetic), one
           launch per size bucket, lanes per chain autotuned per bucket,
           bucket launches concurrent                         -> `value`
  plain    the reference's launch shape: FP32 grid maps, one launch over the
           whole batch sized for its largest ligand, 32 lanes per chain
           (monte_carlo.cu:2877-2882 launches 32 threads per chain)
  plain64  FP32 grid maps, size buckets, 64 lanes per chain
Post-processing of the batched form (clustering, final score, and with
refine_step 3 the refinement against the explicit receptor) is timed per
batch after the docking steps.  The best pose of sampled ligands is
re-scored by the FP64 oracle to report the FP16 form's energy error.  The
CPU baseline is the FP32 C++ restatement of the reference's CPU global
search, built -O3 -march=native on the measuring host, on a bounded sample
of the same batch.

--config 2 / 3 / 4 run the other benchmark configurations instead, each
printing its own JSON line:
  2  screening-scale batching: 5,000-atom receptor, 10,000 ligands (atoms
     N(28, 7) in [8, 60], torsions Poisson(5) in [0, 15]), exhaustiveness 512,
     max_step 40, autotuned lanes, FP16 maps, Vina then Vinardo
  3  mixed-precision maps: 8,000-atom receptor, 30 A box (81^3), 4,096
     ligands sized as in 2; FP16 vs FP32 maps docking at exhaustiveness 128,
     max_step 20, and the energy error of 10^6 random poses against the FP64
     oracle (Vina maps: AD4 maps are not built)
  4  differentiable energy: 65,536 random poses over 2,048 ligands sized as
     in 2 against a 5,000-atom receptor, forward and backward through the
     torch.autograd binding, with central differences (h = 1e-3) of the FP64
     oracle on a sample
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "uni-dock_amd"))
sys.path.insert(0, ROOT)

from unidock_tools import _lib, synth  # noqa: E402

HBM_PEAK = 8.0e12
FP32_PEAK = 157.3e12        # MI355X published dense FP32 vector peak (packed FP32 instructions)
FP32_PEAK_UNPACKED = 78.6e12  # without packed FP32 instructions (the HIP sources are compiled without them)
# HBM bytes per search-kernel launch measured by FETCH_SIZE / WRITE_SIZE
# counter runs of this benchmark (copy-kernel calibrated), one table row per
# kernel and workgroup count
TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "search_kernel_hbm_traffic_batched_docking.md")
LIGAND_TYPES = [0, 1, 2, 3, 8, 9, 10]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(n_lig, seed):
    rec = synth.make_receptor(2000, seed=seed)
    ligs = synth.make_ligand_set(n_lig, seed=seed + 1, atoms=(10, 40), tors=(0, 10))
    return rec, ligs


def global_steps_for(ligs, max_step):
    # Vina::global_search_gpu (vina.h:192-204): 70 * 3 * (50 + max(N + 10 dof)) / 2, capped by max_step
    heur = max(l.n_atoms + 10 * (6 + l.n_tors) for l in ligs)
    return min(int(70 * 3 * (50 + heur) / 2), max_step)


def setup_ctx(rec, ligs, precision):
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINA)
    ctx.set_box(rec.cavity, (22.5, 22.5, 22.5), 0.375)
    ctx.compute_maps(rec.xyz, rec.xs, LIGAND_TYPES, precision)
    ctx.set_ligands(synth.LigandBatch(ligs).arrays())
    return ctx


def time_form(ctx, n_lig, exh, gsteps, lanes, steps, warmup, seed, single_launch=False):
    for _ in range(warmup):
        ctx.dock(exhaustiveness=exh, global_steps=gsteps, seed=seed, lanes=lanes, fetch=False,
                 single_launch=single_launch)
    if warmup:
        ctx.fetch_dock()
    kms = 0.0
    per_launch = None
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.dock(exhaustiveness=exh, global_steps=gsteps, seed=seed, lanes=lanes, fetch=False,
                 single_launch=single_launch)
        P, e, ev = ctx.fetch_dock()
        ms, nl = ctx.last_kernel_time()
        kms += ms
        li = ctx.last_launches()
        if per_launch is None:
            per_launch = li
        else:
            for a, b in zip(per_launch, li):
                a["ms"] += b["ms"]
    wall = time.perf_counter() - t0
    for a in per_launch:
        a["ms"] /= steps
    return dict(wall=wall, ligs_per_s=n_lig * steps / wall, kernel_ms=kms / steps, launches=per_launch,
                poses=P, energies=e, evals=ev)


def energy_flop(lig):
    """FP32 operations of the energy part of one evaluation in the search
    kernel (eval_energy), which every evaluation performs, counted from the
    arithmetic of each part: forward kinematics 31 for the root frame and 129
    per torsion node (origin and axis from the parent frame 33, quaternion
    product 31, normalisation 8, rotation matrix 31, half-angle sine/cosine
    26); coordinates and gyration sum 27 per atom; the grid term 60 per typed
    atom (cell coordinates 6, out-of-box penalty 6, trilinear value 21,
    gradient 12, map scale 4, curl 8, grid factor 3); pair pass A 23 per pair
    (difference 3, r^2 5, sample index 2, value and force-scale interpolation
    6, curl 6, sum 1).  Evaluations whose line-search trial is accepted also
    run the gradient part (pass B 14 per pair and 6 per atom, torque 12 per
    atom and 26 per node), not counted here: the figure is a lower bound."""
    n_typed = int(np.sum(np.asarray(lig.xs) >= 0))
    n_pairs = len(np.asarray(lig.pairs).reshape(-1, 2))
    return 31 + 129 * lig.n_tors + 27 * lig.n_atoms + 60 * n_typed + 23 * n_pairs


def algorithmic_bytes(ligs, exh, evals, bytes_per_value, sel=None):
    # 8 corner gathers per typed atom per energy+gradient evaluation
    n_typed = np.repeat([int(np.sum(l.xs >= 0)) for l in ligs], exh).astype(np.float64)
    ev = evals.astype(np.float64)
    if sel is not None:
        mask = np.repeat(sel, exh)
        n_typed, ev = n_typed[mask], ev[mask]
    return float(np.sum(ev * n_typed) * 8 * bytes_per_value)


def bucket_mask(ligs, atom_cap):
    lo = max([c for c in _lib.BUCKET_ATOMS if c < atom_cap], default=0)
    return np.array([lo < l.n_atoms <= atom_cap for l in ligs])


def cpu_threads():
    """Host threads of the CPU baseline: the cores this process may run on,
    capped at the host cores allotted to one GPU (OMP_NUM_THREADS, 16 when
    unset)."""
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(len(os.sched_getaffinity(0)), share))


def cpu_baseline(rec, ligs, exh, gsteps, n_sample, threads):
    import tempfile
    from oracle import oracle
    from oracle.oracle import Oracle
    flags = "-O3 -march=native"
    try:
        oracle.load(oracle.build_native(tempfile.mkdtemp(prefix="oracle_native_")))
    except (OSError, subprocess.CalledProcessError) as exc:  # no compiler on this host: the in-tree -O3 build
        log(f"[bench] native oracle build failed ({exc}); timing the in-tree -O3 build")
        flags = "-O3"
    o = Oracle(32, 0)
    o.set_grid(rec.cavity, np.array([22.5, 22.5, 22.5]), 0.375)
    o.build_maps(rec.xyz, rec.xs, LIGAND_TYPES)
    idx = []
    for l in ligs[:n_sample]:
        idx.append(o.add_ligand(l))
    lsteps = [(25 + l.n_atoms) // 3 for l in ligs[:n_sample]]
    t0 = time.perf_counter()
    be, ev = o.dock_cpu(idx, exh, 181129, gsteps, lsteps, threads)
    wall = time.perf_counter() - t0
    return dict(value=n_sample / wall, unit="ligands/s", cores=threads, kind="port",
                sample=f"{n_sample} ligands of the batch x exhaustiveness {exh} x {gsteps} MC steps, "
                       f"FP32 C++ restatement of the reference CPU global search (g++ {flags}), {wall:.1f} s wall",
                evals_per_s=ev / wall)


def time_postprocess(ctx, refine_step, max_modes=9):
    """Wall time of ud_postprocess over the last dock's batch (clustering,
    refinement when refine_step > 0, final scores, RMSD bounds)."""
    import ctypes
    n = ctx.n_ligands
    modes = (_lib.UdMode * (n * max_modes))()
    cnt = np.zeros(n, dtype=np.int32)
    t0 = time.perf_counter()
    ctx._ck(ctx.L.ud_postprocess(ctx.h, max_modes, 1.0, refine_step, modes,
                                 cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return time.perf_counter() - t0


def fp64_error(rec, ligs, ctx16, res, exh, n_sample):
    """Best pose per sampled ligand: FP16-grid energy vs FP64 oracle (own FP64 maps)."""
    from oracle.oracle import Oracle
    o = Oracle(64, 0)
    o.set_grid(rec.cavity, np.array([22.5, 22.5, 22.5]), 0.375)
    o.build_maps(rec.xyz, rec.xs, LIGAND_TYPES)
    sel = np.linspace(0, len(ligs) - 1, n_sample).astype(int)
    pl, P = [], []
    for li in sel:
        o.add_ligand(ligs[li])
        e = res["energies"][li * exh:(li + 1) * exh]
        c = int(np.argmin(e))
        pl.append(li)
        P.append(res["poses"][li * exh + c])
    e16, _ = ctx16.score_grad(np.array(pl, dtype=np.int32), np.array(P))
    errs = []
    for k, li in enumerate(sel):
        e64 = o.eval_deriv(k, np.array(P[k], dtype=np.float64))[0]
        errs.append(abs(float(e16[k]) - e64))
    return float(np.max(errs)), float(np.mean(errs))


def random_poses(rng, ligs, pl, lo, hi):
    """Translation uniform in the box, orientation uniform on SO(3) (normalised
    4-D Gaussian), torsions U[-pi, pi)."""
    P = np.zeros((len(pl), _lib.UD_POSE_FLOATS), dtype=np.float64)
    P[:, 0:3] = lo + rng.random((len(pl), 3)) * (hi - lo)
    q = rng.normal(size=(len(pl), 4))
    P[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
    for i, li in enumerate(pl):
        nt = ligs[li].n_tors
        P[i, 7:7 + nt] = rng.uniform(-np.pi, np.pi, nt)
    return P


def config2(args):
    rec = synth.make_receptor(5000, seed=2000)
    ligs = synth.make_ligand_set(10000, seed=2001, dist="screening")
    types = sorted({int(t) for l in ligs for t in l.xs if t >= 0})
    exh, gsteps = 512, global_steps_for(ligs, 40)
    log(f"[config2] {len(ligs)} ligands, atoms {min(l.n_atoms for l in ligs)}-{max(l.n_atoms for l in ligs)}, "
        f"torsions up to {max(l.n_tors for l in ligs)}, global_steps {gsteps}")
    res = {}
    for name, sf in (("vina", _lib.UD_SF_VINA), ("vinardo", _lib.UD_SF_VINARDO)):
        ctx = _lib.Context(0)
        ctx.set_scoring(sf)
        ctx.set_box(rec.cavity, (22.5, 22.5, 22.5), 0.375)
        ctx.compute_maps(rec.xyz, rec.xs, types, _lib.UD_GRID_F16)
        ctx.set_ligands(synth.LigandBatch(ligs).arrays())
        t0 = time.perf_counter()
        tuned = ctx.autotune(exhaustiveness=exh, global_steps=2, sample_ligands=1024)
        t_tune = time.perf_counter() - t0
        r = time_form(ctx, len(ligs), exh, gsteps, 0, 1, 0, 181129)
        ev = float(r["evals"].sum())
        res[name] = dict(ligands_per_s=r["ligs_per_s"], wall_s=r["wall"], kernel_ms=r["kernel_ms"],
                         evals_per_s=ev / (r["kernel_ms"] * 1e-3), autotune_s=t_tune,
                         tuned_lanes=dict(zip([str(c) for c in _lib.BUCKET_ATOMS], [int(x) for x in tuned])),
                         launches=r["launches"])
        log(f"[config2] {name}: {r['ligs_per_s']:.1f} ligands/s ({r['wall']:.1f} s), lanes {res[name]['tuned_lanes']}")
        ctx.close()
    print(json.dumps({"metric": "ligands/sec, screening-scale batch (10,000 ligands, exhaustiveness 512, max_step 40)",
                      "value": res["vina"]["ligands_per_s"], "unit": "ligands/s", "n_gpus": 1,
                      "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded receptor and ligands)",
                      "config": {"workload": "5,000-atom receptor, 10,000 ligands N(28,7) atoms / Poisson(5) "
                                             "torsions, 22.5 A box (61^3 fp16 maps), exhaustiveness 512, max_step 40",
                                 "global_steps": gsteps},
                      "vina": res["vina"], "vinardo": res["vinardo"]}))


def config3(args):
    """AutoDock 4.2 maps at 81^3 (type maps + electrostatic + desolvation):
    docking throughput of 4,096 ligands with FP16 and FP32 maps, and the
    per-pose energy error of both against the FP64 oracle on 10^6 random
    poses.  The oracle samples one 13,424-entry table per atom pair, as the
    reference does (precalculate_byatom), so the error poses are drawn over a
    32-ligand subset."""
    from oracle.oracle import Oracle
    rec = synth.make_receptor(8000, seed=3000)
    # AD4.2's 12-6 walls (4 A C-C contact) leave no bound pose in the
    # 3 A-packed synthetic receptor: clear a 7 A pocket around the cavity
    keep = np.linalg.norm(rec.xyz - rec.cavity, axis=1) > 7.0
    rec = synth.Receptor(xyz=rec.xyz[keep], xs=rec.xs[keep], edge=rec.edge, cavity=rec.cavity)
    rec_ad, rec_q = synth.ad4_receptor(rec, seed=3003)
    ligs = [synth.ad4_ligand(l, seed=3004 + i)
            for i, l in enumerate(synth.make_ligand_set(4096, seed=3001, dist="screening"))]
    types = sorted({_lib.ad4_grid_type(int(t)) for l in ligs for t in l.ad} - {-1})
    box = (30.0, 30.0, 30.0)
    exh, gsteps = 128, global_steps_for(ligs, 20)
    out = {}
    ctxs = {}
    for name, prec in (("fp16", _lib.UD_GRID_F16), ("fp32", _lib.UD_GRID_F32)):
        ctx = _lib.Context(0)
        ctx.set_scoring(_lib.UD_SF_AD4)
        npts = ctx.set_box(rec.cavity, box, 0.375)
        ctx.compute_maps_ad4(rec.xyz, rec_ad, rec_q, types, prec)
        map_ms = ctx.last_kernel_time()[0]
        ctx.set_ligands(synth.LigandBatch(ligs).arrays())
        ctx.autotune(exhaustiveness=exh, global_steps=2, sample_ligands=1024)
        r = time_form(ctx, len(ligs), exh, gsteps, 0, 1, 1, 181129)
        out[name] = dict(ligands_per_s=r["ligs_per_s"], kernel_ms=r["kernel_ms"], map_build_ms=map_ms)
        docked = (r["poses"], r["energies"])
        log(f"[config3] AD4.2 {name} maps ({npts}, {len(types) + 2} maps built in {map_ms:.1f} ms): "
            f"{r['ligs_per_s']:.1f} ligands/s")
        ctxs[name] = ctx
    # per-pose energy error on random poses vs the FP64 oracle (its own FP64 maps)
    n_pose = args.poses
    rng = np.random.default_rng(3002)
    sub = rng.choice(len(ligs), 32, replace=False)
    pl_sub = rng.integers(0, len(sub), n_pose).astype(np.int32)
    pl = sub[pl_sub].astype(np.int32)
    o = Oracle(64, 2)
    n, geo = o.set_grid(rec.cavity, np.array(box), 0.375)
    lo = geo[:3]
    hi = lo + geo[6:9] * (np.array(n) - 1)
    P = random_poses(rng, ligs, pl, lo, hi)
    t0 = time.perf_counter()
    o.build_maps_ad4(rec.xyz, rec_ad, rec_q, types)
    t1 = time.perf_counter()
    for k in sub:
        o.add_ligand(ligs[int(k)])
    nth = min(16, os.cpu_count() or 1)
    e64 = o.eval_batch(pl_sub, P, (1000.0, 1000.0, 1000.0), nth)
    # the docked poses (every chain's best of the last FP32 dock) of the subset
    dk_sub = np.repeat(np.arange(len(sub), dtype=np.int32), exh)
    dk_pl = sub[dk_sub].astype(np.int32)
    dk_P = docked[0].reshape(len(ligs), exh, -1)[sub].reshape(-1, docked[0].shape[1]).astype(np.float64)
    dk_e64 = o.eval_batch(dk_sub, dk_P, (1000.0, 1000.0, 1000.0), nth)
    log(f"[config3] FP64 oracle: maps in {t1 - t0:.1f} s, {n_pose} + {len(dk_pl)} poses in {time.perf_counter() - t1:.1f} s")
    errs = {}
    for name, ctx in ctxs.items():
        e = np.zeros(n_pose, dtype=np.float32)
        for a in range(0, n_pose, 1 << 17):
            b = min(n_pose, a + (1 << 17))
            e[a:b], _ = ctx.score_grad(pl[a:b], P[a:b])
        d = np.abs(e.astype(np.float64) - e64)
        # poses with an atom outside the box carry the 1e6 / A penalty: report
        # those relative, the others in kcal/mol
        inside = e64 < 1e4
        rel = d[~inside] / np.abs(e64[~inside]) if np.any(~inside) else np.zeros(1)
        low = np.abs(e64) < 100.0  # poses without steep clashes (AD4.2's capped 12-6 cores are 1e5)
        mx = lambda x: float(x.max()) if x.size else None
        mn = lambda x: float(x.mean()) if x.size else None
        # docked poses: the search's own results, where the 0.1 kcal/mol bound matters
        ed = np.zeros(len(dk_pl), dtype=np.float32)
        ed[:], _ = ctx.score_grad(dk_pl, dk_P)
        dd = np.abs(ed.astype(np.float64) - dk_e64)
        bound = dk_e64 < 0.0
        errs[name] = dict(n_inside=int(inside.sum()), max_abs=mx(d[inside]), mean_abs=mn(d[inside]),
                          p99_abs=float(np.percentile(d[inside], 99)), max_rel_outside=float(rel.max()),
                          n_below_100=int(low.sum()), max_abs_below_100=mx(d[low]),
                          mean_abs_below_100=mn(d[low]),
                          max_rel_inside=float((d[inside] / np.maximum(1.0, np.abs(e64[inside]))).max()),
                          docked_poses=int(len(dd)), docked_bound=int(bound.sum()), docked_max_abs=mx(dd[bound]),
                          docked_mean_abs=mn(dd[bound]), docked_max_abs_all=mx(dd))
        errs[name]["docked_e64_quantiles"] = [float(x) for x in np.percentile(dk_e64, [0, 10, 50, 90])]
        errs[name]["docked_max_rel"] = float((dd / np.maximum(1.0, np.abs(dk_e64))).max())
        log(f"[config3] {name} error vs FP64: {inside.sum()} random in-box poses max relative "
            f"{errs[name]['max_rel_inside']:.2e} ({errs[name]['n_below_100']} with |E| < 100); {len(dd)} docked "
            f"poses (E quantiles 0/10/50/90 %: {np.round(errs[name]['docked_e64_quantiles'], 3)}): max "
            f"{float(dd.max()):.4f}, mean {float(dd.mean()):.5f} kcal/mol, max relative {errs[name]['docked_max_rel']:.2e}"
            + (f"; E < 0: {int(bound.sum())}, max {errs[name]['docked_max_abs']:.4f}" if bound.any() else ""))
        ctx.close()
    print(json.dumps({"metric": "ligands/sec and per-pose energy error, FP16 vs FP32 AD4.2 maps (81^3)",
                      "value": out["fp16"]["ligands_per_s"], "unit": "ligands/s", "n_gpus": 1,
                      "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded receptor and ligands)",
                      "config": {"workload": "8,000-atom receptor with a 7 A pocket cleared around the cavity "
                                             "(%d atoms), 30 A box (81^3), 4,096 ligands N(28,7) / Poisson(5), "
                                             "AD4.2 scoring: %d affinity maps + electrostatic + desolvation, "
                                             "exhaustiveness 128, max_step 20" % (len(rec.xs), len(types)),
                                 "poses_for_error": n_pose, "ligands_for_error": 32},
                      "docking": out, "energy_error_vs_fp64": errs}))


def config4(args):
    import torch
    from oracle.oracle import Oracle
    from unidock_tools.torch_energy import DockingEnergy
    rec = synth.make_receptor(5000, seed=4000)
    ligs = synth.make_ligand_set(2048, seed=4001, dist="screening")
    box = np.array([22.5, 22.5, 22.5])
    model = DockingEnergy(rec.xyz, rec.xs, ligs, rec.cavity, box)
    rng = np.random.default_rng(4002)
    pl = np.repeat(np.arange(len(ligs), dtype=np.int32), 32)
    o = Oracle(64, 0)
    n, geo = o.set_grid(rec.cavity, box, 0.375)
    lo = geo[:3]
    hi = lo + geo[6:9] * (np.array(n) - 1)
    P = random_poses(rng, ligs, pl, lo, hi)
    poses = torch.tensor(P, dtype=torch.float32, device="cuda", requires_grad=True)
    w = torch.tensor(_lib.VINA_WEIGHTS, dtype=torch.float64, requires_grad=True)
    for _ in range(args.warmup + 1):
        e = model(poses, pl, w)
        e.sum().backward()
    torch.cuda.synchronize()
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        poses.grad = None
        w.grad = None
        e = model(poses, pl, w)
        e.sum().backward()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    log(f"[config4] {len(pl)} poses forward + backward in {dt * 1e3:.2f} ms")
    # central differences (h = 1e-3) of the FP64 oracle on a sample of in-box poses
    t1 = time.perf_counter()
    o.build_maps(rec.xyz, rec.xs, sorted({int(t) for l in ligs for t in l.xs if t >= 0}))
    sel = rng.choice(len(pl), 64, replace=False)
    for l in ligs:
        o.add_ligand(l)
    g = poses.grad.detach().cpu().numpy()
    rel = []
    for i in sel:
        li = int(pl[i])
        nt = ligs[li].n_tors
        x = P[i].copy()
        fd = np.zeros(7 + nt)
        for j in range(7 + nt):
            a, b = x.copy(), x.copy()
            a[j] += 1e-3
            b[j] -= 1e-3
            qa, qb = a[3:7] / np.linalg.norm(a[3:7]), b[3:7] / np.linalg.norm(b[3:7])
            a[3:7], b[3:7] = qa, qb
            fd[j] = (o.eval_deriv(li, a)[0] - o.eval_deriv(li, b)[0]) / 2e-3
        sc = max(1.0, np.abs(fd).max())
        rel.append(float(np.abs(g[i, :7 + nt] - fd).max() / sc))
    rel = np.array(rel)
    log(f"[config4] FD check on {len(sel)} poses: median {np.median(rel):.2e}, 90th pct {np.percentile(rel, 90):.2e} "
        f"(max-norm error / max(1, |g|)) in {time.perf_counter() - t1:.1f} s")
    print(json.dumps({"metric": "poses/sec, energy + pose and weight gradients (torch.autograd)",
                      "value": len(pl) / dt, "unit": "poses/s", "n_gpus": 1, "higher_is_better": True,
                      "dtype": "f32", "data": "synthetic (seeded receptor, ligands and poses)", "steps": steps,
                      "ms_per_step": dt * 1e3,
                      "config": {"workload": "65,536 random poses over 2,048 ligands N(28,7) / Poisson(5), "
                                             "5,000-atom receptor, 22.5 A box (61^3 fp32 maps), forward + backward"},
                      "fd_check_h_1e-3": {"sample": len(sel), "median_rel": float(np.median(rel)),
                                          "p90_rel": float(np.percentile(rel, 90)), "max_rel": float(rel.max())}}))


def measured_traffic(kernel, workgroups):
    """HBM bytes of one launch from the committed counter table, or None."""
    try:
        lines = open(TRAFFIC_TABLE).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].strip("`") == kernel and cells[1] == str(workgroups):
            return float(cells[4].split()[0]) * 1e9
    return None


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB in one elementwise kernel: the
    reference point for the FETCH_SIZE / WRITE_SIZE counters (it reads and
    writes exactly mib MiB)."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert torch.equal(dst[-8:], src[-8:])
    print(json.dumps({"copy_kernel_mib": mib, "launches": 3, "bytes_each_way": n * 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ligands", type=int, default=1024)
    ap.add_argument("--exhaustiveness", type=int, default=128)
    ap.add_argument("--max-step", type=int, default=20)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per chain of the batched form (0 = per bucket)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=512)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--config", type=int, default=1, choices=(1, 2, 3, 4),
                    help="benchmark configuration (1 = the metric's; 2-4 print their own line)")
    ap.add_argument("--poses", type=int, default=1000000, help="random poses of the --config 3 error check")
    ap.add_argument("--copy-kernel-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_kernel_mib:
        copy_calibration(args.copy_kernel_mib)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the ligand batch would be split across "
              "GPUs by the host with no collective in the search", file=sys.stderr)
        sys.exit(2)
    if args.config != 1:
        {2: config2, 3: config3, 4: config4}[args.config](args)
        return

    t0 = time.perf_counter()
    rec, ligs = make_workload(args.ligands, args.seed)
    gsteps = global_steps_for(ligs, args.max_step)
    log(f"[bench] workload ready in {time.perf_counter() - t0:.1f}s: {len(ligs)} ligands, "
        f"atoms {min(l.n_atoms for l in ligs)}-{max(l.n_atoms for l in ligs)}, global_steps {gsteps}")
    exh = args.exhaustiveness

    ctx16 = setup_ctx(rec, ligs, _lib.UD_GRID_F16)
    tuned = tuned_pb = None
    if args.lanes == 0:
        t1 = time.perf_counter()
        # whole buckets, 2 MC steps: a sample must fill the GPU as the real
        # launch does, or lane counts with more workgroups win by occupancy
        tuned = [int(x) for x in ctx16.autotune(exhaustiveness=exh, global_steps=2, sample_ligands=len(ligs))]
        tuned_pb = [int(x) for x in ctx16.bucket_layout()[1]]
        log(f"[bench] autotuned (lanes, per-pair LDS buffer) per bucket "
            f"{dict(zip(_lib.BUCKET_ATOMS, zip(tuned, tuned_pb)))} in {time.perf_counter() - t1:.1f}s (untimed set-up)")
    log("[bench] batched FP16 form")
    b = time_form(ctx16, len(ligs), exh, gsteps, args.lanes, args.steps, args.warmup, 181129)
    ev_lig = b["evals"].reshape(len(ligs), exh).sum(axis=1)
    for a in b["launches"]:
        a["evals"] = int(ev_lig[bucket_mask(ligs, a["atom_cap"])].sum())
    log(f"[bench] batched: {b['ligs_per_s']:.1f} ligands/s, kernel {b['kernel_ms']:.1f} ms/step, "
        f"launches {b['launches']}")
    post = {}
    for rs in (0, 3):
        post[f"refine_step_{rs}"] = time_postprocess(ctx16, rs)
    log(f"[bench] post-processing per batch: {post}")
    plain = plain64 = None
    if not args.no_plain:
        ctx32 = setup_ctx(rec, ligs, _lib.UD_GRID_F32)
        log("[bench] plain FP32 form (one launch per batch, 32 lanes per chain)")
        plain = time_form(ctx32, len(ligs), exh, gsteps, 32, args.steps, args.warmup, 181129, single_launch=True)
        log(f"[bench] plain: {plain['ligs_per_s']:.1f} ligands/s, kernel {plain['kernel_ms']:.1f} ms/step")
        log("[bench] plain64 FP32 form (size buckets, 64 lanes per chain)")
        plain64 = time_form(ctx32, len(ligs), exh, gsteps, 64, args.steps, args.warmup, 181129)
        log(f"[bench] plain64: {plain64['ligs_per_s']:.1f} ligands/s, kernel {plain64['kernel_ms']:.1f} ms/step")
        ctx32.close()

    evals = float(b["evals"].sum())
    # dominant kernel: the longest launch of the batched form (one size bucket)
    dom = max(b["launches"], key=lambda a: a["ms"])
    sel = bucket_mask(ligs, dom["atom_cap"])
    bytes_dom = algorithmic_bytes(ligs, exh, b["evals"], 2, sel)
    achieved = bytes_dom / (dom["ms"] * 1e-3)
    dom_kernel = f"void ud::search_kernel<{dom['lanes']}, 1, true, 16>"
    roof = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=achieved / HBM_PEAK,
                traffic=measured_traffic(dom_kernel, dom["n_workgroups"]))
    # issue roofline of the same launch: FP32 operations of the evaluations'
    # energy part (a lower bound of the work done) per second vs the FP32 peak
    ev_dom = b["evals"].reshape(len(ligs), exh)[sel]
    flop_dom = float(np.sum(ev_dom.sum(axis=1) * np.array([energy_flop(l) for l, s_ in zip(ligs, sel) if s_])))
    fl = flop_dom / (dom["ms"] * 1e-3)
    issue = dict(bound="valu", achieved=fl / 1e12, peak=FP32_PEAK / 1e12, unit="TFLOP/s", frac=fl / FP32_PEAK,
                 frac_unpacked=fl / FP32_PEAK_UNPACKED, flop_per_eval_mean=flop_dom / float(ev_dom.sum()),
                 counted="energy part of each evaluation (lower bound)")
    log(f"[bench] dominant launch: bucket <= {dom['atom_cap']} atoms, {dom['lanes']} lanes, "
        f"{dom['n_ligands']} ligands, {dom['ms']:.2f} ms, {bytes_dom / 1e6:.1f} MB algorithmic")
    err_max, err_mean = fp64_error(rec, ligs, ctx16, b, exh, 64)
    log(f"[bench] FP16 best-pose energy error vs FP64: max {err_max:.4f} mean {err_mean:.5f} kcal/mol")
    cpu = None
    if not args.no_cpu:
        threads = cpu_threads()
        log(f"[bench] CPU baseline on {threads} threads")
        cpu = cpu_baseline(rec, ligs, exh, gsteps, args.cpu_sample, threads)
        log(f"[bench] cpu: {cpu['value']:.3f} ligands/s")

    out = {
        "metric": "ligands/sec on 1 MI355X, batched FP16-grid vs plain, at <=0.1 kcal/mol err",
        "value": round(b["ligs_per_s"], 3),
        "unit": "ligands/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * b["wall"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded receptor and ligands)",
        "config": {
            "workload": "batched docking: 2,000-atom receptor, 1,024 ligands (10-40 atoms, 0-10 torsions), "
                        "box 22.5 A / 0.375 A (61^3), Vina, exhaustiveness 128, max_step 20",
            "n_ligands": len(ligs), "exhaustiveness": exh, "global_steps": gsteps, "grid": "61^3 fp16 + per-map scale",
            "lanes_per_chain": args.lanes or "autotuned per size bucket",
        },
        "roofline": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in roof.items()},
        "issue_roofline": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in issue.items()},
        "cpu_baseline": cpu,
        "evals_per_s": evals / (b["kernel_ms"] * 1e-3),
        "kernel_ms_per_step": b["kernel_ms"],
        "launches": b["launches"],
        "tuned_lanes": dict(zip([str(c) for c in _lib.BUCKET_ATOMS], tuned)) if tuned else None,
        "tuned_pair_buffer": dict(zip([str(c) for c in _lib.BUCKET_ATOMS], tuned_pb)) if tuned else None,
        "dominant_kernel": {"name": f"search_kernel<{dom['lanes']}, fp16 maps, MC>",
                            "bucket_atoms": dom["atom_cap"], "avg_ms": dom["ms"],
                            "algorithmic_bytes": bytes_dom},
        "plain_fp32_ligands_per_s": plain["ligs_per_s"] if plain else None,
        "plain_form": "FP32 maps, one launch over the batch sized for its largest ligand, 32 lanes per chain",
        "speedup_vs_plain": (b["ligs_per_s"] / plain["ligs_per_s"]) if plain else None,
        "plain64_fp32_ligands_per_s": plain64["ligs_per_s"] if plain64 else None,
        "speedup_vs_plain64": (b["ligs_per_s"] / plain64["ligs_per_s"]) if plain64 else None,
        "postprocess_s_per_batch": post,
        "end_to_end_ligands_per_s": {k: len(ligs) / (b["wall"] / args.steps + v) for k, v in post.items()},
        "fp16_best_pose_err_vs_fp64_kcal": {"max": err_max, "mean": err_mean, "sample": 64},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
