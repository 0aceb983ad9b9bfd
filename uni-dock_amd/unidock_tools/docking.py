"""In-process batched docking with the interface of the reference's
run_unidock (unidock_tools/modules/docking/unidock.py:155-201).

The reference builds an `unidock` command line and runs the binary in a
subprocess; here the same arguments drive the library directly on one GPU:
receptor and ligands are parsed (PDBQT / SDF), maps are built on the GPU
(for scoring "ad4" the AutoDock 4.2 maps the reference has AutoGrid write,
unidock.py:47-53, gen_grid.py:37-68),
every ligand is docked in one batch, the modes are clustered, refined and
rescored, and each input file gets `<stem>_out.sdf` / `<stem>_out.pdbqt` in
output_dir with the reference's result fields.  Returns the written files
and, per file, the mode energies in output order.
"""
from __future__ import annotations

import os
from pathlib import Path
from typing import List, Optional, Tuple

import numpy as np

from . import _lib

# --search_mode presets (main.cpp:640-652)
SEARCH_MODES = {"fast": (128, 20), "balance": (384, 40), "balanced": (384, 40), "detail": (512, 40),
                "detailed": (512, 40)}
SCORING = {"vina": _lib.UD_SF_VINA, "vinardo": _lib.UD_SF_VINARDO, "ad4": _lib.UD_SF_AD4}


def global_steps_for(ligands, max_step: int) -> int:
    """Vina::global_search_gpu (vina.h:192-204): 70 * 3 * (50 + max(N + 10 dof)) / 2
    Monte Carlo steps for the batch, capped by max_step (> 0)."""
    heur = max(l.n_atoms + 10 * (6 + l.n_tors) for l in ligands)
    g = int(70 * 3 * (50 + heur) / 2)
    return min(g, max_step) if max_step > 0 else g


def read_scores(ligand_file: Path) -> List[float]:
    """ENERGY of every "> <Uni-Dock RESULT>" field of an output SDF, or the
    first number of every REMARK VINA RESULT line of an output PDBQT."""
    scores = []
    lines = Path(ligand_file).read_text().splitlines()
    for i, ln in enumerate(lines):
        if ln.startswith("> <Uni-Dock RESULT>"):
            scores.append(float(lines[i + 1].partition("LOWER_BOUND=")[0][len("ENERGY="):]))
        elif ln.startswith("REMARK VINA RESULT:"):
            scores.append(float(ln.split()[3]))
    return scores


def run_unidock(
        receptor: Path,
        ligands: List[Path],
        output_dir: Path,
        center_x: float,
        center_y: float,
        center_z: float,
        size_x: float = 22.5,
        size_y: float = 22.5,
        size_z: float = 22.5,
        scoring: str = "vina",
        num_modes: int = 10,
        search_mode: str = "",
        exhaustiveness: int = 256,
        max_step: int = 10,
        energy_range: float = 3.0,
        refine_step: int = 5,
        bias_file: Optional[Path] = None,
        seed: int = 181129,
        score_only: bool = False,
        local_only: bool = False,
        multi_bias: bool = False,
        debug: bool = False,
        device: int = 0,
        precision: int = _lib.UD_GRID_F32,
        min_rmsd: float = 1.0,
) -> Tuple[List[Path], List[List[float]]]:
    """Dock `ligands` (PDBQT or SDF files) into `receptor` (PDBQT) and write
    `<stem>_out.<ext>` per ligand file into output_dir.

    Same arguments and defaults as the reference; biases are not built
    (NotImplementedError).  score_only writes output_dir/scores.txt
    ("REMARK <file> <affinity>" per ligand) and returns the inputs with their
    scores; local_only optimises the input pose instead of searching."""
    if scoring.lower() not in SCORING:
        raise NotImplementedError(f"scoring '{scoring}' is not built (vina, vinardo and ad4 are)")
    if bias_file is not None or multi_bias:
        raise NotImplementedError("biased docking is not built")
    if search_mode:
        if search_mode not in SEARCH_MODES:
            raise ValueError(f"unknown search_mode '{search_mode}'")
        exhaustiveness, max_step = SEARCH_MODES[search_mode]
    ligands = [Path(p) for p in ligands]
    output_dir = Path(output_dir)
    os.makedirs(output_dir, exist_ok=True)
    ls = _lib.LigandSet()
    owner = []  # (input file, ligand indices)
    for path in ligands:
        text = path.read_text()
        if path.suffix.lower() == ".sdf":
            first, cnt = ls.add_sdf(text)
            owner.append((path, list(range(first, first + cnt))))
        elif path.suffix.lower() == ".pdbqt":
            owner.append((path, [ls.add_pdbqt(text)]))
        else:
            raise ValueError(f"{path}: ligands are read from .pdbqt or .sdf files")
    sf = SCORING[scoring.lower()]
    parsed = ls.ligands(sf)
    ctx = _lib.Context(device)
    try:
        ctx.set_scoring(sf)
        ctx.set_box((center_x, center_y, center_z), (size_x, size_y, size_z), 0.375)
        if sf == _lib.UD_SF_AD4:
            rec_xyz, rec_ad, rec_q = _lib.receptor_ad4_from_pdbqt(Path(receptor).read_text())
            types = sorted({_lib.ad4_grid_type(int(t)) for l in parsed for t in l.ad} - {-1})
            ctx.compute_maps_ad4(rec_xyz, rec_ad, rec_q, types, precision)
        else:
            rec_xyz, rec_xs = _lib.receptor_from_pdbqt(Path(receptor).read_text())
            types = sorted({_lib.grid_type(int(t)) for l in parsed for t in l.xs} - {-1})
            ctx.compute_maps(rec_xyz, rec_xs, types, precision)
        ctx.set_ligands(ls.arrays(sf))
        if score_only:
            return _score_only(ctx, parsed, owner, output_dir)
        if local_only:
            modes = _local_only(ctx, parsed)
        else:
            ctx.autotune(exhaustiveness=exhaustiveness, global_steps=2)
            ctx.dock(exhaustiveness=exhaustiveness, global_steps=global_steps_for(parsed, max_step), seed=seed,
                     fetch=False)
            modes = ctx.postprocess(max_modes=num_modes, min_rmsd=min_rmsd, refine_step=refine_step)
        results, scores = [], []
        for path, idx in owner:
            parts = []
            file_scores = []
            for li in idx:
                m = modes[li]
                if not m:
                    continue
                xyz = ctx.pose_coords(np.full(len(m), li, np.int32), np.array([x["pose"] for x in m]))
                xyz = xyz[:, :parsed[li].n_atoms]
                if path.suffix.lower() == ".sdf":
                    parts.append(ls.write_sdf(li, m, xyz, how_many=num_modes, energy_range=energy_range))
                else:
                    parts.append(ls.write_pdbqt(li, m, xyz, how_many=num_modes, energy_range=energy_range))
            if not parts:
                continue
            out = output_dir / f"{path.stem}_out{path.suffix.lower()}"
            out.write_text("".join(parts))
            results.append(out)
            file_scores.extend(read_scores(out))
            scores.append(file_scores)
        return results, scores
    finally:
        ctx.close()


def _input_poses(parsed):
    """The input conformation of each ligand: root at its origin, identity
    orientation, zero torsions."""
    P = np.zeros((len(parsed), _lib.UD_POSE_FLOATS), dtype=np.float32)
    for i, l in enumerate(parsed):
        P[i, 0:3] = l.origin[0]
        P[i, 3] = 1.0
    return P


def _affinity(ctx, lig, inter, intra):
    """Final score of a pose scored on its own: Vina / Vinardo divide inter +
    intra - intra by 1 + w (num_tors) (vina.cpp:404, conf_independent.cpp:
    153-156); AD4.2 adds w_rot TORSDOF to inter (vina.cpp:1210-1227,
    conf_independent.cpp:183-187).  Returns (energy, conf_independent)."""
    if ctx.sf == _lib.UD_SF_AD4:
        c = ctx.weights[-1] * float(lig.torsdof)
        return inter + c, c
    rot = 5 * ctx.weights[-1] / 0.1 - 1  # internal rotational weight (vina.cpp:404)
    y = 1.0 + 0.1 * (rot + 1.0) * float(lig.num_tors) / 5.0
    return inter / y, inter / y - inter


def _score_only(ctx, parsed, owner, output_dir):
    """Vina::score (vina.cpp:1168-1243) of the input poses."""
    n = len(parsed)
    pl = np.arange(n, dtype=np.int32)
    inter, intra = ctx.score_terms(pl, _input_poses(parsed))
    lines, scores = [], []
    for path, idx in owner:
        s = []
        for li in idx:
            s.append(_affinity(ctx, parsed[li], float(inter[li]), float(intra[li]))[0])
        scores.append(s)
        lines.append(f"REMARK {path.name} {s[0]:.3f}")
    (output_dir / "scores.txt").write_text("\n".join(lines) + "\n")
    return [p for p, _ in owner], scores


def _local_only(ctx, parsed):
    """Local optimisation of the input poses (the reference's --local_only
    path of Vina::optimize, vina.cpp:1379-1430), reported as one mode each."""
    n = len(parsed)
    pl = np.arange(n, dtype=np.int32)
    P, e, _ = ctx.optimize(pl, _input_poses(parsed), (1000.0, 1000.0, 1000.0))
    inter, intra = ctx.score_terms(pl, P)
    modes = []
    for i in range(n):
        energy, ci = _affinity(ctx, parsed[i], float(inter[i]), float(intra[i]))
        unbound = -float(intra[i]) if ctx.sf == _lib.UD_SF_AD4 else float(intra[i])
        modes.append([dict(pose=P[i], energy=energy, inter=float(inter[i]), intra=float(intra[i]),
                           conf_independent=ci, search_energy=float(e[i]), rmsd_lb=0.0, rmsd_ub=0.0,
                           unbound=unbound, chain=i)])
    return modes
