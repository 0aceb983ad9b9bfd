"""Shared scenes for the parity tests (small sizes the oracle finishes in seconds)."""
import functools
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "uni-dock_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from unidock_tools import synth  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

LIGAND_TYPES = [0, 1, 2, 3, 8, 9, 10]  # XS types the synthetic generator draws


@functools.lru_cache(maxsize=None)
def small_receptor(n_atoms=400, edge=24.0, seed=5):
    """A small receptor at the same number density as the 2,000-atom / 40 A one."""
    rng = np.random.default_rng(seed)
    from unidock_tools import _lib
    xyz = None
    for sub in range(64):
        try:
            xyz = _lib.synth_receptor_xyz(n_atoms, edge, 3.0, seed * 1000 + sub)
            break
        except _lib.UdError:
            continue
    xs = synth._draw_types(rng, n_atoms)
    cav = synth.find_cavity(xyz, edge, 12.0)
    # open a pocket around the cavity so that docked poses are not clash
    # dominated (there BFGS trajectories are chaotic in FP32)
    keep = np.linalg.norm(xyz - cav, axis=1) > 5.0
    return xyz[keep], xs[keep], cav


BOX = 16.0  # search box edge of the test scenes (A)


def box_of(center, size=BOX):
    return np.asarray(center, dtype=np.float64), np.array([size, size, size], dtype=np.float64)


@functools.lru_cache(maxsize=None)
def small_ligands(n=6, seed=11):
    rng = np.random.default_rng(seed)
    ligs = []
    for i in range(n):
        na = int(rng.integers(8, 30))
        nt = int(rng.integers(0, 8))
        ligs.append(synth.make_ligand(na, nt, seed=int(rng.integers(0, 2**31 - 1))))
    return tuple(ligs)


def oracle_scene(prec=64, sf=0, size=BOX, with_maps=True):
    xyz, xs, cav = small_receptor()
    o = Oracle(prec=prec, sf=sf)
    c, s = box_of(cav, size)
    o.set_grid(c, s, 0.375)
    if with_maps:
        o.build_maps(xyz, xs, LIGAND_TYPES)
    ligs = small_ligands()
    for l in ligs:
        o.add_ligand(l)
    return o, ligs, (xyz, xs, cav)


def random_poses(rng, lig, center, spread, n):
    """Poses with the root inside the box, uniform orientation, torsions U(-pi, pi)."""
    out = np.zeros((n, 7 + 48))  # UD_POSE_FLOATS
    for i in range(n):
        out[i, 0:3] = center + rng.uniform(-spread, spread, 3)
        q = rng.normal(size=4)
        out[i, 3:7] = q / np.linalg.norm(q)
        out[i, 7:7 + lig.n_tors] = rng.uniform(-np.pi, np.pi, lig.n_tors)
    return out


# ---------------------------------------------------------------- AutoDock 4.2 scenes
@functools.lru_cache(maxsize=None)
def small_receptor_ad4():
    """small_receptor with AutoDock types and partial charges: (xyz, ad, q, cav)."""
    xyz, xs, cav = small_receptor()
    rec = synth.Receptor(xyz=xyz, xs=xs, edge=24.0, cavity=cav)
    ad, q = synth.ad4_receptor(rec, seed=17)
    return xyz, ad, q, cav


@functools.lru_cache(maxsize=None)
def small_ligands_ad4():
    return tuple(synth.ad4_ligand(l, seed=100 + i) for i, l in enumerate(small_ligands()))


def ad4_map_types(ligs):
    from unidock_tools import _lib
    return sorted({_lib.ad4_grid_type(int(t)) for l in ligs for t in l.ad} - {-1})


def oracle_scene_ad4(prec=64, size=10.0):
    """AD4.2 oracle with maps built from the small receptor over a `size` box."""
    xyz, ad, q, cav = small_receptor_ad4()
    ligs = small_ligands_ad4()
    o = Oracle(prec=prec, sf=2)
    c, s = box_of(cav, size)
    o.set_grid(c, s, 0.375)
    o.build_maps_ad4(xyz, ad, q, ad4_map_types(ligs))
    for l in ligs:
        o.add_ligand(l)
    return o, ligs, (xyz, ad, q, cav)
