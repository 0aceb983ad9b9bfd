"""The Vina-style C++ host (uni-dock_amd/host/unidock_vina.{h,cpp}) over the C
ABI, compiled with g++ together with tests/cpp/vina_host_client.cpp.

  build    the host and a client compile and link with -Wall -Werror (CPU)
  dock     the reference's 1iep fixture (test_unidock.py: box centre
           (15.19, 53.903, 16.917), 20 A) through set_receptor /
           set_ligand_from_file_gpu / compute_vina_maps / global_search_gpu /
           write_poses_gpu: every mode's energy terms equal the Python
           binding's for the same calls (same library and inputs, fixed
           launch configuration), the written PDBQT carries them, and the
           best score lies in the reference test's range (-20 <= E <= -1)
  ad4      AutoGrid map files written from GPU-built AD4.2 maps of 1iep and
           read back by load_maps: the modes equal the Python binding's
           load_ad4_maps + dock
  errors   the reference's error messages (vina.h:156-166, vina.cpp:491-510)
  modes    the --search_mode presets (main.cpp:644-656), unknown names
           leaving exhaustiveness and max_step unchanged (CPU)
"""
import json
import os
import subprocess

import numpy as np
import pytest

from unidock_tools import _lib, docking

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.dirname(_lib.lib_path())
GOLDEN = os.path.join(HERE, "golden")
REC = os.path.join(GOLDEN, "1iep_receptor.pdbqt")
LIG = os.path.join(GOLDEN, "1iep_ligand.pdbqt")
CENTER = (15.19, 53.903, 16.917)


def _build(tmp_path):
    exe = str(tmp_path / "vina_host_client")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "uni-dock_amd", "host"), os.path.join(ROOT, "uni-dock_amd", "host", "unidock_vina.cpp"),
           os.path.join(HERE, "cpp", "vina_host_client.cpp"), "-o", exe,
           "-L", LIBDIR, "-lunidock_amd", f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, *args):
    r = subprocess.run([exe, *[str(a) for a in args]], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def _terms(modes):
    return np.array([[m["energy"], m["inter"], m["intra"], m["conf_independent"], m["unbound"]] for m in modes],
                    dtype=np.float32)


def test_build(tmp_path):
    _build(tmp_path)


def test_search_mode_presets(tmp_path):
    exe = _build(tmp_path)
    got = _run(exe, "modes", "fast", "balance", "balanced", "detail", "detailed", "Fast", "")["modes"]
    assert got == [[1, 128, 20], [1, 384, 40], [1, 384, 40], [1, 512, 40], [1, 512, 40], [0, 8, 0], [0, 8, 0]]
    # the Python binding's presets are the same table
    assert {k: list(v) for k, v in docking.SEARCH_MODES.items()} == {
        "fast": [128, 20], "balance": [384, 40], "balanced": [384, 40], "detail": [512, 40], "detailed": [512, 40]}


@pytest.mark.gpu
def test_dock_1iep_matches_python_binding(tmp_path):
    print("[gpu] C++ host docks 1iep", flush=True)
    exe = _build(tmp_path)
    exh, max_step, seed, n_poses, refine = 32, 10, 181129, 9, 3
    out = tmp_path / "1iep_out.pdbqt"
    got = _run(exe, "dock", REC, LIG, *CENTER, 20.0, exh, max_step, seed, 0.375, out, "vina", n_poses, refine)
    # the same calls through the Python binding
    ctx = _lib.Context(0)
    try:
        ctx.set_scoring(_lib.UD_SF_VINA)
        ctx.set_box(CENTER, (20.0, 20.0, 20.0), 0.375)
        xyz, xs = _lib.receptor_from_pdbqt(open(REC).read())
        ls = _lib.LigandSet()
        ls.add_pdbqt(open(LIG).read())
        parsed = ls.ligands(_lib.UD_SF_VINA)
        ctx.set_ligands(ls.arrays(_lib.UD_SF_VINA))
        types = sorted({_lib.grid_type(int(t)) for l in parsed for t in l.xs} - {-1})
        ctx.compute_maps(xyz, xs, types, _lib.UD_GRID_F32)
        ctx.dock(exhaustiveness=exh, global_steps=docking.global_steps_for(parsed, max_step), seed=seed, fetch=False)
        modes = ctx.postprocess(max_modes=n_poses, min_rmsd=1.0, refine_step=refine)[0]
    finally:
        ctx.close()
    print(f"[gpu] C++ host: {got['n_modes']} modes, best {got['energies'][0][0]:.3f}", flush=True)
    assert got["n_modes"] == len(modes) >= 1
    np.testing.assert_array_equal(np.array(got["energies"], dtype=np.float32), _terms(modes))
    assert -20.0 <= got["energies"][0][0] <= -1.0
    written = docking.read_scores(out)
    np.testing.assert_allclose(written, [m["energy"] for m in modes], rtol=0, atol=6e-4)


@pytest.mark.gpu
def test_ad4_autogrid_maps_match_python_binding(tmp_path):
    print("[gpu] C++ host: AD4.2 from AutoGrid files", flush=True)
    exe = _build(tmp_path)
    ls = _lib.LigandSet()
    ls.add_pdbqt(open(LIG).read())
    arr = ls.arrays(_lib.UD_SF_AD4)
    types = sorted({_lib.ad4_grid_type(int(t)) for t in arr["atom_ad"]} - {-1})
    center = np.round(np.array(CENTER), 3)
    prefix = str(tmp_path / "rec")
    c1 = _lib.Context(0)
    try:
        c1.set_scoring(_lib.UD_SF_AD4)
        n = np.array(c1.set_box(center, (20.25, 20.25, 20.25), 0.375))
        assert np.all(n % 2 == 1)  # an even number of voxels, as AutoGrid writes
        xyz, ad, q = _lib.receptor_ad4_from_pdbqt(open(REC).read())
        c1.compute_maps_ad4(xyz, ad, q, types, _lib.UD_GRID_F32)
        names = {t: _lib.AD_TYPE_NAMES[t] for t in types}
        names[_lib.UD_AD4_MAP_ELEC], names[_lib.UD_AD4_MAP_DSOLV] = "e", "d"
        for t, nm in names.items():
            _lib.write_ad4_map(f"{prefix}.{nm}.map", center, n - 1, 0.375, c1.get_map(t))
    finally:
        c1.close()
    exh, max_step, seed, n_poses = 32, 10, 7, 9
    got = _run(exe, "ad4", LIG, prefix, exh, max_step, seed, n_poses, tmp_path / "1iep_ad4_out.pdbqt")
    c2 = _lib.Context(0)
    try:
        c2.set_scoring(_lib.UD_SF_AD4)
        c2.load_ad4_maps(prefix, arr["atom_ad"])
        c2.set_ligands(arr)
        parsed = ls.ligands(_lib.UD_SF_AD4)
        c2.dock(exhaustiveness=exh, global_steps=docking.global_steps_for(parsed, max_step), seed=seed, fetch=False)
        modes = c2.postprocess(max_modes=n_poses, min_rmsd=1.0, refine_step=0)[0]
    finally:
        c2.close()
    print(f"[gpu] C++ host AD4.2: {got['n_modes']} modes, best {got['energies'][0][0]:.3f}", flush=True)
    assert got["n_modes"] == len(modes) >= 1
    np.testing.assert_array_equal(np.array(got["energies"], dtype=np.float32), _terms(modes))


@pytest.mark.gpu
def test_reference_error_messages(tmp_path):
    exe = _build(tmp_path)
    e = _run(exe, "errors", REC, LIG)["errors"]
    assert e == [
        "ERROR: Scoring function ad5 not implemented (choices: vina, vinardo or ad4)",
        "ERROR: Cannot do the global search. Ligand(s) was(ere) not initialized.",
        "ERROR: Cannot compute Vina or Vinardo affinity maps. The (rigid) receptor was not initialized.",
        "ERROR: Grid box dimensions must be greater than 0 Angstrom.",
        "ERROR: Cannot do the global search. Affinity maps were not initialized.",
        "ERROR: Exhaustiveness must be 1 or greater",
        "ERROR: no poses for ligand 0",
        "ERROR: load_maps reads AutoGrid maps of the ad4 scoring function",
        "ERROR: Could not open /nonexistent.pdbqt for reading.",
    ]
