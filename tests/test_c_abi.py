"""The C ABI from a C translation unit (tests/c/abi_client.c compiled with gcc
against include/unidock_amd.h): struct layouts equal the ctypes mirror the
Python binding uses, and a C host docks the reference's 1iep fixture through
the ABI (GPU)."""
import json
import os
import subprocess

import pytest

from unidock_tools import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.dirname(_lib.lib_path())


def _build(tmp_path, link):
    exe = str(tmp_path / "abi_client")
    cmd = ["gcc", "-std=c99", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "c", "abi_client.c"), "-o", exe]
    if link:
        cmd += ["-L", LIBDIR, "-lunidock_amd", f"-Wl,-rpath,{LIBDIR}"]
    else:
        cmd += ["-Wl,--unresolved-symbols=ignore-all"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def _ctypes_layout():
    out = {"sizeof.ud_search_params": __import__("ctypes").sizeof(_lib.UdSearchParams),
           "sizeof.ud_mode": __import__("ctypes").sizeof(_lib.UdMode),
           "sizeof.ud_ligand_batch": __import__("ctypes").sizeof(_lib.UdLigandBatch),
           "sizeof.ud_launch_info": __import__("ctypes").sizeof(_lib.UdLaunchInfo)}
    for cname, S in (("ud_search_params", _lib.UdSearchParams), ("ud_mode", _lib.UdMode),
                     ("ud_ligand_batch", _lib.UdLigandBatch), ("ud_launch_info", _lib.UdLaunchInfo)):
        for f, _t in S._fields_:
            d = getattr(S, f)
            out[f"{cname}.{f}"] = [d.offset, d.size]
    return out


def test_struct_layouts_match_ctypes(tmp_path):
    exe = _build(tmp_path, link=False)
    c = json.loads(subprocess.run([exe, "layout"], check=True, capture_output=True, text=True).stdout)
    py = _ctypes_layout()
    for k, v in py.items():
        assert k in c, k
        assert c[k] == v, (k, c[k], v)
    # every field the C header declares is mirrored
    assert sorted(k for k in c if "." in k and not k.startswith("sizeof")) == \
        sorted(k for k in py if not k.startswith("sizeof"))
    assert c["UD_POSE_FLOATS"] == _lib.UD_POSE_FLOATS and c["UD_GRAD_FLOATS"] == _lib.UD_GRAD_FLOATS
    for k in ("UD_MAX_ATOMS", "UD_MAX_TORSIONS", "UD_MAX_PAIRS", "UD_N_BUCKETS", "UD_MAX_LAUNCHES", "UD_MAX_MODES",
              "UD_NOALLOC_MAX_POSES"):
        assert c[k] == getattr(_lib, k), k


@pytest.mark.gpu
def test_c_host_docks_1iep(tmp_path):
    """A C host: receptor and ligand from the reference's PDBQT fixture, maps,
    dock (exhaustiveness 64, 20 steps), post-processing; the best mode lies in
    the reference test's score range (test_unidock.py: -20 <= E <= -1) and its
    components equal ud_score_terms of its pose."""
    print("[gpu] C host docks 1iep", flush=True)
    exe = _build(tmp_path, link=True)
    g = os.path.join(HERE, "golden")
    r = subprocess.run([exe, "dock", os.path.join(g, "1iep_receptor.pdbqt"), os.path.join(g, "1iep_ligand.pdbqt"),
                        "15.19", "53.903", "16.917", "20"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    print(f"[gpu] C host: {out}", flush=True)
    assert out["n_modes"] >= 1 and -20.0 <= out["energy"] <= -1.0
    assert out["inter"] == pytest.approx(out["rescored_inter"], rel=1e-4, abs=1e-3)
    assert out["intra"] == pytest.approx(out["rescored_intra"], rel=1e-4, abs=1e-3)
