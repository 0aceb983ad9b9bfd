"""CPU-side checks of the C ABI: the library loads, exports every declared
entry point, rejects bad input with status codes, and its host-only helpers
work.  No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import ROOT
from unidock_tools import _lib


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "unidock_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(ud_[a-z_0-9]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 18
    for s in syms:
        assert hasattr(L, s), s


def test_default_search_params():
    L = _lib.load()
    p = _lib.UdSearchParams()
    L.ud_default_search_params(ctypes.byref(p))
    assert p.exhaustiveness == 8 and p.global_steps == 20 and p.seed == 181129
    assert list(p.hunt_cap) == [10, 10, 10] and list(p.authentic_v) == [1000, 1000, 1000]
    assert p.temperature == pytest.approx(1.2) and p.mutation_amplitude == 2
    assert p.chains_per_group == 0


def test_null_and_invalid_arguments_return_codes():
    L = _lib.load()
    assert L.ud_create(0, None, None) == _lib.UD_ERR_INVALID
    assert L.ud_destroy(None) == _lib.UD_ERR_INVALID
    assert L.ud_set_scoring(None, 0, None, 0) == _lib.UD_ERR_INVALID
    assert L.ud_dock(None, None, None, None, None) == _lib.UD_ERR_INVALID
    assert L.ud_last_error(None) == b"null context"
    out = np.zeros(3)
    assert L.ud_synth_receptor(0, 10.0, 3.0, 1, 10, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == \
        _lib.UD_ERR_INVALID


def test_synthetic_receptor_spacing_and_determinism():
    a = _lib.synth_receptor_xyz(300, 24.0, 3.0, 12)
    b = _lib.synth_receptor_xyz(300, 24.0, 3.0, 12)
    assert np.array_equal(a, b)
    assert a.min() >= 0 and a.max() < 24.0
    d = np.linalg.norm(a[:, None] - a[None], axis=-1)
    d[np.arange(300), np.arange(300)] = 1e9
    assert d.min() >= 3.0


def test_jammed_receptor_reports_capacity():
    out = np.zeros(3 * 2000)
    rc = _lib.load().ud_synth_receptor(2000, 20.0, 3.0, 1, 10**7, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == _lib.UD_ERR_CAPACITY


def test_ligand_layout_from_generator():
    from unidock_tools import synth
    lig = synth.make_ligand(24, 5, seed=1)
    assert lig.n_atoms == 24 and lig.n_tors == 5 and lig.num_tors == 5.0
    assert lig.parent[0] == -1 and all(lig.parent[k] < k for k in range(1, len(lig.parent)))
    assert lig.abegin[0] == 0 and lig.aend[-1] == 24
    np.testing.assert_allclose(np.linalg.norm(lig.axis[1:], axis=1), 1.0, atol=1e-12)
    # no pair inside a rigid node and none within three bonds
    node_of = np.repeat(np.arange(len(lig.parent)), lig.aend - lig.abegin)
    assert all(node_of[a] != node_of[b] for a, b in lig.pairs)
    arr = synth.LigandBatch([lig]).arrays()
    assert arr["atom_local"].shape == (24, 3) and arr["pairs"].shape[1] == 2
