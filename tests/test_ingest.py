"""PDBQT ingestion (ud_ligset_*, ud_receptor_from_pdbqt) on the reference's own
test fixtures (unidock/test/ligands/1iep_ligand.pdbqt and
unidock/test/receptor/1iep_receptor.pdbqt, copied to tests/golden/)."""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from unidock_tools import _lib

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
LIGAND = open(os.path.join(GOLDEN, "1iep_ligand.pdbqt")).read()
RECEPTOR = open(os.path.join(GOLDEN, "1iep_receptor.pdbqt")).read()

# X-Score type indices (atom_constants.h:84-120)
C_H, C_P, N_P, N_D, N_A, N_DA, O_P, O_D, O_A, O_DA, S_P = range(11)


@pytest.fixture(scope="module")
def lig():
    s = _lib.LigandSet()
    assert s.add_pdbqt(LIGAND) == 0
    return s.ligands()[0]


def test_tree_and_counts(lig):
    # ROOT of 7 atoms and 7 BRANCH blocks in pre-order (TORSDOF 7)
    assert lig.n_atoms == 41 and lig.n_tors == 7
    assert int(lig.heavy.sum()) == 37
    assert lig.parent.tolist() == [-1, 0, 1, 2, 0, 4, 5, 6]
    assert lig.abegin.tolist() == [0, 7, 9, 15, 21, 25, 31, 32]
    assert lig.aend.tolist() == [7, 9, 15, 21, 25, 31, 32, 41]
    # every torsion is between heavy atoms with further heavy neighbours
    assert lig.num_tors == 7.0
    # the file lists atoms in pre-order already
    assert lig.serial.tolist() == list(range(1, 42))


def test_xscore_types_follow_the_chemistry(lig):
    # carbons bonded to N / O are polar; N bonded to a polar H are donors,
    # NA without one acceptors; the carbonyl O (OA) is an acceptor
    expect = {1: C_P, 2: C_H, 3: C_H, 4: C_H, 5: C_P, 6: C_H, 7: C_H, 8: N_D, 9: -1, 10: C_P, 11: C_P, 12: C_H,
              13: N_A, 14: C_P, 15: N_A, 16: C_H, 17: C_H, 18: C_H, 19: C_P, 20: N_A, 21: C_P, 22: N_D, 23: C_P,
              24: -1, 25: O_A, 32: C_P, 33: N_D, 34: C_P, 35: C_P, 36: -1, 37: C_P, 38: C_P, 39: N_D, 40: C_P,
              41: -1}
    for k in range(26, 32):
        expect[k] = C_H  # the central benzene ring
    got = dict(zip(lig.serial.tolist(), lig.xs.tolist()))
    assert got == expect


def _bonds(coords, heavy, ad_cov):
    n = len(coords)
    d2 = np.sum((coords[:, None] - coords[None]) ** 2, axis=2)
    return [[j for j in range(n) if j != i and d2[i, j] < (1.1 * (ad_cov[i] + ad_cov[j])) ** 2] for i in range(n)]


def test_pair_rule(lig):
    """Pairs: atoms in different rigid bodies (a branch's first axis atom
    belongs with the branch), more than 3 bonds apart, both typed."""
    node = np.zeros(lig.n_atoms, dtype=int)
    for k, (a, b) in enumerate(zip(lig.abegin, lig.aend)):
        node[a:b] = k
    cov = np.where(lig.heavy == 1, 0.76, 0.37)  # C / N / O radii 0.73-0.77, H 0.37
    bonds = _bonds(lig.coords, lig.heavy, cov)
    axis_begin = {}
    for k in range(1, len(lig.parent)):
        first = lig.abegin[k]
        axis_begin[k] = min((j for j in bonds[first] if node[j] == lig.parent[k]),
                            key=lambda j: np.sum((lig.coords[j] - lig.coords[first]) ** 2))

    def fixed(i, j):
        return node[i] == node[j] or axis_begin.get(node[j]) == i or axis_begin.get(node[i]) == j

    def within3(i):
        seen, front = {i}, {i}
        for _ in range(3):
            front = {v for u in front for v in bonds[u]} - seen
            seen |= front
        return seen

    expect = []
    for i in range(lig.n_atoms):
        near = within3(i)
        for j in range(i + 1, lig.n_atoms):
            if not fixed(i, j) and j not in near and lig.xs[i] >= 0 and lig.xs[j] >= 0:
                expect.append((i, j))
    assert sorted(map(tuple, lig.pairs.tolist())) == expect


def test_identity_pose_reproduces_input(lig):
    """Root at its origin, identity orientation, zero torsions: forward
    kinematics of the oracle gives back the input coordinates."""
    o = Oracle(64, 0)
    o.set_grid(lig.origin[0], np.array([30.0, 30.0, 30.0]), 0.375)
    o.add_ligand(lig)
    pose = np.zeros(39)
    pose[0:3] = lig.origin[0]
    pose[3] = 1.0
    _, _, coords, _ = o.eval_deriv(0, pose)
    np.testing.assert_allclose(coords, lig.coords, atol=2e-4)


def test_receptor_types():
    xyz, xs = _lib.receptor_from_pdbqt(RECEPTOR)
    heavy_lines = [ln for ln in RECEPTOR.splitlines()
                   if ln.startswith(("ATOM", "HETATM")) and ln[77:79].strip() not in ("H", "HD")]
    assert len(xyz) == len(heavy_lines) == 2229
    assert set(np.unique(xs)) <= {C_H, C_P, N_P, N_D, N_A, N_DA, O_P, O_D, O_A, O_DA, S_P}
    # backbone N carries a polar H in this file: donors dominate the nitrogens
    counts = np.bincount(xs, minlength=11)
    assert counts[N_D] > 10 * counts[N_P] and counts[C_P] > 0 and counts[C_H] > 0


def test_errors_are_reported():
    s = _lib.LigandSet()
    with pytest.raises(_lib.UdError) as ei:
        s.add_pdbqt("ROOT\nATOM      1  C1  LIG L   1      15.822  54.641  12.322  1.00  0.00     0.038 Qq\nENDROOT\n")
    assert ei.value.code == _lib.UD_ERR_INVALID and "Qq" in str(ei.value)
    with pytest.raises(_lib.UdError):
        s.add_pdbqt("BRANCH 1 2\n")
    assert len(s) == 0


def test_write_pdbqt_round_trip(lig):
    """Modes written as PDBQT (MODEL / REMARK VINA RESULT / input lines with
    new coordinates / ENDMDL): with the input coordinates the atom lines come
    back unchanged, the count follows how_many and energy_range, and a written
    model parses back to the same ligand."""
    s = _lib.LigandSet()
    s.add_pdbqt(LIGAND)
    modes = [dict(energy=-9.5, inter=-12.0, intra=-0.5, rmsd_lb=0.0, rmsd_ub=0.0, unbound=-0.4),
             dict(energy=-8.0, inter=-10.0, intra=-0.6, rmsd_lb=1.5, rmsd_ub=2.5, unbound=-0.4),
             dict(energy=-5.0, inter=-7.0, intra=-0.6, rmsd_lb=3.0, rmsd_ub=4.0, unbound=-0.4)]
    coords = np.stack([lig.coords] * 3).astype(np.float32)
    shifted = coords.copy()
    shifted[1] += np.float32([1.0, -2.0, 0.5])
    text = s.write_pdbqt(0, modes, shifted, how_many=9, energy_range=3.0)
    blocks = text.split("ENDMDL\n")
    assert len(blocks) == 3 and blocks[-1] == ""  # the third mode is 4.5 above the best
    first = blocks[0].splitlines()
    assert first[0] == "MODEL 1"
    assert first[1] == "REMARK VINA RESULT:    -9.500      0.000      0.000"
    assert first[2] == "REMARK INTER + INTRA:         -12.500"
    body = [ln for ln in first if not ln.startswith(("MODEL", "REMARK VINA", "REMARK INTER", "REMARK INTRA",
                                                     "REMARK UNBOUND"))]
    src = [ln.rstrip("\n") for ln in LIGAND.splitlines()]
    assert body == src
    assert len(s.write_pdbqt(0, modes, shifted, how_many=1).split("ENDMDL\n")) == 2
    # the second model parses back with its shifted coordinates
    second = "\n".join(ln for ln in blocks[1].splitlines() if not ln.startswith(("MODEL", "REMARK VINA", "REMARK I",
                                                                                  "REMARK U")))
    t = _lib.LigandSet()
    t.add_pdbqt(second)
    back = t.ligands()[0]
    np.testing.assert_allclose(back.coords, shifted[1], atol=6e-4)
    assert back.pairs.tolist() == lig.pairs.tolist() and back.xs.tolist() == lig.xs.tolist()
