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
    """Pairs (model::initialize_pairs, model.cpp:592-630): atoms of variable
    mobility, more than 3 bonds apart, both typed.  Mobility as the reference
    builds it from the BRANCH records (postprocess_branch,
    parse_pdbqt.cpp:498-533): the second atom of "BRANCH a b" lies on the
    rotation axis and is stored with the parent's atoms; a and b are fixed
    relative to every atom stored with the branch."""
    node = np.zeros(lig.n_atoms, dtype=int)
    for k, (a, b) in enumerate(zip(lig.abegin, lig.aend)):
        node[a:b] = k
    idx = {int(s): i for i, s in enumerate(lig.serial)}
    branches = [tuple(int(x) for x in ln.split()[1:3]) for ln in LIGAND.splitlines() if ln.startswith("BRANCH")]
    body = node.copy()
    axis = {}
    for a, b in branches:
        ia, ib = idx[a], idx[b]
        k = node[ib]  # the branch's own rigid body
        body[ib] = node[ia]
        axis[k] = {ia, ib}
    assert len(axis) == len(lig.parent) - 1

    def fixed(i, j):
        return body[i] == body[j] or i in axis.get(body[j], ()) or j in axis.get(body[i], ())

    cov = np.where(lig.heavy == 1, 0.76, 0.37)  # C / N / O radii 0.73-0.77, H 0.37
    bonds = _bonds(lig.coords, lig.heavy, cov)

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
    pose = np.zeros(_lib.UD_POSE_FLOATS)
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


# ---------------------------------------------------------------- SDF (unidock/test/ligands/1a30_ligand.sdf)
SDF = open(os.path.join(GOLDEN, "1a30_ligand.sdf")).read()


def _sdf_fields(text):
    """fragInfo, torsionInfo and atomInfo of the fixture, read directly."""
    out, cur = {}, None
    for ln in text.splitlines():
        if ln.startswith(">"):
            cur = ln.split("<")[1].split(">")[0]
            out[cur] = []
        elif cur is not None:
            if not ln.strip():
                cur = None
            else:
                out[cur].append(ln)
    frags = [[int(x) for x in ln.split()] for ln in out["fragInfo"]]
    tors = [[int(x) for x in ln.split()] for ln in out["torsionInfo"]]
    info = {int(ln[:3].split()[0]): ln[13:].strip() for ln in out["atomInfo"]}
    return frags, tors, info


@pytest.fixture(scope="module")
def sdf_lig():
    s = _lib.LigandSet()
    assert s.add_sdf(SDF) == (0, 1)
    return s.ligands()[0]


def test_sdf_tree_from_fragments(sdf_lig):
    """The tree follows fragInfo / torsionInfo: fragment 0 is the root, each
    torsion opens the fragment on its far side; atoms in no fragment (here the
    non-polar hydrogens) are not part of the ligand."""
    frags, tors, info = _sdf_fields(SDF)
    lig = sdf_lig
    assert lig.n_atoms == sum(len(f) for f in frags) == 31
    assert lig.n_tors == len(tors) == 13
    nodes = [set(lig.serial[a:b].tolist()) for a, b in zip(lig.abegin, lig.aend)]
    assert sorted(map(sorted, nodes)) == sorted(map(sorted, frags))
    assert nodes[0] == set(frags[0])
    # every torsion joins a node to its parent, the far atom being in the child
    frag_of = {a: i for i, f in enumerate(frags) for a in f}
    node_of_frag = {i: next(k for k, n in enumerate(nodes) if n == set(f)) for i, f in enumerate(frags)}
    for a, b, fa, fb in tors:
        assert frag_of[a] == fa and frag_of[b] == fb
        ka, kb = node_of_frag[fa], node_of_frag[fb]
        assert lig.parent[kb] == ka or lig.parent[ka] == kb
    # pre-order: parents before children
    assert all(lig.parent[k] < k for k in range(1, len(lig.parent)))
    assert set(lig.serial.tolist()) == set(info)  # atomInfo lists exactly the kept atoms


def test_sdf_types_from_atominfo(sdf_lig):
    frags, tors, info = _sdf_fields(SDF)
    got = dict(zip(sdf_lig.serial.tolist(), sdf_lig.xs.tolist()))
    for serial, ad in info.items():
        if ad in ("HD", "H"):
            assert got[serial] == -1
        elif ad == "OA":
            assert got[serial] == O_A  # carboxylate / carbonyl oxygens: acceptors
        elif ad == "N":
            assert got[serial] == N_D  # each nitrogen carries a polar hydrogen
        else:
            assert ad == "C" and got[serial] in (C_H, C_P)
    # carbons bonded to N / O are polar
    assert got[11] == C_P and got[3] == C_P and got[22] == C_H and got[24] == C_H


def test_sdf_identity_pose_reproduces_input(sdf_lig):
    lig = sdf_lig
    o = Oracle(64, 0)
    o.set_grid(lig.origin[0], np.array([30.0, 30.0, 30.0]), 0.375)
    o.add_ligand(lig)
    pose = np.zeros(_lib.UD_POSE_FLOATS)
    pose[0:3] = lig.origin[0]
    pose[3] = 1.0
    _, _, coords, _ = o.eval_deriv(0, pose)
    np.testing.assert_allclose(coords, lig.coords, atol=2e-4)
    # the root is atom 11, the first (and only) atom of fragment 0
    np.testing.assert_allclose(lig.origin[0], lig.coords[list(lig.serial).index(11)])


def test_sdf_pair_rule(sdf_lig):
    """As for PDBQT, with the branch axis atoms from torsionInfo: the far atom
    of each torsion is stored with the near fragment's atoms."""
    frags, tors, info = _sdf_fields(SDF)
    lig = sdf_lig
    idx = {int(s): i for i, s in enumerate(lig.serial)}
    node = np.zeros(lig.n_atoms, dtype=int)
    for k, (a, b) in enumerate(zip(lig.abegin, lig.aend)):
        node[a:b] = k
    body = node.copy()
    axis = {}
    for a, b, fa, fb in tors:
        ia, ib = idx[a], idx[b]
        if lig.parent[node[ia]] == node[ib]:  # oriented parent -> child
            ia, ib = ib, ia
        body[ib] = node[ia]
        axis[node[ib]] = {ia, ib}
    cov = np.where(lig.heavy == 1, 0.76, 0.37)
    bonds = _bonds(lig.coords, lig.heavy, cov)

    def fixed(i, j):
        return body[i] == body[j] or i in axis.get(body[j], ()) or j in axis.get(body[i], ())

    def within3(i):
        seen, front = {i}, {i}
        for _ in range(3):
            front = {v for u in front for v in bonds[u]} - seen
            seen |= front
        return seen

    expect = [(i, j) for i in range(lig.n_atoms) for j in range(i + 1, lig.n_atoms)
              if not fixed(i, j) and j not in within3(i) and lig.xs[i] >= 0 and lig.xs[j] >= 0]
    assert sorted(map(tuple, lig.pairs.tolist())) == expect


def test_sdf_write_round_trip(sdf_lig):
    s = _lib.LigandSet()
    s.add_sdf(SDF)
    lig = sdf_lig
    modes = [dict(energy=-7.25, inter=-9.0, intra=-0.75, rmsd_lb=0.0, rmsd_ub=0.0, unbound=-0.5),
             dict(energy=-6.5, inter=-8.0, intra=-0.8, rmsd_lb=1.25, rmsd_ub=2.5, unbound=-0.5)]
    coords = np.stack([lig.coords, lig.coords + np.float32([0.5, -1.0, 2.0])]).astype(np.float32)
    text = s.write_sdf(0, modes, coords, how_many=9, energy_range=3.0)
    recs = text.split("$$$$\n")
    assert len(recs) == 3 and recs[-1] == ""
    first = recs[0].splitlines()
    k = first.index("> <Uni-Dock RESULT> ")
    assert first[k + 1] == "ENERGY=   -7.250  LOWER_BOUND=    0.000  UPPER_BOUND=    0.000"
    assert first[k + 2] == "INTER + INTRA:          -9.750"
    assert first[k + 5] == "UNBOUND:                -0.500"
    # the input record comes back unchanged with the input coordinates
    src = [ln for ln in SDF.splitlines() if "$$$$" not in ln]
    assert first[:k] == src
    # the second record parses back with its shifted coordinates
    t = _lib.LigandSet()
    t.add_sdf(recs[1])
    back = t.ligands()[0]
    np.testing.assert_allclose(back.coords, coords[1], atol=6e-5)
    assert back.pairs.tolist() == lig.pairs.tolist() and back.xs.tolist() == lig.xs.tolist()
    # atoms outside the ligand (non-polar hydrogens) keep their input coordinates
    assert recs[1].splitlines()[4 + 29] == src[4 + 29]
    with pytest.raises(_lib.UdError) as ei:
        s.write_pdbqt(0, modes, coords)
    assert ei.value.code == _lib.UD_ERR_STATE


def test_sdf_without_fragments_is_rigid():
    text = SDF.split(">  <fragInfo>")[0] + "$$$$\n"
    s = _lib.LigandSet()
    s.add_sdf(text)
    lig = s.ligands()[0]
    assert lig.n_atoms == 49 and lig.n_tors == 0 and lig.serial.tolist() == list(range(1, 50))
    assert lig.pairs.size == 0


def test_sdf_multiple_records_and_errors():
    s = _lib.LigandSet()
    assert s.add_sdf(SDF + SDF) == (0, 2)
    assert len(s) == 2
    bad = SDF.replace("11 10 0 6", "11 10 0 60")
    with pytest.raises(_lib.UdError) as ei:
        s.add_sdf(SDF + bad)
    assert "record 2" in str(ei.value)
    assert len(s) == 2  # all or nothing
    with pytest.raises(_lib.UdError):
        s.add_sdf(SDF.replace(" 49 48  0", " xx 48  0"))
