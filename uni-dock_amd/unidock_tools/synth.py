"""Seeded synthetic receptors and ligands for the docking workloads.

Receptor: heavy atoms with protein-like X-Score types placed by seeded random
sequential addition (rejection sampling) in a cube with a minimum spacing.
Ligand: a random tree of heavy atoms with 1.5 A bonds and tetrahedral bond
angles, a chosen number of rotatable (non-terminal) bonds, split into rigid
fragments that form the torsion tree.

The ligand batch layout mirrors how the reference flattens a parsed ligand
(unidock/src/cuda/monte_carlo.cu:2487-2631 and unidock/src/lib/tree.h):
  * atoms are grouped by rigid node, nodes in depth-first pre-order;
  * the root origin is the first root atom (parse_pdbqt.cpp:547-548);
  * a child node's origin is its atom on the rotatable bond and its axis
    points from the parent-side atom to it (parse_pdbqt.cpp:519-521,
    tree.h:101-109);
  * local coordinates are relative to the node origin (parse_pdbqt.cpp:206-211);
  * intramolecular pairs are atom pairs in different rigid nodes, not fixed
    across a rotor axis and more than three bonds apart (model.cpp:592-630).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# X-Score type ids (unidock/src/lib/atom_constants.h:84-120)
XS_C_H, XS_C_P, XS_N_P, XS_N_D, XS_N_A, XS_N_DA = 0, 1, 2, 3, 4, 5
XS_O_P, XS_O_D, XS_O_A, XS_O_DA, XS_S_P = 6, 7, 8, 9, 10
XS_NAMES = ["C_H", "C_P", "N_P", "N_D", "N_A", "N_DA", "O_P", "O_D", "O_A", "O_DA", "S_P",
            "P_P", "F_H", "Cl_H", "Br_H", "I_H", "Si", "At", "Met_D"]

# protein-like type mix: C_H 40%, C_P 22%, N_P 10%, N_D 8%, O_A 12%, O_DA 6%, S_P 2%
TYPE_MIX = ((XS_C_H, 0.40), (XS_C_P, 0.22), (XS_N_P, 0.10), (XS_N_D, 0.08),
            (XS_O_A, 0.12), (XS_O_DA, 0.06), (XS_S_P, 0.02))

TETRA = math.acos(-1.0 / 3.0)  # 109.47 degrees
BOND = 1.5


def _draw_types(rng: np.random.Generator, n: int) -> np.ndarray:
    t = np.array([a for a, _ in TYPE_MIX], dtype=np.int32)
    p = np.array([b for _, b in TYPE_MIX], dtype=np.float64)
    return rng.choice(t, size=n, p=p / p.sum()).astype(np.int32)


@dataclass
class Receptor:
    xyz: np.ndarray      # (n, 3) float64
    xs: np.ndarray       # (n,) int32
    edge: float
    cavity: np.ndarray   # (3,) centre of the largest empty sphere found


def make_receptor(n_atoms: int = 2000, seed: int = 0, edge: float | None = None,
                  min_dist: float = 3.0, box: float = 22.5) -> Receptor:
    """Random sequential addition (exact, over admissible lattice sites; see
    ud_synth_receptor in the C ABI).

    The 40 A cube holds 2,000 atoms at 3.0 A spacing.  Larger receptors use
    edge = 1.03 * 40 * (n / 2000)^(1/3): at equal number density a larger cube
    has relatively less boundary, which would push 8,000 atoms past the
    random-sequential-addition jamming limit.
    """
    from . import _lib
    if edge is None:
        edge = 40.0 if n_atoms <= 2000 else 1.03 * 40.0 * (n_atoms / 2000.0) ** (1.0 / 3.0)
    # 2,000 atoms in a 40 A cube sit close to the jamming limit: a draw that
    # jams early is retried with the next sub-seed (deterministic per seed)
    pts = None
    for sub in range(64):
        try:
            pts = _lib.synth_receptor_xyz(n_atoms, edge, min_dist, seed * 1000 + sub)
            break
        except _lib.UdError:
            continue
    if pts is None:
        raise RuntimeError(f"could not place {n_atoms} atoms at {min_dist} A spacing in a {edge:.1f} A cube")
    rng = np.random.default_rng(seed)
    xs = _draw_types(rng, n_atoms)
    cav = find_cavity(pts, edge, box)
    return Receptor(xyz=pts, xs=xs, edge=edge, cavity=cav)


def find_cavity(xyz: np.ndarray, edge: float, box: float) -> np.ndarray:
    """Centre of the largest empty sphere among candidate points on a 0.5 A
    lattice, restricted so that the docking box stays inside the cube."""
    from scipy.spatial import cKDTree
    lo = box / 2.0
    hi = edge - box / 2.0
    if hi <= lo:
        lo = hi = edge / 2.0
    ax = np.arange(lo, hi + 1e-9, 0.5)
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    d, _ = cKDTree(xyz).query(g)
    return g[int(np.argmax(d))].copy()


@dataclass
class Ligand:
    coords: np.ndarray          # (n, 3) input coordinates, atoms in node order
    xs: np.ndarray              # (n,) int32
    heavy: np.ndarray           # (n,) int32
    parent: np.ndarray          # (nodes,) int32, -1 for the root
    abegin: np.ndarray          # (nodes,) int32
    aend: np.ndarray            # (nodes,) int32
    origin: np.ndarray          # (nodes, 3)
    axis: np.ndarray            # (nodes, 3), zero for the root
    pairs: np.ndarray           # (P, 2) int32, a < b
    bonds: np.ndarray           # (B, 2) int32 (node-ordered atom ids)
    rotatable: np.ndarray       # (B,) bool
    num_tors: float = 0.0
    # AutoDock 4.2 view (ad4_ligand): AutoDock types, partial charges, TORSDOF
    ad: np.ndarray | None = None
    charge: np.ndarray | None = None
    torsdof: float = 0.0

    @property
    def n_atoms(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_tors(self) -> int:
        return int(self.parent.shape[0]) - 1

    @property
    def local(self) -> np.ndarray:
        node_of = np.zeros(self.n_atoms, dtype=np.int64)
        for k in range(len(self.parent)):
            node_of[self.abegin[k]:self.aend[k]] = k
        return self.coords - self.origin[node_of]


def _unit(v):
    return v / np.linalg.norm(v)


def _random_tree_coords(rng, n_atoms, n_tors):
    """Random tree topology + coordinates with 1.5 A bonds, tetrahedral angles."""
    for _attempt in range(200):
        nbr = [[] for _ in range(n_atoms)]
        pos = np.zeros((n_atoms, 3))
        pos[0] = 0.0
        placed = 1
        fail = False
        # prefer extending chains so that enough non-terminal bonds exist
        while placed < n_atoms:
            cands = [i for i in range(placed) if len(nbr[i]) < 4]
            if not cands:
                fail = True
                break
            leaves = [i for i in cands if len(nbr[i]) <= 1]
            ok = False
            for _t in range(30):
                if leaves and rng.random() < 0.6:
                    p = int(rng.choice(leaves))
                else:
                    p = int(rng.choice(cands))
                if not nbr[p]:
                    d = _unit(rng.normal(size=3))
                else:
                    # tetrahedral angle to the bond towards p's first neighbour,
                    # random dihedral about it
                    q = nbr[p][0]
                    b = _unit(pos[p] - pos[q])
                    ref = _unit(np.cross(b, rng.normal(size=3)))
                    phi = rng.uniform(0, 2 * math.pi)
                    perp = math.cos(phi) * ref + math.sin(phi) * np.cross(b, ref)
                    d = math.cos(math.pi - TETRA) * b + math.sin(math.pi - TETRA) * perp
                    # keep roughly tetrahedral to p's other bonds
                    if any(np.dot(d, _unit(pos[o] - pos[p])) > math.cos(math.radians(100.0)) for o in nbr[p]):
                        continue
                newp = pos[p] + BOND * d
                # clash test against atoms more than two bonds away
                near = {p}
                for o in nbr[p]:
                    near.add(o)
                dd = ((pos[:placed] - newp) ** 2).sum(1)
                bad = False
                for j in np.nonzero(dd < 2.6 ** 2)[0]:
                    if int(j) not in near:
                        bad = True
                        break
                if bad:
                    continue
                pos[placed] = newp
                nbr[p].append(placed)
                nbr[placed].append(p)
                placed += 1
                ok = True
                break
            if not ok:
                fail = True
                break
        if fail:
            continue
        internal = [(i, j) for i in range(n_atoms) for j in nbr[i] if i < j and len(nbr[i]) > 1 and len(nbr[j]) > 1]
        if len(internal) >= n_tors or _attempt > 150:
            return pos, nbr, internal
    raise RuntimeError("ligand generator failed")


def make_ligand(n_atoms: int, n_tors: int, seed: int) -> Ligand:
    rng = np.random.default_rng(seed)
    pos, nbr, internal = _random_tree_coords(rng, n_atoms, n_tors)
    n_tors = min(n_tors, len(internal))
    rot_idx = rng.choice(len(internal), size=n_tors, replace=False) if n_tors else np.zeros(0, int)
    rot = {tuple(internal[i]) for i in rot_idx}
    xs = _draw_types(rng, n_atoms)
    # rigid fragments: connected components without rotatable bonds
    frag = -np.ones(n_atoms, dtype=np.int64)
    nf = 0
    for s in range(n_atoms):
        if frag[s] >= 0:
            continue
        stack = [s]
        frag[s] = nf
        while stack:
            a = stack.pop()
            for b in nbr[a]:
                if frag[b] < 0 and (min(a, b), max(a, b)) not in rot:
                    frag[b] = nf
                    stack.append(b)
        nf += 1
    # fragment graph; root = fragment minimising the tree height
    fadj = [[] for _ in range(nf)]
    for (a, b) in rot:
        fa, fb = frag[a], frag[b]
        fadj[fa].append((fb, a, b))
        fadj[fb].append((fa, b, a))

    def height(r):
        best, seen, st = 0, {r}, [(r, 0)]
        while st:
            f, h = st.pop()
            best = max(best, h)
            for (g, _, _) in fadj[f]:
                if g not in seen:
                    seen.add(g)
                    st.append((g, h + 1))
        return best
    root = min(range(nf), key=lambda f: (height(f), -int((frag == f).sum()), f))
    # pre-order DFS over fragments; each child records (parent frag, parent-side atom, child-side atom)
    order = []
    info = {root: (-1, -1, -1)}

    def dfs(f):
        order.append(f)
        for (g, a_here, b_there) in sorted(fadj[f], key=lambda t: t[2]):
            if g not in info:
                info[g] = (f, a_here, b_there)
                dfs(g)
    dfs(root)
    # atom order: node by node; root starts with its lowest atom, child nodes
    # with their rotatable-bond atom
    new_of_old = -np.ones(n_atoms, dtype=np.int64)
    old_order = []
    abegin, aend = [], []
    for f in order:
        atoms = sorted(int(i) for i in np.nonzero(frag == f)[0])
        first = info[f][2] if info[f][0] >= 0 else atoms[0]
        atoms.remove(first)
        atoms = [first] + atoms
        abegin.append(len(old_order))
        for a in atoms:
            new_of_old[a] = len(old_order)
            old_order.append(a)
        aend.append(len(old_order))
    old_order = np.array(old_order)
    coords = pos[old_order]
    xs_n = xs[old_order].astype(np.int32)
    fidx = {f: k for k, f in enumerate(order)}
    nn = len(order)
    parent = np.full(nn, -1, dtype=np.int32)
    origin = np.zeros((nn, 3))
    axis = np.zeros((nn, 3))
    axis_begin_atom = np.full(nn, -1, dtype=np.int64)
    for k, f in enumerate(order):
        pf, a_par, b_child = info[f]
        if pf < 0:
            origin[k] = coords[0]
            continue
        parent[k] = fidx[pf]
        origin[k] = pos[b_child]
        axis[k] = _unit(pos[b_child] - pos[a_par])
        axis_begin_atom[k] = new_of_old[a_par]
    # bonds in the new numbering
    bonds, rflags = [], []
    for a in range(n_atoms):
        for b in nbr[a]:
            if a < b:
                bonds.append((new_of_old[a], new_of_old[b]))
                rflags.append((a, b) in rot)
    bonds = np.array(bonds, dtype=np.int32).reshape(-1, 2)
    rflags = np.array(rflags, dtype=bool)
    # topological distance (bonds) for 1-2/1-3/1-4 exclusion (model.cpp:601-603)
    adj = [[] for _ in range(n_atoms)]
    for a, b in bonds:
        adj[a].append(int(b))
        adj[b].append(int(a))
    node_of = np.zeros(n_atoms, dtype=np.int64)
    for k in range(nn):
        node_of[abegin[k]:aend[k]] = k
    pairs = []
    for i in range(n_atoms):
        dist = {i: 0}
        frontier = [i]
        for d in (1, 2, 3):
            nxt = []
            for a in frontier:
                for b in adj[a]:
                    if b not in dist:
                        dist[b] = d
                        nxt.append(b)
            frontier = nxt
        for j in range(i + 1, n_atoms):
            if node_of[i] == node_of[j]:
                continue  # same rigid body: DISTANCE_FIXED
            # rotor-axis atom is fixed relative to the child node (parse_pdbqt.cpp:456-467)
            if axis_begin_atom[node_of[j]] == i or axis_begin_atom[node_of[i]] == j:
                continue
            if j in dist:
                continue
            pairs.append((i, j))
    pairs = np.array(pairs, dtype=np.int32).reshape(-1, 2)
    lig = Ligand(coords=coords, xs=xs_n, heavy=np.ones(n_atoms, dtype=np.int32), parent=parent,
                 abegin=np.array(abegin, dtype=np.int32), aend=np.array(aend, dtype=np.int32),
                 origin=origin, axis=axis, pairs=pairs, bonds=bonds, rotatable=rflags)
    lig.num_tors = float(num_tors_vina(lig))
    return lig


def num_tors_vina(lig: Ligand) -> float:
    """conf_independent_inputs::num_tors (conf_independent.cpp:79-103): half a
    count per heavy endpoint of each rotatable bond whose other end has more
    than one heavy neighbour."""
    n = lig.n_atoms
    heavy_deg = np.zeros(n, dtype=np.int64)
    for (a, b) in lig.bonds:
        if lig.heavy[b]:
            heavy_deg[a] += 1
        if lig.heavy[a]:
            heavy_deg[b] += 1
    acc = 0.0
    for (a, b), r in zip(lig.bonds, lig.rotatable):
        if not r:
            continue
        if lig.heavy[a] and lig.heavy[b] and heavy_deg[b] > 1:
            acc += 0.5
        if lig.heavy[b] and lig.heavy[a] and heavy_deg[a] > 1:
            acc += 0.5
    return acc


# ---------------------------------------------------------------- AutoDock 4.2
# AutoDock type indices (atom_constants.h:46-81)
AD_C, AD_A, AD_N, AD_O, AD_P, AD_S, AD_H, AD_F, AD_I, AD_NA, AD_OA, AD_SA, AD_HD = range(13)
# synthetic AutoDock type per X-Score type: acceptor N -> NA, O -> OA, S -> SA
_XS_TO_AD = {XS_C_H: AD_C, XS_C_P: AD_C, XS_N_P: AD_N, XS_N_D: AD_N, XS_N_A: AD_NA, XS_N_DA: AD_NA,
             XS_O_P: AD_OA, XS_O_D: AD_OA, XS_O_A: AD_OA, XS_O_DA: AD_OA, XS_S_P: AD_SA}
# partial-charge mean and spread per AutoDock type (Gasteiger-like magnitudes)
_AD_CHARGE = {AD_C: (0.05, 0.08), AD_A: (0.0, 0.04), AD_N: (-0.25, 0.08), AD_NA: (-0.35, 0.08),
              AD_OA: (-0.40, 0.08), AD_SA: (-0.15, 0.05), AD_HD: (0.22, 0.04)}


def _ad_charges(rng, ad):
    return np.array([rng.normal(*_AD_CHARGE[int(t)]) for t in ad], dtype=np.float64)


def ad4_ligand(lig: Ligand, seed: int) -> Ligand:
    """The AutoDock 4.2 view of a synthetic ligand: AutoDock types from the
    X-Score types (30 % of the carbons aromatic "A"), half of the terminal
    atoms bonded to an N or O turned into polar hydrogens "HD" (not heavy),
    seeded partial charges and TORSDOF = torsions.  Coordinates, tree and
    pairs are the ligand's own (every atom is AutoDock-typed)."""
    from dataclasses import replace
    rng = np.random.default_rng(seed)
    ad = np.array([_XS_TO_AD.get(int(t), AD_C) for t in lig.xs], dtype=np.int32)
    ad[(ad == AD_C) & (rng.random(lig.n_atoms) < 0.3)] = AD_A
    deg = np.zeros(lig.n_atoms, dtype=np.int64)
    for a, b in lig.bonds:
        deg[a] += 1
        deg[b] += 1
    heavy = lig.heavy.copy()
    xs = lig.xs.copy()
    polar = {AD_N, AD_NA, AD_OA}
    for a, b in lig.bonds:
        for h, x in ((a, b), (b, a)):
            if deg[h] == 1 and int(ad[x]) in polar and ad[h] != AD_HD and rng.random() < 0.5:
                ad[h] = AD_HD
                heavy[h] = 0
                xs[h] = -1
    return replace(lig, ad=ad, charge=_ad_charges(rng, ad), torsdof=float(lig.n_tors), heavy=heavy, xs=xs)


def ad4_receptor(rec: Receptor, seed: int, hd_fraction: float = 0.08):
    """AutoDock types and partial charges for a synthetic receptor (X-Score
    types mapped as for ligands, carbons 30 % "A", a fraction of the atoms
    polar hydrogens "HD"): (ad int32 [n], charge float64 [n])."""
    rng = np.random.default_rng(seed)
    n = len(rec.xs)
    ad = np.array([_XS_TO_AD.get(int(t), AD_C) for t in rec.xs], dtype=np.int32)
    ad[(ad == AD_C) & (rng.random(n) < 0.3)] = AD_A
    ad[rng.random(n) < hd_fraction] = AD_HD
    return ad, _ad_charges(rng, ad)


@dataclass
class LigandBatch:
    """Flat arrays in the C-ABI layout (see include/unidock_amd.h)."""
    ligands: list = field(default_factory=list)

    def arrays(self) -> dict:
        ls = self.ligands
        na = np.array([l.n_atoms for l in ls])
        nn = np.array([len(l.parent) for l in ls])
        npairs = np.array([len(l.pairs) for l in ls])
        off = lambda a: np.concatenate([[0], np.cumsum(a)]).astype(np.int32)
        return dict(
            n_ligands=len(ls),
            atom_offset=off(na), node_offset=off(nn), pair_offset=off(npairs),
            atom_local=np.concatenate([l.local for l in ls]).astype(np.float32).reshape(-1, 3),
            atom_xs=np.concatenate([l.xs for l in ls]).astype(np.int32),
            atom_heavy=np.concatenate([l.heavy for l in ls]).astype(np.int32),
            node_parent=np.concatenate([l.parent for l in ls]).astype(np.int32),
            node_abegin=np.concatenate([l.abegin for l in ls]).astype(np.int32),
            node_aend=np.concatenate([l.aend for l in ls]).astype(np.int32),
            node_origin=np.concatenate([l.origin for l in ls]).astype(np.float32).reshape(-1, 3),
            node_axis=np.concatenate([l.axis for l in ls]).astype(np.float32).reshape(-1, 3),
            pairs=np.concatenate([l.pairs for l in ls]).astype(np.int32).reshape(-1, 2),
            num_tors=np.array([l.num_tors for l in ls], dtype=np.float32),
            **({} if any(l.ad is None for l in ls) else dict(
                atom_ad=np.concatenate([l.ad for l in ls]).astype(np.int32),
                atom_charge=np.concatenate([l.charge for l in ls]).astype(np.float32),
                torsdof=np.array([l.torsdof for l in ls], dtype=np.float32))),
        )


def make_ligand_set(n: int, seed: int, atoms=(10, 40), tors=(0, 10), dist: str = "uniform") -> list:
    """Ligand sizes: 'uniform' draws atoms ~ U[a0, a1], torsions ~ U[t0, t1];
    'screening' draws atoms ~ N(28, 7) clipped to [8, 60] and torsions ~
    Poisson(5) clipped to [0, 15]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        if dist == "uniform":
            na = int(rng.integers(atoms[0], atoms[1] + 1))
            nt = int(rng.integers(tors[0], tors[1] + 1))
        else:
            na = int(np.clip(round(rng.normal(28, 7)), 8, 60))
            nt = int(np.clip(rng.poisson(5), 0, 15))
        out.append(make_ligand(na, nt, seed=int(rng.integers(0, 2**31 - 1))))
    return out
