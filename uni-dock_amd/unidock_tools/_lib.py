"""ctypes binding of the C ABI (include/unidock_amd.h).

The binding loads the in-tree libunidock_amd.so and raises if it is missing:
there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _build

UD_OK, UD_ERR_INVALID, UD_ERR_CAPACITY, UD_ERR_HIP, UD_ERR_STATE, UD_ERR_NOMEM = range(6)
UD_SF_VINA, UD_SF_VINARDO, UD_SF_AD4 = 0, 1, 2
UD_AD4_MAP_ELEC, UD_AD4_MAP_DSOLV = 35, 36  # AD4.2 electrostatic / desolvation map indices
UD_GRID_F32, UD_GRID_F16 = 0, 1
UD_MAX_ATOMS = 300
UD_MAX_TORSIONS = 48
UD_MAX_PAIRS = 8128
UD_POSE_FLOATS = 7 + UD_MAX_TORSIONS
UD_GRAD_FLOATS = 6 + UD_MAX_TORSIONS
UD_WEIGHT_FLOATS = 7  # dE/dw per pose, in the order of the scoring weights
UD_N_BUCKETS = 8
UD_MAX_LAUNCHES = 2 * UD_N_BUCKETS
UD_MAX_MODES = 64  # output modes per ligand of ud_postprocess
UD_NOALLOC_MAX_POSES = 65536
BUCKET_ATOMS = (16, 32, 48, 64, 96, 128, 192, 300)

VINA_WEIGHTS = (-0.035579, -0.005156, 0.840245, -0.035069, -0.587439, 50.0, 0.05846)  # vina.h:129-132
VINARDO_WEIGHTS = (-0.045, 0.8, -0.035, -0.600, 50.0, 0.05846)                      # vina.h:133-135
AD4_WEIGHTS = (0.1662, 0.1209, 0.1406, 0.1322, 50.0, 0.2983)                         # vina.h:136-138
DEFAULT_WEIGHTS = {UD_SF_VINA: VINA_WEIGHTS, UD_SF_VINARDO: VINARDO_WEIGHTS, UD_SF_AD4: AD4_WEIGHTS}
# AutoDock 4.2 type names by index (atom_constants.h:46-81)
AD_TYPE_NAMES = ("C", "A", "N", "O", "P", "S", "H", "F", "I", "NA", "OA", "SA", "HD", "Mg", "Mn", "Zn", "Ca", "Fe",
                 "Cl", "Br", "Si", "At", "G0", "G1", "G2", "G3", "CG0", "CG1", "CG2", "CG3", "W", "OXA", "NXA",
                 "OXD", "NXD")

_c_int_p = ctypes.POINTER(ctypes.c_int32)
_c_float_p = ctypes.POINTER(ctypes.c_float)


class UdLigandBatch(ctypes.Structure):
    _fields_ = [("n_ligands", ctypes.c_int32),
                ("atom_offset", _c_int_p), ("node_offset", _c_int_p), ("pair_offset", _c_int_p),
                ("atom_local", _c_float_p), ("atom_xs", _c_int_p), ("atom_heavy", _c_int_p),
                ("node_parent", _c_int_p), ("node_abegin", _c_int_p), ("node_aend", _c_int_p),
                ("node_origin", _c_float_p), ("node_axis", _c_float_p),
                ("pairs", _c_int_p), ("num_tors", _c_float_p),
                ("atom_ad", _c_int_p), ("atom_charge", _c_float_p), ("torsdof", _c_float_p)]


class UdLaunchInfo(ctypes.Structure):
    _fields_ = [("atom_cap", ctypes.c_int32), ("lanes", ctypes.c_int32), ("n_ligands", ctypes.c_int32),
                ("n_workgroups", ctypes.c_int32), ("pair_buffer", ctypes.c_int32), ("lds_bytes", ctypes.c_int32),
                ("max_torsions", ctypes.c_int32), ("ms", ctypes.c_double)]


class UdMode(ctypes.Structure):
    _fields_ = [("pose", ctypes.c_float * UD_POSE_FLOATS), ("energy", ctypes.c_float), ("inter", ctypes.c_float),
                ("intra", ctypes.c_float), ("conf_independent", ctypes.c_float), ("search_energy", ctypes.c_float),
                ("rmsd_lb", ctypes.c_float), ("rmsd_ub", ctypes.c_float), ("unbound", ctypes.c_float),
                ("chain", ctypes.c_int32)]


class UdSearchParams(ctypes.Structure):
    _fields_ = [("exhaustiveness", ctypes.c_int32), ("global_steps", ctypes.c_int32),
                ("local_steps", ctypes.c_int32), ("lanes_per_chain", ctypes.c_int32),
                ("seed", ctypes.c_uint64),
                ("hunt_cap", ctypes.c_float * 3), ("authentic_v", ctypes.c_float * 3),
                ("temperature", ctypes.c_float), ("mutation_amplitude", ctypes.c_float),
                ("chains_per_group", ctypes.c_int32), ("first_ls_f0_zero", ctypes.c_int32),
                ("single_launch", ctypes.c_int32)]


_LIB = None


def lib_path() -> str:
    return _build.LIB


def load() -> ctypes.CDLL:
    """Load libunidock_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                           "(no CPU fallback exists)")
    # One HIP runtime per process: with PyTorch installed, load it first, so
    # that this library binds to the HIP runtime torch ships (libamdhip64 is
    # resolved by soname).  The other order maps two runtimes into the
    # process and torch can then not open the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    vp, ip, fp, dp = ctypes.c_void_p, _c_int_p, _c_float_p, ctypes.POINTER(ctypes.c_double)
    sig = {
        "ud_default_search_params": (None, [ctypes.POINTER(UdSearchParams)]),
        "ud_create": (ctypes.c_int, [ctypes.c_int, vp, ctypes.POINTER(vp)]),
        "ud_destroy": (ctypes.c_int, [vp]),
        "ud_last_error": (ctypes.c_char_p, [vp]),
        "ud_set_scoring": (ctypes.c_int, [vp, ctypes.c_int, dp, ctypes.c_int]),
        "ud_set_box": (ctypes.c_int, [vp, dp, dp, ctypes.c_double, ip]),
        "ud_compute_maps": (ctypes.c_int, [vp, ctypes.c_int, fp, ip, ctypes.c_int, ip, ctypes.c_int]),
        "ud_set_maps": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, ctypes.c_int]),
        "ud_get_map": (ctypes.c_int, [vp, ctypes.c_int, fp]),
        "ud_set_ligands": (ctypes.c_int, [vp, ctypes.POINTER(UdLigandBatch)]),
        "ud_score_grad": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, fp, fp, fp]),
        "ud_score_grad_device": (ctypes.c_int, [vp, ctypes.c_int, ip, vp, fp, vp, vp]),
        "ud_score_grad_device_noalloc": (ctypes.c_int, [vp, ctypes.c_int, ip, vp, fp, vp, vp]),
        "ud_energy_grad": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, fp, fp, fp, fp]),
        "ud_energy_grad_device": (ctypes.c_int, [vp, ctypes.c_int, ip, vp, fp, vp, vp, vp]),
        "ud_optimize": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, fp, ctypes.c_int, fp, ip]),
        "ud_dock": (ctypes.c_int, [vp, ctypes.POINTER(UdSearchParams), fp, fp, ctypes.POINTER(ctypes.c_int64)]),
        "ud_fetch_dock_results": (ctypes.c_int, [vp, fp, fp, ctypes.POINTER(ctypes.c_int64)]),
        "ud_score_terms": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, fp, fp]),
        "ud_pose_coords": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, fp]),
        "ud_last_kernel_time": (ctypes.c_int, [vp, dp, ip]),
        "ud_set_lanes": (ctypes.c_int, [vp, ctypes.c_int]),
        "ud_last_launches": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(UdLaunchInfo), ip]),
        "ud_set_bucket_lanes": (ctypes.c_int, [vp, ip]),
        "ud_set_bucket_pair_buffer": (ctypes.c_int, [vp, ip]),
        "ud_get_bucket_layout": (ctypes.c_int, [vp, ip, ip]),
        "ud_autotune": (ctypes.c_int, [vp, ctypes.POINTER(UdSearchParams), ctypes.c_int, ip]),
        "ud_postprocess": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(UdMode), ip]),
        "ud_refine": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, ctypes.c_int, fp, ip]),
        "ud_score_terms_receptor": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, fp, fp]),
        "ud_score_grad_receptor": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, fp, ctypes.c_double, fp, fp]),
        "ud_ligset_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
        "ud_ligset_destroy": (ctypes.c_int, [vp]),
        "ud_ligset_last_error": (ctypes.c_char_p, [vp]),
        "ud_ligset_add_pdbqt": (ctypes.c_int, [vp, ctypes.c_char_p, ip]),
        "ud_ligset_add_sdf": (ctypes.c_int, [vp, ctypes.c_char_p, ip, ip]),
        "ud_ligset_count": (ctypes.c_int, [vp]),
        "ud_ligset_batch": (ctypes.c_int, [vp, ctypes.POINTER(UdLigandBatch)]),
        "ud_ligset_batch_sf": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(UdLigandBatch)]),
        "ud_receptor_ad4_from_pdbqt": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ip, fp, ip, fp]),
        "ud_compute_maps_ad4": (ctypes.c_int, [vp, ctypes.c_int, fp, ip, fp, ctypes.c_int, ip, ctypes.c_int]),
        "ud_ligset_atoms": (ctypes.c_int, [vp, ctypes.c_int, ip, fp, ip]),
        "ud_receptor_from_pdbqt": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ip, fp, ip]),
        "ud_ligset_write_sdf": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(UdMode), fp, ctypes.c_int,
                                               ctypes.c_double, ctypes.c_char_p, ctypes.c_size_t,
                                               ctypes.POINTER(ctypes.c_size_t)]),
        "ud_ligset_write_pdbqt": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(UdMode), fp, ctypes.c_int,
                                                 ctypes.c_double, ctypes.c_char_p, ctypes.c_size_t,
                                                 ctypes.POINTER(ctypes.c_size_t)]),
        "ud_cluster_modes": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, fp, fp, ctypes.c_double, ctypes.c_int, ip, ip]),
        "ud_synth_receptor": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                             ctypes.c_longlong, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class UdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"unidock_amd error {code}: {msg}")
        self.code = code


class ParsedLigand:
    """One ligand of a LigandSet, with the fields of synth.Ligand (batch order)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


class LigandSet:
    """Ligands parsed from PDBQT or SDF text by the library (ud_ligset_*)."""

    def __init__(self):
        self.L = load()
        h = ctypes.c_void_p()
        rc = self.L.ud_ligset_create(ctypes.byref(h))
        if rc != UD_OK:
            raise UdError(rc, "ud_ligset_create failed")
        self.h = h

    def close(self):
        if self.h:
            self.L.ud_ligset_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_pdbqt(self, text: str) -> int:
        idx = ctypes.c_int32()
        rc = self.L.ud_ligset_add_pdbqt(self.h, text.encode(), ctypes.byref(idx))
        if rc != UD_OK:
            raise UdError(rc, self.L.ud_ligset_last_error(self.h).decode())
        return idx.value

    def add_sdf(self, text: str) -> tuple:
        """Add every record of SDF text; returns (index of the first, number added)."""
        first, cnt = ctypes.c_int32(), ctypes.c_int32()
        rc = self.L.ud_ligset_add_sdf(self.h, text.encode(), ctypes.byref(first), ctypes.byref(cnt))
        if rc != UD_OK:
            raise UdError(rc, self.L.ud_ligset_last_error(self.h).decode())
        return first.value, cnt.value

    def __len__(self):
        return self.L.ud_ligset_count(self.h)

    def arrays(self, sf: int = UD_SF_VINA) -> dict:
        """The batch in the dict layout Context.set_ligands takes (copies),
        with the pair list of scoring function sf."""
        b = UdLigandBatch()
        rc = self.L.ud_ligset_batch_sf(self.h, sf, ctypes.byref(b))
        if rc != UD_OK:
            raise UdError(rc, "ud_ligset_batch_sf failed")
        n = b.n_ligands
        ao = np.ctypeslib.as_array(b.atom_offset, (n + 1,)).copy()
        no = np.ctypeslib.as_array(b.node_offset, (n + 1,)).copy()
        po = np.ctypeslib.as_array(b.pair_offset, (n + 1,)).copy()
        na, nn, npr = int(ao[-1]), int(no[-1]), int(po[-1])
        get = lambda ptr, m: np.ctypeslib.as_array(ptr, (m,)).copy() if m > 0 else np.zeros(0, np.int32)
        return dict(
            n_ligands=n, atom_offset=ao, node_offset=no, pair_offset=po,
            atom_local=get(b.atom_local, 3 * na).reshape(-1, 3), atom_xs=get(b.atom_xs, na),
            atom_heavy=get(b.atom_heavy, na), node_parent=get(b.node_parent, nn),
            node_abegin=get(b.node_abegin, nn), node_aend=get(b.node_aend, nn),
            node_origin=get(b.node_origin, 3 * nn).reshape(-1, 3), node_axis=get(b.node_axis, 3 * nn).reshape(-1, 3),
            pairs=get(b.pairs, 2 * npr).reshape(-1, 2), num_tors=get(b.num_tors, n),
            atom_ad=get(b.atom_ad, na), atom_charge=get(b.atom_charge, na), torsdof=get(b.torsdof, n),
        )

    def _write(self, fn, lig, modes, coords, how_many, energy_range) -> str:
        arr = (UdMode * max(1, len(modes)))()
        for k, m in enumerate(modes):
            for f in ("energy", "inter", "intra", "conf_independent", "search_energy", "rmsd_lb", "rmsd_ub",
                      "unbound"):
                setattr(arr[k], f, float(m.get(f, 0.0)))
            arr[k].chain = int(m.get("chain", 0))
        c = np.ascontiguousarray(coords, dtype=np.float32)
        n = ctypes.c_size_t()
        f = getattr(self.L, fn)
        rc = f(self.h, lig, len(modes), arr, _p(c, ctypes.c_float), how_many, energy_range, None, 0, ctypes.byref(n))
        if rc != UD_OK:
            raise UdError(rc, f"{fn} failed")
        buf = ctypes.create_string_buffer(n.value + 1)
        rc = f(self.h, lig, len(modes), arr, _p(c, ctypes.c_float), how_many, energy_range, buf, n.value + 1,
               ctypes.byref(n))
        if rc != UD_OK:
            raise UdError(rc, f"{fn} failed")
        return buf.value.decode()

    def write_pdbqt(self, lig: int, modes: list, coords, how_many=9, energy_range=3.0) -> str:
        """PDBQT text of `lig` (read from PDBQT) with the given modes (dicts as
        Context.postprocess returns) and their atom coordinates
        coords[mode][atom][3] (batch order)."""
        return self._write("ud_ligset_write_pdbqt", lig, modes, coords, how_many, energy_range)

    def write_sdf(self, lig: int, modes: list, coords, how_many=9, energy_range=3.0) -> str:
        """SDF text of `lig` (read from SDF) with the given modes, one record per
        mode with the "> <Uni-Dock RESULT>" field."""
        return self._write("ud_ligset_write_sdf", lig, modes, coords, how_many, energy_range)

    def ligands(self, sf: int = UD_SF_VINA) -> list:
        a = self.arrays(sf)
        out = []
        for l in range(a["n_ligands"]):
            a0, a1 = a["atom_offset"][l], a["atom_offset"][l + 1]
            k0, k1 = a["node_offset"][l], a["node_offset"][l + 1]
            p0, p1 = a["pair_offset"][l], a["pair_offset"][l + 1]
            n_at = ctypes.c_int32()
            self.L.ud_ligset_atoms(self.h, l, ctypes.byref(n_at), None, None)
            xyz = np.zeros((n_at.value, 3), dtype=np.float32)
            serial = np.zeros(n_at.value, dtype=np.int32)
            self.L.ud_ligset_atoms(self.h, l, ctypes.byref(n_at), _p(xyz, ctypes.c_float), _p(serial, ctypes.c_int32))
            out.append(ParsedLigand(
                n_atoms=int(a1 - a0), n_tors=int(k1 - k0 - 1), local=a["atom_local"][a0:a1].astype(np.float64),
                xs=a["atom_xs"][a0:a1], heavy=a["atom_heavy"][a0:a1], parent=a["node_parent"][k0:k1],
                abegin=a["node_abegin"][k0:k1], aend=a["node_aend"][k0:k1],
                origin=a["node_origin"][k0:k1].astype(np.float64), axis=a["node_axis"][k0:k1].astype(np.float64),
                pairs=a["pairs"][p0:p1], num_tors=float(a["num_tors"][l]), coords=xyz.astype(np.float64),
                serial=serial, ad=a["atom_ad"][a0:a1], charge=a["atom_charge"][a0:a1].astype(np.float64),
                torsdof=float(a["torsdof"][l])))
        return out


def receptor_from_pdbqt(text: str):
    """Heavy receptor atoms with X-Score types: (xyz float32 [n, 3], xs int32 [n])."""
    L = load()
    n = ctypes.c_int32()
    rc = L.ud_receptor_from_pdbqt(text.encode(), 0, ctypes.byref(n), None, None)
    if rc != UD_OK:
        raise UdError(rc, "receptor PDBQT could not be parsed")
    xyz = np.zeros((n.value, 3), dtype=np.float32)
    xs = np.zeros(n.value, dtype=np.int32)
    rc = L.ud_receptor_from_pdbqt(text.encode(), n.value, ctypes.byref(n), _p(xyz, ctypes.c_float), _p(xs, ctypes.c_int32))
    if rc != UD_OK:
        raise UdError(rc, "receptor PDBQT could not be parsed")
    return xyz, xs


def receptor_ad4_from_pdbqt(text: str):
    """Every receptor atom with its AutoDock type and partial charge:
    (xyz float32 [n, 3], ad int32 [n], charge float32 [n])."""
    L = load()
    n = ctypes.c_int32()
    rc = L.ud_receptor_ad4_from_pdbqt(text.encode(), 0, ctypes.byref(n), None, None, None)
    if rc != UD_OK:
        raise UdError(rc, "receptor PDBQT could not be parsed")
    xyz = np.zeros((n.value, 3), dtype=np.float32)
    ad = np.zeros(n.value, dtype=np.int32)
    q = np.zeros(n.value, dtype=np.float32)
    rc = L.ud_receptor_ad4_from_pdbqt(text.encode(), n.value, ctypes.byref(n), _p(xyz, ctypes.c_float),
                                      _p(ad, ctypes.c_int32), _p(q, ctypes.c_float))
    if rc != UD_OK:
        raise UdError(rc, "receptor PDBQT could not be parsed")
    return xyz, ad, q


def ad4_grid_type(t: int) -> int:
    """AutoDock type whose affinity map an atom of type t reads: CG carbons
    the C map, G atoms none (ad4cache.cpp:223-240)."""
    if t < 0 or t >= 35 or 22 <= t <= 25:
        return -1
    if 26 <= t <= 29:
        return 0
    return t


def read_ad4_map(path):
    """One AutoGrid map file (read_ad4_map, ad4cache.cpp:332-389): returns
    (center[3], n_voxels[3], spacing, values x-fastest).  The values start on
    line 7; line 4 holds SPACING, line 5 NELEMENTS (even counts), line 6
    CENTER."""
    with open(path) as f:
        lines = f.read().splitlines()
    spacing = float(lines[3].split()[1])
    nv = [int(x) for x in lines[4].split()[1:4]]
    if any(n % 2 for n in nv):
        raise ValueError(f"{path}: number of voxels (NELEMENTS) must be even")
    center = [float(x) for x in lines[5].split()[1:4]]
    vals = np.array([float(x) for x in lines[6:] if x.strip()], dtype=np.float64)
    if vals.size != (nv[0] + 1) * (nv[1] + 1) * (nv[2] + 1):
        raise ValueError(f"{path}: {vals.size} values for {nv} voxels")
    return np.array(center), np.array(nv), spacing, vals


def write_ad4_map(path, center, n_voxels, spacing, values, receptor="receptor.pdbqt"):
    """An AutoGrid-format map file (the layout ad4cache::write writes,
    ad4cache.cpp:470-520): six header lines, then one value per line."""
    nv = [int(x) for x in n_voxels]
    with open(path, "w") as f:
        f.write("GRID_PARAMETER_FILE unidock_amd.gpf\nGRID_DATA_FILE unidock_amd.maps.fld\n")
        f.write(f"MACROMOLECULE {receptor}\nSPACING {spacing:.3f}\n")
        f.write(f"NELEMENTS {nv[0]} {nv[1]} {nv[2]}\n")
        f.write(f"CENTER {center[0]:.3f} {center[1]:.3f} {center[2]:.3f}\n")
        for v in np.asarray(values).reshape(-1):
            f.write(f"{v:.4g}\n")  # std::setprecision(4)


def cluster_modes(heavy_xyz, energy, min_rmsd=1.0, max_size=9):
    """add_to_output_container over candidates in order (host only); kept indices."""
    L = load()
    x = np.ascontiguousarray(heavy_xyz, dtype=np.float32)
    e = np.ascontiguousarray(energy, dtype=np.float32)
    n = x.shape[0]
    nh = x.shape[1] if x.ndim == 3 else (x.shape[1] // 3 if x.ndim == 2 else 0)
    kept = np.zeros(max_size, dtype=np.int32)
    nk = ctypes.c_int32()
    rc = L.ud_cluster_modes(n, nh, _p(x, ctypes.c_float), _p(e, ctypes.c_float), min_rmsd, max_size,
                            _p(kept, ctypes.c_int32), ctypes.byref(nk))
    if rc != UD_OK:
        raise UdError(rc, "ud_cluster_modes failed")
    return kept[:nk.value].tolist()


def synth_receptor_xyz(n_atoms: int, edge: float, min_dist: float, seed: int) -> np.ndarray:
    L = load()
    out = np.zeros(n_atoms * 3, dtype=np.float64)
    rc = L.ud_synth_receptor(n_atoms, edge, min_dist, seed, 10**8, _p(out, ctypes.c_double))
    if rc != UD_OK:
        raise UdError(rc, "receptor sampler failed (density too high?)")
    return out.reshape(-1, 3)


def grid_type(t: int) -> int:
    """XS type whose map an atom of type t reads: CG carbons read the plain
    carbon maps, G atoms and untyped atoms none (cache.cpp:172-191)."""
    if t < 0 or t >= 36 or t in (21, 24, 27, 30):
        return -1
    if t in (19, 22, 25, 28):
        return 0
    if t in (20, 23, 26, 29):
        return 1
    return t


class Context:
    """One device context (ud_ctx) on a HIP stream."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.L = load()
        h = ctypes.c_void_p()
        rc = self.L.ud_create(device, ctypes.c_void_p(stream or 0), ctypes.byref(h))
        if rc != UD_OK:
            raise UdError(rc, "ud_create failed (no HIP device?)")
        self.h = h
        self.n_ligands = 0
        self._keep = None

    def close(self):
        if self.h:
            self.L.ud_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc != UD_OK:
            raise UdError(rc, self.L.ud_last_error(self.h).decode())

    # ---- set-up
    def set_scoring(self, sf: int = UD_SF_VINA, weights=None):
        w = np.array(weights if weights is not None else DEFAULT_WEIGHTS[sf], dtype=np.float64)
        self._ck(self.L.ud_set_scoring(self.h, sf, _p(w, ctypes.c_double), len(w)))
        self.weights = tuple(float(x) for x in w)
        self.sf = sf

    def set_box(self, center, size, granularity=0.375):
        c = np.array(center, dtype=np.float64)
        s = np.array(size, dtype=np.float64)
        n = np.zeros(3, dtype=np.int32)
        self._ck(self.L.ud_set_box(self.h, _p(c, ctypes.c_double), _p(s, ctypes.c_double), granularity,
                                   _p(n, ctypes.c_int32)))
        self.npts = tuple(int(x) for x in n)
        return self.npts

    def compute_maps(self, rec_xyz, rec_xs, types, precision=UD_GRID_F32):
        x = np.ascontiguousarray(rec_xyz, dtype=np.float32).reshape(-1)
        t = np.ascontiguousarray(rec_xs, dtype=np.int32)
        ty = np.ascontiguousarray(types, dtype=np.int32)
        self._ck(self.L.ud_compute_maps(self.h, len(t), _p(x, ctypes.c_float), _p(t, ctypes.c_int32), len(ty),
                                        _p(ty, ctypes.c_int32), precision))
        self.map_types = [int(v) for v in ty]

    def compute_maps_ad4(self, rec_xyz, rec_ad, rec_charge, types, precision=UD_GRID_F32):
        """AD4.2 affinity maps of the listed AutoDock types plus the
        electrostatic and desolvation maps (UD_SF_AD4 scoring)."""
        x = np.ascontiguousarray(rec_xyz, dtype=np.float32).reshape(-1)
        t = np.ascontiguousarray(rec_ad, dtype=np.int32)
        q = np.ascontiguousarray(rec_charge, dtype=np.float32)
        ty = np.ascontiguousarray(types, dtype=np.int32)
        self._ck(self.L.ud_compute_maps_ad4(self.h, len(t), _p(x, ctypes.c_float), _p(t, ctypes.c_int32),
                                            _p(q, ctypes.c_float), len(ty), _p(ty, ctypes.c_int32), precision))
        self.map_types = [int(v) for v in ty] + [UD_AD4_MAP_ELEC, UD_AD4_MAP_DSOLV]

    def load_ad4_maps(self, prefix, ad_types, precision=UD_GRID_F32):
        """AutoGrid maps <prefix>.<type>.map of the needed AutoDock types plus
        <prefix>.e.map and <prefix>.d.map (ad4cache::read, ad4cache.cpp:
        391-440): sets the box from the files and uploads the maps."""
        types = sorted({ad4_grid_type(int(t)) for t in ad_types} - {-1})
        names = [f"{prefix}.{AD_TYPE_NAMES[t]}.map" for t in types] + [f"{prefix}.e.map", f"{prefix}.d.map"]
        geo, vals = None, []
        for nm in names:
            center, nv, spacing, v = read_ad4_map(nm)
            if geo is None:
                geo = (center, nv, spacing)
            elif not (np.allclose(center, geo[0]) and np.array_equal(nv, geo[1]) and spacing == geo[2]):
                raise ValueError(f"{nm}: grid differs from {names[0]}")
            vals.append(v)
        center, nv, spacing = geo
        self.set_box(center, nv * spacing, spacing)
        self.set_maps(types + [UD_AD4_MAP_ELEC, UD_AD4_MAP_DSOLV], np.concatenate(vals), precision)

    def set_maps(self, types, values, precision=UD_GRID_F32):
        ty = np.ascontiguousarray(types, dtype=np.int32)
        v = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        self._ck(self.L.ud_set_maps(self.h, len(ty), _p(ty, ctypes.c_int32), _p(v, ctypes.c_float), precision))
        self.map_types = [int(x) for x in ty]

    def get_map(self, xs_type: int) -> np.ndarray:
        out = np.zeros(int(np.prod(self.npts)), dtype=np.float32)
        self._ck(self.L.ud_get_map(self.h, xs_type, _p(out, ctypes.c_float)))
        return out

    def set_ligands(self, arrays: dict):
        a = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in arrays.items()}
        b = UdLigandBatch()
        b.n_ligands = a["n_ligands"]
        for k in ("atom_offset", "node_offset", "pair_offset", "atom_xs", "atom_heavy", "node_parent",
                  "node_abegin", "node_aend", "pairs"):
            a[k] = np.ascontiguousarray(a[k], dtype=np.int32).reshape(-1)
            setattr(b, k, _p(a[k], ctypes.c_int32))
        for k in ("atom_local", "node_origin", "node_axis", "num_tors"):
            a[k] = np.ascontiguousarray(a[k], dtype=np.float32).reshape(-1)
            setattr(b, k, _p(a[k], ctypes.c_float))
        # AD4.2 typing (optional for Vina / Vinardo)
        if a.get("atom_ad") is not None:
            a["atom_ad"] = np.ascontiguousarray(a["atom_ad"], dtype=np.int32).reshape(-1)
            b.atom_ad = _p(a["atom_ad"], ctypes.c_int32)
        for k in ("atom_charge", "torsdof"):
            if a.get(k) is not None:
                a[k] = np.ascontiguousarray(a[k], dtype=np.float32).reshape(-1)
                setattr(b, k, _p(a[k], ctypes.c_float))
        self._ck(self.L.ud_set_ligands(self.h, ctypes.byref(b)))
        self._keep = a
        self.n_ligands = a["n_ligands"]
        self.n_tors = np.diff(a["node_offset"]) - 1
        self.n_atoms = np.diff(a["atom_offset"])

    # ---- compute
    def score_grad(self, pose_ligand, poses, v=(1000.0, 1000.0, 1000.0)):
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        P = np.ascontiguousarray(poses, dtype=np.float32).reshape(len(pl), UD_POSE_FLOATS)
        vv = np.array(v, dtype=np.float32)
        e = np.zeros(len(pl), dtype=np.float32)
        g = np.zeros((len(pl), UD_GRAD_FLOATS), dtype=np.float32)
        self._ck(self.L.ud_score_grad(self.h, len(pl), _p(pl, ctypes.c_int32), _p(P, ctypes.c_float),
                                      _p(vv, ctypes.c_float), _p(e, ctypes.c_float), _p(g, ctypes.c_float)))
        return e, g

    def energy_grad(self, pose_ligand, poses, v=(1000.0, 1000.0, 1000.0)):
        """Energy, dE/d(position, quaternion w-x-y-z, torsions) of E(q / |q|)
        and dE/d(term weights) (ud_energy_grad)."""
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        P = np.ascontiguousarray(poses, dtype=np.float32).reshape(len(pl), UD_POSE_FLOATS)
        vv = np.array(v, dtype=np.float32)
        e = np.zeros(len(pl), dtype=np.float32)
        gp = np.zeros((len(pl), UD_POSE_FLOATS), dtype=np.float32)
        gw = np.zeros((len(pl), UD_WEIGHT_FLOATS), dtype=np.float32)
        self._ck(self.L.ud_energy_grad(self.h, len(pl), _p(pl, ctypes.c_int32), _p(P, ctypes.c_float),
                                       _p(vv, ctypes.c_float), _p(e, ctypes.c_float), _p(gp, ctypes.c_float),
                                       _p(gw, ctypes.c_float)))
        return e, gp, gw

    def energy_grad_device(self, pose_ligand, d_poses: int, d_energy: int, d_pose_grad: int, d_weight_grad: int,
                           v=(1000.0, 1000.0, 1000.0)):
        """ud_energy_grad_device: device addresses, asynchronous on the context's stream."""
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        vv = np.array(v, dtype=np.float32)
        self._ck(self.L.ud_energy_grad_device(self.h, len(pl), _p(pl, ctypes.c_int32), ctypes.c_void_p(d_poses),
                                              _p(vv, ctypes.c_float), ctypes.c_void_p(d_energy),
                                              ctypes.c_void_p(d_pose_grad), ctypes.c_void_p(d_weight_grad)))

    def score_grad_device(self, pose_ligand, d_poses: int, d_energy: int, d_grad: int, v=(1000.0, 1000.0, 1000.0)):
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        vv = np.array(v, dtype=np.float32)
        self._ck(self.L.ud_score_grad_device(self.h, len(pl), _p(pl, ctypes.c_int32), ctypes.c_void_p(d_poses),
                                             _p(vv, ctypes.c_float), ctypes.c_void_p(d_energy),
                                             ctypes.c_void_p(d_grad)))

    def score_grad_device_noalloc(self, pose_ligand, d_poses: int, d_energy: int, d_grad: int,
                                  v=(1000.0, 1000.0, 1000.0)):
        """ud_score_grad_device_noalloc: as score_grad_device, with no device
        allocation (at most UD_NOALLOC_MAX_POSES poses)."""
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        vv = np.array(v, dtype=np.float32)
        self._ck(self.L.ud_score_grad_device_noalloc(self.h, len(pl), _p(pl, ctypes.c_int32), ctypes.c_void_p(d_poses),
                                                     _p(vv, ctypes.c_float), ctypes.c_void_p(d_energy),
                                                     ctypes.c_void_p(d_grad)))

    def optimize(self, pose_ligand, poses, v=(1000.0, 1000.0, 1000.0), max_steps=0):
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        P = np.array(poses, dtype=np.float32).reshape(len(pl), UD_POSE_FLOATS).copy()
        vv = np.array(v, dtype=np.float32)
        e = np.zeros(len(pl), dtype=np.float32)
        ev = np.zeros(len(pl), dtype=np.int32)
        self._ck(self.L.ud_optimize(self.h, len(pl), _p(pl, ctypes.c_int32), _p(P, ctypes.c_float),
                                    _p(vv, ctypes.c_float), max_steps, _p(e, ctypes.c_float), _p(ev, ctypes.c_int32)))
        return P, e, ev

    def _params(self, exhaustiveness, global_steps, seed, local_steps, lanes, chains_per_group=0,
                first_ls_f0_zero=False, single_launch=False):
        p = UdSearchParams()
        self.L.ud_default_search_params(ctypes.byref(p))
        p.exhaustiveness = exhaustiveness
        p.global_steps = global_steps
        p.seed = seed
        p.local_steps = local_steps
        p.lanes_per_chain = lanes
        p.chains_per_group = chains_per_group
        p.first_ls_f0_zero = int(bool(first_ls_f0_zero))
        p.single_launch = int(bool(single_launch))
        return p

    def dock(self, exhaustiveness=8, global_steps=20, seed=181129, local_steps=0, lanes=0, fetch=True,
             chains_per_group=0, first_ls_f0_zero=False, single_launch=False):
        """ud_dock; first_ls_f0_zero selects the reference GPU's first-line-search
        rule (f0 = 0) instead of the CPU rule (f0 = f(x)); single_launch runs the
        whole batch as one launch (the reference's launch shape) instead of one
        launch per size bucket."""
        p = self._params(exhaustiveness, global_steps, seed, local_steps, lanes, chains_per_group, first_ls_f0_zero,
                         single_launch)
        self._ck(self.L.ud_dock(self.h, ctypes.byref(p), None, None, None))
        self._last_dock = (self.n_ligands, exhaustiveness)
        if fetch:
            return self.fetch_dock()

    def fetch_dock(self):
        n = self._last_dock[0] * self._last_dock[1]
        P = np.zeros((n, UD_POSE_FLOATS), dtype=np.float32)
        e = np.zeros(n, dtype=np.float32)
        ev = np.zeros(n, dtype=np.int64)
        self._ck(self.L.ud_fetch_dock_results(self.h, _p(P, ctypes.c_float), _p(e, ctypes.c_float),
                                              _p(ev, ctypes.c_int64)))
        return P, e, ev

    def score_terms(self, pose_ligand, poses):
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        P = np.ascontiguousarray(poses, dtype=np.float32).reshape(len(pl), UD_POSE_FLOATS)
        a = np.zeros(len(pl), dtype=np.float32)
        b = np.zeros(len(pl), dtype=np.float32)
        self._ck(self.L.ud_score_terms(self.h, len(pl), _p(pl, ctypes.c_int32), _p(P, ctypes.c_float),
                                       _p(a, ctypes.c_float), _p(b, ctypes.c_float)))
        return a, b

    def score_terms_receptor(self, pose_ligand, poses):
        """Final-score components against the explicit receptor (ud_score_terms_receptor)."""
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        P = np.ascontiguousarray(poses, dtype=np.float32).reshape(len(pl), UD_POSE_FLOATS)
        a = np.zeros(len(pl), dtype=np.float32)
        b = np.zeros(len(pl), dtype=np.float32)
        self._ck(self.L.ud_score_terms_receptor(self.h, len(pl), _p(pl, ctypes.c_int32), _p(P, ctypes.c_float),
                                                _p(a, ctypes.c_float), _p(b, ctypes.c_float)))
        return a, b

    def score_grad_receptor(self, pose_ligand, poses, v=(1000.0, 1000.0, 1000.0), slope=1e6):
        """Energy and change-space gradient against the explicit receptor (ud_score_grad_receptor)."""
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        P = np.ascontiguousarray(poses, dtype=np.float32).reshape(len(pl), UD_POSE_FLOATS)
        vv = np.array(v, dtype=np.float32)
        e = np.zeros(len(pl), dtype=np.float32)
        g = np.zeros((len(pl), UD_GRAD_FLOATS), dtype=np.float32)
        self._ck(self.L.ud_score_grad_receptor(self.h, len(pl), _p(pl, ctypes.c_int32), _p(P, ctypes.c_float),
                                               _p(vv, ctypes.c_float), slope, _p(e, ctypes.c_float),
                                               _p(g, ctypes.c_float)))
        return e, g

    def refine(self, pose_ligand, poses, refine_step=3):
        """Refine poses against the explicit receptor (ud_refine); returns
        (refined poses, energies, within flags)."""
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        P = np.array(poses, dtype=np.float32).reshape(len(pl), UD_POSE_FLOATS)
        e = np.zeros(len(pl), dtype=np.float32)
        w = np.zeros(len(pl), dtype=np.int32)
        self._ck(self.L.ud_refine(self.h, len(pl), _p(pl, ctypes.c_int32), _p(P, ctypes.c_float), refine_step,
                                  _p(e, ctypes.c_float), _p(w, ctypes.c_int32)))
        return P, e, w.astype(bool)

    def pose_coords(self, pose_ligand, poses):
        pl = np.ascontiguousarray(pose_ligand, dtype=np.int32)
        P = np.ascontiguousarray(poses, dtype=np.float32).reshape(len(pl), UD_POSE_FLOATS)
        out = np.zeros((len(pl), UD_MAX_ATOMS, 3), dtype=np.float32)
        self._ck(self.L.ud_pose_coords(self.h, len(pl), _p(pl, ctypes.c_int32), _p(P, ctypes.c_float),
                                       _p(out, ctypes.c_float)))
        return out

    def postprocess(self, max_modes=9, min_rmsd=1.0, refine_step=3):
        """Cluster, refine (refine_step > 0, the reference CLI default 3) and
        rescore the last dock's chains; one list of mode dicts per ligand."""
        n = self.n_ligands
        modes = (UdMode * (n * max_modes))()
        cnt = np.zeros(n, dtype=np.int32)
        self._ck(self.L.ud_postprocess(self.h, max_modes, min_rmsd, refine_step, modes, _p(cnt, ctypes.c_int32)))
        out = []
        for l in range(n):
            lst = []
            for k in range(int(cnt[l])):
                m = modes[l * max_modes + k]
                lst.append(dict(pose=np.array(m.pose[:], dtype=np.float32), energy=m.energy, inter=m.inter,
                                intra=m.intra, conf_independent=m.conf_independent,
                                search_energy=m.search_energy, rmsd_lb=m.rmsd_lb, rmsd_ub=m.rmsd_ub,
                                unbound=m.unbound, chain=int(m.chain)))
            out.append(lst)
        return out

    def set_lanes(self, lanes: int):
        self._ck(self.L.ud_set_lanes(self.h, lanes))

    def last_kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_int32()
        self._ck(self.L.ud_last_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def last_launches(self):
        """Per-launch (bucket) timing of the last dock: list of dicts."""
        info = (UdLaunchInfo * UD_MAX_LAUNCHES)()
        n = ctypes.c_int32()
        self._ck(self.L.ud_last_launches(self.h, UD_MAX_LAUNCHES, info, ctypes.byref(n)))
        return [dict(atom_cap=i.atom_cap, lanes=i.lanes, n_ligands=i.n_ligands, n_workgroups=i.n_workgroups,
                     pair_buffer=i.pair_buffer, lds_bytes=i.lds_bytes, max_torsions=i.max_torsions, ms=i.ms)
                for i in info[:n.value]]

    def set_bucket_lanes(self, lanes):
        a = np.zeros(UD_N_BUCKETS, dtype=np.int32)
        a[:len(lanes)] = lanes
        self._ck(self.L.ud_set_bucket_lanes(self.h, _p(a, ctypes.c_int32)))

    def set_bucket_pair_buffer(self, pair_buffer):
        """Per-bucket LDS layout: 1 per-pair force buffer, 0 per-atom pair loops, -1 default."""
        a = np.full(UD_N_BUCKETS, -1, dtype=np.int32)
        a[:len(pair_buffer)] = pair_buffer
        self._ck(self.L.ud_set_bucket_pair_buffer(self.h, _p(a, ctypes.c_int32)))

    def bucket_layout(self):
        """(lanes per chain, pair-buffer setting) per size bucket."""
        lanes = np.zeros(UD_N_BUCKETS, dtype=np.int32)
        pb = np.zeros(UD_N_BUCKETS, dtype=np.int32)
        self._ck(self.L.ud_get_bucket_layout(self.h, _p(lanes, ctypes.c_int32), _p(pb, ctypes.c_int32)))
        return lanes, pb

    def autotune(self, exhaustiveness=8, global_steps=2, seed=181129, local_steps=0, sample_ligands=1 << 20):
        """Pick lanes per chain per size bucket by timing short docks; returns the chosen lanes."""
        p = self._params(exhaustiveness, global_steps, seed, local_steps, 0)
        out = np.zeros(UD_N_BUCKETS, dtype=np.int32)
        self._ck(self.L.ud_autotune(self.h, ctypes.byref(p), sample_ligands, _p(out, ctypes.c_int32)))
        return out
