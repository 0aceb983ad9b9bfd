"""ctypes wrapper of the CPU oracle (liboracle.so built from oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always to check the GPU's results or as the timed
CPU baseline, never by the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

POSE = 7 + 48  # pose stride shared with the product's C ABI (UD_POSE_FLOATS)
SF_VINA, SF_VINARDO, SF_AD4 = 0, 1, 2
AD4_WEIGHTS = (0.1662, 0.1209, 0.1406, 0.1322, 50.0, 0.2983)  # vina.h:136-138
AD4_MAP_ELEC, AD4_MAP_DSOLV = 35, 36  # map indices of the electrostatic / desolvation maps
AD4_TABLE_N = 13424  # sz(32 * 20.48^2) + 3 samples per AD4.2 pair table


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def build_native(out_dir):
    """An -O3 -march=native build of oracle.cpp for the machine it runs on (the
    timed CPU baseline); returns the library path."""
    out = os.path.join(out_dir, "liboracle_native.so")
    subprocess.run(["g++", "-O3", "-march=native", "-std=c++17", "-fPIC", "-shared", "-o", out,
                    os.path.join(HERE, "oracle.cpp"), "-lpthread"], check=True)
    return out


_L = None
_L_PATH = None


def load(path=None):
    """Load the oracle library (the in-tree build, or `path`); later Oracle
    objects use the library loaded last."""
    global _L, _L_PATH
    path = path or _L_PATH or LIB
    if _L is not None and path == _L_PATH:
        return _L
    if path == LIB and not os.path.exists(LIB):
        build()
    L = ctypes.CDLL(path)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    sig = {
        "orc_philox": (None, [u32p, u32p, u32p]),
        "orc_new": (vp, [ctypes.c_int, ctypes.c_int, dp]),
        "orc_free": (None, [vp]),
        "orc_pair_table": (None, [vp, ctypes.c_int, ctypes.c_int, dp, dp, dp]),
        "orc_sf_eval": (ctypes.c_double, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
        "orc_set_grid": (None, [vp, dp, dp, ctypes.c_double, ip, dp]),
        "orc_build_maps": (None, [vp, ctypes.c_int, dp, ip, ctypes.c_int, ip]),
        "orc_get_map": (None, [vp, ctypes.c_int, dp]),
        "orc_set_map": (None, [vp, ctypes.c_int, dp]),
        "orc_grid_eval": (ctypes.c_double, [vp, ctypes.c_int, dp, ctypes.c_double, dp]),
        "orc_add_ligand": (ctypes.c_int, [vp, ctypes.c_int, dp, ip, ip, ctypes.c_int, ip, ip, ip, dp, dp,
                                          ctypes.c_int, ip, ctypes.c_double]),
        "orc_eval_deriv": (ctypes.c_double, [vp, ctypes.c_int, dp, dp, dp, dp, dp]),
        "orc_bfgs": (ctypes.c_double, [vp, ctypes.c_int, dp, dp, ctypes.c_int, ip]),
        "orc_mc_chain": (ctypes.c_double, [vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                                           ctypes.c_int, dp, ip, dp]),
        "orc_random_start": (None, [vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, dp]),
        "orc_mutate": (None, [vp, ctypes.c_int, dp, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32,
                              ctypes.c_uint32]),
        "orc_score_terms": (None, [vp, ctypes.c_int, dp, dp, dp]),
        "orc_score_terms_nc": (None, [vp, ctypes.c_int, dp, dp, dp]),
        "orc_refine": (ctypes.c_double, [vp, ctypes.c_int, dp, ctypes.c_int, ip]),
        "orc_eval_batch": (None, [vp, ctypes.c_int, ip, dp, dp, ctypes.c_int, dp]),
        "orc_nc_atom": (ctypes.c_double, [vp, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, dp]),
        "orc_eval_deriv_nc": (ctypes.c_double, [vp, ctypes.c_int, dp, dp, ctypes.c_double, dp]),
        "orc_conf_independent": (ctypes.c_double, [vp, ctypes.c_double, ctypes.c_double]),
        "orc_set_ad4": (ctypes.c_int, [vp, ctypes.c_int, ip, dp, ctypes.c_double]),
        "orc_build_maps_ad4": (None, [vp, ctypes.c_int, dp, ip, dp, ctypes.c_int, ip]),
        "orc_ad4_terms": (None, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 dp]),
        "orc_ad4_kind": (None, [ctypes.c_int, dp]),
        "orc_pair_table_ad4": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, dp,
                                              dp, dp]),
        "orc_bfgs_quadratic": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_double, dp,
                                                 ctypes.c_int, ctypes.c_int, ip, ip, dp, dp, ip, dp, dp, dp]),
        "orc_set_ls_f0_zero": (None, [vp, ctypes.c_int]),
        "orc_metropolis": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "orc_dock_cpu": (ctypes.c_double, [vp, ctypes.c_int, ip, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ip,
                                           ctypes.c_int, dp, ctypes.POINTER(ctypes.c_longlong), dp]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype = r
        f.argtypes = a
    _L = L
    _L_PATH = path
    return L


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def ad4_kind(t):
    """(radius, depth, hb_depth, hb_r, solvation, volume) of AutoDock type t."""
    out = np.zeros(6)
    load().orc_ad4_kind(t, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def philox(ctr, key):
    L = load()
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    L.orc_philox(c, k, o)
    return list(o)


def bfgs_quadratic(A, b, x0, c=0.0, max_steps=20, prec=64, f0_zero=False):
    """The oracle's BFGS on f(x) = x^T A x / 2 - b^T x + c with x moved by
    x + alpha p.  Returns a dict: x, f (final), evals, steps and per step
    alpha, f, trials, H (after the step's update), s = alpha p, y = g_new - g."""
    L = load()
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    ev, st = ctypes.c_int(), ctypes.c_int()
    al, fv = np.zeros(max_steps), np.zeros(max_steps)
    tr = np.zeros(max_steps, dtype=np.int32)
    H, s, y = np.zeros(max_steps * n * n), np.zeros(max_steps * n), np.zeros(max_steps * n)
    pd = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    f = L.orc_bfgs_quadratic(prec, n, pd(A), _d(b), c, pd(x), max_steps, int(f0_zero), ctypes.byref(ev),
                             ctypes.byref(st), pd(al), pd(fv), tr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             pd(H), pd(s), pd(y))
    k = st.value
    return dict(x=x, f=f, evals=ev.value, steps=k, alpha=al[:k], fs=fv[:k], trials=tr[:k],
                H=H[:k * n * n].reshape(k, n, n), s=s[:k * n].reshape(k, n), y=y[:k * n].reshape(k, n))


def metropolis(old_f, new_f, temperature, u):
    return bool(load().orc_metropolis(old_f, new_f, temperature, u))


class Oracle:
    """prec = 64 (FP64 oracle) or 32 (FP32 restatement of the reference CPU path)."""

    def __init__(self, prec=64, sf=0, weights=None):
        from_w = weights
        if from_w is None:
            from_w = {SF_VINA: (-0.035579, -0.005156, 0.840245, -0.035069, -0.587439, 50.0, 0.05846),
                      SF_VINARDO: (-0.045, 0.8, -0.035, -0.600, 50.0, 0.05846),
                      SF_AD4: AD4_WEIGHTS}[sf]
        self.L = load()
        self.w = np.array(from_w, dtype=np.float64)
        self.h = self.L.orc_new(prec, sf, _d(self.w))
        self.prec = prec
        self.sf = sf
        self.ligs = []

    def __del__(self):
        try:
            self.L.orc_free(self.h)
        except Exception:
            pass

    def pair_table(self, t1, t2):
        e = np.zeros(2051); dor = np.zeros(2051); fast = np.zeros(2051)
        self.L.orc_pair_table(self.h, t1, t2, e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              dor.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              fast.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return e, dor, fast

    def sf_eval(self, t1, t2, r):
        return self.L.orc_sf_eval(self.h, t1, t2, r)

    def set_grid(self, center, size, granularity=0.375):
        n = np.zeros(3, dtype=np.int32)
        geo = np.zeros(9)
        self.L.orc_set_grid(self.h, _d(center), _d(size), granularity, n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            geo.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        self.n = tuple(int(x) for x in n)
        self.geo = geo
        return self.n, geo

    def build_maps(self, rec_xyz, rec_xs, types):
        x = np.ascontiguousarray(rec_xyz, dtype=np.float64)
        t = np.ascontiguousarray(rec_xs, dtype=np.int32)
        ty = np.ascontiguousarray(types, dtype=np.int32)
        self.L.orc_build_maps(self.h, len(t), _d(x), _i(t), len(ty), _i(ty))

    def get_map(self, t):
        out = np.zeros(int(np.prod(self.n)))
        self.L.orc_get_map(self.h, t, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def set_map(self, t, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self.L.orc_set_map(self.h, t, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def grid_eval(self, t, loc, v=1000.0, deriv=True):
        d = np.zeros(3)
        e = self.L.orc_grid_eval(self.h, t, _d(loc), v,
                                 d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if deriv else None)
        return e, d

    def add_ligand(self, lig, pairs=None):
        """lig: unidock_tools.synth.Ligand (or any object with the same fields).
        With AD4.2 scoring the ligand's AutoDock types, charges and TORSDOF
        (lig.ad, lig.charge, lig.torsdof) are used and `pairs` (default
        lig.pairs) is the pair list of AD-typed atoms."""
        pr = np.asarray(lig.pairs if pairs is None else pairs, dtype=np.int32).reshape(-1, 2)
        idx = self.L.orc_add_ligand(self.h, lig.n_atoms, _d(lig.local), _i(lig.xs), _i(lig.heavy), len(lig.parent),
                                    _i(lig.parent), _i(lig.abegin), _i(lig.aend), _d(lig.origin), _d(lig.axis),
                                    len(pr), _i(pr), float(lig.num_tors))
        if self.sf == SF_AD4:
            rc = self.L.orc_set_ad4(self.h, idx, _i(lig.ad), _d(lig.charge), float(lig.torsdof))
            assert rc == 0
        self.ligs.append(lig)
        return idx

    def build_maps_ad4(self, rec_xyz, rec_ad, rec_charge, types):
        """AD4.2 affinity maps of the listed AD types plus the electrostatic
        (AD4_MAP_ELEC) and desolvation (AD4_MAP_DSOLV) maps."""
        x = np.ascontiguousarray(rec_xyz, dtype=np.float64)
        t = np.ascontiguousarray(rec_ad, dtype=np.int32)
        ty = np.ascontiguousarray(types, dtype=np.int32)
        self.L.orc_build_maps_ad4(self.h, len(t), _d(x), _i(t), _d(rec_charge), len(ty), _i(ty))

    def ad4_terms(self, t1, q1, t2, q2, r):
        """(vdw, hb, elec with charges, solvation, weighted sum) of two atoms at r."""
        out = np.zeros(5)
        self.L.orc_ad4_terms(self.h, t1, q1, t2, q2, r, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def pair_table_ad4(self, t1, q1, t2, q2):
        e = np.zeros(AD4_TABLE_N); dor = np.zeros(AD4_TABLE_N); fast = np.zeros(AD4_TABLE_N)
        n = self.L.orc_pair_table_ad4(self.h, t1, q1, t2, q2, e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      dor.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      fast.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert n == AD4_TABLE_N
        return e, dor, fast

    def eval_deriv(self, lig, pose, v=(1000.0, 1000.0, 1000.0)):
        n_t = self.ligs[lig].n_tors
        p = np.zeros(POSE)
        p[:len(pose)] = pose
        g = np.zeros(6 + 48)
        coords = np.zeros(3 * self.ligs[lig].n_atoms)
        terms = np.zeros(2)
        e = self.L.orc_eval_deriv(self.h, lig, _d(p), _d(v), g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  coords.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  terms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return e, g[:6 + n_t], coords.reshape(-1, 3), terms

    def bfgs(self, lig, pose, v=(1000.0, 1000.0, 1000.0), max_steps=None):
        if max_steps is None:
            max_steps = (25 + self.ligs[lig].n_atoms) // 3
        p = np.zeros(POSE)
        p[:len(pose)] = pose
        ev = ctypes.c_int()
        pp = np.ascontiguousarray(p)
        e = self.L.orc_bfgs(self.h, lig, pp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _d(v), max_steps,
                            ctypes.byref(ev))
        return e, pp, ev.value

    def mc_chain(self, lig, seed, chain, global_steps, local_steps=None):
        if local_steps is None:
            local_steps = (25 + self.ligs[lig].n_atoms) // 3
        best = np.zeros(POSE)
        tr = np.zeros(global_steps)
        ev = ctypes.c_int()
        e = self.L.orc_mc_chain(self.h, lig, seed, chain, global_steps, local_steps,
                                best.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(ev),
                                tr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return e, best, ev.value, tr

    def random_start(self, lig, seed, chain):
        p = np.zeros(POSE)
        self.L.orc_random_start(self.h, lig, seed, chain, p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return p

    def mutate(self, lig, pose, rgyr, seed, chain, step):
        p = np.zeros(POSE)
        p[:len(pose)] = pose
        self.L.orc_mutate(self.h, lig, p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), rgyr, seed, chain, step)
        return p

    def score_terms(self, lig, pose):
        p = np.zeros(POSE)
        p[:len(pose)] = pose
        a = ctypes.c_double()
        b = ctypes.c_double()
        self.L.orc_score_terms(self.h, lig, _d(p), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def eval_batch(self, ligs, poses, v=(1000.0, 1000.0, 1000.0), n_threads=16):
        """Energies of many poses (poses[i] in the POSE-float layout) on n_threads threads."""
        lg = np.ascontiguousarray(ligs, dtype=np.int32)
        P = np.zeros((len(lg), POSE))
        P[:, :np.asarray(poses).shape[1]] = poses
        e = np.zeros(len(lg))
        self.L.orc_eval_batch(self.h, len(lg), _i(lg), _d(P), _d(v), n_threads,
                              e.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return e

    def refine(self, lig, pose, refine_step=3):
        """Refinement against the explicit receptor; returns (energy, pose, within)."""
        p = np.zeros(POSE)
        p[:len(pose)] = pose
        w = ctypes.c_int()
        e = self.L.orc_refine(self.h, lig, p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), refine_step,
                              ctypes.byref(w))
        return e, p, bool(w.value)

    def nc_atom(self, t, loc, slope=1e6, v=1000.0, deriv=True):
        """One atom of XS type t at loc against the explicit receptor atoms."""
        d = np.zeros(3)
        e = self.L.orc_nc_atom(self.h, t, _d(loc), slope, v,
                               d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if deriv else None)
        return e, d

    def score_terms_nc(self, lig, pose):
        p = np.zeros(POSE)
        p[:len(pose)] = pose
        a = ctypes.c_double()
        b = ctypes.c_double()
        self.L.orc_score_terms_nc(self.h, lig, _d(p), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def eval_deriv_nc(self, lig, pose, v=(1000.0, 1000.0, 1000.0), slope=1e6):
        n_t = self.ligs[lig].n_tors
        p = np.zeros(POSE)
        p[:len(pose)] = pose
        g = np.zeros(6 + 48)
        e = self.L.orc_eval_deriv_nc(self.h, lig, _d(p), _d(v), slope, g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return e, g[:6 + n_t]

    def conf_independent(self, num_tors, x):
        return self.L.orc_conf_independent(self.h, num_tors, x)

    def set_ls_f0_zero(self, on):
        """First line search of every BFGS against f0 = 0 (the reference GPU's
        bfgs_warp) instead of f(x) (bfgs.h)."""
        self.L.orc_set_ls_f0_zero(self.h, int(bool(on)))

    def dock_cpu(self, ligs, exh, seed, global_steps, local_steps, n_threads, poses=False):
        """Every chain of the ligands: best energies and the evaluation count
        (and the best poses, len(ligs) * exh x POSE, with poses=True)."""
        ligs = np.ascontiguousarray(ligs, dtype=np.int32)
        ls = np.ascontiguousarray(local_steps, dtype=np.int32)
        be = np.zeros(len(ligs) * exh)
        ev = ctypes.c_longlong()
        P = np.zeros((len(ligs) * exh, POSE)) if poses else None
        self.L.orc_dock_cpu(self.h, len(ligs), _i(ligs), exh, seed, global_steps, _i(ls), n_threads,
                            be.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(ev),
                            P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if poses else None)
        return (be, ev.value, P) if poses else (be, ev.value)
