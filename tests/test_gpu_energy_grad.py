"""Differentiable energy: ud_energy_grad through the C ABI and the
torch.autograd binding (unidock_tools.torch_energy) against the FP64 oracle.

  pose gradient    E(position, q / |q|, torsions): central differences of the
                   FP64 oracle's energy (h = 1e-6, and h = 1e-3 with a bound
                   that allows for the kinks of trilinear interpolation at
                   cell faces); tolerance 3e-3 * max(1, |g|_inf) as for the
                   change-space gradient
  weight gradient  central differences over each term weight of the FP64
                   oracle, its maps and tables rebuilt per weight set
                   (h = 1e-3; the energy is linear in the weights before curl);
                   tolerance 2e-3 * max(1, |dE/dw|_inf)
  Euler identity   with curl off (v = FLT_MAX) and every atom inside the box
                   the energy is homogeneous of degree 1 in the weights:
                   sum_k w_k dE/dw_k = E (1e-4 relative)
"""
import numpy as np
import pytest

from helpers import LIGAND_TYPES, box_of, random_poses, small_ligands, small_receptor
from oracle.oracle import Oracle
from unidock_tools import _lib, synth

pytestmark = pytest.mark.gpu

EDGE = 10.0  # box edge (A): small enough for FP64 map rebuilds per weight set
FLT_MAX = 3.4028234663852886e38


def _say(msg):
    print(msg, flush=True)


def _oracle(weights, sf=0):
    xyz, xs, cav = small_receptor()
    o = Oracle(prec=64, sf=sf, weights=weights)
    c, s = box_of(cav, EDGE)
    o.set_grid(c, s, 0.375)
    o.build_maps(xyz, xs, LIGAND_TYPES)
    for lig in small_ligands():
        o.add_ligand(lig)
    return o


@pytest.fixture(scope="module")
def eg():
    xyz, xs, cav = small_receptor()
    ligs = small_ligands()
    _say("[gpu] energy_grad: creating context")
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINA)
    c, s = box_of(cav, EDGE)
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps(xyz, xs, LIGAND_TYPES, _lib.UD_GRID_F32)
    ctx.set_ligands(synth.LigandBatch(list(ligs)).arrays())
    rng = np.random.default_rng(3)
    pl, P = [], []
    for li, lig in enumerate(ligs):
        for p in random_poses(rng, lig, cav, 1.0, 4):
            # a non-unit quaternion: the function normalises it
            p[3:7] *= rng.uniform(0.5, 2.0)
            pl.append(li)
            P.append(p)
    o64 = _oracle(_lib.VINA_WEIGHTS)
    lo = o64.geo[:3]
    hi = lo + o64.geo[6:9] * (np.array(o64.n) - 1)
    yield dict(ctx=ctx, ligs=ligs, pl=np.array(pl, dtype=np.int32), P=np.array(P), o64=o64, lo=lo, hi=hi, cav=cav,
               xyz=xyz, xs=xs)
    ctx.close()


def _unit(p):
    q = p.copy()
    q[3:7] /= np.linalg.norm(q[3:7])
    return q


def test_matches_score_grad_at_unit_quaternions(eg):
    _say("[gpu] energy_grad vs score_grad")
    ctx, pl = eg["ctx"], eg["pl"]
    P = np.array([_unit(p) for p in eg["P"]])
    e, gp, gw = ctx.energy_grad(pl, P)
    e2, g2 = ctx.score_grad(pl, P)
    # the kernel renormalises the FP32-rounded unit quaternion: the pose moves
    # by ~1e-7 rad, which changes the energy by about |g| * 1e-6
    for i, li in enumerate(pl):
        nt = eg["ligs"][li].n_tors
        sc = max(1.0, np.abs(g2[i]).max())
        assert abs(e[i] - e2[i]) <= 1e-6 * abs(e2[i]) + 2e-6 * sc, (i, e[i], e2[i])
        np.testing.assert_allclose(gp[i, :3], g2[i, :3], rtol=0, atol=1e-4 * sc)
        np.testing.assert_allclose(gp[i, 7:7 + nt], g2[i, 6:6 + nt], rtol=0, atol=1e-4 * sc)
        assert np.all(gp[i, 7 + nt:] == 0)
        # tangent to the unit sphere
        assert abs(np.dot(gp[i, 3:7], P[i, 3:7])) <= 1e-5 * max(1.0, np.abs(gp[i, 3:7]).max())
    assert np.all(gw[:, 6] == 0)  # rot


def _fd_pose(o, li, p, nt, h, v=(1000.0, 1000.0, 1000.0)):
    def E(x):
        return o.eval_deriv(li, _unit(x), v)[0]
    idx = list(range(7 + nt))
    out = np.zeros(len(idx))
    for j in idx:
        a, b = p.copy(), p.copy()
        a[j] += h
        b[j] -= h
        out[j] = (E(a) - E(b)) / (2 * h)
    return out


def test_pose_gradient_central_differences(eg):
    _say("[gpu] energy_grad pose gradient vs FP64 central differences")
    ctx, pl, P, o = eg["ctx"], eg["pl"], eg["P"], eg["o64"]
    e, gp, _ = ctx.energy_grad(pl, P)
    n_kinked = n_comp = 0
    for i, li in enumerate(pl):
        nt = eg["ligs"][li].n_tors
        e64 = o.eval_deriv(li, _unit(P[i]))[0]
        assert e[i] == pytest.approx(e64, rel=1e-4, abs=1e-4 + 1e6 * 4e-6 * eg["ligs"][li].n_atoms)
        g = gp[i, :7 + nt].astype(np.float64)
        sc = max(1.0, np.abs(g).max())
        fd = _fd_pose(o, li, P[i], nt, 1e-6)
        np.testing.assert_allclose(g, fd, rtol=0, atol=3e-3 * sc, err_msg=f"pose {i}")
        fd3 = _fd_pose(o, li, P[i], nt, 1e-3)
        bad = np.abs(g - fd3) > 3e-3 * sc
        n_kinked += int(bad.sum())
        n_comp += len(g)
        assert np.all(np.abs(g - fd3) <= 0.25 * sc), (i, g, fd3)
    # a component whose +-1e-3 stencil crosses a cell face of some atom's map
    # differs by the gradient jump there; those are rare
    assert n_kinked <= 0.15 * n_comp, (n_kinked, n_comp)


def test_weight_gradient_central_differences(eg):
    _say("[gpu] energy_grad weight gradient vs FP64 central differences")
    ctx, pl, P = eg["ctx"], eg["pl"], eg["P"]
    _, _, gw = ctx.energy_grad(pl, P)
    w0 = np.array(_lib.VINA_WEIGHTS, dtype=np.float64)
    h = 1e-3
    fd = np.zeros((len(pl), 6))
    for k in range(6):
        wp, wm = w0.copy(), w0.copy()
        wp[k] += h
        wm[k] -= h
        op, om = _oracle(wp), _oracle(wm)
        for i, li in enumerate(pl):
            x = _unit(P[i])
            fd[i, k] = (op.eval_deriv(li, x)[0] - om.eval_deriv(li, x)[0]) / (2 * h)
    for i in range(len(pl)):
        sc = max(1.0, np.abs(fd[i]).max())
        np.testing.assert_allclose(gw[i, :6], fd[i], rtol=0, atol=2e-3 * sc, err_msg=f"pose {i}")
    # the terms are not all zero on these poses: gauss1, gauss2, repulsion, hydrophobic
    assert np.all(np.abs(fd[:, :4]).max(axis=0) > 0.1)


def test_weight_gradient_euler_identity(eg):
    _say("[gpu] energy_grad Euler identity (curl off)")
    ctx, pl, P = eg["ctx"], eg["pl"], eg["P"]
    v = (FLT_MAX, FLT_MAX, FLT_MAX)
    e, _, gw = ctx.energy_grad(pl, P, v)
    coords = ctx.pose_coords(pl, np.array([_unit(p) for p in P]))
    w = np.array(_lib.VINA_WEIGHTS[:6], dtype=np.float64)
    n_in = 0
    for i, li in enumerate(pl):
        c = coords[i, :eg["ligs"][li].n_atoms]
        if np.all(c >= eg["lo"]) and np.all(c <= eg["hi"]):
            n_in += 1
            assert float(np.dot(w, gw[i, :6])) == pytest.approx(float(e[i]), rel=1e-4, abs=1e-4)
    assert n_in >= 6


def test_requires_compute_maps(eg):
    ctx = _lib.Context(0)
    try:
        ctx.set_scoring(_lib.UD_SF_VINA)
        c, s = box_of(eg["cav"], EDGE)
        n = ctx.set_box(c, s, 0.375)
        ctx.set_maps(LIGAND_TYPES, np.zeros(len(LIGAND_TYPES) * int(np.prod(n)), dtype=np.float32))
        ctx.set_ligands(synth.LigandBatch(list(eg["ligs"])).arrays())
        with pytest.raises(_lib.UdError) as ei:
            ctx.energy_grad(eg["pl"][:1], eg["P"][:1])
        assert ei.value.code == _lib.UD_ERR_STATE
    finally:
        ctx.close()


def test_torch_autograd_function(eg):
    _say("[gpu] torch autograd energy")
    import torch
    from unidock_tools.torch_energy import DockingEnergy
    c, s = box_of(eg["cav"], EDGE)
    model = DockingEnergy(eg["xyz"], eg["xs"], list(eg["ligs"]), c, s)
    pl, P = eg["pl"], eg["P"]
    e_ref, gp_ref, gw_ref = eg["ctx"].energy_grad(pl, P)
    poses = torch.tensor(P[:, :7 + 8], dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.tensor(_lib.VINA_WEIGHTS, dtype=torch.float64, requires_grad=True)
    e = model(poses, pl, w)
    np.testing.assert_allclose(e.detach().cpu().numpy(), e_ref, rtol=1e-6, atol=1e-6)
    r = torch.linspace(-1.0, 2.0, len(pl), dtype=torch.float64, device="cuda")
    (e * r).sum().backward()
    rn = r.cpu().numpy()[:, None]
    np.testing.assert_allclose(poses.grad.cpu().numpy(), rn * gp_ref[:, :7 + 8], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(w.grad.numpy(), (rn * gw_ref).sum(0), rtol=1e-4, atol=1e-4)
    # new weights rebuild the maps; the change follows the weight gradient
    w2 = w.detach().clone()
    w2[0] *= 1.01
    e2 = model(poses.detach(), pl, w2).cpu().numpy()
    pred = e_ref + gw_ref[:, 0] * float(w2[0] - w.detach()[0])
    np.testing.assert_allclose(e2, pred, rtol=0, atol=1e-3 * max(1.0, np.abs(e_ref).max()))


def _vinardo_ctx(precision):
    xyz, xs, cav = small_receptor()
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINARDO)
    c, s = box_of(cav, EDGE)
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps(xyz, xs, LIGAND_TYPES, precision)
    ctx.set_ligands(synth.LigandBatch(list(small_ligands())).arrays())
    return ctx


def test_vinardo_weight_gradient(eg):
    """Vinardo (gauss1, repulsion, hydrophobic, hydrogen, glue): dE/dw against
    central differences of the FP64 oracle with maps and tables rebuilt per
    weight set, and Euler's identity with curl off for poses in the box."""
    _say("[gpu] energy_grad Vinardo weight gradient")
    pl, P = eg["pl"], np.array([_unit(p) for p in eg["P"]])
    ctx = _vinardo_ctx(_lib.UD_GRID_F32)
    try:
        _, _, gw = ctx.energy_grad(pl, P)
        ev, _, gwv = ctx.energy_grad(pl, P, (FLT_MAX,) * 3)
    finally:
        ctx.close()
    w0 = np.array(_lib.VINARDO_WEIGHTS, dtype=np.float64)
    h = 1e-3
    fd = np.zeros((len(pl), 5))
    for k in range(5):
        wp, wm = w0.copy(), w0.copy()
        wp[k] += h
        wm[k] -= h
        op, om = _oracle(wp, sf=1), _oracle(wm, sf=1)
        for i, li in enumerate(pl):
            fd[i, k] = (op.eval_deriv(li, P[i])[0] - om.eval_deriv(li, P[i])[0]) / (2 * h)
    for i in range(len(pl)):
        s_ = max(1.0, np.abs(fd[i]).max())
        np.testing.assert_allclose(gw[i, :5], fd[i], rtol=0, atol=2e-3 * s_, err_msg=f"pose {i}")
    assert np.all(gw[:, 5:] == 0)
    assert np.all(np.abs(fd[:, :3]).max(axis=0) > 0.1)
    n_in = 0
    for i, li in enumerate(pl):
        c = eg["ctx"].pose_coords(pl[i:i + 1], P[i:i + 1])[0, :eg["ligs"][li].n_atoms]
        if np.all(c >= eg["lo"]) and np.all(c <= eg["hi"]):
            n_in += 1
            assert float(np.dot(w0[:5], gwv[i, :5])) == pytest.approx(float(ev[i]), rel=1e-4, abs=1e-4)
    assert n_in >= 6


def test_fp16_maps_energy_grad(eg):
    """The FP16-map instantiation: energy and pose gradient within the FP16
    maps' own error of the FP32-map results (the maps differ by at most half
    an FP16 step times the map scale), weight gradients likewise (the per-term
    maps are FP32 in both; only the curl slope reads the FP16 map)."""
    _say("[gpu] energy_grad FP16 maps")
    pl, P = eg["pl"], np.array([_unit(p) for p in eg["P"]])
    e32, g32, w32 = eg["ctx"].energy_grad(pl, P)
    xyz, xs, cav = small_receptor()
    ctx = _lib.Context(0)
    try:
        ctx.set_scoring(_lib.UD_SF_VINA)
        c, s = box_of(cav, EDGE)
        ctx.set_box(c, s, 0.375)
        ctx.compute_maps(xyz, xs, LIGAND_TYPES, _lib.UD_GRID_F16)
        ctx.set_ligands(synth.LigandBatch(list(eg["ligs"])).arrays())
        e16, g16, w16 = ctx.energy_grad(pl, P)
        es16, gs16 = ctx.score_grad(pl, P)
    finally:
        ctx.close()
    # the search's evaluation on the same maps, up to the renormalised
    # quaternion (as test_matches_score_grad_at_unit_quaternions bounds it)
    for i in range(len(pl)):
        assert abs(e16[i] - es16[i]) <= 1e-6 * abs(es16[i]) + 2e-6 * max(1.0, np.abs(gs16[i]).max()), i
    for i in range(len(pl)):
        sc = max(1.0, abs(float(e32[i])))
        assert abs(e16[i] - e32[i]) <= 2e-3 * sc, (i, e16[i], e32[i])
        gs = max(1.0, np.abs(g32[i]).max())
        np.testing.assert_allclose(g16[i], g32[i], rtol=0, atol=5e-3 * gs)
        ws = max(1.0, np.abs(w32[i]).max())
        np.testing.assert_allclose(w16[i], w32[i], rtol=0, atol=2e-3 * ws)
