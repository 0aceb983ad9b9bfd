"""ud_score_grad_device_noalloc (the allocation-free variant of
ud_score_grad_device) against ud_score_grad_device, through the C ABI: the
same kernel on the same inputs, so energies and gradients are identical.  Two
calls in a row, the second with fewer poses (the scratch array keeps the
first call's contents), then a size beyond the maximum."""
import numpy as np
import pytest

from helpers import LIGAND_TYPES, box_of, random_poses, small_ligands, small_receptor
from unidock_tools import _lib, synth

pytestmark = pytest.mark.gpu


def test_noalloc_score_grad_matches_score_grad_device():
    import torch
    print("[gpu] noalloc: context", flush=True)
    xyz, xs, cav = small_receptor()
    ligs = small_ligands()
    ctx = _lib.Context(0)
    ctx.set_scoring(_lib.UD_SF_VINA)
    c, s = box_of(cav)
    ctx.set_box(c, s, 0.375)
    ctx.compute_maps(xyz, xs, LIGAND_TYPES, _lib.UD_GRID_F16)
    ctx.set_ligands(synth.LigandBatch(list(ligs)).arrays())
    rng = np.random.default_rng(31)
    for n_per in (11, 3):  # the second call is smaller than the first
        pl, P = [], []
        for li, lig in enumerate(ligs):
            for p in random_poses(rng, lig, cav, 3.0, n_per):
                pl.append(li)
                P.append(p)
        pl = np.array(pl, dtype=np.int32)
        perm = rng.permutation(len(pl))  # poses of one ligand need not be adjacent
        pl = pl[perm]
        P = np.array(P, dtype=np.float32)[perm]
        n = len(pl)
        print(f"[gpu] noalloc: {n} poses", flush=True)
        dP = torch.tensor(P, device="cuda")
        e_ref = torch.full((n,), float("nan"), device="cuda")
        g_ref = torch.full((n, _lib.UD_GRAD_FLOATS), float("nan"), device="cuda")
        e_na = torch.full((n,), float("nan"), device="cuda")
        g_na = torch.full((n, _lib.UD_GRAD_FLOATS), float("nan"), device="cuda")
        torch.cuda.synchronize()
        ctx.score_grad_device(pl, dP.data_ptr(), e_ref.data_ptr(), g_ref.data_ptr())
        ctx.score_grad_device_noalloc(pl, dP.data_ptr(), e_na.data_ptr(), g_na.data_ptr())
        torch.cuda.synchronize()
        assert torch.isfinite(e_ref).all()
        assert torch.equal(e_na, e_ref)
        assert torch.equal(g_na, g_ref)
        # and the host entry point
        e_h, g_h = ctx.score_grad(pl, P)
        np.testing.assert_array_equal(e_na.cpu().numpy(), e_h)
    with pytest.raises(_lib.UdError) as ei:
        big = np.zeros(_lib.UD_NOALLOC_MAX_POSES + 1, dtype=np.int32)
        ctx.score_grad_device_noalloc(big, dP.data_ptr(), e_na.data_ptr(), g_na.data_ptr())
    assert ei.value.code == _lib.UD_ERR_CAPACITY
    ctx.close()
