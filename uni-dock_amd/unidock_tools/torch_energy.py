"""Docking energy as a differentiable PyTorch function (PyTorch-ROCm binding).

The energy of a pose is model::eval_deriv's (unidock/src/lib/model.cpp:954-981):
grid term of every ligand atom plus intramolecular pairs, with curl caps v.
One HIP launch (ud_energy_grad_device) returns the energy together with its
gradient in the pose's own coordinates (position, quaternion w-x-y-z,
torsions) and in the scoring-term weights; backward scales those by the
incoming gradient.  The quaternion is normalised inside, so the function is
E(position, q / |q|, torsions) and its quaternion gradient is tangent to the
unit sphere at q / |q|, divided by |q|.

Usage::

    model = DockingEnergy(rec_xyz, rec_xs, ligands, center, size)
    w = torch.tensor(_lib.VINA_WEIGHTS, dtype=torch.float64, requires_grad=True)
    poses = torch.tensor(P, device="cuda", requires_grad=True)   # (n, 7 + T)
    e = model(poses, pose_ligand, w)
    e.sum().backward()      # poses.grad, w.grad

Weights that differ from the ones the maps were built with rebuild the
tables and maps (set_scoring + compute_maps) before the evaluation.

With sf = UD_SF_AD4 the receptor types are AutoDock types, receptor_charge
gives the receptor's partial charges, the ligands carry AutoDock types and
charges, and the weights are AD4.2's (vdW, H-bond, electrostatic,
desolvation, glue, torsional; the last two have zero gradient).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


class _EnergyFn(torch.autograd.Function):
    @staticmethod
    def forward(fctx, poses, weights, model, pose_ligand):
        model._use_weights(weights)
        n, k = poses.shape
        if k > _lib.UD_POSE_FLOATS:
            raise ValueError(f"poses have {k} columns, at most {_lib.UD_POSE_FLOATS}")
        P = torch.zeros((n, _lib.UD_POSE_FLOATS), dtype=torch.float32, device=poses.device)
        P[:, :k] = poses.detach()
        e = torch.empty(n, dtype=torch.float32, device=poses.device)
        gp = torch.empty((n, _lib.UD_POSE_FLOATS), dtype=torch.float32, device=poses.device)
        gw = torch.empty((n, _lib.UD_WEIGHT_FLOATS), dtype=torch.float32, device=poses.device)
        model._check_stream()
        model.ctx.energy_grad_device(pose_ligand, P.data_ptr(), e.data_ptr(), gp.data_ptr(), gw.data_ptr(), model.v)
        fctx.save_for_backward(gp[:, :k], gw[:, :weights.numel()])
        fctx.dtypes = (poses.dtype, weights.dtype)
        fctx.w_device = weights.device
        return e.to(poses.dtype)

    @staticmethod
    def backward(fctx, ge):
        gp, gw = fctx.saved_tensors
        g = ge.to(torch.float32)[:, None]
        d_poses = (g * gp).to(fctx.dtypes[0]) if fctx.needs_input_grad[0] else None
        d_w = None
        if fctx.needs_input_grad[1]:
            d_w = (g * gw).sum(0).to(device=fctx.w_device, dtype=fctx.dtypes[1])
        return d_poses, d_w, None, None


class DockingEnergy(torch.nn.Module):
    """Energy of poses of a ligand batch against one receptor, on one GPU.

    receptor_xyz (n, 3) and receptor_xs (n,) X-Score types (AutoDock types
    with sf = UD_SF_AD4, plus receptor_charge); ligands: a list of ligand
    objects (unidock_tools.synth.Ligand or ParsedLigand) or an arrays dict of
    LigandBatch / LigandSet; center and size of the search box (A).
    """

    def __init__(self, receptor_xyz, receptor_xs, ligands, center, size, granularity=0.375,
                 sf: int = _lib.UD_SF_VINA, weights=None, precision: int = _lib.UD_GRID_F32, device: int = 0,
                 v=(1000.0, 1000.0, 1000.0), receptor_charge=None):
        super().__init__()
        from . import synth
        self.device = torch.device("cuda", device)
        self._stream = torch.cuda.current_stream(self.device).cuda_stream
        self.ctx = _lib.Context(device, self._stream)
        self.sf = sf
        self.precision = precision
        self.v = tuple(float(x) for x in v)
        self._rec = (np.ascontiguousarray(receptor_xyz, dtype=np.float32),
                     np.ascontiguousarray(receptor_xs, dtype=np.int32))
        arrays = ligands if isinstance(ligands, dict) else synth.LigandBatch(list(ligands)).arrays()
        if sf == _lib.UD_SF_AD4:
            if receptor_charge is None:
                raise ValueError("AD4.2 scoring needs receptor_charge")
            self._rec_q = np.ascontiguousarray(receptor_charge, dtype=np.float32)
            self.types = sorted({_lib.ad4_grid_type(int(t)) for t in np.asarray(arrays["atom_ad"])} - {-1})
        else:
            self.types = sorted({_lib.grid_type(int(t)) for t in np.asarray(arrays["atom_xs"])} - {-1})
        self.ctx.set_box(center, size, granularity)
        w0 = _lib.DEFAULT_WEIGHTS[sf]
        self._w = None
        self._use_weights(torch.tensor(w0 if weights is None else weights, dtype=torch.float64))
        self.ctx.set_ligands(arrays)

    def _use_weights(self, weights: torch.Tensor):
        w = tuple(float(x) for x in weights.detach().cpu().reshape(-1))
        if w == self._w:
            return
        self.ctx.set_scoring(self.sf, w)
        if self.sf == _lib.UD_SF_AD4:
            self.ctx.compute_maps_ad4(self._rec[0], self._rec[1], self._rec_q, self.types, self.precision)
        else:
            self.ctx.compute_maps(self._rec[0], self._rec[1], self.types, self.precision)
        self._w = w

    def _check_stream(self):
        if torch.cuda.current_stream(self.device).cuda_stream != self._stream:
            raise RuntimeError("DockingEnergy runs on the torch stream that was current when it was created")

    def forward(self, poses: torch.Tensor, pose_ligand, weights: torch.Tensor | None = None) -> torch.Tensor:
        if poses.device != self.device:
            raise ValueError(f"poses must be on {self.device}")
        pl = np.ascontiguousarray(torch.as_tensor(pose_ligand).cpu().numpy(), dtype=np.int32)
        if len(pl) != poses.shape[0]:
            raise ValueError("pose_ligand must have one entry per pose")
        if weights is None:
            weights = torch.tensor(self._w, dtype=torch.float64)
        return _EnergyFn.apply(poses, weights, self, pl)
