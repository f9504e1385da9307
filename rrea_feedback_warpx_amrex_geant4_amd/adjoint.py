"""Field sensitivity of the deposited current: the drift segment's forward
model and its hand-written adjoint (``rrea_drift_forward`` /
``rrea_drift_adjoint``) behind a ``torch.autograd.Function``.

The drift step is the charged step's one-substep branch with the discrete
events and the elastic deflection frozen out (push + CSDA loss + net-chord
deposition); see ``include/rrea_hip.h``.  ``DriftProblem`` holds one
problem -- tables, geometry and density, the ambient field and Btheta, the
initial particles and the step controls -- and ``drift_current(er, ez,
problem)`` maps the gridded Maxwell fields (Er on r-faces, (nr+1)*nz; Ez on
z-faces, nr*(nz+1); float64 device tensors) to the padded net-chord grid
summed over all steps, (jr, jz), differentiably: its backward runs the
reverse sweep with checkpoints every ``checkpoint_every`` steps and recompute
in between, so the memory held between forward and backward is
ceil(steps / checkpoint_every) states per particle.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .config import ParticleTile, RzFields
from .transport import Coupling, DeviceTile, InteractionTables


@dataclass
class DriftConfig:
    steps: int = 64
    dt_s: float = 1.0e-12
    checkpoint_every: int = 8
    low_energy_cutoff_eV: float = 1.0e4
    species: str = "electron"

    def struct(self) -> _lib.DriftConfig:
        if self.species not in ("electron", "positron"):
            raise ValueError("species must be electron or positron")
        return _lib.DriftConfig(0 if self.species == "electron" else 1, int(self.steps), int(self.checkpoint_every),
                                0, float(self.dt_s), float(self.low_energy_cutoff_eV))


class DriftProblem:
    """One drift problem on the device.  ``fields`` gives the geometry, the
    density law, the ambient fields and Btheta (its own Maxwell Er/Ez are the
    default parameter values, ``er0``/``ez0``)."""

    def __init__(self, tables: InteractionTables, fields: RzFields, particles: ParticleTile, cfg: DriftConfig,
                 device="cuda"):
        self.tables = tables
        self.fields = fields
        self.cfg = cfg
        self.coupling = Coupling(fields, tables)
        f = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)  # noqa: E731
        self.amb_er, self.amb_ez, self.btheta = f(fields.amb_er), f(fields.amb_ez), f(fields.btheta)
        self.er0, self.ez0 = f(fields.er), f(fields.ez)
        self.tile0 = DeviceTile.from_host(particles, device)
        nrp, nzp = fields.padded_shape()
        self.jr_size, self.jz_size = (nrp + 1) * nzp, nrp * (nzp + 1)
        self.n_ck = -(-cfg.steps // cfg.checkpoint_every)
        self.device = device

    def _stream(self, stream):
        return torch.cuda.current_stream().cuda_stream if stream is None else stream

    def set_fields(self, er: torch.Tensor, ez: torch.Tensor, stream=None) -> None:
        """Bind Maxwell Er / Ez (device float64 face arrays) to the coupling."""
        F = self.fields
        if er.numel() != (F.nr + 1) * F.nz or ez.numel() != F.nr * (F.nz + 1):
            raise ValueError("field arrays do not match the grid")
        er, ez = er.detach().contiguous(), ez.detach().contiguous()
        if er.dtype != torch.float64 or ez.dtype != torch.float64 or not er.is_cuda or not ez.is_cuda:
            raise ValueError("fields must be float64 device tensors")
        self._bound = (er, ez)
        _lib.check(_lib.lib().rrea_fields_update(self.coupling.h, C.cast(er.data_ptr(), _lib.P_D),
                                                 C.cast(ez.data_ptr(), _lib.P_D),
                                                 C.cast(self.amb_er.data_ptr(), _lib.P_D),
                                                 C.cast(self.amb_ez.data_ptr(), _lib.P_D),
                                                 C.cast(self.btheta.data_ptr(), _lib.P_D), 1,
                                                 self._stream(stream)), "rrea_fields_update")

    def forward(self, er, ez, gjr=None, gjz=None, checkpoints: bool = True, deposit: bool = True, stream=None):
        """Run the drift from the initial particles.  Returns a dict: the
        final tile, alive flags, the grids (jr, jz) when ``deposit``, the
        per-particle losses when loss weights are given, and the checkpoints."""
        self.set_fields(er, ez, stream)
        tile = self.tile0.clone()
        n = len(tile)
        dev = self.device
        out = {"tile": tile, "alive": torch.zeros(n, dtype=torch.uint8, device=dev)}
        if deposit:
            out["jr"] = torch.zeros(self.jr_size, dtype=torch.float64, device=dev)
            out["jz"] = torch.zeros(self.jz_size, dtype=torch.float64, device=dev)
        if checkpoints:
            out["ck"] = torch.empty(self.n_ck * 6 * n, dtype=torch.float64, device=dev)
            out["ck_alive"] = torch.empty(self.n_ck * n, dtype=torch.uint8, device=dev)
        if gjr is not None:
            out["loss"] = torch.zeros(n, dtype=torch.float64, device=dev)
            gjr, gjz = gjr.contiguous(), gjz.contiguous()
        p = lambda k: out[k].data_ptr() if k in out else None  # noqa: E731
        t = tile.struct()
        c = self.cfg.struct()
        _lib.check(_lib.lib().rrea_drift_forward(
            self.tables.h, self.coupling.h, C.byref(t), C.byref(c), p("ck"), p("ck_alive"), p("alive"), p("jr"),
            p("jz"), gjr.data_ptr() if gjr is not None else None, gjz.data_ptr() if gjz is not None else None,
            p("loss"), self._stream(stream)), "rrea_drift_forward")
        return out

    def adjoint(self, er, ez, ck, ck_alive, gjr, gjz, stream=None):
        """dL/d(Er faces), dL/d(Ez faces) for L = sum(jr * gjr) + sum(jz * gjz)."""
        self.set_fields(er, ez, stream)
        F = self.fields
        g_er = torch.zeros((F.nr + 1) * F.nz, dtype=torch.float64, device=self.device)
        g_ez = torch.zeros(F.nr * (F.nz + 1), dtype=torch.float64, device=self.device)
        gjr, gjz = gjr.contiguous(), gjz.contiguous()
        t = self.tile0.struct()
        c = self.cfg.struct()
        _lib.check(_lib.lib().rrea_drift_adjoint(
            self.tables.h, self.coupling.h, C.byref(t), C.byref(c), ck.data_ptr(), ck_alive.data_ptr(),
            gjr.data_ptr(), gjz.data_ptr(), g_er.data_ptr(), g_ez.data_ptr(), self._stream(stream)),
            "rrea_drift_adjoint")
        return g_er, g_ez

    def loss(self, er, ez, gjr, gjz) -> float:
        """L = sum(jr * gjr) + sum(jz * gjz), summed per particle first (an
        order-independent value, for finite differences)."""
        out = self.forward(er, ez, gjr, gjz, checkpoints=False, deposit=False)
        return float(out["loss"].sum().item())


class DriftCurrent(torch.autograd.Function):
    """(er, ez) -> (jr, jz) of a DriftProblem, with the adjoint as backward."""

    @staticmethod
    def forward(ctx, er, ez, problem: DriftProblem):
        out = problem.forward(er, ez)
        ctx.problem = problem
        ctx.ck, ctx.ck_alive = out["ck"], out["ck_alive"]
        ctx.save_for_backward(er, ez)
        return out["jr"], out["jz"]

    @staticmethod
    def backward(ctx, g_jr, g_jz):
        er, ez = ctx.saved_tensors
        p = ctx.problem
        if g_jr is None:
            g_jr = torch.zeros(p.jr_size, dtype=torch.float64, device=er.device)
        if g_jz is None:
            g_jz = torch.zeros(p.jz_size, dtype=torch.float64, device=er.device)
        g_er, g_ez = p.adjoint(er, ez, ctx.ck, ctx.ck_alive, g_jr.to(torch.float64), g_jz.to(torch.float64))
        # the gradients take the shapes of the fields the caller passed (flat or gridded)
        return g_er.view_as(er), g_ez.view_as(ez), None


def drift_current(er: torch.Tensor, ez: torch.Tensor, problem: DriftProblem):
    """The net-chord current (jr, jz) the drift deposits over all steps."""
    return DriftCurrent.apply(er, ez, problem)
