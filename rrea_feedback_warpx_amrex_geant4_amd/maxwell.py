"""The field side of the feedback loop on the device: the step's net-chord
deposit as physical face currents, the axisymmetric TM Maxwell leapfrog they
drive, its energy ledger, and the hand-back of the total field to the
transport's gather.

Mirrors what RreaAmrexAdvance::AdvanceOneStep does for the kinetic current
(native_amrex_rrea/src/RreaAmrexAdvance.cpp:694-846, without the
low-energy fluid):

- AssembleTotalFaceCurrent (:828-873): ConvertPaddedIndexCurrentToPhysicalRZ
  of the coupling's net-chord grid with pad_cells 2 and qe per index flux
  (RreaWarpXCoupling.cpp:715-717);
- the leapfrog B -> Silver-Mueller -> E with that current (:771-781;
  RreaMaxwellTMRZ.H:221-230), the Joule and boundary-Poynting work as the mean
  over E^n and E^{n+1} (:782-794), the energy balance (:822-832);
- CommitMaxwellTotalField: the scattered field plus the ambient is what the
  next step's gather reads (rrea_fields_update with device pointers).

With one coupling substep per outer step (no fluid to resolve) the
reference's B restagger is 0.5 (dt - dt) = 0, so a step is exactly one
MaxwellStep.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .config import RzFields

QE = 1.602176634e-19  # RreaConstants.H:6
PAD_CELLS = 2         # the net-chord grid's pad (RreaWarpXCoupling.cpp:417-444)


def _p(t):
    return None if t is None else t.data_ptr()


class MaxwellField:
    """Device Er, Ez, Btheta (the scattered field) over ``fields``' mesh, with
    the static ambient it rides on and the physical face currents of the
    last conversion."""

    def __init__(self, fields: RzFields, device="cuda"):
        f64 = lambda a: torch.as_tensor(a, dtype=torch.float64, device=device).contiguous()  # noqa: E731
        self.fields = fields
        self.geom = _lib.RzGeometry(fields.nr, fields.nz, fields.dr, fields.dz, 0.0, fields.z_lo)
        self.er, self.ez, self.bt = f64(fields.er).clone(), f64(fields.ez).clone(), f64(fields.btheta).clone()
        self.amb_er, self.amb_ez = f64(fields.amb_er), f64(fields.amb_ez)
        self.jr = torch.zeros_like(self.er)
        self.jz = torch.zeros_like(self.ez)
        self.ledger_out = torch.zeros(4, dtype=torch.float64, device=device)
        # the energy ledger of the steps so far (RreaAmrexAdvance.cpp:782-832)
        self.joule_work_j = 0.0
        self.poynting_out_j = 0.0
        self.energy_start_j = None

    @staticmethod
    def _stream(stream):
        return torch.cuda.current_stream().cuda_stream if stream is None else stream

    def current_from_deposit(self, jr_index: torch.Tensor, jz_index: torch.Tensor, stream: int | None = None):
        """ConvertPaddedIndexCurrentToPhysicalRZ into self.jr, self.jz."""
        _lib.check(_lib.lib().rrea_current_to_physical(C.byref(self.geom), PAD_CELLS, QE, _p(jr_index),
                                                       _p(jz_index), _p(self.jr), _p(self.jz),
                                                       self._stream(stream)), "rrea_current_to_physical")

    def step(self, dt_s: float, current: bool = True, sigma_r=None, sigma_z=None, stream: int | None = None):
        """One MaxwellStep with the converted current (or none)."""
        j = (self.jr, self.jz) if current else (None, None)
        amb = (self.amb_er, self.amb_ez) if sigma_r is not None or sigma_z is not None else (None, None)
        _lib.check(_lib.lib().rrea_maxwell_step(C.byref(self.geom), _p(self.er), _p(self.ez), _p(self.bt), dt_s,
                                                _p(sigma_r), _p(sigma_z), _p(j[0]), _p(j[1]), _p(amb[0]),
                                                _p(amb[1]), self._stream(stream)), "rrea_maxwell_step")

    def ledger(self, current: bool = True, stream: int | None = None) -> torch.Tensor:
        """Device [field energy relative to the ambient, Joule power, boundary
        Poynting power, Jz moment] of the present state."""
        j = (self.jr, self.jz) if current else (None, None)
        _lib.check(_lib.lib().rrea_maxwell_ledger(C.byref(self.geom), _p(self.er), _p(self.ez), _p(self.bt),
                                                  _p(self.amb_er), _p(self.amb_ez), _p(j[0]), _p(j[1]),
                                                  _p(self.ledger_out), self._stream(stream)), "rrea_maxwell_ledger")
        return self.ledger_out

    def feedback_step(self, jr_index: torch.Tensor, jz_index: torch.Tensor, dt_s: float, coupling=None,
                      ledger: bool = False, stream: int | None = None):
        """The kinetic current of a step drives one Maxwell step; the total
        field is handed to ``coupling`` (a transport.Coupling) for the next
        step's gather.  With ``ledger`` the Joule and boundary work of the
        Ampere update are accumulated as the reference forms them (one host
        read per step)."""
        self.current_from_deposit(jr_index, jz_index, stream)
        if ledger:
            s = self._stream(stream)
            L = _lib.lib()
            if self.energy_start_j is None:
                self.energy_start_j = float(self.ledger(True, s)[0].item())
            _lib.check(L.rrea_maxwell_advance_b(C.byref(self.geom), _p(self.er), _p(self.ez), _p(self.bt), dt_s, s),
                       "rrea_maxwell_advance_b")
            _lib.check(L.rrea_maxwell_silver_mueller(C.byref(self.geom), _p(self.er), _p(self.ez), _p(self.bt), s),
                       "rrea_maxwell_silver_mueller")
            before = self.ledger(True, s).clone()
            _lib.check(L.rrea_maxwell_advance_e(C.byref(self.geom), _p(self.er), _p(self.ez), _p(self.bt), dt_s,
                                                None, None, _p(self.jr), _p(self.jz), None, None, s),
                       "rrea_maxwell_advance_e")
            after = self.ledger(True, s)
            b, a = before.cpu().tolist(), after.cpu().tolist()
            self.joule_work_j += 0.5 * dt_s * (b[1] + a[1])
            self.poynting_out_j += 0.5 * dt_s * (b[2] + a[2])
        else:
            self.step(dt_s, True, stream=stream)
        if coupling is not None:
            self.commit(coupling, stream)

    def commit(self, coupling, stream: int | None = None):
        """CommitMaxwellTotalField: the transport gathers Maxwell + ambient."""
        dp = lambda t: C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))  # noqa: E731
        _lib.check(_lib.lib().rrea_fields_update(coupling.h, dp(self.er), dp(self.ez), dp(self.amb_er),
                                                 dp(self.amb_ez), dp(self.bt), 1, self._stream(stream)),
                   "rrea_fields_update")

    def energy_balance_residual_j(self) -> float:
        """Delta U(ambient-relative) + Joule work + Poynting out since the
        first ledgered step (RreaAmrexAdvance.cpp:826-832); zero to round-off
        up to the magnetic energy's mixed-time staggering."""
        u = float(self.ledger(False)[0].item())
        return u - (self.energy_start_j or 0.0) + self.joule_work_j + self.poynting_out_j
