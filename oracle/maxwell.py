"""numpy wrappers of the oracle's field side (rrea_maxwell_oracle.cpp) and of
the reference-compiled pins of the same functions (oracle/_ref) -- TEST
INFRASTRUCTURE: only tests/ import this.

Arrays are float64 in the reference's layouts: Er (nr+1)*nz, Ez nr*(nz+1),
Btheta (nr+1)*(nz+1); the padded index-space deposit of pad cells is
jr (nr+pad+1)*(nz+2 pad), jz (nr+pad)*(nz+2 pad+1).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import binding

P_D = C.POINTER(C.c_double)
QE = 1.602176634e-19  # RreaConstants.H:6

_typed_lib = False
_typed_ref = False


def _p(a):
    return None if a is None else a.ctypes.data_as(P_D)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _lib():
    global _typed_lib
    L = binding.lib()
    if not _typed_lib:
        L.orc_current_to_physical.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                              P_D, P_D, P_D, P_D]
        L.orc_maxwell_advance_b.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, P_D, P_D, P_D, C.c_double]
        L.orc_maxwell_silver_mueller.argtypes = [C.c_int, C.c_int, P_D, P_D, P_D]
        e_args = [C.c_int, C.c_int, C.c_double, C.c_double, P_D, P_D, P_D, C.c_double] + [P_D] * 6
        L.orc_maxwell_advance_e.argtypes = e_args
        L.orc_maxwell_step.argtypes = e_args
        L.orc_maxwell_field_energy.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double] + [P_D] * 5
        L.orc_maxwell_field_energy.restype = C.c_double
        L.orc_maxwell_mixed_magnetic_energy.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, P_D, P_D]
        L.orc_maxwell_mixed_magnetic_energy.restype = C.c_double
        L.orc_maxwell_joule_power.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double] + [P_D] * 7
        L.orc_maxwell_joule_power.restype = C.c_double
        L.orc_maxwell_poynting_power.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double] + [P_D] * 5
        L.orc_maxwell_poynting_power.restype = C.c_double
        _typed_lib = True
    return L


def _ref():
    global _typed_ref
    R = binding.ref()
    if R is not None and not _typed_ref:
        R.ref_current_to_physical.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                              P_D, P_D, P_D, P_D]
        R.ref_maxwell.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P_D, P_D, P_D, C.c_double] \
            + [P_D] * 6
        R.ref_maxwell_diagnostics.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double] + [P_D] * 8
        R.ref_maxwell_mixed_magnetic_energy.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, P_D, P_D]
        R.ref_maxwell_mixed_magnetic_energy.restype = C.c_double
        _typed_ref = True
    return R


def shapes(nr, nz):
    return (nr + 1) * nz, nr * (nz + 1), (nr + 1) * (nz + 1)


def current_to_physical(nr, nz, dr, dz, jr_index, jz_index, pad=2, q=QE, use_ref=False):
    n_er, n_ez, _ = shapes(nr, nz)
    jr, jz = np.zeros(n_er), np.zeros(n_ez)
    a, b = _f64(jr_index), _f64(jz_index)
    fn = _ref().ref_current_to_physical if use_ref else _lib().orc_current_to_physical
    fn(nr, nz, dr, dz, pad, q, _p(a), _p(b), _p(jr), _p(jz))
    return jr, jz


class Field:
    """The scattered TM field (Er, Ez, Btheta) of an nr x nz mesh."""

    def __init__(self, nr, nz, dr, dz, er=None, ez=None, bt=None):
        n_er, n_ez, n_bt = shapes(nr, nz)
        self.nr, self.nz, self.dr, self.dz = nr, nz, dr, dz
        self.er = np.zeros(n_er) if er is None else np.array(er, dtype=np.float64)
        self.ez = np.zeros(n_ez) if ez is None else np.array(ez, dtype=np.float64)
        self.bt = np.zeros(n_bt) if bt is None else np.array(bt, dtype=np.float64)

    def copy(self):
        return Field(self.nr, self.nz, self.dr, self.dz, self.er, self.ez, self.bt)

    def _args(self):
        return self.nr, self.nz, self.dr, self.dz

    # op: 0 step, 1 advance B, 2 Silver-Mueller, 3 advance E
    def apply(self, op, dt=0.0, sigma_r=None, sigma_z=None, jr=None, jz=None, amb_er=None, amb_ez=None,
              use_ref=False):
        opt = [_f64(x) for x in (sigma_r, sigma_z, jr, jz, amb_er, amb_ez)]
        if use_ref:
            _ref().ref_maxwell(op, *self._args(), _p(self.er), _p(self.ez), _p(self.bt), dt, *[_p(x) for x in opt])
            return self
        L = _lib()
        if op == 0:
            L.orc_maxwell_step(*self._args(), _p(self.er), _p(self.ez), _p(self.bt), dt, *[_p(x) for x in opt])
        elif op == 1:
            L.orc_maxwell_advance_b(*self._args(), _p(self.er), _p(self.ez), _p(self.bt), dt)
        elif op == 2:
            L.orc_maxwell_silver_mueller(self.nr, self.nz, _p(self.er), _p(self.ez), _p(self.bt))
        else:
            L.orc_maxwell_advance_e(*self._args(), _p(self.er), _p(self.ez), _p(self.bt), dt,
                                    *[_p(x) for x in opt])
        return self

    def energy(self, amb_er=None, amb_ez=None, bt=None):
        b = self.bt if bt is None else _f64(bt)
        return _lib().orc_maxwell_field_energy(*self._args(), _p(self.er), _p(self.ez), _p(b), _p(_f64(amb_er)),
                                               _p(_f64(amb_ez)))

    def electric_energy(self, amb_er=None, amb_ez=None):
        return self.energy(amb_er, amb_ez, bt=np.zeros_like(self.bt))

    def mixed_magnetic_energy(self, bt_previous):
        return _lib().orc_maxwell_mixed_magnetic_energy(*self._args(), _p(self.bt), _p(_f64(bt_previous)))

    def joule(self, jr, jz, amb_er=None, amb_ez=None):
        m = C.c_double(0.0)
        p = _lib().orc_maxwell_joule_power(*self._args(), _p(self.er), _p(self.ez), _p(_f64(amb_er)),
                                           _p(_f64(amb_ez)), _p(_f64(jr)), _p(_f64(jz)), C.byref(m))
        return p, m.value

    def poynting(self, amb_er=None, amb_ez=None):
        return _lib().orc_maxwell_poynting_power(*self._args(), _p(self.er), _p(self.ez), _p(self.bt),
                                                 _p(_f64(amb_er)), _p(_f64(amb_ez)))

    def ref_diagnostics(self, amb_er=None, amb_ez=None, jr=None, jz=None):
        out = np.zeros(4)
        _ref().ref_maxwell_diagnostics(*self._args(), _p(self.er), _p(self.ez), _p(self.bt), _p(_f64(amb_er)),
                                       _p(_f64(amb_ez)), _p(_f64(jr)), _p(_f64(jz)), _p(out))
        return out

    def divergence(self, fr, fz):
        """MaxwellTMRZ::DivergenceRZ (RreaMaxwellTMRZ.H:124-131) at every cell,
        (nz, nr)."""
        nr, nz = self.nr, self.nz
        fr = np.asarray(fr).reshape(nz, nr + 1)
        fz = np.asarray(fz).reshape(nz + 1, nr)
        i = np.arange(nr)
        return ((i + 1) * fr[:, 1:] - i * fr[:, :-1]) / ((i + 0.5) * self.dr) + (fz[1:, :] - fz[:-1, :]) / self.dz
