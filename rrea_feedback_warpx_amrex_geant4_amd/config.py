"""Host-side mirrors of the reference's transport configuration and of the
coupling state the per-step hook reads.

``RreaParticleInteractionConfig`` mirrors the struct of the same name
(native_amrex_rrea/warpx_overlay/Source/Rrea/RreaParticleInteraction.H:1576-1596)
with the same field names and defaults; ``StepConfig`` adds the per-call
values the hook receives (step, dt) and the coupling getters it calls
(RngSeed, LowEnergyCutoffEv, PhotonCutoffEv).

``RzFields`` is the replicated Maxwell + ambient field state that
``RreaAmrexAdvance::GatherTotalFieldRZ`` (native_amrex_rrea/src/RreaAmrexAdvance.cpp:532-579)
gathers from, plus the density law of ``TransportDensityRatioAtZ``
(warpx_overlay/Source/Rrea/RreaWarpXCoupling.cpp:79-95): either a constant
ratio or an altitude profile of absolute density.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class RreaParticleInteractionConfig:
    density_ratio: float = 1.0
    max_optical_depth_per_substep: float = 0.05
    max_subcycles: int = 256
    strict_transport_guards: bool = True
    debug_disable_hard_moller_secondaries: bool = False
    transport_max_gs_collisions_per_substep: float = 32.0
    transport_gs_total_collision_cap: float = 1024.0
    transport_max_field_impulse_fraction: float = 0.1
    transport_field_impulse_floor_eV: float = 1.0e4
    rng_seed: int = 0


@dataclass
class StepConfig(RreaParticleInteractionConfig):
    """One call of the per-step hook: the interaction config plus the
    coupling values it reads."""

    step: int = 0
    dt_s: float = 1.0e-12
    low_energy_cutoff_eV: float = 1.0e4
    photon_cutoff_eV: float = 5.0e4
    # launch choice, not physics: route every particle through the complete
    # step instead of the fast-path specialisation (results are identical)
    complete_path_only: bool = False

    def validate(self, hard_moller_threshold_eV: float, brems_cutoff_eV: float) -> None:
        # RreaCouplingConfig.cpp:1331-1390: the charged cutoff must sit below the
        # hard-Moller threshold and the photon cutoff is the bundle's brems cutoff.
        if not (self.low_energy_cutoff_eV < hard_moller_threshold_eV):
            raise ValueError("low_energy_cutoff_eV must be below the hard-Moller threshold")
        if self.photon_cutoff_eV != brems_cutoff_eV:
            raise ValueError("photon_cutoff_eV must equal the bundle's brems_absolute_cutoff_eV")
        if not (self.max_optical_depth_per_substep > 0.0) or self.max_subcycles <= 0:
            raise ValueError("invalid optical-depth subcycle controls")
        if self.max_subcycles > 16383:
            raise ValueError("max_subcycles must be <= 16383 for RNG index packing")


@dataclass
class RzFields:
    """RZ mesh, the replicated field arrays and the density law.

    Layouts (RreaMaxwellTMRZ.H:12-14, 62-91):
      er, amb_er : (nr+1)*nz,     face f = i + (nr+1)*j
      ez, amb_ez : nr*(nz+1),     face f = i + nr*j
      btheta     : (nr+1)*(nz+1), node c = i + (nr+1)*j
    """

    nr: int
    nz: int
    dr: float
    dz: float
    z_lo: float
    er: np.ndarray
    ez: np.ndarray
    amb_er: np.ndarray
    amb_ez: np.ndarray
    btheta: np.ndarray
    density_ratio: float = 1.0
    profile_alt_m: np.ndarray | None = None
    profile_rho: np.ndarray | None = None
    altitude_at_z0_m: float = 0.0

    @property
    def r_hi(self) -> float:
        return self.nr * self.dr

    @property
    def z_hi(self) -> float:
        return self.z_lo + self.nz * self.dz

    def padded_shape(self) -> tuple[int, int]:
        """Net-chord deposit grid: +2 cells in r, 2 pad cells each side in z
        (RreaWarpXCoupling.cpp:417-444)."""
        return self.nr + 2, self.nz + 4

    @staticmethod
    def uniform(nr, nz, dr, dz, z_lo=0.0, er_v_per_m=0.0, ez_v_per_m=0.0, density_ratio=1.0):
        """Uniform ambient field, zero Maxwell field."""
        er = np.zeros((nr + 1) * nz)
        ez = np.zeros(nr * (nz + 1))
        return RzFields(nr, nz, dr, dz, z_lo, er, ez,
                        np.full((nr + 1) * nz, float(er_v_per_m)), np.full(nr * (nz + 1), float(ez_v_per_m)),
                        np.zeros((nr + 1) * (nz + 1)), density_ratio=density_ratio)


@dataclass
class ParticleTile:
    """One species' SoA tile as the reference hook sees it: the pre-push
    snapshot (r, z, theta, ux, uy, uz), the macro weight and the AMReX id/cpu
    that form the RNG key (RreaTransportStep.H:731-733)."""

    r: np.ndarray
    z: np.ndarray
    theta: np.ndarray
    ux: np.ndarray
    uy: np.ndarray
    uz: np.ndarray
    w: np.ndarray
    id: np.ndarray
    cpu: np.ndarray
    extra: dict = field(default_factory=dict)

    def __len__(self):
        return len(self.r)

    def as_dict(self) -> dict:
        return {"r": self.r, "z": self.z, "theta": self.theta, "ux": self.ux, "uy": self.uy, "uz": self.uz,
                "w": self.w, "id": self.id, "cpu": self.cpu}

    def take(self, idx) -> "ParticleTile":
        return ParticleTile(*(getattr(self, k)[idx] for k in ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")))
