"""Synthetic inputs of the benchmark configurations (seeded, reproducible).

* ``c1_plain``      1,000 e- at 7.2 MeV along the force in a uniform
                    E_z = -400 kV/m, air at 1.225 kg/m^3, RZ 50 m x 400 m at 1 m.
* ``c2_at_scale``   2^24 macroparticles (70 % e-, 25 % gamma, 5 % e+); K ~ Exp(7.3 MeV)
                    clipped to [10 keV, 100 MeV]; e-/e+ within 30 deg of the force,
                    photons isotropic; rho = 1.225 exp(-h / 8.4 km) for h = 4 km + z,
                    z in [0, 10 km]; |E_z| piecewise linear 0 -> 500 kV/m (ramp over the
                    first 2 km, flat to 8 km, ramp to 0 at 10 km); RZ 256 x 1024.
* ``c3_instances``  per-instance parameters of the 256-instance batched set.

The force on an electron is -e E, so E_z < 0 drives electrons towards +z.
"""

from __future__ import annotations

import numpy as np

from .config import ParticleTile, RzFields

ME_C2_EV = 510998.95
C_M_PER_S = 299792458.0
REFERENCE_DENSITY = 1.20479


def proper_velocity(kinetic_eV: np.ndarray) -> np.ndarray:
    """u = gamma v, RreaElectronProperVelocityFromKineticEv (RreaSeedSchedule.H:30-43)."""
    gamma = 1.0 + kinetic_eV / ME_C2_EV
    return C_M_PER_S * np.sqrt(np.maximum(gamma * gamma - 1.0, 0.0))


def _cone_directions(rng, n, axis_z_sign, half_angle_rad):
    cos_max = np.cos(half_angle_rad)
    ct = 1.0 - rng.random(n) * (1.0 - cos_max)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    phi = 2.0 * np.pi * rng.random(n)
    return st * np.cos(phi), st * np.sin(phi), axis_z_sign * ct


def c1_plain(n: int = 1000):
    """Config 1: returns (fields, electrons)."""
    fields = RzFields.uniform(50, 400, 1.0, 1.0, z_lo=0.0, ez_v_per_m=-4.0e5,
                              density_ratio=1.225 / REFERENCE_DENSITY)
    k = np.full(n, 7.2e6)
    u = proper_velocity(k)
    tile = ParticleTile(
        r=np.zeros(n), z=np.full(n, 1.0), theta=np.zeros(n),
        ux=np.zeros(n), uy=np.zeros(n), uz=u.copy(), w=np.ones(n),
        id=np.arange(1, n + 1, dtype=np.int64), cpu=np.zeros(n, dtype=np.int32))
    return fields, tile


def c2_fields(nr: int = 256, nz: int = 1024, r_max_m: float = 2560.0, z_max_m: float = 10000.0,
              peak_v_per_m: float = 5.0e5) -> RzFields:
    dr = r_max_m / nr
    dz = z_max_m / nz

    def ez_profile(z):
        z = np.asarray(z, dtype=np.float64)
        ramp_up = np.clip(z / 2000.0, 0.0, 1.0)
        ramp_dn = np.clip((10000.0 - z) / 2000.0, 0.0, 1.0)
        return -peak_v_per_m * np.minimum(ramp_up, ramp_dn)

    # Ez on z-faces: value at the face z_j = j dz, f = i + nr*j
    zf = np.arange(nz + 1) * dz
    ez_faces = np.repeat(ez_profile(zf)[:, None], nr, axis=1).reshape(-1)
    alt = np.linspace(4000.0, 14000.0, 1001)
    rho = 1.225 * np.exp(-alt / 8400.0)
    return RzFields(nr, nz, dr, dz, 0.0,
                    np.zeros((nr + 1) * nz), np.zeros(nr * (nz + 1)),
                    np.zeros((nr + 1) * nz), ez_faces, np.zeros((nr + 1) * (nz + 1)),
                    density_ratio=1.0, profile_alt_m=alt, profile_rho=rho, altitude_at_z0_m=4000.0)


def c2_species_counts(total: int):
    n_e = int(round(0.70 * total))
    n_p = int(round(0.05 * total))
    n_g = total - n_e - n_p
    return n_e, n_g, n_p


def c2_electrons(n: int, fields: RzFields, seed: int = 2, first_id: int = 1, charge_sign: float = -1.0):
    """Charged particles of config 2 (electrons by default; charge_sign=+1 for
    positrons): exponential energies, a 30-degree cone about the force."""
    rng = np.random.default_rng(seed)
    k = np.clip(rng.exponential(7.3e6, n), 1.0e4, 1.0e8)
    # The force -q E_z points +z for electrons in E_z < 0.
    dx, dy, dz = _cone_directions(rng, n, 1.0 if charge_sign < 0 else -1.0, np.deg2rad(30.0))
    u = proper_velocity(k)
    r = fields.r_hi * np.sqrt(rng.random(n)) * (1.0 - 1e-9)
    z = fields.z_lo + (fields.z_hi - fields.z_lo) * rng.random(n) * (1.0 - 1e-12)
    theta = 2.0 * np.pi * rng.random(n) - np.pi
    return ParticleTile(r=r, z=z, theta=theta, ux=u * dx, uy=u * dy, uz=u * dz, w=np.ones(n),
                        id=np.arange(first_id, first_id + n, dtype=np.int64),
                        cpu=np.zeros(n, dtype=np.int32))


def c3_instances(n_instances: int = 256, seed: int = 3):
    """Per-instance |E| [V/m], density [kg/m^3] and initial electron count."""
    rng = np.random.default_rng(seed)
    e_mag = rng.uniform(2.5e5, 1.0e6, n_instances)
    rho = rng.uniform(0.3, 1.225, n_instances)
    n0 = np.floor(10.0 ** rng.uniform(3.0, 6.0, n_instances)).astype(np.int64)
    return e_mag, rho, n0


def c3_batch(n_instances: int = 256, seed: int = 3, scale: float = 1.0, nr: int = 64, nz: int = 256,
             cell_m: float = 1.0, params: tuple | None = None):
    """The batched configuration: per instance a uniform ambient E_z = -|E|
    (the force on e- along +z) over air of constant density rho on its own
    nr x nz RZ grid of cell_m cells, and N0 * scale seed electrons (at least
    one) with K ~ U[1, 10] MeV in a 30-degree cone about +z, started in a
    5 m disk at z in [1, 5] m.  ``params`` = (|E|, rho, N0) arrays replaces
    the seeded draw of the instances' parameters.  Returns (fields list,
    tiles list, seeds)."""
    e_mag, rho, n0 = c3_instances(n_instances, seed) if params is None else params
    n_instances = len(e_mag)
    rng = np.random.default_rng(seed + 1)
    fields, tiles, seeds = [], [], []
    next_id = 1
    for i in range(n_instances):
        f = RzFields.uniform(nr, nz, cell_m, cell_m, z_lo=0.0, ez_v_per_m=-float(e_mag[i]),
                             density_ratio=float(rho[i]) / REFERENCE_DENSITY)
        n = max(1, int(round(n0[i] * scale)))
        k = rng.uniform(1.0e6, 1.0e7, n)
        dx, dy, dz = _cone_directions(rng, n, 1.0, np.deg2rad(30.0))
        u = proper_velocity(k)
        r = 5.0 * np.sqrt(rng.random(n))
        z = 1.0 + 4.0 * rng.random(n)
        theta = 2.0 * np.pi * rng.random(n) - np.pi
        tiles.append(ParticleTile(r=r, z=z, theta=theta, ux=u * dx, uy=u * dy, uz=u * dz, w=np.ones(n),
                                  id=np.arange(next_id, next_id + n, dtype=np.int64),
                                  cpu=np.zeros(n, dtype=np.int32)))
        next_id += n
        fields.append(f)
        seeds.append(20240623 + 7919 * i)
    return fields, tiles, seeds


def c4_drift(n: int = 1 << 20, nr: int = 128, nz: int = 512, cell_m: float = 0.25, seed: int = 4):
    """The differentiation configuration: a 128 x 512 RZ grid of 0.25 m cells
    over air at 1.225 kg/m^3 with an ambient E_z = -400 kV/m, the Maxwell
    E_r, E_z a smooth seeded perturbation of up to 50 kV/m (the parameters),
    and n electrons with K ~ U[1, 20] MeV in a 30-degree cone about the force
    (+z), spread over r < 0.8 R and the middle 80 % of z.  Returns
    (fields, electrons)."""
    rng = np.random.default_rng(seed)
    f = RzFields.uniform(nr, nz, cell_m, cell_m, z_lo=0.0, ez_v_per_m=-4.0e5,
                         density_ratio=1.225 / REFERENCE_DENSITY)
    r_max, z_max = nr * cell_m, nz * cell_m
    # a few smooth modes on the face centres
    def smooth(rr, zz):
        v = np.zeros_like(rr)
        for _ in range(4):
            kr, kz = rng.uniform(0.5, 3.0) * np.pi / r_max, rng.uniform(0.5, 6.0) * np.pi / z_max
            v += rng.uniform(-1.0, 1.0) * np.cos(kr * rr + rng.uniform(0, 2 * np.pi)) * np.sin(
                kz * zz + rng.uniform(0, 2 * np.pi))
        return 5.0e4 * v / 4.0
    ir, jr_ = np.meshgrid(np.arange(nr + 1), np.arange(nz), indexing="xy")
    f.er = smooth(ir * cell_m, (jr_ + 0.5) * cell_m).ravel()
    iz, jz_ = np.meshgrid(np.arange(nr), np.arange(nz + 1), indexing="xy")
    f.ez = smooth((iz + 0.5) * cell_m, jz_ * cell_m).ravel()
    k = rng.uniform(1.0e6, 2.0e7, n)
    dx, dy, dz = _cone_directions(rng, n, 1.0, np.deg2rad(30.0))
    u = proper_velocity(k)
    r = 0.8 * r_max * np.sqrt(rng.random(n))
    z = z_max * (0.1 + 0.8 * rng.random(n))
    theta = 2.0 * np.pi * rng.random(n) - np.pi
    tile = ParticleTile(r=r, z=z, theta=theta, ux=u * dx, uy=u * dy, uz=u * dz, w=np.ones(n),
                        id=np.arange(1, n + 1, dtype=np.int64), cpu=np.zeros(n, dtype=np.int32))
    return f, tile


def push_photons(r, theta, z, ux, uy, uz, dt_s):
    """Straight-line photon push at c along u / |u| in Cartesian, back to
    (r, theta, z): the pushed position the photon step reads as its chord end
    (WarpX's own push for the photon species)."""
    n = np.sqrt(ux * ux + uy * uy + uz * uz)
    s = C_M_PER_S * dt_s / np.where(n > 0.0, n, 1.0)
    x = r * np.cos(theta) + ux * s
    y = r * np.sin(theta) + uy * s
    return np.hypot(x, y), np.arctan2(y, x), z + uz * s


def c2_photons(n: int, fields: RzFields, dt_s: float = 1.0e-12, seed: int = 4, first_id: int = 1):
    """Photons of config 2: exponential energies clipped to [10 keV, 100 MeV],
    isotropic directions, u = E * dir; ``extra`` holds the pushed position."""
    rng = np.random.default_rng(seed)
    k = np.clip(rng.exponential(7.3e6, n), 1.0e4, 1.0e8)
    mu = 2.0 * rng.random(n) - 1.0
    phi = 2.0 * np.pi * rng.random(n)
    st = np.sqrt(1.0 - mu * mu)
    ux, uy, uz = k * st * np.cos(phi), k * st * np.sin(phi), k * mu
    r = fields.r_hi * np.sqrt(rng.random(n)) * (1.0 - 1e-9)
    z = fields.z_lo + (fields.z_hi - fields.z_lo) * rng.random(n) * (1.0 - 1e-12)
    theta = 2.0 * np.pi * rng.random(n) - np.pi
    tile = ParticleTile(r=r, z=z, theta=theta, ux=ux, uy=uy, uz=uz, w=np.ones(n),
                        id=np.arange(first_id, first_id + n, dtype=np.int64), cpu=np.zeros(n, dtype=np.int32))
    rc, thc, zc = push_photons(r, theta, z, ux, uy, uz, dt_s)
    tile.extra = {"r_cur": rc, "theta_cur": thc, "z_cur": zc}
    return tile
