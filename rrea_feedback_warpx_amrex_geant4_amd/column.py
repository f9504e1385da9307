"""The feedback column: one avalanche instance with every part of the
coupling's step on the device, in the reference's order.

Per step (RreaWarpXCoupling::AfterParticlePushBeforeBoundaries,
warpx_overlay/Source/Rrea/RreaWarpXCoupling.cpp:210-252, and
RreaAmrexAdvance::AdvanceOneStep for the kinetic current):

1. injection at the column top: ``rate_per_s * dt`` physical electrons as
   ``inject_macros`` macroparticles, kinetic energies from a power law
   E^-index over [e_min, e_max], uniform over r < ``inject_radius_m``, in a
   cone about the force (the seed-injection of the production decks; the
   overlay's own injector is WarpX input, not part of the hook);
2. ``Simulation.step``: the photon push, ``Apply`` (electrons, photons,
   positrons), the newborn drain, the commit of its survivors, compaction;
3. the field side: the step's net-chord deposit as physical current, one TM
   Maxwell step, the total field back to the gather (``MaxwellField``);
4. every ``control_every`` steps, population control of the charged species
   above 1.25 x their target (``population.thin_tile``), as the controller's
   interval does (RreaCouplingPopulation.cpp:649-684).

The atmosphere is exponential (rho = 1.225 exp(-h / 8.4 km), h = 4 km + z
over a 10 km column); the ambient E_z ramps to ``peak_v_per_m`` and is
signed so the force on electrons points down the column (-z), from the
injection plane toward the ground.
"""

from __future__ import annotations

import numpy as np
import torch

from . import population
from .config import ParticleTile, RzFields, StepConfig
from .maxwell import MaxwellField
from .simulation import ELECTRON, ME_C2_EV, POSITRON, Simulation
from .transport import InteractionTables

C_LIGHT = 299792458.0


def column_fields(peak_v_per_m: float = 4.5e5, nr: int = 256, nz: int = 1024, r_max_m: float = 2560.0,
                  z_max_m: float = 10000.0) -> RzFields:
    """The column's mesh, density profile and ambient field: E_z > 0 (force
    on electrons -z), ramping over the top and bottom 2 km."""
    dr, dz = r_max_m / nr, z_max_m / nz
    zf = np.arange(nz + 1) * dz
    prof = peak_v_per_m * np.minimum(np.clip(zf / 2000.0, 0, 1), np.clip((z_max_m - zf) / 2000.0, 0, 1))
    ez = np.repeat(prof[:, None], nr, axis=1).reshape(-1)
    alt = np.linspace(4000.0, 4000.0 + z_max_m, 1001)
    rho = 1.225 * np.exp(-alt / 8400.0)
    return RzFields(nr, nz, dr, dz, 0.0, np.zeros((nr + 1) * nz), np.zeros(nr * (nz + 1)),
                    np.zeros((nr + 1) * nz), ez, np.zeros((nr + 1) * (nz + 1)), density_ratio=1.0,
                    profile_alt_m=alt, profile_rho=rho, altitude_at_z0_m=4000.0)


def power_law_energies(rng, n: int, index: float, e_min: float, e_max: float) -> np.ndarray:
    """Inverse-CDF samples of E^-index on [e_min, e_max] (index != 1)."""
    a = 1.0 - index
    u = rng.random(n)
    return (e_min ** a + u * (e_max ** a - e_min ** a)) ** (1.0 / a)


def injection_tile(rng, n: int, weight: float, radius_m: float, z_m, first_id: int, index: float = 1.5,
                   e_min_eV: float = 1.0e6, e_max_eV: float = 1.0e8) -> ParticleTile:
    """One step's injected electrons at the column top: kinetic energies from
    E^-index over [e_min, e_max], uniform over r < radius, in a 30-degree cone
    about -z (the force on electrons), ids from first_id on.  Every draw comes
    from ``rng`` in a fixed order, so a replica generator replays it."""
    k = power_law_energies(rng, n, index, e_min_eV, e_max_eV)
    g = 1.0 + k / ME_C2_EV
    u = C_LIGHT * np.sqrt((g - 1.0) * (g + 1.0))
    mu = -(1.0 - (1.0 - np.cos(np.deg2rad(30.0))) * rng.random(n))
    phi = 2.0 * np.pi * rng.random(n)
    st = np.sqrt(1.0 - mu * mu)
    r = radius_m * np.sqrt(rng.random(n))
    theta = 2.0 * np.pi * rng.random(n) - np.pi
    return ParticleTile(r=r, z=np.broadcast_to(np.asarray(z_m, dtype=np.float64), (n,)).copy(), theta=theta,
                        ux=u * st * np.cos(phi), uy=u * st * np.sin(phi), uz=u * mu, w=np.full(n, weight),
                        id=np.arange(first_id, first_id + n, dtype=np.int64), cpu=np.zeros(n, dtype=np.int32))


class FeedbackColumn:
    def __init__(self, tables: InteractionTables, fields: RzFields | None = None, dt_s: float = 1.0e-9,
                 rate_per_s: float = 1.0e14, inject_macros: int = 256, inject_radius_m: float = 500.0,
                 inject_top_offset_m: float = 50.0, spectrum_index: float = 1.5, e_min_eV: float = 1.0e6,
                 e_max_eV: float = 1.0e8, targets=(1 << 20, 1 << 18), control_every: int = 10,
                 capacity: int = 1 << 22, rng_seed: int = 20240623, feedback: bool = True, seed: int = 1):
        self.fields = fields if fields is not None else column_fields()
        self.cfg = StepConfig(rng_seed=rng_seed, dt_s=dt_s)
        self.rng = np.random.default_rng(seed)
        self.rate, self.inject_macros = rate_per_s, inject_macros
        self.inject_radius, self.inject_z = inject_radius_m, self.fields.z_hi - inject_top_offset_m
        self.index, self.e_min, self.e_max = spectrum_index, e_min_eV, e_max_eV
        self.targets = targets
        self.control_every = control_every
        empty = ParticleTile(*(np.zeros(0) for _ in range(7)), id=np.zeros(0, dtype=np.int64),
                             cpu=np.zeros(0, dtype=np.int32))
        self.sim = Simulation(tables, self.fields, self.cfg, empty, capacity=capacity, next_id=1,
                              secondary_capacity=max(1 << 14, capacity // 16), compact_every=1)
        self.maxwell = MaxwellField(self.fields) if feedback else None
        self.injected = 0
        self.removed = np.zeros(2, dtype=np.int64)
        self.steps = 0

    def _inject(self) -> None:
        n = self.inject_macros
        if n <= 0:
            return
        store = self.sim.electrons
        first = int(self.sim.next_id.item())
        tile = injection_tile(self.rng, n, self.rate * self.cfg.dt_s / n, self.inject_radius, self.inject_z, first,
                              self.index, self.e_min, self.e_max)
        store.load(tile)
        self.sim.next_id += n
        self.sim.counts[ELECTRON] = store.n
        self.injected += n

    def step(self) -> None:
        self._inject()
        self.sim.step()
        if self.maxwell is not None:
            g = self.sim.out[0]
            self.maxwell.feedback_step(g.jr, g.jz, self.cfg.dt_s, coupling=self.sim.coupling)
        self.steps += 1
        if self.control_every > 0 and self.steps % self.control_every == 0:
            self.control()

    def control(self) -> None:
        """Population control of electrons and positrons above 1.25 x target."""
        for k, (sp, target) in enumerate(((ELECTRON, self.targets[0]), (POSITRON, self.targets[1]))):
            store = self.sim.stores[sp]
            live = store.live()
            if population.action(live, target, 0.0) != "thin":
                continue
            res = population.thin_tile(store.tile(), self.fields, target, self.cfg.rng_seed, self.sim.step_index,
                                       species=0 if sp == ELECTRON else 1)
            self.removed[k] += res.removed
        self.sim.compact()

    def run(self, steps: int) -> None:
        for _ in range(steps):
            self.step()

    def final_window_sensitivity(self, weights: torch.Tensor, steps: int = 8, checkpoint_every: int = 4,
                                 dt_s: float | None = None):
        """The adjoint of the final window's current with respect to the
        Maxwell field: the live electrons' drift segment (push, CSDA loss,
        deposit; discrete events frozen) over ``steps`` steps from the column's
        present state and field, L = sum(J_z * weights) over the padded jz
        grid, and dL/dE_r, dL/dE_z on the Maxwell faces (adjoint.DriftCurrent).
        Returns (L, dL/dEr, dL/dEz)."""
        import dataclasses

        from . import _lib
        from .adjoint import DriftConfig, DriftProblem, drift_current
        e = self.sim.electrons
        h = e.host()
        live = h["id"] != _lib.INVALID_ID
        keys = ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")
        parts = ParticleTile(*(h[k][live] for k in keys))
        f = self.fields
        if self.maxwell is not None:
            f = dataclasses.replace(f, er=self.maxwell.er.cpu().numpy(), ez=self.maxwell.ez.cpu().numpy(),
                                    btheta=self.maxwell.bt.cpu().numpy())
        prob = DriftProblem(self.sim.tables, f, parts,
                            DriftConfig(steps=steps, dt_s=dt_s or self.cfg.dt_s, checkpoint_every=checkpoint_every))
        er = prob.er0.clone().requires_grad_(True)
        ez = prob.ez0.clone().requires_grad_(True)
        _, jz = drift_current(er, ez, prob)
        loss = (jz * weights).sum()
        loss.backward()
        return loss.detach(), er.grad, ez.grad

    def observe(self) -> dict:
        e, p, g = (self.sim.stores[s] for s in (0, 1, 2))
        out = {"steps": self.steps, "electrons": e.live(), "positrons": p.live(), "photons": g.live(),
               "injected": self.injected, "removed_electrons": int(self.removed[0]),
               "removed_positrons": int(self.removed[1])}
        if self.maxwell is not None:
            out["max_scattered_ez_v_per_m"] = float(self.maxwell.ez.abs().max().item())
        w = e.w[: e.n][e.id[: e.n] != -(2 ** 63)]
        out["electron_weight"] = float(w.sum().item()) if w.numel() else 0.0
        return out


class BatchedColumn:
    """The feedback configuration's batch: ``len(peaks)`` independent columns
    of differing E_z peaks stepped together.  Per step: injection into every
    instance, the photon push, the batched species steps and drain
    (rrea_transport_step_batched, rrea_newborn_drain_batched), the commit of
    the survivors into the batched stores (each particle keeps its instance),
    compaction, then each instance's deposit drives its own Maxwell field,
    handed back to its own fields handle; every ``control_every`` steps each
    instance's charged species are thinned above 1.25 x target on its own
    mesh.  Ids are drawn from one counter for the whole batch, so an
    instance's particle ids (the RNG keys) are not those a separate run of
    that instance would assign."""

    def __init__(self, tables: InteractionTables, peaks_v_per_m, nr: int = 256, nz: int = 1024,
                 dt_s: float = 1.0e-9, rate_per_s: float = 1.0e14, inject_macros: int = 256,
                 targets=(1 << 18, 1 << 15), control_every: int = 10, capacity: int = 1 << 22,
                 rng_seed: int = 20240623, seed: int = 1, inject_radius_m: float = 500.0,
                 spectrum_index: float = 1.5, e_min_eV: float = 1.0e6, e_max_eV: float = 1.0e8):
        from .simulation import PHOTON, ParticleStore
        from .transport import Batch, Coupling, DrainOutputs, InstanceGrids, StepOutputs
        self.tables = tables
        self.fields = [column_fields(float(p), nr=nr, nz=nz) for p in peaks_v_per_m]
        self.n_inst = len(self.fields)
        self.cfg = StepConfig(rng_seed=0, dt_s=dt_s)
        self.seeds = [rng_seed + 7919 * i for i in range(self.n_inst)]
        self.couplings = [Coupling(f, tables) for f in self.fields]
        self.grids = InstanceGrids(self.fields)
        self.batch = Batch(self.couplings, self.grids.grids, self.seeds)
        self.maxwell = [MaxwellField(f) for f in self.fields]
        self.stores = [ParticleStore(s, capacity, batched=True) for s in (ELECTRON, POSITRON, PHOTON)]
        sec_cap = max(1 << 14, capacity // 16)
        self.out = [StepOutputs(0, self.fields[0], secondary_capacity=sec_cap, per_particle=False)]
        for _ in range(2):
            self.out.append(StepOutputs(0, self.fields[0], secondary_capacity=sec_cap, per_particle=False,
                                        share_grid=self.out[0]))
        self.drain = DrainOutputs(4 * sec_cap, self.out[0])
        self.next_id = torch.ones(1, dtype=torch.int64, device="cuda")
        self.counts = torch.zeros(3, dtype=torch.int64, device="cuda")
        self.rng = np.random.default_rng(seed)
        self.rate, self.inject_macros, self.radius = rate_per_s, inject_macros, inject_radius_m
        self.index, self.e_min, self.e_max = spectrum_index, e_min_eV, e_max_eV
        self.targets, self.control_every = targets, control_every
        self.steps = 0
        self.step_index = 0
        self.removed = np.zeros(2, dtype=np.int64)
        self.injected = 0
        self._photon = PHOTON

    def _inject(self):
        n = self.inject_macros
        if n <= 0:
            return
        m = n * self.n_inst
        first = int(self.next_id.item())
        z_top = np.repeat([f.z_hi - 50.0 for f in self.fields], n)
        tile = injection_tile(self.rng, m, self.rate * self.cfg.dt_s / n, self.radius, z_top, first, self.index,
                              self.e_min, self.e_max)
        e = self.stores[ELECTRON]
        e.load(tile, np.repeat(np.arange(self.n_inst, dtype=np.int32), n))
        self.next_id += m
        self.counts[ELECTRON] = e.n
        self.injected += m

    def seed_population(self, per_instance: int, seed: int = 7) -> None:
        """A start population at the configuration's scale: ``per_instance``
        electrons in every instance, drawn on the device (torch's generator)
        from the injection's spectrum, cone and radius, spread uniformly over
        the column's height, with the injection's macro weight; appended in
        instance order behind the store's count, ids from the batch's
        counter."""
        import math
        e = self.stores[ELECTRON]
        m = per_instance * self.n_inst
        if e.n + m > e.capacity:
            raise ValueError("particle store capacity exceeded")
        g = torch.Generator(device="cuda").manual_seed(seed)

        def rnd():
            return torch.rand(m, generator=g, dtype=torch.float64, device="cuda")
        a = 1.0 - self.index
        k = (self.e_min ** a + rnd() * (self.e_max ** a - self.e_min ** a)) ** (1.0 / a)
        gam = 1.0 + k / ME_C2_EV
        u = C_LIGHT * torch.sqrt((gam - 1.0) * (gam + 1.0))
        mu = -(1.0 - (1.0 - math.cos(math.radians(30.0))) * rnd())
        phi = 2.0 * math.pi * rnd()
        st = torch.sqrt(1.0 - mu * mu)
        f0 = self.fields[0]
        sl = slice(e.n, e.n + m)
        e.r[sl] = self.radius * torch.sqrt(rnd())
        e.theta[sl] = 2.0 * math.pi * rnd() - math.pi
        e.z[sl] = f0.z_lo + (f0.z_hi - f0.z_lo) * (1.0 - 1.0e-9) * rnd()
        e.ux[sl] = u * st * torch.cos(phi)
        e.uy[sl] = u * st * torch.sin(phi)
        e.uz[sl] = u * mu
        e.w[sl] = self.rate * self.cfg.dt_s / max(self.inject_macros, 1)
        first = int(self.next_id.item())
        e.id[sl] = torch.arange(first, first + m, dtype=torch.int64, device="cuda")
        e.cpu[sl] = 0
        e.inst[sl] = torch.arange(self.n_inst, dtype=torch.int32, device="cuda").repeat_interleave(per_instance)
        e.n += m
        self.next_id += m
        self.counts[ELECTRON] = e.n

    def live(self) -> int:
        """Live macroparticles of all species and instances."""
        from . import _lib
        return sum(int((s.id[: s.n] != _lib.INVALID_ID).sum().item()) for s in self.stores)

    def step(self):
        from . import _lib
        self._inject()
        self.grids.zero_()
        for o in self.out:
            o.secondary_count.zero_()
        cfg = StepConfig(rng_seed=0, step=self.step_index, dt_s=self.cfg.dt_s)
        e, p, g = (self.stores[s] for s in (ELECTRON, POSITRON, self._photon))
        if g.n:
            g.photon_tile().push(cfg.dt_s)
        for name, store, out in (("electron", e, self.out[0]), ("photon", g, self.out[1]),
                                 ("positron", p, self.out[2])):
            if store.n == 0:
                continue
            tile = store.photon_tile() if name == "photon" else store.tile()
            self.batch.step(self.tables, name, tile, store.inst[: store.n], cfg, out, check=False)
        self.batch.drain(self.tables, cfg, self.out, self.drain, check=False)
        structs = (_lib.ParticleStore * 3)(*[s.struct() for s in self.stores])
        err = self.out[0].error
        _lib.check(_lib.lib().rrea_commit_survivors(self.drain.survivors.data_ptr(),
                                                    self.drain.survivor_count.data_ptr(), structs,
                                                    self.counts.data_ptr(), self.next_id.data_ptr(), 0,
                                                    err.data_ptr(), torch.cuda.current_stream().cuda_stream),
                   "rrea_commit_survivors")
        counts = self.counts.cpu().numpy()
        self.out[0].check_error()
        for s, c in zip(self.stores, counts):
            s.n = int(c)
        self.compact()
        for i, m in enumerate(self.maxwell):
            gi = self.grids.grids[i]
            m.feedback_step(gi.jr, gi.jz, cfg.dt_s, coupling=self.couplings[i])
        self.step_index += 1
        self.steps += 1
        if self.control_every > 0 and self.steps % self.control_every == 0:
            self.control()

    def compact(self):
        import ctypes as C

        from . import _lib
        L = _lib.lib()
        for s in self.stores:
            if s.n == 0:
                continue
            t = s.photon_tile().struct() if s.photon else s.tile().struct()
            live = C.c_int64(0)
            _lib.check(L.rrea_tile_compact(self.couplings[0].h, C.byref(t), 1 if s.photon else 0, s.inst.data_ptr(),
                                           C.byref(live), torch.cuda.current_stream().cuda_stream),
                       "rrea_tile_compact")
            s.n = int(live.value)
        self.counts.copy_(torch.tensor([s.n for s in self.stores], dtype=torch.int64))

    def control(self):
        """Each instance's electrons and positrons thinned on its own mesh."""
        from . import _lib
        for k, (sp, target) in enumerate(((ELECTRON, self.targets[0]), (POSITRON, self.targets[1]))):
            store = self.stores[sp]
            if store.n == 0:
                continue
            inst = store.inst[: store.n]
            valid = store.id[: store.n] != _lib.INVALID_ID
            per = torch.bincount(inst[valid].long(), minlength=self.n_inst).cpu().numpy()
            for i in range(self.n_inst):
                if population.action(int(per[i]), target, 0.0) != "thin":
                    continue
                idx = torch.nonzero(valid & (inst == i)).flatten()
                sub = {q: getattr(store, q)[: store.n][idx].contiguous()
                       for q in ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")}
                sid = (torch.abs(sub["id"]) ^ (sub["cpu"].to(torch.int64) << 40)).contiguous()
                res = population.thin(sub["r"], sub["z"], sub["theta"], sub["ux"], sub["uy"], sub["uz"], sub["w"],
                                      sid, self.fields[i], target, self.seeds[i], self.step_index,
                                      species=0 if sp == ELECTRON else 1, id=sub["id"])
                store.w[: store.n][idx] = sub["w"]
                store.id[: store.n][idx] = sub["id"]
                self.removed[k] += res.removed
        self.compact()

    def run(self, steps: int):
        for _ in range(steps):
            self.step()

    def observe(self) -> dict:
        from . import _lib
        out = {"steps": self.steps, "injected": self.injected, "removed_electrons": int(self.removed[0]),
               "removed_positrons": int(self.removed[1])}
        for name, sp in (("electrons", ELECTRON), ("positrons", POSITRON), ("photons", self._photon)):
            s = self.stores[sp]
            valid = s.id[: s.n] != _lib.INVALID_ID
            out[name] = torch.bincount(s.inst[: s.n][valid].long(), minlength=self.n_inst).cpu().tolist()
        out["max_scattered_ez_v_per_m"] = [float(m.ez.abs().max().item()) for m in self.maxwell]
        return out

    def close(self):
        self.batch.close()
