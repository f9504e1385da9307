"""The WarpX-side loop around the transport hook, kept on the device.

One ``Simulation.step`` is what one WarpX step does to the three RREA species
around ``RreaParticleInteraction::Apply``:

1. the photon push (``rrea_photon_push``: WarpX's straight push, whose pushed
   point is the photon step's chord end; charged species need none, as the
   hook rebuilds their chord from the pre-push snapshot, which the fused
   in-place tiles hold);
2. ``Apply``: the electron, photon and positron steps in that order
   (RreaParticleInteraction.cpp:57-59), each staging its ordered secondaries,
   then the newborn residual advance of ``DrainAndCommitSecondaries`` (:60);
3. the commit of the drain's survivors (``rrea_commit_survivors``: the
   reference's CreateLocal* -> AddNParticles, RreaSecondaryCommit.cpp:594-629);
4. the removal of invalidated particles (``rrea_tile_compact``:
   deleteInvalidParticles, RreaSecondaryCommit.cpp:578-592), every
   ``compact_every`` steps.  Transport does not depend on it: invalidated
   particles are skipped wherever they sit, and the live particles keep their
   relative order, which is all the secondary order depends on.

The J grid is cleared at the start of every step, as the coupling re-arms its
net-chord grid per step (RreaWarpXCoupling.cpp:417-444).  Its consumer (the
Maxwell solve) is outside this library.
"""

from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import torch

from . import _lib
from .config import ParticleTile, RzFields, StepConfig
from .transport import (Coupling, DeviceTile, DrainOutputs, InteractionTables, PhotonTile, RreaParticleInteraction,
                        StepOutputs)

ELECTRON, POSITRON, PHOTON = 0, 1, 2
ME_C2_EV = 510998.95


class ParticleStore:
    """One species' particles with room to grow (rrea_particle_store): device
    arrays of ``capacity`` particles, the first ``n`` in use.  Photon stores
    also hold the pre-push point (prev_r, prev_z, prev_theta)."""

    def __init__(self, species: int, capacity: int, device="cuda", batched: bool = False):
        self.species = species
        self.capacity = int(capacity)
        f = lambda: torch.zeros(self.capacity, dtype=torch.float64, device=device)  # noqa: E731
        self.r, self.z, self.theta, self.ux, self.uy, self.uz, self.w = (f() for _ in range(7))
        self.id = torch.full((self.capacity,), _lib.INVALID_ID, dtype=torch.int64, device=device)
        self.cpu = torch.zeros(self.capacity, dtype=torch.int32, device=device)
        # a batched store also holds each particle's instance (rrea_particle_store.instance)
        self.inst = torch.zeros(self.capacity, dtype=torch.int32, device=device) if batched else None
        self.photon = species == PHOTON
        if self.photon:
            self.prev_r, self.prev_z, self.prev_theta = f(), f(), f()
        self.n = 0

    def load(self, tile: ParticleTile, instance: np.ndarray | None = None) -> None:
        """Append host particles (r, z, theta, u, w, id, cpu; and their
        instances in a batched store) behind the store's count."""
        m = len(tile.r)
        if self.n + m > self.capacity:
            raise ValueError("particle store capacity exceeded")
        if self.inst is not None:
            if instance is None or len(instance) != m:
                raise ValueError("a batched store needs each particle's instance")
            self.inst[self.n:self.n + m].copy_(torch.as_tensor(np.ascontiguousarray(instance, dtype=np.int32)))
        for k in ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu"):
            dst = getattr(self, k)
            dst[self.n:self.n + m].copy_(torch.as_tensor(np.ascontiguousarray(getattr(tile, k)), dtype=dst.dtype))
        if self.photon:
            for k in ("r", "z", "theta"):
                getattr(self, "prev_" + k)[self.n:self.n + m].copy_(getattr(self, k)[self.n:self.n + m])
        self.n += m

    def tile(self) -> DeviceTile:
        """The in-use particles (the current state) as a DeviceTile view."""
        return DeviceTile(*(getattr(self, k)[: self.n] for k in ("r", "z", "theta", "ux", "uy", "uz", "w", "id",
                                                                  "cpu")))

    def photon_tile(self) -> PhotonTile:
        return PhotonTile(self.prev_r[: self.n], self.prev_z[: self.n], self.prev_theta[: self.n], self.tile())

    def full_tile(self) -> DeviceTile:
        """All ``capacity`` rows as a DeviceTile (a graph-captured store)."""
        return DeviceTile(*(getattr(self, k) for k in ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")))

    def full_photon_tile(self) -> PhotonTile:
        return PhotonTile(self.prev_r, self.prev_z, self.prev_theta, self.full_tile())

    def struct(self) -> _lib.ParticleStore:
        p = lambda t: t.data_ptr()  # noqa: E731
        if self.photon:
            t = _lib.Tile(0, p(self.prev_r), p(self.prev_z), p(self.prev_theta), None, None, None, p(self.w),
                          p(self.id), p(self.cpu), p(self.r), p(self.z), p(self.theta), p(self.ux), p(self.uy),
                          p(self.uz))
        else:
            cur = [p(getattr(self, k)) for k in ("r", "z", "theta", "ux", "uy", "uz")]
            t = _lib.Tile(0, *cur, p(self.w), p(self.id), p(self.cpu), *cur)
        return _lib.ParticleStore(t, None if self.inst is None else p(self.inst), self.capacity)

    def live(self) -> int:
        return int((self.id[: self.n] != _lib.INVALID_ID).sum().item())

    def host(self) -> dict:
        """The in-use particles on the host (every array, invalidated ones included)."""
        keys = ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")
        return {k: getattr(self, k)[: self.n].cpu().numpy() for k in keys}


class Simulation:
    """One simulation's three species stepped on the device (the loop in the
    module docstring).  ``electrons``, ``photons`` and ``positrons`` are the
    initial particles (ParticleTile; photon positions are those before the
    first push).  New particles take ids from ``next_id`` on, in the
    reference's commit order."""

    def __init__(self, tables: InteractionTables, fields: RzFields, cfg: StepConfig, electrons: ParticleTile,
                 photons: ParticleTile | None = None, positrons: ParticleTile | None = None,
                 capacity: int | None = None, next_id: int | None = None, secondary_capacity: int | None = None,
                 fifo_capacity: int | None = None, compact_every: int = 1, device="cuda"):
        self.tables = tables
        self.fields = fields
        self.cfg = cfg
        self.coupling = Coupling(fields, tables)
        n0 = len(electrons.r) + (len(photons.r) if photons is not None else 0) + (
            len(positrons.r) if positrons is not None else 0)
        cap = capacity if capacity is not None else max(1 << 12, 8 * n0)
        self.stores = [ParticleStore(s, cap, device) for s in (ELECTRON, POSITRON, PHOTON)]
        ids = [electrons.id]
        self.stores[ELECTRON].load(electrons)
        if photons is not None:
            self.stores[PHOTON].load(photons)
            ids.append(photons.id)
        if positrons is not None:
            self.stores[POSITRON].load(positrons)
            ids.append(positrons.id)
        first = int(max(int(np.max(i)) for i in ids if len(i)) + 1) if n0 else 1
        self.next_id = torch.tensor([next_id if next_id is not None else first], dtype=torch.int64, device=device)
        self.counts = torch.tensor([s.n for s in self.stores], dtype=torch.int64, device=device)
        sec_cap = secondary_capacity if secondary_capacity is not None else max(1 << 12, cap // 4)
        self.out = [StepOutputs(0, fields, secondary_capacity=sec_cap, per_particle=False, device=device)]
        for _ in range(2):
            self.out.append(StepOutputs(0, fields, secondary_capacity=sec_cap, per_particle=False,
                                        share_grid=self.out[0], device=device))
        self.drain = DrainOutputs(fifo_capacity if fifo_capacity is not None else 4 * sec_cap, self.out[0],
                                  device=device)
        self.compact_every = compact_every
        self.step_index = cfg.step
        self.created = np.zeros(3, dtype=np.int64)  # committed survivors per species

    @property
    def electrons(self) -> ParticleStore:
        return self.stores[ELECTRON]

    @property
    def photons(self) -> ParticleStore:
        return self.stores[PHOTON]

    @property
    def positrons(self) -> ParticleStore:
        return self.stores[POSITRON]

    def step(self, stream: int | None = None) -> None:
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        L = _lib.lib()
        cfg = dataclasses.replace(self.cfg, step=self.step_index)
        g = self.out[0]
        for t in (g.jr, g.jz, g.rho):
            t.zero_()
        for o in self.out:
            o.secondary_count.zero_()
        ph = self.photons
        if ph.n:
            ph.photon_tile().push(cfg.dt_s, stream)
        apply = RreaParticleInteraction
        e_out, p_out, g_out = self.out[0], self.out[2], self.out[1]
        apply.apply_electrons(self.tables, self.coupling, self.electrons.tile(), cfg, e_out, stream=stream,
                              check=False)
        apply.apply_photons(self.tables, self.coupling, ph.photon_tile(), cfg, g_out, stream=stream, check=False)
        apply.apply_positrons(self.tables, self.coupling, self.positrons.tile(), cfg, p_out, stream=stream,
                              check=False)
        apply.drain_secondaries(self.tables, self.coupling, cfg, [e_out, g_out, p_out], self.drain, stream=stream,
                                check=False)
        before = self.counts.clone()
        stores = (_lib.ParticleStore * 3)(*[s.struct() for s in self.stores])
        _lib.check(L.rrea_commit_survivors(self.drain.survivors.data_ptr(), self.drain.survivor_count.data_ptr(),
                                           stores, self.counts.data_ptr(), self.next_id.data_ptr(), 0,
                                           g.error.data_ptr(), stream), "rrea_commit_survivors")
        counts = self.counts.cpu().numpy()  # the new store sizes, one round trip
        g.check_error()
        self.created += counts - before.cpu().numpy()
        for s, c in zip(self.stores, counts):
            s.n = int(c)
        self.step_index += 1
        if self.compact_every > 0 and (self.step_index - self.cfg.step) % self.compact_every == 0:
            self.compact(stream)

    def compact(self, stream: int | None = None) -> None:
        """Drop invalidated particles from every store (order kept)."""
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        L = _lib.lib()
        for s in self.stores:
            if s.n == 0:
                continue
            t = s.photon_tile().struct() if s.photon else s.tile().struct()
            live = C.c_int64(0)
            _lib.check(L.rrea_tile_compact(self.coupling.h, C.byref(t), 1 if s.photon else 0, None, C.byref(live),
                                           stream), "rrea_tile_compact")
            s.n = int(live.value)
        self.counts.copy_(torch.tensor([s.n for s in self.stores], dtype=torch.int64))

    def run(self, steps: int, observe=None, every: int = 1) -> list:
        """Step ``steps`` times; ``observe(self)`` every ``every`` steps (and
        after the last) collects what it returns."""
        rec = []
        for k in range(steps):
            self.step()
            if observe is not None and ((k + 1) % every == 0 or k + 1 == steps):
                rec.append(observe(self))
        return rec

    def close(self) -> None:
        self.coupling = None


def observables(particles: dict, runaway_eV: float = 1.0e6) -> dict:
    """Population observables of one species' host arrays (``ParticleStore.host``
    or the oracle's): live count, runaway count (kinetic energy above
    ``runaway_eV``) and their mean kinetic energy and z."""
    live = particles["id"] != _lib.INVALID_ID
    ux, uy, uz = (particles[k][live] for k in ("ux", "uy", "uz"))
    u2 = ux * ux + uy * uy + uz * uz
    c2 = 299792458.0 ** 2
    k = ME_C2_EV * (u2 / c2) / (np.sqrt(1.0 + u2 / c2) + 1.0)
    run = k > runaway_eV
    z = particles["z"][live]
    return {"live": int(live.sum()), "runaways": int(run.sum()),
            "mean_runaway_eV": float(k[run].mean()) if run.any() else 0.0,
            "mean_runaway_z": float(z[run].mean()) if run.any() else 0.0}


class BatchedAvalanche:
    """Independent avalanche instances stepped together: batched stores of
    the three species, each particle keeping its instance, each instance with
    its own fields handle, J grid and RNG seed.  A step is what
    ``Simulation.step`` does for one instance -- the photon push, the species
    steps in Apply's order, the newborn drain, the commit of the survivors --
    for every instance at once (rrea_transport_step_batched,
    rrea_newborn_drain_batched, rrea_commit_survivors).  ``step()`` runs one
    from the host; ``graph(steps)`` captures ``steps`` whole steps, commit
    included, into one HIP graph (rrea_batch_graph_* with the stores'
    counts).  Either way the stores grow on the device; ``sync_counts()``
    reads their lengths back."""

    def __init__(self, tables: InteractionTables, fields_list, electrons_list, seeds, cfg: StepConfig,
                 capacity=(1 << 20, 1 << 16, 1 << 18), secondary_capacity: int = 1 << 16,
                 fifo_capacity: int | None = None, device="cuda"):
        from .transport import Batch, InstanceGrids
        self.tables = tables
        self.fields = list(fields_list)
        self.cfg = cfg
        self.couplings = [Coupling(f, tables) for f in self.fields]
        self.grids = InstanceGrids(self.fields)
        self.batch = Batch(self.couplings, self.grids.grids, seeds)
        self.stores = [ParticleStore(s, c, device, batched=True) for s, c in zip((ELECTRON, POSITRON, PHOTON),
                                                                                  capacity)]
        sizes = [len(t.r) for t in electrons_list]
        cat = ParticleTile(*(np.concatenate([getattr(t, k) for t in electrons_list]) for k in
                             ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")))
        self.stores[ELECTRON].load(cat, np.repeat(np.arange(len(sizes), dtype=np.int32), sizes))
        self.next_id = torch.tensor([int(cat.id.max()) + 1 if len(cat.id) else 1], dtype=torch.int64, device=device)
        self.counts = torch.tensor([s.n for s in self.stores], dtype=torch.int64, device=device)
        f0 = self.fields[0]
        self.out = [StepOutputs(0, f0, secondary_capacity=secondary_capacity, per_particle=False, device=device)]
        for _ in range(2):
            self.out.append(StepOutputs(0, f0, secondary_capacity=secondary_capacity, per_particle=False,
                                        share_grid=self.out[0], device=device))
        self.drain = DrainOutputs(fifo_capacity or 4 * secondary_capacity, self.out[0], device=device)
        self._stores = (_lib.ParticleStore * 3)(*[s.struct() for s in self.stores])
        self.step_index = cfg.step

    def check_error(self) -> None:
        self.out[0].check_error()

    @property
    def electrons(self) -> ParticleStore:
        return self.stores[ELECTRON]

    @property
    def photons(self) -> ParticleStore:
        return self.stores[PHOTON]

    @property
    def positrons(self) -> ParticleStore:
        return self.stores[POSITRON]

    def step(self, stream: int | None = None) -> None:
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        cfg = dataclasses.replace(self.cfg, step=self.step_index)
        self.grids.zero_()
        for o in self.out:
            o.secondary_count.zero_()
        # every step runs over the stores' whole capacity: rows past a store's
        # length hold the invalid id and are skipped, so the commit grows the
        # stores on the device and the step needs no host copy of their lengths
        ph, el, po = self.photons, self.electrons, self.positrons
        ph.full_photon_tile().push(cfg.dt_s, stream)
        e_out, g_out, p_out = self.out
        for name, store, o in (("electron", el, e_out), ("photon", ph, g_out), ("positron", po, p_out)):
            if store.capacity:
                tile = store.full_photon_tile() if store.photon else store.full_tile()
                self.batch.step(self.tables, name, tile, store.inst, cfg, o, stream=stream, check=False)
        self.batch.drain(self.tables, cfg, [e_out, g_out, p_out], self.drain, stream=stream, check=False)
        L = _lib.lib()
        _lib.check(L.rrea_commit_survivors(self.drain.survivors.data_ptr(), self.drain.survivor_count.data_ptr(),
                                           self._stores, self.counts.data_ptr(), self.next_id.data_ptr(), 0,
                                           e_out.error.data_ptr(), stream), "rrea_commit_survivors")
        self.step_index += 1

    def graph(self, steps: int, record_capacity: int | None = None, max_generations: int = 4):
        """``steps`` whole steps, commit included, captured into one graph;
        ``launch(first_step)`` runs them and grows the stores on the device
        (read ``counts`` after a synchronisation: ``sync_counts``)."""
        from .transport import BatchGraph
        tiles = {"electron": (self.electrons.full_tile(), self.electrons.inst),
                 "positron": (self.positrons.full_tile(), self.positrons.inst),
                 "photon": (self.photons.full_photon_tile(), self.photons.inst)}
        cap = record_capacity or self.drain.capacity
        return BatchGraph(self.tables, self.batch, tiles, self.cfg, cap,
                          clear=(self.grids.jr, self.grids.jz, self.grids.rho), steps=steps,
                          max_generations=max_generations, push_photons=True,
                          survivor_capacity=cap, counts=self.counts, next_id=self.next_id)

    def sort(self, energy_bits: int = -1, stream: int | None = None) -> None:
        """WarpX's particle sort for the batch (rrea_batch_tile_sort): each
        store's rows reordered by instance, cell on the instance's mesh and
        energy band, invalidated particles behind the live ones (the
        stores keep their lengths).  Reads the lengths back first."""
        self.sync_counts()
        for st in self.stores:
            if st.n == 0:
                continue
            tile = st.photon_tile() if st.photon else st.tile()
            self.batch.sort(tile, st.inst[: st.n], st.photon, energy_bits, stream)

    def sync_counts(self) -> None:
        """The stores' lengths from the device (after graph launches)."""
        for s, c in zip(self.stores, self.counts.cpu().numpy()):
            s.n = int(c)

    def live(self) -> list:
        """Live particles per store (electrons, positrons, photons), lengths read back first."""
        self.sync_counts()
        return [s.live() for s in self.stores]
