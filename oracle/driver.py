"""The oracle's multi-step loop -- test infrastructure.

A CPU restatement of the WarpX-side loop around the hook that
``rrea_feedback_warpx_amrex_geant4_amd.simulation.Simulation`` runs on the
device, built from the oracle's own step, photon step and newborn drain:
photon push, the electron / photon / positron steps in Apply's order
(RreaParticleInteraction.cpp:57-59), the drain (:60), the commit of the
drain's survivors (CreateLocal* -> CommitLocalParticleBatch -> AddNParticles,
RreaSecondaryCommit.cpp:594-629, RreaCouplingSources.cpp:239-370) and the
removal of invalidated particles (RreaSecondaryCommit.cpp:578-592).

The commit's id assignment and insertion order are WarpX's (AddNParticles
appends in the order given, ids from one counter); the reference pins them
only through its integration runs, so they are defined here the same way as
the device loop: per species behind the live particles in FIFO order, ids
consecutive in the commit order electrons, photons, positrons.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from . import binding as ob

ELECTRON, POSITRON, PHOTON = 0, 1, 2
ME_C2_EV = 510998.95
C_M_PER_S = 299792458.0
INVALID_ID = np.iinfo(np.int64).min
KEYS = ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")
# statuses that invalidate the particle: escape, demoted, zero weight, and the
# species' consumption (annihilated e+, absorbed photon)
CHARGED_KILL = (1, 2, 3, 6)
PHOTON_KILL = (1, 2, 3, 7)


def _empty():
    d = {k: np.zeros(0) for k in KEYS[:7]}
    d["id"] = np.zeros(0, dtype=np.int64)
    d["cpu"] = np.zeros(0, dtype=np.int32)
    return d


def _as_arrays(tile) -> dict:
    if tile is None:
        return _empty()
    return {k: np.array(getattr(tile, k), dtype=(np.int64 if k == "id" else np.int32 if k == "cpu" else np.float64))
            for k in KEYS}


def proper_velocity(k):
    """RreaElectronProperVelocityFromKineticEv (RreaSeedSchedule.H:30-43)."""
    gamma = 1.0 + k / ME_C2_EV
    return C_M_PER_S * np.sqrt(np.maximum(gamma * gamma - 1.0, 0.0))


def commit_arrays(sv: np.ndarray, species: int) -> dict:
    """CommitLocalParticleBatch's particles (RreaCouplingSources.cpp:280-307)
    from one species' survivors, in their order, ids not yet assigned."""
    sel = sv[sv["species"] == species]
    keep = sel["weight"] > 0.0
    if species == PHOTON:
        keep &= sel["energy_eV"] > 0.0
    sel = sel[keep]
    px, py, pz = sel["dir_x"], sel["dir_y"], sel["dir_z"]
    e = sel["energy_eV"]
    norm = np.sqrt(px * px + py * py + pz * pz)
    pos = norm > 0.0
    safe = np.where(pos, norm, 1.0)
    if species == PHOTON:
        scale = np.where(pos, e / safe, e)
        ux, uy, uz = px * scale, py * scale, np.where(pos, pz * scale, e)
    else:
        um = proper_velocity(e)
        live = e > 0.0
        ux = np.where(live, np.where(pos, um * px / safe, 0.0), px)
        uy = np.where(live, np.where(pos, um * py / safe, 0.0), py)
        uz = np.where(live, np.where(pos, um * pz / safe, um), pz)
    return {"r": np.hypot(sel["x"], sel["y"]), "z": sel["z"].copy(), "theta": np.arctan2(sel["y"], sel["x"]),
            "ux": ux, "uy": uy, "uz": uz, "w": sel["weight"].copy(),
            "id": np.zeros(len(sel), dtype=np.int64), "cpu": np.zeros(len(sel), dtype=np.int32)}


class OracleSimulation:
    """The oracle's version of the device Simulation (same arguments)."""

    def __init__(self, otables, fields, cfg, electrons, photons=None, positrons=None, next_id=None,
                 nthreads: int = 8):
        self.t = otables
        self.fields = fields
        self.cfg = cfg
        self.nthreads = nthreads
        self.sp = {ELECTRON: _as_arrays(electrons), POSITRON: _as_arrays(positrons), PHOTON: _as_arrays(photons)}
        ids = np.concatenate([self.sp[s]["id"] for s in self.sp])
        self.next_id = int(next_id) if next_id is not None else (int(ids.max()) + 1 if len(ids) else 1)
        self.step_index = cfg.step
        self.created = np.zeros(3, dtype=np.int64)
        # the last step's net-chord deposit (index space), species steps and drain together
        self.jr = self.jz = None

    def _push_photons(self, dt):
        g = self.sp[PHOTON]
        live = g["id"] != INVALID_ID
        g["prev_r"], g["prev_z"], g["prev_theta"] = g["r"].copy(), g["z"].copy(), g["theta"].copy()
        ux, uy, uz = g["ux"], g["uy"], g["uz"]
        n = np.sqrt(ux * ux + uy * uy + uz * uz)
        s = (C_M_PER_S * dt) / np.where(n > 0.0, n, 1.0)
        x = g["r"] * np.cos(g["theta"]) + ux * s
        y = g["r"] * np.sin(g["theta"]) + uy * s
        g["r"] = np.where(live, np.hypot(x, y), g["r"])
        g["theta"] = np.where(live, np.arctan2(y, x), g["theta"])
        g["z"] = np.where(live, g["z"] + uz * s, g["z"])

    def _writeback(self, d, o, kill):
        st = o["status"]
        ok = st == 0
        for k in ("r", "z", "theta", "ux", "uy", "uz"):
            d[k] = np.where(ok, o[k], d[k])
        d["id"] = np.where(np.isin(st, kill), INVALID_ID, d["id"])

    def _deposit(self, o):
        if o.get("jr") is None:
            return
        if self.jr is None:
            self.jr, self.jz = o["jr"].copy(), o["jz"].copy()
        else:
            self.jr += o["jr"]
            self.jz += o["jz"]

    def step(self) -> dict:
        cfg = dataclasses.replace(self.cfg, step=self.step_index)
        self._push_photons(cfg.dt_s)
        self.jr = self.jz = None
        staged = []
        e = self.sp[ELECTRON]
        if len(e["r"]):
            o = ob.charged_step("electron", self.t, self.fields, cfg, e, self.nthreads, sec_cap=max(1 << 12, len(e["r"])))
            if o["code"]:
                raise ob.OracleError(f"electron step {o['code']}")
            self._deposit(o)
            self._writeback(e, o, CHARGED_KILL)
            staged.append(o["secondaries"])
        g = self.sp[PHOTON]
        if len(g["r"]):
            pin = {"r": g["prev_r"], "z": g["prev_z"], "theta": g["prev_theta"], "r_cur": g["r"], "z_cur": g["z"],
                   "theta_cur": g["theta"], "ux": g["ux"], "uy": g["uy"], "uz": g["uz"], "w": g["w"], "id": g["id"],
                   "cpu": g["cpu"]}
            o = ob.photon_step(self.t, self.fields, cfg, pin, self.nthreads, sec_cap=max(1 << 12, 4 * len(g["r"])))
            if o["code"]:
                raise ob.OracleError(f"photon step {o['code']}")
            self._writeback(g, o, PHOTON_KILL)
            staged.append(o["secondaries"])
        p = self.sp[POSITRON]
        if len(p["r"]):
            o = ob.charged_step("positron", self.t, self.fields, cfg, p, self.nthreads,
                                sec_cap=max(1 << 12, 4 * len(p["r"])))
            if o["code"]:
                raise ob.OracleError(f"positron step {o['code']}")
            self._deposit(o)
            self._writeback(p, o, CHARGED_KILL)
            staged.append(o["secondaries"])
        ini = np.concatenate(staged) if staged else np.zeros(0, dtype=ob.SECONDARY_DTYPE)
        d = ob.newborn_drain(self.t, self.fields, cfg, ini, cap=max(1 << 12, 16 * len(ini)))
        if d["code"]:
            raise ob.OracleError(f"drain {d['code']}")
        self._deposit(d)
        # deleteInvalidParticles, then AddNParticles of each species' batch
        for s, arr in self.sp.items():
            keep = arr["id"] != INVALID_ID
            for k in list(arr):
                arr[k] = arr[k][keep]
        new = {s: commit_arrays(d["survivors"], s) for s in (ELECTRON, PHOTON, POSITRON)}
        for s in (ELECTRON, PHOTON, POSITRON):
            m = len(new[s]["r"])
            new[s]["id"] = np.arange(self.next_id, self.next_id + m, dtype=np.int64)
            self.next_id += m
            self.created[s] += m
            arr = self.sp[s]
            for k in KEYS:
                arr[k] = np.concatenate([arr[k], new[s][k]])
            if s == PHOTON:
                for k in ("r", "z", "theta"):
                    arr["prev_" + k] = arr[k].copy()
        self.step_index += 1
        return {"staged": len(ini), "fifo": len(d["work"]), "survivors": len(d["survivors"])}

    def run(self, steps: int, observe=None, every: int = 1) -> list:
        rec = []
        for k in range(steps):
            self.step()
            if observe is not None and ((k + 1) % every == 0 or k + 1 == steps):
                rec.append(observe(self))
        return rec
