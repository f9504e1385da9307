"""Host-side mirror of the reference's per-step transport interface over the
HIP C-ABI.

``RreaParticleInteraction(config).apply_electrons(...)`` plays the role of
``RreaParticleInteraction::Apply``'s electron phase
(native_amrex_rrea/warpx_overlay/Source/Rrea/RreaParticleInteraction.cpp:12-62,
RreaElectronTransport.cpp:9-902): same configuration names, same meaning of
the pre-push snapshot, same fail-closed behaviour (a device error raises and
the caller commits nothing).  ``InteractionTables`` plays
``RreaInteractionTables::Load`` from the converted bundle and ``Coupling``
holds the field and density state of ``RreaWarpXCoupling``.

PyTorch provides device memory and the stream; all transport arithmetic is in
``librrea_hip.so``.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .config import RreaParticleInteractionConfig, RzFields, StepConfig

HERE = os.path.dirname(os.path.abspath(__file__))
BUNDLE_PATH = os.path.join(HERE, "data", "schema6_air.npz")
PAIR_BANK_PATH = os.path.join(HERE, "data", "pair_bank_air.npz")

RATE_KEYS = ["e_col", "e_soft", "e_moller", "e_brems", "e_eltr",
             "p_col", "p_soft", "p_bhabha", "p_brems", "p_eltr", "p_annih",
             "g_compton", "g_pe", "g_pairn", "g_pairt"]

SECONDARY_DTYPE = np.dtype([
    ("parent_index", "<i8"), ("species", "<i4"), ("process", "<i4"),
    ("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
    ("dir_x", "<f8"), ("dir_y", "<f8"), ("dir_z", "<f8"),
    ("weight", "<f8"), ("energy_eV", "<f8"), ("birth_fraction", "<f8"),
    ("root_key", "<u8"), ("lineage", "<u8"), ("root_species", "<i4"), ("generation", "<i4"),
    ("instance", "<i4"), ("reserved", "<i4"),
])
assert SECONDARY_DTYPE.itemsize == C.sizeof(_lib.Secondary)


def load_bundle(path: str = BUNDLE_PATH, pair_bank_path: str = PAIR_BANK_PATH) -> dict:
    """The schema-6 production bundle as converted by ``tables.convert``,
    with the pair final-state bank's training records."""
    out = {}
    for p in (path, pair_bank_path):
        with np.load(p, allow_pickle=False) as data:
            out.update({k: data[k] for k in data.files})
    return out


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_lib.P_D)


class InteractionTables:
    """Device-resident schema-6 tables (RreaInteractionTables)."""

    def __init__(self, bundle: dict | None = None):
        bundle = load_bundle() if bundle is None else bundle
        self._keep = []
        L = _lib.lib()
        im = _lib.TableImage()

        def arr(a, dtype=np.float64):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a

        for i, key in enumerate(RATE_KEYS):
            x, y = arr(bundle[key + "_x"]), arr(bundle[key + "_y"])
            im.rate[i] = _lib.RateColumn(_dptr(x), _dptr(y), len(x), int(bundle[key + "_mode"]))
        for field, key in [("moller", "moller"), ("bhabha", "bhabha"), ("electron_brems", "ebrems"),
                           ("positron_brems", "pbrems")]:
            prim, start = arr(bundle[key + "_primary"]), arr(bundle[key + "_start"], np.int64)
            u, v = arr(bundle[key + "_u"]), arr(bundle[key + "_v"])
            setattr(im, field, _lib.CdfGroups(_dptr(prim), len(prim), start.ctypes.data_as(_lib.P_I64),
                                              _dptr(u), _dptr(v)))
        for field, key in [("gs_electron", "gs_e"), ("gs_positron", "gs_p"), ("wentzel_electron", "wz_e"),
                           ("wentzel_positron", "wz_p")]:
            e, n, u, omc = (arr(bundle[key + s]) for s in ("_e", "_n", "_u", "_omc"))
            setattr(im, field, _lib.AngleTable(_dptr(e), len(e), _dptr(n), len(n), _dptr(u), len(u), _dptr(omc)))
        for field, key in [("elastic_electron", "el_e"), ("elastic_positron", "el_p")]:
            e, mfp, eta = arr(bundle[key + "_e"]), arr(bundle[key + "_mfp"]), arr(bundle[key + "_eta"])
            setattr(im, field, _lib.ElasticTable(_dptr(e), _dptr(mfp), _dptr(eta), len(e)))
        s = np.asarray(bundle["scalars"], dtype=np.float64)
        im.hard_moller_secondary_threshold_eV = s[0]
        im.brems_absolute_cutoff_eV = s[1]
        im.reference_density_kg_m3 = s[2]
        im.charged_energy_range_eV[:] = [s[3], s[4]]
        im.photon_energy_range_eV[:] = [s[5], s[6]]
        im.gs_few_collision_threshold_n = s[7]
        im.gs_isotropic_above_path_n = s[8]
        im.gs_wentzel_transition_energy_eV = s[9]
        pp, ps = arr(bundle["pe_primary"]), arr(bundle["pe_start"], np.int64)
        pu, pb = arr(bundle["pe_u"]), arr(bundle["pe_binding"])
        im.photoelectric = _lib.PhotoelectricTable(_dptr(pp), len(pp), ps.ctypes.data_as(_lib.P_I64), _dptr(pu),
                                                   _dptr(pb))
        bc, be = arr(bundle["pair_block_channel"], np.int64), arr(bundle["pair_block_energy"])
        bs, bf, bd = arr(bundle["pair_block_start"], np.int64), arr(bundle["pair_frac"]), arr(bundle["pair_dir"])
        im.pair = _lib.PairBank(bc.ctypes.data_as(_lib.P_I64), _dptr(be), len(be), bs.ctypes.data_as(_lib.P_I64),
                                _dptr(bf), _dptr(bd))
        self.scalars = s
        self.n_charged = len(bundle["e_col_x"])
        h = C.c_void_p()
        _lib.check(L.rrea_tables_create(C.byref(im), C.byref(h)), "rrea_tables_create")
        self.h = h
        self._keep = []

    @property
    def hard_moller_threshold_eV(self) -> float:
        return float(self.scalars[0])

    @property
    def brems_cutoff_eV(self) -> float:
        return float(self.scalars[1])

    @property
    def reference_density(self) -> float:
        return float(self.scalars[2])

    def hm_mean(self) -> np.ndarray:
        out = np.zeros(self.n_charged)
        _lib.check(_lib.lib().rrea_tables_hm_mean(self.h, _dptr(out), len(out)), "rrea_tables_hm_mean")
        return out

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().rrea_tables_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Coupling:
    """Field and density state the step reads (RreaWarpXCoupling services)."""

    def __init__(self, fields: RzFields, tables: InteractionTables, stream: int | None = None):
        L = _lib.lib()
        self.fields = fields
        g = _lib.RzGeometry(fields.nr, fields.nz, fields.dr, fields.dz, 0.0, fields.z_lo)
        h = C.c_void_p()
        _lib.check(L.rrea_fields_create(C.byref(g), C.byref(h)), "rrea_fields_create")
        self.h = h
        self.update(fields, stream)
        if fields.profile_alt_m is not None:
            alt = np.ascontiguousarray(fields.profile_alt_m, dtype=np.float64)
            rho = np.ascontiguousarray(fields.profile_rho, dtype=np.float64)
            _lib.check(L.rrea_density_set_profile(self.h, _dptr(alt), _dptr(rho), len(alt), fields.altitude_at_z0_m,
                                                  tables.reference_density), "rrea_density_set_profile")
        else:
            _lib.check(L.rrea_density_set_constant(self.h, fields.density_ratio), "rrea_density_set_constant")

    def set_block(self, threads: int) -> None:
        """The fast kernels' workgroup size on this handle (128, 256 or 512;
        rrea_fields_set_block): a tuning knob, results unchanged."""
        _lib.check(_lib.lib().rrea_fields_set_block(self.h, int(threads)), "rrea_fields_set_block")

    def update(self, fields: RzFields, stream: int | None = None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64)
                for a in (fields.er, fields.ez, fields.amb_er, fields.amb_ez, fields.btheta)]
        _lib.check(_lib.lib().rrea_fields_update(self.h, *[_dptr(a) for a in arrs], 0, stream),
                   "rrea_fields_update")

    def normalize_sources(self, sources: torch.Tensor, dt_s: float, diagnostics: bool = True,
                          stream: int | None = None) -> None:
        """RreaEventAccumulator::NormalizeAndSync on raw source grids, in place."""
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().rrea_sources_normalize(self.h, sources.data_ptr(), dt_s, int(diagnostics), stream),
                   "rrea_sources_normalize")

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().rrea_fields_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class DeviceTile:
    """Device SoA of one species (float64, int64 ids, int32 cpu)."""

    r: torch.Tensor
    z: torch.Tensor
    theta: torch.Tensor
    ux: torch.Tensor
    uy: torch.Tensor
    uz: torch.Tensor
    w: torch.Tensor
    id: torch.Tensor
    cpu: torch.Tensor

    @staticmethod
    def from_host(tile, device="cuda") -> "DeviceTile":
        f = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)  # noqa: E731
        return DeviceTile(*(f(getattr(tile, k), torch.float64) for k in ("r", "z", "theta", "ux", "uy", "uz", "w")),
                          f(tile.id, torch.int64), f(tile.cpu, torch.int32))

    def __len__(self):
        return self.r.numel()

    def clone(self) -> "DeviceTile":
        return DeviceTile(*(getattr(self, k).clone() for k in
                            ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")))

    def struct(self) -> _lib.Tile:
        """The fused in-place tile (r..uz are both the captured state and the writeback)."""
        p = [getattr(self, k).data_ptr() for k in ("r", "z", "theta", "ux", "uy", "uz")]
        return _lib.Tile(len(self), *p, self.w.data_ptr(), self.id.data_ptr(), self.cpu.data_ptr(), *p)

    def sort_by_cell(self, coupling: "Coupling", stream: int | None = None, energy_bits: int = -1) -> None:
        """Reorder the tile by cell and energy band (rrea_tile_sort_banded), in place."""
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        t = self.struct()
        _lib.check(_lib.lib().rrea_tile_sort_banded(coupling.h, C.byref(t), 0, energy_bits, stream), "rrea_tile_sort")


@dataclass
class PhotonTile:
    """Device SoA of the photon species: the captured pre-push point
    (prev_r, prev_z, prev_theta) and the tile after WarpX's push (``cur``:
    pushed r, z, theta, u = E dir, w, id, cpu)."""

    prev_r: torch.Tensor
    prev_z: torch.Tensor
    prev_theta: torch.Tensor
    cur: DeviceTile

    @staticmethod
    def from_host(tile, device="cuda") -> "PhotonTile":
        """``tile`` holds the pre-push point in r/z/theta and the pushed point in
        extra["r_cur"/"z_cur"/"theta_cur"]."""
        f = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)  # noqa: E731
        cur = DeviceTile.from_host(tile, device)
        cur.r, cur.z, cur.theta = f(tile.extra["r_cur"]), f(tile.extra["z_cur"]), f(tile.extra["theta_cur"])
        return PhotonTile(f(tile.r), f(tile.z), f(tile.theta), cur)

    def __len__(self):
        return len(self.cur)

    def struct(self) -> _lib.Tile:
        c = self.cur
        return _lib.Tile(len(c), self.prev_r.data_ptr(), self.prev_z.data_ptr(), self.prev_theta.data_ptr(), None,
                         None, None, c.w.data_ptr(), c.id.data_ptr(), c.cpu.data_ptr(), c.r.data_ptr(), c.z.data_ptr(),
                         c.theta.data_ptr(), c.ux.data_ptr(), c.uy.data_ptr(), c.uz.data_ptr())

    def sort_by_cell(self, coupling: "Coupling", stream: int | None = None, energy_bits: int = -1) -> None:
        """Reorder the photon tile by cell and energy (rrea_tile_sort_banded), in place."""
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        t = self.struct()
        _lib.check(_lib.lib().rrea_tile_sort_banded(coupling.h, C.byref(t), 1, energy_bits, stream), "rrea_tile_sort")

    def push(self, dt_s: float, stream: int | None = None) -> None:
        """WarpX's photon push (rrea_photon_push): prev <- current, current += c dt u/|u|."""
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        t = self.struct()
        _lib.check(_lib.lib().rrea_photon_push(C.byref(t), dt_s, stream), "rrea_photon_push")


class StepOutputs:
    """Device outputs of one step (allocated once, reused)."""

    def __init__(self, n: int, fields: RzFields, secondary_capacity: int, device="cuda", per_particle: bool = True,
                 share_grid: "StepOutputs | None" = None, grid_buffers: "tuple | None" = None,
                 sources: bool = False):
        """``share_grid``: deposit into another output's J grid, tallies and
        material sources, as every species of one step deposits into the
        coupling's one net-chord grid (RreaWarpXCoupling.cpp:417-444).
        ``grid_buffers``: (jr, jz, rho) device views to deposit into instead of
        grids of its own.  ``sources``: accumulate the material sources
        (RREA_SOURCE_FIELDS x nr x nz raw weights) into ``self.sources``."""
        nrp, nzp = fields.padded_shape()
        self.status = torch.zeros(n, dtype=torch.uint8, device=device) if per_particle else None
        self.ion_pairs = torch.zeros(n, dtype=torch.float64, device=device) if per_particle else None
        self.substeps = torch.zeros(n, dtype=torch.int32, device=device) if per_particle else None
        self.secondaries = torch.zeros(max(secondary_capacity, 1) * SECONDARY_DTYPE.itemsize, dtype=torch.uint8,
                                       device=device)
        self.secondary_capacity = secondary_capacity
        self.secondary_count = torch.zeros(1, dtype=torch.int64, device=device)
        self.sources = None
        if sources and share_grid is None:
            self.sources = torch.zeros(_lib.SOURCE_FIELDS * fields.nr * fields.nz, dtype=torch.float64, device=device)
        if share_grid is not None:
            self.jr, self.jz, self.rho = share_grid.jr, share_grid.jz, share_grid.rho
            self.tallies, self.error = share_grid.tallies, share_grid.error
            self.sources = share_grid.sources
        elif grid_buffers is not None:
            self.jr, self.jz, self.rho = grid_buffers[:3]
            if (self.jr.numel(), self.jz.numel(), self.rho.numel()) != ((nrp + 1) * nzp, nrp * (nzp + 1), nrp * nzp):
                raise ValueError("grid buffers do not match the padded grid")
            if len(grid_buffers) > 3 and grid_buffers[3] is not None:
                if grid_buffers[3].numel() != _lib.SOURCE_FIELDS * fields.nr * fields.nz:
                    raise ValueError("source buffer does not match the grid")
                self.sources = grid_buffers[3]
            self.tallies = torch.zeros(_lib.TALLY_COUNT, dtype=torch.float64, device=device)
            self.error = torch.zeros(2, dtype=torch.int32, device=device)
        else:
            self.jr = torch.zeros((nrp + 1) * nzp, dtype=torch.float64, device=device)
            self.jz = torch.zeros(nrp * (nzp + 1), dtype=torch.float64, device=device)
            self.rho = torch.zeros(nrp * nzp, dtype=torch.float64, device=device)
            self.tallies = torch.zeros(_lib.TALLY_COUNT, dtype=torch.float64, device=device)
            self.error = torch.zeros(2, dtype=torch.int32, device=device)

    def reset(self):
        for t in (self.secondary_count, self.jr, self.jz, self.rho, self.tallies, self.error, self.sources):
            if t is not None:
                t.zero_()

    def struct(self, order: bool) -> _lib.StepOut:
        p = lambda t: None if t is None else t.data_ptr()  # noqa: E731
        return _lib.StepOut(p(self.status), p(self.ion_pairs), p(self.substeps), p(self.secondaries),
                            self.secondary_capacity, p(self.secondary_count), p(self.jr), p(self.jz), p(self.rho),
                            p(self.tallies), p(self.error), int(order), p(self.sources))

    def secondaries_host(self) -> np.ndarray:
        n = min(int(self.secondary_count.item()), self.secondary_capacity)
        raw = self.secondaries[: n * SECONDARY_DTYPE.itemsize].cpu().numpy()
        return np.frombuffer(raw.tobytes(), dtype=SECONDARY_DTYPE, count=n)

    def check_error(self):
        code, index = (int(v) for v in self.error.cpu().tolist())
        if code:
            raise _lib.RreaHipError(f"device transport error: {_lib.ERRORS.get(code, code)} at particle {index}")


class DrainOutputs:
    """Device buffers of the newborn residual advance (rrea_newborn_drain):
    the FIFO, its per-entry status and ion-pair weight, and the survivors.
    The J grid, tallies and error word are those of ``grid`` (the step's
    StepOutputs), as the reference books newborn chords into the same
    coupling scratch (RreaSecondaryCommit.cpp:627)."""

    def __init__(self, capacity: int, grid: StepOutputs, device="cuda"):
        self.capacity = capacity
        self.work = torch.zeros(max(capacity, 1) * SECONDARY_DTYPE.itemsize, dtype=torch.uint8, device=device)
        self.work_count = torch.zeros(1, dtype=torch.int64, device=device)
        self.status = torch.zeros(max(capacity, 1), dtype=torch.uint8, device=device)
        self.ion_pairs = torch.zeros(max(capacity, 1), dtype=torch.float64, device=device)
        self.substeps = torch.zeros(max(capacity, 1), dtype=torch.int32, device=device)
        self.survivors = torch.zeros(max(capacity, 1) * SECONDARY_DTYPE.itemsize, dtype=torch.uint8, device=device)
        self.survivor_count = torch.zeros(1, dtype=torch.int64, device=device)
        self.grid = grid

    def stage(self, staged) -> int:
        """Concatenate the steps' ordered secondaries into the FIFO, in the
        order given (Apply's: electrons, photons, positrons).  ``staged`` is a
        list of StepOutputs, or host records (SECONDARY_DTYPE) already in
        that order."""
        size = SECONDARY_DTYPE.itemsize
        if isinstance(staged, np.ndarray):
            n = len(staged)
            if n > self.capacity:
                raise ValueError("the step's secondaries exceed the FIFO capacity")
            if n:
                raw = np.frombuffer(np.ascontiguousarray(staged, dtype=SECONDARY_DTYPE).tobytes(), dtype=np.uint8)
                self.work[: n * size].copy_(torch.from_numpy(raw.copy()))
            return n
        at = 0
        counts = torch.cat([out.secondary_count for out in staged]).cpu().tolist()  # one round trip
        for out, count in zip(staged, counts):
            n = min(int(count), out.secondary_capacity)
            if at + n > self.capacity:
                raise ValueError("the step's secondaries exceed the FIFO capacity")
            if n:
                self.work[at * size:(at + n) * size].copy_(out.secondaries[: n * size])
            at += n
        return at

    def struct(self, n_initial: int) -> _lib.DrainIO:
        g = self.grid
        return _lib.DrainIO(self.work.data_ptr(), self.capacity, n_initial, self.work_count.data_ptr(),
                            self.status.data_ptr(), self.ion_pairs.data_ptr(), self.substeps.data_ptr(),
                            self.survivors.data_ptr(),
                            self.capacity, self.survivor_count.data_ptr(), g.jr.data_ptr(), g.jz.data_ptr(),
                            g.rho.data_ptr(), g.tallies.data_ptr(), g.error.data_ptr(),
                            g.sources.data_ptr() if g.sources is not None else None)

    def work_host(self) -> np.ndarray:
        n = int(self.work_count.item())
        raw = self.work[: n * SECONDARY_DTYPE.itemsize].cpu().numpy()
        return np.frombuffer(raw.tobytes(), dtype=SECONDARY_DTYPE, count=n)

    def survivors_host(self) -> np.ndarray:
        n = int(self.survivor_count.item())
        raw = self.survivors[: n * SECONDARY_DTYPE.itemsize].cpu().numpy()
        return np.frombuffer(raw.tobytes(), dtype=SECONDARY_DTYPE, count=n)


class InstanceGrids:
    """The J grids (and, with ``sources``, the material-source grids) of many
    instances in contiguous device buffers, so one fill clears them all;
    ``grids[i]`` is instance i's StepOutputs (no per-particle outputs) over
    views of them."""

    def __init__(self, fields_list, device="cuda", sources: bool = False):
        shapes = [f.padded_shape() for f in fields_list]
        sizes = [((a + 1) * b, a * (b + 1), a * b, _lib.SOURCE_FIELDS * f.nr * f.nz if sources else 0)
                 for (a, b), f in zip(shapes, fields_list)]
        tot = [sum(s[k] for s in sizes) for k in range(4)]
        self.jr, self.jz, self.rho, self.sources = (torch.zeros(t, dtype=torch.float64, device=device) for t in tot)
        self.grids = []
        at = [0, 0, 0, 0]
        bufs = (self.jr, self.jz, self.rho, self.sources)
        for f, sz in zip(fields_list, sizes):
            views = [buf[at[k]:at[k] + sz[k]] for k, buf in enumerate(bufs)]
            if not sources:
                views[3] = None
            self.grids.append(StepOutputs(0, f, secondary_capacity=0, device=device, per_particle=False,
                                          grid_buffers=tuple(views)))
            at = [at[k] + sz[k] for k in range(4)]

    def zero_(self):
        for t in (self.jr, self.jz, self.rho, self.sources):
            t.zero_()


class Batch:
    """Independent simulation instances stepped together (rrea_batch_*): one
    Coupling (fields, density), one J grid (a StepOutputs) and one RNG seed
    per instance.  A batched tile is a DeviceTile holding every instance's
    particles plus ``inst``, each particle's instance index (int32)."""

    SPECIES = {"electron": 0, "positron": 1, "photon": 2}

    def __init__(self, couplings, grids, seeds):
        n = len(couplings)
        if not (n == len(grids) == len(seeds)) or n == 0:
            raise ValueError("one coupling, grid and seed per instance")
        self.couplings, self.grids = list(couplings), list(grids)
        arr = lambda vals: (C.c_void_p * n)(*vals)  # noqa: E731
        fh = arr([c.h.value if isinstance(c.h, C.c_void_p) else c.h for c in couplings])
        jr = arr([g.jr.data_ptr() for g in grids])
        jz = arr([g.jz.data_ptr() for g in grids])
        rho = arr([g.rho.data_ptr() for g in grids])
        sd = (C.c_uint64 * n)(*[int(x) for x in seeds])
        h = C.c_void_p()
        _lib.check(_lib.lib().rrea_batch_create(fh, n, jr, jz, rho, sd, C.byref(h)), "rrea_batch_create")
        self.h = h
        if all(g.sources is not None for g in grids):
            src = arr([g.sources.data_ptr() for g in grids])
            _lib.check(_lib.lib().rrea_batch_bind_sources(self.h, src), "rrea_batch_bind_sources")
        w = C.c_void_p()
        _lib.check(_lib.lib().rrea_batch_workspace(self.h, C.byref(w)), "rrea_batch_workspace")
        self.workspace = w

    def step(self, tables: InteractionTables, species: str, tile: DeviceTile, inst: torch.Tensor, cfg: StepConfig,
             out: StepOutputs, order_secondaries: bool = True, stream: int | None = None, check: bool = True):
        """One batched step of one species over the instances' particles (in place)."""
        if inst.dtype != torch.int32 or inst.numel() != len(tile):
            raise ValueError("inst must be int32, one per particle")
        t = tile.struct()
        c = _step_config(cfg)
        o = out.struct(order_secondaries)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().rrea_transport_step_batched(tables.h, self.h, self.SPECIES[species], C.byref(t),
                                                          inst.data_ptr(), C.byref(c), C.byref(o), stream),
                   "rrea_transport_step_batched")
        if check:
            out.check_error()

    def set_block(self, threads: int) -> None:
        """The batched fast kernels' workgroup size (on the batch's workspace
        handle; graphs created afterwards take it)."""
        _lib.check(_lib.lib().rrea_fields_set_block(self.workspace, int(threads)), "rrea_fields_set_block")

    def sort(self, tile, inst: torch.Tensor, photon: bool = False, energy_bits: int = -1,
             stream: int | None = None) -> None:
        """rrea_batch_tile_sort: the batched tile (a DeviceTile, or a
        PhotonTile with ``photon``) and its instance array reordered by
        instance, cell on the instance's mesh and energy band, in place."""
        if inst.dtype != torch.int32 or inst.numel() != len(tile):
            raise ValueError("inst must be int32, one per particle")
        t = tile.struct()
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().rrea_batch_tile_sort(self.h, C.byref(t), 1 if photon else 0, inst.data_ptr(),
                                                   int(energy_bits), stream), "rrea_batch_tile_sort")

    def drain(self, tables: InteractionTables, cfg: StepConfig, staged, out: "DrainOutputs",
              stream: int | None = None, check: bool = True):
        """The newborn drain of a batched step: every record advances in its instance."""
        n0 = out.stage(staged)
        io = out.struct(n0)
        c = _step_config(cfg)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().rrea_newborn_drain_batched(tables.h, self.h, C.byref(c), C.byref(io), stream),
                   "rrea_newborn_drain_batched")
        if check:
            out.grid.check_error()

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().rrea_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchGraph:
    """Whole batched steps captured in one HIP graph (rrea_batch_graph_*): per
    step the listed grids cleared, the optional photon push, the species
    steps in Apply's order and the device-driven newborn drain, with no
    host round trip inside.  ``tiles`` maps "electron", "positron" and
    "photon" to (tile, inst) pairs of the batch's layout.  The records
    buffer holds each step's ordered secondaries, then the whole FIFO; the
    survivors are the drain's.  With ``counts`` (device int64[3], the tiles'
    lengths) and ``next_id`` (device int64) each step also commits the
    survivors into the tiles, which are then stores whose length is the
    capacity (the avalanche).  ``launch(first_step)`` enqueues the graph's
    ``steps`` steps; outputs are read after a synchronisation."""

    def __init__(self, tables: InteractionTables, batch: Batch, tiles: dict, cfg: StepConfig, record_capacity: int,
                 clear=(), steps: int = 1, max_generations: int = 8, push_photons: bool = False,
                 survivor_capacity: int | None = None, counts: torch.Tensor | None = None,
                 next_id: torch.Tensor | None = None, cpu: int = 0, device="cuda"):
        size = SECONDARY_DTYPE.itemsize
        self.record_capacity = int(record_capacity)
        self.survivor_capacity = int(survivor_capacity if survivor_capacity is not None else record_capacity)
        self.records = torch.zeros(max(self.record_capacity, 1) * size, dtype=torch.uint8, device=device)
        self.record_count = torch.zeros(1, dtype=torch.int64, device=device)
        self.survivors = torch.zeros(max(self.survivor_capacity, 1) * size, dtype=torch.uint8, device=device)
        self.survivor_count = torch.zeros(1, dtype=torch.int64, device=device)
        self.tallies = torch.zeros(_lib.TALLY_COUNT, dtype=torch.float64, device=device)
        self.error = torch.zeros(2, dtype=torch.int32, device=device)
        self.steps = int(steps)
        clear = list(clear)
        if len(clear) > 4 or any(t.dtype != torch.float64 or not t.is_contiguous() for t in clear):
            raise ValueError("at most four contiguous float64 buffers to clear")
        d = _lib.BatchGraphDesc()
        structs = []
        for name, (tile, inst) in tiles.items():
            k = Batch.SPECIES[name]
            if inst.dtype != torch.int32 or inst.numel() != len(tile):
                raise ValueError("inst must be int32, one per particle")
            t = tile.struct()
            structs.append(t)
            d.tiles[k] = C.cast(C.pointer(t), C.c_void_p)
            d.instance[k] = inst.data_ptr()
        d.config = _step_config(cfg)
        d.records, d.record_capacity, d.record_count = (self.records.data_ptr(), self.record_capacity,
                                                        self.record_count.data_ptr())
        d.survivors, d.survivor_capacity, d.survivor_count = (self.survivors.data_ptr(), self.survivor_capacity,
                                                              self.survivor_count.data_ptr())
        d.tallies, d.error = self.tallies.data_ptr(), self.error.data_ptr()
        for k, t in enumerate(clear):
            d.clear[k] = t.data_ptr()
            d.clear_n[k] = t.numel()
        d.max_generations, d.steps, d.push_photons = int(max_generations), self.steps, 1 if push_photons else 0
        if counts is not None:
            if next_id is None or counts.dtype != torch.int64 or counts.numel() != 3 or next_id.dtype != torch.int64:
                raise ValueError("counts: device int64[3] and next_id: device int64")
            d.counts, d.next_id, d.cpu = counts.data_ptr(), next_id.data_ptr(), int(cpu)
        self._keep = (tiles, clear, counts, next_id)
        h = C.c_void_p()
        _lib.check(_lib.lib().rrea_batch_graph_create(tables.h, batch.h, C.byref(d), C.byref(h)),
                   "rrea_batch_graph_create")
        self.h = h

    def launch(self, first_step: int, stream: int | None = None):
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().rrea_batch_graph_launch(self.h, int(first_step), stream), "rrea_batch_graph_launch")

    def records_host(self) -> np.ndarray:
        n = min(int(self.record_count.item()), self.record_capacity)
        raw = self.records[: n * SECONDARY_DTYPE.itemsize].cpu().numpy()
        return np.frombuffer(raw.tobytes(), dtype=SECONDARY_DTYPE, count=n)

    def survivors_host(self) -> np.ndarray:
        n = int(self.survivor_count.item())
        raw = self.survivors[: n * SECONDARY_DTYPE.itemsize].cpu().numpy()
        return np.frombuffer(raw.tobytes(), dtype=SECONDARY_DTYPE, count=n)

    def check_error(self):
        code, index = (int(v) for v in self.error.cpu().tolist())
        if code:
            raise _lib.RreaHipError(f"device transport error: {_lib.ERRORS.get(code, code)} at index {index}")

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().rrea_batch_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _step_config(cfg: StepConfig) -> _lib.StepConfigC:
    return _lib.StepConfigC(
        cfg.rng_seed, cfg.step, cfg.dt_s, cfg.low_energy_cutoff_eV, cfg.photon_cutoff_eV,
        cfg.max_optical_depth_per_substep, cfg.max_subcycles, int(cfg.strict_transport_guards),
        int(cfg.debug_disable_hard_moller_secondaries), cfg.transport_max_gs_collisions_per_substep,
        cfg.transport_gs_total_collision_cap, cfg.transport_max_field_impulse_fraction,
        cfg.transport_field_impulse_floor_eV, int(cfg.complete_path_only), 0)


class RreaParticleInteraction:
    """Per-step transport hook (RreaParticleInteraction)."""

    def __init__(self, config: RreaParticleInteractionConfig | None = None):
        self.config = config if config is not None else RreaParticleInteractionConfig()

    def step_config(self, step: int, dt_s: float, low_energy_cutoff_eV: float, photon_cutoff_eV: float) -> StepConfig:
        c = self.config
        return StepConfig(
            density_ratio=c.density_ratio, max_optical_depth_per_substep=c.max_optical_depth_per_substep,
            max_subcycles=c.max_subcycles, strict_transport_guards=c.strict_transport_guards,
            debug_disable_hard_moller_secondaries=c.debug_disable_hard_moller_secondaries,
            transport_max_gs_collisions_per_substep=c.transport_max_gs_collisions_per_substep,
            transport_gs_total_collision_cap=c.transport_gs_total_collision_cap,
            transport_max_field_impulse_fraction=c.transport_max_field_impulse_fraction,
            transport_field_impulse_floor_eV=c.transport_field_impulse_floor_eV,
            rng_seed=c.rng_seed, step=step, dt_s=dt_s, low_energy_cutoff_eV=low_energy_cutoff_eV,
            photon_cutoff_eV=photon_cutoff_eV)

    @staticmethod
    def apply_electrons(tables: InteractionTables, coupling: Coupling, prev: DeviceTile, cfg: StepConfig,
                        out: StepOutputs, current: DeviceTile | None = None, order_secondaries: bool = True,
                        stream: int | None = None, check: bool = True) -> None:
        """One electron step (TransportStep::ApplyElectronTransport).  ``prev`` is
        the pre-push snapshot; the written-back state goes to ``current``
        (default: in place over ``prev``)."""
        RreaParticleInteraction._apply("rrea_transport_step_electrons", tables, coupling, prev, cfg, out, current,
                                       order_secondaries, stream, check)

    @staticmethod
    def apply_positrons(tables: InteractionTables, coupling: Coupling, prev: DeviceTile, cfg: StepConfig,
                        out: StepOutputs, current: DeviceTile | None = None, order_secondaries: bool = True,
                        stream: int | None = None, check: bool = True) -> None:
        """One positron step (TransportStep::ApplyPositronTransport), same
        arguments as ``apply_electrons``."""
        RreaParticleInteraction._apply("rrea_transport_step_positrons", tables, coupling, prev, cfg, out, current,
                                       order_secondaries, stream, check)

    @staticmethod
    def apply_photons(tables: InteractionTables, coupling: Coupling, tile: PhotonTile, cfg: StepConfig,
                      out: StepOutputs, order_secondaries: bool = True, stream: int | None = None,
                      check: bool = True) -> None:
        """One photon step (TransportStep::ApplyPhotonTransport) over a pushed
        photon tile; writes back the photon tile's current state in place."""
        n = len(tile)
        c = tile.cur
        if n and not all(t.is_contiguous() and t.numel() == n and t.dtype == torch.float64 for t in (
                tile.prev_r, tile.prev_z, tile.prev_theta, c.r, c.z, c.theta, c.ux, c.uy, c.uz, c.w)):
            raise ValueError("photon tile tensors must be contiguous float64 of one length")
        t = tile.struct()
        cfgc = _step_config(cfg)
        o = out.struct(order_secondaries)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().rrea_transport_step_photons(tables.h, coupling.h, C.byref(t), C.byref(cfgc),
                                                           C.byref(o), stream), "rrea_transport_step_photons")
        if check:
            out.check_error()

    @staticmethod
    def drain_secondaries(tables: InteractionTables, coupling: Coupling, cfg: StepConfig, staged, out: DrainOutputs,
                          stream: int | None = None, check: bool = True) -> None:
        """The newborn residual advance of TransportStep::DrainAndCommitSecondaries
        (RreaSecondaryCommit.cpp:338-468): stage the species' ordered
        secondaries (Apply's order) into the FIFO and fly every newborn
        through the rest of the step.  Survivors are left in ``out.survivors``
        for the caller's AddNParticles."""
        n0 = out.stage(staged)
        io = out.struct(n0)
        c = _step_config(cfg)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().rrea_newborn_drain(tables.h, coupling.h, C.byref(c), C.byref(io), stream),
                   "rrea_newborn_drain")
        if check:
            out.grid.check_error()

    @staticmethod
    def _apply(entry: str, tables: InteractionTables, coupling: Coupling, prev: DeviceTile, cfg: StepConfig,
               out: StepOutputs, current: DeviceTile | None, order_secondaries: bool, stream: int | None,
               check: bool) -> None:
        cur = prev if current is None else current
        n = len(prev)
        if n and not all(t.is_contiguous() and t.numel() == n for t in (
                prev.r, prev.z, prev.theta, prev.ux, prev.uy, prev.uz, prev.w, prev.id, prev.cpu,
                cur.r, cur.z, cur.theta, cur.ux, cur.uy, cur.uz)):
            raise ValueError("tile tensors must be contiguous and of one length")
        if n and (prev.r.dtype != torch.float64 or prev.id.dtype != torch.int64 or prev.cpu.dtype != torch.int32):
            raise ValueError("tile dtypes must be float64 / int64 id / int32 cpu")
        tile = _lib.Tile(n, prev.r.data_ptr(), prev.z.data_ptr(), prev.theta.data_ptr(), prev.ux.data_ptr(),
                         prev.uy.data_ptr(), prev.uz.data_ptr(), prev.w.data_ptr(), prev.id.data_ptr(),
                         prev.cpu.data_ptr(), cur.r.data_ptr(), cur.z.data_ptr(), cur.theta.data_ptr(),
                         cur.ux.data_ptr(), cur.uy.data_ptr(), cur.uz.data_ptr())
        c = _step_config(cfg)
        o = out.struct(order_secondaries)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(getattr(_lib.lib(), entry)(tables.h, coupling.h, C.byref(tile), C.byref(c), C.byref(o), stream),
                   entry)
        if check:
            out.check_error()


class ConcurrentApply:
    """One Apply with its three species passes on three streams at once.

    The electron, photon and positron steps of one Apply read only their own
    tile and the (unchanging) fields, and write their own secondaries and,
    through device atomics, the shared J grid, sources and tallies; none
    reads another's output before the drain.  Run concurrently they give the
    sequential Apply's results (the J and source sums up to FP64 summation
    order).  Each pass gets its own fields handle (the argument block and the
    ordering workspace belong to a handle) and its own stream; every pass
    orders its secondaries with the count on the device, so one host thread
    enqueues all three without a round trip.  The photon and positron passes
    wait for the end of the electron fast-path kernel
    (rrea_fields_signal_fast_path), so they run under the latency-bound
    electron hand-over rather than competing with the electron kernel;
    ``early`` names the passes that start with the electron pass instead
    (``("positron",)``: the small positron pass and its own hand-over run
    under the electron fast kernel, and the tail after it holds only the
    electron and photon hand-overs).  The drain then runs on the caller's
    stream after all three."""

    def __init__(self, tables: InteractionTables, fields: RzFields, early: tuple = ()):
        if not set(early) <= {"photon", "positron"}:
            raise ValueError(f"early passes must be 'photon' and/or 'positron', not {early}")
        self.tables = tables
        self.early = tuple(early)
        self.couplings = [Coupling(fields, tables) for _ in range(3)]  # electrons, photons, positrons
        self.streams = [torch.cuda.Stream(), torch.cuda.Stream()]      # default priority
        _lib.check(_lib.lib().rrea_fields_signal_fast_path(self.couplings[0].h, 1), "rrea_fields_signal_fast_path")

    def run(self, electrons: DeviceTile, photons: "PhotonTile", positrons: DeviceTile, cfg: StepConfig,
            out_e: StepOutputs, out_g: StepOutputs, out_p: StepOutputs, drain: "DrainOutputs | None" = None,
            check: bool = False, push_photons: bool = False, device_drain: int = 0, pre_push: bool = False) -> None:
        """``push_photons``: WarpX's photon push for the next step
        (rrea_photon_push) right after the photon pass, on its stream, so it
        runs under the other passes instead of after the drain (nothing
        after the photon pass reads the photon tile).  ``pre_push``: the push
        of this step instead, on the photon stream right before the photon
        pass, so it runs with that pass (under the electron hand-over unless
        the photon pass is early; nothing before the photon pass reads the
        photon tile); photons the caller commits
        after the drain are then pushed by the next step's run."""
        if push_photons and pre_push:
            raise ValueError("push_photons and pre_push each push the photons once per step: choose one")
        main = torch.cuda.current_stream()
        ready = torch.cuda.Event()
        ready.record(main)
        for s in self.streams:
            s.wait_event(ready)
        apply = RreaParticleInteraction
        cg, cp = self.couplings[1], self.couplings[2]
        L = _lib.lib()

        def photon_pass():
            if pre_push:
                photons.push(cfg.dt_s, self.streams[0].cuda_stream)
            apply.apply_photons(self.tables, cg, photons, cfg, out_g, 2, self.streams[0].cuda_stream, False)
            if push_photons:
                photons.push(cfg.dt_s, self.streams[0].cuda_stream)

        def positron_pass():
            apply.apply_positrons(self.tables, cp, positrons, cfg, out_p, None, 2, self.streams[1].cuda_stream, False)

        passes = (("photon", photon_pass, self.streams[0]), ("positron", positron_pass, self.streams[1]))
        for name, run, _ in passes:
            if name in self.early:
                run()
        apply.apply_electrons(self.tables, self.couplings[0], electrons, cfg, out_e, order_secondaries=2,
                              stream=main.cuda_stream, check=False)
        for name, run, s in passes:
            if name not in self.early:
                _lib.check(L.rrea_stream_wait_fast_path(self.couplings[0].h, s.cuda_stream),
                           "rrea_stream_wait_fast_path")
                run()
        for s in self.streams:
            done = torch.cuda.Event()
            done.record(s)
            main.wait_event(done)
        if drain is not None and device_drain > 0:
            # the drain with its counts on the device, up to device_drain generations
            outs = (out_e, out_g, out_p)
            recs = (C.c_void_p * 3)(*[o.secondaries.data_ptr() for o in outs])
            cnts = (C.c_void_p * 3)(*[o.secondary_count.data_ptr() for o in outs])
            caps = (C.c_int64 * 3)(*[o.secondary_capacity for o in outs])
            io = drain.struct(0)
            c = _step_config(cfg)
            _lib.check(_lib.lib().rrea_newborn_drain_staged(self.tables.h, self.couplings[0].h, C.byref(c), recs,
                                                            cnts, caps, 3, C.byref(io), int(device_drain),
                                                            main.cuda_stream), "rrea_newborn_drain_staged")
        elif drain is not None:
            apply.drain_secondaries(self.tables, self.couplings[0], cfg, [out_e, out_g, out_p], drain,
                                    stream=main.cuda_stream, check=False)
        if check:
            out_e.check_error()

    def set_block(self, threads: int) -> None:
        """The fast kernels' workgroup size for all three passes."""
        for c in self.couplings:
            c.set_block(threads)

    def close(self):
        pass


GROUP_RESULT_BITS = {"required": 1, "present": 2, "finite": 4, "valid": 8}


def validate_secondary_groups(groups, low_energy_cutoff_eV: float, stream: int | None = None):
    """RreaValidateSecondaryGroupConservation over an array of group records
    (OV/RreaParticleInteraction.H:1099-1252), evaluated on the device.

    ``groups`` is a ctypes array of ``_lib.SecondaryGroup``; returns the
    result words (bit 1 required, 2 present, 4 finite, 8 valid) and the
    (charge, energy) residuals per group, both as numpy arrays."""
    n = len(groups)
    buf = torch.empty(max(n, 1) * C.sizeof(_lib.SecondaryGroup), dtype=torch.uint8)
    if n:
        C.memmove(buf.data_ptr(), C.addressof(groups), n * C.sizeof(_lib.SecondaryGroup))
    dev = buf.cuda()
    result = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    residuals = torch.zeros(2 * max(n, 1), dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream if stream is None else stream
    _lib.check(_lib.lib().rrea_validate_secondary_groups(dev.data_ptr(), n, low_energy_cutoff_eV,
                                                         result.data_ptr(), residuals.data_ptr(), s),
               "rrea_validate_secondary_groups")
    torch.cuda.synchronize()
    return result[:n].cpu().numpy(), residuals[:2 * n].cpu().numpy().reshape(n, 2)
