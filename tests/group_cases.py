"""Interaction groups of the reference's secondary-group smoke test
(native_amrex_rrea/src/RreaSecondaryGroupSmoke.cpp:282-523), as plain
records, and a builder into either C layout (the oracle's orc_group or the
library's rrea_secondary_group, which are the same)."""
import numpy as np

ME = 510998.95  # electron rest energy, eV (RreaSpeciesRestEnergyEv)
CUT = 1000.0  # low_cut_eV of the smoke test
EPS = float(np.finfo(float).eps)

# RreaSecondaryProcess (OV/RreaParticleInteraction.H:34-48)
CONTINUOUS, SEED, MOLLER, BHABHA, BREMS, COMPTON, PHOTOELECTRIC, PAIR_NUCLEAR, PAIR_TRIPLET, ANNIHILATION, \
    PHOTON_ESCAPE, PARTICLE_ESCAPE, DEMOTION = range(13)
E, P, G = 0, 1, 2  # species
ION, POS_ION, LOW_E = 0, 1, 2  # material source kinds
MISSING, DETAILED, EXPLICIT_ZERO = 0, 1, 2


def group(process, parent=E, remove=False, still_valid=False, mode=DETAILED, w=0.0, before=0.0, after=0.0,
          local=0.0, escaped=0.0, esc_q=0.0, particles=(), sources=(), survives=None):
    """A group record; ``local`` and ``escaped`` are per unit weight, as the
    smoke test's ``detailed`` helper takes them (:313-328)."""
    return dict(process=process, parent_species=parent, remove_parent=int(remove), parent_still_valid=int(still_valid),
                closure_mode=mode, parent_survives=int(not remove if survives is None else survives),
                parent_weight=w, parent_energy_before_eV=before, parent_energy_after_eV=after,
                local_energy_eV=w * local, escaped_energy_eV=w * escaped, escaped_charge_e=esc_q,
                particles=list(particles), sources=list(sources))


W = 2.0
VALID = {
    "moller": group(MOLLER, E, w=W, before=2.0e6, after=1.2e6, particles=[(E, W, 0.8e6)], sources=[(ION, W, 0.8e6)]),
    "bhabha": group(BHABHA, P, w=W, before=2.0e6, after=1.2e6, particles=[(E, W, 0.8e6)], sources=[(POS_ION, W, 0.0)]),
    "electron_brems": group(BREMS, E, w=W, before=2.0e6, after=1.4e6, particles=[(G, W, 0.6e6)]),
    "positron_brems": group(BREMS, P, w=W, before=2.0e6, after=1.4e6, particles=[(G, W, 0.6e6)]),
    "compton": group(COMPTON, G, w=W, before=1.0e6, after=0.7e6, particles=[(E, W, 0.3e6)],
                     sources=[(ION, W, 0.3e6)]),
    "photoelectric": group(PHOTOELECTRIC, G, remove=True, w=W, before=100.0e3, after=0.0, local=10.0e3,
                           particles=[(E, W, 90.0e3)], sources=[(ION, W, 90.0e3)]),
    "nuclear_pair": group(PAIR_NUCLEAR, G, remove=True, w=W, before=2.0 * ME + 300.0e3, after=0.0, local=50.0e3,
                          particles=[(E, W, 100.0e3), (P, W, 150.0e3)]),
    "triplet": group(PAIR_TRIPLET, G, remove=True, w=W, before=2.0 * ME + 600.0e3, after=0.0, local=50.0e3,
                     particles=[(E, W, 100.0e3), (P, W, 200.0e3), (E, W, 250.0e3)], sources=[(ION, W, 250.0e3)]),
    "annihilation": group(ANNIHILATION, P, remove=True, w=W, before=200.0e3, after=0.0,
                          particles=[(G, W, ME + 100.0e3), (G, W, ME + 100.0e3)], sources=[(POS_ION, W, 0.0)]),
    "demotion": group(DEMOTION, E, remove=True, w=W, before=500.0, after=0.0, local=500.0,
                      sources=[(LOW_E, W, 0.0)]),
    "zero_weight_cleanup": group(DEMOTION, remove=True, mode=EXPLICIT_ZERO),
    "electron_escape": group(PARTICLE_ESCAPE, E, remove=True, still_valid=True, w=W, before=750.0e3, after=0.0,
                             escaped=750.0e3 + ME, esc_q=-W),
    "positron_escape": group(PARTICLE_ESCAPE, P, remove=True, still_valid=True, w=W, before=750.0e3, after=0.0,
                             escaped=750.0e3 + ME, esc_q=W),
}


def _with(name, **kw):
    g = dict(VALID[name])
    g.update(kw)
    return g


# (group, expected result bits) -- :480-523
INVALID = {
    "zero_escape_bypass": (_with("electron_escape", closure_mode=EXPLICIT_ZERO, parent_weight=0.0,
                                 parent_energy_before_eV=0.0, escaped_energy_eV=0.0, escaped_charge_e=0.0,
                                 parent_survives=0), 1 | 2),
    "bad_escaped_charge": (_with("electron_escape", escaped_charge_e=-W + 1.0), 1 | 2 | 4),
    "bad_escaped_energy": (_with("positron_escape", escaped_energy_eV=W * (750.0e3 + ME) + 1.0), 1 | 2 | 4),
    "bad_residual": (_with("compton", local_energy_eV=1.0), 1 | 2 | 4),
    "missing": (_with("compton", closure_mode=MISSING), 1),
    "nonfinite": (_with("compton", parent_energy_before_eV=float("nan")), 1 | 2),
}


def build(cls, entry_cls, g):
    """One record of the C layout ``cls`` (entries of ``entry_cls``)."""
    out = cls()
    for k in ("process", "parent_species", "remove_parent", "parent_still_valid", "closure_mode", "parent_survives",
              "parent_weight", "parent_energy_before_eV", "parent_energy_after_eV", "local_energy_eV",
              "escaped_energy_eV", "escaped_charge_e"):
        setattr(out, k, g[k])
    out.n_particles = len(g["particles"])
    out.n_sources = len(g["sources"])
    for i, (k, w, e) in enumerate(g["particles"]):
        out.particles[i] = entry_cls(k, 0, w, e)
    for i, (k, w, e) in enumerate(g["sources"]):
        out.sources[i] = entry_cls(k, 0, w, e)
    return out


def input_energy(g):
    """The closure's input side: parent total energy plus the target
    electron's rest energy per ionization or positive-ion source."""
    m = 0.0 if g["parent_species"] == G else ME
    e = g["parent_weight"] * (g["parent_energy_before_eV"] + m)
    return e + sum(w * ME for k, w, _ in g["sources"] if k in (ION, POS_ION))


def random_groups(rng, n):
    """Groups of every detailed kind with random weights and energies, closed
    in exact arithmetic, each with the number of ulps of its input energy
    planted into its local term: 0 for about two thirds, 1 to 40 (inside the
    64-ulp tolerance with room for the sums' rounding) or 100 to 400
    (outside it) for the rest.  Returns (group, ulps) pairs."""
    out = []
    for _ in range(n):
        w = float(rng.uniform(0.1, 1e4))
        kind = int(rng.integers(0, 9))
        before = float(10.0 ** rng.uniform(3.5, 8))
        t = float(before * rng.uniform(0.0, 1.0))
        if kind == 0:
            g = group(MOLLER, E, w=w, before=before, after=before - t, particles=[(E, w, t)],
                      sources=[(ION, w, t)])
        elif kind == 1:
            g = group(BHABHA, P, w=w, before=before, after=before - t, particles=[(E, w, t)],
                      sources=[(POS_ION, w, 0.0)])
        elif kind == 2:
            g = group(BREMS, int(rng.integers(0, 2)), w=w, before=before, after=before - t, particles=[(G, w, t)])
        elif kind == 3:
            g = group(COMPTON, G, w=w, before=before, after=before - t, particles=[(E, w, t)],
                      sources=[(ION, w, t)])
        elif kind == 4:
            b = float(rng.uniform(0.0, 1.0)) * t
            g = group(PHOTOELECTRIC, G, remove=True, w=w, before=t, local=b, particles=[(E, w, t - b)],
                      sources=[(ION, w, t - b)])
        elif kind == 5:
            e1 = float(rng.uniform(0, 1)) * t
            g = group(PAIR_NUCLEAR, G, remove=True, w=w, before=2.0 * ME + t,
                      particles=[(E, w, e1), (P, w, t - e1)])
        elif kind == 6:
            g = group(ANNIHILATION, P, remove=True, w=w, before=t,
                      particles=[(G, w, ME + 0.5 * t), (G, w, ME + 0.5 * t)], sources=[(POS_ION, w, 0.0)])
        elif kind == 7:
            g = group(DEMOTION, E, remove=True, w=w, before=t, local=t, sources=[(LOW_E, w, 0.0)])
        else:
            s = int(rng.integers(0, 2))
            g = group(PARTICLE_ESCAPE, s, remove=True, still_valid=True, w=w, before=t, escaped=t + ME,
                      esc_q=-w if s == 0 else w)
        for k, sw, se in g["sources"]:
            if k == ION and se < CUT and (E, sw, se) in g["particles"]:
                # a sub-cutoff recoil is deposited; its source emits the electron into the fluid
                g["particles"].remove((E, sw, se))
                g["local_energy_eV"] += sw * se
        ulps = 0
        u = rng.uniform()
        if u < 0.17:
            ulps = int(rng.integers(1, 41))
        elif u < 0.34:
            ulps = int(rng.integers(100, 401))
        g["local_energy_eV"] += ulps * EPS * input_energy(g)
        out.append((g, ulps))
    return out
