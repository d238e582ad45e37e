"""Per-grid tiling of the three-sweep Hpsi: a committed table keyed by grid
length and band class (tuning.json, "hpsi"), and the autotuner that fills it.

    knobs = lookup(n, max_bands)                 # None where the table has no entry
    best, times, default = autotune_hpsi(L, nst) # times every knob combination

The knobs are inqamd_tuning's (include/inq_amd.h): z-tile shape (8 y x 8
bands or 16 y x 4 bands), x-y plane threads (1 or 2 x plane_threads(L)),
bands per chunk of the sweeps and bands per projector workgroup.  The
z-tile shape, chunk and projector bands change no result bit (the sweeps'
arithmetic per column is the same for every shape, and the projector's point
chunking follows from the band count alone); the x-y thread count selects
another instantiation of the plane kernel, whose floating-point contraction
may differ in the last bits.  The table is a speed choice.
"""
import itertools
import json
import os

import torch

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

# band classes of the table: (lowest, highest) max_bands
BAND_CLASSES = ((1, 32), (33, 128), (129, 1 << 30))
KNOBS = ("z_tile_y", "xy_threads", "band_chunk", "projector_bands")


def band_class(nst):
    for lo, hi in BAND_CLASSES:
        if lo <= nst <= hi:
            return f"{lo}-{hi}" if hi < (1 << 30) else f"{lo}+"
    raise ValueError(f"no band class for {nst} bands")


def load(path=_PATH):
    with open(path) as f:
        return json.load(f).get("hpsi", {})


def lookup(n, nst, path=_PATH):
    """The committed knobs for a cubic grid of length n[0] and nst bands
    (max_bands), or None (non-cubic grid, or no entry)."""
    n = tuple(int(x) for x in n)
    if len(set(n)) != 1:
        return None
    entry = load(path).get(str(n[0]), {}).get(band_class(int(nst)))
    if not entry:
        return None
    knobs = {k: int(entry[k]) for k in KNOBS if k in entry}
    # a band chunk wider than the handle's bands is meaningless: keep the default
    if knobs.get("band_chunk", 0) > nst:
        knobs["band_chunk"] = 0
    return knobs


def chunk_candidates(L, nst):
    """Band chunks worth timing: intermediates of 32-384 MiB per chunk (the
    sweeps' intermediate of one chunk is 16 L^3 bytes per band) and 16, 32
    or 64 bands, multiples of 8, at most half the bands (two chunks are in
    flight), and 0 (the library's choice)."""
    per_band = 16.0 * L ** 3
    out = {0}
    sizes = [max(8, int(mib * (1 << 20) / per_band) // 8 * 8)
             for mib in (32, 48, 64, 96, 128, 192, 384)]
    for b in sizes + [16, 32, 64]:
        if 2 * b <= nst:
            out.add(b)
    return sorted(out)


def _time(fn, reps):
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    fn()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tuning_system(L, seed=0):
    """The system an entry is tuned on: a cubic cell of L points at the
    config cells' spacing with the diamond cell's atom density (8 atoms per
    32^3), at seeded uniform positions."""
    from . import systems
    natoms = max(2, int(round(8 * (L / 32.0) ** 3)))
    return systems.uneven_batch_system(900 + L, L, natoms, seed=seed)


def autotune_hpsi(L, nst, reps=10, z_tile_y=(8, 16), xy_threads=(1, 2), projector_bands=(16, 32, 64),
                  band_chunk=None, system=None):
    """Times Hpsi of nst bands on an L^3 system for every knob combination;
    returns (best knobs, {knob tuple: ms}, ms with the library defaults)."""
    from .hamiltonian import ks_hamiltonian
    s = system if system is not None else tuning_system(L)
    ham = ks_hamiltonian(s, max_bands=nst, tuning=None)
    dev = torch.device("cuda", torch.cuda.current_device())
    g = torch.Generator(device=dev).manual_seed(L * 1000 + nst)
    psi = torch.randn(tuple(s.n) + (nst,), dtype=torch.complex128, device=dev, generator=g)
    out = torch.empty_like(psi)
    chunks = band_chunk if band_chunk is not None else chunk_candidates(L, nst)
    times = {}
    for combo in itertools.product(z_tile_y, xy_threads, chunks, projector_bands):
        ham.tuning = dict(zip(KNOBS, combo))
        times[combo] = _time(lambda: ham(psi, out=out), reps)
    best = min(times, key=times.get)
    # the winner against the library's defaults, alternately, three times each:
    # a winner less than 1 % faster keeps the defaults (all knobs 0)
    zero = (0, 0, 0, 0)
    tb, td = [], []
    for _ in range(3):
        ham.tuning = dict(zip(KNOBS, best))
        tb.append(_time(lambda: ham(psi, out=out), reps))
        ham.tuning = dict(zip(KNOBS, zero))
        td.append(_time(lambda: ham(psi, out=out), reps))
    default = min(td)
    times[best] = min(tb)
    if not times[best] < 0.99 * default:
        times[zero] = default
        best = zero
    ham.close()
    return dict(zip(KNOBS, best)), times, default


def table_entry(best, times, default):
    """A tuning.json entry: the knobs and what they measured against the
    library defaults."""
    e = dict(best)
    t = times[tuple(best[k] for k in KNOBS)]
    e["measured"] = f"{t:.4f} ms vs {default:.4f} ms with the defaults ({len(times)} combinations)"
    return e


def save(table, path=_PATH):
    """Writes the "hpsi" section of tuning.json, keeping the other sections."""
    with open(path) as f:
        doc = json.load(f)
    doc["hpsi"] = table
    with open(path, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
        f.write("\n")


def representative_bands(cls):
    lo, hi = cls
    return {(1, 32): 16, (33, 128): 64}.get(cls, 256)


def grid_table(lengths, reps=10, log=print):
    """Autotunes every (grid length, band class) and returns the table."""
    table = {}
    for L in lengths:
        for cls in BAND_CLASSES:
            nst = representative_bands(cls)
            if L >= 96 and nst == 256:
                nst = 512
            best, times, default = autotune_hpsi(L, nst, reps=reps)
            table.setdefault(str(L), {})[band_class(nst)] = table_entry(best, times, default)
            log(f"L={L} nst={nst}: {table[str(L)][band_class(nst)]}")
    return table


def check_table_consistency(table):
    """Every entry names valid knob values (used by the CPU tests)."""
    for L, classes in table.items():
        int(L)
        for cls, e in classes.items():
            assert e["z_tile_y"] in (0, 8, 16)
            assert e["xy_threads"] in (0, 1, 2)
            assert e["band_chunk"] >= 0 and e["band_chunk"] % 8 == 0
            assert e["projector_bands"] in (0, 16, 32, 64)
            assert cls in {band_class(lo) for lo, _ in BAND_CLASSES}
    return True


__all__ = ["lookup", "autotune_hpsi", "grid_table", "save", "load", "band_class",
           "chunk_candidates", "tuning_system", "check_table_consistency"]
