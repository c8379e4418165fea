"""Launch-parameter autotuner of the tiled step.

The reference's tile split (ntilex, ntiley, ntilez) is the user's and is
kept; on the GPU the unit of work is a work item: a run of cell-sorted
particles of one tile whose field and current blocks sit in LDS.  Its shape
and the launch of the persistent kernel are set by the keys of
``pxrhip_set_tuning``:

* ``fused.chunk``     -- particles per work item (rows of the sorting grid
  are added while the LDS blocks fit);
* ``fused.wg_per_cu`` with ``fused.lds_bytes`` -- workgroups per CU and the
  LDS each may use (2 x 79616, 1 x 155648, 3 x 52224 bytes; 0 workgroups:
  the library's automatic choice between the first and the last);
* ``fused.spread``    -- lanes on particles S apart in the deposition;
* ``fused.sstride``   -- the largest such S (``mixed.sstride`` for a
  mixed-precision species; -1: the library's choice by shape order);
* ``fused.order``     -- the queue hands out the largest work items first.

Every candidate is applied on top of the tuning found on entry, so a key one
candidate sets does not carry over to the next.

``autotune`` times the gather + push + deposit stage of a Simulation for
each candidate with the library's stream timers and keeps the fastest.  The
trial steps are real steps of the simulation: launch parameters change the
order of floating-point atomic additions only, not the algorithm.

Results are cached in a JSON table keyed by the workload's signature (grid,
tiles, method, precision, particles per cell and the device), so a later
run of the same workload applies the stored choice without trials.  The
table lives at ``$PICSAR_AMD_TUNE_CACHE`` or ``~/.cache/picsar_amd/
autotune.json``.
"""
from __future__ import annotations

import ctypes as C
import itertools
import json
import os
import tempfile
from typing import Dict, Iterable, List, Optional

from . import _lib

KEYS = ("fused.chunk", "fused.wg_per_cu", "fused.lds_bytes", "fused.spread",
        "mixed.wg_per_cu", "mixed.lds_bytes", "fused.order", "fused.sstride", "mixed.sstride")
TIMER_BASE = 65000  # timer slots used here (the library has 65536)

DEFAULT_CANDIDATES: List[Dict[str, int]] = [
    dict(zip(("fused.chunk", "fused.wg_per_cu", "fused.lds_bytes", "fused.spread"),
             (ch, wg, lds, sp)))
    for ch, (wg, lds), sp in itertools.product((1024, 2048, 4096, 8192),
                                               ((2, 79616), (1, 155648), (3, 52224)), (1,))
] + [{"fused.chunk": 4096, "fused.wg_per_cu": 2, "fused.lds_bytes": 79616, "fused.spread": 0},
      {"fused.chunk": 4096, "fused.wg_per_cu": 0, "fused.lds_bytes": 52224, "fused.spread": 1},
      {"fused.chunk": 4096, "fused.wg_per_cu": 0, "fused.lds_bytes": 52224, "fused.spread": 1,
       "fused.order": 1}] + [
    {"fused.chunk": 4096, "fused.wg_per_cu": 0, "fused.lds_bytes": 52224, "fused.spread": 1,
     "fused.sstride": ss} for ss in (0, 4, 8, 16, 32)]


def get_tuning(keys: Iterable[str] = KEYS) -> Dict[str, int]:
    out = {}
    for k in keys:
        v = C.c_int64(0)
        _lib.call("pxrhip_get_tuning", k.encode(), C.byref(v))
        out[k] = v.value
    return out


def set_tuning(values: Dict[str, int]) -> None:
    for k, v in values.items():
        _lib.call("pxrhip_set_tuning", k.encode(), int(v))


def _time_stage(sim, steps: int) -> float:
    """Milliseconds per step of gather + push + deposit (all species), over
    ``steps`` whole steps."""
    total = 0.0
    for s in range(steps):
        a, b = TIMER_BASE + 2 * s, TIMER_BASE + 2 * s + 1
        _lib.call("pxrhip_timer_record", a)
        sim.gather_push_deposit()
        _lib.call("pxrhip_timer_record", b)
        sim.current_bcs()
        sim.field_step()
        sim.particle_bcs_and_sorting()
        sim.it += 1
    for s in range(steps):
        ms = C.c_double(0.0)
        _lib.call("pxrhip_timer_elapsed", TIMER_BASE + 2 * s, TIMER_BASE + 2 * s + 1, C.byref(ms))
        total += ms.value
    return total / steps


def _for_precision(c: Dict[str, int], precision: str) -> Dict[str, int]:
    """Mixed-precision species are launched with the mixed.* keys."""
    if precision != "mixed":
        return c
    ren = {"fused.wg_per_cu": "mixed.wg_per_cu", "fused.lds_bytes": "mixed.lds_bytes",
           "fused.sstride": "mixed.sstride"}
    return {ren.get(k, k): v for k, v in c.items()}


def cache_path() -> str:
    return os.environ.get("PICSAR_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "picsar_amd",
                                       "autotune.json"))


def signature(sim, device: str = "") -> str:
    """The workload as the tuner sees it: grid and tiles, method, precision,
    particles per cell of each species (two significant digits) and the
    device.  Equal signatures share a tuning."""
    g = sim.grid
    cells = max(1, sim.nx * sim.ny * sim.nz)
    ppc = [float(f"{sp.npart / cells:.2g}") for sp in sim.species]
    parts = [f"n={sim.nx}x{sim.ny}x{sim.nz}", f"tiles={g.ntx}x{g.nty}x{g.ntz}",
             f"order={sim.order}", f"lo={int(sim.l_lower_order_in_v)}", f"pusher={sim.pusher}",
             f"depo={sim.depo}", f"precision={getattr(sim, 'precision', 'fp64')}",
             f"ppc={ppc}", f"device={device}"]
    return ";".join(parts)


def device_name() -> str:
    """gfx name and CU count of the current device (empty without one)."""
    try:
        v = C.c_int64(0)
        _lib.call("pxrhip_get_tuning", b"device.num_cus", C.byref(v))
        return f"gfx950/{v.value}cu"
    except Exception:
        return ""


def load_cache(path: Optional[str] = None) -> Dict[str, dict]:
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def store_cache(sig: str, entry: dict, path: Optional[str] = None) -> None:
    """Add or replace one entry; the file is replaced atomically."""
    path = path or cache_path()
    data = load_cache(path)
    data[sig] = entry
    d = os.path.dirname(path) or "."
    os.makedirs(d, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=d, prefix=".autotune.", suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def autotune(sim, candidates: Optional[List[Dict[str, int]]] = None, steps: int = 2,
             verbose: bool = False, cache: bool = True,
             cache_file: Optional[str] = None) -> Dict[str, object]:
    """Try every candidate on ``sim`` for ``steps`` steps (after one step to
    warm the candidate up), apply the fastest and return it with the
    timings.  Raises if any trial step skipped a particle.  With ``cache``,
    a stored result for the same signature is applied without trials, and a
    new result is stored (only for the default candidate list)."""
    sig = signature(sim, device_name())
    if cache and candidates is None:
        hit = load_cache(cache_file).get(sig)
        if hit and isinstance(hit.get("best"), dict):
            best = _for_precision(hit["best"], getattr(sim, "precision", "fp64"))
            set_tuning(best)
            return {"best": best, "ms_per_step": hit.get("ms_per_step"), "cached": True,
                    "signature": sig, "trials": []}
    cands = DEFAULT_CANDIDATES if candidates is None else candidates
    cands = [_for_precision(c, getattr(sim, "precision", "fp64")) for c in cands]
    before = get_tuning()
    results = []
    try:
        for c in cands:
            set_tuning(before)
            set_tuning(c)
            _time_stage(sim, 1)
            ms = _time_stage(sim, steps)
            results.append((ms, c))
            if verbose:
                print(f"autotune: {c} -> {ms:.3f} ms/step")
        sim.check()
    except Exception:
        set_tuning(before)
        raise
    best_ms, best = min(results, key=lambda r: r[0])
    set_tuning(before)
    set_tuning(best)
    out = {"best": best, "ms_per_step": best_ms, "cached": False, "signature": sig,
           "trials": [{"tuning": c, "ms_per_step": ms} for ms, c in results]}
    if cache and candidates is None:
        store_cache(sig, {"best": best, "ms_per_step": best_ms}, cache_file)
    return out
