"""The autotuner's cache (picsar_amd.autotune): signatures tell workloads
apart, entries round-trip through the JSON table, and a stored result is
applied without trial steps.  Host-side only: the library's tuning keys are
set and read, no kernel runs."""
import json
import types

import pytest


def fake_sim(nx=64, ppc=(8.0, 8.0), precision="mixed", order=3):
    g = types.SimpleNamespace(ntx=8, nty=8, ntz=8)
    cells = nx ** 3
    species = [types.SimpleNamespace(npart=int(p * cells)) for p in ppc]
    return types.SimpleNamespace(grid=g, nx=nx, ny=nx, nz=nx, order=order, l_lower_order_in_v=True,
                                 pusher=1, depo=0, precision=precision, species=species)


def test_signature_tells_workloads_apart():
    from picsar_amd.autotune import signature
    a = signature(fake_sim(), "dev")
    assert a == signature(fake_sim(), "dev")
    assert a != signature(fake_sim(nx=128), "dev")
    assert a != signature(fake_sim(ppc=(16.0, 8.0)), "dev")
    assert a != signature(fake_sim(precision="fp64"), "dev")
    assert a != signature(fake_sim(order=2), "dev")
    assert a != signature(fake_sim(), "other")


def test_cache_round_trip(tmp_path):
    from picsar_amd.autotune import load_cache, store_cache
    path = str(tmp_path / "t" / "autotune.json")
    assert load_cache(path) == {}
    store_cache("a", {"best": {"fused.chunk": 2048}, "ms_per_step": 1.5}, path)
    store_cache("b", {"best": {"fused.chunk": 4096}, "ms_per_step": 2.5}, path)
    store_cache("a", {"best": {"fused.chunk": 8192}, "ms_per_step": 1.0}, path)
    data = json.load(open(path))
    assert data == load_cache(path)
    assert data["a"]["best"] == {"fused.chunk": 8192} and data["b"]["ms_per_step"] == 2.5
    (tmp_path / "bad.json").write_text("not json")
    assert load_cache(str(tmp_path / "bad.json")) == {}


def test_cached_result_is_applied_without_trials(tmp_path, monkeypatch):
    from picsar_amd import autotune as at
    try:
        from picsar_amd import _lib
        _lib.load()
    except Exception as e:  # the library must be built for the tuning keys
        pytest.skip(f"libpxrhip.so not loadable: {e}")
    path = str(tmp_path / "autotune.json")
    sim = fake_sim()
    sig = at.signature(sim, at.device_name())
    at.store_cache(sig, {"best": {"fused.chunk": 2048, "fused.wg_per_cu": 3,
                                  "fused.lds_bytes": 52224, "fused.spread": 1},
                         "ms_per_step": 3.0}, path)
    monkeypatch.setattr(at, "_time_stage", lambda *a, **k: pytest.fail("trial step ran"))
    before = at.get_tuning()
    try:
        res = at.autotune(sim, cache_file=path)
        assert res["cached"] and res["signature"] == sig and res["trials"] == []
        got = at.get_tuning()
        assert got["fused.chunk"] == 2048
        # mixed precision: the launch keys go to the mixed.* names
        assert got["mixed.wg_per_cu"] == 3 and got["mixed.lds_bytes"] == 52224
    finally:
        at.set_tuning(before)


def test_trials_start_from_the_entry_tuning(monkeypatch):
    """Each candidate is applied on top of the tuning found on entry (a key
    set by one candidate does not carry over to the next), the fastest is
    applied the same way, and a mixed-precision run sets the lane stride
    under its mixed.* name.  The trial timer is replaced: no kernel runs."""
    from picsar_amd import autotune as at
    try:
        from picsar_amd import _lib
        _lib.load()
    except Exception as e:
        pytest.skip(f"libpxrhip.so not loadable: {e}")
    sim = fake_sim(precision="mixed")
    sim.check = lambda: None
    before = at.get_tuning()
    seen = []

    def fake_time(sim_, steps):
        t = at.get_tuning()
        seen.append(t)
        return 1.0 if t["fused.chunk"] == 2048 else 2.0

    monkeypatch.setattr(at, "_time_stage", fake_time)
    cands = [{"fused.order": 1 - before["fused.order"]}, {"fused.chunk": 2048},
             {"fused.sstride": 8}]
    try:
        res = at.autotune(sim, candidates=cands, cache=False)
        # two timings per candidate (warm-up and trial)
        assert len(seen) == 6
        assert seen[2]["fused.order"] == before["fused.order"]  # no carry-over
        assert seen[4]["mixed.sstride"] == 8 and seen[4]["fused.sstride"] == before["fused.sstride"]
        assert res["best"] == {"fused.chunk": 2048}
        got = at.get_tuning()
        assert got["fused.chunk"] == 2048
        assert got["fused.order"] == before["fused.order"]
        assert got["mixed.sstride"] == before["mixed.sstride"]
    finally:
        at.set_tuning(before)
