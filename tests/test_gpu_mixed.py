"""Mixed precision on the GPU against the FP64 oracle.

A mixed-precision species keeps FP64 positions and FP32 momenta, gamma and
weights; the gather, the push and the per-particle current are FP32; the
LDS current block accumulates in 64-bit fixed point (default), FP64,
compensated FP32 or plain FP32; the global current is FP64.  The bar
of the method is 1e-5 relative error against the FP64 oracle on J and on
the momenta after one step; several steps are compared at a looser bar.
The discrete continuity of the mixed Esirkepov current is checked with
FP64 charge densities at the old and new positions.
"""
import ctypes as C

import numpy as np
import pytest

from tests.test_gpu_step import build, rel

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5


@pytest.fixture(scope="module")
def torch():
    from picsar_amd import _lib
    _lib.call("pxrhip_init", 0)  # raises unless a gfx950 device is present
    return None


def set_tuning(key, value):
    from picsar_amd import _lib
    old = C.c_int64(0)
    _lib.call("pxrhip_get_tuning", key.encode(), C.byref(old))
    _lib.call("pxrhip_set_tuning", key.encode(), value)
    return old.value


def mixed_build(torch, orc, order, lo, pusher, depo, **kw):
    return build(torch, orc, order, lo, pusher, depo, precision="mixed", id_scale=2.0 ** -15,
                 **kw)


def errors(sim, ref):
    out = {}
    for k in ("jx", "jy", "jz", "ex", "ey", "ez", "bx", "by", "bz"):
        out[k] = rel(sim.fields[k].numpy(), ref.f[k])
    for sp, rs in zip(sim.species, ref.species):
        a = sp.arrays()
        ia, ib = np.argsort(a["w"]), np.argsort(rs["w"].astype(np.float32))
        for k in ("ux", "uy", "uz"):
            out[f"{sp.name}.{k}"] = rel(a[k][ia].astype(np.float64), rs[k][ib])
        for k in ("x", "y", "z"):
            out[f"{sp.name}.{k}.abs"] = np.abs(a[k][ia] - rs[k][ib]).max()
    return out


def unique_float32_weights(ref):
    # weights carry particle ids; they must stay distinct in float32
    for rs in ref.species:
        assert len(np.unique(rs["w"].astype(np.float32))) == len(rs["w"])


@pytest.mark.parametrize("order,lo,pusher,depo", [
    (1, True, "boris", "direct"),
    (2, True, "vay", "esirkepov"),
    (3, True, "vay", "esirkepov"),
    (3, False, "boris", "esirkepov"),
    (3, True, "vay", "direct"),
])
@pytest.mark.parametrize("jacc", [3, 2, 1, 0])
def test_one_step_against_fp64_oracle(torch, orc, order, lo, pusher, depo, jacc):
    """jacc: the LDS current in fixed point (3, the default), FP64 (2),
    compensated FP32 (1), FP32 (0)."""
    old = set_tuning("mixed.jacc", jacc)
    try:
        sim, ref, _ = mixed_build(torch, orc, order, lo, pusher, depo, ppc=2)
        unique_float32_weights(ref)
        # the oracle runs on the same (FP32-rounded) momenta and weights
        for sp, rs in zip(sim.species, ref.species):
            for k in ("ux", "uy", "uz", "w"):
                rs[k][...] = rs[k].astype(np.float32).astype(np.float64)
            rs["gaminv"][...] = 1.0 / np.sqrt(1 + rs["ux"] ** 2 + rs["uy"] ** 2 + rs["uz"] ** 2)
        sim.step(1)
        ref.step()
    finally:
        set_tuning("mixed.jacc", old)
    err = errors(sim, ref)
    for k in ("jx", "jy", "jz", "ex", "ey", "ez", "bx", "by", "bz"):
        assert err[k] <= MIXED_TOL, (k, err[k])
    for k, v in err.items():
        if k.endswith(("ux", "uy", "uz")):
            assert v <= MIXED_TOL, (k, v)
        if k.endswith(".abs"):
            assert v <= 1e-7, (k, v)


@pytest.mark.parametrize("order,depo", [(1, "direct"), (3, "esirkepov")])
def test_relativistic_step_against_fp64_oracle(torch, orc, order, depo):
    """Electrons at u ~ 3 (v up to 0.99 c) cross cells on most steps: the
    FP64 path of the shape differences and the wide Esirkepov box carry the
    mixed-precision step, which stays within the same bounds on J, E, B and
    the momenta.  Positions move by up to 0.55 cells per step here, and the
    FP32 gamma of the push puts a few FP32 ulps of that displacement into
    them: 4e-7 cells (1e-7 for the slow plasma of the other tests)."""
    sim, ref, _ = mixed_build(torch, orc, order, True, "vay", depo, ppc=2, uth=(3.0, 0.3),
                              seed=11)
    unique_float32_weights(ref)
    for sp, rs in zip(sim.species, ref.species):
        for k in ("ux", "uy", "uz", "w"):
            rs[k][...] = rs[k].astype(np.float32).astype(np.float64)
        rs["gaminv"][...] = 1.0 / np.sqrt(1 + rs["ux"] ** 2 + rs["uy"] ** 2 + rs["uz"] ** 2)
    sim.step(1)
    ref.step()
    err = errors(sim, ref)
    for k in ("jx", "jy", "jz", "ex", "ey", "ez", "bx", "by", "bz"):
        assert err[k] <= MIXED_TOL, (k, err[k])
    for k, v in err.items():
        if k.endswith(("ux", "uy", "uz")):
            assert v <= MIXED_TOL, (k, v)
        if k.endswith(".abs"):
            assert v <= 4e-7, (k, v)


@pytest.mark.parametrize("order,depo", [(1, "direct"), (2, "esirkepov"), (3, "esirkepov")])
def test_cut_current_blocks_against_fp64_oracle(torch, orc, order, depo):
    """Work items built with no motion room around the current block of
    their particles' old positions (mixed.jmargin 0) under a
    16 KB cap on the LDS current block, and two gather+push+deposit passes
    without re-binning, so that in the second the particles (u_th = 0.6 c)
    have left the sorting cells their work items were built from: the
    current boxes that no longer fit are deposited by k_fused_cut through a
    block cut down around the old positions, and the particles whose box
    leaves it through global memory.  J is the oracle's at 1e-5 (the oracle
    is given the FP32-rounded momenta and weights)."""
    from picsar_amd import _lib
    old_m = set_tuning("mixed.jmargin", 0)
    old_j = set_tuning("fused.j_lds_bytes", 16384)
    cut = 0
    try:
        sim, ref, _ = mixed_build(torch, orc, order, True, "vay", depo, ppc=2, uth=(0.6, 0.6),
                                  seed=11)
        unique_float32_weights(ref)
        for sp, rs in zip(sim.species, ref.species):
            for k in ("ux", "uy", "uz", "w"):
                rs[k][...] = rs[k].astype(np.float32).astype(np.float64)
            rs["gaminv"][...] = 1.0 / np.sqrt(1 + rs["ux"] ** 2 + rs["uy"] ** 2 + rs["uz"] ** 2)
        sim.gather_push_deposit()
        for sp in sim.species:
            sim.gather_push_deposit_species(sp)  # J accumulates
            ni, no = C.c_int64(0), C.c_int64(0)
            _lib.call("pxrhip_tiles_stats", C.byref(ni), C.byref(no))
            cut += no.value
        sim.check()
    finally:
        set_tuning("mixed.jmargin", old_m)
        set_tuning("fused.j_lds_bytes", old_j)
    ref.gather_push_deposit(zero_j=True)
    ref.gather_push_deposit(zero_j=False)
    assert cut > 0, "no work item was cut"
    err = errors(sim, ref)
    for k in ("jx", "jy", "jz"):
        assert err[k] <= MIXED_TOL, (k, err[k])
    for k, v in err.items():
        if k.endswith(("ux", "uy", "uz")):
            assert v <= MIXED_TOL, (k, v)


@pytest.mark.parametrize("jmargin", [0, 1, -1])
def test_motion_room_settings_give_the_same_steps(torch, orc, jmargin):
    """Three steps of a fast and a slow species with each motion room
    around the work items' current blocks (mixed.jmargin 0, 1 node per side,
    or -1: automatic, the default) stay within the ten-step bound of the
    oracle."""
    old_m = set_tuning("mixed.jmargin", jmargin)
    try:
        sim, ref, _ = mixed_build(torch, orc, 3, True, "vay", "esirkepov", ppc=2,
                                  uth=(0.6, 0.02), seed=5)
        for sp, rs in zip(sim.species, ref.species):
            for k in ("ux", "uy", "uz", "w"):
                rs[k][...] = rs[k].astype(np.float32).astype(np.float64)
            rs["gaminv"][...] = 1.0 / np.sqrt(1 + rs["ux"] ** 2 + rs["uy"] ** 2 + rs["uz"] ** 2)
        for _ in range(3):
            sim.step(1)
            ref.step()
    finally:
        set_tuning("mixed.jmargin", old_m)
    err = errors(sim, ref)
    for k in ("jx", "jy", "jz", "ex", "ey", "ez", "bx", "by", "bz"):
        assert err[k] <= 1e-4, (k, err[k])


def test_ten_steps_against_fp64_oracle(torch, orc):
    sim, ref, _ = mixed_build(torch, orc, 3, True, "vay", "esirkepov", ppc=2)
    for sp, rs in zip(sim.species, ref.species):
        for k in ("ux", "uy", "uz", "w"):
            rs[k][...] = rs[k].astype(np.float32).astype(np.float64)
        rs["gaminv"][...] = 1.0 / np.sqrt(1 + rs["ux"] ** 2 + rs["uy"] ** 2 + rs["uz"] ** 2)
    for _ in range(10):
        sim.step(1)
        ref.step()
    err = errors(sim, ref)
    for k in ("jx", "jy", "jz", "ex", "ey", "ez", "bx", "by", "bz"):
        assert err[k] <= 1e-4, (k, err[k])
    for k, v in err.items():
        if k.endswith(("ux", "uy", "uz")):
            assert v <= 1e-4, (k, v)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_mixed_esirkepov_continuity(torch, orc, order):
    """(rho_new - rho_old)/dt + div J with J from the mixed-precision tiled
    kernel and FP64 rho at the positions before and after the push: zero
    to FP32 round-off (relative to rho/dt), although particles move only a
    small fraction of a cell (the factored shape differences)."""
    from picsar_amd.devmem import DeviceArray, DeviceGrid
    from picsar_amd.picsarpy import picsar
    sim, ref, _ = mixed_build(torch, orc, order, True, "vay", "esirkepov", ppc=2,
                              uth=(0.02, 0.002), field_amp=0.0)
    n, g = (sim.nx, sim.ny, sim.nz), sim.ng

    def rho_now():
        r = DeviceGrid(*n, g)
        for sp in sim.species:
            a = sp.arrays()
            picsar.pxr_depose_rho_n(r, sp.npart, *[DeviceArray.from_numpy(a[k]) for k in
                                                   ("x", "y", "z")],
                                    DeviceArray.from_numpy(a["w"].astype(np.float64)), sp.q,
                                    sim.xmin, sim.ymin, sim.zmin, sim.dx, sim.dy, sim.dz, *n,
                                    g, g, g, order, order, order)
        return r.numpy()
    r0 = rho_now()
    sim.gather_push_deposit()
    sim.check()
    r1 = rho_now()
    J = [sim.fields[k].numpy() for k in ("jx", "jy", "jz")]
    div = np.zeros_like(r0)
    div[1:, :, :] += (J[0][1:, :, :] - J[0][:-1, :, :]) / sim.dx
    div[:, 1:, :] += (J[1][:, 1:, :] - J[1][:, :-1, :]) / sim.dy
    div[:, :, 1:] += (J[2][:, :, 1:] - J[2][:, :, :-1]) / sim.dz
    res = (r1 - r0) / sim.dt + div
    scale = np.abs(r1).max() / sim.dt
    assert np.abs(res[1:, 1:, 1:]).max() <= 1e-6 * scale, np.abs(res[1:, 1:, 1:]).max() / scale


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_kinetic_energy_of_a_cold_species(torch, orc, precision):
    """compute_kin_energy (kin_energy.F90:33-143): m c^2 sum (gamma - 1) w.  A
    mixed-precision species sums gamma - 1 = u^2 / (1 + sqrt(1 + u^2)) in FP64
    from its FP32 momenta, which stays exact for ions at u_th = 0.001 c (from
    an FP32 gaminv the sum would be off by percents)."""
    kw = dict(precision="mixed", id_scale=2.0 ** -15, ppc=2) if precision == "mixed" else {}
    sim, ref, _ = build(torch, orc, 3, True, "vay", "esirkepov", uth=(0.01, 0.001), **kw)
    sim.step(2)
    want = 0.0
    for sp in sim.species:
        a = sp.arrays()
        usq = sum(a[k].astype(np.float64) ** 2 for k in ("ux", "uy", "uz"))
        want += sp.m * (usq / (1.0 + np.sqrt(1.0 + usq)) * a["w"].astype(np.float64)).sum()
    got = sim.kinetic_energy()
    # FP64: the reference's 1/gaminv - 1 resolves gamma - 1 ~ 5e-7 to ~1e-16/5e-7;
    # mixed: the same FP64 formula on the same momenta, summed in another order
    tol = 1e-8 if precision == "fp64" else 1e-12
    assert abs(got - want) <= tol * want, (got, want)


@pytest.mark.parametrize("order,depo", [(1, "direct"), (3, "esirkepov")])
def test_step_against_unrounded_fp64_oracle(torch, orc, order, depo):
    """The oracle runs on the FP64 momenta and gammas the mixed-precision
    species was built from, not on their FP32 roundings: the error includes
    the FP32 storage of the momenta, gamma and weights.  Same bars: 1e-5
    after one step, 1e-4 after ten."""
    sim, ref, _ = mixed_build(torch, orc, order, True, "vay", depo, ppc=2, seed=21)
    unique_float32_weights(ref)
    sim.step(1)
    ref.step()
    err = errors(sim, ref)
    for k in ("jx", "jy", "jz", "ex", "ey", "ez", "bx", "by", "bz"):
        assert err[k] <= MIXED_TOL, (k, err[k])
    for k, v in err.items():
        if k.endswith(("ux", "uy", "uz")):
            assert v <= MIXED_TOL, (k, v)
        if k.endswith(".abs"):
            assert v <= 1e-6, (k, v)
    for _ in range(9):
        sim.step(1)
        ref.step()
    err = errors(sim, ref)
    for k in ("jx", "jy", "jz", "ex", "ey", "ez", "bx", "by", "bz"):
        assert err[k] <= 1e-4, (k, err[k])
