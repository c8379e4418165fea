"""The oracle against the golden vectors written by the reference's own kernels.

tests/golden/*.npz were produced by tests/golden/gen_golden.py from the PICSAR
FP64 leaf routines compiled from their sources (oracle/build_ref.sh).  The C
restatement follows the reference's scalar routines operation by operation and
is compiled without FP contraction, so everything except the default
vectorised gathers (which group their sums differently,
field_gathering_o1_3d.F90:2194-2240 vs :180) must agree bit for bit.
"""
import numpy as np
import pytest


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("lo", [1, 0])
def test_gather(orc, golden, ggrid, order, lo):
    d = golden("gather")
    g = ggrid
    n = len(d["xp"])
    e = [np.zeros(n) for _ in range(6)]
    F = [np.asfortranarray(d[k]) for k in ("exg", "eyg", "ezg", "bxg", "byg", "bzg")]
    bad = orc.geteb3d_energy_conserving(d["xp"], d["yp"], d["zp"], *e, g["xmin"], g["ymin"],
                                        g["zmin"], g["dx"], g["dy"], g["dz"], g["nx"], g["ny"],
                                        g["nz"], g["ng"], g["ng"], g["ng"], order, order, order,
                                        *F, bool(lo))
    assert bad == 0
    out = np.stack(e)
    # scalar routines: identical operation order
    np.testing.assert_array_equal(out, d[f"o{order}_lo{lo}_a1"])
    # default vectorised routines: same sums, different grouping
    assert _rel(out, d[f"o{order}_lo{lo}_a0"]) < 1e-14


def test_push(orc, golden):
    d = golden("push")
    E = [d[k] for k in ("ex", "ey", "ez", "bx", "by", "bz")]
    q, m, dt = float(d["q"]), float(d["m"]), float(d["dt"])
    u = [d[k].copy() for k in ("ux", "uy", "uz")]
    g = d["gaminv"].copy()
    orc.boris_push_u_3d(*u, g, *E, q, m, dt)
    np.testing.assert_array_equal(np.stack(u + [g]), d["boris"])
    for which in (0, 1, 2):
        u = [d[k].copy() for k in ("ux", "uy", "uz")]
        g = d["gaminv"].copy()
        orc.ebcancelpush3d(*u, g, *E, q, m, dt, which)
        np.testing.assert_array_equal(np.stack(u + [g]), d[f"vay{which}"])
    g = np.zeros_like(d["gaminv"])
    orc.set_gamma(d["ux"], d["uy"], d["uz"], g)
    np.testing.assert_array_equal(g, d["set_gamma"])
    x = [d[k].copy() for k in ("x", "y", "z")]
    orc.pushxyz(*x, d["ux"], d["uy"], d["uz"], d["gaminv"], dt)
    np.testing.assert_array_equal(np.stack(x), d["pushxyz"])


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("algo", [0, 3])
def test_deposit(orc, golden, ggrid, order, algo):
    d = golden("deposit")
    g = ggrid
    shape = d[f"o{order}_a{algo}"].shape[1:]
    j = [np.zeros(shape, order="F") for _ in range(3)]
    args = (*j, d["xp"], d["yp"], d["zp"], d["ux"], d["uy"], d["uz"], d["gaminv"], d["w"],
            float(d["q"]), g["xmin"], g["ymin"], g["zmin"], float(d["dt"]), g["dx"], g["dy"],
            g["dz"], g["nx"], g["ny"], g["nz"], g["ng"], g["ng"], g["ng"], order)
    bad = orc.depose_esirkepov(*args) if algo == 0 else orc.depose_direct(*args, False)
    assert bad == 0
    np.testing.assert_array_equal(np.stack(j), d[f"o{order}_a{algo}"])


def test_fdtd(orc, golden, ggrid):
    d = golden("fdtd")
    g = ggrid
    nx, ny, nz, ng = g["nx"], g["ny"], g["nz"], g["ng"]
    F = [np.asfortranarray(d[f"in{i}"]) for i in range(9)]
    lo, hi = (-1, -1, -1), (nx + 1, ny + 1, nz + 1)
    mudt, sx, sy, sz = 0.45, 0.45 / g["dx"], 0.45 / g["dy"], 0.45 / g["dz"]
    A = [f.copy(order="F") for f in F]
    orc.push_em3d_evec(lo, hi, *A, mudt, sx, sy, sz, nx, ny, nz, ng, ng, ng)
    np.testing.assert_array_equal(np.stack(A[:3]), d["yee_e"])
    A = [f.copy(order="F") for f in F]
    orc.push_em3d_bvec(lo, hi, *A[:6], 0.5 * sx, 0.5 * sy, 0.5 * sz, nx, ny, nz, ng, ng, ng)
    np.testing.assert_array_equal(np.stack(A[3:6]), d["yee_b"])
    for no in (2, 4, 6):
        c = orc.fd_weights(no)
        np.testing.assert_array_equal(c, d[f"fdw{no}"])
        A = [f.copy(order="F") for f in F]
        orc.push_em3d_evec_norder(*A, mudt, sx * c, sy * c, sz * c, nx, ny, nz, no, no, no,
                                  ng, ng, ng, 1, 1, 1)
        np.testing.assert_array_equal(np.stack(A[:3]), d[f"no{no}_e"])
        A = [f.copy(order="F") for f in F]
        orc.push_em3d_bvec_norder(*A[:6], 0.5 * sx * c, 0.5 * sy * c, 0.5 * sz * c, nx, ny, nz,
                                  no, no, no, ng, ng, ng, 1, 1, 1)
        np.testing.assert_array_equal(np.stack(A[3:6]), d[f"no{no}_b"])
    np.testing.assert_array_equal(orc.fd_weights(8), d["fdw8"])
    np.testing.assert_array_equal(orc.fd_weights(16), d["fdw16"])


@pytest.mark.parametrize("case", ["", "2"])
def test_sort(orc, golden, case):
    d = golden("sort")
    P = [d[f"in{i}"].copy() for i in range(8)]
    bad = orc.particle_bin_sorting(*P, *d["box" + case])
    assert bad == 0
    np.testing.assert_array_equal(np.stack(P), d["out" + case])


@pytest.mark.parametrize("order", [1, 2, 3])
def test_rho(orc, golden, ggrid, order):
    """The oracle's charge density is pxr_depose_rho_n bit for bit; the
    scalar routines group the order-2 products differently (last ulp)."""
    d = golden("rho")
    g = ggrid
    rho = np.zeros(d[f"o{order}_n"].shape, order="F")
    bad = orc.depose_rho(rho, d["xp"], d["yp"], d["zp"], d["w"], float(d["q"]), g["xmin"],
                         g["ymin"], g["zmin"], g["dx"], g["dy"], g["dz"], g["nx"], g["ny"],
                         g["nz"], g["ng"], g["ng"], g["ng"], order)
    assert bad == 0
    np.testing.assert_array_equal(rho, d[f"o{order}_n"])
    assert _rel(rho, d[f"o{order}_scalar"]) < 1e-15


@pytest.mark.parametrize("case", range(7))
def test_laser_profile(orc, golden, case):
    """The oracle's antenna amplitudes are gaussian_profile / hanning_profile
    (laser_pusher_manager_3d.F90:368-478) to the last few ulps (complex exp
    and division may round differently)."""
    d = golden("laser")
    win, tord, pol, t, qr, qi = d[f"case{case}"]
    amp = orc.laser_profile(int(win), d["xx"], d["yy"], float(d["emax1"]), float(d["emax2"]),
                            d["pv1"], d["pv2"], float(d["k0"]), complex(qr, qi),
                            float(d["tau"]), float(t), float(d["t_peak"]), int(tord), float(pol))
    ref = d[f"amp{case}"]
    scale = max(np.abs(ref).max(), 1e-300)
    assert np.abs(amp - ref).max() <= 8 * np.finfo(float).eps * scale


def test_field_divergence(orc, golden):
    """calc_field_div and calc_field_divB (diags.F90:129-197) bit for bit,
    including the nodes outside their loops, which keep their value."""
    d = golden("div")
    nx, ny, nz, g = (int(d[k]) for k in ("nx", "ny", "nz", "ng"))
    h = tuple(float(d[k]) for k in ("dx", "dy", "dz"))
    F = [np.asfortranarray(d[k]) for k in ("fx", "fy", "fz")]
    for name, fn in (("dive", orc.calc_field_div), ("divb", orc.calc_field_divb)):
        out = np.asfortranarray(np.full(F[0].shape, 7.0))
        fn(out, *F, nx, ny, nz, g, g, g, *h)
        np.testing.assert_array_equal(out, d[name])


# One shape order per axis: the reference's generic routines, which the
# dispatchers call for unequal orders, order 0 (deposition), field_gathe_algo
# 2 and current_depo_algo 1.  Each case is stored with the fixture.
def test_generic_gather(orc, golden, ggrid):
    """pxr_gete3d_n/pxr_getb3d_n_energy_conserving bit for bit, including
    their single-precision 1./6. and 2./3. in the cubic weights (about 1e-7
    away from the order-3 routines' gather)."""
    d = golden("generic")
    g = ggrid
    n = len(d["g_xp"])
    F = [np.asfortranarray(d[k]) for k in ("exg", "eyg", "ezg", "bxg", "byg", "bzg")]
    for c, (nox, noy, noz, lo, nod, l4) in enumerate(d["gather_cases"]):
        e = [np.zeros(n) for _ in range(6)]
        bad = orc.geteb3d_n(d["g_xp"], d["g_yp"], d["g_zp"], *e, g["xmin"], g["ymin"],
                            g["zmin"], g["dx"], g["dy"], g["dz"], g["nx"], g["ny"], g["nz"],
                            g["ng"], g["ng"], g["ng"], int(nox), int(noy), int(noz), *F,
                            l4symtry=l4, l_lower_order_in_v=lo, l_nodal=nod)
        assert bad == 0
        np.testing.assert_array_equal(np.stack(e), d[f"gather{c}"], err_msg=f"case {c}")


def test_generic_esirkepov(orc, golden, ggrid):
    d = golden("generic")
    g = ggrid
    for c, o in enumerate(d["esirk_cases"]):
        j = [np.zeros(d[f"esirk{c}"].shape[1:], order="F") for _ in range(3)]
        bad = orc.depose_esirkepov_n(*j, d["d_xp"], d["d_yp"], d["d_zp"], d["ux"], d["uy"],
                                     d["uz"], d["gaminv"], d["w"], float(d["q"]), g["xmin"],
                                     g["ymin"], g["zmin"], float(d["dt"]), g["dx"], g["dy"],
                                     g["dz"], g["nx"], g["ny"], g["nz"], g["ng"], g["ng"],
                                     g["ng"], *map(int, o))
        assert bad == 0
        np.testing.assert_array_equal(np.stack(j), d[f"esirk{c}"], err_msg=f"case {c}")


def test_generic_rho(orc, golden, ggrid):
    d = golden("generic")
    g = ggrid
    for c, (nox, noy, noz, lpw, l4) in enumerate(d["rho_cases"]):
        rho = np.zeros(d[f"rho{c}"].shape, order="F")
        bad = orc.depose_rho_n(rho, d["r_xp"], d["r_yp"], d["r_zp"], d["w"], 0.7, g["xmin"],
                               g["ymin"], g["zmin"], g["dx"], g["dy"], g["dz"], g["nx"], g["ny"],
                               g["nz"], g["ng"], g["ng"], g["ng"], int(nox), int(noy), int(noz),
                               l_particles_weight=lpw, l4symtry=l4)
        assert bad == 0
        np.testing.assert_array_equal(rho, d[f"rho{c}"], err_msg=f"case {c}")
