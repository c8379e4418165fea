"""Invariants that pin the oracle's particle boundaries (tests/pic_oracle.py,
OracleState.particle_bcs) independently of its own formulas.

The reference's two exchanges (particle_boundaries.F90):
  * partcom 0, its default (control_file.F90:139):
    particle_bcs_tiles_and_mpi_3d (:2107-3031), whose SELECT CASE on the
    face knows absorbing (1) and reinjecting (3) and sends every other value
    to the periodic wrap (:2308-2333);
  * partcom 1, 2: particle_bcs_tiles then particle_bcs_mpi_blocking
    (:1040-1371), where a reflecting face (2) moves the particle one cell
    back in and reverses its normal momentum.

Each check below is one a plausible slip in the oracle would fail: a wrong
sign or side, a wrap by the wrong length, an axis mix-up, a reinjection
clamped to the wrong tile, a removal of particles that did not cross an
absorbing face, or a reflecting face treated as such under partcom 0.
"""
import numpy as np
import pytest

from tests.pic_oracle import OracleState

NX, NY, NZ = 12, 10, 14
D = (1.0, 0.8, 1.25)
X0 = (-2.0, 0.5, 3.0)
L = (NX * D[0], NY * D[1], NZ * D[2])


def make_state(pbound, partcom, nt=(3, 2, 2)):
    st = OracleState(None, NX, NY, NZ, 3, *D, *X0, 0.4, 1, True, "boris", "esirkepov")
    st.pbound = tuple(pbound)
    st.partcom = partcom
    st.nt = nt
    return st


def particles(rng, n=4000, overshoot=0.45):
    """Positions inside the box and just outside every face (by less than
    `overshoot` cells, as one step of a particle slower than c moves it);
    "old" positions (before the push) inside the box, in the tile the
    particle starts from."""
    x = [rng.uniform(X0[a], X0[a] + L[a], n) for a in range(3)]
    old = [v.copy() for v in x]
    kind = rng.integers(0, 7, n)  # 0: inside; 1..6: out through face kind-1
    for i in np.nonzero(kind)[0]:
        f = kind[i] - 1
        a, side = f // 2, f % 2
        dd = rng.uniform(1e-9, overshoot) * D[a]
        if side == 0:
            old[a][i] = X0[a] + rng.uniform(0.0, 0.5) * D[a]
            x[a][i] = X0[a] - dd
        else:
            old[a][i] = X0[a] + L[a] - rng.uniform(1e-9, 0.5) * D[a]
            x[a][i] = X0[a] + L[a] + dd
    u = [rng.standard_normal(n) * 0.3 for _ in range(3)]
    w = 1.0 + np.arange(n) * 1e-6
    return x, old, u, w, kind


def run(pbound, partcom, seed=5, **kw):
    rng = np.random.default_rng(seed)
    x, old, u, w, kind = particles(rng, **kw)
    st = make_state(pbound, partcom)
    st.add_species(-1.0, 1.0, *x, *u, w)
    s = st.species[0]
    s["old"] = tuple(v.copy() for v in old)
    before = {k: s[k].copy() for k in ("x", "y", "z", "ux", "uy", "uz", "w")}
    st.particle_bcs()
    return st, before, kind


def inside(s):
    ok = np.ones(len(s["x"]), dtype=bool)
    for a, k in enumerate(("x", "y", "z")):
        ok &= (s[k] >= X0[a]) & (s[k] < X0[a] + L[a])
    return ok


def by_weight(s, before):
    """Indices into `before` of the particles kept in s (weights are ids)."""
    return np.searchsorted(before["w"], s["w"])


@pytest.mark.parametrize("partcom", [0, 1])
def test_periodic_faces_wrap_by_whole_lengths(partcom):
    st, b, kind = run((0,) * 6, partcom)
    s = st.species[0]
    assert len(s["x"]) == len(b["x"])  # nothing removed
    assert inside(s).all()
    for a, k in enumerate(("x", "y", "z")):
        shift = (s[k] - b[k]) / L[a]
        assert np.all(np.isin(np.round(shift), (-1.0, 0.0, 1.0)))
        assert np.abs(shift - np.round(shift)).max() < 1e-12
        moved = np.round(shift) != 0
        # exactly the particles that crossed a face of this axis moved
        np.testing.assert_array_equal(moved, (kind == 2 * a + 1) | (kind == 2 * a + 2))
    for k in ("ux", "uy", "uz"):
        np.testing.assert_array_equal(s[k], b[k])


@pytest.mark.parametrize("partcom", [0, 1])
def test_absorbing_faces_remove_exactly_the_particles_outside(partcom):
    st, b, kind = run((1,) * 6, partcom)
    s = st.species[0]
    outside = np.zeros(len(b["x"]), dtype=bool)
    for a, k in enumerate(("x", "y", "z")):
        outside |= (b[k] < X0[a]) | (b[k] >= X0[a] + L[a])
    np.testing.assert_array_equal(np.sort(s["w"]), np.sort(b["w"][~outside]))
    for k in ("x", "y", "z", "ux", "uy", "uz"):
        np.testing.assert_array_equal(s[k], b[k][~outside])


def test_one_absorbing_face_removes_only_its_crossers():
    # absorbing at y_max only: particles out through the other faces wrap
    pb = (0, 0, 0, 1, 0, 0)
    st, b, kind = run(pb, 0)
    s = st.species[0]
    removed = np.setdiff1d(b["w"], s["w"])
    np.testing.assert_array_equal(np.sort(removed), np.sort(b["w"][kind == 4]))
    assert inside(s).all()


def test_reflecting_faces_reflect_under_partcom_1():
    st, b, kind = run((2,) * 6, 1)
    s = st.species[0]
    assert len(s["x"]) == len(b["x"])  # reflection keeps every particle
    assert inside(s).all()
    for a, (k, uk) in enumerate((("x", "ux"), ("y", "uy"), ("z", "uz"))):
        crossed = (kind == 2 * a + 1) | (kind == 2 * a + 2)
        # the normal momentum is reversed exactly, the others untouched
        np.testing.assert_array_equal(s[uk][crossed], -b[uk][crossed])
        np.testing.assert_array_equal(s[uk][~crossed], b[uk][~crossed])
        # the particle is back inside, within one cell of the face it crossed
        lo = kind == 2 * a + 1
        hi = kind == 2 * a + 2
        assert np.all(s[k][lo] - X0[a] < D[a]) and np.all(s[k][lo] >= X0[a])
        assert np.all(X0[a] + L[a] - s[k][hi] <= D[a]) and np.all(s[k][hi] < X0[a] + L[a])
        # and moved by less than one cell
        assert np.abs(s[k] - b[k]).max() <= D[a] * (1 + 1e-12)


def test_reflecting_is_periodic_under_partcom_0():
    """particle_bcs_tiles_and_mpi_3d has no reflecting branch: its faces wrap."""
    st_r, b, _ = run((2,) * 6, 0)
    st_p, _, _ = run((0,) * 6, 0)
    for k in ("x", "y", "z", "ux", "uy", "uz", "w"):
        np.testing.assert_array_equal(st_r.species[0][k], st_p.species[0][k])
    st_1, _, _ = run((2,) * 6, 1)
    assert not np.array_equal(st_1.species[0]["ux"], st_r.species[0]["ux"])


@pytest.mark.parametrize("partcom", [0, 1])
def test_reinjecting_faces_stop_particles_in_their_tile(partcom):
    st, b, kind = run((3,) * 6, partcom)
    s = st.species[0]
    assert len(s["x"]) == len(b["x"])
    crossed = kind > 0
    for uk in ("ux", "uy", "uz"):
        assert np.all(s[uk][crossed] == 0.0)
        np.testing.assert_array_equal(s[uk][~crossed], b[uk][~crossed])
    old = st.species[0]["old"]
    for a, k in enumerate(("x", "y", "z")):
        bounds, n0 = st.tile_bounds(a)
        # the tile the particle came from, by brute force over the bounds
        told = np.array([next(t for t, (lo, hi) in enumerate(bounds) if lo <= v < hi)
                         for v in old[a]])
        lo = np.array([bounds[t][0] for t in told]) + 0.5 * D[a]
        hi = np.array([bounds[t][1] for t in told]) - 0.5 * D[a]
        assert np.all(s[k][crossed] >= lo[crossed] - 1e-12)
        assert np.all(s[k][crossed] <= hi[crossed] + 1e-12)
        np.testing.assert_array_equal(s[k][~crossed], b[k][~crossed])
    assert inside(s).all()


def test_mixed_faces_conserve_particles_without_absorption():
    pb = (3, 2, 0, 2, 3, 0)
    for partcom in (0, 1):
        st, b, kind = run(pb, partcom, seed=11)
        assert len(st.species[0]["x"]) == len(b["x"])
        assert inside(st.species[0]).all()
