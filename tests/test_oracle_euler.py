"""The oracle's Euler physics (include/solver/EulerPhysics.hpp and the
amr_solver step / dt over its fields), pinned by closed forms and invariants
that do not restate its expressions:

- primitive -> conservative: rho and rho*u exact, E to rounding;
- F(U, U) is the physical flux f(U) (consistency) for every direction;
- a stationary contact (u = 0, equal p) has the closed-form dissipative flux;
- getMaxSpeed is |u_d| + sqrt(gamma p / rho);
- a constant state is preserved exactly (uniform and mixed trees);
- rows of a state varying along x only stay identical (y fluxes cancel);
- totals of rho, rho*u, E are conserved on periodic uniform trees;
- one step on a uniform tree agrees with an independent dense periodic grid
  computation (numpy) to rounding.

Rounding-level expression order of the flux follows EulerPhysics.hpp:72-125
as written; no check above pins it (the reference cannot be built here).
"""
import numpy as np
import pytest

from oracle import OracleEuler, euler_flux, euler_max_speed, euler_prim_to_cons

GAMMA = 1.4


def random_prim(rng, rank, n):
    prim = np.zeros((n, rank + 2))
    prim[:, 0] = rng.uniform(0.5, 2.0, n)
    prim[:, 1 : rank + 1] = rng.uniform(-1.0, 1.0, (n, rank))
    prim[:, rank + 1] = rng.uniform(0.5, 3.0, n)
    return prim


def physical_flux(rank, U, d, gamma):
    rho = U[0]
    vel = U[1 : rank + 1] / rho
    E = U[rank + 1]
    p = (gamma - 1.0) * (E - 0.5 * rho * np.dot(vel, vel))
    f = np.zeros(rank + 2)
    f[0] = rho * vel[d]
    f[1 : rank + 1] = rho * vel * vel[d]
    f[1 + d] += p
    f[rank + 1] = (E + p) * vel[d]
    return f


@pytest.mark.parametrize("rank", [2, 3])
def test_prim_to_cons(rank):
    rng = np.random.default_rng(rank)
    prim = random_prim(rng, rank, 200)
    cons = euler_prim_to_cons(rank, prim, GAMMA)
    assert np.array_equal(cons[:, 0], prim[:, 0])
    assert np.array_equal(cons[:, 1 : rank + 1], prim[:, :1] * prim[:, 1 : rank + 1])
    E = prim[:, rank + 1] / (GAMMA - 1.0) + 0.5 * prim[:, 0] * (prim[:, 1 : rank + 1] ** 2).sum(1)
    np.testing.assert_allclose(cons[:, rank + 1], E, rtol=4e-16, atol=0)


@pytest.mark.parametrize("rank", [2, 3])
def test_flux_consistency(rank):
    rng = np.random.default_rng(10 + rank)
    cons = euler_prim_to_cons(rank, random_prim(rng, rank, 100), GAMMA)
    for U in cons:
        for d in range(rank):
            F = euler_flux(rank, U, U, d, GAMMA)
            np.testing.assert_allclose(F, physical_flux(rank, U, d, GAMMA), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("rank", [2, 3])
def test_stationary_contact(rank):
    rhoL, rhoR, p = 1.0, 0.25, 0.7
    for d in range(rank):
        primL = np.zeros(rank + 2)
        primL[0], primL[rank + 1] = rhoL, p
        primR = primL.copy()
        primR[0] = rhoR
        UL = euler_prim_to_cons(rank, primL, GAMMA)
        UR = euler_prim_to_cons(rank, primR, GAMMA)
        F = euler_flux(rank, UL, UR, d, GAMMA)
        smax = max(np.sqrt(GAMMA * p / rhoL), np.sqrt(GAMMA * p / rhoR))
        np.testing.assert_allclose(F[0], -0.5 * smax * (rhoR - rhoL), rtol=1e-15)
        mom = np.zeros(rank)
        mom[d] = p
        np.testing.assert_array_equal(F[1 : rank + 1], mom)
        np.testing.assert_allclose(F[rank + 1], -0.5 * smax * (UR[rank + 1] - UL[rank + 1]), rtol=1e-15)
        # swapping the states reverses the dissipative mass flux
        assert euler_flux(rank, UR, UL, d, GAMMA)[0] == -F[0]


@pytest.mark.parametrize("rank", [2, 3])
def test_max_speed(rank):
    rng = np.random.default_rng(20 + rank)
    prim = random_prim(rng, rank, 50)
    cons = euler_prim_to_cons(rank, prim, GAMMA)
    for P, U in zip(prim, cons):
        for d in range(rank):
            exp = abs(P[1 + d]) + np.sqrt(GAMMA * P[rank + 1] / P[0])
            np.testing.assert_allclose(euler_max_speed(rank, U, d, GAMMA), exp, rtol=1e-14)


def build(rank, n, levels, cap_extra=64, lengths=(1.0, 1.0, 1.0), flags=None):
    depth = 7 if rank == 2 else 5
    e = OracleEuler(rank, n, 1, depth, 2 * ((1 << (rank * max(levels, 1))) + cap_extra), lengths)
    e.set_params(GAMMA, 0.3, threads=4)
    for _ in range(levels):
        e.refine_all()
    if flags is not None:
        e.reconstruct(flags(e))
    return e


def field_from_prim(e, prim_fn):
    cen = e.cell_centers()
    prim = prim_fn(cen)  # (leaves, ni, nv)
    cons = euler_prim_to_cons(e.rank, prim.reshape(-1, e.nv), GAMMA).reshape(prim.shape)
    return np.ascontiguousarray(np.moveaxis(cons, -1, 0))


@pytest.mark.parametrize("rank,n,mixed", [(2, 8, False), (2, 8, True), (3, 4, False), (3, 4, True)])
def test_constant_state_preserved(rank, n, mixed):
    def flags(e):
        f = np.zeros(e.size(), dtype=np.int8)
        f[::3] = 1
        return f

    e = build(rank, n, 2, flags=flags if mixed else None)

    def prim_fn(cen):
        p = np.zeros(cen.shape[:-1] + (rank + 2,))
        p[..., 0] = 1.0
        p[..., 1] = 0.5
        p[..., rank + 1] = 2.0
        return p

    u0 = field_from_prim(e, prim_fn)
    e.set_interior(u0)
    e.halo_exchange()
    _, ex = e.advance_batch(5)
    assert ex == 5
    assert np.array_equal(e.get_interior(), u0.reshape(e.get_interior().shape))


def test_rows_stay_identical():
    e = build(2, 8, 3)

    def prim_fn(cen):
        x = cen[..., 0]
        p = np.zeros(cen.shape[:-1] + (4,))
        p[..., 0] = 1.0 + 0.2 * np.sin(2 * np.pi * x)
        p[..., 1] = 0.3 * np.cos(2 * np.pi * x)
        p[..., 3] = 1.0 + 0.1 * np.sin(4 * np.pi * x)
        return p

    e.set_interior(field_from_prim(e, prim_fn))
    e.halo_exchange()
    e.advance_batch(10)
    u = e.get_interior()  # (nv, leaves, y, x)
    cen = e.cell_centers().reshape(e.size(), 8, 8, 2)
    # every cell equals the cell with the same x in the first row of leaf 0's column
    xs = {}
    for v in range(4):
        for i in range(e.size()):
            for yy in range(8):
                for xx in range(8):
                    key = (v, cen[i, yy, xx, 0])
                    val = u[v, i, yy, xx]
                    assert xs.setdefault(key, val) == val or (np.isnan(val) and np.isnan(xs[key]))
    assert np.all(u[2] == 0.0)  # y momentum stays zero


@pytest.mark.parametrize("rank,n", [(2, 8), (3, 4)])
def test_conservation_uniform(rank, n):
    e = build(rank, n, 2)
    rng = np.random.default_rng(5)
    amp = rng.uniform(0.05, 0.2, 3)

    def prim_fn(cen):
        s = np.sin(2 * np.pi * cen[..., 0]) * np.cos(2 * np.pi * cen[..., 1])
        p = np.zeros(cen.shape[:-1] + (rank + 2,))
        p[..., 0] = 1.0 + amp[0] * s
        p[..., 1] = amp[1] * s
        p[..., 2] = -amp[2] * s
        p[..., rank + 1] = 1.0 + amp[0] * s
        return p

    u0 = field_from_prim(e, prim_fn)
    e.set_interior(u0)
    e.halo_exchange()
    e.advance_batch(20)
    tot0 = u0.reshape(rank + 2, -1).sum(1)
    tot = e.get_interior().reshape(rank + 2, -1).sum(1)
    scale = np.abs(u0).reshape(rank + 2, -1).sum(1)
    assert np.all(np.abs(tot - tot0) <= 1e-13 * scale)


def dense_step(U, dx, dt, gamma):
    """One periodic Rusanov step on a dense 2D grid U[var, y, x], written from
    the conservation law (physical fluxes, local Lax-Friedrichs speed)."""
    def phys(U, d):
        rho = U[0]
        u, v = U[1] / rho, U[2] / rho
        p = (gamma - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v))
        vd = u if d == 0 else v
        F = np.stack([rho * vd, U[1] * vd, U[2] * vd, (U[3] + p) * vd])
        F[1 + d] += p
        return F, np.abs(vd) + np.sqrt(gamma * p / rho)

    out = U.copy()
    for d in (0, 1):
        ax = 2 - d  # array axis of physical dim d in U[var, y, x]
        F, s = phys(U, d)
        Fr, sr, Ur = np.roll(F, -1, axis=ax), np.roll(s, -1, axis=ax - 1), np.roll(U, -1, axis=ax)
        face = 0.5 * (F + Fr) - 0.5 * np.maximum(s, sr) * (Ur - U)  # flux through face i + 1/2
        out -= dt / dx * (face - np.roll(face, 1, axis=ax))
    return out


def test_dense_grid_one_step():
    import brute

    level, n = 3, 8
    e = build(2, n, level)

    def prim_fn(cen):
        x, y = cen[..., 0], cen[..., 1]
        p = np.zeros(cen.shape[:-1] + (4,))
        g = np.exp(-((x - 0.4) ** 2 + (y - 0.6) ** 2) / 0.02)
        p[..., 0] = 1.0 + 0.3 * g
        p[..., 1] = 0.2 * np.sin(2 * np.pi * y)
        p[..., 2] = -0.1 * np.cos(2 * np.pi * x)
        p[..., 3] = 1.0 + 0.5 * g
        return p

    u0 = field_from_prim(e, prim_fn)
    e.set_interior(u0)
    e.halo_exchange()
    dt = e.compute_dt()
    e.step(dt)
    ncell = n << level
    coords = np.array([brute.deinterleave(int(k) >> 6, 2) for k in e.keys()]) >> (7 - level)
    D0 = np.zeros((4, ncell, ncell))
    D1 = np.zeros((4, ncell, ncell))
    got = e.get_interior()
    for i, (cx, cy) in enumerate(coords):
        D0[:, cy * n : (cy + 1) * n, cx * n : (cx + 1) * n] = u0.reshape(4, e.size(), n, n)[:, i]
        D1[:, cy * n : (cy + 1) * n, cx * n : (cx + 1) * n] = got[:, i]
    # dt: cfl * min over cells and dims of dx / (|u_d| + c)
    rho = D0[0]
    pr = (GAMMA - 1.0) * (D0[3] - 0.5 * (D0[1] ** 2 + D0[2] ** 2) / rho)
    c = np.sqrt(GAMMA * pr / rho)
    smax = max((np.abs(D0[1] / rho) + c).max(), (np.abs(D0[2] / rho) + c).max())
    np.testing.assert_allclose(dt, 0.3 * (1.0 / ncell) / smax, rtol=1e-14)
    np.testing.assert_allclose(D1, dense_step(D0, 1.0 / ncell, dt, GAMMA), rtol=1e-12, atol=1e-13)
