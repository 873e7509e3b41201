"""The oracle's physics against the reference's own functions, compiled from
its headers (include/solver/EulerPhysics.hpp:39-155, AdvectionPhysics.hpp:
45-66; oracle/_ref, built by `make -C oracle ref` where the reference
sources are present): primitiveToConservative, rusanovFlux and getMaxSpeed
for 2D and 3D Euler states, and the advection rusanovFlux, bit for bit on
seeded random states and edge cases.  This pins the rounding order of the
Euler flux (the GPU kernels are pinned against the oracle by the GPU tests).
Skipped where the reference library is not built.  The libraries are opened
by fixtures when a test of this module runs, not when the module is
collected, so a run that selects none of these tests (the GPU tests) never
maps them."""
import ctypes

import numpy as np
import pytest

import oracle

REF = None
REFM = None


@pytest.fixture(scope="module", autouse=True)
def _reference_libraries():
    """Opens oracle/_ref's libraries for the tests of this module."""
    global REF, REFM
    REF = oracle.load_ref()
    REFM = oracle.load_ref_morton()
    yield
    REF = REFM = None


@pytest.fixture
def ref():
    if REF is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return REF


@pytest.fixture
def refm():
    if REFM is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return REFM


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ref_prim_to_cons(rank, prim, gamma):
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.zeros(rank + 2)
    assert REF.ref_euler_prim_to_cons(rank, _p(prim), _p(out), gamma) == 0
    return out


def ref_flux(rank, UL, UR, d, gamma):
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    out = np.zeros(rank + 2)
    assert REF.ref_euler_flux(rank, _p(UL), _p(UR), _p(out), d, gamma) == 0
    return out


def ref_max_speed(rank, U, d, gamma):
    U = np.ascontiguousarray(U, dtype=np.float64)
    return REF.ref_euler_max_speed(rank, _p(U), d, gamma)


def random_prims(rng, rank, count):
    """Positive densities and pressures over several decades, velocities of
    both signs (including exact zeros)."""
    prim = np.empty((count, rank + 2))
    prim[:, 0] = 10.0 ** rng.uniform(-3, 3, count)
    prim[:, 1:rank + 1] = rng.normal(0.0, 3.0, (count, rank))
    zero = rng.random((count, rank)) < 0.1
    prim[:, 1:rank + 1][zero] = 0.0
    prim[:, rank + 1] = 10.0 ** rng.uniform(-3, 3, count)
    return prim


def same_bits(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.uint64), np.asarray(b, dtype=np.float64).view(np.uint64))


@pytest.mark.parametrize("rank", [2, 3])
@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0])
def test_prim_to_cons_matches_reference(ref, rank, gamma):
    rng = np.random.default_rng(100 + rank)
    prims = random_prims(rng, rank, 400)
    ours = oracle.euler_prim_to_cons(rank, prims, gamma)
    for k in range(len(prims)):
        assert same_bits(ours[k], ref_prim_to_cons(rank, prims[k], gamma)), k


@pytest.mark.parametrize("rank", [2, 3])
@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0])
def test_rusanov_flux_and_max_speed_match_reference(ref, rank, gamma):
    rng = np.random.default_rng(200 + rank)
    cons = oracle.euler_prim_to_cons(rank, random_prims(rng, rank, 600), gamma)
    for k in range(0, len(cons), 2):
        UL, UR = cons[k], cons[k + 1]
        for d in range(rank):
            assert same_bits(oracle.euler_flux(rank, UL, UR, d, gamma), ref_flux(rank, UL, UR, d, gamma)), (k, d)
            assert same_bits(oracle.euler_max_speed(rank, UL, d, gamma), ref_max_speed(rank, UL, d, gamma)), (k, d)
    # identical states, a state against its mirror, the acoustic pulse's
    # background and peak
    for UL, UR in [(cons[0], cons[0]), (cons[1], cons[1] * np.r_[1.0, -np.ones(rank), 1.0])]:
        for d in range(rank):
            assert same_bits(oracle.euler_flux(rank, UL, UR, d, gamma), ref_flux(rank, UL, UR, d, gamma))
    bg = oracle.euler_prim_to_cons(rank, np.r_[0.5, np.zeros(rank), 1.0], gamma)
    pk = oracle.euler_prim_to_cons(rank, np.r_[2.5, np.zeros(rank), 11.0], gamma)
    for d in range(rank):
        assert same_bits(oracle.euler_flux(rank, bg, pk, d, gamma), ref_flux(rank, bg, pk, d, gamma))


@pytest.mark.parametrize("rank", [2, 3])
def test_advection_flux_matches_reference(ref, rank):
    """The reference's AdvectionPhysics has the fixed velocity (1, 0.5, 0);
    the oracle's flux takes the velocity component as an argument."""
    rng = np.random.default_rng(300 + rank)
    vel = (1.0, 0.5, 0.0)
    vals = np.concatenate([rng.normal(0, 1, 500), 10.0 ** rng.uniform(-300, 300, 100), [0.0, -0.0, 1.0, -1.0]])
    for k in range(0, len(vals) - 1):
        for d in range(rank):
            ours = oracle.advection_flux(vals[k], vals[k + 1], vel[d])
            assert same_bits(ours, REF.ref_advection_flux(rank, vals[k], vals[k + 1], d)), (k, d)


@pytest.mark.parametrize("rank,depth,levels", [(2, 7, 3), (2, 7, 5), (3, 5, 2), (3, 5, 3)])
def test_uniform_tree_keys_match_reference_encoding(refm, rank, depth, levels):
    """Every key of a uniform tree equals morton_id<depth, rank>::encode of
    its cell: libmorton's encode of the coordinates at the 2^depth
    resolution, << 6, | level (the reference's dependency, not a retyped
    interleave)."""
    o = oracle.OracleTree(rank, 4, 1, depth, (1 << (rank * levels)) * 2 + 64)
    for _ in range(levels):
        o.refine_all()
    step = 1 << (depth - levels)
    n = 1 << levels
    exp = []
    for z in range(n if rank == 3 else 1):
        for y in range(n):
            for x in range(n):
                exp.append(REFM.ref_morton_key(rank, x * step, y * step, z * step, levels))
    np.testing.assert_array_equal(o.keys(), np.sort(np.array(exp, dtype=np.uint64)))
