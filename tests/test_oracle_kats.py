"""Pins of the CPU oracle against the reference's own tests and doctests, closed
forms, special cases and brute force (no GPU needed).

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a swapped operand) fails it.
"""
import math

import numpy as np
import pytest

from cellular_raza_amd import (
    BoundLennardJones,
    CartesianCuboid,
    CellSpecificInteraction,
    CellSpecificInteractionLangevin,
    InstanceSpec,
    MiePotential,
    MorsePotential,
    NewtonDamped,
    NoInteraction,
    RodMechanics,
    SpeciesMorsePotential,
)
from oracle import oracle as O


# ---------------------------------------------------------------------------
# Philox4x32-10: known-answer vectors of the Random123 distribution (kat_vectors)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize(
    "ctr,key,expected",
    [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ],
)
def test_philox_kat(ctr, key, expected):
    assert O.philox4x32_10(ctr, key) == expected


def test_gaussians_are_standard_normal():
    z = np.concatenate([O.gaussians(7, a, s, 3) for a in range(4000) for s in range(5)])
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02
    # fourth moment of N(0, 1) is 3
    assert abs((z ** 4).mean() - 3.0) < 0.15


# ---------------------------------------------------------------------------
# Grid, decomposition, stencil, voxel index
# ---------------------------------------------------------------------------
def test_grid_doctest_kat():
    """cartesian_cuboid_n.rs:167-182: [2,3,1]..[10,10,20], range 2 -> [4, 3, 9]."""
    d = CartesianCuboid.from_boundaries_and_interaction_range([2.0, 3.0, 1.0], [10.0, 10.0, 20.0], 2.0)
    n, dx = O.grid(d, 3)
    assert n == [4, 3, 9]
    assert dx == [8.0 / 4.0, 7.0 / 3.0, 19.0 / 9.0]


def test_grid_closed_form_random():
    rng = np.random.default_rng(0)
    for _ in range(200):
        lo = rng.uniform(-50, 50, 3)
        hi = lo + rng.uniform(0.1, 100, 3)
        r = rng.uniform(0.05, 30)
        d = CartesianCuboid.from_boundaries_and_interaction_range(lo, hi, r)
        n, dx = O.grid(d, 3)
        for k in range(3):
            f = math.floor((hi[k] - lo[k]) / r)
            assert n[k] == max(f, 1)
            if f > 0:
                assert dx[k] == (hi[k] - lo[k]) / f and dx[k] >= r
            else:
                assert math.isinf(dx[k])


def test_decomp_res_invariants():
    """cartesian_cuboid_n.rs:670-699 (debug-build sweep)."""
    for n_voxel in range(1, 500):
        for n_regions in range(1, 100):
            res = O.decomp_res(n_voxel, n_regions)
            assert res is not None, (n_voxel, n_regions)
            n, m, avg = res
            assert n + m == n_regions
            assert n * avg + m * (avg - 1) == n_voxel


def test_generate_subdomains_cover():
    """cartesian_cuboid_n.rs:380-402: 4 subdomains cover all 5^3 voxels (checked through
    the threaded engine: every agent is owned by exactly one voxel after stepping)."""
    d = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [100.0] * 3, 20.0)
    assert O.grid(d, 3)[0] == [5, 5, 5]
    rng = np.random.default_rng(3)
    pos = rng.uniform(1e-6, 100, (300, 3))
    spec = [InstanceSpec(d, NoInteraction(), NewtonDamped(0.0, 1.0))]
    s = O.OracleSim(spec, dt=0.1, n_subdomains=4)
    s.set_agents(pos, rng.uniform(-5, 5, (300, 3)))
    s.step(20)
    vox, start, order = s.get_voxels()
    assert start[-1] == 300 and sorted(order.tolist()) == list(range(300))


def test_neighbor_stencil_brute_force():
    """get_neighbor_voxel_indices (cartesian_cuboid_n.rs:641-662) vs direct enumeration."""
    for n in ([3, 4], [1, 5], [3, 3, 3], [2, 5, 4], [1, 1, 6]):
        D = len(n)
        V = int(np.prod(n))
        for plain in range(V):
            idx = np.unravel_index(plain, n)
            want = []
            for q in range(V):
                j = np.unravel_index(q, n)
                if q != plain and all(abs(int(a) - int(b)) <= 1 for a, b in zip(idx, j)):
                    want.append(q)
            assert O.neighbors(n, plain) == want
            assert len(want) <= 3 ** D - 1


def test_voxel_index_brute_force_and_faces():
    d = CartesianCuboid.from_boundaries_and_interaction_range([-3.0, 1.0, 0.5], [17.0, 12.0, 40.0], 2.5)
    n, dx = O.grid(d, 3)
    rng = np.random.default_rng(4)
    pts = rng.uniform(d.min, d.max, (3000, 3))
    # points planted on voxel faces
    faces = np.array([[d.min[k] + i * dx[k] for k in range(3)] for i in range(1, 6)])
    pts = np.concatenate([pts, faces])
    for p in pts:
        st, idx = O.voxel_index(d, p)
        assert st == 0
        assert idx == [int(math.trunc((p[k] - d.min[k]) / dx[k])) for k in range(3)]
    # initial placement: p == min fails the strict check, p == max has no voxel
    assert O.voxel_index(d, [d.min[0], 5.0, 5.0], initial=True)[0] == 2  # BoundaryError
    assert O.voxel_index(d, [d.max[0], 5.0, 5.0], initial=True)[0] == 3  # IndexError
    # per-step rule: values in (-1, 0) truncate to 0 (num-traits to_usize)
    st, idx = O.voxel_index(d, [d.min[0] - 0.5 * dx[0], 5.0, 5.0])
    assert st == 0 and idx[0] == 0


# ---------------------------------------------------------------------------
# Potentials
# ---------------------------------------------------------------------------
def morse_V(r, R, lam, V0):
    """Morse potential of the doc comment (interaction.rs:198-200)."""
    return V0 * (1.0 - math.exp(-lam * (r - R))) ** 2


def test_morse_special_cases_and_derivative():
    pot = MorsePotential(radius=1.0, potential_stiffness=1.3, cutoff=3.0, strength=0.7)
    R = 2.0
    # equilibrium d = R_own + R_ext: exactly zero
    st, fo, fe = O.pair_force(pot, [0.0, 0.0, 0.0], [R, 0.0, 0.0])
    assert st == 0 and np.all(fo == 0.0)
    # beyond the cutoff and at zero distance: zero
    assert np.all(O.pair_force(pot, [0.0, 0.0], [3.0001, 0.0])[1] == 0)
    assert np.all(O.pair_force(pot, [1.0, 1.0], [1.0, 1.0])[1] == 0)
    rng = np.random.default_rng(5)
    for _ in range(200):
        own = rng.uniform(-2, 2, 3)
        ext = own + rng.normal(size=3) * rng.uniform(0.3, 1.0)
        st, fo, fe = O.pair_force(pot, own, ext)
        r = np.linalg.norm(own - ext)
        # Newton's third law, bitwise
        assert np.array_equal(fe, -fo)
        if r > 3.0:
            assert np.all(fo == 0.0)
            continue
        # force = -dV/dr along (own - ext) / r
        h = 1e-6
        dV = (morse_V(r + h, R, 1.3, 0.7) - morse_V(r - h, R, 1.3, 0.7)) / (2 * h)
        want = -dV * (own - ext) / r
        assert np.allclose(fo, want, rtol=1e-6, atol=1e-9)


def test_species_morse_scales_strength():
    A = [[1.0, 0.2], [0.2, 1.0]]
    sp = SpeciesMorsePotential(radius=1.0, potential_stiffness=1.0, cutoff=2.5, strength=0.5, species_matrix=A)
    base = MorsePotential(radius=1.0, potential_stiffness=1.0, cutoff=2.5, strength=1.0)
    own, ext = [0.0, 0.0, 0.0], [1.7, 0.3, -0.4]
    f1 = O.pair_force(base, own, ext)[1]
    for a in (0, 1):
        for b in (0, 1):
            fs = O.pair_force(sp, own, ext, a, b)[1]
            assert np.allclose(fs, f1 * 0.5 * A[a][b], rtol=1e-15, atol=0)


def mie_U(r, sigma, n, m, eps):
    C = n / (n - m) * (n / m) ** (m / (n - m))
    return C * eps * ((sigma / r) ** n - (sigma / r) ** m)


def test_mie_reduces_to_lennard_jones_and_errors():
    pot = MiePotential(radius=1.0, strength=1.0, bound=1e9, cutoff=2.5, en=12.0, em=6.0)
    sigma = (6.0 / 12.0) ** (1.0 / 6.0) * 2.0
    assert math.isclose(sigma, 1.781797436, rel_tol=1e-9)
    rng = np.random.default_rng(6)
    for r in rng.uniform(1.6, 2.5, 50):
        st, fo, fe = O.pair_force(pot, [r, 0.0, 0.0], [0.0, 0.0, 0.0])
        # 12-6 Lennard-Jones force with C = 4
        lj = 24.0 * 1.0 / r * (2 * (sigma / r) ** 12 - (sigma / r) ** 6)
        assert math.isclose(fo[0], lj, rel_tol=1e-12)
        h = 1e-6
        dU = (mie_U(r + h, sigma, 12, 6, 1.0) - mie_U(r - h, sigma, 12, 6, 1.0)) / (2 * h)
        assert math.isclose(fo[0], -dU, rel_tol=1e-6)
    # minimum of the potential at r = 2^(1/6) sigma = R_own + R_ext
    st, fo, _ = O.pair_force(pot, [2.0, 0, 0], [0, 0, 0])
    assert abs(fo[0]) < 1e-12
    # r == 0 -> CalcError; beyond the cutoff -> zero; bound caps the force
    assert O.pair_force(pot, [1.0, 2.0, 3.0], [1.0, 2.0, 3.0])[0] == 1
    assert np.all(O.pair_force(pot, [2.6, 0, 0], [0, 0, 0])[1] == 0)
    capped = MiePotential(radius=1.0, strength=1.0, bound=3.0, cutoff=2.5, en=12.0, em=6.0)
    assert O.pair_force(capped, [1.2, 0, 0], [0, 0, 0])[1][0] == 3.0


def test_bound_lennard_jones_structure():
    """interaction.rs:130-137: val = 4 eps / r (12 (s/r)^11 - 6 (s/r)^5), force on own
    = -dir * q * min(bound / r, val); zero crossing at r = 2^(1/6) sigma."""
    pot = BoundLennardJones(epsilon=0.8, sigma=1.5, bound=1e9, cutoff=4.0)
    r0 = 2 ** (1 / 6) * 1.5
    st, fo, fe = O.pair_force(pot, [r0, 0.0], [0.0, 0.0])
    assert abs(fo[0]) < 1e-13
    for r in (1.2, 1.5, 1.9, 2.5, 3.9):
        st, fo, fe = O.pair_force(pot, [r, 0.0], [0.0, 0.0])
        val = 4 * 0.8 / r * (12 * (1.5 / r) ** 11 - 6 * (1.5 / r) ** 5)
        assert math.isclose(fo[0], -val, rel_tol=1e-13)
        assert np.array_equal(fe, -fo)
    # beyond cutoff: q = 0; bound/r caps from above
    assert np.all(O.pair_force(pot, [4.1, 0], [0, 0])[1] == 0)
    capped = BoundLennardJones(epsilon=0.8, sigma=1.5, bound=2.0, cutoff=4.0)
    assert math.isclose(O.pair_force(capped, [1.0, 0], [0, 0])[1][0], -2.0 / 1.0)


def test_cell_sorting_interaction_cases():
    """CellSpecificInteraction: strength = s * clamp(1/sig^2 - 1/sig^4, +-(4 + 1/sig)),
    sig = r / (R + R); attraction only between equal species; zero beyond
    relative_interaction_range * (R + R); the example variant has the opposite sign."""
    R, s = 6.0, 2.0
    bench = CellSpecificInteraction(cell_radius=R, potential_strength=s, relative_interaction_range=1.5)
    ex = CellSpecificInteraction(cell_radius=R, potential_strength=s, relative_interaction_range=1.5,
                                 variant="example")
    for sig in (0.5, 1.0, math.sqrt(2.0), 1.4):
        r = sig * 2 * R
        want = s * max(min(1 / sig ** 2 - 1 / sig ** 4, 4 + 1 / sig), -(4 + 1 / sig))
        _, same, _ = O.pair_force(bench, [r, 0, 0], [0, 0, 0], 0, 0)
        _, diff, _ = O.pair_force(bench, [r, 0, 0], [0, 0, 0], 0, 1)
        _, same_ex, _ = O.pair_force(ex, [r, 0, 0], [0, 0, 0], 1, 1)
        assert math.isclose(same[0], want, rel_tol=1e-13, abs_tol=1e-15)
        assert math.isclose(diff[0], min(want, 0.0), rel_tol=1e-13, abs_tol=1e-15)
        assert same_ex[0] == -same[0]
    # zero beyond the interaction range, clamp below 0.316 R
    assert np.all(O.pair_force(bench, [1.5 * 2 * R + 0.01, 0, 0], [0, 0, 0])[1] == 0)
    near = O.pair_force(bench, [0.1, 0, 0], [0, 0, 0])[1]
    clamp = O.pair_force(bench, [0.3162277660168379 * R, 0, 0], [0, 0, 0])[1]
    assert np.array_equal(near, clamp)
    assert np.all(O.pair_force(bench, [1, 2, 3], [1, 2, 3])[1] == 0)


def test_cell_sorting_langevin_regimes():
    """main-langevin.rs:44-92.  Contact (r <= R = 2 cell_radius): linear in the
    overlap, slope = repulsion, independent of species, zero at R.  Tail (R < r <=
    cutoff): F c / ((r - R)(c - r)) = a exp(-(r / R - 1)^2), so its logarithm is a
    parabola in r with second difference -2 h^2 / R^2 and the factor a is the (own,
    ext) entry of the attraction matrix (a transposed matrix fails).  Beyond the
    cutoff and at r = 0: zero, no error.  Sign: the example returns (-force, force),
    so in contact the own agent is pushed towards the other (f_own . (own - ext) <
    0) and the attraction term pushes it away."""
    R1, rep, cut = 6.0, 5.0, 20.0
    A = [[1.5, 0.015], [0.4, 0.075]]  # asymmetric on purpose: A[0][1] != A[1][0]
    pot = CellSpecificInteractionLangevin(cell_radius=R1, attraction_rr=0, attraction_bb=0, attraction_rb=0,
                                          repulsion=rep, cutoff=cut, attraction_matrix=A)
    R = 2 * R1
    u = np.array([0.48, -0.6, 0.64])  # unit vector
    assert math.isclose(np.linalg.norm(u), 1.0, rel_tol=1e-15)

    def f(r, a=0, b=0):
        st, fo, fe = O.pair_force(pot, r * u, [0.0, 0.0, 0.0], a, b)
        assert st == 0 and np.array_equal(fe, -fo)
        return fo

    # contact: F(r) = -rep (R - r) u for every species pair
    for r in (0.5, 3.0, 7.5, 11.0):
        for a, b in ((0, 0), (0, 1), (1, 1)):
            fo = f(r, a, b)
            assert np.allclose(fo, -rep * (R - r) * u, rtol=1e-14, atol=1e-14)
            assert np.dot(fo, u) < 0
    slope = (np.dot(f(4.0), u) - np.dot(f(9.0), u)) / 5.0
    assert math.isclose(slope, -rep, rel_tol=1e-13)
    assert np.all(np.abs(f(R)) < 1e-15)
    # continuity at R and at the cutoff
    assert np.abs(f(R + 1e-9)).max() < 1e-8 and np.abs(f(R - 1e-9)).max() < 1e-8
    assert np.abs(f(cut - 1e-9)).max() < 1e-9
    # tail: log(F c / ((r - R)(c - r)) / a) is -(r / R - 1)^2
    h = 0.5
    for a, b in ((0, 0), (0, 1), (1, 0), (1, 1)):
        rs = 13.0 + h * np.arange(5)
        g = np.array([np.dot(f(r, a, b), u) * cut / ((r - R) * (cut - r)) for r in rs])
        assert np.all(g > 0)  # the tail term pushes the own agent along +u: the (-force, force) sign
        lg = np.log(g)
        d2 = lg[2:] - 2 * lg[1:-1] + lg[:-2]
        assert np.allclose(d2, -2 * h * h / R ** 2, rtol=1e-9)
        assert np.allclose(lg, np.log(A[a][b]) - (rs / R - 1) ** 2, rtol=0, atol=1e-13)
    # beyond the cutoff and at zero distance
    assert np.all(f(cut + 1e-9) == 0)
    assert np.all(f(0.0) == 0)


# ---------------------------------------------------------------------------
# Rods
# ---------------------------------------------------------------------------
def test_nearest_point_doctest_and_unit_test():
    """interaction.rs:478-496 (doctest) and 821-851 (test_closest_points)."""
    d, x, t = O.nearest_point([0.0, 0.0], [1.0, 0.0], [-1.0, 2.0])
    assert abs(d - 1 / math.sqrt(2)) < 1e-10 and abs(t - 0.25) < 1e-10
    assert abs(x[0] - 0.5) < 1e-10 and abs(x[1] - 0.5) < 1e-10
    p1, p2 = [0.0, 0.0], [2.0, 0.0]
    for q, r, dist in (([0.5, 1.0], [0.5, 0.0], 1.0), ([-1.0, 2.0], p1, math.sqrt(5)),
                       ([3.0, -2.0], p2, math.sqrt(5))):
        d, x, _ = O.nearest_point(q, p1, p2)
        assert d == dist and list(x) == r


def test_rod_internal_forces():
    mech = RodMechanics(n_vertices=5, diffusion_constant=0.0, spring_tension=10.0, rigidity=2.0,
                        spring_length=0.5, damping=1.5)
    straight = np.array([[i * 0.5, 0.0] for i in range(5)])
    zero = np.zeros_like(straight)
    dx, dv = O.rod_increment(mech, straight, zero, zero)
    assert np.abs(dv).max() < 1e-14  # at rest length and straight: no internal force
    # stretched spring: equal and opposite, magnitude k (l' - l) / 2 on each end
    two = RodMechanics(n_vertices=2, diffusion_constant=0.0, spring_tension=10.0, rigidity=2.0,
                       spring_length=0.5, damping=0.0)
    dx, dv = O.rod_increment(two, np.array([[0.0, 0.0], [0.7, 0.0]]), np.zeros((2, 2)), np.zeros((2, 2)))
    assert np.allclose(dv, [[10 * 0.2 / 2, 0.0], [-10 * 0.2 / 2, 0.0]], atol=1e-13)
    # single bend of angle alpha at the middle vertex of three: F = 2 eta tan(alpha/2),
    # -F/2, +F, -F/2 along the bisector; the total is zero
    three = RodMechanics(n_vertices=3, diffusion_constant=0.0, spring_tension=0.0, rigidity=2.0,
                         spring_length=0.5, damping=0.0)
    alpha = 0.3
    pos = np.array([[-0.5, 0.0], [0.0, 0.0], [0.5 * math.cos(alpha), 0.5 * math.sin(alpha)]])
    dx, dv = O.rod_increment(three, pos, np.zeros((3, 2)), np.zeros((3, 2)))
    assert np.allclose(dv.sum(axis=0), 0.0, atol=1e-14)
    assert math.isclose(np.linalg.norm(dv[1]), 2 * 2.0 * math.tan(alpha / 2), rel_tol=1e-12)
    assert np.allclose(dv[0], dv[2], atol=1e-14) and np.allclose(dv[0], -0.5 * dv[1], atol=1e-14)
    # damping: -gamma v, and dx = v
    v = np.random.default_rng(1).normal(size=(5, 2))
    dx, dv = O.rod_increment(mech, straight, v, zero)
    assert np.array_equal(dx, v) and np.allclose(dv, -1.5 * v, atol=1e-14)


def test_rod_interaction_symmetry():
    pot = MorsePotential(radius=0.5, potential_stiffness=2.0, cutoff=1.5, strength=1.0)
    a = np.array([[i * 0.5, 0.0] for i in range(8)])
    b = a + np.array([0.3, 0.9])
    st, fo, fe = O.rod_pair_force(pot, a, b)
    assert st == 0
    # parallel offset rods: own forces point away from the other rod (repulsive y < 0)
    assert fo[:, 1].sum() < 0
    # every vertex-segment evaluation returns (f, -f) split over a segment's two ends:
    # total force on own equals minus total on ext
    assert np.allclose(fo.sum(axis=0), -fe.sum(axis=0), atol=1e-13)


# ---------------------------------------------------------------------------
# Engine: weighting, integrators, ring semantics
# ---------------------------------------------------------------------------
def _two_body(dist, weighting="reference", n_sub=1, axis_len=40.0, rng_range=10.0):
    d = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [axis_len, axis_len], rng_range)
    pot = MorsePotential(radius=2.0, potential_stiffness=1.0, cutoff=9.0, strength=1.0)
    s = O.OracleSim([InstanceSpec(d, pot, NewtonDamped(1.0, 1.0))], dt=0.01, weighting=weighting,
                    n_subdomains=n_sub)
    return s, pot


def test_pair_weighting_same_adjacent_and_cross_subdomain():
    """update_mechanics.rs:178-220: a same-voxel pair gets the full force, an adjacent-voxel
    pair inside one subdomain half of it, a pair across subdomains the full force."""
    s, pot = _two_body(3.0)
    full = O.pair_force(pot, [5.0, 5.0], [8.0, 5.0])[1]
    s.set_agents(np.array([[5.0, 5.0], [8.0, 5.0]]))  # same voxel (dx = 10)
    f = s.compute_forces()
    assert np.allclose(f[0], full, rtol=1e-15)
    s.set_agents(np.array([[8.0, 5.0], [11.0, 5.0]]))  # adjacent voxels
    f = s.compute_forces()
    half = O.pair_force(pot, [8.0, 5.0], [11.0, 5.0])[1] * 0.5
    assert np.allclose(f[0], half, rtol=1e-15) and np.allclose(f[1], -half, rtol=1e-15)
    sp, _ = _two_body(3.0, weighting="physical")
    sp.set_agents(np.array([[8.0, 5.0], [11.0, 5.0]]))
    assert np.allclose(sp.compute_forces()[0], 2 * half, rtol=1e-15)
    # 4 x 4 voxels split into two subdomains along the lexicographic order: rows 0-1 | 2-3
    s2, _ = _two_body(3.0, n_sub=2)
    s2.set_agents(np.array([[18.0, 5.0], [21.0, 5.0]]))
    assert np.allclose(s2.compute_forces()[0], 2 * half, rtol=1e-15)


@pytest.mark.parametrize("weighting,k", [("reference", 1), ("reference", 2), ("reference", 7), ("physical", 1)])
def test_forces_sum_to_zero_every_weighting(weighting, k):
    """The part of the pair weighting that the mathematics fixes: with an interaction
    symmetric under exchange of the agents, each pair delivers opposite forces to its
    two agents under every weighting (same voxel: both halves to both; adjacent voxel
    in one subdomain: half of its own evaluation to each; across subdomains: both
    halves to both), so the forces sum to zero.  Delivering a half to one agent only
    leaves |F| / 2 per such pair."""
    wl_dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [24.0] * 3, 2.5)
    pot = SpeciesMorsePotential(radius=1.0, potential_stiffness=1.0, cutoff=2.5, strength=0.5,
                                species_matrix=[[1.0, 0.2], [0.2, 1.0]])
    rng = np.random.default_rng(17)
    pos = rng.uniform(1e-6, 24.0, (2500, 3))
    sp = (rng.random(2500) < 0.5).astype(np.int32)
    s = O.OracleSim([InstanceSpec(wl_dom, pot, NewtonDamped(1.0, 1.0))], dt=0.01, weighting=weighting,
                    n_subdomains=k)
    s.set_agents(pos, species=sp)
    f = s.compute_forces()
    scale = np.abs(f).sum(axis=0)
    assert np.all(scale > 0)
    assert np.all(np.abs(f.sum(axis=0)) <= 1e-12 * scale), (f.sum(axis=0), scale)


def test_euler_damped_closed_form():
    """No force, damping g: Euler gives v_k = v0 (1 - g dt)^k exactly in real arithmetic."""
    d = CartesianCuboid.from_boundaries_and_interaction_range([-1e4, -1e4], [1e4, 1e4], 100.0)
    g, dt = 0.7, 0.01
    s = O.OracleSim([InstanceSpec(d, NoInteraction(), NewtonDamped(g, 1.0))], dt=dt, mechanics_solver_order=0)
    s.set_agents(np.array([[0.0, 0.0]]), np.array([[3.0, -2.0]]))
    s.step(100)
    p, v = s.get_state()
    assert np.allclose(v[0], np.array([3.0, -2.0]) * (1 - g * dt) ** 100, rtol=1e-12)
    # x_k = x0 + dt v0 sum_{j<k} (1 - g dt)^j
    geo = (1 - (1 - g * dt) ** 100) / (g * dt)
    assert np.allclose(p[0], dt * np.array([3.0, -2.0]) * geo, rtol=1e-12)


def test_adams_bashforth_history_semantics():
    """Damped free motion (dx = v, dv = -g v), stepped by a recursion written from the
    solver description rather than from the oracle: chili pushes the new increment
    before reading the ring (solvers.rs:178-186), so the first step of order 1 and 2
    is AB2 on [dx, dx] (Euler), order 1 stays Euler, and order 2 afterwards is AB3 on
    [dx_k, first, second] with the ring iterated oldest first ([dx_{k-1}, dx_k]) or
    newest first ([dx_k, dx_{k-1}]).  A textbook AB3 or a history read before the push
    fails this test."""
    d = CartesianCuboid.from_boundaries_and_interaction_range([-1e4] * 3, [1e4] * 3, 100.0)
    g, dt, v0 = 0.5, 0.01, 2.0
    for order, hist in ((0, "oldest_first"), (1, "oldest_first"), (2, "oldest_first"), (2, "newest_first")):
        s = O.OracleSim([InstanceSpec(d, NoInteraction(), NewtonDamped(g, 1.0))], dt=dt,
                        mechanics_solver_order=order, history_order=hist)
        s.set_agents(np.array([[0.0, 0.0, 0.0]]), np.array([[v0, 0.0, 0.0]]))
        xs, vs, dxh, dvh = [0.0], [v0], [], []
        for k in range(30):
            s.step(1)
            p, v = s.get_state()
            dx, dv = vs[-1], -g * vs[-1]
            dxh.append(dx)
            dvh.append(dv)
            if order <= 1 or len(dxh) == 1:
                cx, cv = dx, dv
            elif hist == "oldest_first":
                cx = 23 / 12 * dx - 16 / 12 * dxh[-2] + 5 / 12 * dx
                cv = 23 / 12 * dv - 16 / 12 * dvh[-2] + 5 / 12 * dv
            else:
                cx = 23 / 12 * dx - 16 / 12 * dx + 5 / 12 * dxh[-2]
                cv = 23 / 12 * dv - 16 / 12 * dv + 5 / 12 * dvh[-2]
            xs.append(xs[-1] + dt * cx)
            vs.append(vs[-1] + dt * cv)
            assert math.isclose(p[0][0], xs[-1], rel_tol=1e-13), (order, hist, k)
            assert math.isclose(v[0][0], vs[-1], rel_tol=1e-13), (order, hist, k)


def test_ring_doctest_order():
    """aux_storage.rs:859-877: push 1, push 3 -> iteration [1, 3] (oldest first).
    Free drift pushes dx = v = 1 then dx = 1; a damped run pushes distinct values."""
    d = CartesianCuboid.from_boundaries_and_interaction_range([-1e4] * 2, [1e4] * 2, 100.0)
    s = O.OracleSim([InstanceSpec(d, NoInteraction(), NewtonDamped(0.5, 1.0))], dt=1.0,
                    mechanics_solver_order=2)
    s.set_agents(np.array([[0.0, 0.0]]), np.array([[1.0, 0.0]]))
    s.step(1)
    dx, dv, n = s.get_history()
    assert n == 1 and dx[0][0][0] == 1.0
    v1 = s.get_state()[1][0][0]
    s.step(1)
    dx, dv, n = s.get_history()
    assert n == 2 and dx[0][0][0] == 1.0 and dx[1][0][0] == v1 and v1 != 1.0
    s.step(1)
    dx, dv, n = s.get_history()
    assert n == 2 and dx[0][0][0] == v1


def test_threaded_baseline_matches_structure():
    """k subdomains reproduce the single-subdomain run up to the cross-subdomain weights:
    with a potential whose range never crosses a subdomain face the runs agree."""
    d = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [100.0, 100.0], 10.0)
    pot = MorsePotential(radius=0.5, potential_stiffness=1.0, cutoff=1.0, strength=0.1)
    rng = np.random.default_rng(8)
    # subdomains are lexicographic chunks, i.e. slabs along the first axis; agents sit
    # in the middle of slabs 1 and 6 only, so no interacting pair crosses a slab face
    pos = np.stack([rng.choice([14.0, 16.0, 64.0, 66.0], 400), rng.uniform(1, 99, 400)], axis=1)
    pos[:, 0] += rng.uniform(-0.4, 0.4, 400)
    runs = []
    for k in (1, 2, 5):
        s = O.OracleSim([InstanceSpec(d, pot, NewtonDamped(1.0, 1.0))], dt=0.01, n_subdomains=k)
        s.set_agents(pos)
        s.step(10)
        runs.append(s.get_state()[0])
    assert np.allclose(runs[0], runs[1], rtol=0, atol=1e-12)
    assert np.allclose(runs[0], runs[2], rtol=0, atol=1e-12)


def test_langevin_msd_oracle():
    """brownian_diffusion_constant_approx.rs:260-283 on the oracle with Philox noise."""
    L, n, dt, steps, D, damping = 200.0, 160, 1e-3, 200, 1.0, 10.0
    from cellular_raza_amd import Langevin

    dom = CartesianCuboid.from_boundaries_and_n_voxels([0.0] * 3, [L] * 3, [3] * 3)
    kT = D * damping
    s = O.OracleSim([InstanceSpec(dom, NoInteraction(), Langevin(1.0, damping, kT))], dt=dt, noise="philox",
                    seed=1)
    x0 = np.full(3, L / 2)
    s.set_agents(np.tile(x0, (n, 1)))
    s.step(steps)
    t = steps * dt
    expected = -3 * kT / damping ** 2 * (1 - math.exp(-damping * t)) * (3 - math.exp(-damping * t)) \
        + 6 * kT * t / damping
    d2 = ((s.get_state()[0] - x0) ** 2).sum(axis=1)
    assert abs(expected - d2.mean()) <= 3.5 * d2.std()


# ---------------------------------------------------------------------------
# Neighbour sensing (bacterial_rods/src/main.rs:48-78 accumulator): brute force
# over all pairs, with radii inside the voxel stencil, and the strict inequality
# ---------------------------------------------------------------------------
def test_neighbor_counts_brute_force_points():
    from scipy.spatial.distance import cdist

    from cellular_raza_amd import workloads

    wl = workloads.c2_sorting_3d(n=3000, seed=3)
    o = O.OracleSim(wl.instances, dt=wl.dt)
    o.set_agents(wl.pos, wl.vel, wl.species)
    d = cdist(wl.pos, wl.pos)
    np.fill_diagonal(d, np.inf)
    for radius in (0.8, 1.6, 2.5):  # voxel edge >= 2.5: every pair within radius is in the stencil
        assert np.array_equal(o.neighbor_counts(radius), (d < radius).sum(1)), radius
    o.close()


def test_neighbor_counts_brute_force_rods():
    from scipy.spatial.distance import cdist

    from cellular_raza_amd import RodInteraction, workloads

    rng = np.random.default_rng(4)
    n, nv = 300, 6
    pos = workloads.straight_rods(rng, n, 2, 30.0, nv, 0.5)
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [30.0, 30.0], 4.0)
    spec = [InstanceSpec(dom, RodInteraction(MorsePotential(radius=0.5, potential_stiffness=2.0, cutoff=1.5,
                                                            strength=0.1)),
                         RodMechanics(n_vertices=nv, diffusion_constant=0.0, spring_tension=10.0, rigidity=2.0,
                                      spring_length=0.5, damping=1.5))]
    o = O.OracleSim(spec, dt=0.002)
    o.set_agents(pos)
    flat = pos.reshape(-1, 2)
    rod = np.repeat(np.arange(n), nv)
    d = cdist(flat, flat)
    d[rod[:, None] == rod[None, :]] = np.inf
    for radius in (0.4, 1.0):  # rod half length 1.25 + radius <= voxel edge 4: stencil covers
        want = (d < radius).reshape(n, nv, n * nv).sum((1, 2))
        assert np.array_equal(o.neighbor_counts(radius), want), radius
    o.close()


def test_neighbor_counts_strict_inequality():
    """Points at exact distance 1: norm < radius excludes radius 1, includes the next double."""
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0, 0.0], [8.0, 8.0], 2.0)
    spec = [InstanceSpec(dom, NoInteraction(), NewtonDamped(damping_constant=1.0))]
    o = O.OracleSim(spec, dt=0.01, mechanics_solver_order=1)
    o.set_agents(np.array([[1.5, 1.5], [2.5, 1.5], [2.5, 2.5]]))
    assert o.neighbor_counts(1.0).tolist() == [0, 0, 0]
    assert o.neighbor_counts(math.nextafter(1.0, 2.0)).tolist() == [1, 2, 1]
    assert o.neighbor_counts(math.nextafter(math.sqrt(2.0), 2.0)).tolist() == [2, 2, 2]
    o.close()


# ---------------------------------------------------------------------------
# DomainForce (SubDomainForce): the reference's spring test, and the rod gel
# force against its closed form
# ---------------------------------------------------------------------------
def test_domain_force_spring_single_particle():
    """cellular_raza/tests/domain_force_spring.rs:71-112: one NewtonDamped2D particle
    (mass 0.5, no damping) on a spring of strength 0.0032 along x, dt 0.001 for 10
    time units, default solver order: x(t) within 1e-3 of x0 cos(omega t) at every
    save point (every 0.1)."""
    from cellular_raza_amd import DomainSpring

    k, m, dt, x0 = 0.0032, 0.5, 0.001, 10.0
    dom = CartesianCuboid.from_boundaries_and_n_voxels([-77.0, -77.0], [77.0, 77.0], [3, 3])
    spec = [InstanceSpec(dom, NoInteraction(), NewtonDamped(damping_constant=0.0, mass=m),
                         DomainSpring(strength=[k, 0.0]))]
    o = O.OracleSim(spec, dt=dt, mechanics_solver_order=2)
    o.set_agents(np.array([[x0, 0.0]]), np.zeros((1, 2)))
    omega = math.sqrt(k / m)
    for it in range(100, 10001, 100):
        o.step(100)
        assert abs(o.get_state()[0][0, 0] - x0 * math.cos(omega * it * dt)) < 1e-3, it
    assert o.get_state()[0][0, 1] == 0.0
    o.close()


def test_domain_force_rod_gel_closed_form():
    """CartesianSubDomainRods (bacterial_rods.rs:447-477): every vertex gets
    (0, 0, -gel); a moving vertex within the friction distance of a face also gets
    -(v/|v|) gel friction; a resting one does not."""
    from cellular_raza_amd import RodGelForce, RodInteraction

    gel, fric, dist = 0.7, 0.3, 0.5
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [20.0] * 3, 5.0)
    spec = [InstanceSpec(dom, RodInteraction(NoInteraction()),
                         RodMechanics(n_vertices=3, diffusion_constant=0.0, spring_tension=0.0, rigidity=0.0,
                                      spring_length=1.0, damping=1.0),
                         RodGelForce(gel_pressure=gel, surface_friction=fric, surface_friction_distance=dist))]
    pos = np.array([[[0.3, 10.0, 10.0], [1.3, 10.0, 10.0], [2.3, 10.0, 10.0]],      # vertex 0 near x = 0
                    [[10.0, 10.0, 19.8], [11.0, 10.0, 19.8], [12.0, 10.0, 19.8]]])  # all near z = 20
    vel = np.array([[[3.0, 4.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]],
                    [[0.0, 0.0, 0.0], [0.0, 2.0, 0.0], [1.0, 1.0, 1.0]]])
    o = O.OracleSim(spec, dt=0.01, mechanics_solver_order=1)
    o.set_agents(pos.reshape(2, -1), vel.reshape(2, -1))
    f = o.compute_forces().reshape(2, 3, 3)
    want = np.zeros((2, 3, 3))
    want[..., 2] = -gel
    near = [[True, False, False], [True, True, True]]
    for a in range(2):
        for v in range(3):
            q = np.linalg.norm(vel[a, v])
            if near[a][v] and q != 0:
                want[a, v] += -(vel[a, v] / q) * gel * fric
    assert np.allclose(f, want, rtol=0, atol=1e-15)
    o.close()


# ---------------------------------------------------------------------------
# Cycle of the rods example: RodMechanics::divide's doctest, the identifier
# bookkeeping of update_cell_cycle_4, the growth-rate reaction
# ---------------------------------------------------------------------------
def _rod_cycle_sim(pos, spring_length, cycle, L=40.0, irange=5.0, sp=None):
    from cellular_raza_amd import RodInteraction

    n, nv, dim = pos.shape
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * dim, [L] * dim, irange)
    spec = [InstanceSpec(dom, RodInteraction(NoInteraction()),
                         RodMechanics(n_vertices=nv, diffusion_constant=0.0, spring_tension=0.1, rigidity=0.05,
                                      spring_length=spring_length, damping=0.0))]
    o = O.OracleSim(spec, dt=0.1, mechanics_solver_order=2)
    o.set_agents(pos.reshape(n, -1))
    o.set_rod_cycle(cycle)
    return o


def test_rod_divide_doctest():
    """bacterial_rods.rs:560-585: 7 vertices 0.5 apart, radius 0.25: the mother's last
    vertex and the daughter's first are 2 radius apart.  Closed form: the mother keeps
    v * l from the start, the daughter v * l from the end, l = (3 - 2 r) / 6 / 2; both
    spring lengths (1/2 - r / (7 * 0.5)) * 0.5; identifiers Division(voxel, 2) and
    Division(voxel, 3) (the voxel counted its initial rod), parent Initial(0)."""
    from cellular_raza_amd import RodGrowthCycle

    pos = np.zeros((1, 7, 2))
    pos[0, :, 0] = 10.0 + 0.5 * np.arange(7)
    pos[0, :, 1] = 10.0
    cyc = RodGrowthCycle(growth_base_rate=0.0, division_threshold=0.4, division_radius=0.25, sensing_radius=0.5)
    o = _rod_cycle_sim(pos, 0.5, cyc)
    o.step(1)
    assert o.n_agents == 2
    p = o.get_state()[0].reshape(2, 7, 2)
    l = (3.0 - 0.5) / 6 / 2
    assert np.allclose(p[0, :, 0], 10.0 + l * np.arange(7), rtol=0, atol=1e-12)
    assert np.allclose(p[1, :, 0], 13.0 - l * np.arange(7)[::-1], rtol=0, atol=1e-12)
    assert abs(np.linalg.norm(p[0, 6] - p[1, 0]) - 0.5) < 1e-3
    st = o.cycle_state()
    assert np.allclose(st["spring_length"], (0.5 - 0.25 / 3.5) * 0.5, rtol=0, atol=1e-15)
    vox = st["identifier"][0, 1]
    assert st["identifier"].tolist() == [[1, vox, 2], [1, vox, 3]]
    assert st["parent"].tolist() == [[0, 0, 0], [0, 0, 0]]
    o.close()


def test_rod_growth_reacts_to_neighbours():
    """react_to_neighbors (bacterial_rods/src/main.rs:66-73): no neighbour -> base rate;
    n vertex pairs within 2 R -> ((base (8 - n)) / 8).max(0); the spring length grows
    by rate * dt."""
    from cellular_raza_amd import RodGrowthCycle

    base, dt = 0.2, 0.1
    pos = np.zeros((3, 3, 2))
    pos[:, :, 0] = 0.5 * np.arange(3)
    pos[0] += [5.0, 5.0]
    pos[1] += [5.0, 5.4]    # parallel at 0.4: 3 vertex pairs within 0.5 for each rod
    pos[2] += [20.0, 20.0]  # alone
    cyc = RodGrowthCycle(growth_base_rate=base, division_threshold=10.0, division_radius=0.25, sensing_radius=0.5)
    o = _rod_cycle_sim(pos, 0.5, cyc)
    o.step(1)
    st = o.cycle_state()
    want = np.array([(base * (8.0 - 3.0)) / 8.0, (base * (8.0 - 3.0)) / 8.0, base])
    assert np.array_equal(st["growth_rate"], want)
    assert np.array_equal(st["spring_length"], 0.5 + want * dt)
    o.close()


# ---------------------------------------------------------------------------
# CycleEvent::Remove (update_cycle.rs:41-55)
# ---------------------------------------------------------------------------
def _sorting_sim(n, seed, order=2):
    rng = np.random.default_rng(seed)
    L = 30.0
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [L] * 3, 3.0)
    pot = MorsePotential(radius=1.0, potential_stiffness=1.5, cutoff=3.0, strength=0.5)
    spec = [InstanceSpec(dom, pot, NewtonDamped(damping_constant=1.0, mass=1.0))]
    pos = rng.uniform(1e-9, L, size=(n, 3))
    o = O.OracleSim(spec, dt=0.01, mechanics_solver_order=order)
    o.set_agents(pos, rng.normal(scale=0.1, size=(n, 3)))
    return o, spec, rng


def test_remove_keeps_voxel_order_state_and_history():
    """The removed agents leave their voxel's list and the others keep their order
    (Vec::retain), state and history; agent indices are compacted in order.  The
    expected lists are the old ones filtered and renumbered in numpy."""
    o, _, rng = _sorting_sim(600, 21)
    o.step(7)
    vox0, start0, order0 = o.get_voxels()
    p0, v0 = o.get_state()
    dx0, dv0 = o.get_history()[:2]
    rm = rng.random(600) < 0.3
    new_index = np.cumsum(~rm) - 1
    o.remove_agents(rm)
    assert o.n_agents == int((~rm).sum())
    vox1, start1, order1 = o.get_voxels()
    want_order = new_index[order0[~rm[order0]]]
    assert np.array_equal(order1, want_order)
    counts0 = np.diff(start0.astype(np.int64))
    removed_per_voxel = np.bincount(vox0[rm], minlength=len(counts0))
    assert np.array_equal(np.diff(start1.astype(np.int64)), counts0 - removed_per_voxel)
    assert np.array_equal(vox1, vox0[~rm])
    p1, v1 = o.get_state()
    assert np.array_equal(p1, p0[~rm]) and np.array_equal(v1, v0[~rm])
    dx1, dv1 = o.get_history()[:2]
    assert np.array_equal(dx1, dx0[:, ~rm]) and np.array_equal(dv1, dv0[:, ~rm])
    o.remove_agents(np.zeros(o.n_agents, dtype=bool))  # no-op
    assert o.n_agents == int((~rm).sum())
    with pytest.raises(O.OracleError):
        o.remove_agents(np.ones(o.n_agents, dtype=bool))
    o.close()


def test_remove_isolated_agent_leaves_the_others_unchanged():
    """An agent beyond every cutoff exerts no force: removing it (AB2, with history)
    leaves the other agents' trajectories bitwise equal to the run that keeps it,
    while removing an interacting agent changes them."""
    rng = np.random.default_rng(22)
    pos = rng.uniform(1e-9, 30.0, size=(300, 3))
    pos[0] = [0.5, 0.5, 0.5]  # a corner, more than 3 cutoffs from everyone
    pos = pos[(np.linalg.norm(pos - pos[0], axis=1) > 9.0) | (np.arange(300) == 0)]
    vel = np.zeros_like(pos)
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [30.0] * 3, 3.0)
    pot = MorsePotential(radius=1.0, potential_stiffness=1.5, cutoff=3.0, strength=0.5)
    spec = [InstanceSpec(dom, pot, NewtonDamped(damping_constant=1.0, mass=1.0))]

    def run(index):
        o = O.OracleSim(spec, dt=0.01, mechanics_solver_order=2)
        o.set_agents(pos, vel)
        o.step(5)
        if index is not None:
            rm = np.zeros(len(pos), dtype=bool)
            rm[index] = True
            o.remove_agents(rm)
        o.step(10)
        p = o.get_state()[0]
        o.close()
        return p

    base = run(None)
    assert np.array_equal(run(0), base[1:])
    d = np.linalg.norm(pos[:, None] - pos[None], axis=2) + np.eye(len(pos)) * 1e9
    j = int(np.argmin(d.min(axis=1)))
    assert d.min() < 3.0
    assert not np.array_equal(run(j), np.delete(base, j, axis=0))


def test_rebinning_keeps_stayers_and_appends_movers():
    """sort_cells_in_voxels_step_1/2 (update_mechanics.rs:418-513) with one
    subdomain: a voxel's new list is its stayers in their old order followed by the
    arrivals in order of old voxel, then old position -- computed here from the
    lists before and the voxels after each step."""
    rng = np.random.default_rng(24)
    L = 30.0
    dom = CartesianCuboid.from_boundaries_and_interaction_range([0.0] * 3, [L] * 3, 3.0)
    pot = MorsePotential(radius=1.0, potential_stiffness=1.5, cutoff=3.0, strength=0.5)
    spec = [InstanceSpec(dom, pot, NewtonDamped(damping_constant=0.1, mass=1.0))]
    o = O.OracleSim(spec, dt=0.02, mechanics_solver_order=1)
    n = 3000
    o.set_agents(rng.uniform(1.0, L - 1.0, size=(n, 3)), rng.normal(scale=8.0, size=(n, 3)))
    moved = 0
    for _ in range(4):
        vox0, start0, order0 = o.get_voxels()
        o.step(1)
        vox1, start1, order1 = o.get_voxels()
        pos_in_old = np.empty(n, dtype=np.int64)
        pos_in_old[order0] = np.arange(n)
        for w in np.flatnonzero(np.diff(start1.astype(np.int64))):
            old = order0[start0[w]:start0[w + 1]]
            stay = [a for a in old if vox1[a] == w]
            arrivals = sorted((a for a in np.flatnonzero((vox1 == w) & (vox0 != w))), key=lambda a: pos_in_old[a])
            assert list(order1[start1[w]:start1[w + 1]]) == stay + arrivals, w
            moved += len(arrivals)
    assert moved > 100
    o.close()
