"""Pins of the integrator oracle (oracle/md.c): known-answer vectors for the
RNG, SETTLE against an independent iterative SHAKE, and closed forms of the
leapfrog Langevin scheme (src/CudaLangevinThermostatIntegrator.cu:138-167,
576-682)."""
import numpy as np
import pytest

from apocharmm_amd import systems

O = pytest.importorskip("oracle.oracle")


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32 with 10 rounds
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        assert [int(v) for v in O.philox4x32_10(ctr, key)] == list(want)


def test_langevin_normals_distribution():
    z = np.array([O.langevin_normals(2003, s, a) for s in range(40) for a in range(1000)])
    assert z.shape == (40000, 3)
    assert abs(z.mean()) < 0.02
    assert abs(z.var() - 1.0) < 0.02
    # kurtosis of a normal is 3
    assert abs((z ** 4).mean() / z.var() ** 2 - 3.0) < 0.1
    # counter independence: neighbouring atoms/steps/seeds are uncorrelated
    a = np.array([O.langevin_normals(7, 5, i)[0] for i in range(5000)])
    b = np.array([O.langevin_normals(7, 6, i)[0] for i in range(5000)])
    c = np.array([O.langevin_normals(8, 5, i)[0] for i in range(5000)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.05 and abs(np.corrcoef(a, c)[0, 1]) < 0.05


@pytest.fixture(scope="module")
def water1000():
    sp = systems.config1()
    return np.asarray(sp["waters"], np.int32), np.asarray(sp["xyz"], np.float64), sp


def test_settle_equals_iterative_shake(water1000):
    w, x0, _ = water1000
    rng = np.random.default_rng(5)
    x1 = x0 + rng.normal(0, 0.02, x0.shape)
    xs = O.settle(w, x0, x1.copy())
    xk = O.shake_water(w, x0, x1.copy(), tol=1e-14)
    assert np.abs(xs - xk).max() < 1e-11


def test_settle_geometry_and_momentum(water1000):
    w, x0, _ = water1000
    rng = np.random.default_rng(6)
    x1 = x0 + rng.normal(0, 0.03, x0.shape)
    xs = O.settle(w, x0, x1.copy())
    d = lambda x, i, j: np.linalg.norm(x[w[:, i]] - x[w[:, j]], axis=1)
    geom = O.TIP3P_GEOM
    assert np.abs(d(xs, 0, 1) - geom[2]).max() < 1e-10
    assert np.abs(d(xs, 0, 2) - geom[2]).max() < 1e-10
    assert np.abs(d(xs, 1, 2) - geom[3]).max() < 1e-10
    m = np.array([geom[0], geom[1], geom[1]])
    com = lambda x: (x[w] * m[None, :, None]).sum(1) / m.sum()
    assert np.abs(com(xs) - com(x1)).max() < 1e-12
    # the canonical triangle: ra + rb is the O to H-H midpoint height
    ra, rb, rc = O.settle_geometry()
    assert abs((ra + rb) ** 2 + rc ** 2 - geom[2] ** 2) < 1e-12
    assert abs(ra * geom[0] - 2 * geom[1] * rb) < 1e-12


def test_settle_of_rigid_input_is_identity(water1000):
    w, x0, _ = water1000
    xr = O.settle(w, x0, x0.copy())          # make exactly rigid at the TIP3P geometry
    xs = O.settle(w, xr, xr.copy())
    assert np.abs(xs - xr).max() < 1e-12


def _free_particles(n, mass=12.0):
    invm = np.full(n, 1.0 / mass)
    return invm, np.zeros((0, 3), np.int32)


def test_leapfrog_harmonic_closed_form():
    """gamma = 0: x_{n+1} - 2 x_n + x_{n-1} = -dt^2 k x_n / m, so
    x_n = A cos(n theta) with cos(theta) = 1 - dt^2 omega^2 / 2."""
    invm, w = _free_particles(1, mass=4.0)
    lt = O.Langevin(invm, w, dt_ps=0.002, temperature=0.0, friction_ps=0.0, seed=1)
    k = 30.0
    x = np.array([[1.0, 0.0, 0.0]])
    v = np.zeros((1, 3))
    dprev = lt.init_backstep(k * x, x, v)
    xs = [x[0, 0]]
    for s in range(400):
        lt.step(k * x, x, dprev, v, s)
        xs.append(x[0, 0])
    theta = np.arccos(1.0 - lt.dt ** 2 * k * invm[0] / 2.0)
    # x_1 = x_0 (1 - dt^2 w^2/2) fixes the phase: x_n = cos(n theta)
    ref = np.cos(theta * np.arange(len(xs)))
    assert np.abs(np.array(xs) - ref).max() < 1e-9


def test_init_then_step_returns_initial_velocity():
    """With gamma = 0 the back-step makes the first on-step velocity equal the
    initial one whatever the force; with gamma > 0 and T = 0 it is alpha v."""
    rng = np.random.default_rng(3)
    n = 50
    invm = 1.0 / rng.uniform(1, 20, n)
    w = np.zeros((0, 3), np.int32)
    for gamma_ps, fac in ((0.0, 1.0), (150.0, None)):
        lt = O.Langevin(invm, w, dt_ps=0.002, temperature=0.0, friction_ps=gamma_ps, seed=1)
        x = rng.normal(size=(n, 3))
        v0 = rng.normal(size=(n, 3)) * 0.01
        g = rng.normal(size=(n, 3))
        v = v0.copy()
        dprev = lt.init_backstep(g, x, v)
        lt.step(g, x, dprev, v, 0)
        alpha = (1 - lt.gamma / 2) / (1 + lt.gamma / 2)
        want = v0 * (1.0 if fac else alpha)
        assert np.abs(v - want).max() < 1e-12 * np.abs(v0).max() + 1e-15


def test_langevin_equipartition_free_particles():
    """Free particles: the half-step displacement has variance dt^2 kT/m and
    the on-step velocity kT/m exactly, so both kinetic energies average
    (3/2) N kT."""
    n = 4000
    rng = np.random.default_rng(9)
    invm = 1.0 / rng.uniform(1, 16, n)
    w = np.zeros((0, 3), np.int32)
    T = 300.0
    lt = O.Langevin(invm, w, dt_ps=0.002, temperature=T, friction_ps=100.0, seed=2003)
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    g = np.zeros((n, 3))
    dprev = lt.init_backstep(g, x, v)
    ke = []
    for s in range(300):
        ke.append(lt.step(g, x, dprev, v, s))
    ke = np.array(ke[100:])
    want = 1.5 * n * O.KBOLTZ * T
    assert abs(ke[:, 0].mean() / want - 1) < 0.01
    assert abs(ke[:, 1].mean() / want - 1) < 0.01


def test_remove_com_and_centering():
    rng = np.random.default_rng(4)
    n = 30
    invm = 1.0 / rng.uniform(1, 16, n)
    d = rng.normal(size=(n, 3))
    O.Langevin(invm, np.zeros((0, 3), np.int32), 0.001, 300, 0, 0).remove_com(d)
    assert np.abs(((1 / invm)[:, None] * d).sum(0)).max() < 1e-12
    box = np.array([10.0, 12.0, 14.0])
    gptr = np.array([0, 3, 5, 6], np.int32)
    gatom = np.arange(6, dtype=np.int32)
    x = np.array([[5.5, 0, 0], [5.8, 0, 0], [5.9, 0, 0],      # centroid 5.73 > 5 -> -10
                  [-7.5, 6.5, 0], [-7.0, 6.1, 0],             # x -7.25 -> +10, y 6.3 > 6 -> -12
                  [1.0, -6.9, 7.1]])                          # y -6.9 -> +12, z 7.1 -> -14
    xc = O.center(gptr, gatom, box, x.copy())
    shift = xc - x
    assert np.allclose(shift[:3], [-10, 0, 0])
    assert np.allclose(shift[3:5], [10, -12, 0])
    assert np.allclose(shift[5], [0, 12, -14])
    assert np.array_equal(O.center(gptr, gatom, box, xc.copy()), xc)


def _shake_case(rng, m, stretch=0.08):
    """A heavy atom with m hydrogens: reference positions on bonds of 1.09 A,
    then an unconstrained move of up to `stretch` A per coordinate."""
    ref = np.zeros((m + 1, 3))
    for k in range(m):
        d = rng.normal(size=3)
        ref[k + 1] = ref[0] + 1.09 * d / np.linalg.norm(d)
    ref += rng.normal(size=3)
    new = ref + rng.uniform(-stretch, stretch, size=ref.shape)
    invm = np.array([1 / 12.011] + [1 / 1.008] * m)
    g = np.array([[0] + list(range(1, m + 1)) + [-1] * (3 - m)], np.int32)
    return ref, new, invm, g, np.array([1.09 ** 2])


def test_shake_single_bond_matches_closed_form(oracle):
    """One bond: the iterative solver reaches the root of the quadratic that
    the reference's Shake2Kernel evaluates in closed form
    (src/CudaHolonomicConstraint.cu:324-360)."""
    rng = np.random.default_rng(31)
    for _ in range(20):
        ref, new, invm, g, r2 = _shake_case(rng, 1)
        x = new.copy()
        assert oracle.shake_groups(g, r2, invm, ref, x) >= 0
        p = new[0] - new[1]
        r0 = ref[0] - ref[1]
        rp, rr = p @ r0, r0 @ r0
        lam = (-rp + np.sqrt(rp * rp + rr * (r2[0] - p @ p))) / (rr * (invm[0] + invm[1]))
        np.testing.assert_allclose(x[0], new[0] + invm[0] * lam * r0, rtol=0, atol=1e-12)
        np.testing.assert_allclose(x[1], new[1] - invm[1] * lam * r0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("m", [2, 3])
def test_shake_group_invariants(oracle, m):
    """Two or three bonds on one heavy atom: every bond length is met, the
    total momentum of the correction is zero, and each hydrogen moves only
    along its own reference bond (the heavy atom takes the mass-weighted
    rest), which fixes the solution of the constraint equations."""
    rng = np.random.default_rng(40 + m)
    for _ in range(20):
        ref, new, invm, g, r2 = _shake_case(rng, m)
        x = new.copy()
        assert oracle.shake_groups(g, r2, invm, ref, x) >= 0
        for k in range(1, m + 1):
            assert abs(np.sum((x[0] - x[k]) ** 2) - r2[0]) <= 1e-13 * r2[0]
            dk = x[k] - new[k]
            bond = ref[0] - ref[k]
            assert np.linalg.norm(np.cross(dk, bond)) <= 1e-12 * max(np.linalg.norm(dk), 1e-30)
        mom = ((x - new) / invm[:, None]).sum(0)
        assert np.abs(mom).max() < 1e-12


def test_langevin_step_with_shake_keeps_bond_lengths(oracle):
    """The oracle integrator constrains hydrogen groups with the waters: after
    the initial constraint, a back-step and five steps every SHAKE bond and
    every water geometry holds to FP64 rounding."""
    rng = np.random.default_rng(5)
    ref, _, invm, g, r2 = _shake_case(rng, 3, 0.0)
    w = np.array([[0.0, 0.0, 0.0], [0.9572, 0, 0], [-0.24, 0.9266, 0.0]]) + 4.0
    x = np.vstack([ref, w])
    invmass = np.concatenate([invm, [1 / 15.9994, 1 / 1.008, 1 / 1.008]])
    v = rng.normal(size=x.shape) * 0.01
    lt = oracle.Langevin(invmass, [[4, 5, 6]], 0.002, 300.0, 1.0, 3, shake=(g, r2))
    lt.init_constrain(x, v)
    grad = rng.normal(size=x.shape)
    dprev = lt.init_backstep(grad, x, v)
    for s in range(5):
        lt.step(rng.normal(size=x.shape), x, dprev, v, s)
        for k in range(1, 4):
            assert abs(np.sum((x[0] - x[k]) ** 2) - r2[0]) <= 1e-12 * r2[0]
        assert abs(np.linalg.norm(x[4] - x[5]) - 0.9572) < 1e-10


def test_shake_processes_every_group_when_one_misses_tol(oracle):
    """A tolerance below FP64 rounding makes the first group miss it; the
    solver still constrains every later group (it returns -1 and carries on,
    rather than leaving them as they were)."""
    rng = np.random.default_rng(11)
    cases = [_shake_case(rng, m) for m in (3, 1, 2)]
    n0 = 0
    refs, news, invms, groups = [], [], [], []
    for ref, new, invm, g, r2 in cases:
        m = len(ref) - 1
        refs.append(ref + 40.0)  # box-scale coordinates: rounding ~1e-14 of r^2
        news.append(new + 40.0)
        invms.append(invm)
        groups.append([n0 + k for k in range(m + 1)] + [-1] * (3 - m))
        n0 += m + 1
    ref, new, invm = np.vstack(refs), np.vstack(news), np.concatenate(invms)
    g = np.array(groups, np.int32)
    r2 = np.full(len(g), 1.09 ** 2)
    x = new.copy()
    assert oracle.shake_groups(g, r2, invm, ref, x, tol=1e-17, maxit=50) == -1
    for grp in g:
        for k in range(1, 4):
            if grp[k] >= 0:
                assert abs(np.sum((x[grp[0]] - x[grp[k]]) ** 2) / r2[0] - 1) < 1e-12
