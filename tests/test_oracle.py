"""CPU tests: pin the oracle (golden vectors from the reference core, the
compiled reference itself, closed forms, independent root finding, schedule
independence, invariants) and check the host-side pieces."""
import hashlib
import math
import os

import numpy as np
import pytest

import oracle_ctypes as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_fsm_vectors.npz")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64 if a.dtype == np.float64 else np.uint32)


# ---------------------------------------------------------------- golden vectors
@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


@pytest.mark.parametrize("name", ["a", "b", "c", "d"])
def test_oracle_matches_reference_golden_vectors(gold, name):
    h = tuple(gold[f"{name}_h"])
    inp = gold[f"{name}_input"]
    for sn in range(1, 9):
        a = inp.copy()
        O.sweep_ref64(a, sn, h)
        assert np.array_equal(bits(a), bits(gold[f"{name}_sn{sn}"])), sn
    a = inp.copy()
    for sn in range(1, 9):
        O.sweep_ref64(a, sn, h)
    assert np.array_equal(bits(a), bits(gold[f"{name}_pass1"]))
    a = inp.copy()
    passes, _ = O.solve_ref64(a, h, tol=1e-12, max_passes=200)
    assert passes == int(gold[f"{name}_passes"])
    assert np.array_equal(bits(a), bits(gold[f"{name}_converged"]))


def test_c0_point_source_golden(gold):
    a = O.gen_point(64, 64, 64, 12345)
    passes, md = O.solve_ref64(a, tol=1e-12)
    assert passes == int(gold["c0_passes"])
    assert hashlib.sha256(a.tobytes()).digest() == gold["c0_sha256"].tobytes()
    assert np.array_equal(a.reshape(-1)[gold["c0_sample_idx"]], gold["c0_sample_val"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference absent)")
@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_compiled_reference(seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = (int(v) for v in rng.integers(2, 14, 3))
    h = tuple(float(v) for v in rng.choice([0.5, 1.0, 1.7], 3))
    a = O.random_field(X, Y, Z, 1000 + seed)
    for sn in range(1, 9):
        b, c = a.copy(), a.copy()
        O.sweep_ref64(b, sn, h)
        O.ref_run_sn(c, sn, h)
        assert np.array_equal(bits(b), bits(c)), (X, Y, Z, h, sn)


# ---------------------------------------------------------------- Godunov solves
def quad_root(ms, ds):
    """Largest root of sum(((u - m_i)/d_i)^2) = 1 by numpy.roots (independent
    of the closed forms under test)."""
    c2 = sum(1.0 / d ** 2 for d in ds)
    c1 = sum(-2.0 * m / d ** 2 for m, d in zip(ms, ds))
    c0 = sum(m * m / d ** 2 for m, d in zip(ms, ds)) - 1.0
    r = np.roots([c2, c1, c0])
    return float(np.max(r.real))


def godunov_brute(mx, my, mz, dx, dy, dz):
    """Godunov upwind value by trying the 1-, 2-, 3-axis solutions in order of
    the sorted minima and keeping the first consistent one."""
    pairs = sorted([(mx, dx), (my, dy), (mz, dz)], key=lambda p: p[0])
    for k in (1, 2, 3):
        ms = [p[0] for p in pairs[:k]]
        ds = [p[1] for p in pairs[:k]]
        if any(math.isinf(m) for m in ms):
            break
        u = ms[0] + ds[0] if k == 1 else quad_root(ms, ds)
        if k == 3 or u <= pairs[k][0]:
            return u
    return math.inf


@pytest.mark.parametrize("fn", ["ref", "d64", "f32"])
def test_godunov_closed_forms(fn):
    g = {"ref": O.lib().orc_godunov_ref, "d64": O.lib().orc_godunov_d64,
         "f32": O.lib().orc_godunov_f32}[fn]
    tol = 1e-12 if fn != "f32" else 2e-6
    inf = math.inf
    for m in (0.0, 3.0, 250.0):
        for hh in (1.0, 0.5):
            assert abs(g(1e9, m, inf, inf, hh, hh, hh) - (m + hh)) <= tol * max(1, m)
            assert abs(g(1e9, m, m, inf, hh, hh, hh) - (m + hh / math.sqrt(2))) <= tol * max(1, m)
            assert abs(g(1e9, m, m, m, hh, hh, hh) - (m + hh / math.sqrt(3))) <= tol * max(1, m)
    # all neighbours infinite -> cur; cur below the candidate wins
    assert g(7.0, inf, inf, inf, 1, 1, 1) == 7.0
    assert g(0.25, 2.0, 3.0, 4.0, 1, 1, 1) == 0.25


@pytest.mark.parametrize("fn", ["ref", "d64", "f32"])
def test_godunov_general_roots_vs_numpy(fn):
    """Random anisotropic cases against an independent root finder: a
    dropped term, swapped spacing or wrong sign fails here."""
    g = {"ref": O.lib().orc_godunov_ref, "d64": O.lib().orc_godunov_d64,
         "f32": O.lib().orc_godunov_f32}[fn]
    rng = np.random.default_rng(5)
    for it in range(3000):
        base = float(rng.uniform(0, 300))
        m = [base + float(v) for v in rng.uniform(0, 2.0, 3)]
        if it % 5 == 0:
            m[int(rng.integers(3))] = math.inf
        d = [float(v) for v in rng.uniform(0.3, 2.0, 3)]
        if it % 2:
            d = [d[0]] * 3
        want = godunov_brute(m[0], m[1], m[2], *d)
        got = g(1e9, m[0], m[1], m[2], *d)
        tol = (1e-9 if fn != "f32" else 4e-5) * max(1.0, base)
        assert abs(got - want) <= tol, (m, d, got, want)


@pytest.mark.parametrize("fn", ["d64", "f32"])
def test_godunov_iso_2d_3d_boundary(fn):
    """The isotropic difference form picks the 3-D root when
    2h^2 - a^2 > (2b - a)^2 (a = m1 - m0, b = m2 - m0), i.e. when the 2-D
    root exceeds m2, and takes one square root. Checked where the choice is
    closest: m2 just below, at and just above the 2-D root, against the
    independent root finder; a wrong test (e.g. 2b + a, or b instead of
    2b - a) picks the other root there and fails."""
    g = {"d64": O.lib().orc_godunov_d64, "f32": O.lib().orc_godunov_f32}[fn]
    rng = np.random.default_rng(17)
    for it in range(2000):
        h = float(rng.choice([1.0, 0.5, 1.5]))
        m0 = float(rng.uniform(0, 200))
        a = float(rng.uniform(0, 0.99 * h))
        u2 = m0 + 0.5 * (a + math.sqrt(2 * h * h - a * a))
        for rel in (-1e-3, -1e-6, 0.0, 1e-6, 1e-3, 0.05):
            m2 = u2 + rel * h
            if m2 < m0 + a:
                continue
            m = [m0, m0 + a, m2]
            rng.shuffle(m)
            want = godunov_brute(m[0], m[1], m[2], h, h, h)
            got = g(1e9, m[0], m[1], m[2], h, h, h)
            tol = (1e-9 if fn == "d64" else 4e-5) * max(1.0, m0)
            assert abs(got - want) <= tol, (m, h, got, want)
    # away from the boundary the roots differ by far more than the
    # tolerance: for a = 0.6 h the 2-D root is m0 + 0.94 h; with
    # m2 = m0 + 1.4 h it is the answer, with m2 = m0 + 0.75 h the 3-D root
    # (lower, and above m2) is
    for h in (1.0, 0.5):
        m0, a = 10.0, 0.6 * h
        u2 = m0 + 0.5 * (a + math.sqrt(2 * h * h - a * a))
        got2 = g(1e9, m0, m0 + a, m0 + 1.4 * h, h, h, h)
        assert abs(got2 - u2) <= 1e-5
        got3 = g(1e9, m0, m0 + a, m0 + 0.75 * h, h, h, h)
        assert abs(got3 - godunov_brute(m0, m0 + a, m0 + 0.75 * h, h, h, h)) <= 1e-5
        assert m0 + 0.75 * h < got3 < u2 - 1e-3 * h


def test_godunov_invariants():
    L = O.lib()
    rng = np.random.default_rng(9)
    for _ in range(2000):
        m = rng.uniform(0, 5, 3)
        d = rng.uniform(0.5, 1.5, 3)
        u = L.orc_godunov_d64(1e9, *m, *d)
        # shift invariance G(m + c) = G(m) + c
        c = float(rng.uniform(-3, 100))
        assert abs(L.orc_godunov_d64(1e9, *(m + c), *d) - (u + c)) < 1e-10
        # monotone and non-expansive: raising one minimum never lowers G and
        # raises it by at most that amount
        k = int(rng.integers(3))
        e = float(rng.uniform(0, 0.5))
        m2 = m.copy()
        m2[k] += e
        u2 = L.orc_godunov_d64(1e9, *m2, *d)
        assert u - 1e-12 <= u2 <= u + e + 1e-12
        # result never above the current value, and above the smallest minimum
        assert L.orc_godunov_d64(2.0, *m, *d) <= 2.0
        assert u > min(m)


# ---------------------------------------------------------------- sweeps
@pytest.mark.parametrize("sn", range(1, 9))
def test_hyperplane_equals_lexicographic_gauss_seidel(sn):
    """Schedule independence (the hyperplane sweep reads only level L-1 new
    and L+1 old values), bitwise, anisotropic spacing."""
    for seed, (X, Y, Z) in enumerate([(2, 2, 2), (5, 3, 4), (9, 7, 6), (16, 4, 11)]):
        a = O.random_field(X, Y, Z, 50 + seed + 10 * sn)
        b = a.copy()
        O.sweep_ref64(a, sn, (0.8, 1.1, 1.3))
        O.sweep_lex64(b, sn, (0.8, 1.1, 1.3))
        assert np.array_equal(bits(a), bits(b))


def test_ordering_directions_match_reference_table():
    """A single source at the centre of a 1-D line along each axis: ordering
    sn reaches exactly the cells downstream of the source in its direction
    ((x,y,z) signs: 1 +++, 2 -+-, 3 ++-, 4 -++, 5 ---, 6 +-+, 7 --+, 8 +--),
    plus the one upstream neighbour, which reads the source as an old value."""
    table = {1: (1, 1, 1), 2: (-1, 1, -1), 3: (1, 1, -1), 4: (-1, 1, 1),
             5: (-1, -1, -1), 6: (1, -1, 1), 7: (-1, -1, 1), 8: (1, -1, -1)}
    for sn, signs in table.items():
        for axis in range(3):
            dims = [1, 1, 1]
            dims[axis] = 9
            X, Y, Z = dims
            a = O.padded(X, Y, Z)
            a[1:-1, 1:-1, 1:-1] = 1e6
            c = [1, 1, 1]
            c[axis] = 5
            a[c[2], c[1], c[0]] = 0.0
            O.sweep_ref64(a, sn)
            line = O.interior(a).reshape(-1)
            s = signs[axis]
            for p in range(9):
                off = (p + 1) - 5
                want = abs(off) if (abs(off) <= 1 or np.sign(off) == s) else 1e6
                assert line[p] == want, (sn, axis, p)


def test_point_source_axis_exactness_and_lipschitz():
    a = O.gen_point(40, 36, 30, 77)
    O.solve_ref64(a)
    u = O.interior(a)
    k, j, i = np.argwhere(u == 0)[0][::-1]
    for n in range(1, 8):
        for (dx, dy, dz) in ((n, 0, 0), (-n, 0, 0), (0, n, 0), (0, -n, 0), (0, 0, n), (0, 0, -n)):
            x, y, z = k + dx, j + dy, i + dz
            if 0 <= x < 40 and 0 <= y < 36 and 0 <= z < 30:
                assert u[z, y, x] == float(n)
    for ax in range(3):
        assert np.abs(np.diff(u, axis=ax)).max() <= 1.0 + 1e-12
    # converged field is a fixed point of v = min(init, G(v))
    init = O.gen_point(40, 36, 30, 77)
    assert O.lib().orc_residual_f64(a, init, 40, 36, 30, 1.0, 1.0, 1.0) < 1e-9


def test_solution_is_distance_like():
    """Two-source field: the converged value is within the known first-order
    Godunov error of the Euclidean distance and never below it along axes."""
    X = Y = Z = 24
    a = O.padded(X, Y, Z)
    a[1:-1, 1:-1, 1:-1] = 1e6
    srcs = [(4, 5, 6), (18, 17, 15)]
    for (x, y, z) in srcs:
        a[z + 1, y + 1, x + 1] = 0.0
    O.solve_ref64(a)
    u = O.interior(a)
    zz, yy, xx = np.mgrid[0:Z, 0:Y, 0:X]
    eu = np.min([np.sqrt((xx - x) ** 2 + (yy - y) ** 2 + (zz - z) ** 2) for (x, y, z) in srcs],
                axis=0)
    # first-order upwind: over-estimates away from the sources, may dip a
    # little below where the two fronts meet
    assert np.min(u - eu) > -0.05
    assert np.max(u - eu) < 0.1 * np.max(eu)
    # the sources' own axes are exact
    assert u[6, 5, 4 + 3] == 3.0 and u[15, 17 - 2, 18] == 2.0


# ---------------------------------------------------------------- FP32 + judge
def test_f32_oracle_close_to_f64_and_judge():
    a = O.gen_spheres(48, 40, 36, 3, n=6, r=(2.0, 6.0))
    f = a.astype(np.float32)
    init32 = f.copy()
    O.solve_f32(f)
    b = a.copy()
    O.solve_d64(b, tol=0.0)
    err = np.abs(O.interior(f).astype(np.float64) - O.interior(b))
    assert err.max() < 5e-5
    assert O.judge_f32(f, init32) == 0
    # a region lowered below the truth is flagged (at least at its border)
    bad = f.copy()
    bad[10:16, 10:16, 10:16] -= 0.75
    n, fl = O.judge_f32(bad, init32, want_flags=True)
    assert n > 0 and fl[10:16, 10:16, 10:16].any()
    # a single raised point is flagged exactly there
    bad = f.copy()
    bad[20, 21, 22] += 0.5
    n, fl = O.judge_f32(bad, init32, want_flags=True)
    assert fl[20, 21, 22] == 1


# ---------------------------------------------------------------- pre/post
def test_update_distance_known_answers():
    loc = np.array([-1, 0, 0, 0, 1, 1, 0, 0], np.int32)
    d = np.array([0.0, 0.0, np.nan, np.inf, 0.0, 2.5, -3.0, 7.0])
    O.lib().orc_update_distance(loc, d, d.size)
    assert d[0] == 0.0  # ghost location untouched
    assert d[1] == 90000 and d[2] == 90000 and d[4] == 90000
    assert d[3] == np.inf and d[5] == 2.5 and d[6] == -3.0 and d[7] == 7.0


def test_set_negative_inside_known_answers():
    loc = np.array([-1, 1, 1, 0, 1], np.int32)
    d = np.array([5.0, 3.0, np.inf, 4.0, 0.5])
    O.lib().orc_set_negative_inside(loc, d, d.size)
    assert list(d) == [5.0, -1.0, np.inf, 4.0, -1.0]


def test_adjust_boundary_is_edge_padding():
    X, Y, Z = 5, 4, 3
    a = O.padded(X, Y, Z)
    a[1:-1, 1:-1, 1:-1] = np.arange(X * Y * Z, dtype=np.float64).reshape(Z, Y, X)
    O.lib().orc_adjust_boundary(a, X, Y, Z)
    want = np.pad(O.interior(a), 1, mode="edge")
    assert np.array_equal(a, want)


# ---------------------------------------------------------------- generators
def splitmix64(s):
    s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = s
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return s, z ^ (z >> 31)


def u01_stream(seed):
    s = seed
    while True:
        s, z = splitmix64(s)
        yield (z >> 11) * 2.0 ** -53


def test_generators_independent_recomputation():
    X, Y, Z, seed = 30, 26, 22, 8
    # spheres
    g = u01_stream(seed)
    sph = []
    for _ in range(64):
        cx, cy, cz = next(g) * (X - 1), next(g) * (Y - 1), next(g) * (Z - 1)
        r = 4 + 28 * next(g)
        sph.append((cx, cy, cz, r))
    zz, yy, xx = np.mgrid[0:Z, 0:Y, 0:X].astype(np.float64)
    s = np.min([np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2) - r
                for cx, cy, cz, r in sph], axis=0)
    want = np.where(s <= 0, 0.0, np.where(s <= 1.5, s, 1e6))
    got = O.interior(O.gen_spheres(X, Y, Z, seed))
    assert np.allclose(got, want, rtol=0, atol=1e-12)
    # terrain
    g = u01_stream(seed)
    hills = [(5 + 35 * next(g), 10 + 50 * next(g), next(g) * (X - 1), next(g) * (Y - 1))
             for _ in range(8)]
    yy2, xx2 = np.mgrid[0:Y, 0:X].astype(np.float64)
    H = sum(A * np.exp(-((xx2 - cx) ** 2 + (yy2 - cy) ** 2) / (2 * S * S)) for A, S, cx, cy in hills)
    off = zz - H[None]
    want = np.where(off <= 0, 0.0, np.where(off <= 1.5, off, 1e6))
    got = O.interior(O.gen_terrain(X, Y, Z, seed))
    assert np.allclose(got, want, rtol=0, atol=1e-9)
    # point
    g = u01_stream(seed)
    k, j, i = 1 + int(next(g) * X), 1 + int(next(g) * Y), 1 + int(next(g) * Z)
    p = O.gen_point(X, Y, Z, seed)
    assert p[i, j, k] == 0.0 and np.count_nonzero(O.interior(p) == 0) == 1
    assert np.all(np.isinf(p[0])) and np.all(np.isinf(p[:, :, -1]))


def test_one_unchanged_ordering_is_the_fixed_point():
    """The device skips a field's orderings after its first ordering that
    changes nothing. Pinned here on the FP32 oracle sweep: after such an
    ordering, two more full passes (every direction) change no bit, on
    terrain, sphere, point-source and anisotropic fields."""
    cases = [(O.gen_terrain(20, 17, 13, seed=3, amp=(1.0, 6.0), sig=(2.0, 9.0)), (1.0, 1.0, 1.0)),
             (O.gen_spheres(19, 14, 16, seed=4, n=4, r=(1.5, 4.0)), (1.0, 1.0, 1.0)),
             (O.gen_point(15, 18, 12, seed=5), (1.0, 1.0, 1.0)),
             (O.gen_point(16, 11, 14, seed=6), (0.5, 1.25, 0.75))]
    for a0, h in cases:
        a = a0.astype(np.float32)
        for k in range(200):
            prev = a.copy()
            O.sweep_f32(a, k % 8 + 1, h)
            if np.array_equal(prev.view(np.int32), a.view(np.int32)):
                break
        else:
            raise AssertionError("no unchanged ordering")
        assert k > 0  # the first ordering changed something: not a trivial field
        fixed = a.copy()
        for k2 in range(k + 1, k + 17):
            O.sweep_f32(a, k2 % 8 + 1, h)
        assert np.array_equal(fixed.view(np.int32), a.view(np.int32))


def test_speed_sweep_unit_and_scaled_slowness():
    """Slowness 1 is the unit-speed sweep bit for bit; slowness 2 is the
    unit-speed sweep with doubled spacings bit for bit (dx * 2 is exact)."""
    rng = np.random.default_rng(8)
    for X, Y, Z, h in ((9, 7, 6, (1.0, 1.0, 1.0)), (8, 10, 5, (0.5, 0.75, 1.25))):
        a = O.gen_spheres(X, Y, Z, seed=int(rng.integers(100)), n=3, r=(1.0, 2.5))
        for sn in range(1, 9):
            b, c = a.copy(), a.copy()
            O.sweep_ref64_speed(b, np.ones_like(a), sn, h)
            O.sweep_ref64(c, sn, h)
            assert np.array_equal(b.view(np.uint64), c.view(np.uint64))
            b, c = a.copy(), a.copy()
            O.sweep_ref64_speed(b, np.full_like(a, 2.0), sn, h)
            O.sweep_ref64(c, sn, tuple(2 * v for v in h))
            assert np.array_equal(b.view(np.uint64), c.view(np.uint64))


def test_speed_layered_medium_closed_form():
    """A source plane z = 1 in a medium layered in z: the travel time is the
    sum of slowness x dz over the layers crossed, u(z) = u(z-1) + dz s(z),
    exactly (every update is the 1-D one)."""
    X, Y, Z, dz = 6, 5, 12, 0.5
    a = O.padded(X, Y, Z)
    a[1:-1, 1:-1, 1:-1] = 1e6
    a[1, 1:-1, 1:-1] = 0.0
    s = np.ones_like(a)
    layers = np.random.default_rng(3).uniform(0.25, 4.0, size=Z)
    for k in range(Z):
        s[k + 1] = layers[k]
    O.solve_ref64_speed(a, s, (1.0, 1.0, dz), tol=0.0)
    want = [0.0]
    for k in range(1, Z):
        want.append(want[-1] + dz * layers[k])
    for k in range(Z):
        assert np.all(a[k + 1, 1:-1, 1:-1] == want[k]), k


def test_speed_monotone_in_slowness():
    """Comparison principle: a pointwise larger slowness gives a pointwise
    larger (or equal) converged travel time."""
    rng = np.random.default_rng(4)
    a = O.gen_point(14, 11, 9, seed=9)
    s1 = rng.uniform(0.5, 2.0, size=a.shape)
    s2 = s1 * rng.uniform(1.0, 1.5, size=a.shape)
    u1, u2 = a.copy(), a.copy()
    O.solve_ref64_speed(u1, s1, tol=0.0)
    O.solve_ref64_speed(u2, s2, tol=0.0)
    inner = (slice(1, -1),) * 3
    assert np.all(u1[inner] <= u2[inner])
    assert np.any(u1[inner] < u2[inner])


# ---------------------------------------------------------------- FP64 polish
@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.7, 1.3)])
def test_polish_limit_is_the_fp64_solution(h):
    """The polish (u = min(init, G64(u)), which may raise values) started
    from the FP32 solution converges to the monotone FP64 difference-form
    solve's limit (a different iteration of the same fixed-point equation)."""
    a = O.gen_spheres(30, 26, 22, seed=91, n=5, r=(2.0, 5.0))
    u64 = a.copy()
    O.solve_d64(u64, h=h, tol=0.0, max_passes=200)
    v = a.astype(np.float32)
    O.solve_f32(v, h=h)
    w = v.astype(np.float64)
    passes, _ = O.polish64(w, a, h=h, tol=0.0, max_passes=200)
    assert passes < 200
    # it stopped on an ordering that changed nothing: the fixed point
    again = w.copy()
    _, md = O.polish64(again, a, h=h, tol=0.0, max_passes=1)
    assert md == 0.0 and np.array_equal(bits(again), bits(w))
    diff = np.abs(O.interior(w) - O.interior(u64))
    assert diff.max() <= 1e-12 * max(1.0, float(np.abs(O.interior(u64)).max()))


def test_polish_raises_a_planted_region_where_min_sweeps_cannot():
    """A box planted 0.5 h below the FP64 solution comes back to it under the
    polish; the monotone min(u, G(u)) difference-form sweeps leave it low
    (which a planted monotone update in the polish would do too)."""
    a = O.gen_spheres(34, 30, 26, seed=92, n=4, r=(2.0, 5.0))
    u64 = a.copy()
    O.solve_d64(u64, tol=0.0, max_passes=200)
    bad = u64.copy()
    bad[8:14, 9:15, 10:18] -= 0.5
    pol = bad.copy()
    O.polish64(pol, a, tol=0.0, max_passes=200)
    assert np.abs(O.interior(pol) - O.interior(u64)).max() <= 1e-9
    mono = bad.copy()
    O.solve_d64(mono, tol=0.0, max_passes=200)
    assert np.abs(O.interior(mono) - O.interior(u64)).max() >= 0.4


def test_polish_point_source_axis_exact():
    """Along the axis lines through a point source the polished FP32 field
    holds n h exactly (1-D updates from the exact source)."""
    X, Y, Z = 40, 36, 30
    a = O.gen_point(X, Y, Z, seed=93)
    i, j, k = (int(v) for v in np.argwhere(a == 0.0)[0])
    want, _, pp = O.solve_f32_polished(a)
    assert pp >= 1
    for kk in range(1, X + 1):
        assert want[i, j, kk] == abs(kk - k)
    for jj in range(1, Y + 1):
        assert want[i, jj, k] == abs(jj - j)
    for ii in range(1, Z + 1):
        assert want[ii, j, k] == abs(ii - i)


def test_polish_cuts_the_fp32_storage_error():
    """On a field with distances near 100 h the FP32 fixed point carries the
    storage rounding accumulated along characteristics; the polish brings it
    to the 1e-5 h polish tolerance and below the FP32 error."""
    a = O.gen_spheres(96, 96, 96, seed=94, n=3, r=(2.0, 4.0))
    u64 = a.copy()
    O.solve_d64(u64, tol=1e-12, max_passes=200)
    raw = a.astype(np.float32)
    O.solve_f32(raw)
    pol, _, pp = O.solve_f32_polished(a)
    e_raw = np.abs(O.interior(raw).astype(np.float64) - O.interior(u64)).max()
    e_pol = np.abs(O.interior(pol).astype(np.float64) - O.interior(u64)).max()
    print(f"FP32 {e_raw:.3e} h, polished {e_pol:.3e} h ({pp} polish passes)")
    assert e_pol <= 1e-5 and e_pol < e_raw


# ------------------------------------------- spatially varying speed, FP32 path
def _smooth_slowness(shape, seed, lo=0.5, hi=2.0):
    """Padded slowness: a few random plane waves mapped into [lo, hi]."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    f = np.zeros(shape)
    for _ in range(4):
        k = rng.uniform(-0.4, 0.4, size=3)
        f += np.sin(k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 6.3))
    f = (f - f.min()) / (f.max() - f.min())
    return lo + (hi - lo) * f


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.75, 1.25)])
def test_speed_f32_scaled_slowness_is_scaled_spacing(h):
    """FP32: slowness 2 (and 0.5) everywhere is the unit-speed sweep with
    doubled (halved) spacings bit for bit: the spacings d s and the weights
    1 / (d s)^2 = (1 / d^2) / s^2 are exact power-of-two scalings. A weight
    scaled by 1/s instead of 1/s^2, or the spacing left unscaled, fails."""
    a = O.gen_spheres(10, 8, 7, seed=12, n=3, r=(1.0, 2.5)).astype(np.float32)
    for sc in (2.0, 0.5):
        s = np.full(a.shape, sc, np.float32)
        for sn in range(1, 9):
            b, c = a.copy(), a.copy()
            O.sweep_f32_speed(b, s, sn, h)
            O.sweep_f32(c, sn, tuple(sc * v for v in h))
            assert np.array_equal(b.view(np.uint32), c.view(np.uint32)), (sc, sn)
        b = a.copy()
        O.sweep_f32_speed(b, np.ones_like(a), 1, h)
        c = O.sweep_f32(a.copy(), 1, h)
        assert np.array_equal(b.view(np.uint32), c.view(np.uint32))


@pytest.mark.parametrize("axis", [0, 2])
def test_speed_f32_layered_medium_exact_recursion(axis):
    """A source plane in a medium layered along one axis, with anisotropic
    spacings: every update is the 1-D one, so the FP32 result is exactly the
    recursion u_k = u_{k-1} + fl(d_axis s_k) in FP32 along that axis (the
    other axes' spacings must not enter)."""
    X, Y, Z = 7, 6, 11
    h = (0.75, 1.0, 0.5)
    a = O.padded(X, Y, Z, dtype=np.float32)
    a[1:-1, 1:-1, 1:-1] = 1e6
    s = np.ones(a.shape, np.float32)
    n = (Z, Y, X)[axis]
    layers = np.random.default_rng(5 + axis).uniform(0.25, 4.0, size=n).astype(np.float32)
    d = np.float32(h[2 - axis])
    if axis == 0:
        a[1, 1:-1, 1:-1] = 0.0
        for k in range(n):
            s[k + 1] = layers[k]
    else:
        a[1:-1, 1:-1, 1] = 0.0
        for k in range(n):
            s[:, :, k + 1] = layers[k]
    O.solve_f32_speed(a, s, h)
    want = [np.float32(0.0)]
    for k in range(1, n):
        want.append(np.float32(want[-1] + np.float32(d * layers[k])))
    for k in range(n):
        got = a[k + 1, 1:-1, 1:-1] if axis == 0 else a[1:-1, 1:-1, k + 1]
        assert np.all(got == want[k]), k


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.6, 0.8, 1.1)])
def test_speed_d64_difference_form_matches_reference_form(h):
    """The FP64 difference form with scaled spacings converges to the
    reference-form speed solve (orc_solve_ref64_speed, itself pinned by the
    unit / doubled / layered checks above) to 1e-10 relative: the same
    equation written two ways. A wrong per-point scaling of the weights
    moves the 2-D / 3-D roots by far more."""
    a = O.gen_spheres(18, 15, 13, seed=31, n=4, r=(1.5, 3.5))
    s = _smooth_slowness(a.shape, 2)
    r64 = a.copy()
    O.solve_ref64_speed(r64, s, h, tol=0.0, max_passes=200)
    d64 = a.copy()
    O.solve_d64_speed(d64, s, h, tol=0.0, max_passes=200)
    inner = (slice(1, -1),) * 3
    ref, got = r64[inner], d64[inner]
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(got))
    assert np.max(np.abs(got[fin] - ref[fin]) / np.maximum(ref[fin], min(h))) < 1e-10
    # the slowness matters: the unit-speed solve is far away
    u1 = a.copy()
    O.solve_ref64(u1, h, tol=0.0)
    assert np.max(np.abs(u1[inner][fin] - ref[fin])) > 1.0


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.6, 0.8, 1.1)])
def test_speed_polish_limit_and_f32_accuracy(h):
    """The FP64 speed polish from the FP32 speed solution converges to the
    FP64 speed solution (reference form) to 1e-9 h; the FP32 speed solve
    alone is within 1e-4 h of it on this small field."""
    # an init exact in FP32, so that both solves start from the same values
    a = O.gen_spheres(20, 17, 15, seed=44, n=4, r=(1.5, 3.5)).astype(np.float32).astype(np.float64)
    s32 = _smooth_slowness(a.shape, 6).astype(np.float32)
    s64 = s32.astype(np.float64)
    r64 = a.copy()
    O.solve_ref64_speed(r64, s64, h, tol=0.0, max_passes=200)
    u = a.astype(np.float32)
    O.solve_f32_speed(u, s32, h)
    inner = (slice(1, -1),) * 3
    fin = np.isfinite(r64[inner])
    err32 = np.max(np.abs(u[inner][fin].astype(np.float64) - r64[inner][fin]))
    assert err32 < 1e-4 * min(h)
    v = u.astype(np.float64)
    init = a.astype(np.float32).astype(np.float64)
    p, _ = O.polish64_speed(v, init, s32, h, tol=0.0, max_passes=100)
    assert p < 100
    assert np.max(np.abs(v[inner][fin] - r64[inner][fin])) < 1e-9 * min(h)
    # unit-slowness polish of the same field lands elsewhere
    w = u.astype(np.float64)
    O.polish64_speed(w, init, None, h, tol=0.0, max_passes=100)
    assert np.max(np.abs(w[inner][fin] - r64[inner][fin])) > 0.5


def test_speed_judge_uses_the_slowness():
    """The speed judge passes the FP64 speed solution rounded to FP32, flags
    a value planted 1e-3 below it, and the unit-speed judge rejects the same
    solution (the slowness enters the judged equation)."""
    h = (1.0, 1.0, 1.0)
    a = O.gen_spheres(16, 14, 12, seed=52, n=3, r=(1.5, 3.0))
    s32 = _smooth_slowness(a.shape, 9).astype(np.float32)
    r64 = a.copy()
    O.solve_ref64_speed(r64, s32.astype(np.float64), h, tol=0.0, max_passes=200)
    v = r64.astype(np.float32)
    init = a.astype(np.float32)
    assert O.judge_f32_speed(v, init, s32, h) == 0
    assert O.judge_f32_speed(v, init, None, h) > 100
    w = v.copy()
    w[7, 6, 9] -= np.float32(1e-3)
    assert O.judge_f32_speed(w, init, s32, h) >= 1


# ------------------------------------------------------- fixed-point storage
FXI = 0xFFFFFFFF


def test_fx_bits_range_bound():
    """k = 32 - ceil(log2(X + Y + Z)): the largest converged offset, at most
    (X + Y + Z - 3) h, stays below 2^32 - 1 quanta, and one more bit would
    not (the quantum is the finest that always fits)."""
    for X, Y, Z in [(1, 1, 1), (5, 4, 3), (64, 64, 64), (256, 256, 128), (512, 512, 512),
                    (812, 812, 812), (1000, 1, 1), (31, 1, 1), (32, 1, 1)]:
        k = O.fx_bits(X, Y, Z)
        s = X + Y + Z
        assert (s - 3) * 2 ** k < FXI
        assert s * 2 ** (k + 1) > 2 ** 32
    assert O.fx_bits(512, 512, 512) == 21 and O.fx_bits(256, 256, 128) == 22


@pytest.mark.parametrize("k", [21, 24, 30])
def test_godunov_fx_closed_forms(k):
    """1-D updates are exact (m + 2^k); equal minima in 2-D / 3-D give
    m + h/sqrt2, m + h/sqrt3 (h = 2^k quanta) to half a quantum plus the FP32
    rounding of the root offset (the nearest quantum up to k = 23); +inf
    neighbours and saturation."""
    g = O.lib().orc_godunov_fx
    H = 2 ** k
    for m in (0, 12345, 3 * 2 ** 29):
        if m + 2 * H >= FXI:
            continue
        assert g(FXI, m, FXI, FXI, k) == m + H
        for n in (2, 3):
            t = H / math.sqrt(n)
            got = g(FXI, m, m, FXI if n == 2 else m, k) - m
            assert abs(got - t) <= 0.5 + 2 * t * 2.0 ** -24, (k, n, got, t)
            if k <= 23:
                assert got == round(t)
        # cur below the candidate wins
        assert g(m, m, m, m, k) == m
    assert g(77, FXI, FXI, FXI, k) == 77          # no finite neighbour
    assert g(FXI, FXI - 5, FXI, FXI, k) == FXI    # saturates at +inf


@pytest.mark.parametrize("k", [21, 24])
def test_godunov_fx_general_roots_vs_numpy(k):
    """Random sorted minima within 2 h: the fixed-point update is the
    independent root (numpy.roots) to 1.5 quanta plus FP32 rounding of the
    root offset; a dropped term or wrong branch rule is off by many quanta."""
    g = O.lib().orc_godunov_fx
    H = 2 ** k
    rng = np.random.default_rng(17)
    for it in range(4000):
        base = int(rng.integers(0, 2 ** 31))
        off = [int(v) for v in rng.uniform(0, 2.0 * H, 3)]
        m = [base + o for o in off]
        if it % 7 == 0:
            m[int(rng.integers(3))] = FXI
        got = g(FXI, m[0], m[1], m[2], k)
        fin = [(v - base) / H if v != FXI else math.inf for v in m]
        want = godunov_brute(fin[0], fin[1], fin[2], 1.0, 1.0, 1.0) * H + base
        t = want - min(x for x in m if x != FXI)
        assert abs(got - want) <= 1.5 + 4 * t * 2.0 ** -24, (m, got, want)


def test_fx_encode_decode_round_trip():
    rng = np.random.default_rng(2)
    v = rng.uniform(-40, 900, 5000).astype(np.float32)
    v[:5] = [np.inf, np.nan, 1e6, -40.0, 0.0]
    k = O.fx_bits(512, 512, 512)
    base = float(np.nanmin(v[np.isfinite(v)]))
    assert float(O.lib().orc_fx_min(np.ascontiguousarray(v), v.size)) == base
    d = 2.0 ** -k
    q = O.fx_encode(v, base, d)
    assert q[0] == FXI and q[1] == FXI and q[2] == FXI   # +inf, NaN, beyond the range
    r = np.rint((v.astype(np.float64) - base) / d)
    ok = np.isfinite(r) & (r < FXI)
    assert np.array_equal(q[ok], r[ok].astype(np.uint32))
    back = O.fx_decode(q, base, d)
    assert np.all(np.isinf(back[~ok]))
    assert np.all(np.abs(back[ok].astype(np.float64) - v[ok]) <=
                  d / 2 + np.abs(v[ok]) * 2.0 ** -24)


@pytest.mark.parametrize("h", [1.0, 0.37])
def test_fixed_solve_close_to_fp64_and_point_axis_exact(h):
    """Converged fixed-point solves are within 5e-5 h of the FP64 difference
    form's solution (a different arithmetic), on spheres and a random field;
    along the axes through a point source the values are n h exactly."""
    for a in (O.gen_spheres(70, 53, 41, 9, h=h, n=6, r=(2.0, 8.0)),
              O.random_field(48, 40, 36, 3)):
        if h != 1.0:
            a = a.copy()
            a[1:-1, 1:-1, 1:-1] *= h
        start = np.ascontiguousarray(O.interior(a)).astype(np.float32)
        ref = start.astype(np.float64)
        pad = np.full((ref.shape[0] + 2, ref.shape[1] + 2, ref.shape[2] + 2), np.inf)
        pad[1:-1, 1:-1, 1:-1] = ref
        O.solve_d64(pad, h=(h, h, h), tol=0.0)
        got, p, _, _ = O.solve_fixed(start, h=h)
        err = np.abs(got.astype(np.float64) - O.interior(pad)).max()
        assert err <= 5e-5 * h, err
    X, Y, Z = 37, 29, 23
    src = (11, 7, 5)   # (z, y, x)
    start = np.full((Z, Y, X), 1e6, np.float32)
    start[src] = 0.0
    got, _, (k, base, d), q = O.solve_fixed(start, h=h)
    qi = O.interior(q)
    for n in range(1, X - src[2]):
        assert qi[src[0], src[1], src[2] + n] == n * 2 ** k
    for n in range(1, Z - src[0]):
        assert qi[src[0] + n, src[1], src[2]] == n * 2 ** k


def test_fixed_solve_shift_invariant():
    """Adding a constant to the start field (values on a 2^-8 grid, so the
    shifted FP32 values are exact) adds it to every stored result: the base
    takes it and the quanta are equal."""
    start = np.ascontiguousarray(O.interior(O.gen_spheres(40, 33, 29, 5, n=5, r=(2.0, 6.0))))
    start = (np.round(start * 256.0) / 256.0).astype(np.float32)
    _, _, (k, b0, d), q0 = O.solve_fixed(start)
    _, _, (_, b1, _), q1 = O.solve_fixed(start - np.float32(37.25))
    assert b1 == b0 - 37.25
    assert np.array_equal(q0, q1)


def test_judge_fx_converged_passes_planted_fails():
    start = np.ascontiguousarray(O.interior(O.gen_spheres(30, 26, 22, 8, n=4, r=(2.0, 5.0))))
    start = start.astype(np.float32)
    _, _, (k, base, d), q = O.solve_fixed(start)
    init = np.full(q.shape, np.inf, np.float32)
    init[1:-1, 1:-1, 1:-1] = start
    tol = 2e-6 + 4 * d
    assert O.judge_fx(q, init, 1.0, base, d, tol) == 0
    bad = q.copy()
    bad[10, 10, 10] -= np.uint32(int(tol / d) + 10)   # too low: below its own update
    n, fl = O.judge_fx(bad, init, 1.0, base, d, tol, want_flags=True)
    assert n >= 1 and fl[10, 10, 10] == 1
    # a quantum size that disagrees with the stored values (a wrong d) fails
    assert O.judge_fx(q, init, 1.0, base, 1.01 * d, tol) > 0
