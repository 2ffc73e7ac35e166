"""Pins for the CPU restatement (oracle/wc_oracle.c).  The reference ships no tests or golden
vectors, so each oracle function is checked against what the mathematics fixes: closed forms
evaluated independently in float64, normalisation integrals, invariants, and O(N^2) brute force
on small scenes.  A plausible transcription slip (dropped factor, wrong sign, swapped operand,
wrong neighbourhood order) fails at least one of these."""
import math

import numpy as np
import pytest

from watercube_amd import scenes
from watercube_amd import _lib as L


def _params(preset=L.WC_PRESET_DAM_BREAK_16K, **kw):
    p = L.preset(preset)
    for k, v in kw.items():
        if isinstance(v, (list, tuple)):
            getattr(p, k)[:] = v
        else:
            setattr(p, k, v)
    return p


# ---------------------------------------------------------------- kernel constants and kernels
def test_kernel_constants_closed_form():
    for h in (0.03, 0.016, 0.004, 0.04):
        p = _params(kernel_radius=h)
        L.derive(p)
        hd = float(np.float32(h))
        assert p.poly6_const == np.float32(315.0 / (64.0 * math.pi * hd ** 9))
        assert p.spiky_const == np.float32(-45.0 / (math.pi * hd ** 6))
        assert p.visc_const == np.float32(45.0 / (math.pi * hd ** 6))
        assert p.bin_size == np.float32(np.float32(1.0) / np.float32(p.grid_res))


def test_poly6_is_normalised(oracle):
    """Integral of W over the ball of radius h is 1 (Mueller et al. poly6 normalisation)."""
    p = _params()
    h = p.kernel_radius
    r = np.linspace(0.0, h, 4001)
    w = np.array([oracle.poly6(p, float(x)) for x in r], np.float64)
    integral = np.trapezoid(w * 4.0 * math.pi * r ** 2, r)
    assert abs(integral - 1.0) < 1e-4


def test_poly6_matches_float64(oracle):
    p = _params()
    h = float(np.float32(p.kernel_radius))
    for frac in (0.0, 0.3, 0.77, 0.999):
        r = float(np.float32(frac * h))
        exact = 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3
        # h*h - r*r cancels near r = h: float32 keeps ~ulp(h^2)/(h^2 - r^2) relative accuracy, cubed
        cond = 3 * 6e-8 * h * h / (h * h - r * r)
        assert oracle.poly6(p, r) == pytest.approx(exact, rel=2e-6 + cond)


def test_spiky_is_gradient_of_spiky_kernel(oracle):
    """spikyKernel(r, d) equals grad W_spiky with W_spiky = 15/(pi h^6)(h - d)^3:
    -45/(pi h^6)(h - d)^2 r/d, checked against a finite difference of W_spiky."""
    p = _params()
    h = float(np.float32(p.kernel_radius))
    wsp = lambda d: 15.0 / (math.pi * h ** 6) * (h - d) ** 3
    for d, direction in ((0.2 * h, (1, 0, 0)), (0.6 * h, (0.6, -0.8, 0.0)), (0.9 * h, (0.48, 0.6, 0.64))):
        u = np.array(direction, np.float64)
        u /= np.linalg.norm(u)
        r = (u * d).astype(np.float32)
        g = oracle.spiky(p, r, float(np.linalg.norm(r.astype(np.float64))))
        eps = 1e-7
        dwd = (wsp(d + eps) - wsp(d - eps)) / (2 * eps)
        np.testing.assert_allclose(g, dwd * u, rtol=1e-4, atol=1e-3)


def test_viscosity_laplacian(oracle):
    """Wvis(r) = 45/(pi h^6)(h - r), the Laplacian of Mueller's viscosity kernel."""
    p = _params()
    h = float(np.float32(p.kernel_radius))
    for r in (0.0, 0.5 * h, 0.99 * h):
        assert oracle.wvis(p, r) == pytest.approx(45.0 / (math.pi * h ** 6) * (h - r), rel=2e-6)


@pytest.mark.parametrize("eos", [L.WC_EOS_CUBIC, L.WC_EOS_LINEAR])
def test_equation_of_state(oracle, eos):
    p = _params(eos=eos, rest_pressure=7.0)
    for rho in (500.0, 1000.0, 1234.5):
        x = rho / p.rest_density
        exact = 7.0 + p.stiffness * ((x ** 3 - 1.0) if eos == L.WC_EOS_CUBIC else (rho - p.rest_density))
        assert oracle.eos(p, rho) == pytest.approx(exact, rel=1e-5, abs=1e-3)


# ---------------------------------------------------------------- grid
def test_cell_of_cell_centres_and_clamps(oracle):
    p = _params()
    R, b = p.grid_res, p.bin_size
    for c in [(0, 0, 0), (5, 7, 9), (R - 1, R - 1, R - 1), (R - 1, 0, 3)]:
        pos = [(ci + 0.5) * b for ci in c]
        assert oracle.cell(p, pos) == c
        assert oracle.cell_key(p, pos) == c[2] * R * R + c[1] * R + c[0]
    assert oracle.cell(p, [-0.5, -1e-9, float("nan")]) == (0, 0, 0)
    assert oracle.cell(p, [1.0, 7.0, 1e30]) == (R - 1, R - 1, R - 1)


def test_sort_run_matches_numpy(oracle):
    p, parts, _, _ = scenes.workload("dam_break_16k")
    rng = np.random.default_rng(3)
    parts = parts[rng.permutation(len(parts))]
    st = oracle.OracleState(p, parts)
    st.sort_run()
    R = p.grid_res
    c = np.clip((parts["position"] / np.float32(p.bin_size)).astype(np.int64), 0, R - 1)
    keys = c[:, 2] * R * R + c[:, 1] * R + c[:, 0]
    assert np.array_equal(st.keys[: len(parts)], keys)
    counts = np.bincount(keys, minlength=R ** 3)
    assert np.array_equal(st.counts, counts)
    assert st.counts.sum() == p.num_particles
    assert np.array_equal(st.offsets, np.concatenate([[0], np.cumsum(counts)[:-1]]))
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(st.sorted[: len(parts)], order)
    assert st.buf2.tobytes() == parts[order].tobytes()


def test_linear_scan(oracle):
    rng = np.random.default_rng(1)
    counts = rng.integers(0, 20, 10007).astype(np.uint32)
    got = oracle.linear_scan(counts)
    assert np.array_equal(got, np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32))


# ---------------------------------------------------------------- density
def _single(pos, vel=(0, 0, 0), rho=0.0, pr=0.0):
    a = np.zeros(1, L.PARTICLE_DTYPE)
    a["position"] = pos
    a["velocity"] = vel
    a["density"] = rho
    a["pressure"] = pr
    return a


def test_isolated_particle_density(oracle):
    p = _params(num_particles=1)
    st = oracle.OracleState(p, _single([0.5, 0.5, 0.5]))
    st.sort_run()
    st.density()
    h = float(np.float32(p.kernel_radius))
    exact = p.particle_mass * 315.0 / (64.0 * math.pi * h ** 9) * h ** 6
    assert st.buf2["density"][0] == pytest.approx(exact, rel=1e-6)
    assert st.buf2["pressure"][0] == pytest.approx(p.stiffness * (st.buf2["density"][0] - p.rest_density), rel=1e-6)


def test_pair_density(oracle):
    p = _params(num_particles=2)
    h = float(np.float32(p.kernel_radius))
    d = 0.4 * h
    parts = np.concatenate([_single([0.5, 0.5, 0.5]), _single([0.5 + d, 0.5, 0.5])])
    st = oracle.OracleState(p, parts)
    st.sort_run()
    st.density()
    W = lambda r: 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3
    exact = p.particle_mass * (W(0) + W(d))
    np.testing.assert_allclose(st.buf2["density"], [exact, exact], rtol=2e-6)


def test_wall_density_terms(oracle):
    p = _params(preset=L.WC_PRESET_REFERENCE_DEFAULT)
    h = float(np.float32(p.kernel_radius))
    m = p.particle_mass
    W = lambda r: 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3
    # x near the low wall only
    assert oracle.wall_density(p, [0.3 * h, 0.5, 0.5]) == pytest.approx(4 * m * W(0.3 * h), rel=2e-6)
    # z near the high wall: the reference's z branch tests p.y, so nothing is added unless y is high too
    assert oracle.wall_density(p, [0.5, 0.5, 1.0 - 0.3 * h]) == 0.0
    got = oracle.wall_density(p, [0.5, 1.0 - 0.5 * h, 1.0 - 0.3 * h])
    assert got == pytest.approx(4 * m * (W(0.5 * h) + W(0.3 * h)), rel=2e-6)


def _neighbourhood_index(ci, cj):
    d = np.asarray(cj) - np.asarray(ci)
    if np.any(np.abs(d) > 1):
        return None
    return (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1)


def _brute_force_order(oracle, p, ps, i):
    """All j != i in the 27 cells around i, ordered as the reference enumerates them."""
    ci = oracle.cell(p, ps["position"][i])
    cand = []
    for j in range(len(ps)):
        if j == i:
            continue
        b = _neighbourhood_index(ci, oracle.cell(p, ps["position"][j]))
        if b is not None:
            cand.append((b, j))
    cand.sort()
    return [j for _, j in cand]


def test_density_brute_force_bit_exact(oracle):
    """O(N^2) density in NEIGHBORHOOD order, evaluated with numpy float32 scalars, matches the
    grid-based pass bit for bit on a small jittered block."""
    p, parts = scenes.scaled_block((9, 8, 7), 0.015, seed=11)
    st = oracle.OracleState(p, parts)
    st.sort_run()
    st.density()
    ps = st.buf2
    f = np.float32
    h, m, c6 = f(p.kernel_radius), f(p.particle_mass), f(p.poly6_const)
    for i in range(0, len(ps), 7):
        pi = ps["position"][i]
        a0 = h * h - f(0) * f(0)
        rho = m * (((a0 * a0) * a0) * c6)
        for j in _brute_force_order(oracle, p, ps, i):
            r = pi - ps["position"][j]
            dist = np.sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2], dtype=f)
            if dist >= h:
                continue
            a = h * h - dist * dist
            rho = rho + m * (((a * a) * a) * c6)
        assert ps["density"][i].tobytes() == f(rho).tobytes(), i


def test_lattice_interior_density_near_rest(oracle):
    p, parts, _, _ = scenes.workload("dam_break_16k")
    st = oracle.OracleState(p, parts)
    st.sort_run()
    st.density()
    pos = st.buf2["position"]
    h = p.kernel_radius
    inner = np.all((pos > 0.03 + h) & (pos < np.array([0.48, 0.48, 0.24]) - h), axis=1)
    rho = st.buf2["density"][inner]
    assert inner.sum() > 1000
    assert abs(np.median(rho) / p.rest_density - 1.0) < 0.05


# ---------------------------------------------------------------- update
def test_free_fall_closed_form(oracle):
    p = _params(num_particles=1)
    dt = np.float32(2e-4)
    st = oracle.OracleState(p, _single([0.5, 0.9, 0.5]))
    n = 50
    st.step(float(dt), n)
    g = float(np.float32(p.gravity[1]))
    v = st.buf1["velocity"][0]
    x = st.buf1["position"][0]
    # symplectic Euler: v_n = n g dt, y_n = y0 + g dt^2 n(n+1)/2
    assert v[1] == pytest.approx(n * g * float(dt), rel=1e-5)
    assert x[1] == pytest.approx(0.9 + g * float(dt) ** 2 * n * (n + 1) / 2, rel=1e-6)
    assert v[0] == 0.0 and v[2] == 0.0


def test_wall_clamp_and_damping(oracle):
    p = _params(num_particles=1, gravity=(0.0, 0.0, 0.0))
    dt = 1e-3
    st = oracle.OracleState(p, _single([0.0015, 0.5, 0.9985], vel=(-2.0, 0.0, 2.0)))
    st.step(dt, 1)
    x, v = st.buf1["position"][0], st.buf1["velocity"][0]
    assert x[0] == np.float32(p.border) and v[0] == pytest.approx(2.0 * p.wall_damping, rel=1e-6)
    assert x[2] == np.float32(np.float32(1.0) - np.float32(p.border))
    assert v[2] == pytest.approx(-2.0 * p.wall_damping, rel=1e-6)


def test_speed_clamp(oracle):
    p = _params(num_particles=1, gravity=(0.0, -1e9, 0.0))
    st = oracle.OracleState(p, _single([0.5, 0.5, 0.5]))
    st.step(1e-3, 1)
    assert st.buf1["velocity"][0][1] == -np.float32(p.max_speed)


def test_pair_forces_closed_form(oracle):
    """Two particles at distance d with known density/pressure: pressure force along r with the
    spiky magnitude m P 45/(pi h^6)(h-d)^2, viscosity m (dv/rho) mu 45/(pi h^6)(h-d)."""
    p = _params(num_particles=2, gravity=(0.0, 0.0, 0.0))
    h = float(np.float32(p.kernel_radius))
    d = 0.5 * h
    rho, pr = 1010.0, 2500.0
    parts = np.concatenate([_single([0.5, 0.5, 0.5], vel=(0, 0, 0), rho=rho, pr=pr),
                            _single([0.5 + d, 0.5, 0.5], vel=(0, 0.1, 0), rho=rho, pr=pr)])
    st = oracle.OracleState(p, parts)
    st.sort_run()          # density is not recomputed: update reads the given rho and p
    dt = 1e-5
    st.update(dt)
    m, mu = p.particle_mass, p.viscosity
    P = (pr + pr) / (2 * rho)
    fp = m * P * 45.0 / (math.pi * h ** 6) * (h - d) ** 2   # pushes particle 0 towards -x
    fv = m * (0.1 / rho) * mu * 45.0 / (math.pi * h ** 6) * (h - d)
    v0 = st.buf1["velocity"][0]
    assert v0[0] == pytest.approx(-fp / rho * dt, rel=1e-4)
    assert v0[1] == pytest.approx(fv / rho * dt, rel=1e-4)
    v1 = st.buf1["velocity"][1]
    assert v1[0] == pytest.approx(fp / rho * dt, rel=1e-4)


def test_uniform_velocity_has_no_viscous_force(oracle):
    p, parts = scenes.scaled_block((6, 6, 6), 0.015, seed=2)
    p.gravity[:] = (0.0, 0.0, 0.0)
    p.stiffness = 0.0  # no pressure force
    parts["velocity"] = (0.3, -0.2, 0.1)
    st = oracle.OracleState(p, parts)
    st.step(1e-4, 1)
    np.testing.assert_array_equal(st.buf1["velocity"], np.float32([0.3, -0.2, 0.1]) * np.ones((len(parts), 1), np.float32))


def test_wall_force_quirk(oracle):
    """wallForces passes vec3.length() == 3 as the distance: spiky(r, 3) * 0.01."""
    p = _params(preset=L.WC_PRESET_REFERENCE_DEFAULT)
    h = float(np.float32(p.kernel_radius))
    x = 0.25 * h
    f = oracle.wall_forces(p, [x, 0.5, 0.5])
    exact = (-45.0 / (math.pi * h ** 6)) * (h - 3.0) ** 2 * (-x / 3.0) * 0.01
    assert f[0] == pytest.approx(exact, rel=1e-5)
    assert f[1] == 0.0 and f[2] == 0.0


def test_mouse_force(oracle):
    p = _params(preset=L.WC_PRESET_REFERENCE_DEFAULT, mouse_force=1)
    p.camera_position[:] = (0.5, 0.5, -2.0)
    p.mouse_ray_direction[:] = (0.0, 0.0, 1.0)
    h = float(np.float32(p.kernel_radius))
    # particle 0.5h off the ray: distance to ray = 0.5h, toMouse = pos - camera
    pos = np.array([0.5 + 0.5 * h, 0.5, 0.3], np.float64)
    pr = 40.0
    f = oracle.mouse_force(p, pos, pr)
    tm = pos - np.array([0.5, 0.5, -2.0])
    dd = 0.5 * h
    exact = -p.particle_mass * pr * (-45.0 / (math.pi * h ** 6)) * (h - dd) ** 2 * tm / dd * 1e-5
    np.testing.assert_allclose(f, exact, rtol=2e-3)
    assert np.all(oracle.mouse_force(p, [0.9, 0.9, 0.3], pr) == 0.0)  # farther than h from the ray


def test_update_brute_force_bit_exact(oracle):
    """O(N^2) forces + integration in NEIGHBORHOOD order, in numpy float32 scalars, match the
    grid-based update pass bit for bit."""
    p, parts = scenes.scaled_block((7, 6, 6), 0.015, seed=5)
    st = oracle.OracleState(p, parts)
    st.step(2e-4, 3)  # some motion and pressure
    st.sort_run()
    st.density()
    ps = st.buf2.copy()
    dt = np.float32(2e-4)
    st.update(float(dt))
    f = np.float32
    h, m, cs, cv, mu = f(p.kernel_radius), f(p.particle_mass), f(p.spiky_const), f(p.visc_const), f(p.viscosity)
    g = np.array(p.gravity[:], f)
    for i in range(0, len(ps), 5):
        pi, vi, rhoi, pri = ps["position"][i], ps["velocity"][i], ps["density"][i], ps["pressure"][i]
        pf = np.zeros(3, f)
        vf = np.zeros(3, f)
        for j in _brute_force_order(oracle, p, ps, i):
            r = pi - ps["position"][j]
            dist = np.sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2], dtype=f)
            if dist >= h:
                continue
            P = (pri + ps["pressure"][j]) / (f(2.0) * ps["density"][j])
            if P > 0:
                d = dist + f(1e-16)
                a = h - d
                w = ((a * a) * (r / d)) * cs
                pf = pf - (m * P) * w
            vf = vf + (m * ((ps["velocity"][j] - vi) / ps["density"][j])) * ((h - dist) * cv)
        ext = g * rhoi
        ext = ext + (np.zeros(3, f) + np.zeros(3, f))
        vf = vf * mu
        acc = ((pf + vf) + ext) / (rhoi + f(1e-16))
        v = np.minimum(np.maximum(vi + acc * dt, -f(p.max_speed)), f(p.max_speed))
        x = pi + v * dt
        lo, hi = f(p.border), f(1.0) - f(p.border)
        for c in range(3):
            if x[c] < lo:
                v[c] = v[c] * -f(p.wall_damping)
                x[c] = lo
            elif x[c] > hi:
                v[c] = v[c] * -f(p.wall_damping)
                x[c] = hi
        assert st.buf1["velocity"][i].tobytes() == v.astype(f).tobytes(), i
        assert st.buf1["position"][i].tobytes() == x.astype(f).tobytes(), i


def test_planted_mistake_is_caught(oracle):
    """The brute-force order check would catch a neighbourhood order slip: summing the same terms
    in plain slot order differs in the last bits for some particle."""
    p, parts = scenes.scaled_block((9, 8, 7), 0.015, seed=11)
    st = oracle.OracleState(p, parts)
    st.sort_run()
    st.density()
    ps = st.buf2
    f = np.float32
    h, m, c6 = f(p.kernel_radius), f(p.particle_mass), f(p.poly6_const)
    differs = 0
    for i in range(0, len(ps), 3):
        pi = ps["position"][i]
        a0 = h * h
        rho = m * (((a0 * a0) * a0) * c6)
        for j in sorted(_brute_force_order(oracle, p, ps, i)):  # wrong order: by slot only
            r = pi - ps["position"][j]
            dist = np.sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2], dtype=f)
            if dist < h:
                a = h * h - dist * dist
                rho = rho + m * (((a * a) * a) * c6)
        differs += ps["density"][i].tobytes() != f(rho).tobytes()
    assert differs > 0


def test_cutoff_by_squared_distance_is_exact():
    """The kernels accept a neighbour when d2 < cut2, with cut2 the smallest float whose
    correctly rounded sqrt is >= h; the reference tests length(r) >= h (density.comp:117,
    update.comp:173).  sqrtf is monotone, so the two agree for every float d2: checked on
    every float within 4096 ulp of cut2 for each preset's h."""
    f32 = np.float32
    for h in (0.03, 0.016, 0.004, 0.04, 0.0123):
        h = f32(h)
        cut2 = f32(h * h)
        while np.sqrt(np.nextafter(cut2, f32(0))) >= h:
            cut2 = np.nextafter(cut2, f32(0))
        while np.sqrt(cut2) < h:
            cut2 = np.nextafter(cut2, f32(np.inf))
        bits = cut2.view(np.uint32).astype(np.int64) + np.arange(-4096, 4097)
        d2 = bits.astype(np.uint32).view(np.float32)
        assert np.array_equal(np.sqrt(d2) >= h, d2 >= cut2)


def test_advect_single_step_outcomes(oracle):
    """advect.comp:20-59: x + v dt, then the clamp at the shader's own border 0.01 and damping 0.3
    (the parameter block's border and damping are set to other values and must not be used)."""
    p = _params(num_particles=4, border=0.002, wall_damping=0.5)
    parts = np.zeros(4, L.PARTICLE_DTYPE)
    parts["position"] = [[0.5, 0.5, 0.5], [0.02, 0.5, 0.5], [0.5, 0.995, 0.5], [0.5, 0.5, 0.0105]]
    parts["velocity"] = [[0.25, -0.5, 1.0], [-1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, -0.01]]
    parts["density"] = [1.0, 2.0, 3.0, 4.0]
    parts["pressure"] = [5.0, 6.0, 7.0, 8.0]
    st = oracle.OracleState(p, parts)
    st.advect(0.1)
    x, v = st.buf1["position"], st.buf1["velocity"]
    f = np.float32
    # interior: exact sums at these values (0.5 + 0.025 etc. rounded once each)
    assert x[0].tolist() == [f(f(0.5) + f(f(0.25) * f(0.1))), f(f(0.5) + f(f(-0.5) * f(0.1))),
                             f(f(0.5) + f(f(1.0) * f(0.1)))]
    assert v[0].tolist() == [f(0.25), f(-0.5), f(1.0)]
    # low wall: x = 0.02 - 0.1 < 0.01 -> 0.01, v = -1 * -0.3
    assert x[1][0] == f(0.01) and v[1][0] == f(f(-1.0) * f(-0.3))
    # high wall: 0.995 + 0.1 > size - 0.01 -> size - 0.01, v = 1 * -0.3
    assert x[2][1] == f(f(p.size) - f(0.01)) and v[2][1] == f(-0.3)
    # 0.0105 - 0.001 = 0.0095 < 0.01: clamped; the parameter block's border 0.002 would not clamp
    assert x[3][2] == f(0.01) and v[3][2] == f(f(-0.01) * f(-0.3))
    # density and pressure pass through
    assert st.buf1["density"].tolist() == [1.0, 2.0, 3.0, 4.0]
    assert st.buf1["pressure"].tolist() == [5.0, 6.0, 7.0, 8.0]
