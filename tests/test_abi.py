"""C-ABI checks that need no GPU: the library loads, exports every symbol include/watercube.h
declares, the ctypes mirror of every struct matches the C layout, presets and parameter
derivation, the host-side scene generators, and the error paths of a GPU-less host."""
import ctypes as C
import math
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from watercube_amd import _lib as L
from watercube_amd import scenes

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_symbols():
    names = []
    for h in HEADERS:
        text = h.read_text()
        names += re.findall(r"WC_API\s+[\w\s\*]+?\b(wc_\w+)\s*\(", text)
    return names


def test_every_declared_symbol_is_exported(wclib):
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(wclib.lib(), n)]
    assert not missing, missing
    # and the ctypes table covers exactly the declared API
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_exported_symbols_via_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (wc_\w+)", out.stdout))
    assert set(declared_symbols()) <= exported


LAYOUT_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "watercube.h"
#define F(s, f) printf(#s "." #f " %zu\n", offsetof(s, f));
int main(void) {
    printf("wc_particle %zu\nwc_params %zu\nwc_device_view %zu\n", sizeof(wc_particle), sizeof(wc_params),
           sizeof(wc_device_view));
    F(wc_particle, position) F(wc_particle, density) F(wc_particle, velocity) F(wc_particle, pressure)
    F(wc_params, num_particles) F(wc_params, gravity) F(wc_params, poly6_const) F(wc_params, eos)
    F(wc_params, mouse_force) F(wc_params, camera_position) F(wc_params, time_scale) F(wc_params, reserved)
    F(wc_device_view, cell_range) F(wc_device_view, stream) F(wc_device_view, num_cells)
    return 0;
}
"""


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(LAYOUT_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                                check=True).stdout.split("\n") if line)
    out = {k: int(v) for k, v in out.items()}
    assert out["wc_particle"] == 32 == L.PARTICLE_DTYPE.itemsize
    assert out["wc_params"] == C.sizeof(L.wc_params)
    assert out["wc_device_view"] == C.sizeof(L.wc_device_view)
    for key, off in out.items():
        if "." not in key:
            continue
        s, f = key.split(".")
        if s == "wc_particle":
            assert L.PARTICLE_DTYPE.fields[f][1] == off, key
        else:
            assert getattr(getattr(L, s), f).offset == off, key


def test_abi_version_and_target(wclib):
    assert wclib.lib().wc_abi_version() == 1
    assert wclib.lib().wc_target_arch() == b"gfx950"


def test_presets_and_derivation():
    p = L.preset(L.WC_PRESET_REFERENCE_DEFAULT)
    # src/core/Fluid.cpp:9-27 and 206-216
    assert (p.num_particles, p.grid_res) == (80000, 21)
    assert p.kernel_radius == np.float32(0.04) and p.particle_mass == np.float32(0.08)
    assert (p.rest_density, p.viscosity, p.stiffness, p.rest_pressure) == (500.0, 200.0, 100.0, 0.0)
    assert list(p.gravity) == [0.0, -900.0, 0.0]
    assert p.eos == L.WC_EOS_CUBIC and p.wall_density == 1 and p.wall_force == 1
    assert p.wall_damping == np.float32(0.3) and p.border == np.float32(0.001) and p.max_speed == 50.0
    assert p.bin_size == np.float32(np.float32(1.0) / np.float32(21))
    q = L.preset(L.WC_PRESET_BLOCK_1M)
    assert (q.num_particles, q.grid_res) == (1_000_000, 62)
    assert q.bin_size >= q.kernel_radius
    assert q.eos == L.WC_EOS_LINEAR and q.wall_density == 0 and q.wall_force == 0
    r = L.preset(L.WC_PRESET_BLOCK_SPHERE_17M)
    assert r.num_particles == 16_777_216 + 1_000_000 and r.grid_res == 250 and r.bin_size >= r.kernel_radius
    bad = L.wc_params()
    assert L.lib().wc_params_preset(C.byref(bad), 99) == L.WC_ERR_INVALID_ARG
    assert b"preset" in L.lib().wc_last_error()


class SplitMix64:
    """Independent restatement of the generator contract written in include/watercube.h."""

    def __init__(self, seed):
        self.s = seed & (2 ** 64 - 1)

    def next(self):
        M = 2 ** 64 - 1
        self.s = (self.s + 0x9E3779B97F4A7C15) & M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    def jitter(self, amp):
        f = np.float32
        u = f(self.next() >> 40) * f(1.0 / 16777216.0)
        return (f(2.0) * u - f(1.0)) * f(amp)


def test_block_generator_bit_exact():
    nx, ny, nz, s, j, seed = 5, 4, 3, 0.015, 0.0015, 42
    got = scenes.block(nx, ny, nz, s, j, seed, origin=(0.1, 0.2, 0.3), velocity=(0.0, -2.0, 0.5))
    rng = SplitMix64(seed)
    f = np.float32
    k = 0
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                base = [f(o) + (f(i) + f(0.5)) * f(s) for o, i in zip((0.1, 0.2, 0.3), (x, y, z))]
                jit = [rng.jitter(j) for _ in range(3)]
                want = np.array([b + q for b, q in zip(base, jit)], f)
                assert got["position"][k].tobytes() == want.tobytes()
                k += 1
    assert np.all(got["velocity"] == np.float32([0.0, -2.0, 0.5]))
    assert np.all(got["density"] == 0) and np.all(got["pressure"] == 0)


def test_block_generator_properties():
    a = scenes.block(32, 32, 16, 0.015, 0.0015, 0)
    b = scenes.block(32, 32, 16, 0.015, 0.0015, 0)
    c = scenes.block(32, 32, 16, 0.015, 0.0015, 1)
    assert a.tobytes() == b.tobytes() and a.tobytes() != c.tobytes()
    lattice = (np.stack(np.meshgrid(np.arange(32), np.arange(32), np.arange(16), indexing="ij"), -1)
               .transpose(2, 1, 0, 3).reshape(-1, 3) + 0.5) * 0.015
    dev = a["position"] - lattice
    assert np.all(np.abs(dev) <= 0.0015 * (1 + 1e-5))
    assert abs(dev.mean()) < 1e-4 and dev.std() > 0.0015 / 2.5  # uniform in [-j, j]: std = j/sqrt(3)


def test_sphere_generator():
    n = 5000
    s = scenes.sphere(n, (0.5, 0.6, 0.5), 0.01, 0.0, seed=3, velocity=(0, -2, 0))
    assert len(s) == n
    d = np.linalg.norm(s["position"] - np.float32([0.5, 0.6, 0.5]), axis=1)
    r = (3 * n / (4 * np.pi)) ** (1 / 3) * 0.01
    assert d.max() < r * 1.1 and np.median(d) < r
    assert len({tuple(x) for x in s["position"]}) == n
    assert np.all(s["velocity"][:, 1] == -2.0)


def test_reference_scene_layout():
    n = 1000
    r = 0.01
    a = scenes.reference_scene(n, r, 0)
    d = int(np.ceil(np.cbrt(n)))
    idx = np.arange(n)
    lattice = np.stack([idx % d, (idx // d) % d, idx // (d * d)], 1) * np.float32(r * 1.75)
    assert np.all(np.abs(a["position"] - lattice) <= r * 1.75 / 4 * (1 + 1e-5))


def test_no_device_error_path():
    """On a host without a GPU, wc_create fails with WC_ERR_NO_DEVICE and a message; with one it
    would succeed.  Either way nothing crashes and the ABI reports through status codes."""
    p = L.preset(L.WC_PRESET_DAM_BREAK_16K)
    ctx = C.c_void_p()
    rc = L.lib().wc_create(C.byref(p), C.byref(ctx))
    if rc == L.WC_OK:
        L.lib().wc_destroy(ctx)
    else:
        assert rc == L.WC_ERR_NO_DEVICE
        assert L.lib().wc_last_error() != b""
    assert L.lib().wc_create(None, C.byref(ctx)) == L.WC_ERR_INVALID_ARG
    assert L.lib().wc_destroy(None) == L.WC_OK


def test_cutoff_threshold_equivalence():
    """The kernels compare d2 >= t* instead of sqrtf(d2) >= h.  t* is the smallest float whose
    square root rounds to >= h; sqrt is monotonic, so the tests agree for every d2."""
    for h in (0.03, 0.016, 0.004, 0.04):
        hf = np.float32(h)
        t = np.float32(hf * hf)
        while np.sqrt(np.nextafter(t, np.float32(0))) >= hf:
            t = np.nextafter(t, np.float32(0))
        while np.sqrt(t) < hf:
            t = np.nextafter(t, np.float32(np.inf))
        assert np.sqrt(t) >= hf and np.sqrt(np.nextafter(t, np.float32(0))) < hf
        # every float in a window around t gives the same answer both ways
        xs = (t.view(np.uint32) + np.arange(-2000, 2000, dtype=np.int64)).astype(np.uint32).view(np.float32)
        assert np.array_equal(np.sqrt(xs) >= hf, xs >= t)


def test_gravity_direction_closed_forms():
    """Fluid::updateGravity (src/core/Fluid.cpp:259-263): transpose(toMat3(q)) * (0, -1, 0).
    Closed forms for the identity, half turns and quarter turns, and agreement with a float64
    rotation matrix built independently from the quaternion for random orientations."""
    from watercube_amd.fluid import gravity_direction

    s = math.sqrt(0.5)
    assert gravity_direction((1, 0, 0, 0)).tolist() == [0.0, -1.0, 0.0]
    assert gravity_direction((0, 1, 0, 0)).tolist() == [0.0, 1.0, 0.0]    # half turn about x
    assert gravity_direction((0, 0, 0, 1)).tolist() == [0.0, 1.0, 0.0]    # half turn about z
    assert gravity_direction((0, 0, 1, 0)).tolist() == [0.0, -1.0, 0.0]   # half turn about y: unchanged
    g = gravity_direction((s, 0, 0, s))   # quarter turn about z: R^T maps -y to -x
    assert np.allclose(g, [-1.0, 0.0, 0.0], atol=1e-6)
    g = gravity_direction((s, s, 0, 0))   # quarter turn about x: R^T maps -y to +z
    assert np.allclose(g, [0.0, 0.0, 1.0], atol=1e-6)
    rng = np.random.default_rng(5)
    for _ in range(100):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        want = R.T @ np.array([0.0, -1.0, 0.0])
        assert np.allclose(gravity_direction(q.astype(np.float32)), want, atol=3e-6)
