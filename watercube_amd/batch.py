"""Many independent scenes advanced together on one GPU (``wc_create_batch``).

Every pass of a step is one launch over all scenes and a step is one captured HIP graph; see
``include/watercube.h``.  ``scenes_256`` builds the 256-scene workload: particle counts drawn
log-uniform in [4,096, 262,144] from a seed, each scene a seeded block, ball or two-column
fill at the 16,384-particle dam break's constants scaled to the scene's lattice spacing
(h = 2 s, m = rho0 s^3, dt = 2e-4 s / 0.015).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from . import scenes as S

FILL_KINDS = ("block", "ball", "two_columns")


def scene_params(n: int, spacing: float) -> L.wc_params:
    p = L.preset(L.WC_PRESET_DAM_BREAK_16K)
    p.num_particles = n
    p.kernel_radius = np.float32(2.0 * spacing)
    p.particle_mass = np.float32(1000.0 * spacing ** 3)
    p.particle_radius = np.float32(p.kernel_radius / 4)
    p.grid_res = max(1, int(math.floor(1.0 / float(p.kernel_radius))))
    return L.derive(p)


def fill(kind: str, n_target: int, seed: int):
    """(particles, spacing) of one scene of about n_target particles."""
    if kind == "block":
        # the dam break's 2:2:1 block, 0.48 x 0.48 x 0.24, in the corner
        nz = max(2, round((n_target / 4) ** (1 / 3)))
        nx = 2 * nz
        s = 0.48 / nx
        return S.block(nx, nx, nz, s, 0.1 * s, seed), s
    if kind == "ball":
        # a ball of the block's volume, centred at y = 0.6, falling from rest
        s = (0.48 * 0.48 * 0.24 / n_target) ** (1 / 3)
        return S.sphere(n_target, (0.5, 0.6, 0.5), s, 0.1 * s, seed), s
    if kind == "two_columns":
        # two 1:2:1 columns in opposite corners (a double dam break)
        nz = max(2, round((n_target / 4) ** (1 / 3)))
        nx, ny = nz, 2 * nz
        s = 0.25 / nx
        a = S.block(nx, ny, nz, s, 0.1 * s, seed)
        b = S.block(nx, ny, nz, s, 0.1 * s, seed + 7919, origin=(1.0 - 0.25, 0.0, 1.0 - 0.25))
        return np.concatenate([a, b]), s
    raise KeyError(kind)


def scenes_256(seed: int = 0, num_scenes: int = 256, n_min: int = 4096, n_max: int = 262144):
    """[(params, particles, dt, kind)] of the batched workload."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(num_scenes):
        n_target = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        kind = FILL_KINDS[int(rng.integers(0, len(FILL_KINDS)))]
        parts, spacing = fill(kind, n_target, seed * 1_000_003 + s)
        p = scene_params(len(parts), spacing)
        dt = float(np.float32(2e-4 * spacing / 0.015))
        out.append((p, parts, dt, kind))
    return out


class SceneBatch:
    """Host handle of a multi-scene context."""

    def __init__(self, scenes):
        self.params = [p.copy() for p, *_ in scenes]
        self._parts = [np.ascontiguousarray(x[1], dtype=L.PARTICLE_DTYPE) for x in scenes]
        self.dt = np.array([x[2] for x in scenes], np.float32)
        self.ctx = None

    @property
    def num_scenes(self) -> int:
        return len(self.params)

    @property
    def total_particles(self) -> int:
        return int(sum(p.num_particles for p in self.params))

    def setup(self) -> "SceneBatch":
        arr = (L.wc_params * self.num_scenes)(*self.params)
        ctx = C.c_void_p()
        L.check(L.lib().wc_create_batch(arr, self.num_scenes, C.byref(ctx)), "wc_create_batch")
        self.ctx = ctx
        for s, parts in enumerate(self._parts):
            L.check(L.lib().wc_scene_upload(self.ctx, s, parts.ctypes.data, len(parts)), "wc_scene_upload")
        self._parts = None
        self.set_dt(self.dt)
        return self

    def set_dt(self, dt) -> None:
        self.dt = np.ascontiguousarray(dt, np.float32)
        L.check(L.lib().wc_set_scene_dt(self.ctx, self.dt.ctypes.data_as(C.POINTER(C.c_float)), self.num_scenes),
                "wc_set_scene_dt")

    def step(self, nsteps: int = 1) -> None:
        L.check(L.lib().wc_step_scenes(self.ctx, nsteps), "wc_step_scenes")

    def synchronize(self) -> None:
        L.check(L.lib().wc_synchronize(self.ctx), "wc_synchronize")

    def tune(self, reps: int = 3) -> dict:
        nv = L.lib().wc_num_variants()
        md, mu, mdl, mul = ((C.c_double * nv)() for _ in range(4))
        t = L.wc_tuning()
        L.check(L.lib().wc_tune(self.ctx, C.c_float(0.0), reps, C.byref(t), md, mu, mdl, mul), "wc_tune")
        return {"density_variant": t.density_variant, "update_variant": t.update_variant,
                "neighbour_list": t.neighbour_list, "density_ms": list(md), "update_ms": list(mu),
                "density_list_ms": list(mdl), "update_list_ms": list(mul)}

    def scene_particles(self, s: int, which: int = 1) -> np.ndarray:
        n = int(self.params[s].num_particles)
        out = np.empty(n, L.PARTICLE_DTYPE)
        L.check(L.lib().wc_scene_download(self.ctx, s, which, out.ctypes.data, n), "wc_scene_download")
        return out

    def scene_cells(self, s: int):
        r = int(self.params[s].grid_res)
        counts = np.empty(r ** 3, np.uint32)
        offsets = np.empty(r ** 3, np.uint32)
        L.check(L.lib().wc_scene_get_cells(self.ctx, s, L.u32ptr(counts), L.u32ptr(offsets), r ** 3),
                "wc_scene_get_cells")
        return counts, offsets

    def latency_ms(self) -> np.ndarray:
        out = np.zeros(self.num_scenes, np.float64)
        L.check(L.lib().wc_scene_latency(self.ctx, out.ctypes.data_as(C.POINTER(C.c_double)), self.num_scenes),
                "wc_scene_latency")
        return out

    def close(self) -> None:
        if self.ctx is not None:
            L.lib().wc_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self.setup() if self.ctx is None else self

    def __exit__(self, *exc):
        self.close()
