"""Python host side mirroring the reference's ``core::Fluid`` and ``core::Sort`` stages.

Method names follow src/core/Fluid.h and src/core/Sort.h; every call goes through the C ABI of
``libwatercube.so`` (include/watercube.h) and runs on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .scenes import reference_scene


def gravity_direction(q_wxyz) -> np.ndarray:
    """Fluid::updateGravity's direction (src/core/Fluid.cpp:259-263) for the fluid's orientation
    quaternion q = (w, x, y, z): transpose(toMat3(q)) * (0, -1, 0) (wc_gravity_direction)."""
    q = (C.c_float * 4)(*[float(v) for v in q_wxyz])
    out = (C.c_float * 3)()
    L.check(L.lib().wc_gravity_direction(q, out), "wc_gravity_direction")
    return np.array(out[:], np.float32)


class Sort:
    """``core::Sort`` (src/core/Sort.h:20-71): the grid build of one simulation context.

    ``run()`` is ``Sort::run(buffer1, buffer2)`` (src/core/Sort.cpp:254-267); the getters return
    host copies of what ``getCountBuffer`` / ``getOffsetBuffer`` / ``getSortedBuffer``
    (src/core/Sort.h:35-37) hold on the device.
    """

    def __init__(self, fluid: "Fluid"):
        self._fluid = fluid

    @property
    def num_cells(self) -> int:
        r = self._fluid.params.grid_res
        return r * r * r

    def run(self) -> None:
        L.check(L.lib().wc_sort_run(self._fluid.ctx), "wc_sort_run")

    def getCountBuffer(self) -> np.ndarray:  # noqa: N802 (reference name)
        out = np.empty(self.num_cells, np.uint32)
        L.check(L.lib().wc_get_counts(self._fluid.ctx, L.u32ptr(out), out.size), "wc_get_counts")
        return out

    def getOffsetBuffer(self) -> np.ndarray:  # noqa: N802
        out = np.empty(self.num_cells, np.uint32)
        L.check(L.lib().wc_get_offsets(self._fluid.ctx, L.u32ptr(out), out.size), "wc_get_offsets")
        return out

    def getSortedBuffer(self) -> np.ndarray:  # noqa: N802
        n = self._fluid.params.num_particles
        out = np.empty(n, np.uint32)
        L.check(L.lib().wc_get_sorted(self._fluid.ctx, L.u32ptr(out), n), "wc_get_sorted")
        return out

    def getCellKeys(self) -> np.ndarray:  # noqa: N802
        n = self._fluid.params.num_particles
        out = np.empty(n, np.uint32)
        L.check(L.lib().wc_get_cell_keys(self._fluid.ctx, L.u32ptr(out), n), "wc_get_cell_keys")
        return out


class Fluid:
    """``core::Fluid`` (src/core/Fluid.h:32-141), simulation half, on one GPU.

    ``setup()`` allocates the context and uploads the initial particles (src/core/Fluid.cpp:203-235);
    ``update(time)`` is one ``Fluid::update`` (src/core/Fluid.cpp:342-354) with
    ``dt = float(time) * time_scale`` as src/core/Fluid.cpp:308 computes it.
    """

    def __init__(self, params: L.wc_params, particles: np.ndarray | None = None):
        self.params = params.copy()
        self._initial = particles
        self.ctx = None
        self.sort_: Sort | None = None

    # -- reference API ----------------------------------------------------------------------
    def numParticles(self) -> int:  # noqa: N802
        return int(self.params.num_particles)

    def generateInitialParticles(self, seed: int = 0) -> np.ndarray:  # noqa: N802
        return reference_scene(self.params.num_particles, self.params.particle_radius, seed)

    def setup(self) -> "Fluid":
        if self._initial is None:
            self._initial = self.generateInitialParticles()
        parts = np.ascontiguousarray(self._initial, dtype=L.PARTICLE_DTYPE)
        if parts.shape[0] != self.params.num_particles:
            raise ValueError(f"{parts.shape[0]} particles for num_particles={self.params.num_particles}")
        ctx = C.c_void_p()
        L.check(L.lib().wc_create(C.byref(self.params), C.byref(ctx)), "wc_create")
        self.ctx = ctx
        L.check(L.lib().wc_upload(self.ctx, parts.ctypes.data, parts.shape[0]), "wc_upload")
        self.sort_ = Sort(self)
        return self

    def update(self, time: float) -> None:
        dt = np.float32(np.float32(time) * np.float32(self.params.time_scale))
        self.step(float(dt), 1)

    def runDensityProg(self) -> None:  # noqa: N802  (src/core/Fluid.cpp:268-289, on buffer 2)
        L.check(L.lib().wc_density(self.ctx), "wc_density")

    def runUpdateProg(self, dt: float) -> None:  # noqa: N802  (src/core/Fluid.cpp:294-321)
        L.check(L.lib().wc_update(self.ctx, C.c_float(dt)), "wc_update")

    def runAdvectProg(self, dt: float) -> None:  # noqa: N802  (src/core/Fluid.cpp:326-336, on buffer 1)
        L.check(L.lib().wc_advect(self.ctx, C.c_float(dt)), "wc_advect")

    def updateGravity(self, q_wxyz, strength: float) -> None:  # noqa: N802  (src/core/Fluid.cpp:259-263)
        """Rotating gravity: direction from the fluid's orientation q = (w, x, y, z), times
        `strength` (src/core/Fluid.cpp:310), pushed to the context."""
        d = gravity_direction(q_wxyz)
        for a in range(3):
            self.params.gravity[a] = float(np.float32(d[a]) * np.float32(strength))
        if self.ctx:
            L.check(L.lib().wc_set_params(self.ctx, C.byref(self.params)), "wc_set_params")

    # -- extras -----------------------------------------------------------------------------
    def step(self, dt: float, nsteps: int = 1) -> None:
        L.check(L.lib().wc_step(self.ctx, C.c_float(dt), nsteps), "wc_step")

    def synchronize(self) -> None:
        L.check(L.lib().wc_synchronize(self.ctx), "wc_synchronize")

    def upload(self, particles: np.ndarray) -> None:
        """Replace buffer 1 with `particles` (structured 32-byte records)."""
        parts = np.ascontiguousarray(particles, dtype=L.PARTICLE_DTYPE)
        L.check(L.lib().wc_upload(self.ctx, parts.ctypes.data, parts.shape[0]), "wc_upload")

    def particles(self, which: int = 1) -> np.ndarray:
        n = self.numParticles()
        out = np.empty(n, L.PARTICLE_DTYPE)
        L.check(L.lib().wc_download(self.ctx, which, out.ctypes.data, n), "wc_download")
        return out

    def profile(self, dt: float, nsteps: int) -> dict:
        ms = (C.c_double * L.WC_NUM_PASSES)()
        L.check(L.lib().wc_profile_steps(self.ctx, C.c_float(dt), nsteps, ms), "wc_profile_steps")
        return {L.lib().wc_pass_name(k).decode(): ms[k] for k in range(L.WC_NUM_PASSES)}

    def tune(self, dt: float, reps: int = 5) -> dict:
        """Time every variant of the density and force passes on the current state, without and
        with neighbour lists, and keep the fastest (``wc_tune``); the particle state is left as
        it was."""
        nv = L.lib().wc_num_variants()
        md, mu, mdl, mul = ((C.c_double * nv)() for _ in range(4))
        t = L.wc_tuning()
        L.check(L.lib().wc_tune(self.ctx, C.c_float(dt), reps, C.byref(t), md, mu, mdl, mul), "wc_tune")
        return {"density_variant": t.density_variant, "update_variant": t.update_variant,
                "neighbour_list": t.neighbour_list, "density_ms": list(md), "update_ms": list(mu),
                "density_list_ms": list(mdl), "update_list_ms": list(mul), "variants": L.variants()}

    def set_tuning(self, density_variant: int, update_variant: int, neighbour_list: int = 0) -> None:
        t = L.wc_tuning(density_variant, update_variant, neighbour_list)
        L.check(L.lib().wc_set_tuning(self.ctx, C.byref(t)), "wc_set_tuning")

    def neighbour_counts(self) -> np.ndarray:
        """Accepted neighbours per sorted particle from the last density pass with neighbour
        lists (0xFFFFFFFF: more than the list holds)."""
        n = self.params.num_particles
        out = np.empty(n, np.uint32)
        L.check(L.lib().wc_get_neighbour_counts(self.ctx, L.u32ptr(out), n), "wc_get_neighbour_counts")
        return out

    def get_tuning(self) -> tuple[int, int, int]:
        t = L.wc_tuning()
        L.check(L.lib().wc_get_tuning(self.ctx, C.byref(t)), "wc_get_tuning")
        return t.density_variant, t.update_variant, t.neighbour_list

    def track_ids(self, enable: bool = True) -> None:
        L.check(L.lib().wc_track_ids(self.ctx, 1 if enable else 0), "wc_track_ids")

    def ids(self, which: int = 1) -> np.ndarray:
        n = self.numParticles()
        out = np.empty(n, np.uint32)
        L.check(L.lib().wc_get_ids(self.ctx, which, L.u32ptr(out), n), "wc_get_ids")
        return out

    def max_density(self) -> float:
        v = C.c_float()
        L.check(L.lib().wc_max_density(self.ctx, C.byref(v)), "wc_max_density")
        return float(v.value)

    def close(self) -> None:
        if self.ctx is not None:
            L.lib().wc_destroy(self.ctx)
            self.ctx = None

    def __enter__(self) -> "Fluid":
        return self.setup() if self.ctx is None else self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass
