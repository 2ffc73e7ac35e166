"""ctypes wrapper of the CPU restatement (``oracle/libwc_oracle.so``, built from wc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the benchmark's CPU-baseline
leg, as the checker.  The library under test never uses it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "libwc_oracle.so"

# Same 32-byte record and parameter block as the library's C ABI (include/watercube.h).
PARTICLE_DTYPE = np.dtype(
    {
        "names": ["position", "density", "velocity", "pressure"],
        "formats": [("<f4", (3,)), "<f4", ("<f4", (3,)), "<f4"],
        "offsets": [0, 12, 16, 28],
        "itemsize": 32,
    }
)

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return SO


def lib():
    global _lib
    if _lib is None:
        if not SO.exists():
            build()
        L = C.CDLL(str(SO))
        vp, u32p, f32p = C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_float)
        sig = {
            "wco_poly6": (C.c_float, [vp, C.c_float]),
            "wco_spiky": (None, [vp, f32p, C.c_float, f32p]),
            "wco_wvis": (C.c_float, [vp, C.c_float]),
            "wco_cell": (None, [vp, f32p, C.POINTER(C.c_int)]),
            "wco_cell_key": (C.c_uint32, [vp, f32p]),
            "wco_wall_density": (C.c_float, [vp, f32p]),
            "wco_wall_forces": (None, [vp, f32p, f32p]),
            "wco_mouse_force": (None, [vp, f32p, C.c_float, f32p]),
            "wco_eos": (C.c_float, [vp, C.c_float]),
            "wco_count": (None, [vp, vp, C.c_uint32, u32p, u32p]),
            "wco_linear_scan": (None, [u32p, C.c_uint32, u32p]),
            "wco_reorder": (None, [vp, vp, C.c_uint32, u32p, u32p, u32p, vp, u32p]),
            "wco_sort_run": (None, [vp, vp, vp, u32p, u32p, u32p, u32p]),
            "wco_density": (None, [vp, vp, u32p, u32p, C.c_int]),
            "wco_update": (None, [vp, vp, u32p, u32p, vp, C.c_float, C.c_int]),
            "wco_step": (None, [vp, vp, vp, u32p, u32p, u32p, u32p, C.c_float, C.c_uint32, C.c_int]),
            "wco_step_timed": (None, [vp, vp, vp, u32p, u32p, u32p, u32p, C.c_float, C.c_uint32, C.c_int,
                                      C.POINTER(C.c_double)]),
            "wco_advect": (None, [vp, vp, C.c_uint32, C.c_float]),
            "wco_max_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _u32(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def _pp(params):
    return C.cast(C.byref(params), C.c_void_p)


class OracleState:
    """Host buffers of one simulated scene: buffer 1/2 and the grid arrays."""

    def __init__(self, params, particles: np.ndarray):
        self.params = params
        n = int(params.num_particles)
        r = int(params.grid_res)
        self.buf1 = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE).copy()
        self.buf2 = self.buf1.copy()
        self.keys = np.zeros(max(n, 1), np.uint32)
        self.sorted = np.zeros(max(n, 1), np.uint32)
        self.counts = np.zeros(r * r * r, np.uint32)
        self.offsets = np.zeros(r * r * r, np.uint32)

    def sort_run(self):
        lib().wco_sort_run(_pp(self.params), self.buf1.ctypes.data, self.buf2.ctypes.data, _u32(self.keys),
                           _u32(self.counts), _u32(self.offsets), _u32(self.sorted))

    def density(self, threads: int = 0):
        lib().wco_density(_pp(self.params), self.buf2.ctypes.data, _u32(self.counts), _u32(self.offsets),
                          threads or lib().wco_max_threads())

    def update(self, dt: float, threads: int = 0):
        lib().wco_update(_pp(self.params), self.buf2.ctypes.data, _u32(self.counts), _u32(self.offsets),
                         self.buf1.ctypes.data, C.c_float(dt), threads or lib().wco_max_threads())

    def advect(self, dt: float):
        """advect.comp:20-59 on buffer 1, in place."""
        lib().wco_advect(_pp(self.params), self.buf1.ctypes.data, self.buf1.shape[0], C.c_float(dt))

    def step(self, dt: float, nsteps: int = 1, threads: int = 0):
        lib().wco_step(_pp(self.params), self.buf1.ctypes.data, self.buf2.ctypes.data, _u32(self.keys),
                       _u32(self.counts), _u32(self.offsets), _u32(self.sorted), C.c_float(dt), nsteps,
                       threads or lib().wco_max_threads())

    def step_timed(self, dt: float, nsteps: int, threads: int):
        ms = (C.c_double * 3)()
        lib().wco_step_timed(_pp(self.params), self.buf1.ctypes.data, self.buf2.ctypes.data, _u32(self.keys),
                             _u32(self.counts), _u32(self.offsets), _u32(self.sorted), C.c_float(dt), nsteps,
                             threads, ms)
        return {"sort": ms[0], "density": ms[1], "update": ms[2]}


# -- scalar helpers (for pinning tests) ---------------------------------------------------------
def poly6(params, r):
    return lib().wco_poly6(_pp(params), r)


def spiky(params, r, d):
    out = (C.c_float * 3)()
    lib().wco_spiky(_pp(params), _f3(r), d, out)
    return np.array(out[:], np.float32)


def wvis(params, r):
    return lib().wco_wvis(_pp(params), r)


def cell(params, pos):
    out = (C.c_int * 3)()
    lib().wco_cell(_pp(params), _f3(pos), out)
    return tuple(out[:])


def cell_key(params, pos):
    return int(lib().wco_cell_key(_pp(params), _f3(pos)))


def wall_density(params, pos):
    return lib().wco_wall_density(_pp(params), _f3(pos))


def wall_forces(params, pos):
    out = (C.c_float * 3)()
    lib().wco_wall_forces(_pp(params), _f3(pos), out)
    return np.array(out[:], np.float32)


def mouse_force(params, pos, pressure):
    out = (C.c_float * 3)()
    lib().wco_mouse_force(_pp(params), _f3(pos), pressure, out)
    return np.array(out[:], np.float32)


def eos(params, rho):
    return lib().wco_eos(_pp(params), rho)


def linear_scan(counts: np.ndarray) -> np.ndarray:
    counts = np.ascontiguousarray(counts, np.uint32)
    out = np.zeros_like(counts)
    lib().wco_linear_scan(_u32(counts), counts.size, _u32(out))
    return out
