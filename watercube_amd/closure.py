"""Learned closure for large steps: data generation, training and evaluation on the GPU.

The network (``wc_mlp_*`` in include/watercube.h) maps a particle's 32 nearest neighbours
(relative positions and velocities), its density, the step multiplier K and the filled fraction
of its neighbour places to a correction of its acceleration.  With the correction, one step of
K * dt is meant to follow K plain steps of dt.

Training targets come from plain small-dt trajectories of the 1,000,000-particle block's
constants (h = 0.016, spacing 0.008, dt = 1e-4) on smaller scenes: for a particle at time t,

    target = ((v(t + K dt) - v(t)) / (K dt) - a_plain(t)) / 9.81 m/s^2,

the gap between its mean acceleration over the next K small steps and the plain acceleration a
single large step would use.  Particles within 2h of a wall are left out (walls clamp positions
and flip velocities, which no acceleration describes); at run time the correction applies only
away from walls as well.

    python -m watercube_amd.closure train --out watercube_amd/weights/closure.npy
    python -m watercube_amd.closure evaluate --weights watercube_amd/weights/closure.npy
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib as L
from . import scenes as S
from .fluid import Fluid

G = 9.81
NFEAT = 208
WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"
DEFAULT_WEIGHTS = WEIGHTS_DIR / "closure.npy"


class DeviceArray:
    """A device buffer owned by Python (wc_device_alloc)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        L.check(L.lib().wc_device_alloc(max(self.nbytes, 16), C.byref(p)), "wc_device_alloc")
        self.ptr = p

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        L.check(L.lib().wc_device_upload(self.ptr, a.ctypes.data, a.nbytes), "wc_device_upload")
        return self

    def download(self, rows: int | None = None) -> np.ndarray:
        shape = self.shape if rows is None else (rows,) + self.shape[1:]
        out = np.empty(shape, self.dtype)
        L.check(L.lib().wc_device_download(out.ctypes.data, self.ptr, out.nbytes), "wc_device_download")
        return out

    def offset(self, rows: int) -> C.c_void_p:
        row_bytes = int(np.prod(self.shape[1:])) * self.dtype.itemsize
        return C.c_void_p(self.ptr.value + rows * row_bytes)

    def free(self):
        if self.ptr:
            L.lib().wc_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.free()
        except Exception:
            pass


class ClosureNet:
    """Handle of one network (``wc_mlp``)."""

    def __init__(self, max_batch: int = 8192, seed: int = 0):
        h = C.c_void_p()
        L.check(L.lib().wc_mlp_create(max_batch, C.c_uint64(seed), C.byref(h)), "wc_mlp_create")
        self.h = h
        self.max_batch = max_batch

    @staticmethod
    def num_params() -> int:
        return L.lib().wc_mlp_num_params()

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params(), np.float32)
        L.check(L.lib().wc_mlp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size),
                "wc_mlp_get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        L.check(L.lib().wc_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size), "wc_mlp_set_params")

    def save(self, path) -> None:
        np.save(path, self.params(), allow_pickle=False)

    def load(self, path) -> "ClosureNet":
        self.set_params(np.load(path, allow_pickle=False))
        return self

    def predict(self, x: DeviceArray, y: DeviceArray, n: int) -> None:
        L.check(L.lib().wc_mlp_predict(self.h, x.ptr, y.ptr, n), "wc_mlp_predict")

    def forward(self, x: DeviceArray, y: DeviceArray, n: int) -> None:
        L.check(L.lib().wc_mlp_forward(self.h, x.ptr, y.ptr, n), "wc_mlp_forward")

    def gradient(self, x: DeviceArray, t: DeviceArray, n: int):
        g = np.empty(self.num_params(), np.float32)
        loss = C.c_double()
        L.check(L.lib().wc_mlp_gradient(self.h, x.ptr, t.ptr, n, g.ctypes.data_as(C.POINTER(C.c_float)),
                                        C.byref(loss)), "wc_mlp_gradient")
        return g, loss.value

    def train_step(self, x_ptr, t_ptr, n: int, lr: float, want_loss: bool = False) -> float | None:
        loss = C.c_double()
        L.check(L.lib().wc_mlp_train_step(self.h, x_ptr, t_ptr, n, C.c_float(lr),
                                          C.byref(loss) if want_loss else None), "wc_mlp_train_step")
        return loss.value if want_loss else None

    def profile_predict(self, x: DeviceArray, y: DeviceArray, n: int, reps: int = 10) -> float:
        ms = C.c_double()
        L.check(L.lib().wc_mlp_profile_predict(self.h, x.ptr, y.ptr, n, reps, C.byref(ms)), "wc_mlp_profile_predict")
        return ms.value

    def profile_train(self, x: DeviceArray, t: DeviceArray, n: int, lr: float = 1e-4, reps: int = 10) -> float:
        """Average ms of one training step over `reps` steps (they update the parameters)."""
        ms = C.c_double()
        L.check(L.lib().wc_mlp_profile_train(self.h, x.ptr, t.ptr, n, C.c_float(lr), reps, C.byref(ms)),
                "wc_mlp_profile_train")
        return ms.value

    def close(self):
        if self.h:
            L.lib().wc_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def mlp_flops_per_particle() -> int:
    """Multiply-adds x 2 of one forward pass, unpadded (195 -> 128 -> 128 -> 128 -> 3)."""
    return 2 * (195 * 128 + 128 * 128 + 128 * 128 + 128 * 3)


def mlp_train_flops_per_particle() -> int:
    """Multiply-adds x 2 of one training step per row, unpadded: the forward pass, every layer's
    weight gradient, and the input gradients of layers 2-4 (layer 1's is not needed)."""
    weights = 195 * 128 + 128 * 128 + 128 * 128 + 128 * 3
    return 2 * (weights + weights + (128 * 128 + 128 * 128 + 128 * 3))


# ---------------------------------------------------------------------------------------------
# training scenes: the 1,000,000-particle block's constants on 64k..262k-particle fills
# ---------------------------------------------------------------------------------------------
def training_scene(n_target: int, kind: str, seed: int):
    p = L.preset(L.WC_PRESET_BLOCK_1M)
    s = 0.008
    if kind == "block":
        nz = max(2, round((n_target / 4) ** (1 / 3)))
        parts = S.block(2 * nz, 2 * nz, nz, s, 0.1 * s, seed)
    elif kind == "ball":
        parts = S.sphere(n_target, (0.5, 0.6, 0.5), s, 0.1 * s, seed)
    elif kind == "tall_column":
        # full height of the 1,000,000-particle block (100 layers), narrower
        nx = max(2, round(math.sqrt(n_target / 100)))
        parts = S.block(nx, 100, nx, s, 0.1 * s, seed)
    elif kind == "corner_cube":
        nx = max(2, round(n_target ** (1 / 3)))
        parts = S.block(nx, nx, nx, s, 0.1 * s, seed)
    else:
        nx = max(2, round((n_target / 4) ** (1 / 3)))
        a = S.block(nx, 2 * nx, nx, s, 0.1 * s, seed)
        b = S.block(nx, 2 * nx, nx, s, 0.1 * s, seed + 7919, origin=(1.0 - nx * s, 0.0, 1.0 - nx * s))
        parts = np.concatenate([a, b])
    p.num_particles = len(parts)
    return p, parts


def _plain_sample(f: Fluid, n: int, dt: float, K: int, nsamp: int, rng, xbuf: DeviceArray, abuf: DeviceArray):
    """One snapshot of training samples at the current state; advances the fluid by K steps."""
    f.sort_.run()
    f.runDensityProg()
    ids_t = f.ids(2)
    vel_t = f.particles(2)["velocity"]
    # one plain step that also reports the plain acceleration
    L.check(L.lib().wc_closure_update(f.ctx, C.c_float(dt), None, C.c_float(0.0), abuf.ptr), "wc_closure_update")
    f.synchronize()
    acc = abuf.download(n)
    interior = np.nonzero(acc[:, 3] > 0.5)[0]
    if interior.size == 0:
        f.step(dt, K - 1)
        return None
    rows = rng.choice(interior, size=min(nsamp, interior.size), replace=False).astype(np.uint32)
    rows.sort()
    L.check(L.lib().wc_closure_features(f.ctx, C.c_float(K), L.u32ptr(rows), rows.size, xbuf.ptr),
            "wc_closure_features")
    x = xbuf.download(rows.size)
    if K > 1:
        f.step(dt, K - 1)
    ids_end = f.ids(1)
    vel_end = f.particles(1)["velocity"]
    slot_of = np.empty(n, np.int64)
    slot_of[ids_end] = np.arange(n)
    v_end = vel_end[slot_of[ids_t[rows]]]
    t = np.zeros((rows.size, 4), np.float32)
    t[:, :3] = ((v_end - vel_t[rows]) / np.float32(K * dt) - acc[rows, :3]) / np.float32(G)
    return x, t


def make_training_set(num_traj: int = 200, steps: int = 400, every: int = 20, per_snapshot: int = 1024,
                      seed: int = 0, ks=(2, 4), log=print):
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    kinds = ("block", "ball", "two_columns", "tall_column", "corner_cube")
    t0 = time.perf_counter()
    for j in range(num_traj):
        n_target = int(round(math.exp(rng.uniform(math.log(64 * 1024), math.log(262144)))))
        kind = kinds[j % len(kinds)]
        p, parts = training_scene(n_target, kind, seed * 100003 + j)
        n = p.num_particles
        dt = 1e-4
        xbuf = DeviceArray((per_snapshot, NFEAT), np.uint16)
        abuf = DeviceArray((n, 4), np.float32)
        with Fluid(p, parts) as f:
            f.track_ids(True)
            step = 0
            snap = 0
            while step < steps:
                K = ks[snap % len(ks)]
                if step % every == 0 and step + K <= steps:
                    out = _plain_sample(f, n, dt, K, per_snapshot, rng, xbuf, abuf)
                    if out is not None:
                        xs.append(out[0])
                        ts.append(out[1])
                    step += K
                    snap += 1
                else:
                    nxt = min(steps, (step // every + 1) * every)
                    f.step(dt, nxt - step)
                    step = nxt
            f.synchronize()
        xbuf.free()
        abuf.free()
        if (j + 1) % 10 == 0 or j == 0:
            log(f"trajectory {j + 1}/{num_traj} ({kind}, {n} particles): {sum(len(a) for a in xs)} samples, "
                f"{time.perf_counter() - t0:.1f} s")
    return np.concatenate(xs), np.concatenate(ts)


def train(x: np.ndarray, t: np.ndarray, epochs: int = 4, batch: int = 8192, lr: float = 2e-3, seed: int = 0,
          val_frac: float = 0.05, log=print) -> tuple[ClosureNet, dict]:
    rng = np.random.default_rng(seed)
    n = len(x)
    perm = rng.permutation(n)
    nval = max(batch, int(n * val_frac))
    val, tr = perm[:nval], perm[nval:]
    net = ClosureNet(max_batch=max(batch, nval), seed=seed)
    xd_val = DeviceArray((nval, NFEAT), np.uint16).upload(x[val])
    td_val = DeviceArray((nval, 4), np.float32).upload(t[val])
    baseline = float(np.mean(t[val][:, :3] ** 2))
    _, val0 = net.gradient(xd_val, td_val, nval)
    log(f"samples: {len(tr)} train, {nval} validation; loss of predicting zero {baseline:.5f}, "
        f"initial network {val0:.5f}")
    xd = DeviceArray((len(tr), NFEAT), np.uint16)
    td = DeviceArray((len(tr), 4), np.float32)
    steps_per_epoch = len(tr) // batch
    total = epochs * steps_per_epoch
    history = []
    step = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = tr[rng.permutation(len(tr))]
        xd.upload(x[order])
        td.upload(t[order])
        for b in range(steps_per_epoch):
            cur_lr = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(1, total)))
            net.train_step(xd.offset(b * batch), td.offset(b * batch), batch, cur_lr)
            step += 1
        _, vloss = net.gradient(xd_val, td_val, nval)
        history.append(vloss)
        log(f"epoch {ep + 1}/{epochs}: validation loss {vloss:.5f} ({vloss / baseline:.3f} of predicting zero), "
            f"{time.perf_counter() - t0:.1f} s")
    for a in (xd, td, xd_val, td_val):
        a.free()
    return net, {"baseline_loss": baseline, "initial_loss": val0, "val_loss": history, "samples": int(n),
                 "steps": step, "batch": batch, "epochs": epochs}


# ---------------------------------------------------------------------------------------------
# evaluation on the 1,000,000-particle block
# ---------------------------------------------------------------------------------------------
def _densities_by_id(f: Fluid) -> np.ndarray:
    f.sort_.run()
    f.runDensityProg()
    ids = f.ids(2)
    rho = f.particles(2)["density"]
    out = np.empty_like(rho)
    out[ids] = rho
    return out


def _state_by_id(f: Fluid) -> np.ndarray:
    ids = f.ids(1)
    rec = f.particles(1)
    out = np.empty_like(rec)
    out[ids] = rec
    return out


def _closure_run(f: Fluid, net: "ClosureNet | None", dt: float, K: int, macro: int, judge_tol: float,
                 smooth: bool, substeps: bool) -> L.wc_closure_stats:
    L.check(L.lib().wc_closure_set_smoothing(f.ctx, int(smooth)), "wc_closure_set_smoothing")
    L.check(L.lib().wc_closure_set_integrator(f.ctx, int(substeps)), "wc_closure_set_integrator")
    st = L.wc_closure_stats()
    L.check(L.lib().wc_closure_run(f.ctx, net.h if net is not None else None, C.c_float(dt), K, macro,
                                   C.c_float(judge_tol), C.byref(st)), "wc_closure_run")
    return st


# Large-step methods compared against K plain steps of dt: (network?, substep integrator?)
METHODS = {
    "plain_large": (False, False),    # one symplectic Euler step of K dt (graph step)
    "frozen_force": (False, True),    # K steps of dt under the force frozen at the step's start
    "learned": (True, False),         # one step of K dt with the learned correction
    "learned_frozen": (True, True),   # K frozen-force steps of dt with the learned correction
}


def one_step_errors(net: ClosureNet, p, parts: np.ndarray, dt: float, K: int, starts=(0, 30, 60, 90),
                    variants=None, smooth: bool = True) -> dict:
    """From states along a plain small-dt run: one step of K*dt by each method (judge disabled)
    against K plain steps of dt.  RMS velocity and position errors over the particles at least
    2h from every wall, per start and averaged, and their ratios to the plain large step's."""
    h = float(p.kernel_radius)
    size = float(p.size)

    def fluid(state):
        f = Fluid(p, state).setup()
        f.track_ids(True)
        if variants:
            f.set_tuning(*variants)
        return f

    rows = []
    with fluid(parts) as run:
        done = 0
        for s0 in starts:
            if s0 > done:
                run.step(dt, s0 - done)
                done = s0
            state = _state_by_id(run)
            pos0 = state["position"]
            interior = np.all((pos0 > 2 * h) & (pos0 < size - 2 * h), axis=1)
            with fluid(state) as a:
                a.step(dt, K)
                ref = _state_by_id(a)
            out = {}
            for m, (use_net, sub) in METHODS.items():
                with fluid(state) as b:
                    if m == "plain_large":
                        b.step(dt * K, 1)
                    else:
                        _closure_run(b, net if use_net else None, dt, K, 1, 1e30, smooth, sub)
                    out[m] = _state_by_id(b)

            def rms(x, field):
                d = (x[field] - ref[field])[interior]
                return float(np.sqrt(np.mean(np.sum(d.astype(np.float64) ** 2, axis=1))))

            rows.append({"start": s0, "interior": int(interior.sum()),
                         **{m: {"vel_rms": rms(x, "velocity"), "pos_rms": rms(x, "position")} for m, x in out.items()}})
    mean = {k: {m: float(np.mean([r[k][m] for r in rows])) for m in ("vel_rms", "pos_rms")} for k in METHODS}
    base = mean["plain_large"]
    ratios = {k: {"vel": mean[k]["vel_rms"] / max(base["vel_rms"], 1e-30),
                  "pos": mean[k]["pos_rms"] / max(base["pos_rms"], 1e-30)} for k in METHODS}
    return {"K": K, "per_start": rows, "mean": mean, "ratio_to_plain_large": ratios,
            "vel_err_ratio": ratios["learned"]["vel"], "pos_err_ratio": ratios["learned"]["pos"]}


def _sampled_plain(fluid, dt: float, K: int, macro: int, rho0: float):
    """Plain steps of K dt taken one at a time, reading the density maximum of every state (a
    step's density pass measures the state the step starts from; one more sort and density pass
    measures the last): the largest rise of the maximum over one step and its peak, both / rho0,
    and the final densities by particle identity.  Untimed."""
    maxima = []
    with fluid() as f:
        for _ in range(macro):
            f.step(dt * K, 1)
            maxima.append(f.max_density())
        rho = _densities_by_id(f)
        maxima.append(f.max_density())
    m = np.asarray(maxima, np.float64) / rho0
    return float(max(0.0, np.diff(m).max(initial=0.0))), float(m.max()), rho


def exact_correction_run(fluid, p, dt: float, K: int, macro: int, rho0: float):
    """Upper bound of the learned step.  Every macro step of K dt adds to each interior particle
    the exact value of the network's training target -- the gap between its mean acceleration over
    K plain steps of dt, run from the same state in a second context, and its plain acceleration
    -- so that its velocity after the step equals the K small steps' up to rounding.  No network
    predicts its target better; the error that remains is the one a correction of the acceleration
    cannot remove (the large step's position update).  Returns the largest rise of the density
    maximum over one step, its peak (both / rho0) and the final densities by identity."""
    n = int(p.num_particles)
    corr = DeviceArray((n, 4), np.float32)
    acc = DeviceArray((n, 4), np.float32)
    maxima = []
    c = np.zeros((n, 4), np.float32)
    with fluid() as a, fluid() as b:
        for _ in range(macro):
            a.sort_.run()
            a.runDensityProg()
            maxima.append(a.max_density())
            s0 = a.particles(2)
            # the same sorted state in b: its stable sort keeps the order, its identities are a's slots
            b.upload(s0)
            b.track_ids(True)
            b.sort_.run()
            b.runDensityProg()
            L.check(L.lib().wc_closure_update(b.ctx, C.c_float(dt), None, C.c_float(0.0), acc.ptr),
                    "wc_closure_update")
            b.synchronize()
            a_plain = acc.download()[:, :3]
            if K > 1:
                b.step(dt, K - 1)
            v_end = np.empty((n, 3), np.float32)
            v_end[b.ids(1)] = b.particles(1)["velocity"]
            c[:, :3] = ((v_end - s0["velocity"]) / np.float32(K * dt) - a_plain) / np.float32(G)
            corr.upload(c)
            L.check(L.lib().wc_closure_update(a.ctx, C.c_float(np.float32(dt) * np.float32(K)), corr.ptr,
                                              C.c_float(G), None), "wc_closure_update")
        rho = _densities_by_id(a)
        maxima.append(a.max_density())
    corr.free()
    acc.free()
    m = np.asarray(maxima, np.float64) / rho0
    return float(max(0.0, np.diff(m).max(initial=0.0))), float(m.max()), rho


def evaluate(net: ClosureNet | None, workload: str = "block_1m", horizon: int = 500, ks=(2, 4, 8),
             judge_tol: float = 0.01, smooth: bool = True, exact: bool = False, log=print) -> dict:
    """Plain small-dt run vs the large-step methods of METHODS over `horizon` small steps from the
    same state.  For each method and K:

    * peak_density: the largest density maximum of the run, / rho0, and peak_excess, its excess over
      the small-dt run's -- the density error the comparison is made on: a large step that
      compresses the fluid further than the small-dt run did shows up here, whatever happens to
      individual particles;
    * max_accepted_jump (judged runs): the largest rise of the density maximum over one accepted
      learned step, / rho0 -- the judge's own measure, at most K * judge_tol by construction
      (judge_tol is per small step of dt);
      max_step_jump (plain runs): the same rise over every step, for comparison;
    * mean / p99 / max of |rho - rho_ref| / rho0 per particle at the end of the horizon, particles
      matched by identity with the small-dt run (a trajectory-divergence measure);
    * simulated time per wall-clock second, and its ratio to the small-dt run.

    With `exact`, also the upper bound of any network (exact_correction_run).  Every run uses the
    same tuned neighbour-pass variants; graph capture and tuning stay outside the timed regions."""
    p, parts, dt, _ = S.workload(workload)
    n = int(p.num_particles)
    rho0 = float(p.rest_density)
    with Fluid(p, parts) as f:
        tuning = f.tune(dt, reps=2)
    variants = (tuning["density_variant"], tuning["update_variant"], tuning["neighbour_list"])

    def fluid(state=parts):
        f = Fluid(p, state).setup()
        f.track_ids(True)
        f.set_tuning(*variants)
        return f

    res = {"workload": workload, "particles": n, "dt": dt, "horizon_steps": horizon, "judge_tol": judge_tol,
           "smooth": bool(smooth), "tuning": [int(v) for v in variants]}
    with fluid() as f:
        f.step(dt, 1)  # captures the step graph
        f.synchronize()
        t0 = time.perf_counter()
        f.step(dt, horizon - 1)
        f.synchronize()
        t_plain = (time.perf_counter() - t0) * horizon / (horizon - 1)
    jump_ref, peak_ref, ref = _sampled_plain(fluid, dt, 1, horizon, rho0)
    res["plain"] = {"wall_s": t_plain, "sim_s_per_s": horizon * dt / t_plain, "particle_steps_per_s": n * horizon / t_plain,
                    "max_step_jump": jump_ref, "peak_density": peak_ref}
    log(f"plain dt={dt}: {json.dumps(res['plain'])}")

    def errors(rho, jump_name, jump, peak):
        err = np.abs(rho - ref) / rho0
        return {jump_name: jump, "peak_density": peak, "peak_excess": peak - peak_ref,
                "mean_density_err": float(err.mean()), "p99_density_err": float(np.percentile(err, 99)),
                "max_density_err": float(err.max())}

    for K in ks:
        macro = horizon // K
        entry = {}
        with fluid() as f:
            f.step(dt * K, 1)  # captures the step graph
            f.synchronize()
            t0 = time.perf_counter()
            f.step(dt * K, macro - 1)
            f.synchronize()
            t_large = (time.perf_counter() - t0) * macro / max(1, macro - 1)
        jump, peak, rho = _sampled_plain(fluid, dt, K, macro, rho0)
        entry["plain_large"] = {**errors(rho, "max_step_jump", jump, peak), "wall_s": t_large, "sim_s_per_s": macro * K * dt / t_large,
                                "speedup_vs_plain": (macro * K * dt / t_large) / res["plain"]["sim_s_per_s"]}
        for m, (use_net, sub) in METHODS.items():
            if m == "plain_large" or (use_net and net is None):
                continue
            with fluid() as f:
                st = _closure_run(f, net if use_net else None, dt, K, macro, judge_tol, smooth, sub)
                rho = _densities_by_id(f)
            wall = st.wall_ms * 1e-3
            entry[m] = {**errors(rho, "max_accepted_jump", float(st.max_accepted_jump), float(st.peak_density)),
                        "macro_steps": int(st.macro_steps), "fallbacks": int(st.fallbacks), "wall_s": wall,
                        "sim_s_per_s": macro * K * dt / wall,
                        "speedup_vs_plain": (macro * K * dt / wall) / res["plain"]["sim_s_per_s"],
                        "max_density_jump": float(st.max_density_jump)}
        if exact:
            jump, peak, rho = exact_correction_run(fluid, p, dt, K, macro, rho0)
            entry["exact_correction"] = errors(rho, "max_step_jump", jump, peak)
        if net is not None:
            entry["one_step"] = one_step_errors(net, p, parts, dt, K, variants=variants, smooth=smooth)
        res[f"K{K}"] = entry
        log(f"K={K}: {json.dumps({k: v for k, v in entry.items() if k != 'one_step'})}")
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("train")
    a.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    a.add_argument("--trajectories", type=int, default=200)
    a.add_argument("--steps", type=int, default=400)
    a.add_argument("--every", type=int, default=20)
    a.add_argument("--per-snapshot", type=int, default=1024)
    a.add_argument("--epochs", type=int, default=8)
    a.add_argument("--batch", type=int, default=8192)
    a.add_argument("--lr", type=float, default=2e-3)
    a.add_argument("--ks", default="2,4", help="step multipliers K of the training targets, comma-separated")
    a.add_argument("--log", default=None)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", default=str(DEFAULT_WEIGHTS))
    e.add_argument("--horizon", type=int, default=500)
    e.add_argument("--ks", default="2,4,8", help="step multipliers K to compare, comma-separated")
    e.add_argument("--exact", action="store_true", help="also the exact-correction upper bound")
    e.add_argument("--raw", action="store_true", help="each particle's own correction (no averaging)")
    e.add_argument("--judge-tol", type=float, default=0.01,
                   help="largest accepted rise of the density maximum / rho0, per small step of dt")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if args.cmd == "train":
        t0 = time.perf_counter()
        train_ks = tuple(int(k) for k in args.ks.split(","))
        x, t = make_training_set(args.trajectories, args.steps, args.every, args.per_snapshot, ks=train_ks, log=log)
        log(f"training set: {len(x)} samples in {time.perf_counter() - t0:.1f} s")
        net, info = train(x, t, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
        Path(args.out).parent.mkdir(parents=True, exist_ok=True)
        net.save(args.out)
        info["trajectories"] = args.trajectories
        info["steps_per_trajectory"] = args.steps
        info["ks"] = list(train_ks)
        info["wall_s"] = time.perf_counter() - t0
        log(json.dumps(info))
        if args.log:
            Path(args.log).write_text(json.dumps(info, indent=1))
        ev = evaluate(net, log=log)
        log(json.dumps(ev))
    else:
        net = ClosureNet(max_batch=1024).load(args.weights)
        ks = tuple(int(k) for k in args.ks.split(","))
        log(json.dumps(evaluate(net, horizon=args.horizon, ks=ks, judge_tol=args.judge_tol, smooth=not args.raw,
                                exact=args.exact, log=log)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
