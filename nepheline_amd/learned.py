# SPDX-License-Identifier: Apache-2.0
"""The learned warm start: logging, training and using the contact network.

The network (nepheline_amd/csrc/mlp_hip.hip) maps a contact point's 14
features at the step's start - approach speed along the normal and both
tangents, separation, normal, the two inverse masses, friction,
normal mass and the tangent mass terms - to the impulses the plain solver
converges to (normal and two friction impulses, divided by the normal
mass).  Everything numeric runs in the library on the device: the world logs
features and converged impulses (nph_amd_world_log_contacts), the trainer
normalises them and runs minibatch Adam with the network's own forward,
gradient and Adam kernels (nph_amd_trainer_*), and a world takes the trained
network (nph_amd_world_set_warm_start).  This module is the ctypes binding
and the driver around those calls; device memory is the library's own
(nph_amd_device_alloc), so one HIP runtime serves the process.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import physics as P
from . import scenes as S

FEATURES = 14
INPUTS = 16
OUTPUTS = 4
PARAMS = 6288
NEW_POINTS = 0
ALL_POINTS = 1

_U32 = C.c_uint32
_F32 = C.c_float
_VP = C.c_void_p
_FP = C.POINTER(C.c_float)


class WarmStart(C.Structure):
    _fields_ = [("params", _FP), ("feature_shift", C.c_float * 16),
                ("feature_scale", C.c_float * 16), ("output_scale", C.c_float * 4),
                ("mode", _U32), ("gate", C.c_bool)]


class WarmStats(C.Structure):
    _fields_ = [("residual_max_plain", C.c_float), ("residual_max_learned", C.c_float),
                ("residual_sum_plain", C.c_float), ("residual_sum_learned", C.c_float),
                ("penetration_max", C.c_float), ("contacts_plain", _U32),
                ("contacts_learned", _U32), ("gated", _U32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _bind(a: P.Api):
    L = a.lib
    if getattr(a, "_learned", None) is not None:
        return a._learned

    class F:
        pass

    f = F()

    def b(name, res, args):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
        setattr(f, name.replace("nph_amd_", ""), fn)

    b("nph_amd_mlp_forward", C.c_int, [_VP, _U32, _VP, _VP, _VP])
    b("nph_amd_mlp_grad", C.c_int, [_VP, _VP, _U32, _VP, _F32, _VP, _VP, _VP])
    b("nph_amd_mlp_adam", C.c_int, [_VP, _VP, _VP, _VP, _U32, _F32, _VP])
    b("nph_amd_mlp_gelu", C.c_int, [_VP, _U32, _VP, _VP])
    b("nph_amd_device_alloc", _VP, [C.c_uint64])
    b("nph_amd_device_free", None, [_VP])
    b("nph_amd_device_copy", C.c_int, [_VP, _VP, C.c_uint64])
    b("nph_amd_world_set_warm_start", C.c_int, [_VP, C.POINTER(WarmStart)])
    b("nph_amd_world_measure_residual", None, [_VP, C.c_bool])
    b("nph_amd_world_warm_stats", C.c_int, [_VP, C.POINTER(WarmStats)])
    b("nph_amd_world_log_contacts", C.c_int, [_VP, _U32])
    b("nph_amd_world_contact_log", C.c_int,
      [_VP, C.POINTER(_U32), C.POINTER(_VP), C.POINTER(_VP)])
    b("nph_amd_trainer_create", C.c_int, [_U32, C.c_uint64, C.POINTER(_VP)])
    b("nph_amd_trainer_destroy", None, [_VP])
    b("nph_amd_trainer_add", C.c_int, [_VP, _VP, _VP, _U32, _F32])
    b("nph_amd_trainer_rows", _U32, [_VP])
    b("nph_amd_trainer_prepare", C.c_int, [_VP])
    b("nph_amd_trainer_epoch", C.c_int, [_VP, _U32, _F32, C.POINTER(_F32)])
    b("nph_amd_trainer_loss", C.c_int, [_VP, C.POINTER(_F32)])
    b("nph_amd_trainer_warm_start", C.c_int, [_VP, _FP, C.POINTER(WarmStart)])
    a._learned = f
    return f


def _check(r, what):
    if r != P.OK:
        raise RuntimeError(f"{what} failed: {r}")


class DeviceArray:
    """A device buffer holding a numpy array's bytes (the library's memory)."""

    def __init__(self, a: P.Api, host: np.ndarray | None = None, shape=None, dtype=None):
        self.f = _bind(a)
        if host is not None:
            host = np.ascontiguousarray(host)
            shape, dtype = host.shape, host.dtype
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = self.f.device_alloc(max(self.nbytes, 4))
        if not self.ptr:
            raise MemoryError(f"nph_amd_device_alloc({self.nbytes}) failed")
        if host is not None:
            self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes == self.nbytes
        _check(self.f.device_copy(self.ptr, host.ctypes.data, self.nbytes), "device copy")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self.f.device_copy(out.ctypes.data, self.ptr, self.nbytes), "device copy")
        return out

    def free(self):
        if self.ptr:
            self.f.device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Network:
    """The network's kernels on device buffers, for direct use and tests."""

    def __init__(self, a: P.Api, params: np.ndarray):
        self.api = a
        self.f = _bind(a)
        self.params = DeviceArray(a, np.asarray(params, dtype=np.float32))
        self.grad = DeviceArray(a, np.zeros(PARAMS, np.float32))
        self.m = DeviceArray(a, np.zeros(PARAMS, np.float32))
        self.v = DeviceArray(a, np.zeros(PARAMS, np.float32))
        self.loss = DeviceArray(a, np.zeros(1, np.float32))
        self.t = 0

    def forward(self, xbits: DeviceArray) -> np.ndarray:
        """xbits: (n, 16) BF16 bit patterns on the device -> (n, 4) float32."""
        n = xbits.shape[0]
        y = DeviceArray(self.api, shape=(n, OUTPUTS), dtype=np.float32)
        _check(self.f.mlp_forward(xbits.ptr, n, self.params.ptr, y.ptr, None), "mlp forward")
        return y.numpy()

    def gradient(self, xbits: DeviceArray, target: DeviceArray, scale: float):
        """Fills self.grad and self.loss for one batch (both zeroed first)."""
        self.grad.upload(np.zeros(PARAMS, np.float32))
        self.loss.upload(np.zeros(1, np.float32))
        _check(self.f.mlp_grad(xbits.ptr, target.ptr, xbits.shape[0], self.params.ptr,
                               float(scale), self.grad.ptr, self.loss.ptr, None), "mlp grad")

    def adam(self, lr: float):
        self.t += 1
        _check(self.f.mlp_adam(self.params.ptr, self.grad.ptr, self.m.ptr, self.v.ptr, self.t,
                               float(lr), None), "mlp adam")


class Trainer:
    """The library's trainer: a device training set and Adam state."""

    def __init__(self, a: P.Api, capacity: int, seed: int = 0):
        self.f = _bind(a)
        p = _VP()
        _check(self.f.trainer_create(capacity, seed, C.byref(p)), "trainer create")
        self.ptr = p

    def add_log(self, world: P.World, min_flag: float = 0.0):
        rows, fx, tx = contact_log(world)
        _check(self.f.trainer_add(self.ptr, fx, tx, rows, float(min_flag)), "trainer add")
        return rows

    @property
    def rows(self):
        return int(self.f.trainer_rows(self.ptr))

    def prepare(self):
        _check(self.f.trainer_prepare(self.ptr), "trainer prepare")

    def epoch(self, batch=4096, lr=2e-3):
        loss = C.c_float()
        _check(self.f.trainer_epoch(self.ptr, batch, lr, C.byref(loss)), "trainer epoch")
        return loss.value

    def loss(self):
        loss = C.c_float()
        _check(self.f.trainer_loss(self.ptr, C.byref(loss)), "trainer loss")
        return loss.value

    def warm_start(self, mode=NEW_POINTS, gate=True):
        params = np.zeros(PARAMS, np.float32)
        ws = WarmStart()
        _check(self.f.trainer_warm_start(self.ptr, params.ctypes.data_as(_FP), C.byref(ws)),
               "trainer warm start")
        ws.mode = mode
        ws.gate = gate
        return ws, params

    def close(self):
        if self.ptr:
            self.f.trainer_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_warm_start(world: P.World, ws: WarmStart | None, params: np.ndarray | None = None):
    f = _bind(world.api)
    if ws is None:
        _check(f.world_set_warm_start(world.ptr, None), "set warm start")
        return
    if params is not None:
        ws.params = np.ascontiguousarray(params, dtype=np.float32).ctypes.data_as(_FP)
        world._keep.append(params)
    _check(f.world_set_warm_start(world.ptr, C.byref(ws)), "set warm start")


def measure_residual(world: P.World, enable=True):
    _bind(world.api).world_measure_residual(world.ptr, enable)


def warm_stats(world: P.World) -> WarmStats:
    s = WarmStats()
    _check(_bind(world.api).world_warm_stats(world.ptr, C.byref(s)), "warm stats")
    return s


class BatchWarmStats(C.Structure):
    """nph_amd_batch_warm_stats: sums and maxima over worlds and steps."""
    _fields_ = [("residual_sum_plain", C.c_double), ("residual_sum_learned", C.c_double),
                ("contacts_plain", C.c_double), ("contacts_learned", C.c_double),
                ("residual_max_plain", C.c_double), ("residual_max_learned", C.c_double),
                ("penetration_max", C.c_double), ("gated", C.c_double),
                ("steps", _U32), ("pad", _U32)]


def _bind_batch(a: P.Api):
    if hasattr(a, "batch_set_warm_start"):
        return
    L = a.lib
    a.batch_set_warm_start = P._bind(L, "nph_amd_batch_set_warm_start", C.c_int,
                                     [C.c_void_p, C.POINTER(WarmStart)])
    a.batch_measure_residual = P._bind(L, "nph_amd_batch_measure_residual", None,
                                       [C.c_void_p, C.c_bool])
    a.batch_warm_stats = P._bind(L, "nph_amd_batch_warm_stats", C.c_int,
                                 [C.c_void_p, C.POINTER(BatchWarmStats), C.c_bool])


def batch_set_warm_start(batch: P.Batch, ws: WarmStart | None, params: np.ndarray | None = None):
    _bind_batch(batch.api)
    if ws is None:
        _check(batch.api.batch_set_warm_start(batch.ptr, None), "batch warm start")
        return
    if params is not None:
        params = np.ascontiguousarray(params, dtype=np.float32)
        ws.params = params.ctypes.data_as(_FP)
        batch._keep = getattr(batch, "_keep", []) + [params]
    _check(batch.api.batch_set_warm_start(batch.ptr, C.byref(ws)), "batch warm start")


def batch_measure_residual(batch: P.Batch, enable=True):
    _bind_batch(batch.api)
    batch.api.batch_measure_residual(batch.ptr, enable)


def batch_warm_stats(batch: P.Batch, reset=False) -> BatchWarmStats:
    _bind_batch(batch.api)
    s = BatchWarmStats()
    _check(batch.api.batch_warm_stats(batch.ptr, C.byref(s), reset), "batch warm stats")
    return s


def log_contacts(world: P.World, capacity: int):
    _check(_bind(world.api).world_log_contacts(world.ptr, capacity), "log contacts")


def contact_log(world: P.World):
    rows = _U32()
    fx = _VP()
    tx = _VP()
    _check(_bind(world.api).world_contact_log(world.ptr, C.byref(rows), C.byref(fx), C.byref(tx)),
           "contact log")
    return rows.value, fx, tx


def download_log(world: P.World):
    """The log's rows as numpy arrays (features (n, 16), targets (n, 4))."""
    f = _bind(world.api)
    n, fx, tx = contact_log(world)
    feats = np.empty((n, INPUTS), np.float32)
    targs = np.empty((n, OUTPUTS), np.float32)
    if n:
        _check(f.device_copy(feats.ctypes.data, fx, feats.nbytes), "device copy")
        _check(f.device_copy(targs.ctypes.data, tx, targs.nbytes), "device copy")
    return feats, targs


def log_bin_drops(a: P.Api, trainer: Trainer, bank, seeds, bodies, steps, min_flag=0.0,
                  log_capacity=4_000_000, flush_every=50, progress=None):
    """Runs plain bin drops, appending their contact logs to the trainer
    every `flush_every` steps (the log restarts empty after each)."""
    total = 0
    for seed in seeds:
        w = S.bin_drop(a, bodies, seed, bank)
        log_contacts(w, log_capacity)
        for i in range(steps):
            w.step(S.DT)
            if (i + 1) % flush_every == 0 or i + 1 == steps:
                total += trainer.add_log(w, min_flag)
                log_contacts(w, log_capacity)
        w.close()
        if progress:
            progress(f"logged seed {seed}: {trainer.rows} rows kept of {total}")
    return total


def train(a: P.Api, bank, seeds=(100, 101), bodies=2000, steps=120, epochs=6, batch=4096,
          lr=2e-3, min_flag=1.0, seed=0, capacity=None, log_capacity=4_000_000,
          progress=None):
    """Logs plain bin drops and trains a network on their converged impulses.
    Returns (WarmStart, params, history of epoch losses)."""
    tr = Trainer(a, capacity=capacity or len(seeds) * 4_000_000, seed=seed)
    log_bin_drops(a, tr, bank, seeds, bodies, steps, min_flag=min_flag,
                  log_capacity=log_capacity, progress=progress)
    tr.prepare()
    hist = [tr.loss()]
    for e in range(epochs):
        tr.epoch(batch, lr if e < epochs - 2 else lr * 0.3)
        hist.append(tr.loss())
        if progress:
            progress(f"epoch {e + 1}: loss {hist[-1]:.4f}")
    ws, params = tr.warm_start()
    tr.close()
    return ws, params, hist
