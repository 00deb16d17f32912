# SPDX-License-Identifier: Apache-2.0
"""The world step's OUT_OF_MEMORY contract (world.h:161-166: "the step is
abandoned between stages; the world stays consistent but is not
advanced"), by failure injection in the manner of the reference's
test_sensor.c:390-440: the world's allocator fails after a budget of
allocations, swept from 0 upwards, during a step that has to grow the
world's arrays (forty bodies teleported into a pile make new pairs and
contacts). A budget that fails the step before the solve must leave the
bodies as they were and the world valid, and the step retried with memory
must give exactly the world of a twin that never failed (hash, pair and
contact-id lists, body states). One that fails it after the solve (the
islands' bookkeeping) leaves the bodies where the twin's are, and the next
step finishes the work. Either way the two stay equal for 30 more steps,
and the step succeeds once the budget covers it."""
import ctypes as C

import numpy as np
import pytest

from nepheline_amd import physics as P
from nepheline_amd import scenes as S

pytestmark = pytest.mark.gpu


class Callsite(C.Structure):
    _fields_ = [("file", C.c_char_p), ("line", C.c_uint32)]


AllocFn = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, Callsite)
ReallocFn = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32,
                        Callsite)
FreeFn = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, Callsite)


class Vtable(C.Structure):
    _fields_ = [("alloc", AllocFn), ("realloc", ReallocFn), ("free", FreeFn)]


class Allocator(C.Structure):
    _fields_ = [("vtable", C.POINTER(Vtable)), ("name", C.c_char_p)]


_libc = C.CDLL(None)
_libc.posix_memalign.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_size_t]
_libc.posix_memalign.restype = C.c_int
_libc.free.argtypes = [C.c_void_p]
_libc.memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
_libc.memcpy.restype = C.c_void_p


class FailingAllocator:
    """libc memory; with `budget` set, the allocations (alloc and realloc)
    past the first `budget` fail."""

    def __init__(self):
        self.budget = None
        self.calls = 0
        self.denied = None  # the first refused call's site

        def grant(site):
            self.calls += 1
            ok = self.budget is None or self.calls <= self.budget
            if not ok and self.denied is None:
                self.denied = "%s:%d" % (site.file.decode() if site.file else "?", site.line)
            return ok

        def alloc(_self, size, align, site):
            if not grant(site):
                return None
            p = C.c_void_p()
            return p.value if _libc.posix_memalign(C.byref(p), max(align, 16), max(size, 1)) == 0 \
                else None

        def realloc(_self, ptr, old, new, align, site):
            if not grant(site):
                return None
            p = C.c_void_p()
            if _libc.posix_memalign(C.byref(p), max(align, 16), max(new, 1)) != 0:
                return None
            if ptr:
                _libc.memcpy(p, ptr, min(old, new))
                _libc.free(ptr)
            return p.value

        def free(_self, ptr, _size, _align, _site):
            if ptr:
                _libc.free(ptr)

        self._fns = (AllocFn(alloc), ReallocFn(realloc), FreeFn(free))
        self.vtable = Vtable(*self._fns)
        self.allocator = Allocator(C.pointer(self.vtable), b"failing")

    def address(self):
        return C.addressof(self.allocator)

    def fail_after(self, budget):
        self.calls = 0
        self.denied = None
        self.budget = budget


def _scene(api, fa=None):
    desc = {"alloc": fa.address()} if fa is not None else None
    w = S.bin_drop(api, 160, 41, None, half=4.0, height=8.0, y_lo=1.0, desc=desc)
    if fa is not None:
        w._keep.append(fa)  # the allocator outlives the world
    return w


def _pile(w):
    """Forty bodies put into a tight pile: new pairs and contacts."""
    for i in range(40):
        b = w.bodies[5 + i]  # past the bin's five static bodies
        x = (i % 4) * 0.45 - 0.7
        y = 1.0 + (i // 16) * 0.45
        z = ((i // 4) % 4) * 0.45 - 0.7
        w.api.body_set_transform(w.ptr, b, P.DVec3(x, y, z), P.Quat(0.0, 0.0, 0.0, 1.0))


def _same(a, b):
    assert a.hash() == b.hash()
    pa, ia, _ = P.contacts_as_arrays(a.contacts())
    pb, ib, _ = P.contacts_as_arrays(b.contacts())
    assert pa == pb and ia == ib
    np.testing.assert_array_equal(a.body_state(), b.body_state())


def test_a_step_out_of_memory_is_abandoned_and_retried(gpu_api):
    early = late = 0
    for budget in range(0, 96):
        print(f"out of memory: budget {budget}", flush=True)
        fa = FailingAllocator()
        w = _scene(gpu_api, fa)
        twin = _scene(gpu_api)
        for _ in range(20):
            assert w.api.world_step(w.ptr, S.DT, None) == P.OK
            twin.step(S.DT)
        _pile(w)
        _pile(twin)
        before = w.body_state()
        fa.fail_after(budget)
        r = w.api.world_step(w.ptr, S.DT, None)
        fa.budget = None
        twin.step(S.DT)
        if r == P.OK:
            _same(w, twin)
            break
        print(f"  refused at {fa.denied}", flush=True)
        assert r == P.OUT_OF_MEMORY
        if np.array_equal(w.body_state(), before):
            # Abandoned before the solve: not advanced, consistent, and the
            # retry is the step the twin took.
            early += 1
            assert w.api.world_validate(w.ptr)
            assert w.api.world_step(w.ptr, S.DT, None) == P.OK
        else:
            # Abandoned after the solve, in the bookkeeping that follows it
            # (the reference's islands stage, world.c:350-361, fails there
            # the same way): the bodies moved as the twin's did, and the
            # next step finishes the work.
            late += 1
            np.testing.assert_array_equal(w.body_state(), twin.body_state())
            assert w.api.world_step(w.ptr, S.DT, None) == P.OK
            twin.step(S.DT)
        _same(w, twin)
        assert w.api.world_validate(w.ptr)
        for _ in range(30):
            w.step(S.DT)
            twin.step(S.DT)
        _same(w, twin)
        w.close()
        twin.close()
    else:
        pytest.fail("the step never succeeded")
    print(f"failed before the solve at {early} budgets, after it at {late}", flush=True)
    assert early > 0
