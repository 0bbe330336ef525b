"""The allocation-free field entry point (fk_field_noalloc) against fk_field.

Device memory for its inputs and outputs comes from the HIP runtime the
library itself uses (hipMalloc / hipMemcpy through ctypes); the entry point
allocates nothing and keeps its energy partials in a static per-GPU scratch.
Two calls in a row, the second on a smaller mesh, each compared with fk_field
bit for bit (the same kernel, the energy summed in the same order)."""
import ctypes

import numpy as np
import pytest

from femeko_hip import _lib
from femeko_hip import workloads as W
from femeko_hip.device import DeviceMesh, context

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    _lib.load()  # the HIP runtime libfemeko_hip.so links (soname libamdhip64.so.7) is then loaded
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


class DevArray:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(nbytes, 8)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, _H2D) == 0

    def get(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, _D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _one_call(hip, n, seed, scattered=False):
    pr = W.linear(n, seed=seed)
    if scattered:
        # a node numbering the library renumbers inside: phi arrives in the
        # caller's order and is gathered into the internal one
        pr = W.renumbered(pr, np.random.default_rng(seed).permutation(pr.mesh.nv), "random")
    ctx = context(0)
    L = _lib.load()
    dm = DeviceMesh(pr.mesh, ctx)
    bufs = []
    try:
        assert dm.numbering()[0] == scattered
        dm.set_coefficient(pr.mu)
        dm.assemble()
        dm.set_dirichlet(pr.mask, pr.g)
        dm.solve("f64", rtol=1e-10)
        phi = dm.potential()
        H_ref, Hn_ref, E_ref = dm.field(want_H=True, want_Hn=True, want_energy=True)
        d_phi, d_H, d_Hn, d_E = (DevArray(hip, 8 * k) for k in (dm.nv, 3 * dm.nt, dm.nt, 1))
        bufs = [d_phi, d_H, d_Hn, d_E]
        d_phi.put(phi)
        s = L.fk_field_noalloc(dm.handle, d_phi.ptr, d_H.ptr, d_Hn.ptr, d_E.ptr)
        ctx.check(s, "fk_field_noalloc")
        ctx.synchronize()
        H = d_H.get((dm.nt, 3)).T
        Hn = d_Hn.get(dm.nt)
        E = float(d_E.get(1)[0])
        assert np.array_equal(H, H_ref)
        assert np.array_equal(Hn, Hn_ref)
        assert E == E_ref
        # outputs may be left out
        s = L.fk_field_noalloc(dm.handle, d_phi.ptr, None, None, d_E.ptr)
        ctx.check(s, "fk_field_noalloc (energy only)")
        ctx.synchronize()
        assert float(d_E.get(1)[0]) == E_ref
    finally:
        for b in bufs:
            b.free()
        dm.close()
    return dm.nt


def test_noalloc_field_matches_fk_field_twice():
    hip = _hip()
    big = _one_call(hip, 24, 5)
    small = _one_call(hip, 9, 6)  # a smaller input after a larger one
    assert small < big


def test_noalloc_field_rejects_missing_potential():
    ctx = context(0)
    L = _lib.load()
    pr = W.linear(4)
    dm = DeviceMesh(pr.mesh, ctx)
    try:
        assert L.fk_field_noalloc(dm.handle, None, None, None, None) == _lib.FK_ERR_INVALID_ARG
    finally:
        dm.close()


def test_noalloc_field_on_a_renumbered_mesh():
    hip = _hip()
    _one_call(hip, 10, 8, scattered=True)
