"""The allocation-free div B entry point (mhdg_div_b_noalloc) against mhdg_div_b through the C ABI: calls in a row on
two contexts, the second on a smaller mesh, with and without the extra flux, and a divergence-free flux (no worst
cell); the inputs and outputs are device arrays owned by the caller.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library runs on (already loaded with it: dlopen by soname returns that copy), for the
    caller's own device buffers."""
    from mpcns_uns_amd._lib import lib

    lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def _noalloc(dev, B, extra):
    from mpcns_uns_amd._lib import lib

    hip = _hip()
    H2D, D2H = 1, 2
    bufs = []

    def upload(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
        bufs.append(p)
        assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    try:
        Bd = upload(B)
        Xd = None if extra is None else upload(extra)
        od = upload(np.full(3, np.nan))
        rc = lib().mhdg_div_b_noalloc(dev._h, Bd, Xd, od)
        assert rc == 0, lib().mhdg_last_error(dev._h)
        dev.sync()
        out = np.zeros(3)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), od, 24, D2H) == 0
        return out
    finally:
        for p in bufs:
            hip.hipFree(p)


def test_div_b_noalloc_matches_div_b():
    from mpcns_uns_amd import Device, workloads

    for net in ((12, 48, 3), (6, 12, 2)):  # the second call on a smaller mesh
        w = workloads.build(net=net, hall=False)
        m = w["mesh"]
        print(f"\n[no-alloc] net {net}: {m.n_cells} cells", flush=True)
        dev = Device(m, w["cfg"])
        Phi, PhiAdd = w["state"][2], w["background"][0]
        rng = np.random.default_rng(3)
        B = Phi + 1e-3 * rng.normal(size=Phi.size)  # a flux with divergence
        for extra in (None, PhiAdd):
            ref = dev.div_b(B, extra)
            got = _noalloc(dev, B, extra)
            assert got[0] == ref[0] and got[1] == ref[1] and int(got[2]) == ref[2], (got, ref)
            assert ref[0] > 0 and ref[2] >= 0
        # zero flux: no cell attains a positive divergence
        got = _noalloc(dev, np.zeros_like(Phi), None)
        assert list(got) == [0.0, 0.0, -1.0]
        assert dev.div_b(np.zeros_like(Phi))[2] == -1
        dev.close()
