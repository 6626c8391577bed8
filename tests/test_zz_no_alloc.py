"""The allocation-free field entry point (bem_ext_pressure_noalloc) against
the allocating one it varies (bem_ext_pressure, genExtPressure mesh.cu:1038-1073)
and the oracle, through the C-ABI, on device buffers the test allocates with
the HIP runtime.  Two calls in a row, the second with fewer points, on the
library's stream and on a stream of the caller's."""
import ctypes

import numpy as np
import pytest

import bem_gpu_clean_cpp_amd as bem
import bemcases as C
import oracle_py as O

pytestmark = pytest.mark.gpu

RHO, SPEED = float(np.float32(1.21)), float(np.float32(343.0))


class _Hip:
    """The few HIP runtime calls the test needs as the caller (libbem links
    the same runtime, so these are the library's device and context)."""

    def __init__(self):
        L = ctypes.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.L = L

    def check(self, rc, what):
        assert rc == 0, "%s failed with HIP error %d" % (what, rc)

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.L.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        return p

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        p = self.malloc(a.nbytes)
        self.check(self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1), "hipMemcpy H2D")
        return p

    def to_host(self, p, out):
        self.check(self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2), "hipMemcpy D2H")
        return out


def test_ext_pressure_noalloc_matches_ext_pressure():
    bem.quad_init(6)
    O.quad_init(6)
    bem.set_medium(RHO, SPEED)
    bem.set_source_strength(1.0)
    hip = _Hip()
    bufs = []
    stream = ctypes.c_void_p()
    try:
        m = bem.Mesh().icosphere(3, 0.1)
        k = float(np.float32(2 * np.pi * 1000.0 / SPEED))
        rng = np.random.default_rng(17)
        N = m.num_pnts
        sp = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
        xs = C.sphere_points(18, 50, 0.4)
        src = [0.02, -0.01, 0.03]
        # the allocating entry point, one point at a time (makes the mesh resident)
        ref = np.array([bem.ext_pressure(k, m, src, x, sp) for x in xs], np.complex64)
        # ... and the oracle (incident point source of strength 1 at src)
        p, e, bc = C.mesh_arrays(m)
        d = np.linalg.norm(xs.astype(np.float64) - np.asarray(src, np.float32).astype(np.float64), axis=1)
        inc = np.exp(-1j * k * d) / (4 * np.pi * d)
        fo = O.field(k, p.astype(np.float64), N, e, bc, xs, sp, inc=inc[:, None], dens=RHO, speed=SPEED)[:, 0]
        err = np.max(np.abs(ref - fo)) / np.max(np.abs(fo))
        print("bem_ext_pressure vs oracle: %.3g" % err)
        assert err < 2e-5
        xs_d, sp_d = hip.to_device(xs.astype(np.float32)), hip.to_device(sp)
        out_d = hip.malloc(8 * len(xs))
        bufs += [xs_d, sp_d, out_d]
        # a mesh no allocating call has made resident is refused
        fresh = bem.Mesh().icosphere(2, 0.1)
        with pytest.raises(bem.BemError, match="not resident"):
            bem.ext_pressure_noalloc(k, fresh, src, xs_d, 1, sp_d, out_d, None)
        # call 1: 50 points on the library's stream
        bem.ext_pressure_noalloc(k, m, src, xs_d, len(xs), sp_d, out_d, None)
        bem.sync()
        out1 = hip.to_host(out_d, np.zeros(len(xs), np.complex64))
        print("call 1 (50 points): max |noalloc - ext_pressure| = %.3g" % np.max(np.abs(out1 - ref)))
        assert np.array_equal(out1, ref)
        # call 2: 7 points on the caller's stream, after the scratch held call 1's data
        hip.check(hip.L.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
        out_d2 = hip.malloc(8 * 7)
        bufs.append(out_d2)
        bem.ext_pressure_noalloc(k, m, src, ctypes.c_void_p(xs_d.value + 12 * 10), 7, sp_d, out_d2, stream)
        hip.check(hip.L.hipStreamSynchronize(stream), "hipStreamSynchronize")
        out2 = hip.to_host(out_d2, np.zeros(7, np.complex64))
        print("call 2 (7 points): max |noalloc - ext_pressure| = %.3g" % np.max(np.abs(out2 - ref[10:17])))
        assert np.array_equal(out2, ref[10:17])
        # beyond the scratch array's size: refused with a message, nothing runs
        with pytest.raises(bem.BemError, match="scratch"):
            bem.ext_pressure_noalloc(k, m, src, xs_d, 10 ** 6, sp_d, out_d, None)
    finally:
        bem.sync()
        if stream.value:
            hip.L.hipStreamDestroy(stream)
        for b in bufs:
            hip.L.hipFree(b)
        bem.set_source_strength(0.0)
        bem.set_medium(1.2041, 343.21)
