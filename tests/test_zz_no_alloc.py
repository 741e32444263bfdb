"""fb_mlp_gradients_noalloc: the allocation-free variant of fb_mlp_gradients
(work arrays in the library's fixed scratch array, every argument in device
memory), called through the C ABI and checked against fb_mlp_gradients on
the same batches: gradients bit-identical (the same kernels on the same
inputs), loss to FP32 rounding of the sum.  Two calls in a row, the second
with a smaller batch, so stale scratch contents from the first would show;
a batch above FB_MLP_NOALLOC_MAX_BATCH is refused.  Device buffers come from
the HIP runtime through ctypes."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP = 8576
MAXB = 2048
H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class DeviceArrays:
    """hipMalloc'd float32 buffers, freed on exit."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptrs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        for p in self.ptrs:
            self.hip.hipFree(ctypes.c_void_p(p))

    def alloc(self, n):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(4 * n)) == 0
        self.ptrs.append(p.value)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = self.alloc(a.size)
        assert self.hip.hipMemcpy(ctypes.c_void_p(p), a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes),
                                  H2D) == 0
        return p

    def filled(self, n, value):
        return self.put(np.full(n, value, np.float32))

    def get(self, p, n):
        a = np.empty(n, np.float32)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p), ctypes.c_size_t(a.nbytes),
                                  D2H) == 0
        return a


def _asm():
    from fireball_hip import synth
    from fireball_hip.lib import Assembler
    return Assembler(synth.make_tables([synth.SI]), precision="mixed")


def _batch(rng, n):
    X = rng.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32)
    T = rng.normal(0.0, 0.5, size=(n, 64)).astype(np.float32)
    return X, T


def test_noalloc_gradients_match_allocating_entry_point():
    asm = _asm()
    rng = np.random.default_rng(21)
    params = (rng.standard_normal(NP) * 0.2).astype(np.float32)
    with DeviceArrays() as d:
        dp = d.put(params)
        for n, nme in ((1800, 10), (700, 3)):
            print(f"noalloc batch n={n} nme={nme}", flush=True)
            X, T = _batch(rng, n)
            dX, dT = d.put(X), d.put(T)
            # poison the outputs: every value must be written by the call
            dg, dl = d.filled(NP, np.nan), d.filled(1, np.nan)
            asm.mlp_gradients_noalloc(nme, dp, n, dX, dT, dg, dl)
            asm.synchronize()
            g = d.get(dg, NP)
            loss = float(d.get(dl, 1)[0])
            loss_ref, g_ref = asm.mlp_gradients(nme, params, X, T)
            assert np.all(np.isfinite(g))
            assert np.array_equal(g, g_ref), float(np.max(np.abs(g - g_ref)))
            assert abs(loss - loss_ref) <= 1e-5 * abs(loss_ref), (loss, loss_ref)


def test_noalloc_refuses_batches_above_the_maximum():
    from fireball_hip.lib import FireballError
    asm = _asm()
    n = MAXB + 1
    with DeviceArrays() as d:
        dp, dX, dT = d.filled(NP, 0.0), d.filled(n * 3, 0.0), d.filled(n * 64, 0.0)
        dg, dl = d.filled(NP, 0.0), d.filled(1, 0.0)
        with pytest.raises(FireballError) as e:
            asm.mlp_gradients_noalloc(10, dp, n, dX, dT, dg, dl)
        assert e.value.code == -10
