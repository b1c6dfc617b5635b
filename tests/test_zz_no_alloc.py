"""desv_norm2_device, the entry point that allocates nothing (its partials
live in a fixed device array), against desv_norm2 through the C ABI: two
calls in a row, the second on a smaller state, in both precisions; the
result lands in caller-provided device memory."""
import ctypes as C

import numpy as np
import pytest

import desv
from desv import _capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("complex128", [False, True])
def test_norm2_device_matches_norm2(complex128):
    torch = pytest.importorskip("torch")
    out = torch.full((2,), -1.0, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()            # the fill is done before the engine's stream writes
    rng = np.random.default_rng(7)
    for i, n in enumerate((22, 9)):     # the second call on a smaller state
        h = desv.StateHandle(n, complex128)
        v = 1.7 * (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n))
        h.upload(v)
        ptr = out.data_ptr() + 8 * i
        _capi.check(_capi.lib.desv_norm2_device(h.ptr, C.c_void_p(ptr)))
        h.synchronize()
        got = float(out[i].item())
        want = h.norm2()
        assert got == want, (n, got, want)
        v_state = v.astype(h.dtype).astype(np.complex128)   # the amplitudes as stored, summed in FP64
        ref = np.vdot(v_state, v_state).real
        assert abs(got - ref) <= 1e-9 * ref
        h.close()
    with pytest.raises(ValueError):
        _capi.check(_capi.lib.desv_norm2_device(desv.StateHandle(3, complex128).ptr, None))
