"""pmx_ow_convergence_noalloc (the convergence reductions with their
partials in the library's static per-GPU array) against pmx_ow_convergence
(handle scratch), through the C ABI: two calls in a row on two models, the
second one smaller, each equal to the allocating entry point bit for bit,
so nothing of the first call's larger reduction leaks into the second."""
import ctypes

import numpy as np
import pytest

import oracle.grid as og

from cases import build_device, small_gravity_inputs

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _reductions(md, st, entry):
    from pymrst_amd import _lib
    out7 = np.zeros(7)
    wells = np.zeros(3 * md.nw)
    _lib.check(getattr(md.lib, entry)(md.h, _lib.ptr(st.p), _lib.ptr(out7), _lib.ptr(wells)), entry)
    return out7, wells


def test_noalloc_convergence_equals_the_allocating_entry_point():
    from pymrst_amd.ow import DeviceOWState
    dev = torch.device("cuda")
    dt = 10 * og.DAY
    for dims in ((40, 30, 12), (6, 5, 4)):        # 14,400 cells, then 120
        inp = small_gravity_inputs(nx=dims[0], ny=dims[1], nz=dims[2])
        md, s0 = build_device(inp)
        rng = np.random.default_rng(sum(dims))
        st = s0.copy()
        st.pressure = st.pressure + rng.normal(0, 2e5, st.pressure.size)
        st.sW = np.clip(st.sW + rng.normal(0, 0.05, st.sW.size), 0.0, 1.0)
        d0, d = DeviceOWState.from_host(s0, dev), DeviceOWState.from_host(st, dev)
        md.assemble_device(d0, d, dt, None)
        print(f"noalloc convergence on {dims}", flush=True)
        ref7, refw = _reductions(md, d, "pmx_ow_convergence")
        got7, gotw = _reductions(md, d, "pmx_ow_convergence_noalloc")
        # maxima are order-free; the sums may be grouped differently
        assert np.array_equal(got7[:2], ref7[:2])
        assert np.allclose(got7[2:], ref7[2:], rtol=1e-12, atol=1e-12 * np.abs(ref7[2:]).max())
        assert np.array_equal(gotw, refw)
        assert got7[6] == pytest.approx(st.pressure.sum(), rel=1e-12)
