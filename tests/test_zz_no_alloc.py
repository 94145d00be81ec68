"""mom_gpu_tracer_content_noalloc (the allocation-free variant of
mom_gpu_tracer_content, partial sums in the library's fixed per-GPU array)
called through the C-ABI: its device result equals, bit for bit, the host
result of mom_gpu_tracer_content / _f32 on the same tracer, on two calls in a
row whose second domain is smaller (the second call's partials overwrite the
first's), FP64 and FP32, and both agree with the CPU restatement's
long-double sum to 1e-15."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _content_pair(case, T, dtype):
    from mom5_amd import DeviceArray, OceanGPU
    g = OceanGPU(case.dom, case.grid)
    try:
        dT = DeviceArray.from_host(T.astype(dtype))
        rho = DeviceArray.from_host(case.fields["rho_dzt"])
        out = DeviceArray((1,), np.float64)
        out.upload(np.array([np.nan]))
        print(f"noalloc content: {case.ni}x{case.nj}x{case.nk}, dtype {np.dtype(dtype).name}",
              flush=True)
        rc = g.L.mom_gpu_tracer_content_noalloc(g.ctx, dT.ptr, np.dtype(dtype).itemsize,
                                                 rho.ptr, out.ptr)
        assert rc == 0, g.L.mom_gpu_last_error(g.ctx)
        g.sync()
        got = float(out.to_host()[0])
        want = g.tracer_content(dT, rho)
        return got, want
    finally:
        g.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_content_equals_allocating_entry_point(oracle_lib, dtype):
    from mom5_amd.synth import make_case
    big = make_case(150, 90, 14, nts=0, n_passive=0, barotropic=False, seed=31)
    small = make_case(37, 23, 7, nts=0, n_passive=0, barotropic=False, seed=32)
    for case in (big, small):                # second call: smaller input
        T = case.tracers[0][1]
        got, want = _content_pair(case, T, dtype)
        assert np.float64(got).tobytes() == np.float64(want).tobytes(), (got, want)
        ref = oracle_lib.Oracle(case).tracer_content(T.astype(dtype).astype(np.float64))
        assert abs(got - ref) <= 1e-15 * abs(ref), (got, ref)


def test_noalloc_rejects_bad_dtype():
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.synth import make_case
    case = make_case(20, 12, 4, nts=0, n_passive=0, barotropic=False, seed=33)
    g = OceanGPU(case.dom, case.grid)
    try:
        out = DeviceArray((1,), np.float64)
        T = DeviceArray.from_host(case.tracers[0][1])
        rc = g.L.mom_gpu_tracer_content_noalloc(g.ctx, T.ptr, 2, T.ptr, out.ptr)
        assert rc != 0
        assert b"dtype" in g.L.mom_gpu_last_error(g.ctx)
    finally:
        g.close()
