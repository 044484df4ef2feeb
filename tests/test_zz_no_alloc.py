"""GPU test of mft_fts_partition_function_noalloc, the allocation-free
variant of mft_fts_partition_function: through the C ABI, on two calls in a
row (the second on a smaller grid, so stale partial sums of the first call
sit in the shared scratch beyond the ones the second reads), its Q must equal
the existing entry point's bit for bit (both add the same FP64 block partial
sums in the same order); a grid beyond the 2^24-point limit is rejected
before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _q_with_work(mft, qb, M, gvol, V):
    work = mft.DeviceBuffer(32 * ((M + 255) // 256))
    out = (C.c_double * 2)()
    mft.check(mft.lib().mft_fts_partition_function(out, qb.ptr, C.c_double(gvol), C.c_double(V), C.c_long(M),
                                                   work.ptr, None))
    return complex(out[0], out[1])


def _q_noalloc(mft, qb, M, gvol, V):
    out = (C.c_double * 2)()
    mft.check(mft.lib().mft_fts_partition_function_noalloc(out, qb.ptr, C.c_double(gvol), C.c_double(V),
                                                           C.c_long(M), None))
    return complex(out[0], out[1])


@pytest.mark.parametrize("shapes", [((96, 96, 96), (24, 20, 15)), ((64, 48, 40), (5, 7, 3))])
def test_noalloc_partition_function_equals_work_buffer_variant(gpu, shapes):
    rng = np.random.default_rng(17)
    for call, shape in enumerate(shapes):
        M = int(np.prod(shape))
        print(f"case: noalloc Q call {call} shape {shape} M {M}", flush=True)
        q = (rng.normal(1.0, 0.3, M) + 1j * rng.normal(0.0, 0.2, M)).astype(np.complex64)
        qb = gpu.DeviceBuffer.from_array(q)
        gvol, V = 0.37, 0.37 * M
        ref = _q_with_work(gpu, qb, M, gvol, V)
        got = _q_noalloc(gpu, qb, M, gvol, V)
        assert got == ref, (call, got, ref)
        exact = np.sum(q.astype(np.complex128)) * gvol / V
        assert abs(got - exact) < 1e-12 * abs(exact)
        qb.free()


def test_noalloc_partition_function_rejects_grids_beyond_its_scratch(gpu):
    print("case: noalloc Q size limit", flush=True)
    qb = gpu.DeviceBuffer(8 * 256)
    out = (C.c_double * 2)()
    rc = gpu.lib().mft_fts_partition_function_noalloc(out, qb.ptr, C.c_double(1.0), C.c_double(1.0),
                                                      C.c_long((1 << 24) + 1), None)
    assert rc != 0
    assert b"allocation-free limit" in gpu.lib().mft_last_error()
    qb.free()
