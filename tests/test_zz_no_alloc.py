"""pbm_zipf_trace_noalloc, the entry point that allocates nothing, against
pbm_zipf_trace through the C-ABI: identical traces on two calls in a row,
the second over fewer logical pages than the first (the static CDF then
holds the first call's entries past the second's n_logical), and sizes
beyond the static array rejected."""
import ctypes as C

import numpy as np
import pytest

import pbm
import pbm_oracle as orc
from conftest import log_case

pytestmark = pytest.mark.gpu


def _pair(gpu, n_logical, s, seed, first, count):
    a = gpu.full((count,), -7, dtype=gpu.int32, device="cuda")
    b = gpu.full((count,), -9, dtype=gpu.int32, device="cuda")
    pbm.zipf_trace(a, n_logical, s, seed, first=first)
    pbm.zipf_trace_noalloc(b, n_logical, s, seed, first=first)
    gpu.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_trace_equals_the_allocating_one_on_calls_in_a_row(gpu):
    # first a large table (4,194,304 pages, the buffer-cache workload's),
    # then a much smaller one over the scratch the first left behind
    for n_logical, s, seed, first, count in ((1 << 22, 1.1, 7, 0, 1 << 20),
                                             (1000, 0.8, 3, 12345, 50000)):
        log_case(f"zipf noalloc n_logical={n_logical} count={count}")
        a, b = _pair(gpu, n_logical, s, seed, first, count)
        assert np.array_equal(a, b)
        assert a.min() >= 0 and a.max() < n_logical
        # and both are the oracle's trace
        assert np.array_equal(b[:4096], orc.zipf_trace(n_logical, s, seed, first, 4096))


def test_noalloc_rejects_tables_beyond_its_scratch(gpu):
    out = gpu.zeros((16,), dtype=gpu.int32, device="cuda")
    rc = pbm.lib().pbm_zipf_trace_noalloc((1 << 22) + 1, 1.1, 7, 0, 16,
                                          C.c_void_p(out.data_ptr()), 0, None)
    assert rc == -5  # PBM_ERR_RANGE
    assert "4,194,304" in pbm.lib().pbm_last_error().decode()
