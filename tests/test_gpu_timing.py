"""pbm_time_next_kernel: the next batch call's main kernel records the
caller's events at its own start and end; every call disarms it."""
import numpy as np
import pytest

import pbm
from conftest import log_case

pytestmark = pytest.mark.gpu


def test_kernel_events_time_the_rank_and_free_kernels(gpu):
    T, n = 1 << 20, 262144
    log_case("kernel timing events")
    rank = pbm.launch(pbm.MODE_BATCHED, flags=pbm.LAUNCH_RANK)
    with pbm.Pool(T, seed=42) as pool:
        pool.prefill_bernoulli(0.9, 42)
        before = pool.snapshot()
        ids = gpu.full((n,), -1, dtype=gpu.int32, device="cuda")
        ev = [pbm.KernelEvent() for _ in range(4)]
        pool.time_next_kernel(ev[0], ev[1])
        pool.get_pages(n, ids, None, 0, rank)
        pool.time_next_kernel(ev[2], ev[3])
        pool.free_pages(n, ids, pbm.launch(pbm.MODE_BATCHED))
        gpu.cuda.synchronize()
        alloc_ms, free_ms = ev[0].elapsed_ms(ev[1]), ev[2].elapsed_ms(ev[3])
        assert 0.0 < alloc_ms < 50.0 and 0.0 < free_ms < 50.0
        assert np.array_equal(pool.snapshot(), before)
        for e in ev:
            e.close()


def test_every_call_disarms_the_events(gpu):
    with pbm.Pool(4096, strategy=pbm.LINKED_LIST) as lst, pbm.Pool(4096) as bm:
        a, b = pbm.KernelEvent(), pbm.KernelEvent()
        ids = gpu.full((64,), -1, dtype=gpu.int32, device="cuda")
        # the list pool's launch does not record them, and clears them
        lst.time_next_kernel(a, b)
        lst.get_pages(64, ids, None, 0, pbm.launch(mode=pbm.MODE_PLAIN))
        # an empty call clears them too
        bm.time_next_kernel(a, b)
        bm.get_pages(0, ids)
        bm.get_pages(64, ids, None, 0, pbm.launch(mode=pbm.MODE_PLAIN))
        gpu.cuda.synchronize()
        with pytest.raises(pbm.PbmError):      # never recorded
            a.elapsed_ms(b)
        # the failed HIP call is reported once, by the library: the caller's
        # next HIP work runs clean
        assert int(gpu.full((8,), 3, dtype=gpu.int32, device="cuda").sum().item()) == 24
        a.close()
        b.close()
