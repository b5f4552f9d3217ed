"""aur_create_emitters_noalloc, the allocation-free variant of
aur_create_emitters: through the C ABI its table equals the allocating entry
point's bit for bit on two calls in a row (the second over a smaller scene,
so the fixed scratch holds the larger call's tiles), and the library
allocates no device memory during the call (an allocator hook counts)."""
from __future__ import annotations

import numpy as np
import pytest

from aurora_amd import native
from tests.test_oracle_lights import emissive_scene

pytestmark = pytest.mark.gpu


def _table(fn, n):
    import torch

    buf = torch.full((max(1, n) * 28,), 0xCD, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    count = fn(buf.data_ptr(), n)
    return buf[: count * 28].cpu().numpy().view(native.EMITTER_DTYPE)


def test_noalloc_equals_allocating_entry_point():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    talloc, tfree = native.torch_alloc_hook()
    allocs = []

    def alloc(n):
        allocs.append(n)
        return talloc(n)

    with native.Context(0) as ctx:
        for seed, n in ((11, 300_000), (12, 5_000)):  # second call: fewer tiles than the first
            s = emissive_scene(seed, n)
            ctx.upload_scene(s)
            ctx.rebuild_nmh()
            torch.cuda.synchronize()
            native.set_alloc_hook(alloc, tfree)
            try:
                before = len(allocs)
                got = _table(ctx.emitters_noalloc_ptr, ctx.count)
                assert len(allocs) == before, "the allocation-free entry point allocated"
            finally:
                native.set_alloc_hook()
            ref = _table(ctx.emitters_ptr, ctx.count)
            assert got.shape == ref.shape and len(ref) > n // 4
            assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
            again = _table(ctx.emitters_noalloc_ptr, ctx.count)
            assert np.array_equal(again.view(np.uint8), ref.view(np.uint8))
