"""athx_exclusive_scan_noalloc, the entry point that allocates nothing, on the
GPU through the C ABI: two calls in a row, the second smaller, each equal bit
for bit to athx_exclusive_scan's result (output and total), with no device
memory taken by the context or by PyTorch's allocator during the calls, and
the size limit refused with INVALID_ARGUMENT."""
from __future__ import annotations

import sys

import pytest
import torch

from athenea_amd import TileRasterizer
from athenea_amd._lib import AthxError

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def test_noalloc_scan_matches_the_allocating_scan():
    raster = TileRasterizer(0)
    try:
        g = torch.Generator(device="cuda:0").manual_seed(11)
        for n in (5_000_003, 777_777):
            say("no-alloc scan, n =", n)
            values = torch.randint(0, 1 << 10, (n,), dtype=torch.int32, device="cuda:0", generator=g)
            want, want_total = raster.exclusive_scan(values)
            out = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
            total = torch.full((1,), -1, dtype=torch.int64, device="cuda:0")
            torch.cuda.synchronize()
            held = raster.scratch_bytes()
            torch_held = torch.cuda.memory_allocated()
            raster.exclusive_scan_noalloc(values, out, total)
            torch.cuda.synchronize()
            assert raster.scratch_bytes() == held
            assert torch.cuda.memory_allocated() == torch_held
            assert torch.equal(out, want)
            assert int(total.item()) == int(want_total.item()) == int(values.to(torch.int64).sum().item())
        with pytest.raises(AthxError) as err:
            big = torch.empty((1 << 24) + 1, dtype=torch.int32, device="cuda:0")
            raster.exclusive_scan_noalloc(big, torch.empty_like(big), torch.empty(1, dtype=torch.int64, device="cuda:0"))
        assert err.value.code == "InvalidArgument"
    finally:
        raster.close()
