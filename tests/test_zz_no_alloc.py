"""GPU: tomo_sq_norm_noalloc, the allocation-free squared norm.

It runs tomo_sq_norm's kernels with the same block count, so its result
must equal tomo_sq_norm's bit for bit.  Checked on every storage type, on
two calls in a row whose second input is smaller (the per-GPU scratch array
still holds the first call's partial sums, which the second must not read),
on a stream other than the default one, and against an FP64 sum on the host.
"""
import numpy as np
import pytest
import torch

from tomogpi_amd import lib

pytestmark = pytest.mark.gpu

SIZES = (3_000_001, 777)  # the first caps the partials at 2048 blocks, the second uses 4


def _host_sq_norm(t, scale):
    v = t.float().cpu().numpy().astype(np.float64) / scale
    return float(np.dot(v, v))


@pytest.mark.parametrize("dtype,scale", [(torch.float32, 1.0), (torch.float16, 4.0), (torch.bfloat16, 1.0)])
def test_noalloc_equals_sq_norm_on_two_calls(ctx, dtype, scale):
    g = torch.Generator().manual_seed(11)
    xs = [((torch.rand(n, generator=g) * 2.0 - 1.0) * scale).to(dtype).cuda() for n in SIZES]
    torch.cuda.synchronize()
    got = [lib.sq_norm_noalloc(ctx, x, scale) for x in xs]   # two calls in a row, the second smaller
    ref = [lib.sq_norm(ctx, x, scale) for x in xs]
    for n, a, b, x in zip(SIZES, got, ref, xs):
        print(f"n={n} dtype={dtype} noalloc={a!r} sq_norm={b!r}", flush=True)
        assert a == b
        h = _host_sq_norm(x, scale)
        assert abs(a - h) <= 1e-12 * h


def test_noalloc_on_a_side_stream(ctx):
    s = torch.cuda.Stream()
    g = torch.Generator().manual_seed(12)
    with torch.cuda.stream(s):
        big = torch.rand(1_000_000, generator=g).cuda()
        small = torch.rand(5_000, generator=g).cuda()
        a = lib.sq_norm_noalloc(ctx, big, stream=s)
        b = lib.sq_norm_noalloc(ctx, small, stream=s)
    torch.cuda.synchronize()
    assert a == lib.sq_norm(ctx, big)
    assert b == lib.sq_norm(ctx, small)
