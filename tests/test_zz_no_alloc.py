"""The allocation-free entry point gd_timestep_error_noalloc through the C ABI:
device arrays in, the result in a device double, partial sums in the
library's static device scratch.  Checked against the allocating
gd_timestep_error (the same kernels and block count, so the same bits) and
against numpy, on two calls in a row, the second with a smaller input, and
once more after a call of the allocating entry point."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_timestep_error_matches_allocating_entry():
    import torch

    import gdhip  # noqa: F401
    from gdhip._arrays import context
    from gdhip._lib import check, lib

    L = lib()
    vp = ctypes.c_void_p
    f = L.gd_timestep_error_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [vp, ctypes.c_long, vp, vp, vp, vp, vp]
    ctx = context(0)
    check(L.gd_ctx_set_stream(ctx, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    g = torch.Generator(device="cuda").manual_seed(3)
    for n in (1_048_583, 4_099, 1):
        print("case: noalloc timestep error n=%d" % n, flush=True)
        a = [torch.randn(n, dtype=torch.float64, device="cuda", generator=g) for _ in range(4)]
        out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        ref = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        ptrs = [ctypes.c_void_p(t.data_ptr()) for t in a]
        check(f(ctx, n, *ptrs, ctypes.c_void_p(out.data_ptr())))
        check(L.gd_timestep_error(ctx, n, *ptrs, ctypes.c_void_p(ref.data_ptr()), 0))
        torch.cuda.synchronize()
        got, want = out.item(), ref.item()
        assert got == want, (n, got, want)
        h = [t.cpu().numpy() for t in a]
        exact = float(np.sqrt(np.sum((h[0] - h[2]) ** 2 + (h[1] - h[3]) ** 2)))
        assert abs(got - exact) <= 1e-12 * exact, (n, got, exact)
    # arguments are checked before anything is launched
    with pytest.raises(gdhip.GDError):
        check(f(ctx, -1, *ptrs, ctypes.c_void_p(out.data_ptr())))
    check(L.gd_ctx_set_stream(ctx, None))
