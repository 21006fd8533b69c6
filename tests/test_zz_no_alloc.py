"""The WLS solve with no allocation (lsh_wls_solve_noalloc) against the
allocating one (lsh_wls_solve, reference mode) through the C ABI: the same
kernels on the same numbers, so the solution, the iterations and the residuals
are the same bit for bit -- on two calls in a row, the second on a smaller grid
(the scratch array keeps the first call's contents), from the FGS guess and from
a warm start; a grid beyond the scratch array is refused with an error."""
import numpy as np
import pytest

import lapstack_amd as la

pytestmark = pytest.mark.gpu


def _system(w, h, seed):
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    yy, xx = torch.meshgrid(torch.arange(h, device="cuda", dtype=torch.float32),
                            torch.arange(w, device="cuda", dtype=torch.float32), indexing="ij")
    guide = (0.5 + 0.2 * torch.sin(xx / 17.0) * torch.cos(yy / 13.0)
             + 0.05 * torch.rand((h, w), generator=g, device="cuda")).contiguous()
    guide[:, w // 3:w // 3 + 2] += 0.3  # an edge the weights respect
    conf = torch.rand((h, w), generator=g, device="cuda").contiguous()
    conf[h // 4:h // 2, w // 4:w // 2] = 0.0  # a hole to fill
    d = (4.0 * torch.rand((h, w), generator=g, device="cuda")).contiguous()
    return guide, conf, d


def _compare(w, h, seed, warm=False):
    import torch

    print(f"no-alloc WLS {w}x{h} warm={warm}", flush=True)
    guide, conf, d = _system(w, h, seed)
    lam, sig = 3.0, 0.04
    wm = (d * 0.5).contiguous() if warm else None
    want = torch.empty((h, w), dtype=torch.float32, device="cuda")
    _, rw, _ = la.wls_solve(guide, conf, d, lam, sig, "reference", tol=1e-6, max_iters=200, warm=wm, u_out=want)
    got, rg = la.wls_solve_noalloc(guide, conf, d, lam, sig, tol=1e-6, max_iters=200, warm=wm)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    assert rg["iters"] == rw["iters"] and rg["rel"] == rw["rel"] and rg["true_rel"] == rw["true_rel"]
    assert rg["iters"] > 0 and rg["rel"] < 1e-6
    return got


def test_no_alloc_wls_equals_the_allocating_solve_twice_in_a_row():
    why = la.available()
    if why is not None:
        pytest.fail(f"HIP backend unavailable on a GPU run: {why}")
    _compare(512, 384, 3)           # the largest working grid the configurations run through it
    _compare(300, 200, 4)           # smaller, after it: the array holds the first call's values
    _compare(300, 200, 5, warm=True)


def test_no_alloc_wls_refuses_a_grid_beyond_its_scratch():
    import torch

    guide, conf, d = (torch.zeros((768, 1024), dtype=torch.float32, device="cuda") for _ in range(3))
    with pytest.raises(la.HipError, match="no-alloc"):
        la.wls_solve_noalloc(guide, conf, d, 3.0, 0.04)
