"""kifmm_direct_device_noalloc (the direct sum with the library's fixed per-GPU
scratch, no allocation) against kifmm_direct_device on two calls in a row,
the second on a smaller problem, so it runs over the first call's leftover
scratch contents; sizes beyond the scratch are rejected."""
import numpy as np
import pytest

import kifmm
from kifmm import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run_pair(gpu, kt, coefs, n, m, seed):
    import torch
    pos, nor = W.spheres(n, seed, nsph=6)
    rng = np.random.default_rng(seed + 1)
    tgt = np.concatenate([pos[rng.choice(n, m // 2, replace=False)], W.uniform(m - m // 2, seed + 2)])
    f = kifmm.FMM3d(kt, coefs, 4, 64, 30)
    f.set_points(pos, nor)
    f.setup()
    den = W.densities(n, f.src_dof, seed + 3, kind="normal")
    d_den = torch.from_numpy(den).to(gpu)
    d_chk = torch.from_numpy(np.ascontiguousarray(tgt)).to(gpu)
    a = torch.empty(m * f.trg_dof, dtype=torch.float64, device=gpu)
    b = torch.full((m * f.trg_dof,), np.nan, dtype=torch.float64, device=gpu)
    print("case noalloc kt=%d n=%d targets=%d" % (kt, n, m), flush=True)
    f.direct_device(d_den, d_chk, a)
    f.direct_device_noalloc(d_den, d_chk, b)
    torch.cuda.synchronize()
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert np.all(np.isfinite(b))
    return f, rel(b, a), bool(np.array_equal(a, b))


def test_direct_noalloc_matches_direct(gpu):
    # first call: Stokes double layer, 3 values per target
    f1, e1, same1 = run_pair(gpu, 321, [1.0, 0.25], 20000, 3000, 900)
    # second call, smaller: Laplace single layer over the scratch the first left
    f2, e2, same2 = run_pair(gpu, 111, [1.0, 0.25], 5000, 700, 910)
    print("noalloc vs direct: %.3e (bitwise %s), %.3e (bitwise %s)" % (e1, same1, e2, same2), flush=True)
    assert e1 <= 1e-13 and e2 <= 1e-13


def test_direct_noalloc_rejects_oversize(gpu):
    import torch
    f = kifmm.FMM3d(321, [1.0, 0.25], 4, 64, 30)
    pos, nor = W.spheres(2000, 920, nsph=4)
    f.set_points(pos, nor)
    f.setup()
    m = (1 << 21) // 3 + 1  # m * 3 values exceed the scratch
    d_den = torch.zeros(2000 * 3, dtype=torch.float64, device=gpu)
    d_chk = torch.full((3 * m,), 0.5, dtype=torch.float64, device=gpu)
    d_out = torch.empty(3 * m, dtype=torch.float64, device=gpu)
    with pytest.raises(kifmm.KifmmError) as e:
        f.direct_device_noalloc(d_den, d_chk, d_out)
    assert e.value.code == -2 and "scratch" in str(e.value)
