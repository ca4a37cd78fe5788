"""GPU: several right-hand sides in one pass (kifmm_evaluate_multi).

Each phase runs one launch over all right-hand sides (point kernels in groups
of up to 8 with every source staged once, GEMMs over nrhs times the columns, V
list over nrhs times the boxes).  Every right-hand side's arithmetic is the
one-RHS arithmetic in the same order (the FP32 near-field sums pair their
sources as the one-RHS sweep does, the GEMMs keep the one-RHS K split), so in
both precisions the values equal one-RHS evaluates bit for bit.
"""
import numpy as np
import pytest

import kifmm
from kifmm import workloads as W

pytestmark = pytest.mark.gpu


def one_by_one(f, den):
    return np.stack([f.evaluate(d) for d in den])


@pytest.mark.parametrize("kt,coefs,np_,nrhs", [(111, [1.0, 0.25], 6, 3), (111, [1.0, 0.25], 4, 11),
                                               (321, [1.0, 0.25], 4, 2), (511, [1.0, 0.3], 6, 8)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multi_equals_single(kt, coefs, np_, nrhs, precision):
    n = 12000
    print("case multi-rhs kt=%d np=%d nrhs=%d %s" % (kt, np_, nrhs, precision), flush=True)
    if kt == 321:
        pos, nor = W.spheres(n, 600, nsph=5)
    else:
        pos, nor = W.gaussians(n, 601), None
    f = kifmm.FMM3d(kt, coefs, np_, 32, 30, precision=precision)
    f.set_points(pos, nor)
    f.setup()
    rng = np.random.default_rng(602)
    den = rng.standard_normal((nrhs, n * f.src_dof))
    multi = f.evaluate_multi(den)
    single = one_by_one(f, den)
    r = np.linalg.norm(multi - single) / np.linalg.norm(single)
    print("relative difference %.3e" % r, flush=True)
    assert np.array_equal(multi, single)
    # a smaller nrhs after a larger one reuses the buffers
    again = f.evaluate_multi(den[:2])
    assert np.array_equal(again, multi[:2])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multi_graph_replay(gpu, precision):
    """Graph replay of the 4-RHS pass on a non-default stream, three replays with
    new densities, each against the eager pass on a second handle (bitwise)."""
    import torch
    n, R = 15000, 4
    print("case multi-rhs graph %s" % precision, flush=True)
    pos = W.gaussians(n, 610)
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 64, 30, precision=precision)
    f.set_points(pos)
    f.setup()
    g = kifmm.FMM3d(111, [1.0, 0.25], 6, 64, 30, precision=precision)
    g.set_points(pos)
    g.setup()
    g.set_graph(True)
    st = torch.cuda.Stream(device=gpu)
    d_den = torch.empty(R * n, dtype=torch.float64, device=gpu)
    d_val = torch.empty(R * n, dtype=torch.float64, device=gpu)
    for rep in range(3):
        den = np.random.default_rng(611 + rep).uniform(-1, 1, (R, n))
        with torch.cuda.stream(st):
            d_den.copy_(torch.from_numpy(den.reshape(-1)).to(gpu))
            g.evaluate_multi_device(d_den, d_val, R, stream=st.cuda_stream)
        st.synchronize()
        got = d_val.cpu().numpy().reshape(R, n)
        ref = f.evaluate_multi(den)
        assert np.array_equal(got, ref), rep
