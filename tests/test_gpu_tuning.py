"""Launch shapes of the mixed path and the autotuner that picks them.

The near-field launch forms (one workgroup per chunk in leaf order, the same
largest chunk first, a persistent grid over a global work queue) compute every
chunk alone, so they must agree bit for bit; the chunk-split staging weight
changes how targets share lanes (FP32 partial sums), so it moves values only at
round-off.  The autotuner must keep results within the mixed path's bar of the
FP64 path whatever it picks."""
import numpy as np
import pytest

import kifmm
from kifmm import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("kt,coefs,np_,nrhs", [(111, [1.0, 0.25], 6, 1), (111, [1.0, 0.25], 4, 3),
                                              (511, [1.0, 0.3], 4, 2)])
def test_near_launch_forms_bitwise(gpu, kt, coefs, np_, nrhs):
    n = 15000
    pos = W.gaussians(n, 300 + kt)
    f = kifmm.FMM3d(kt, coefs, np_, 32, 30, precision="mixed")
    f.set_points(pos)
    f.setup()
    den = np.concatenate([W.densities(n, f.src_dof, 301 + r, kind="normal") for r in range(nrhs)])
    out = {}
    for mode in (0, 1, 2):
        print("case kt=%d nrhs=%d near_mode=%d" % (kt, nrhs, mode), flush=True)
        f.set_tuning(near_mode=mode)
        out[mode] = f.evaluate_multi(den) if nrhs > 1 else f.evaluate(den)
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(out[0], out[2])


def test_near_stage_weights(gpu):
    n = 20000
    pos = W.gaussians(n, 310)
    den = W.densities(n, 1, 311)
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 64, 30, precision="mixed")
    f.set_points(pos)
    f.setup()
    base = f.evaluate(den)
    for st in (0, 2):
        print("case near_stage=%d" % st, flush=True)
        f.set_tuning(near_stage=st)
        assert rel(f.evaluate(den), base) < 1e-7


def test_autotune_keeps_accuracy(gpu):
    """autotune on the handle's own tree: the chosen knobs are valid, every
    trial is reported, and the values after tuning (eager and graph replay)
    stay within the mixed bar of the FP64 path."""
    import torch
    n = 60000
    pos = W.gaussians(n, 320)
    den = W.densities(n, 1, 321)
    g = kifmm.FMM3d(111, [1.0, 0.25], 6, 128, 14, precision="fp64")
    g.set_points(pos)
    g.setup()
    ref = g.evaluate(den)
    d_den = torch.from_numpy(den).to(gpu)
    d_val = torch.empty(n, dtype=torch.float64, device=gpu)
    assert g.autotune(d_den, d_val)["trials"] == []  # FP64: nothing to tune
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 128, 14, precision="mixed")
    f.set_points(pos)
    f.setup()
    rep = f.autotune(d_den, d_val)
    torch.cuda.synchronize()
    print("autotune: %s" % rep, flush=True)
    t = rep["tuning"]
    assert t["near_mode"] in (0, 1, 2) and t["near_stage"] in (0, 1, 2) and t["translations"] in (0, 6)
    assert [k for k, _, _ in rep["trials"]] == ["translations"] * 2 + ["near_stage"] * 3 + ["near_mode"] * 3
    assert all(ms > 0 for _, _, ms in rep["trials"])
    eager = f.evaluate(den)
    assert rel(eager, ref) < 3.4e-7
    st = torch.cuda.Stream(device=gpu)
    f.set_graph(True)
    with torch.cuda.stream(st):
        for _ in range(3):
            f.evaluate_device(d_den, d_val, stream=st.cuda_stream)
            st.synchronize()
            assert np.array_equal(d_val.cpu().numpy(), eager)
    f.set_graph(False)


@pytest.mark.parametrize("kt,coefs,np_,gen,tol", [(321, [1.0, 0.25], 4, "spheres", 3e-5),
                                                (511, [1.0, 0.3], 6, "uniform", 1e-6)])
def test_vector_hadamard_forms(gpu, kt, coefs, np_, gen, tol):
    """The per-pair and the halo-sweep Hadamards of the vector kernels (1 or 2
    groups a tile) sum the same products in different orders: each within the
    mixed bar of the FP64 path, and within FP32 round-off of each other."""
    n = 20000
    if gen == "spheres":
        pos, nor = W.spheres(n, 330, nsph=6)
    else:
        pos, nor = W.uniform(n, 331), None
    den = W.densities(n, 3, 332, kind="normal")
    g = kifmm.FMM3d(kt, coefs, np_, 32, 30, precision="fp64")
    g.set_points(pos, nor)
    g.setup()
    ref = g.evaluate(den)
    f = kifmm.FMM3d(kt, coefs, np_, 32, 30, precision="mixed")
    f.set_points(pos, nor)
    f.setup()
    out = {}
    for form in (0, 1, 2):
        print("case kt=%d vec_hadamard=%d" % (kt, form), flush=True)
        f.set_tuning(vec_hadamard=form)
        out[form] = f.evaluate(den)
        print("  vs fp64 %.3e" % rel(out[form], ref), flush=True)
        assert rel(out[form], ref) < tol
    assert rel(out[1], out[0]) < 1e-6 and rel(out[2], out[0]) < 1e-6
    assert np.array_equal(out[1], out[2]) or rel(out[1], out[2]) < 1e-6
