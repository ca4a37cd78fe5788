"""Mixed path with the translations (M2M, L2L) as split-BF16 products on the
matrix cores (kifmm_set_translation 3 or 6; the pseudo-inverse applications
stay FP64): against the FP64 path (itself within 1e-12 of the oracle) and
against the direct sum, beside the default mixed path (FP64 matrix-core
translations).

Bars: BF16x6 (about 24 bits per operand) stays as close to FP64 as the default
mixed path does (3.4e-7 for the Laplace np=6 case, the mixed tolerances of the
vector kernels otherwise) and no less accurate against the direct sum beyond
that difference; BF16x3 (about 16 bits) is reported and held to 10x those."""
import numpy as np
import pytest

import kifmm
import oracle as O
from kifmm import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [(111, 6, [1.0, 0.25], "gauss", 3.4e-7), (321, 4, [1.0, 0.25], "spheres", 3e-5),
         (511, 6, [1.0, 0.3], "uniform", 1e-6)]


@pytest.mark.parametrize("kt,np_,coefs,gen,tol", CASES)
def test_bf16_translations(gpu, kt, np_, coefs, gen, tol):
    n = 20000
    if gen == "spheres":
        pos, nor = W.spheres(n, 191, nsph=6)
    else:
        pos = W.gaussians(n, 192) if gen == "gauss" else W.uniform(n, 193)
        nor = np.zeros_like(pos)
    den = W.densities(n, O.src_dof(kt), 194, kind="normal")
    out, err = {}, {}
    for name, prec, terms in (("fp64", "fp64", 0), ("mixed", "mixed", 0), ("bf16x6", "mixed", 6),
                              ("bf16x3", "mixed", 3)):
        print("case kt=%d np=%d %s" % (kt, np_, name), flush=True)
        f = kifmm.FMM3d(kt, coefs, np_, 32, 30, precision=prec)
        f.set_translation(terms)
        f.set_points(pos, nor)
        f.setup()
        out[name] = f.evaluate(den)
        err[name] = f.check(den, out[name], 2000, seed=7)
        f.close()
    d = {k: rel(out[k], out["fp64"]) for k in ("mixed", "bf16x6", "bf16x3")}
    print("vs fp64: %s  vs direct: %s" % (d, err), flush=True)
    assert d["mixed"] < tol
    assert d["bf16x6"] < tol
    assert err["bf16x6"] <= err["fp64"] + 2 * d["bf16x6"]
    assert d["bf16x3"] < 10 * tol


def test_bf16_translations_multi_rhs_bitwise(gpu):
    """Several right-hand sides in one pass give each right-hand side's
    one-RHS values bit for bit (every column of the BF16 products is computed
    alone)."""
    import torch
    n, R = 12000, 3
    pos = W.gaussians(n, 195)
    den = np.stack([W.densities(n, 1, 196 + r) for r in range(R)])
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 32, 30, precision="mixed")
    f.set_translation(6)
    f.set_points(pos)
    f.setup()
    multi = f.evaluate_multi(den.reshape(-1)).reshape(R, -1)
    for r in range(R):
        assert np.array_equal(multi[r], f.evaluate(den[r])), r


def test_translation_setting_needs_setup(gpu):
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 32, 30, precision="mixed")
    f.set_points(W.uniform(2000, 197))
    f.setup()
    f.set_translation(6)
    with pytest.raises(kifmm.KifmmError) as e:
        f.evaluate(W.densities(2000, 1, 198))
    assert e.value.code == -3
    with pytest.raises(kifmm.KifmmError) as e:
        f.set_translation(4)
    assert e.value.code == -2
