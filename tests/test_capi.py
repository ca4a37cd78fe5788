"""C ABI (include/kifmm.h) without a GPU: the library loads, exports every
declared entry point, validates arguments, reports errors through
kifmm_last_error() instead of aborting, and parses the reference's options-file
format."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import kifmm
from kifmm import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    src = open(os.path.join(ROOT, "include", "kifmm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kifmm_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(kifmm.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(kifmm.SIGNATURES), set(names) ^ set(kifmm.SIGNATURES)


def test_create_rejects_bad_arguments():
    for kw in [dict(kernel_type=999), dict(kernel_type=191), dict(np=2), dict(np=9), dict(pts_max=0),
               dict(max_level=0), dict(max_level=31), dict(precision="half")]:
        with pytest.raises((kifmm.KifmmError, KeyError)) as e:
            kifmm.FMM3d(**kw)
        if isinstance(e.value, kifmm.KifmmError):
            assert e.value.code == -2 and str(e.value)


def test_state_errors():
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 32, 30)
    with pytest.raises(kifmm.KifmmError) as e:
        f.tree()
    assert e.value.code == -3
    f.set_points(W.uniform(100, 1))
    with pytest.raises(kifmm.KifmmError) as e:
        f.evaluate(np.zeros(100))
    assert e.value.code == -3 and "before setup" in str(e.value)


def test_points_outside_root_box_rejected():
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 32, 30)
    pos = W.uniform(50, 2)
    pos[7] = [0.5, 1.5, 0.5]
    f.set_points(pos)
    with pytest.raises(kifmm.KifmmError) as e:
        f.setup_host()
    assert "outside the root box" in str(e.value)
    assert e.value.code == -2  # KIFMM_ERR_ARG: bad input, not a call out of order
    tp = W.uniform(30, 3)
    tp[4] = [-0.25, 0.5, 0.5]
    f.set_points(W.uniform(50, 2), None, tp)
    with pytest.raises(kifmm.KifmmError) as e:
        f.setup_host()
    assert e.value.code == -2 and "target point 4" in str(e.value)


def test_double_layer_needs_normals():
    f = kifmm.FMM3d(321, [1.0, 0.25], 4, 32, 30)
    with pytest.raises(kifmm.KifmmError):
        f.set_points(W.uniform(50, 3), None)


def test_navier_needs_two_coefficients():
    f = kifmm.FMM3d(511, [1.0], 4, 32, 30)
    f.set_points(W.uniform(50, 3))
    with pytest.raises(kifmm.KifmmError):
        f.setup_host()


def test_options_parser_reference_format():
    path = os.path.join(GOLDEN, "options_laplace")
    assert kifmm.options_lookup(path, "-kt") == "111"
    assert kifmm.options_lookup(path, "-fmm3d_np") == "4"
    # the code asks for ptsMax / maxLevel, the files spell them in lower case
    assert kifmm.options_lookup(path, "-fmm3d_let3d_ptsMax") == "64"
    assert kifmm.options_lookup(path, "-fmm3d_let3d_maxLevel") == "30"
    assert kifmm.options_lookup(path, "-do_2nd_eval") == ""          # flag
    assert kifmm.options_lookup(path, "-log_summary") == ""          # PETSc-only flag kept
    assert kifmm.options_lookup(path, "-numchk") == "1"              # value after a comment
    assert kifmm.options_lookup(path, "-balance_octree") is None     # commented out
    assert kifmm.options_lookup(path, "-kifmm_shift") == "-0.5"      # negative value
    f = kifmm.FMM3d.from_options(path)
    assert f.src_dof == 1 and f.trg_dof == 1


def test_options_navier_coefficients():
    f = kifmm.FMM3d.from_options(os.path.join(GOLDEN, "options_navier"))
    assert f.src_dof == 3 and f.trg_dof == 3


def test_options_missing_file_and_option():
    with pytest.raises(kifmm.KifmmError) as e:
        kifmm.FMM3d.from_options(os.path.join(GOLDEN, "no_such_file"))
    assert e.value.code == -4
    with pytest.raises(kifmm.KifmmError) as e:
        kifmm.FMM3d.from_options(os.path.join(GOLDEN, "options_incomplete"))
    assert e.value.code == -4 and "let3d_maxLevel" in str(e.value)


def test_adjoint_state_error():
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 32, 30)
    f.set_points(W.uniform(100, 1))
    with pytest.raises(kifmm.KifmmError) as e:
        f.evaluate_adjoint(np.zeros(100))
    assert e.value.code == -3 and "before setup" in str(e.value)
