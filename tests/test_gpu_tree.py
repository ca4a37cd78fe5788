"""Octree and lists built on the GPU (kifmm_setup's default builder) against
the host builder (kifmm_setup_host), which the CPU tests pin bit for bit to
the oracle: nodes, paths, tags, packed ranges, permutations and all four
lists must be identical, for single problems, forests and the full
8,388,608-point clustered point set."""
import numpy as np
import pytest

import kifmm
from kifmm import workloads as W
from test_host_parity import CASES

pytestmark = pytest.mark.gpu

KEYS = ["parent", "child", "depth", "tag", "src_beg", "src_num", "trg_beg", "trg_num", "src_perm", "trg_perm"]


def both_trees(set_points, kt=111, coefs=(1.0, 0.25), np_=4, pts_max=32, max_level=30):
    out = []
    for on_device in (False, True):
        f = kifmm.FMM3d(kt, list(coefs), np_, pts_max, max_level)
        set_points(f)
        if on_device:
            f.set_tree_builder(True)
            f.setup()
        else:
            f.setup_host()
        out.append((f, f.tree()))
    return out


def assert_same(a, b):
    assert a["parent"].size == b["parent"].size
    assert a["max_depth"] == b["max_depth"]
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["path"], b["path"])
    for name in "uvwx":
        assert np.array_equal(a["lists"][name][0], b["lists"][name][0]), name
        assert np.array_equal(a["lists"][name][1], b["lists"][name][1]), name


@pytest.mark.parametrize("case", sorted(CASES))
def test_device_tree_equals_host(gpu, case):
    pos, tpos, pm, ml = CASES[case]()
    (fh, th), (fd, td) = both_trees(lambda f: f.set_points(pos, np.zeros_like(pos), tpos), pts_max=pm, max_level=ml)
    print("case %s: %d nodes, setup %s" % (case, td["parent"].size, fd.setup_times()), flush=True)
    assert_same(th, td)


def test_device_tree_forest(gpu):
    """Problems of uneven size, one without sources, separate targets: the
    forest numbering (problem p's nodes after those of 0..p-1)."""
    src = [W.uniform(3000, 501), np.zeros((0, 3)), W.gaussians(5000, 502), W.uniform(700, 503),
           W.gaussians(2200, 504)]
    trg = [W.uniform(1000, 601), W.uniform(40, 602), W.gaussians(5000, 502), W.uniform(900, 603),
           W.uniform(10, 604)]
    (fh, th), (fd, td) = both_trees(lambda f: f.set_points_batch(src, [np.zeros_like(s) for s in src], trg),
                                    kt=511, coefs=(1.0, 0.3), pts_max=40, max_level=10)
    assert_same(th, td)


def test_device_tree_clustered_8m(gpu):
    """The 8,388,608-point clustered workload (ptsMax 128, maxLevel 14)."""
    pos = W.gaussians(8388608, W.seed(2, 0))
    (fh, th), (fd, td) = both_trees(lambda f: f.set_points(pos), np_=6, pts_max=128, max_level=14)
    print("8M: host %s  device %s" % (fh.setup_times(), fd.setup_times()), flush=True)
    assert_same(th, td)


def test_device_tree_depth_cap_and_duplicates(gpu):
    """Every point in one depth-maxLevel cell except a few: the cap stops the
    split (leaves above ptsMax) on both builders."""
    rng = np.random.default_rng(7)
    pos = np.concatenate([np.full((500, 3), 0.3) + rng.random((500, 3)) * 1e-12, W.uniform(300, 8)])
    (fh, th), (fd, td) = both_trees(lambda f: f.set_points(pos), pts_max=16, max_level=12)
    assert_same(th, td)


def test_device_tree_rejects_points_outside(gpu):
    pos = W.uniform(1000, 3)
    pos[417, 1] = 1.5
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 32, 30)
    f.set_points(pos)
    with pytest.raises(kifmm.KifmmError) as e:
        f.setup()
    assert e.value.code == -2
    assert "source point 417 lies outside the root box" in str(e.value)
