"""Host half of setup (no GPU): the product's octree, tags, packing and U/V/W/X
lists must equal the oracle's bit for bit, and its operator tables must equal
the oracle's (kernel matrices, FFT spectra) or satisfy the truncated
Moore-Penrose identities and match numpy's LAPACK pseudo-inverse (pinv)."""
import numpy as np
import pytest

import kifmm
import oracle as O
from kifmm import workloads as W
from test_oracle import check_truncated_pinv


def product_tree(pos, tpos=None, kt=111, np_=4, pts_max=32, max_level=30, nor=None):
    f = kifmm.FMM3d(kt, [1.0, 0.25], np_, pts_max, max_level)
    f.set_points(pos, nor if nor is not None else np.zeros_like(pos), tpos)
    f.setup_host()
    return f, f.tree()


CASES = {
    "uniform": lambda: (W.uniform(20000, 1), None, 32, 30),
    "gaussians": lambda: (W.gaussians(30000, 2), None, 16, 30),
    "gaussians_capped": lambda: (W.gaussians(30000, 3), None, 16, 5),
    "spheres": lambda: (W.spheres(20000, 4, nsph=8)[0], None, 24, 30),
    "separate_targets": lambda: (W.uniform(5000, 5), W.gaussians(4000, 6), 20, 30),
    "duplicates": lambda: (np.concatenate([W.uniform(500, 7)] * 3), None, 8, 30),
    "faces": lambda: (np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.25, 0.75, 0.5],
                                [W.HI, W.HI, W.HI], [0.5, 0.0, W.HI]] * 4 + W.uniform(200, 8).tolist()),
                      None, 4, 30),
    "small": lambda: (W.uniform(10, 9), None, 64, 30),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_tree_and_lists_bit_exact(case):
    pos, tpos, pm, ml = CASES[case]()
    f, tr = product_tree(pos, tpos, pts_max=pm, max_level=ml)
    ot = O.Tree(pos, tpos, pts_max=pm, max_level=ml)
    assert tr["parent"].size == ot.n_nodes
    for k in ["parent", "child", "depth", "tag", "src_beg", "src_num", "trg_beg", "trg_num",
              "src_perm", "trg_perm"]:
        assert np.array_equal(tr[k], getattr(ot, k)), k
    assert np.array_equal(tr["path"], ot.path)
    for name in "uvwx":
        assert np.array_equal(tr["lists"][name][0], ot.lists[name][0]), name
        assert np.array_equal(tr["lists"][name][1], ot.lists[name][1]), name


def test_tree_and_lists_bit_exact_clustered_8m():
    """The whole point set of the 8,388,608-point clustered Laplace workload
    (16 Gaussians + 10% uniform, ptsMax 128, maxLevel 14): octree, tags,
    packing and all four lists equal the oracle's bit for bit (about 20 s)."""
    pos = W.gaussians(8388608, W.seed(2, 0))
    f, tr = product_tree(pos, None, np_=6, pts_max=128, max_level=14)
    ot = O.Tree(pos, None, pts_max=128, max_level=14)
    assert tr["parent"].size == ot.n_nodes
    for k in ["parent", "child", "depth", "tag", "src_beg", "src_num", "trg_beg", "trg_num",
              "src_perm", "trg_perm"]:
        assert np.array_equal(tr[k], getattr(ot, k)), k
    assert np.array_equal(tr["path"], ot.path)
    for name in "uvwx":
        assert np.array_equal(tr["lists"][name][0], ot.lists[name][0]), name
        assert np.array_equal(tr["lists"][name][1], ot.lists[name][1]), name


@pytest.mark.parametrize("kt,np_,coefs", [(111, 4, [1, .25]), (111, 6, [1, .25]), (321, 4, [1, .25]),
                                          (321, 6, [1, .25]), (511, 4, [1, .3]), (511, 8, [1, .3])])
def test_operator_tables(kt, np_, coefs):
    f = kifmm.FMM3d(kt, coefs, np_, 32, 30)
    p = W.uniform(100, 11)
    f.set_points(p, p)
    f.setup_host()
    km = O.mm_kernel(kt)
    sd, td = O.src_dof(km), O.trg_dof(km)
    nue = np_ ** 3 - (np_ - 2) ** 3
    nuc = (np_ + 2) ** 3 - np_ ** 3
    eps = 0.1 ** (np_ + 1)
    pu = f.operator(0).reshape(nuc * td, nue * sd).T
    pd = f.operator(1).reshape(nue * td, nue * sd).T
    # the product's Jacobi-SVD pinv satisfies the truncated identities ...
    up, dn = O.numpy_pinv_tables(km, coefs, np_)
    for which, P, Pn in ((0, pu, up), (1, pd, dn)):
        M = O.check_matrix(km, coefs, np_, which)
        check_truncated_pinv(M, P, eps)
    # ... and agrees with numpy's LAPACK pinv up to the conditioning (~1e7 at np 6)
    tol = {4: 1e-9, 6: 1e-7, 8: 1e-5}[np_]
    assert np.linalg.norm(pu - up) / np.linalg.norm(up) < tol
    assert np.linalg.norm(pd - dn) / np.linalg.norm(dn) < tol
    # translation tables: identical to the oracle's (same closed forms, same DFT)
    ops = O.Ops(km, coefs, np_, up, dn)
    m2m = f.operator(2).reshape(8, nue * sd, nuc * td).transpose(0, 2, 1)
    l2l = f.operator(3).reshape(8, nue * sd, nue * td).transpose(0, 2, 1)
    np.testing.assert_allclose(m2m, ops.m2m, rtol=0, atol=1e-15 * np.abs(ops.m2m).max())
    np.testing.assert_allclose(l2l, ops.l2l, rtol=0, atol=1e-15 * np.abs(ops.l2l).max())
    H = (np_ + 1) * (2 * np_) ** 2
    m2l = f.operator(4).view(np.complex128).reshape(343, sd * td, H).transpose(0, 2, 1)
    np.testing.assert_allclose(m2l, ops.m2l, rtol=0, atol=1e-13 * np.abs(ops.m2l).max())


def test_work_counts_match_lists():
    pos = W.gaussians(20000, 12)
    f, tr = product_tree(pos, pts_max=32)
    w = f.work_counts()
    off, lst = tr["lists"]["v"]
    assert w["vlist_pairs"] == len(lst)
    trg_num = tr["trg_num"]
    # exact near pairs = sum over target leaves of n_t * sum of n_s over U items (+ W/X fallbacks)
    u_off, u_lst = tr["lists"]["u"]
    src_num = tr["src_num"]
    u_pairs = sum(float(trg_num[b]) * src_num[u_lst[u_off[b]:u_off[b + 1]]].sum()
                  for b in range(len(trg_num)) if trg_num[b] > 0)
    assert w["near_exact_pairs"] >= u_pairs > 0


def test_forest_is_the_per_problem_trees():
    """Many problems in one handle (set_points_batch): the forest's nodes,
    ranges, permutations and lists are each problem's own tree (the oracle's,
    bit for bit) shifted by the node and point counts of the problems before
    it; no list entry leaves its tree."""
    probs = [W.uniform(3000, 501), W.gaussians(5000, 502), W.uniform(700, 503), W.gaussians(2200, 504)]
    f = kifmm.FMM3d(511, [1.0, 0.3], 4, 40, 10)
    f.set_points_batch(probs)
    f.setup_host()
    tr = f.tree()
    node0 = pt0 = 0
    for pos in probs:
        ot = O.Tree(pos, None, pts_max=40, max_level=10)
        sl = slice(node0, node0 + ot.n_nodes)
        sh = lambda a, by: np.where(a >= 0, a + by, a)  # noqa: E731
        assert np.array_equal(tr["parent"][sl], sh(ot.parent, node0))
        assert np.array_equal(tr["child"][sl], sh(ot.child, node0))
        for k in ["depth", "tag", "src_num", "trg_num"]:
            assert np.array_equal(tr[k][sl], getattr(ot, k)), k
        assert np.array_equal(tr["path"][sl], ot.path)
        assert np.array_equal(tr["src_beg"][sl], sh(ot.src_beg, pt0))
        assert np.array_equal(tr["trg_beg"][sl], sh(ot.trg_beg, pt0))
        assert np.array_equal(tr["src_perm"][pt0:pt0 + len(pos)], ot.src_perm + pt0)
        assert np.array_equal(tr["trg_perm"][pt0:pt0 + len(pos)], ot.trg_perm + pt0)
        for name in "uvwx":
            off, lst = tr["lists"][name]
            ooff, olst = ot.lists[name]
            got = lst[off[node0]:off[node0 + ot.n_nodes]]
            assert np.array_equal(np.diff(off[node0:node0 + ot.n_nodes + 1]), np.diff(ooff)), name
            assert np.array_equal(got, olst + node0), name
        node0 += ot.n_nodes
        pt0 += len(pos)
    assert node0 == tr["parent"].size
