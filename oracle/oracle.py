"""ctypes front end of the KIFMM3d CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  The oracle restates
the reference CPU evaluate path (fmm3d_mpi/fmm3d_eval_mpi.cpp and friends) in
plain C; see ``kifmm_oracle.c`` for the per-function citations.

The truncated pseudo-inverse (common/vecmatop.cpp:129-164, common/svdrep.cpp:38-80)
is computed here with numpy's LAPACK SVD, which is independent of the product's
own SVD.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

P_D = C.POINTER(C.c_double)
P_I = C.POINTER(C.c_int)
P_L = C.POINTER(C.c_int64)


class _Tree(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("max_level", C.c_int),
                ("parent", P_I), ("child", P_I), ("depth", P_I), ("path", P_I), ("tag", P_I),
                ("src_cnt", P_I), ("trg_cnt", P_I),
                ("src_beg", P_I), ("src_num", P_I), ("trg_beg", P_I), ("trg_num", P_I),
                ("src_perm", P_I), ("trg_perm", P_I),
                ("u_off", P_L), ("v_off", P_L), ("w_off", P_L), ("x_off", P_L),
                ("u_lst", P_I), ("v_lst", P_I), ("w_lst", P_I), ("x_lst", P_I),
                ("n_src", C.c_int), ("n_trg", C.c_int)]


class _Ops(C.Structure):
    _fields_ = [("kt_mm", C.c_int), ("np", C.c_int), ("sd", C.c_int), ("td", C.c_int),
                ("coefs", C.c_double * 4), ("ncoefs", C.c_int),
                ("n_ue", C.c_int), ("n_uc", C.c_int), ("n_de", C.c_int), ("n_dc", C.c_int),
                ("pinv_up", P_D), ("pinv_dn", P_D), ("m2m", P_D), ("l2l", P_D), ("m2l", P_D),
                ("deg", C.c_double * 4)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_kernel_matrix.argtypes = [C.c_int, P_D, C.c_int, P_D, P_D, C.c_int, P_D, P_D]
        L.orc_sample_pos.argtypes = [C.c_int, C.c_double, P_D]
        L.orc_regular_pos.argtypes = [C.c_int, C.c_double, P_D]
        L.orc_rfft3.argtypes = [C.c_int, C.c_int, P_D, P_D]
        L.orc_irfft3.argtypes = [C.c_int, C.c_int, P_D, P_D]
        L.orc_tree_build.argtypes = [C.POINTER(_Tree), C.c_int, P_D, C.c_int, P_D, P_D,
                                     C.c_int, C.c_int, C.c_int]
        L.orc_tree_free.argtypes = [C.POINTER(_Tree)]
        L.orc_ops_init.argtypes = [C.POINTER(_Ops), C.c_int, P_D, C.c_int, C.c_int, P_D, P_D]
        L.orc_ops_free.argtypes = [C.POINTER(_Ops)]
        L.orc_check_matrix.argtypes = [C.c_int, P_D, C.c_int, C.c_int, C.c_int, P_D]
        L.orc_evaluate.argtypes = [C.POINTER(_Tree), C.POINTER(_Ops), C.c_int, P_D, C.c_int,
                                   P_D, P_D, P_D, P_D, P_D, P_D, P_D, P_D]
        L.orc_eff_den.argtypes = [C.POINTER(_Ops), C.c_int, P_D, P_D]
        L.orc_m2l_hadamard.argtypes = [C.POINTER(_Ops), C.POINTER(C.c_int), P_D, P_D]
        L.orc_eff_val.argtypes = [C.POINTER(_Ops), P_D, P_D]
        L.orc_direct.argtypes = [C.c_int, P_D, C.c_int, P_D, P_D, P_D, C.c_int, P_D, P_D,
                                 C.c_int]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(P_D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def src_dof(kt):
    return lib().orc_src_dof(kt)


def trg_dof(kt):
    return lib().orc_trg_dof(kt)


def mm_kernel(kt):
    return lib().orc_mm_kernel(kt)


def coefs_arr(coefs):
    c = np.zeros(4)
    c[:len(coefs)] = coefs
    return c


def kernel_matrix(kt, coefs, spos, snor, tpos):
    """Dense (nt*td) x (ns*sd) interaction matrix (kernel3d_mpi.cpp:142-466)."""
    spos, snor, tpos = _f64(spos).reshape(-1, 3), _f64(snor).reshape(-1, 3), _f64(tpos).reshape(-1, 3)
    sd, td = src_dof(kt), trg_dof(kt)
    ns, nt = len(spos), len(tpos)
    out = np.zeros((ns * sd, nt * td))  # column-major (nt*td) x (ns*sd) == C-order transpose
    c = coefs_arr(coefs)
    rc = lib().orc_kernel_matrix(kt, _p(c), ns, _p(spos), _p(snor), nt, _p(tpos), _p(out))
    assert rc == 0
    return out.T.copy()


def sample_pos(np_, R):
    n = np_ ** 3 - (np_ - 2) ** 3
    out = np.zeros((n, 3))
    lib().orc_sample_pos(np_, R, _p(out))
    return out


def regular_pos(np_, R):
    n = (2 * np_) ** 3
    out = np.zeros((n, 3))
    lib().orc_regular_pos(np_, R, _p(out))
    return out


def rfft3(x, n, howmany):
    x = _f64(x)
    h = n // 2 + 1
    out = np.zeros(n * n * h * howmany * 2)
    lib().orc_rfft3(n, howmany, _p(x), _p(out))
    return out


def irfft3(X, n, howmany):
    X = _f64(X)
    out = np.zeros(n ** 3 * howmany)
    lib().orc_irfft3(n, howmany, _p(X), _p(out))
    return out


def pinv_truncated(M, eps):
    """Truncated pseudo-inverse: keep sigma_i >= eps*sigma_0 (common/vecmatop.cpp:129-164)."""
    U, S, VT = np.linalg.svd(M, full_matrices=False)
    keep = S >= eps * S[0]
    return (VT[keep].T / S[keep]) @ U[:, keep].T


def check_matrix(kt_mm, coefs, np_, which):
    sd, td = src_dof(kt_mm), trg_dof(kt_mm)
    nden = np_ ** 3 - (np_ - 2) ** 3
    nchk = ((np_ + 2) ** 3 - np_ ** 3) if which == 0 else nden
    out = np.zeros((nden * sd, nchk * td))
    c = coefs_arr(coefs)
    rc = lib().orc_check_matrix(kt_mm, _p(c), len(coefs), np_, which, _p(out))
    assert rc == 0
    return out.T.copy()


def numpy_pinv_tables(kt_mm, coefs, np_):
    """Oracle-side pinv tables: (P_up, P_dn) with cutoff 0.1^(np+1) (matmgnt3d_mpi.cpp:37-40)."""
    eps = 0.1 ** (np_ + 1)
    return (pinv_truncated(check_matrix(kt_mm, coefs, np_, 0), eps),
            pinv_truncated(check_matrix(kt_mm, coefs, np_, 1), eps))


class Tree:
    """Octree + U/V/W/X lists built by the oracle (Dendro rule + let3d_mpi.cpp)."""

    def __init__(self, spos, tpos=None, ctr=(0.5, 0.5, 0.5), root_level=1, pts_max=64,
                 max_level=30):
        spos = _f64(spos).reshape(-1, 3)
        tpos = spos if tpos is None else _f64(tpos).reshape(-1, 3)
        self._keep = (spos, tpos)
        self._t = _Tree()
        c = _f64(ctr)
        rc = lib().orc_tree_build(C.byref(self._t), len(spos), _p(spos), len(tpos), _p(tpos),
                                  _p(c), root_level, pts_max, max_level)
        if rc != 0:
            raise ValueError("orc_tree_build failed")
        t = self._t
        n = t.n_nodes
        self.n_nodes = n

        def arr(ptr, m, dt=np.int32):
            return np.ctypeslib.as_array(ptr, shape=(m,)).astype(dt).copy() if m > 0 else np.zeros(0, dt)
        self.parent = arr(t.parent, n)
        self.child = arr(t.child, n)
        self.depth = arr(t.depth, n)
        self.path = arr(t.path, 3 * n).reshape(n, 3)
        self.tag = arr(t.tag, n)
        self.src_cnt = arr(t.src_cnt, n)
        self.trg_cnt = arr(t.trg_cnt, n)
        self.src_beg = arr(t.src_beg, n)
        self.src_num = arr(t.src_num, n)
        self.trg_beg = arr(t.trg_beg, n)
        self.trg_num = arr(t.trg_num, n)
        self.src_perm = arr(t.src_perm, t.n_src)
        self.trg_perm = arr(t.trg_perm, t.n_trg)
        self.lists = {}
        for name in "uvwx":
            off = arr(getattr(t, name + "_off"), n + 1, np.int64)
            lst = arr(getattr(t, name + "_lst"), int(off[-1]))
            self.lists[name] = (off, lst)

    def list_of(self, name, i):
        off, lst = self.lists[name]
        return lst[off[i]:off[i + 1]]

    def __del__(self):
        try:
            lib().orc_tree_free(C.byref(self._t))
        except Exception:
            pass


class Ops:
    """Translation operators (matmgnt3d_mpi.cpp) with the given pinv tables."""

    def __init__(self, kt_mm, coefs, np_, pinv_up, pinv_dn):
        self._o = _Ops()
        c = coefs_arr(coefs)
        up = np.asfortranarray(pinv_up, dtype=np.float64)
        dn = np.asfortranarray(pinv_dn, dtype=np.float64)
        rc = lib().orc_ops_init(C.byref(self._o), kt_mm, _p(c), len(coefs), np_,
                                up.ctypes.data_as(P_D), dn.ctypes.data_as(P_D))
        if rc != 0:
            raise ValueError("orc_ops_init failed for kernel %d" % kt_mm)
        o = self._o
        self.sd, self.td, self.np = o.sd, o.td, np_
        self.n_ue, self.n_uc = o.n_ue, o.n_uc
        H = (np_ + 1) * (2 * np_) ** 2
        szuc, szue = o.n_uc * o.td, o.n_ue * o.sd
        szdc, szde = o.n_dc * o.td, o.n_de * o.sd
        self.m2m = np.ctypeslib.as_array(o.m2m, shape=(8, szue, szuc)).transpose(0, 2, 1).copy()
        self.l2l = np.ctypeslib.as_array(o.l2l, shape=(8, szde, szdc)).transpose(0, 2, 1).copy()
        m2l = np.ctypeslib.as_array(o.m2l, shape=(343, H * o.sd * o.td * 2)).copy()
        self.m2l = m2l.view(np.complex128).reshape(343, H, o.sd * o.td)

    def __del__(self):
        try:
            lib().orc_ops_free(C.byref(self._o))
        except Exception:
            pass


def m2l_fft(ops, idx, level, ue):
    """One M2L through the grid: DC of a target box from the UE of the same-level
    source box at offset idx (matmgnt3d_mpi.cpp:276-431)."""
    H = (ops.np + 1) * (2 * ops.np) ** 2
    ue = _f64(ue)
    eff = np.zeros(2 * H * ops.sd)
    val = np.zeros(2 * H * ops.td)
    dc = np.zeros(ops.n_ue * ops.td)
    lib().orc_eff_den(C.byref(ops._o), int(level), _p(ue), _p(eff))
    ix = (C.c_int * 3)(*[int(v) for v in idx])
    lib().orc_m2l_hadamard(C.byref(ops._o), ix, _p(eff), _p(val))
    lib().orc_eff_val(C.byref(ops._o), _p(val), _p(dc))
    return dc


def evaluate(tree, ops, kt, coefs, spos, snor, tpos, den, root_level=1, ctr=(0.5, 0.5, 0.5),
             return_equiv=False):
    """One FMM evaluate (fmm3d_eval_mpi.cpp:396-1515 CPU branches)."""
    spos, snor, tpos, den = (_f64(a) for a in (spos, snor, tpos, den))
    TD = trg_dof(kt)
    out = np.zeros(len(tpos.reshape(-1, 3)) * TD)
    c = coefs_arr(coefs)
    cc = _f64(ctr)
    ue = de = None
    pu = pd = None
    if return_equiv:
        ue = np.zeros(tree.n_nodes * ops.n_ue * ops.sd)
        de = np.zeros(tree.n_nodes * ops.n_ue * ops.sd)
        pu, pd = _p(ue), _p(de)
    rc = lib().orc_evaluate(C.byref(tree._t), C.byref(ops._o), kt, _p(c), root_level, _p(cc),
                            _p(spos), _p(snor), _p(tpos), _p(den), _p(out), pu, pd)
    if rc != 0:
        raise ValueError("orc_evaluate failed")
    if return_equiv:
        return out, ue, de
    return out


def direct(kt, coefs, spos, snor, den, tpos, nthreads=8):
    """Direct sum (dense3d_mpi.cpp:122-129), r < 1e-8 pairs excluded."""
    spos, snor, den, tpos = (_f64(a) for a in (spos, snor, den, tpos))
    TD = trg_dof(kt)
    nt = len(tpos.reshape(-1, 3))
    out = np.zeros(nt * TD)
    c = coefs_arr(coefs)
    rc = lib().orc_direct(kt, _p(c), len(spos.reshape(-1, 3)), _p(spos), _p(snor), _p(den), nt,
                          _p(tpos), _p(out), nthreads)
    assert rc == 0
    return out
