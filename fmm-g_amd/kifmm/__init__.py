"""Python binding of libkifmm.so (the MI355X KIFMM3d evaluate path).

A thin ctypes layer over the C ABI in ``include/kifmm.h``.  ``FMM3d`` mirrors the
reference's ``FMM3d_MPI`` operator interface (fmm3d_mpi/fmm3d_mpi.hpp:36-249,
knlmat3d_mpi.hpp:29-80): configure, ``set_points``, ``setup()``, then
``evaluate()`` any number of times.  Every failure raises ``KifmmError`` with the
library's message; there is no fallback path: if the native library is missing
the import of a handle fails.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libkifmm.so")
_LIB = None

P_D = C.POINTER(C.c_double)
P_F = C.POINTER(C.c_float)
P_I = C.POINTER(C.c_int)
P_L = C.POINTER(C.c_int64)
H = C.c_void_p

PREC = {"fp64": 0, "mixed": 1}
PHASES = ["EvlIni", "EvlUpwComp", "EvlVList", "EvlXList", "EvlCombine", "EvlUList", "EvlFinalize", "VListHadamard"]
TUNING = ["near_mode", "near_stage", "translations", "vec_hadamard"]
SETUP_PARTS = ["tree", "tree_device", "tree_copy", "operators", "tasks", "upload"]
WORK = ["near_pairs", "near_exact_pairs", "s2m_pairs", "s2l_pairs", "vlist_pairs", "n_vsrc",
        "n_vtrg", "n_vgrp", "up_cols", "m2m_cols", "dn_cols", "l2l_cols"]

# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "kifmm_create": (C.c_int, [C.c_int, P_D, C.c_int, C.c_int, C.c_int, C.c_int, P_D, C.c_int,
                               C.c_int, C.POINTER(H)]),
    "kifmm_create_from_options": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(H)]),
    "kifmm_destroy": (C.c_int, [H]),
    "kifmm_set_points": (C.c_int, [H, C.c_int64, P_D, P_D, C.c_int64, P_D]),
    "kifmm_set_points_batch": (C.c_int, [H, C.c_int, P_L, P_D, P_D, P_L, P_D]),
    "kifmm_setup": (C.c_int, [H]),
    "kifmm_setup_host": (C.c_int, [H]),
    "kifmm_src_dof": (C.c_int, [H]),
    "kifmm_trg_dof": (C.c_int, [H]),
    "kifmm_evaluate": (C.c_int, [H, P_D, P_D]),
    "kifmm_evaluate_device": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kifmm_evaluate_multi": (C.c_int, [H, P_D, P_D, C.c_int]),
    "kifmm_evaluate_multi_device": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "kifmm_set_graph": (C.c_int, [H, C.c_int]),
    "kifmm_evaluate_adjoint": (C.c_int, [H, P_D, P_D]),
    "kifmm_field": (C.c_int, [H, P_D, P_D]),
    "kifmm_field_device": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kifmm_evaluate_adjoint_device": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kifmm_checkpoint_size": (C.c_int, [H, P_L]),
    "kifmm_evaluate_checkpoint_device": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kifmm_field_checkpoint_device": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kifmm_direct_device": (C.c_int, [H, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kifmm_direct_device_noalloc": (C.c_int, [H, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kifmm_check": (C.c_int, [H, P_D, P_D, C.c_int, C.c_uint64, P_D]),
    "kifmm_tree_info": (C.c_int, [H, P_L]),
    "kifmm_get_tree": (C.c_int, [H] + [P_I] * 11),
    "kifmm_get_list": (C.c_int, [H, C.c_int, P_L, P_I]),
    "kifmm_get_operator": (C.c_int, [H, C.c_int, P_D, P_L]),
    "kifmm_set_timing": (C.c_int, [H, C.c_int]),
    "kifmm_phase_times": (C.c_int, [H, P_F, C.c_int]),
    "kifmm_phase_name": (C.c_char_p, [C.c_int]),
    "kifmm_work_counts": (C.c_int, [H, P_D, C.c_int]),
    "kifmm_set_tree_builder": (C.c_int, [H, C.c_int]),
    "kifmm_setup_times": (C.c_int, [H, P_D, C.c_int]),
    "kifmm_set_translation": (C.c_int, [H, C.c_int]),
    "kifmm_autotune": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "kifmm_get_tuning": (C.c_int, [H, P_I, C.c_int]),
    "kifmm_set_tuning": (C.c_int, [H, P_I, C.c_int]),
    "kifmm_tuning_report": (C.c_int, [H, P_D, C.c_int]),
    "kifmm_options_lookup": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]),
    "kifmm_last_error": (C.c_char_p, [H]),
}


class KifmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("kifmm error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libkifmm.so (built in-tree by ``make -C fmm-g_amd``)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise KifmmError(-1, "native library not built: %s" % LIB_PATH)
        # PyTorch ships its own HIP runtime.  Loading it before libkifmm makes the
        # library bind to the runtime torch uses, so torch tensors, streams and
        # the library share one runtime whichever module the caller imports first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


def _pd(a):
    return a.ctypes.data_as(P_D) if a is not None else None


def _dev(t, n, what):
    """Device pointer of a float64, contiguous GPU tensor holding exactly n
    values; anything else raises before the library sees the pointer."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor" % what)
    if not t.is_cuda:
        raise TypeError("%s must be a GPU tensor" % what)
    if t.dtype != torch.float64:
        raise TypeError("%s must be float64, got %s" % (what, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % what)
    if t.numel() != n:
        raise ValueError("%s must hold %d values, got %d" % (what, n, t.numel()))
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    return C.c_void_p(stream) if stream else None


def options_lookup(path, name):
    """Value of option `name` in a PETSc-style options file ('' for a flag, None if absent)."""
    buf = C.create_string_buffer(1024)
    rc = lib().kifmm_options_lookup(path.encode(), name.encode(), buf, 1024)
    if rc < 0:
        raise KifmmError(rc, lib().kifmm_last_error(None).decode())
    return buf.value.decode() if rc == 1 else None


class FMM3d:
    """KIFMM3d evaluate operator on one GPU (reference: FMM3d_MPI)."""

    def __init__(self, kernel_type=111, coefs=(1.0, 0.25), np=6, pts_max=64, max_level=30,
                 ctr=(0.5, 0.5, 0.5), root_level=1, precision="fp64", _handle=None):
        L = lib()
        self._h = H()
        if _handle is not None:
            self._h = _handle
        else:
            c = _f64(coefs)
            cc = _f64(ctr, 3)
            rc = L.kifmm_create(int(kernel_type), _pd(c), c.size, int(np), int(pts_max),
                                int(max_level), _pd(cc), int(root_level), PREC[precision],
                                C.byref(self._h))
            if rc != 0:
                raise KifmmError(rc, L.kifmm_last_error(None).decode())
        self.kernel_type = kernel_type
        self.n_src = self.n_trg = 0

    @classmethod
    def from_options(cls, path, prefix=None):
        L = lib()
        h = H()
        rc = L.kifmm_create_from_options(path.encode(), prefix.encode() if prefix else None,
                                         C.byref(h))
        if rc != 0:
            raise KifmmError(rc, L.kifmm_last_error(None).decode())
        return cls(_handle=h)

    def _chk(self, rc):
        if rc != 0:
            raise KifmmError(rc, lib().kifmm_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().kifmm_destroy(self._h)
            self._h = H()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def src_dof(self):
        return lib().kifmm_src_dof(self._h)

    @property
    def trg_dof(self):
        return lib().kifmm_trg_dof(self._h)

    def set_points(self, src_pos, src_nor=None, trg_pos=None):
        sp = _f64(src_pos).reshape(-1)
        tp = sp if trg_pos is None else _f64(trg_pos).reshape(-1)
        sn = None if src_nor is None else _f64(src_nor, sp.size)
        self.n_src, self.n_trg = sp.size // 3, tp.size // 3
        self._chk(lib().kifmm_set_points(self._h, self.n_src, _pd(sp), _pd(sn), self.n_trg, _pd(tp)))

    def set_points_batch(self, src_pos, src_nor=None, trg_pos=None):
        """Many independent problems in one handle: src_pos (and src_nor,
        trg_pos) are lists with one (n_p, 3) array per problem; trg_pos None
        means targets = sources.  Densities and values are then the problems'
        vectors one after another."""
        ns = np.array([len(p) for p in src_pos], np.int64)
        sp = _f64(np.concatenate([np.asarray(p, np.float64).reshape(-1, 3) for p in src_pos])).reshape(-1)
        sn = None
        if src_nor is not None:
            sn = _f64(np.concatenate([np.asarray(p, np.float64).reshape(-1, 3) for p in src_nor]), sp.size)
        if trg_pos is None:
            nt, tp = ns, None
        else:
            nt = np.array([len(p) for p in trg_pos], np.int64)
            tp = _f64(np.concatenate([np.asarray(p, np.float64).reshape(-1, 3) for p in trg_pos])).reshape(-1)
        self.n_src, self.n_trg = int(ns.sum()), int(nt.sum())
        self.problem_src, self.problem_trg = ns, nt
        self._chk(lib().kifmm_set_points_batch(self._h, len(ns), ns.ctypes.data_as(P_L), _pd(sp), _pd(sn),
                                               nt.ctypes.data_as(P_L), _pd(tp)))

    def setup(self):
        self._chk(lib().kifmm_setup(self._h))

    def setup_host(self):
        self._chk(lib().kifmm_setup_host(self._h))

    def evaluate(self, src_den):
        """Host arrays in, host array out (FMM3d_MPI::evaluate)."""
        d = _f64(src_den, self.n_src * self.src_dof)
        out = np.zeros(self.n_trg * self.trg_dof)
        self._chk(lib().kifmm_evaluate(self._h, _pd(d), _pd(out)))
        return out

    def evaluate_device(self, d_src_den, d_trg_val, stream=None):
        """Device-resident float64 tensors (e.g. torch CUDA/HIP tensors)."""
        self._chk(lib().kifmm_evaluate_device(self._h, _dev(d_src_den, self.n_src * self.src_dof, "densities"),
                                              _dev(d_trg_val, self.n_trg * self.trg_dof, "values"),
                                              _stream(stream)))

    def evaluate_multi(self, src_den):
        """nrhs density vectors (rows of src_den) -> nrhs value vectors."""
        d = np.ascontiguousarray(src_den, dtype=np.float64).reshape(-1, self.n_src * self.src_dof)
        out = np.zeros((d.shape[0], self.n_trg * self.trg_dof))
        self._chk(lib().kifmm_evaluate_multi(self._h, _pd(d), _pd(out), d.shape[0]))
        return out

    def evaluate_multi_device(self, d_src_den, d_trg_val, nrhs, stream=None):
        """nrhs right-hand sides in one pass on device tensors (nrhs vectors one
        after another)."""
        self._chk(lib().kifmm_evaluate_multi_device(
            self._h, _dev(d_src_den, nrhs * self.n_src * self.src_dof, "densities"),
            _dev(d_trg_val, nrhs * self.n_trg * self.trg_dof, "values"), int(nrhs), _stream(stream)))

    def evaluate_adjoint(self, trg_in):
        """Transposed pass A^T w: trgDOF per target in, srcDOF per source out."""
        w = _f64(trg_in, self.n_trg * self.trg_dof)
        out = np.zeros(self.n_src * self.src_dof)
        self._chk(lib().kifmm_evaluate_adjoint(self._h, _pd(w), _pd(out)))
        return out

    def evaluate_adjoint_device(self, d_trg_in, d_src_out, stream=None):
        """Transposed pass on device-resident float64 tensors."""
        self._chk(lib().kifmm_evaluate_adjoint_device(self._h, _dev(d_trg_in, self.n_trg * self.trg_dof, "input"),
                                                      _dev(d_src_out, self.n_src * self.src_dof, "output"),
                                                      _stream(stream)))

    def field(self, src_den):
        """Gradient of the potential at the targets, shape (n_trg, 3)."""
        d = _f64(src_den, self.n_src * self.src_dof)
        out = np.zeros(self.n_trg * 3)
        self._chk(lib().kifmm_field(self._h, _pd(d), _pd(out)))
        return out.reshape(-1, 3)

    def field_device(self, d_src_den, d_trg_grad, stream=None):
        """Gradient pass on device-resident float64 tensors (3 per target)."""
        self._chk(lib().kifmm_field_device(self._h, _dev(d_src_den, self.n_src * self.src_dof, "densities"),
                                           _dev(d_trg_grad, self.n_trg * 3, "gradients"),
                                           _stream(stream)))

    def checkpoint_size(self):
        """Doubles of an upward-equivalent-density checkpoint."""
        n = C.c_int64()
        self._chk(lib().kifmm_checkpoint_size(self._h, C.byref(n)))
        return n.value

    def evaluate_checkpoint_device(self, d_src_den, d_trg_val, d_ue, stream=None):
        """evaluate_device that also writes the upward equivalent densities to
        d_ue (checkpoint_size() float64 values)."""
        self._chk(lib().kifmm_evaluate_checkpoint_device(
            self._h, _dev(d_src_den, self.n_src * self.src_dof, "densities"),
            _dev(d_trg_val, self.n_trg * self.trg_dof, "values"), _dev(d_ue, self.checkpoint_size(), "checkpoint"),
            _stream(stream)))

    def field_checkpoint_device(self, d_src_den, d_ue, d_trg_grad, stream=None):
        """field_device with the up pass taken from a checkpoint of the same
        densities."""
        self._chk(lib().kifmm_field_checkpoint_device(
            self._h, _dev(d_src_den, self.n_src * self.src_dof, "densities"),
            _dev(d_ue, self.checkpoint_size(), "checkpoint"), _dev(d_trg_grad, self.n_trg * 3, "gradients"),
            _stream(stream)))

    def direct_device(self, d_src_den, d_chk_pos, d_chk_val, stream=None):
        """FP64 direct sum at the check positions (device tensors)."""
        if d_chk_pos.numel() % 3:
            raise ValueError("check positions must hold 3 values per point")
        n = d_chk_pos.numel() // 3
        self._chk(lib().kifmm_direct_device(self._h, _dev(d_src_den, self.n_src * self.src_dof, "densities"), n,
                                            _dev(d_chk_pos, 3 * n, "check positions"),
                                            _dev(d_chk_val, n * self.trg_dof, "check values"),
                                            _stream(stream)))

    def direct_device_noalloc(self, d_src_den, d_chk_pos, d_chk_val, stream=None):
        """direct_device with the library's fixed per-GPU scratch: allocates
        nothing (check targets * trg_dof <= 2^21)."""
        if d_chk_pos.numel() % 3:
            raise ValueError("check positions must hold 3 values per point")
        n = d_chk_pos.numel() // 3
        self._chk(lib().kifmm_direct_device_noalloc(
            self._h, _dev(d_src_den, self.n_src * self.src_dof, "densities"), n,
            _dev(d_chk_pos, 3 * n, "check positions"), _dev(d_chk_val, n * self.trg_dof, "check values"),
            _stream(stream)))

    def check(self, src_den, trg_val, numchk, seed=0):
        d = _f64(src_den)
        v = _f64(trg_val)
        r = C.c_double()
        self._chk(lib().kifmm_check(self._h, _pd(d), _pd(v), int(numchk), int(seed), C.byref(r)))
        return r.value

    def tree(self):
        info = np.zeros(8, np.int64)
        self._chk(lib().kifmm_tree_info(self._h, info.ctypes.data_as(P_L)))
        n, ns, nt = int(info[0]), int(info[6]), int(info[7])
        names = ["parent", "child", "depth", "path", "tag", "src_beg", "src_num", "trg_beg",
                 "trg_num", "src_perm", "trg_perm"]
        sizes = [n, n, n, 3 * n, n, n, n, n, n, ns, nt]
        arrs = [np.zeros(max(s, 1), np.int32) for s in sizes]
        self._chk(lib().kifmm_get_tree(self._h, *[a.ctypes.data_as(P_I) for a in arrs]))
        out = {k: a[:s] for k, a, s in zip(names, arrs, sizes)}
        out["path"] = out["path"].reshape(n, 3)
        out["max_depth"] = int(info[1])
        out["lists"] = {}
        for kind, name in enumerate("uvwx"):
            off = np.zeros(n + 1, np.int64)
            lst = np.zeros(max(int(info[2 + kind]), 1), np.int32)
            self._chk(lib().kifmm_get_list(self._h, kind, off.ctypes.data_as(P_L),
                                           lst.ctypes.data_as(P_I)))
            out["lists"][name] = (off, lst[:int(info[2 + kind])])
        return out

    def operator(self, which):
        cnt = C.c_int64()
        self._chk(lib().kifmm_get_operator(self._h, which, None, C.byref(cnt)))
        out = np.zeros(cnt.value)
        self._chk(lib().kifmm_get_operator(self._h, which, _pd(out), C.byref(cnt)))
        return out

    def set_graph(self, on=True):
        """Replay evaluate_device from a captured hipGraph (same buffers, same
        non-default stream)."""
        self._chk(lib().kifmm_set_graph(self._h, 1 if on else 0))

    def set_timing(self, on=True):
        self._chk(lib().kifmm_set_timing(self._h, 1 if on else 0))

    def phase_times(self):
        ms = (C.c_float * len(PHASES))()
        self._chk(lib().kifmm_phase_times(self._h, ms, len(PHASES)))
        return dict(zip(PHASES, list(ms)))

    def work_counts(self):
        w = np.zeros(len(WORK))
        self._chk(lib().kifmm_work_counts(self._h, _pd(w), len(WORK)))
        return dict(zip(WORK, w.tolist()))

    def set_tree_builder(self, on_device=True):
        """setup() builds the octree and lists on the GPU (default) or on the host."""
        self._chk(lib().kifmm_set_tree_builder(self._h, 1 if on_device else 0))

    def setup_times(self):
        """Wall times of the last setup, milliseconds, by part."""
        ms = np.zeros(len(SETUP_PARTS))
        self._chk(lib().kifmm_setup_times(self._h, _pd(ms), len(SETUP_PARTS)))
        return dict(zip(SETUP_PARTS, ms.tolist()))

    def set_translation(self, terms):
        """Mixed-path translations: 0 FP64 matrix cores (M2M fused into the
        pseudo-inverse), 3 or 6 split-BF16 products; setup() again afterwards."""
        self._chk(lib().kifmm_set_translation(self._h, int(terms)))

    def autotune(self, d_src_den, d_trg_val, nrhs=1, stream=None):
        """Times the mixed path's launch-shape choices on this tree and keeps the
        fastest (device tensors, nrhs right-hand sides); returns the chosen
        knobs and every trial as (knob, value, ms)."""
        self._chk(lib().kifmm_autotune(self._h, _dev(d_src_den, nrhs * self.n_src * self.src_dof, "densities"),
                                       _dev(d_trg_val, nrhs * self.n_trg * self.trg_dof, "values"), int(nrhs),
                                       _stream(stream)))
        n = lib().kifmm_tuning_report(self._h, None, 0)
        rows = np.zeros(3 * max(n, 1))
        lib().kifmm_tuning_report(self._h, _pd(rows), n)
        trials = [(TUNING[int(rows[3 * i])], int(rows[3 * i + 1]), float(rows[3 * i + 2])) for i in range(n)]
        return {"tuning": self.tuning(), "trials": trials}

    def tuning(self):
        k = (C.c_int * len(TUNING))()
        self._chk(lib().kifmm_get_tuning(self._h, k, len(TUNING)))
        return dict(zip(TUNING, list(k)))

    def set_tuning(self, near_mode=None, near_stage=None, translations=None, vec_hadamard=None):
        cur = self.tuning()
        for name, v in (("near_mode", near_mode), ("near_stage", near_stage), ("translations", translations),
                        ("vec_hadamard", vec_hadamard)):
            if v is not None:
                cur[name] = int(v)
        k = (C.c_int * len(TUNING))(*[cur[n] for n in TUNING])
        self._chk(lib().kifmm_set_tuning(self._h, k, len(TUNING)))
