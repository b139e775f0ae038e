"""Python binding of libtfmm.so (the C ABI in include/tfmm.h) via ctypes.

This is the host-side mirror of the Tapas FMM driver sequence for Python
callers (tests, bench): ``Partition`` (tree build), ``upward``, ``traverse``,
``downward`` and the result accessors, the same phase split as the
reference's tapas_exafmm.cxx:557-641.  It loads the in-tree HIP library and
fails loudly if it is missing; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtfmm.so")

TFMM_OK = 0
FP64 = 0
MIXED = 1
TUNE_M2L_UNIT = 1
TUNE_M2L_XCD_RUN = 2
TUNE_P2P_SCHEDULE = 3
TUNE_GRAPH = 4
TUNE_P2P_WAVES = 5
TUNE_DEVICE_LIMIT = 6
TUNE_CART_M2L = 7
DIRECT_NOALLOC_MAX = 65536
PERIODIC_MAX_IMAGES = 6
P = 10
NTERM = 55
BASIS_SPHERICAL, BASIS_CARTESIAN = 0, 1
CART_TERMS = 220  # P (P+1) (P+2) / 6 real terms of a Cartesian expansion

_lib = None


class TfmmError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("ncrit", ctypes.c_int32), ("mutual", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("theta", ctypes.c_double)]


class BodyDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("stride", ctypes.c_size_t), ("off_x", ctypes.c_size_t),
                ("off_q", ctypes.c_size_t)]


class Stats(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("ncells", ctypes.c_int64), ("nleaves", ctypes.c_int64),
                ("nlevels", ctypes.c_int32), ("max_depth", ctypes.c_int32),
                ("region_min", ctypes.c_double * 3), ("region_max", ctypes.c_double * 3),
                ("n_m2l", ctypes.c_uint64), ("n_m2l_exec", ctypes.c_uint64),
                ("n_p2p_leafpairs", ctypes.c_uint64), ("n_p2p_calls", ctypes.c_uint64),
                ("n_p2p_self", ctypes.c_uint64),
                ("ms_tree", ctypes.c_double), ("ms_upward", ctypes.c_double),
                ("ms_traverse", ctypes.c_double), ("ms_downward", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("ms_p2p", ctypes.c_double),
                ("ms_m2l", ctypes.c_double), ("ms_dtt", ctypes.c_double),
                ("n_m2l_groups", ctypes.c_uint64), ("n_m2l_items", ctypes.c_uint64),
                ("n_m2l_fallback", ctypes.c_uint64), ("ms_m2l_tiles", ctypes.c_double),
                ("n_problems", ctypes.c_int64), ("one_graph", ctypes.c_int64)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if isinstance(v, ctypes.Array) else v
        return out


EXPORTED = [
    "tfmm_create", "tfmm_destroy", "tfmm_last_error", "tfmm_stream", "tfmm_sync",
    "tfmm_set_bodies", "tfmm_set_bodies_device", "tfmm_set_batch", "tfmm_set_batch_device",
    "tfmm_batch_evaluate", "tfmm_num_problems", "tfmm_set_tuning", "tfmm_get_tuning",
    "tfmm_autotune",
    "tfmm_build_tree", "tfmm_set_params",
    "tfmm_upward",
    "tfmm_traverse", "tfmm_downward", "tfmm_evaluate", "tfmm_get_fields", "tfmm_get_stats",
    "tfmm_get_sorted", "tfmm_get_cells", "tfmm_get_expansions", "tfmm_get_lists", "tfmm_direct",
    "tfmm_direct_noalloc", "tfmm_traverse_lists", "tfmm_set_periodic", "tfmm_set_basis",
    "tfmm_get_basis",
    "tfmm_picker_nparam", "tfmm_picker_create", "tfmm_picker_destroy", "tfmm_picker_last_error",
    "tfmm_picker_set_weights", "tfmm_picker_get_weights", "tfmm_picker_set_theta_max",
    "tfmm_picker_train",
    "tfmm_picker_gradient", "tfmm_picker_predict", "tfmm_tree_features", "tfmm_pick_params",
]

NCRIT_CLASSES = (16, 32, 64, 128, 256)


def lib():
    """Load libtfmm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TfmmError("libtfmm.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.tfmm_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.tfmm_destroy.argtypes = [vp]
        L.tfmm_destroy.restype = None
        L.tfmm_last_error.argtypes = [vp]
        L.tfmm_last_error.restype = ctypes.c_char_p
        L.tfmm_stream.argtypes = [vp]
        L.tfmm_stream.restype = vp
        L.tfmm_sync.argtypes = [vp]
        L.tfmm_set_bodies.argtypes = [vp, vp, ctypes.POINTER(BodyDesc)]
        L.tfmm_set_bodies_device.argtypes = [vp, ctypes.c_int64, vp, vp, vp, vp]
        L.tfmm_set_batch.argtypes = [vp, ctypes.c_int, ctypes.POINTER(BodyDesc),
                                     ctypes.POINTER(vp)]
        L.tfmm_set_batch_device.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp]
        L.tfmm_batch_evaluate.argtypes = [vp, ctypes.c_int, ctypes.POINTER(BodyDesc),
                                          ctypes.POINTER(vp), ctypes.POINTER(Params)]
        L.tfmm_num_problems.argtypes = [vp, vp]
        L.tfmm_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.tfmm_set_periodic.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.tfmm_set_basis.argtypes = [vp, ctypes.c_int]
        L.tfmm_get_basis.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.tfmm_get_tuning.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.tfmm_autotune.argtypes = [vp, ctypes.POINTER(Params), ctypes.c_int, vp]
        L.tfmm_build_tree.argtypes = [vp, ctypes.POINTER(Params)]
        L.tfmm_evaluate.argtypes = [vp, ctypes.POINTER(Params)]
        L.tfmm_set_params.argtypes = [vp, ctypes.POINTER(Params)]
        for f in ("tfmm_upward", "tfmm_traverse", "tfmm_downward"):
            getattr(L, f).argtypes = [vp]
        L.tfmm_traverse_lists.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.c_int64]
        L.tfmm_get_fields.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
        L.tfmm_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.tfmm_get_sorted.argtypes = [vp, vp, vp]
        L.tfmm_get_cells.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.tfmm_get_expansions.argtypes = [vp, vp, vp]
        L.tfmm_get_lists.argtypes = [vp, vp, vp, vp, vp]
        L.tfmm_direct.argtypes = [vp, vp, ctypes.c_int64, vp]
        L.tfmm_direct_noalloc.argtypes = [vp, vp, ctypes.c_int64, vp]
        L.tfmm_picker_nparam.restype = ctypes.c_int64
        L.tfmm_picker_nparam.argtypes = []
        L.tfmm_picker_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.tfmm_picker_destroy.argtypes = [vp]
        L.tfmm_picker_destroy.restype = None
        L.tfmm_picker_last_error.argtypes = [vp]
        L.tfmm_picker_last_error.restype = ctypes.c_char_p
        L.tfmm_picker_set_weights.argtypes = [vp, vp, ctypes.c_int64]
        L.tfmm_picker_set_theta_max.argtypes = [vp, ctypes.c_float]
        L.tfmm_picker_get_weights.argtypes = [vp, vp, ctypes.c_int64]
        L.tfmm_picker_train.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                        vp]
        L.tfmm_picker_gradient.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp, vp]
        L.tfmm_picker_predict.argtypes = [vp, vp, ctypes.c_int, vp]
        L.tfmm_tree_features.argtypes = [vp, vp]
        L.tfmm_pick_params.argtypes = [vp, vp, ctypes.POINTER(Params)]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def make_params(ncrit=64, theta=0.4, precision=FP64, mutual=0):
    return Params(int(ncrit), int(mutual), int(precision), 0, float(theta))


class Context:
    """One tfmm_ctx: a device, a stream and the buffers of one problem."""

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.tfmm_create(int(device), ctypes.byref(h))
        if rc != TFMM_OK:
            raise TfmmError("tfmm_create failed with status %d" % rc)
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().tfmm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- helpers ---------------------------------------------------------
    def _check(self, rc, what):
        if rc != TFMM_OK:
            msg = lib().tfmm_last_error(self._h)
            raise TfmmError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    @property
    def stream(self):
        return lib().tfmm_stream(self._h)

    def sync(self):
        self._check(lib().tfmm_sync(self._h), "tfmm_sync")

    # -- bodies ----------------------------------------------------------
    def set_bodies(self, x, y, z, q):
        """Copy host bodies (through the reference's 96-byte Body layout:
        X at offset 0, SRC at 24)."""
        n = len(x)
        body = np.zeros((n, 12), np.float64)  # 96 bytes per body
        body[:, 0] = x
        body[:, 1] = y
        body[:, 2] = z
        body[:, 3] = q
        desc = BodyDesc(n, 96, 0, 24)
        self._check(lib().tfmm_set_bodies(self._h, _ptr(body), ctypes.byref(desc)),
                    "tfmm_set_bodies")
        self.n = n

    def set_bodies_device(self, n, px, py, pz, pq):
        """Use device-resident FP64 SoA arrays (raw device pointers)."""
        self._check(lib().tfmm_set_bodies_device(self._h, int(n), px, py, pz, pq),
                    "tfmm_set_bodies_device")
        self.n = int(n)

    @staticmethod
    def _aos(x, y, z, q):
        body = np.zeros((len(x), 12), np.float64)  # the reference's 96-byte Body
        body[:, 0] = x
        body[:, 1] = y
        body[:, 2] = z
        body[:, 3] = q
        return body

    def set_batch(self, problems):
        """Copy a batch of independent problems, a list of (x, y, z, q),
        evaluated together as one forest (bodies concatenated in order)."""
        bodies = [self._aos(*p) for p in problems]
        descs = (BodyDesc * len(bodies))(*[BodyDesc(len(b), 96, 0, 24) for b in bodies])
        ptrs = (ctypes.c_void_p * len(bodies))(*[b.ctypes.data for b in bodies])
        self._check(lib().tfmm_set_batch(self._h, len(bodies), descs, ptrs), "tfmm_set_batch")
        self.n = sum(len(b) for b in bodies)

    def set_batch_device(self, offsets, px, py, pz, pq):
        """Use device-resident FP64 SoA arrays holding len(offsets) - 1
        problems, problem p = [offsets[p], offsets[p+1])."""
        off = np.ascontiguousarray(offsets, np.int64)
        self._check(lib().tfmm_set_batch_device(self._h, len(off) - 1, _ptr(off), px, py, pz,
                                                pq), "tfmm_set_batch_device")
        self.n = int(off[-1])

    def set_periodic(self, images, cycle=0.0):
        """periodic images (0: none) of period cycle; see tfmm_set_periodic"""
        self._check(lib().tfmm_set_periodic(self._h, int(images), float(cycle)),
                    "tfmm_set_periodic")

    def set_basis(self, basis):
        """BASIS_SPHERICAL or BASIS_CARTESIAN; see tfmm_set_basis"""
        self._check(lib().tfmm_set_basis(self._h, int(basis)), "tfmm_set_basis")

    def basis(self):
        v = ctypes.c_int()
        self._check(lib().tfmm_get_basis(self._h, ctypes.byref(v)), "tfmm_get_basis")
        return v.value

    def set_tuning(self, knob, value):
        self._check(lib().tfmm_set_tuning(self._h, int(knob), float(value)), "tfmm_set_tuning")

    def get_tuning(self, knob):
        v = ctypes.c_double()
        self._check(lib().tfmm_get_tuning(self._h, int(knob), ctypes.byref(v)), "tfmm_get_tuning")
        return v.value

    def autotune(self, params, reps=3):
        """Pick the kernel knobs for these bodies; returns (unit factor, XCD
        run, P2P schedule, ms per evaluation)."""
        out = np.zeros(4)
        self._check(lib().tfmm_autotune(self._h, ctypes.byref(params), int(reps), _ptr(out)),
                    "tfmm_autotune")
        return float(out[0]), int(out[1]), int(out[2]), float(out[3])

    def problem_offsets(self):
        n = lib().tfmm_num_problems(self._h, None)
        self._check(min(n, 0), "tfmm_num_problems")
        off = np.zeros(n + 1, np.int64)
        lib().tfmm_num_problems(self._h, _ptr(off))
        return off

    # -- phases ----------------------------------------------------------
    def build_tree(self, params):
        self._check(lib().tfmm_build_tree(self._h, ctypes.byref(params)), "tfmm_build_tree")

    def set_params(self, params):
        self._check(lib().tfmm_set_params(self._h, ctypes.byref(params)), "tfmm_set_params")

    def upward(self):
        self._check(lib().tfmm_upward(self._h), "tfmm_upward")

    def traverse(self):
        self._check(lib().tfmm_traverse(self._h), "tfmm_traverse")

    def traverse_lists(self, m2l, p2p):
        """Run the M2L and P2P kernels over given (target, source) cell pairs
        instead of the GPU traversal."""
        m = np.ascontiguousarray(m2l, np.int32).reshape(-1, 2)
        p = np.ascontiguousarray(p2p, np.int32).reshape(-1, 2)
        self._check(lib().tfmm_traverse_lists(self._h, _ptr(m), len(m), _ptr(p), len(p)),
                    "tfmm_traverse_lists")

    def downward(self):
        self._check(lib().tfmm_downward(self._h), "tfmm_downward")

    def evaluate(self, params):
        self._check(lib().tfmm_evaluate(self._h, ctypes.byref(params)), "tfmm_evaluate")

    # -- results ---------------------------------------------------------
    def fields(self, order=0):
        out = np.zeros((self.n, 4), np.float64)
        self._check(lib().tfmm_get_fields(self._h, _ptr(out), int(order), 0), "tfmm_get_fields")
        return out

    def fields_to_device(self, dev_ptr, order=0):
        self._check(lib().tfmm_get_fields(self._h, dev_ptr, int(order), 1), "tfmm_get_fields")

    def stats(self):
        s = Stats()
        self._check(lib().tfmm_get_stats(self._h, ctypes.byref(s)), "tfmm_get_stats")
        return s.as_dict()

    def sorted(self):
        k = np.zeros(self.n, np.uint64)
        p = np.zeros(self.n, np.int32)
        self._check(lib().tfmm_get_sorted(self._h, _ptr(k), _ptr(p)), "tfmm_get_sorted")
        return k, p

    def cells(self):
        nc = self.stats()["ncells"]
        out = dict(key=np.zeros(nc, np.uint64), beg=np.zeros(nc, np.int32),
                   nb=np.zeros(nc, np.int32), leaf=np.zeros(nc, np.int32),
                   center=np.zeros((nc, 3)), width=np.zeros(nc))
        self._check(lib().tfmm_get_cells(self._h, _ptr(out["key"]), _ptr(out["beg"]),
                                         _ptr(out["nb"]), _ptr(out["leaf"]),
                                         _ptr(out["center"]), _ptr(out["width"])),
                    "tfmm_get_cells")
        return out

    def expansions(self, want_l=True):
        """(M, L) of every cell: NTERM complex (spherical basis) or
        CART_TERMS real (Cartesian basis) per cell"""
        nc = self.stats()["ncells"]
        if self.basis() == BASIS_CARTESIAN:
            M = np.zeros((nc, CART_TERMS))
            Lx = np.zeros((nc, CART_TERMS)) if want_l else None
        else:
            M = np.zeros((nc, NTERM), np.complex128)
            Lx = np.zeros((nc, NTERM), np.complex128) if want_l else None
        self._check(lib().tfmm_get_expansions(self._h, _ptr(M), _ptr(Lx)),
                    "tfmm_get_expansions")
        return M, Lx

    def lists(self):
        """Executed M2L and P2P list entries, CSR over target cells (a mutual
        pair is an entry in each direction)."""
        nc = self.stats()["ncells"]
        mo = np.zeros(nc + 1, np.int32)
        po = np.zeros(nc + 1, np.int32)
        self._check(lib().tfmm_get_lists(self._h, _ptr(mo), None, _ptr(po), None),
                    "tfmm_get_lists")
        ms = np.zeros(max(1, int(mo[-1])), np.int32)
        ps = np.zeros(max(1, int(po[-1])), np.int32)
        self._check(lib().tfmm_get_lists(self._h, _ptr(mo), _ptr(ms), _ptr(po), _ptr(ps)),
                    "tfmm_get_lists")
        return (mo, ms[:int(mo[-1])]), (po, ps[:int(po[-1])])

    def tree_features(self):
        f = np.zeros(16, np.float32)
        self._check(lib().tfmm_tree_features(self._h, _ptr(f)), "tfmm_tree_features")
        return f

    def pick_params(self, picker=None, precision=FP64):
        p = make_params(64, 0.4, precision)
        self._check(lib().tfmm_pick_params(self._h, None if picker is None else picker.handle,
                                           ctypes.byref(p)), "tfmm_pick_params")
        return p

    def direct(self, targets):
        t = np.ascontiguousarray(targets, np.int64)
        out = np.zeros((len(t), 4), np.float64)
        self._check(lib().tfmm_direct(self._h, _ptr(t), len(t), _ptr(out)), "tfmm_direct")
        return out

    def direct_noalloc(self, targets):
        """tfmm_direct staged in the library's static device array (no
        device allocation; at most DIRECT_NOALLOC_MAX targets)."""
        t = np.ascontiguousarray(targets, np.int64)
        out = np.zeros((len(t), 4), np.float64)
        self._check(lib().tfmm_direct_noalloc(self._h, _ptr(t), len(t), _ptr(out)),
                    "tfmm_direct_noalloc")
        return out


class Picker:
    """The learned (ncrit, theta) picker (an FP32 MLP trained on the GPU)."""

    def __init__(self, device=0, seed=0):
        h = ctypes.c_void_p()
        rc = lib().tfmm_picker_create(int(device), int(seed), ctypes.byref(h))
        if rc != TFMM_OK:
            raise TfmmError("tfmm_picker_create failed with status %d" % rc)
        self._h = h
        self.nparam = int(lib().tfmm_picker_nparam())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().tfmm_picker_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != TFMM_OK:
            msg = lib().tfmm_picker_last_error(self._h)
            raise TfmmError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    @property
    def handle(self):
        return self._h

    def weights(self):
        w = np.zeros(self.nparam, np.float32)
        self._check(lib().tfmm_picker_get_weights(self._h, _ptr(w), self.nparam), "get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(lib().tfmm_picker_set_weights(self._h, _ptr(w), len(w)), "set_weights")

    def set_theta_max(self, theta_max):
        self._check(lib().tfmm_picker_set_theta_max(self._h, float(theta_max)), "set_theta_max")
        self.theta_max = float(theta_max)

    @classmethod
    def load(cls, path, device=0):
        """A picker with the weights (and theta cap) saved by picker_train.py"""
        d = np.load(path, allow_pickle=False)
        pk = cls(device)
        pk.set_weights(d["weights"])
        if "theta_max" in d.files:
            pk.set_theta_max(float(d["theta_max"]))
        return pk

    @staticmethod
    def _batch(X, cls, theta):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 16)
        c = None if cls is None else np.ascontiguousarray(cls, np.int32)
        t = None if theta is None else np.ascontiguousarray(theta, np.float32)
        return X, c, t

    def train(self, X, cls, theta, steps, lr=1e-2):
        X, c, t = self._batch(X, cls, theta)
        loss = np.zeros(steps, np.float32)
        self._check(lib().tfmm_picker_train(self._h, _ptr(X), _ptr(c), _ptr(t), len(X), int(steps),
                                            float(lr), _ptr(loss)), "train")
        return loss

    def gradient(self, X, cls, theta):
        X, c, t = self._batch(X, cls, theta)
        g = np.zeros(self.nparam, np.float32)
        loss = np.zeros(len(X), np.float32)
        self._check(lib().tfmm_picker_gradient(self._h, _ptr(X), _ptr(c), _ptr(t), len(X), _ptr(g),
                                               _ptr(loss)), "gradient")
        return g, loss

    def predict(self, X):
        X, _, _ = self._batch(X, None, None)
        out = np.zeros((len(X), 6), np.float32)
        self._check(lib().tfmm_picker_predict(self._h, _ptr(X), len(X), _ptr(out)), "predict")
        return out


def decode(out, theta_max=0.6):
    """(ncrit, theta) from raw picker outputs (theta capped at theta_max, as
    tfmm_pick_params does)."""
    out = np.atleast_2d(out)
    cls = np.argmax(out[:, :5], axis=1)
    theta = 0.2 + 0.4 / (1.0 + np.exp(-out[:, 5].astype(np.float64)))
    return np.array(NCRIT_CLASSES)[cls], np.minimum(theta, theta_max)


def evaluate(x, y, z, q, ncrit=64, theta=0.4, precision=FP64, device=0):
    """One full FMM evaluation of host bodies; returns (fields, stats)."""
    with Context(device) as ctx:
        ctx.set_bodies(x, y, z, q)
        ctx.evaluate(make_params(ncrit, theta, precision))
        return ctx.fields(), ctx.stats()
