"""ctypes binding of the CPU oracle (liboracle.so) and runner of the compiled
reference (oracle/_ref).  TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product path.
"""
import ctypes
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
P = 10
NTERM = 55

_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u64 = ctypes.POINTER(ctypes.c_uint64)
_i32 = ctypes.POINTER(ctypes.c_int32)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_run.restype = ctypes.c_void_p
        L.orc_run.argtypes = [_d, _d, _d, _d, ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                              ctypes.c_int, ctypes.c_int]
        L.orc_run_mutual.restype = ctypes.c_void_p
        L.orc_run_mutual.argtypes = [_d, _d, _d, _d, ctypes.c_int64, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_run_periodic.restype = ctypes.c_void_p
        L.orc_run_periodic.argtypes = [_d, _d, _d, _d, ctypes.c_int64, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int]
        L.orc_run_basis.restype = ctypes.c_void_p
        L.orc_run_basis.argtypes = [_d, _d, _d, _d, ctypes.c_int64, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int]
        L.orc_run_periodic_basis.restype = ctypes.c_void_p
        L.orc_run_periodic_basis.argtypes = [_d, _d, _d, _d, ctypes.c_int64, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_cart_terms.restype = ctypes.c_int
        L.orc_cart_index.restype = ctypes.c_int
        for name in ("orc_free", "orc_region", "orc_ncells", "orc_max_depth", "orc_sorted",
                     "orc_cells", "orc_expansions", "orc_fields", "orc_counts", "orc_lists",
                     "orc_list_sizes"):
            getattr(L, name).argtypes = None
        L.orc_ncells.restype = ctypes.c_int64
        L.orc_max_depth.restype = ctypes.c_int
        L.orc_finest_key.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def bbox(x, y, z):
    L = lib()
    mn = np.zeros(3)
    mx = np.zeros(3)
    L.orc_bbox(_p(x), _p(y), _p(z), ctypes.c_int64(len(x)), _p(mn), _p(mx))
    return mn, mx


def keys(x, y, z, rmin, rmax):
    L = lib()
    k = np.zeros(len(x), np.uint64)
    L.orc_keys(_p(x), _p(y), _p(z), ctypes.c_int64(len(x)), _p(np.ascontiguousarray(rmin, float)),
               _p(np.ascontiguousarray(rmax, float)), _p(k))
    return k


def finest_key(anchor):
    a = (ctypes.c_int * 3)(*[int(v) for v in anchor])
    return int(lib().orc_finest_key(a))


def cell_region(key, rmin, rmax):
    mn = np.zeros(3)
    mx = np.zeros(3)
    lib().orc_cell_region(ctypes.c_uint64(key), _p(np.asarray(rmin, float)),
                          _p(np.asarray(rmax, float)), _p(mn), _p(mx))
    return mn, mx


def eval_multipole(rho, alpha, beta):
    y = np.zeros(2 * P * P)
    t = np.zeros(2 * P * P)
    lib().orc_eval_multipole(ctypes.c_double(rho), ctypes.c_double(alpha), ctypes.c_double(beta),
                             _p(y), _p(t))
    return y.view(np.complex128), t.view(np.complex128)


def eval_local(rho, alpha, beta):
    y = np.zeros(2 * P * P)
    lib().orc_eval_local(ctypes.c_double(rho), ctypes.c_double(alpha), ctypes.c_double(beta), _p(y))
    return y.view(np.complex128)


def p2m(center, x, y, z, q):
    M = np.zeros(NTERM, np.complex128)
    lib().orc_p2m(_p(M), _p(np.asarray(center, float)), _p(x), _p(y), _p(z), _p(q),
                  ctypes.c_int64(len(x)))
    return M


def m2m(Mc, dX, Mp=None):
    Mp = np.zeros(NTERM, np.complex128) if Mp is None else Mp.copy()
    lib().orc_m2m(_p(Mp), _p(np.ascontiguousarray(Mc, np.complex128)), _p(np.asarray(dX, float)))
    return Mp


def m2l(Mj, dX, Li=None):
    Li = np.zeros(NTERM, np.complex128) if Li is None else Li.copy()
    lib().orc_m2l(_p(Li), _p(np.ascontiguousarray(Mj, np.complex128)), _p(np.asarray(dX, float)))
    return Li


def l2l(Lp, dX, Lc=None):
    Lc = np.zeros(NTERM, np.complex128) if Lc is None else Lc.copy()
    lib().orc_l2l(_p(Lc), _p(np.ascontiguousarray(Lp, np.complex128)), _p(np.asarray(dX, float)))
    return Lc


def l2p(trg, bx, q, L, center):
    t = np.array(trg, float).copy()
    lib().orc_l2p(_p(t), _p(np.asarray(bx, float)), ctypes.c_double(q),
                  _p(np.ascontiguousarray(L, np.complex128)), _p(np.asarray(center, float)))
    return t


# ---------------------------------------------------------------------------
# Cartesian Taylor basis (CN real terms per expansion)
# ---------------------------------------------------------------------------
SPHERICAL, CARTESIAN = 0, 1
CN = P * (P + 1) * (P + 2) // 6


def cart_terms():
    """multi-index (x, y, z) of every Cartesian term, in index order"""
    out = []
    for d in range(P):
        for x in range(d, -1, -1):
            for y in range(d - x, -1, -1):
                out.append((x, y, d - x - y))
    return out


def cart_index(x, y, z):
    return int(lib().orc_cart_index(int(x), int(y), int(z)))


def cart_derivs(dX):
    C = np.zeros(CN)
    lib().orc_cart_derivs(_p(C), _p(np.asarray(dX, float)))
    return C


def cart_powers(dX, c0=1.0):
    C = np.zeros(CN)
    lib().orc_cart_powers(_p(C), _p(np.asarray(dX, float)), ctypes.c_double(c0))
    return C


def cart_p2m(center, x, y, z, q):
    M = np.zeros(CN)
    x, y, z, q = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, q)]
    lib().orc_cart_p2m(_p(M), _p(np.asarray(center, float)), _p(x), _p(y), _p(z), _p(q),
                       ctypes.c_int64(len(x)))
    return M


def cart_m2m(Mc, dX, Mp=None):
    Mp = np.zeros(CN) if Mp is None else np.array(Mp, float)
    lib().orc_cart_m2m(_p(Mp), _p(np.ascontiguousarray(Mc, float)), _p(np.asarray(dX, float)))
    return Mp


def cart_m2l(Mj, dX, Li=None):
    Li = np.zeros(CN) if Li is None else np.array(Li, float)
    lib().orc_cart_m2l(_p(Li), _p(np.ascontiguousarray(Mj, float)), _p(np.asarray(dX, float)))
    return Li


def cart_m2l_mutual(Mi, Mj, dX, Li=None, Lj=None):
    Li = np.zeros(CN) if Li is None else np.array(Li, float)
    Lj = np.zeros(CN) if Lj is None else np.array(Lj, float)
    lib().orc_cart_m2l_mutual(_p(Li), _p(Lj), _p(np.ascontiguousarray(Mi, float)),
                              _p(np.ascontiguousarray(Mj, float)), _p(np.asarray(dX, float)))
    return Li, Lj


def cart_l2l(Lp, dX, Lc=None):
    Lc = np.zeros(CN) if Lc is None else np.array(Lc, float)
    lib().orc_cart_l2l(_p(Lc), _p(np.ascontiguousarray(Lp, float)), _p(np.asarray(dX, float)))
    return Lc


def cart_l2p(trg, bx, q, L, center):
    t = np.array(trg, float).copy()
    lib().orc_cart_l2p(_p(t), _p(np.asarray(bx, float)), ctypes.c_double(q),
                       _p(np.ascontiguousarray(L, float)), _p(np.asarray(center, float)))
    return t


class Run:
    """One oracle evaluation.  phases: 0 tree, 1 +upward, 2 +traversal,
    3 +downward (full FMM); mutual as the reference's --mutual; images > 0:
    periodic images of period cycle (mutual 0); basis SPHERICAL or CARTESIAN."""

    def __init__(self, x, y, z, q, ncrit=64, theta=0.4, phases=3, record_lists=False, mutual=0,
                 images=0, cycle=0.0, basis=SPHERICAL):
        L = lib()
        self.x, self.y, self.z, self.q = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, q)]
        self.n = len(self.x)
        self.basis = int(basis)
        if images > 0:
            h = L.orc_run_periodic_basis(self.x.ctypes.data_as(_d), self.y.ctypes.data_as(_d),
                                         self.z.ctypes.data_as(_d), self.q.ctypes.data_as(_d),
                                         self.n, int(ncrit), float(theta), int(images),
                                         float(cycle), int(phases), int(bool(record_lists)),
                                         self.basis)
        else:
            h = L.orc_run_basis(self.x.ctypes.data_as(_d), self.y.ctypes.data_as(_d),
                                self.z.ctypes.data_as(_d), self.q.ctypes.data_as(_d), self.n,
                                int(ncrit), float(theta), int(phases), int(bool(record_lists)),
                                int(bool(mutual)), self.basis)
        if not h:
            raise ValueError("oracle rejected the input")
        self._h = ctypes.c_void_p(h)
        self.record_lists = record_lists

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.orc_free(self._h)
            self._h = None

    def region(self):
        mn = np.zeros(3)
        mx = np.zeros(3)
        lib().orc_region(self._h, _p(mn), _p(mx))
        return mn, mx

    @property
    def ncells(self):
        return int(lib().orc_ncells(self._h))

    @property
    def max_depth(self):
        return int(lib().orc_max_depth(self._h))

    def sorted(self):
        k = np.zeros(self.n, np.uint64)
        p = np.zeros(self.n, np.int64)
        lib().orc_sorted(self._h, _p(k), _p(p))
        return k, p

    def cells(self):
        nc = self.ncells
        out = dict(key=np.zeros(nc, np.uint64), nb=np.zeros(nc, np.int64),
                   bid=np.zeros(nc, np.int64), leaf=np.zeros(nc, np.int32),
                   center=np.zeros((nc, 3)), width=np.zeros((nc, 3)))
        lib().orc_cells(self._h, _p(out["key"]), _p(out["nb"]), _p(out["bid"]), _p(out["leaf"]),
                        _p(out["center"]), _p(out["width"]))
        return out

    def expansions(self):
        """(M, L): NTERM complex per cell (spherical) or CN real (Cartesian)"""
        nc = self.ncells
        if self.basis == CARTESIAN:
            M = np.zeros((nc, CN))
            Lx = np.zeros((nc, CN))
        else:
            M = np.zeros((nc, NTERM), np.complex128)
            Lx = np.zeros((nc, NTERM), np.complex128)
        lib().orc_expansions(self._h, _p(M), _p(Lx))
        return M, Lx

    def fields(self):
        t = np.zeros((self.n, 4))
        lib().orc_fields(self._h, _p(t))
        return t

    def counts(self):
        c = np.zeros(4, np.uint64)
        lib().orc_counts(self._h, _p(c))
        return dict(m2l=int(c[0]), p2p_leafpairs=int(c[1]), p2p_bodypairs=int(c[2]),
                    p2p_selfpairs=int(c[3]))

    def lists(self):
        """Recorded (target, source) cell pairs; a mutual call gives two."""
        sz = np.zeros(2, np.int64)
        lib().orc_list_sizes(self._h, _p(sz))
        m = np.zeros((int(sz[0]), 2), np.int64)
        p = np.zeros((int(sz[1]), 2), np.int64)
        lib().orc_lists(self._h, _p(m), _p(p))
        return m, p


def direct(x, y, z, q, targets):
    """FP64 direct sum on the target subset (pot, ax, ay, az), normalised by q."""
    x, y, z, q = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, q)]
    t = np.ascontiguousarray(targets, np.int64)
    out = np.zeros((len(t), 4))
    lib().orc_direct(_p(x), _p(y), _p(z), _p(q), ctypes.c_int64(len(x)), _p(t),
                     ctypes.c_int64(len(t)), _p(out))
    return out


def direct_periodic(x, y, z, q, targets, images, cycle):
    """direct sum over the (3^images)^3 copies of the bodies (period cycle)"""
    x, y, z, q = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, q)]
    t = np.ascontiguousarray(targets, np.int64)
    out = np.zeros((len(t), 4))
    lib().orc_direct_periodic(_p(x), _p(y), _p(z), _p(q), ctypes.c_int64(len(x)), _p(t),
                              ctypes.c_int64(len(t)), ctypes.c_int(images),
                              ctypes.c_double(cycle), _p(out))
    return out


def rel_l2(trg, ref):
    """(potential, acceleration) relative L2 errors as in verify.h."""
    trg = np.ascontiguousarray(trg, np.float64)
    ref = np.ascontiguousarray(ref, np.float64)
    e = np.zeros(2)
    lib().orc_rel_l2(_p(trg), _p(ref), ctypes.c_int64(len(trg)), _p(e))
    return float(e[0]), float(e[1])


# ---------------------------------------------------------------------------
# compiled reference (oracle/_ref), available where it has been built
# ---------------------------------------------------------------------------

def ref_binary(count=True):
    p = os.path.join(REF_DIR, "tapas_ref_count" if count else "tapas_ref")
    return p if os.path.exists(p) else None


def write_input(path, x, y, z, q, ncrit, theta):
    with open(path, "wb") as f:
        f.write(struct.pack("<qid", len(x), int(ncrit), float(theta)))
        for a in (x, y, z, q):
            f.write(np.ascontiguousarray(a, "<f8").tobytes())


def read_ref_output(path):
    with open(path, "rb") as f:
        buf = f.read()
    off = 0

    def take(dtype, count):
        nonlocal off
        dt = np.dtype(dtype)
        a = np.frombuffer(buf, dt, count, off)
        off += dt.itemsize * count
        return a

    n = int(take("<i8", 1)[0])
    rmin = take("<f8", 3).copy()
    rmax = take("<f8", 3).copy()
    max_depth = int(take("<i4", 1)[0])
    keys = take("<u8", n).copy()
    ibody = take("<i8", n).copy()
    nc = int(take("<i8", 1)[0])
    rec = np.dtype([("key", "<u8"), ("nb", "<i8"), ("bid", "<i8"), ("leaf", "<i4"),
                    ("center", "<f8", 3), ("width", "<f8", 3)])
    cells = take(rec, nc).copy()
    trg = take("<f8", 4 * n).reshape(n, 4).copy()
    m2l, p2p = [int(v) for v in take("<u8", 2)]
    times = take("<f8", 4).copy()
    has_exp = int(take("<i4", 1)[0])
    M = Lx = None
    if has_exp:
        e = take("<f8", nc * 4 * NTERM).reshape(nc, 2, NTERM, 2)
        M = (e[:, 0, :, 0] + 1j * e[:, 0, :, 1]).copy()
        Lx = (e[:, 1, :, 0] + 1j * e[:, 1, :, 1]).copy()
    return dict(n=n, rmin=rmin, rmax=rmax, max_depth=max_depth, keys=keys, ibody=ibody,
                cells=cells, trg=trg, m2l=m2l, p2p_calls=p2p,
                times=dict(tree=times[0], upward=times[1], traverse=times[2], downward=times[3]),
                M=M, L=Lx)


MPIEXEC = "/opt/conda/bin/mpiexec"


def run_ref_ranks(x, y, z, q, ncrit, theta, ranks, timeout=None):
    """Time the compiled reference with `ranks` MPI ranks (each starting from
    a slice of the bodies); returns the four phase times (maxima over the
    ranks), or None when no MPI launcher is available."""
    exe = ref_binary(count=False)
    if exe is None or not os.path.exists(MPIEXEC):
        return None
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        write_input(fin, x, y, z, q, ncrit, theta)
        env = dict(os.environ)
        env["OMP_NUM_THREADS"] = "1"
        subprocess.run([MPIEXEC, "-n", str(int(ranks)), exe, fin, fout, "0"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=env, timeout=timeout)
        with open(fout, "rb") as f:
            buf = f.read()
        n, p = struct.unpack_from("<qi", buf, 0)
        t = struct.unpack_from("<4d", buf, 12)
        return {"n": n, "ranks": p, "times": dict(zip(("tree", "upward", "traverse", "downward"), t))}


def run_ref(x, y, z, q, ncrit=64, theta=0.4, count=True, dump_expansions=False, timeout=None,
            mutual=0):
    """Run the compiled reference on the bodies (with the given mutual flag);
    returns its dump as a dict."""
    exe = ref_binary(count)
    if exe is None:
        raise FileNotFoundError("oracle/_ref not built")
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        write_input(fin, x, y, z, q, ncrit, theta)
        env = dict(os.environ)
        env.setdefault("OMP_NUM_THREADS", "1")
        subprocess.run([exe, fin, fout, "1" if dump_expansions else "0", str(int(mutual))],
                       check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=env, timeout=timeout)
        return read_ref_output(fout)


def run_ref_cartesian(seed=12345):
    """Run the reference's Cartesian-basis kernels on the fixed cells and
    bodies of oracle/ref_cartesian.cxx; returns inputs and outputs as arrays."""
    exe = os.path.join(REF_DIR, "ref_cartesian")
    if not os.path.exists(exe):
        raise FileNotFoundError("oracle/_ref/ref_cartesian not built")
    with tempfile.TemporaryDirectory() as td:
        fout = os.path.join(td, "out.bin")
        subprocess.run([exe, fout, str(int(seed))], check=True, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE)
        a = np.fromfile(fout, np.float64)
    nc, nb, nt = int(a[0]), int(a[1]), int(a[2])
    p = 3
    w = 3 + 2 * nt + 5
    cells = a[p:p + nc * w].reshape(nc, w)
    p += nc * w
    bod = a[p:p + nb * 12].reshape(nb, 12)
    p += nb * 12
    xp = a[p:p + 6].reshape(2, 3)
    return dict(center=cells[:, :3].copy(), M=cells[:, 3:3 + nt].copy(),
                L=cells[:, 3 + nt:3 + 2 * nt].copy(),
                cell_ibody=cells[:, -5].astype(np.int64), cell_nbody=cells[:, -4].astype(np.int64),
                cell_ichild=cells[:, -3].astype(np.int64),
                cell_nchild=cells[:, -2].astype(np.int64),
                cell_iparent=cells[:, -1].astype(np.int64),
                x=bod[:, :3].copy(), q=bod[:, 3].copy(), trg_in=bod[:, 4:8].copy(),
                trg_out=bod[:, 8:12].copy(), xperiodic=xp.copy())
