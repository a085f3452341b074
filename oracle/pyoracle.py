"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It wraps oracle/_build/liboracle_{fp64,asship}.so, which
restate the reference's MGCG path on the CPU (see top3d_oracle.c for the
file:line citations and how each function is pinned).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

c_int, c_int64, c_float, c_double, c_void_p = (ctypes.c_int, ctypes.c_int64,
                                               ctypes.c_float, ctypes.c_double,
                                               ctypes.c_void_p)


class Record(ctypes.Structure):
    _fields_ = [("loop", c_int), ("c", c_double), ("vol", c_double),
                ("change", c_double), ("relres", c_double), ("cgiter", c_int)]


class CSR(ctypes.Structure):
    _fields_ = [("nrows", c_int), ("nnz", c_int64),
                ("rowOffsets", ctypes.POINTER(c_int)),
                ("colIndex", ctypes.POINTER(c_int)),
                ("vals", ctypes.POINTER(c_double))]


_LIBS = {}


def build_if_missing():
    libs = [os.path.join(_BUILD, f"liboracle_{t}.so") for t in ("fp64", "asship")]
    if not all(os.path.exists(p) for p in libs):
        import subprocess
        subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)


def load(typing="fp64"):
    """The oracle library of one typing ("fp64" or "asship"), or the build at
    the path `typing` when it names a .so file (the tests' deliberately wrong
    copies of one function)."""
    if typing in _LIBS:
        return _LIBS[typing]
    if typing.endswith(".so"):
        lib = ctypes.CDLL(typing)
    else:
        build_if_missing()
        lib = ctypes.CDLL(os.path.join(_BUILD, f"liboracle_{typing}.so"))
    P = c_void_p
    sig = {
        "or_sizeof_ctype": (c_int, []),
        "or_create": (P, [c_int, c_int, c_int, c_int, c_int]),
        "or_destroy": (None, [P]),
        "or_set_bcsym": (None, [P, c_int]),
        "or_set_support": (None, [P, c_int]),
        "or_line_load": (None, [P, P]),
        "or_line_load_profile": (None, [P, P, c_int]),
        "or_padding": (None, [P, c_int, P, P, P, P, P]),
        "or_nfixed": (c_int, [P, c_int]),
        "or_fixed": (None, [P, c_int, P]),
        "or_ke": (None, [c_double, c_int, c_int, P]),
        "or_ke_copy": (None, [P, c_int, P]),
        "or_load": (None, [P, P]),
        "or_matvec0": (None, [P, P, P, P]),
        "or_matvec_l": (None, [P, c_int, P, P, P]),
        "or_prolong": (None, [P, c_int, P, P]),
        "or_restrict": (None, [P, c_int, P, P]),
        "or_invdiag": (None, [P, P, c_int, P]),
        "or_csr_create": (ctypes.POINTER(CSR), [P, c_int]),
        "or_csr_destroy": (None, [P]),
        "or_csr_assemble": (None, [P, c_int, P, P, P]),
        "or_csr_apply": (None, [P, c_int, P, P, P]),
        "or_solver_create": (P, [P]),
        "or_solver_destroy": (None, [P, P]),
        "or_solver_setup": (c_int, [P, P, P]),
        "or_sensitivity": (None, [P, P, P, P, P]),
        "or_filter_grad": (None, [P, None, P]),
        "or_volume": (None, [P, P]),
        "or_change": (c_float, [P, P, c_int64]),
        "or_top3d": (c_int, [c_int, c_int, c_int, c_float, c_float, c_int,
                             c_float, c_int, c_int, P, P, P]),
        "or_write_vtu": (c_int, [P, P, ctypes.c_char_p, c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        if None not in args:
            f.argtypes = args
    lib._ctype = np.float64 if lib.or_sizeof_ctype() == 8 else np.float32
    _LIBS[typing] = lib
    return lib


def _p(a):
    return a.ctypes.data_as(c_void_p)


class Grid:
    """One oracle grid (setupGC) with solver buffers created on demand."""

    def __init__(self, nelx, nely, nelz, nl, yalign=1, typing="fp64", bcsym=False,
                 support="strips"):
        self.lib = load(typing)
        self.ct = self.lib._ctype
        self.dt = self.ct
        self.nelx, self.nely, self.nelz, self.nl = nelx, nely, nelz, nl
        self.h = self.lib.or_create(nelx, nely, nelz, nl, yalign)
        if not self.h:
            raise ValueError("invalid oracle grid")
        self.lib.or_set_bcsym(self.h, int(bcsym))
        if support not in ("strips", "face"):
            raise ValueError(f"unknown support {support!r}")
        self.lib.or_set_support(self.h, int(support == "face"))
        self._solver = None

    def close(self):
        if self._solver:
            self.lib.or_solver_destroy(self.h, self._solver)
            self._solver = None
        if self.h:
            self.lib.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- layout -----------------------------------------------------------
    def padding(self, l):
        vals = [c_int(), c_int(), c_int(), c_int(), c_int64()]
        self.lib.or_padding(self.h, l, *[ctypes.byref(v) for v in vals])
        return tuple(v.value for v in vals)

    def ndof(self, l=0):
        return self.padding(l)[4]

    def nelem(self):
        _, wx, wy, wz, _ = self.padding(0)
        return (wx - 1) * (wy - 1) * (wz - 1)

    def fixed(self, l):
        n = self.lib.or_nfixed(self.h, l)
        out = np.zeros(n, np.int64)
        self.lib.or_fixed(self.h, l, _p(out))
        return out

    def ke(self, l):
        n = 1 << l
        out = np.zeros(576 * n ** 3)
        self.lib.or_ke_copy(self.h, l, _p(out))
        return out.reshape(n ** 3, 24, 24)

    def load(self):
        F = np.zeros(self.ndof(0), self.ct)
        self.lib.or_load(self.h, _p(F))
        return F

    def line_load(self, profile=0):
        """0: constant, 1: half weight at the two end nodes, 2: sine."""
        F = np.zeros(self.ndof(0), self.ct)
        self.lib.or_line_load_profile(self.h, _p(F), profile)
        return F

    # ---- operators --------------------------------------------------------
    def _x(self, x):
        return np.ascontiguousarray(x, self.dt)

    def write_vtu(self, x, path, halo=False):
        """writeDensityVtkFile(WithHalo) of the padded element densities x."""
        if self.lib.or_write_vtu(self.h, _p(self._x(x)), os.fsencode(path), int(halo)) != 0:
            raise OSError(f"cannot write {path}")

    def matvec0(self, x, v):
        v = np.ascontiguousarray(v, self.ct)
        out = np.zeros_like(v)
        self.lib.or_matvec0(self.h, _p(self._x(x)), _p(v), _p(out))
        return out

    def matvec_l(self, l, x, v):
        v = np.ascontiguousarray(v, self.ct)
        out = np.zeros_like(v)
        self.lib.or_matvec_l(self.h, l, _p(self._x(x)), _p(v), _p(out))
        return out

    def prolong(self, l, coarse):
        coarse = np.ascontiguousarray(coarse, self.ct)
        out = np.zeros(self.ndof(l), self.ct)
        self.lib.or_prolong(self.h, l, _p(coarse), _p(out))
        return out

    def restrict(self, l, fine):
        fine = np.ascontiguousarray(fine, self.ct)
        out = np.zeros(self.ndof(l + 1), self.ct)
        self.lib.or_restrict(self.h, l, _p(fine), _p(out))
        return out

    def invdiag(self, x, l):
        out = np.zeros(self.ndof(l))
        self.lib.or_invdiag(self.h, _p(self._x(x)), l, _p(out))
        return out

    def csr(self, l, x):
        """Assembled level-l matrix with symmetric BC elimination (scipy)."""
        import scipy.sparse as sp
        M = self.lib.or_csr_create(self.h, l)
        tmp = np.zeros(max(self.ndof(l), M.contents.nrows))
        self.lib.or_csr_assemble(self.h, l, _p(self._x(x)), M, _p(tmp))
        n, nnz = M.contents.nrows, M.contents.nnz
        ro = np.ctypeslib.as_array(M.contents.rowOffsets, (n + 1,)).copy()
        ci = np.ctypeslib.as_array(M.contents.colIndex, (nnz,)).copy()
        va = np.ctypeslib.as_array(M.contents.vals, (nnz,)).copy()
        self.lib.or_csr_destroy(M)
        return sp.csr_matrix((va, ci, ro), shape=(n, n))

    def sensitivity(self, x, u):
        u = np.ascontiguousarray(u, np.float64)
        c = np.zeros(1, self.dt)
        dc = np.zeros(self.nelem(), self.dt)
        self.lib.or_sensitivity(self.h, _p(self._x(x)), _p(u), _p(c), _p(dc))
        return float(c[0]), dc

    def filter(self, rmin, rho):
        out = np.zeros(self.nelem(), self.dt)
        rmin_t = c_double if self.dt == np.float64 else c_float
        self.lib.or_filter.argtypes = [c_void_p, rmin_t, c_void_p, c_void_p]
        self.lib.or_filter(self.h, rmin_t(rmin), _p(self._x(rho)), _p(out))
        return out

    def filter_grad(self, rmin, v):
        v = self._x(v).copy()
        rmin_t = c_double if self.dt == np.float64 else c_float
        self.lib.or_filter_grad.argtypes = [c_void_p, rmin_t, c_void_p]
        self.lib.or_filter_grad(self.h, rmin_t(rmin), _p(v))
        return v

    def volume(self, xphys):
        rt = c_double if self.dt == np.float64 else c_float
        self.lib.or_volume.restype = rt
        return float(self.lib.or_volume(self.h, _p(self._x(xphys))))

    def change(self, x, xnew):
        """(max |x - xnew| as float, x after the call: compute_change copies
        xnew into x)."""
        x = self._x(x).copy()
        ch = self.lib.or_change(_p(x), _p(self._x(xnew)), x.size)
        return ch, x

    def oc(self, x, dv, dc, g):
        x = self._x(x)
        xnew = np.zeros_like(x)
        gt = c_double if self.dt == np.float64 else c_float
        self.lib.or_oc.argtypes = [c_void_p] * 4 + [c_int64, gt]
        self.lib.or_oc(_p(x), _p(xnew), _p(self._x(dv)), _p(self._x(dc)),
                       x.size, gt(g))
        return xnew

    # ---- solver -----------------------------------------------------------
    def solver(self):
        if not self._solver:
            self._solver = self.lib.or_solver_create(self.h)
        return self._solver

    def setup(self, x):
        return self.lib.or_solver_setup(self.h, self.solver(), _p(self._x(x)))

    def coarse_solve(self, rhs):
        l = self.nl - 1
        rhs = np.ascontiguousarray(rhs, self.ct)
        out = np.zeros(self.ndof(l), self.ct)
        self.lib.or_coarse_solve.argtypes = [c_void_p] * 2 + [c_int] + [c_void_p] * 2
        self.lib.or_coarse_solve(self.h, self.solver(), l, _p(rhs), _p(out))
        return out

    def solve(self, x, b, u, tol=1e-5, nswp=4, maxiter=1000):
        """solveStateMG_halo; u is updated in place (float64)."""
        b = np.ascontiguousarray(b, self.ct)
        it, res = c_int(), c_float()
        tol_t = c_double if self.ct == np.float64 else c_float
        self.lib.or_solve.argtypes = [c_void_p, c_void_p, c_void_p, c_int, tol_t,
                                      c_void_p, c_void_p, c_void_p, c_void_p, c_int]
        self.lib.or_solve.restype = c_int
        xx = self._x(x).copy()
        rc = self.lib.or_solve(self.h, self.solver(), _p(xx), nswp, tol_t(tol),
                               ctypes.byref(it), ctypes.byref(res), _p(b), _p(u),
                               maxiter)
        if rc:
            raise RuntimeError("oracle coarse factorization failed")
        return it.value, res.value

    def solve_times(self):
        """(set-up seconds, CG-loop seconds) of the last solve."""
        a, b = c_double(), c_double()
        self.lib.or_solver_times.argtypes = [c_void_p, c_void_p, c_void_p]
        self.lib.or_solver_times(self.solver(), ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def vcycle(self, x, r):
        """One V-cycle z = M^{-1} r at level 0 (setup must have run)."""
        s = self.solver()
        S = ctypes.cast(s, ctypes.POINTER(ctypes.c_void_p * 4))
        rptr, zptr = S.contents[0], S.contents[3]
        n = self.ndof(0)
        esz = np.dtype(self.ct).itemsize
        ctypes.memmove(rptr, _p(np.ascontiguousarray(r, self.ct)), n * esz)
        om_t = c_double if self.ct == np.float64 else c_float
        self.lib.or_vcycle.argtypes = [c_void_p, c_void_p, c_void_p, c_int, om_t, c_int]
        self.lib.or_vcycle(self.h, s, _p(self._x(x)), 0, om_t(0.6), 4)
        z = np.zeros(n, self.ct)
        ctypes.memmove(_p(z), zptr, n * esz)
        return z


def top3d(nelx, nely, nelz, nl, iters, volfrac=0.2, rmin=1.5, cgtol=1e-5,
          yalign=1, typing="fp64"):
    """The design loop (top3dmgcg); returns (records, xPhys, U)."""
    lib = load(typing)
    g = Grid(nelx, nely, nelz, nl, yalign, typing)
    nelem, ndof = g.nelem(), g.ndof(0)
    g.close()
    recs = (Record * iters)()
    xphys = np.zeros(nelem, lib._ctype)
    U = np.zeros(ndof)
    n = lib.or_top3d(nelx, nely, nelz, volfrac, rmin, nl, cgtol, iters, yalign,
                     ctypes.cast(recs, c_void_p), _p(xphys), _p(U))
    if n < 0:
        raise RuntimeError("oracle design loop failed")
    out = [dict(loop=r.loop, c=r.c, vol=r.vol, change=r.change,
                relres=r.relres, cgiter=r.cgiter) for r in recs[:n]]
    return out, xphys, U
