"""Python binding of libtop3d_hip, the gfx950 MGCG state-solve library.

Host-side mirror of the reference interface (topopt/TopOpt-in-OpenMP-GPU):
`Grid` wraps the reference's gridContext / gpuGrid / SolverDataBuffer
(definitions.h:36-53, gpu_definitions.h:7-37) and exposes its entry points
under their own names (setupGC, solveStateMG_halo, applyStateOperator_stencil,
getComplianceAndSensetivity_halo, applyDensityFilter(Gradient), ...) on torch
tensors that live on the GPU.  `design_loop` is top3dmgcg
(stencil_optimization.c:13-223) driven step by step through the C ABI, and
`Compliance` is the torch.autograd.Function of c(xPhys).

The native library is required: importing this module raises if
lib/libtop3d_hip.so has not been built, and no CPU fallback exists.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libtop3d_hip.so")

c_int, c_double, c_float, c_void_p, c_u64 = (ctypes.c_int, ctypes.c_double, ctypes.c_float,
                                             ctypes.c_void_p, ctypes.c_uint64)

TOP3D_FP64, TOP3D_MIXED = 0, 1
TOP3D_BC_REFERENCE, TOP3D_BC_SYMMETRIC = 0, 1
TOP3D_SUPPORT_STRIPS, TOP3D_SUPPORT_FACE = 0, 1
SUPPORTS = {"strips": TOP3D_SUPPORT_STRIPS, "face": TOP3D_SUPPORT_FACE}
ERRORS = {0: "ok", 1: "hip", 2: "argument", 3: "unmapped", 4: "not SPD", 5: "no convergence",
          6: "non-finite"}


# ---- ABI structs (byte layout of the reference's structs) -------------------
class FixedDofs(ctypes.Structure):
    _fields_ = [("n", c_u64), ("idx", ctypes.POINTER(c_u64))]


class GridContext(ctypes.Structure):
    _fields_ = [("E0", c_double), ("Emin", c_double), ("nu", c_double),
                ("elementSizeX", c_double), ("elementSizeY", c_double),
                ("elementSizeZ", c_double), ("nelx", c_u64), ("nely", c_u64),
                ("nelz", c_u64), ("wrapx", c_u64), ("wrapy", c_u64), ("wrapz", c_u64),
                ("penal", c_double), ("precomputedKE", ctypes.POINTER(ctypes.POINTER(c_double))),
                ("fixedDofs", ctypes.POINTER(FixedDofs))]


class CSRMatrix(ctypes.Structure):
    _fields_ = [("nnz", c_u64), ("nrows", ctypes.c_int32), ("rowOffsets", ctypes.POINTER(c_int)),
                ("colIndex", ctypes.POINTER(c_int)), ("vals", ctypes.POINTER(c_double))]


class CoarseSolverData(ctypes.Structure):
    _fields_ = [(n, c_void_p) for n in ("cholmodCommon", "sparseMatrix", "factoredMatrix", "rhs",
                                        "solution", "Y_workspace", "E_workspace")]


class SolverDataBuffer(ctypes.Structure):
    _fields_ = [("r", c_void_p), ("p", c_void_p), ("q", c_void_p), ("z", c_void_p),
                ("invD", ctypes.POINTER(c_void_p)), ("dmg", ctypes.POINTER(c_void_p)),
                ("rmg", ctypes.POINTER(c_void_p)), ("zmg", ctypes.POINTER(c_void_p)),
                ("coarseMatrices", ctypes.POINTER(CSRMatrix)),
                ("bottomSolver", CoarseSolverData)]


class GpuNode(ctypes.Structure):
    pass


GpuNode._fields_ = [("id", c_int), ("prev", ctypes.POINTER(GpuNode)),
                    ("next", ctypes.POINTER(GpuNode)), ("gc", ctypes.POINTER(GridContext)),
                    ("offset_x", c_int), ("x_global", c_int)]


class GpuGrid(ctypes.Structure):
    _fields_ = [("num_targets", c_int), ("target", ctypes.POINTER(GpuNode))]


class MemoryInfo(ctypes.Structure):
    _fields_ = [("library_bytes", c_u64), ("library_peak_bytes", c_u64),
                ("device_used_bytes", c_u64), ("device_total_bytes", c_u64)]


def memory_info(reset_peak=False):
    """Device memory of the library (now and peak) and of the whole device
    (top3d_memory_info), in bytes."""
    m = MemoryInfo()
    if library().top3d_memory_info(ctypes.byref(m), int(bool(reset_peak))):
        check("top3d_memory_info")
    return {f: getattr(m, f) for f, _ in MemoryInfo._fields_}


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("relres", c_double), ("setup_ms", c_double),
                ("solve_ms", c_double), ("status", c_int)]


# every entry point of include/top3d_hip.h: name -> (restype, argtypes)
P = c_void_p
SIGNATURES = {
    "setupGC": (None, [P, c_int, c_int, c_int, c_int]),
    "freeGC": (None, [P, c_int]),
    "setFixedDof_halo": (None, [P, c_int]),
    "allocateZeroPaddedStateField": (None, [GridContext, c_int, P]),
    "allocateZeroPaddedStateField_MTYPE": (None, [GridContext, c_int, P]),
    "allocateZeroPaddedStateField_STYPE": (None, [GridContext, c_int, P]),
    "computePadding": (None, [P, c_int, P, P, P, P, P]),
    "gridContextToGPUGrid": (None, [P, P, c_int, c_int]),
    "freeGPUGrid": (None, [P, c_int]),
    "getKEsubspace": (None, [P, c_double, c_int]),
    "applyStateOperator_stencil": (None, [P, P, P, P]),
    "applyStateOperatorSubspace_halo": (None, [P, c_int, P, P, P]),
    "getComplianceAndSensetivity_halo": (None, [P, P, P, P, P]),
    "projectToFinerGrid_halo": (None, [P, c_int, P, P]),
    "projectToCoarserGrid_halo": (None, [P, c_int, P, P]),
    "assembleInvertedMatrixDiagonalSubspace_halo": (None, [P, P, c_int, P]),
    "smoothDampedJacobi_halo": (None, [P, P, c_u64, c_double, P, P, P, P]),
    "smoothDampedJacobiSubspace_halo": (None, [P, P, c_int, c_u64, c_double, P, P, P, P]),
    "smoothDampedJacobiSubspaceMatrix_halo": (None, [P, CSRMatrix, c_int, c_u64, c_double,
                                                     P, P, P, P]),
    "solveStateMG_halo": (None, [P, P, c_int, c_int, c_double, P, P, P, P, P]),
    "allocateSolverData": (None, [GridContext, c_int, P]),
    "freeSolverData": (None, [P, c_int]),
    "allocateSubspaceMatrix": (None, [GridContext, c_int, P]),
    "freeSubspaceMatrix": (None, [P]),
    "assembleSubspaceMatrix": (None, [GridContext, c_int, P, CSRMatrix, P]),
    "applyStateOperatorSubspaceMatrix": (None, [GridContext, c_int, CSRMatrix, P, P]),
    "initializeCholmod": (None, [GridContext, c_int, P, CSRMatrix]),
    "finishCholmod": (None, [GridContext, c_int, P, CSRMatrix, c_int]),
    "factorizeSubspaceMatrix": (None, [GridContext, c_int, CoarseSolverData, CSRMatrix]),
    "solveSubspaceMatrix": (None, [GridContext, c_int, CoarseSolverData, P, P]),
    "compute_volume": (c_double, [GridContext, P]),
    "compute_change": (c_float, [P, P, c_int]),
    "update_solution": (None, [P, P, P, P, c_int, c_double]),
    "update_solution_noalloc": (None, [P, P, P, P, c_int, c_double]),
    "enter_data": (None, [GridContext, SolverDataBuffer, GpuGrid, P, P, P, P, P, P, c_int,
                          c_int]),
    "applyDensityFilter": (None, [GridContext, c_double, P, P]),
    "applyDensityFilterGradient": (None, [GridContext, c_double, P]),
    "top3dmgcg": (None, [c_u64, c_u64, c_u64, c_double, c_double, c_u64, c_float, c_u64,
                         c_int, c_int, c_int]),
    "writeDensityVtkFile": (None, [c_int, c_int, c_int, P, ctypes.c_char_p]),
    "writeDensityVtkFileWithHalo": (None, [c_int, c_int, c_int, P, ctypes.c_char_p]),
    "top3d_last_error": (c_int, []),
    "top3d_last_error_string": (ctypes.c_char_p, []),
    "top3d_clear_error": (None, []),
    "top3d_set_precision": (None, [P, c_int]),
    "top3d_set_bc_mode": (None, [P, c_int]),
    "top3d_set_max_iter": (None, [P, c_int]),
    "top3d_set_default_precision": (None, [c_int]),
    "top3d_set_default_bc_mode": (None, [c_int]),
    "top3d_cantilever_load": (None, [P, P]),
    "top3d_line_load": (None, [P, P]),
    "top3d_line_load_profile": (None, [P, P, c_int]),
    "top3d_set_default_support": (None, [c_int]),
    "top3d_get_default_support": (c_int, []),
    "top3d_map": (c_int, [P, c_u64]),
    "top3d_map_alloc": (c_int, [P, c_u64]),
    "top3d_unmap": (c_int, [P]),
    "top3d_update_to": (c_int, [P, c_u64]),
    "top3d_update_from": (c_int, [P, c_u64]),
    "top3d_device_ptr": (P, [P]),
    "top3d_set_stream": (None, [P]),
    "top3d_get_stream": (P, []),
    "top3d_synchronize": (c_int, []),
    "top3d_solver_stats": (None, [P, P]),
    "top3d_compliance_forward": (c_int, [P, P, c_int, c_double, P, P, P, P]),
    "top3d_compliance_backward": (c_int, [P, c_double, P, P, c_double, P]),
    "top3d_dense_spd_inverse": (c_int, [P, c_int, c_int]),
    "top3d_time_kernel": (c_int, [P, c_int, c_int, P, P]),
    "top3d_memory_info": (c_int, [P, c_int]),
    "top3d_reference_padding": (None, [P, c_int, P, P, P, P]),
    "top3d_nodes_from_reference_layout": (None, [P, c_int, P, P]),
    "top3d_nodes_to_reference_layout": (None, [P, c_int, P, P]),
    "top3d_elements_from_reference_layout": (None, [P, P, P]),
    "top3d_elements_to_reference_layout": (None, [P, P, P]),
    "top3d_sweep_timing": (c_int, [P, c_int, P, P, P]),
    "applyStateOperator_stencil_f32": (None, [P, P, P, P]),
    "applyStateOperatorSubspace_halo_f32": (None, [P, c_int, P, P, P]),
}

_lib = None


def library(path=LIB_PATH):
    """Load libtop3d_hip.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libtop3d_hip.so not built ({path}); run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class Top3dError(RuntimeError):
    pass


def check(what=""):
    lib = library()
    code = lib.top3d_last_error()
    if code != 0:
        msg = lib.top3d_last_error_string().decode()
        lib.top3d_clear_error()
        raise Top3dError(f"{what}: {ERRORS.get(code, code)}: {msg}")


def _ptr(t):
    return c_void_p(t.data_ptr())


class Grid:
    """A fine grid with nl multigrid levels, its GPU mirror and solver buffers.

    Arrays are torch tensors on cuda:0 in the reference's padded halo layout
    (node (i,j,k) at i*wrapy*wrapz + k*wrapy + j, DOF 3*node + comp; element
    (i,j,k) at i*(wrapy-1)*(wrapz-1) + k*(wrapy-1) + j).
    """

    def __init__(self, nelx, nely, nelz, nl, precision="fp64", bc="reference", device="cuda:0",
                 support="strips"):
        import torch
        self.torch = torch
        self.lib = library()
        self.device = torch.device(device)
        self.lib.top3d_set_stream(c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        self.nelx, self.nely, self.nelz, self.nl = nelx, nely, nelz, nl
        self.gc = GridContext()
        if support not in SUPPORTS:
            raise ValueError(f"unknown support {support!r}")
        self.support = support
        prev = self.lib.top3d_get_default_support()
        self.lib.top3d_set_default_support(SUPPORTS[support])
        try:
            self.lib.setupGC(ctypes.byref(self.gc), nl, nelx, nely, nelz)
        finally:
            self.lib.top3d_set_default_support(prev)
        check("setupGC")
        self.gpu = GpuGrid()
        self.gpu.num_targets = 1
        self.lib.gridContextToGPUGrid(ctypes.byref(self.gc), ctypes.byref(self.gpu), nl, 0)
        check("gridContextToGPUGrid")
        self.node = ctypes.byref(self.gpu.target[0])
        self.data = SolverDataBuffer()
        self.lib.allocateSolverData(self.gc, nl, ctypes.byref(self.data))
        check("allocateSolverData")
        self.set_precision(precision)
        self.set_bc(bc)
        self._closed = False

    # ---- configuration ----------------------------------------------------
    def set_precision(self, precision):
        self.precision = precision
        self.lib.top3d_set_precision(ctypes.byref(self.data),
                                     TOP3D_MIXED if precision == "mixed" else TOP3D_FP64)

    def set_bc(self, bc):
        self.bc = bc
        self.lib.top3d_set_bc_mode(ctypes.byref(self.data),
                                   TOP3D_BC_SYMMETRIC if bc == "symmetric" else TOP3D_BC_REFERENCE)

    def set_max_iter(self, n):
        self.lib.top3d_set_max_iter(ctypes.byref(self.data), n)

    def close(self):
        if getattr(self, "_closed", True):
            return
        self.lib.top3d_synchronize()
        self.lib.freeSolverData(ctypes.byref(self.data), self.nl)
        self.lib.freeGPUGrid(ctypes.byref(self.gpu), self.nl)
        self.lib.freeGC(ctypes.byref(self.gc), self.nl)
        self._closed = True

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- layout -----------------------------------------------------------
    def padding(self, l):
        v = [c_int(), c_int(), c_int(), c_int(), c_u64()]
        self.lib.computePadding(ctypes.byref(self.gc), l, *[ctypes.byref(q) for q in v])
        return tuple(q.value for q in v)

    def ndof(self, l=0):
        return self.padding(l)[4]

    @property
    def nelem(self):
        return (self.gc.wrapx - 1) * (self.gc.wrapy - 1) * (self.gc.wrapz - 1)

    def fixed_dofs(self, l):
        fd = self.gc.fixedDofs[l]
        return np.ctypeslib.as_array(fd.idx, (fd.n,)).astype(np.int64).copy()

    def _interior_index(self):
        if getattr(self, "_iidx", None) is None:
            self._iidx = self.torch.from_numpy(self.interior_elements()).to(self.device)
        return self._iidx

    def interior_elements(self):
        ey, ez = self.gc.wrapy - 1, self.gc.wrapz - 1
        i, k, j = np.meshgrid(np.arange(1, self.nelx + 1), np.arange(1, self.nelz + 1),
                              np.arange(1, self.nely + 1), indexing="ij")
        return (i * ey * ez + k * ey + j).ravel()

    def zeros(self, l=0, dtype=None):
        return self.torch.zeros(self.ndof(l), dtype=dtype or self.torch.float64, device=self.device)

    def density(self, values=None, fill=0.0):
        """Padded element tensor; `values` (i, k, j order) fill the interior."""
        x = np.zeros(self.nelem)
        idx = self.interior_elements()
        x[idx] = fill if values is None else np.asarray(values, np.float64)
        return self.torch.from_numpy(x).to(self.device)

    def load(self):
        F = np.zeros(self.ndof(0))
        self.lib.top3d_cantilever_load(ctypes.byref(self.gc), F.ctypes.data_as(c_void_p))
        return self.torch.from_numpy(F).to(self.device)

    def line_load(self, profile=0):
        """Edge load of the MATLAB specification; profile 0 constant, 1 half
        weight at the two end nodes, 2 sine."""
        F = np.zeros(self.ndof(0))
        self.lib.top3d_line_load_profile(ctypes.byref(self.gc), F.ctypes.data_as(c_void_p),
                                         profile)
        return self.torch.from_numpy(F).to(self.device)

    def ke(self, l):
        n = 1 << l
        out = np.zeros(576 * n ** 3)
        self.lib.getKEsubspace(out.ctypes.data_as(c_void_p), 0.3, l)
        return out.reshape(n ** 3, 24, 24)

    # ---- operators (reference entry points) -------------------------------
    def _sync_torch(self):
        self.torch.cuda.current_stream(self.device).synchronize()

    def apply(self, x, v, out=None):
        out = self.zeros(0) if out is None else out
        self.lib.applyStateOperator_stencil(self.node, _ptr(x), _ptr(v), _ptr(out))
        check("applyStateOperator_stencil")
        return out

    def apply_f32(self, x, v):
        out = self.zeros(0, self.torch.float32)
        self.lib.applyStateOperator_stencil_f32(self.node, _ptr(x), _ptr(v), _ptr(out))
        check("applyStateOperator_stencil_f32")
        return out

    def apply_level(self, l, x, v):
        out = self.zeros(l)
        self.lib.applyStateOperatorSubspace_halo(self.node, l, _ptr(x), _ptr(v), _ptr(out))
        check("applyStateOperatorSubspace_halo")
        return out

    def apply_level_f32(self, l, x, v):
        out = self.zeros(l, self.torch.float32)
        self.lib.applyStateOperatorSubspace_halo_f32(self.node, l, _ptr(x), _ptr(v), _ptr(out))
        check("applyStateOperatorSubspace_halo_f32")
        return out

    def prolong(self, l, coarse):
        out = self.zeros(l)
        self.lib.projectToFinerGrid_halo(self.node, l, _ptr(coarse), _ptr(out))
        check("projectToFinerGrid_halo")
        return out

    def restrict(self, l, fine):
        out = self.zeros(l + 1)
        self.lib.projectToCoarserGrid_halo(self.node, l, _ptr(fine), _ptr(out))
        check("projectToCoarserGrid_halo")
        return out

    def invdiag(self, x, l):
        out = self.zeros(l)
        self.lib.assembleInvertedMatrixDiagonalSubspace_halo(self.node, _ptr(x), l, _ptr(out))
        check("assembleInvertedMatrixDiagonalSubspace_halo")
        return out

    def jacobi(self, x, nswp, invD, u, b, l=0, omega=0.6):
        tmp = self.zeros(l)
        if l == 0:
            self.lib.smoothDampedJacobi_halo(self.node, _ptr(x), nswp, omega, _ptr(invD),
                                             _ptr(u), _ptr(b), _ptr(tmp))
        else:
            self.lib.smoothDampedJacobiSubspace_halo(self.node, _ptr(x), l, nswp, omega,
                                                     _ptr(invD), _ptr(u), _ptr(b), _ptr(tmp))
        check("smoothDampedJacobi")
        return u

    def sensitivity(self, x, u):
        c = c_double()
        dc = self.torch.zeros(self.nelem, dtype=self.torch.float64, device=self.device)
        self.lib.getComplianceAndSensetivity_halo(self.node, _ptr(x), _ptr(u), ctypes.byref(c),
                                                  _ptr(dc))
        check("getComplianceAndSensetivity_halo")
        return c.value, dc

    def filter(self, rmin, rho):
        out = self.torch.zeros_like(rho)
        self.lib.applyDensityFilter(self.gc, rmin, _ptr(rho), _ptr(out))
        check("applyDensityFilter")
        return out

    def filter_grad(self, rmin, v):
        v = v.clone()
        self.lib.applyDensityFilterGradient(self.gc, rmin, _ptr(v))
        check("applyDensityFilterGradient")
        return v

    def volume(self, xphys):
        v = self.lib.compute_volume(self.gc, _ptr(xphys))
        check("compute_volume")
        return v

    def change(self, x, xnew):
        ch = self.lib.compute_change(_ptr(x), _ptr(xnew), x.numel())
        check("compute_change")
        return ch

    def oc(self, x, xnew, dv, dc, g):
        self.lib.update_solution(_ptr(x), _ptr(xnew), _ptr(dv), _ptr(dc), x.numel(), g)
        check("update_solution")

    def solve(self, x, b, u, tol=1e-5, nswp=4, allow_no_converge=False):
        """solveStateMG_halo; u (float64 tensor) is updated in place."""
        it, res = c_int(), c_float()
        self._sync_torch()
        self.lib.top3d_clear_error()
        self.lib.solveStateMG_halo(ctypes.byref(self.gpu), _ptr(x), nswp, self.nl, tol,
                                   ctypes.byref(self.data), ctypes.byref(it), ctypes.byref(res),
                                   _ptr(b), _ptr(u))
        code = self.lib.top3d_last_error()
        if code == 5 and allow_no_converge:
            self.lib.top3d_clear_error()
        else:
            check("solveStateMG_halo")
        return it.value, res.value

    def stats(self):
        s = SolveStats()
        self.lib.top3d_solver_stats(ctypes.byref(self.data), ctypes.byref(s))
        return dict(iterations=s.iterations, relres=s.relres, setup_ms=s.setup_ms,
                    solve_ms=s.solve_ms, status=s.status)

    KERNELS = {"fine_sweep_f32": 0, "fine_sweep_f64": 1, "fine_apply_f64": 2,
               "level1_sweep_f32": 3, "level1_sweep_f64": 4, "fine_apply_dot_f64": 5}

    def time_kernel(self, name, reps=20, cold=False):
        """(mean ms per launch, algorithmic bytes per launch) of a hot kernel.
        cold=True flushes the memory-side cache before each launch."""
        ms, by = c_double(), c_double()
        which = self.KERNELS[name] | (16 if cold else 0)
        rc = self.lib.top3d_time_kernel(ctypes.byref(self.data), which, reps,
                                        ctypes.byref(ms), ctypes.byref(by))
        if rc:
            check("top3d_time_kernel")
        return ms.value, by.value


    def sweep_timing(self, enable):
        """Start (enable=True) or stop live timing of the fine sweeps inside the
        V-cycles; returns (mean ms per sweep, sweeps timed, bytes per sweep)."""
        ms, cnt, by = c_double(), c_int(), c_double()
        rc = self.lib.top3d_sweep_timing(ctypes.byref(self.data), int(bool(enable)),
                                         ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(by))
        if rc:
            check("top3d_sweep_timing")
        return ms.value, cnt.value, by.value


class DesignState:
    """Arrays of the SIMP/OC loop (stencil_optimization.c:34-126) on the GPU."""

    def __init__(self, grid, volfrac=0.2, rmin=1.5):
        torch = grid.torch
        self.grid, self.rmin = grid, float(np.float32(rmin))
        self.volfrac = float(np.float32(volfrac))  # (double)(float) as in top3d.c:19
        self.F = grid.load()
        self.U = grid.zeros(0)
        self.x = grid.density(fill=self.volfrac)
        self.xPhys = self.x.clone()
        self.xnew = torch.zeros_like(self.x)
        self.dc = torch.zeros_like(self.x)
        self.dv = torch.ones_like(self.x)
        self.dv = grid.filter_grad(self.rmin, self.dv)
        self.loop = 0
        self.change = 1.0


def design_step(st, cgtol=1e-5):
    """One design iteration of top3dmgcg (stencil_optimization.c:150-189)."""
    g = st.grid
    st.loop += 1
    cgiter, cgres = g.solve(st.xPhys, st.F, st.U, tol=float(np.float32(cgtol)))
    vol = g.volume(st.xPhys)
    c, dc = g.sensitivity(st.xPhys, st.U)
    st.dc = g.filter_grad(st.rmin, dc)
    vol /= g.nelx * g.nely * g.nelz
    g.oc(st.x, st.xnew, st.dv, st.dc, vol - st.volfrac)
    st.change = g.change(st.x, st.xnew)
    st.xPhys = g.filter(st.rmin, st.x)
    return dict(loop=st.loop, c=c, vol=vol, change=st.change, relres=cgres, cgiter=cgiter)


def design_loop(grid, iters, volfrac=0.2, rmin=1.5, cgtol=1e-5):
    st = DesignState(grid, volfrac, rmin)
    recs = []
    while st.change > 1e-2 and st.loop < iters:
        recs.append(design_step(st, cgtol))
    return recs, st


def _compliance_fn():
    import torch

    class _Compliance(torch.autograd.Function):
        @staticmethod
        def forward(ctx, xp, grid, F, U, tol, allow_no_converge):
            U = grid.zeros(0) if U is None else U
            c = c_double()
            grid._sync_torch()
            grid.lib.top3d_clear_error()
            rc = grid.lib.top3d_compliance_forward(ctypes.byref(grid.gpu), ctypes.byref(grid.data),
                                                   grid.nl, tol, _ptr(xp), _ptr(F), _ptr(U),
                                                   ctypes.byref(c))
            st = grid.stats()
            if rc == 5 and not allow_no_converge:
                check("top3d_compliance_forward")
            if rc not in (0, 5):
                check("top3d_compliance_forward")
            grid.lib.top3d_clear_error()
            if rc == 5:
                import warnings
                warnings.warn(f"Compliance: CG stopped at {st['iterations']} iterations with "
                              f"relative residual {st['relres']:.3e} above {tol:.1e}",
                              RuntimeWarning, stacklevel=3)
            ctx.iterations, ctx.relres, ctx.converged = st["iterations"], st["relres"], rc == 0
            grid.last_compliance_solve = dict(iterations=st["iterations"], relres=st["relres"],
                                              converged=rc == 0)
            ctx.grid = grid
            ctx.save_for_backward(xp, U.clone())
            return torch.tensor(c.value, dtype=torch.float64, device=xp.device)

        @staticmethod
        def backward(ctx, gout):
            xp, U = ctx.saved_tensors
            grid = ctx.grid
            grad = torch.zeros_like(xp)
            grid._sync_torch()
            rc = grid.lib.top3d_compliance_backward(ctypes.byref(grid.gpu), 0.0, _ptr(xp),
                                                    _ptr(U), float(gout), _ptr(grad))
            if rc:
                check("top3d_compliance_backward")
            return grad, None, None, None, None, None

    return _Compliance


def _filter_fn():
    import torch

    class _DensityFilter(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, grid, rmin):
            ctx.grid, ctx.rmin = grid, rmin
            return grid.filter(rmin, x.detach())

        @staticmethod
        def backward(ctx, gout):
            grid = ctx.grid
            g = torch.zeros_like(gout)
            idx = grid._interior_index()
            g[idx] = gout[idx]
            out = torch.zeros_like(gout)
            out[idx] = grid.filter_grad(ctx.rmin, g)[idx]
            return out, None, None

    return _DensityFilter


_FNS = {}


class Compliance:
    """torch.autograd.Function c(xPhys) = sum_e E(xPhys_e) u_e^T KE u_e with u
    from the MGCG state solve (solveStateMG_halo, stencil_solvers.c:259-397)
    and c from getComplianceAndSensetivity_halo (stencil_methods.c:808-862).

    backward is the self-adjoint sensitivity dc/dxPhys_e = -3 (E0 - Emin)
    xPhys_e^2 u_e^T KE u_e from the forward displacement (no second solve),
    scaled by the incoming gradient.  It is the gradient with respect to
    xPhys only; compose with DensityFilter for the gradient with respect to
    the design variables (stencil_optimization.c:160-162).

    U: optional float64 displacement tensor used as the warm start and
    overwritten with the solution (the reference carries U from one design
    iteration to the next).  A solve that stops at maxIter above `tol`
    raises Top3dError unless allow_no_converge is set, in which case it
    warns; the iteration count and relative residual are kept on
    grid.last_compliance_solve either way."""

    @staticmethod
    def apply(xphys, grid, F, U=None, tol=1e-5, allow_no_converge=False):
        if "c" not in _FNS:
            _FNS["c"] = _compliance_fn()
        return _FNS["c"].apply(xphys, grid, F, U, float(tol), bool(allow_no_converge))


class DensityFilter:
    """torch.autograd.Function xPhys = (H x) / Hs over the interior elements
    (applyDensityFilter, stencil_optimization.c:228-285); backward pulls a
    gradient back as H (g / Hs) (applyDensityFilterGradient,
    stencil_optimization.c:290-395).  Halo elements get zero gradient."""

    @staticmethod
    def apply(x, grid, rmin):
        if "f" not in _FNS:
            _FNS["f"] = _filter_fn()
        return _FNS["f"].apply(x, grid, float(rmin))


def design_step_autograd(st, cgtol=1e-5, allow_no_converge=False):
    """One design iteration of top3dmgcg (stencil_optimization.c:150-189) in
    which torch autograd supplies the OC update's sensitivities: xPhys =
    DensityFilter(x), c = Compliance(xPhys), c.backward() gives
    dc/dx = H (dc/dxPhys / Hs), the filtered sensitivity of
    stencil_optimization.c:160-162.  The state U is carried over as warm
    start, as in the reference."""
    g = st.grid
    st.loop += 1
    xl = st.x.detach().requires_grad_(True)
    xphys = DensityFilter.apply(xl, g, st.rmin)
    if st.loop == 1:
        # the reference's first iteration uses xPhys = x (stencil_optimization.c:104-120)
        xphys = xphys + (st.xPhys - xphys).detach()
    c = Compliance.apply(xphys, g, st.F, st.U, float(np.float32(cgtol)), allow_no_converge)
    c.backward()
    info = g.last_compliance_solve
    st.xPhys = xphys.detach()
    vol = g.volume(st.xPhys) / (g.nelx * g.nely * g.nelz)
    st.dc = xl.grad
    g.oc(st.x, st.xnew, st.dv, st.dc, vol - st.volfrac)
    st.change = g.change(st.x, st.xnew)
    return dict(loop=st.loop, c=float(c.detach()), vol=vol, change=st.change, relres=info["relres"],
                cgiter=info["iterations"] - 1)


def design_loop_autograd(grid, iters, volfrac=0.2, rmin=1.5, cgtol=1e-5):
    st = DesignState(grid, volfrac, rmin)
    recs = []
    while st.change > 1e-2 and st.loop < iters:
        recs.append(design_step_autograd(st, cgtol))
    st.xPhys = grid.filter(st.rmin, st.x)
    return recs, st
