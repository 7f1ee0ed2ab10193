"""ctypes binding of libcpfem_hip.so (declarations mirror include/cpfem.h).

The library is the only compute path: if it is missing, or no HIP device is visible,
``lib()`` raises instead of falling back to anything else.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpfem_hip.so")

OK, EBADARG, EHIP, ELOCAL, EKRYLOV, ENONFINITE, ENOMEM, EDIRECT = range(8)
FP64, MIXED = 0, 1
APPLY_BC = 1
XM, GSS_A, H, TSAT, R, C11, C12, C44, NPARAM = range(9)
PT_CAPPED, PT_NONFINITE = 1, 2
NOALLOC_MAX_DOFS = 196608  # CPFEM_NOALLOC_MAX_DOFS

c_dp = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p


class CpfemMesh(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_cells", ctypes.c_int64),
                ("points", vp), ("cells", vp)]


class CpfemSlip(ctypes.Structure):
    _fields_ = [("n_slip", ctypes.c_int32), ("directions", vp), ("normals", vp)]


class CpfemState(ctypes.Structure):
    _fields_ = [("Fp_inv", vp), ("g", vp), ("slip", vp), ("rot", vp)]


class CpfemStateOut(ctypes.Structure):
    _fields_ = [("Fp_inv", vp), ("g", vp), ("slip", vp)]


class CpfemParams(ctypes.Structure):
    _fields_ = [("ptr", vp * NPARAM), ("stride", ctypes.c_int64 * NPARAM),
                ("gamma0", ctypes.c_double), ("dt", ctypes.c_double), ("tol", ctypes.c_double),
                ("max_sub_step", ctypes.c_int32), ("max_iter", ctypes.c_int32),
                ("self_block", ctypes.c_int32)]


class CpfemVjp(ctypes.Structure):
    _fields_ = [("Pbar", vp), ("sbar_cell", vp), ("Fbar", vp), ("gbar", vp), ("slipbar", vp),
                ("Hbar", vp), ("Fpibar", vp), ("gbar_old", vp), ("pbar", vp), ("Rbar", vp),
                ("mask", ctypes.c_int32), ("accumulate", ctypes.c_int32)]


VJP_H, VJP_FPI, VJP_G, VJP_MAT, VJP_R = 1, 2, 4, 8, 16


class CpfemSolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("outer", ctypes.c_int32),
                ("residual", ctypes.c_double), ("b_norm", ctypes.c_double)]


_SIGS = {
    "cpfem_last_error": (ctypes.c_char_p, []),
    "cpfem_version": (ctypes.c_int, []),
    "cpfem_create": (ctypes.c_int, [ctypes.POINTER(CpfemMesh), ctypes.POINTER(CpfemSlip), ctypes.POINTER(vp)]),
    "cpfem_destroy": (None, [vp]),
    "cpfem_info": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64)]),
    "cpfem_pattern": (ctypes.c_int, [vp, vp, vp]),
    "cpfem_geometry": (ctypes.c_int, [vp, vp, vp]),
    "cpfem_set_dirichlet": (ctypes.c_int, [vp, vp, ctypes.c_int64, vp, vp]),
    "cpfem_bc_x0": (ctypes.c_int, [vp, vp, vp, vp]),
    "cpfem_local": (ctypes.c_int, [vp, vp, ctypes.c_int64, vp, vp, ctypes.POINTER(CpfemState),
                                   ctypes.POINTER(CpfemParams), vp, vp, vp, ctypes.POINTER(CpfemStateOut), vp]),
    "cpfem_assemble": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(CpfemState), ctypes.POINTER(CpfemParams),
                                      vp, vp, vp, ctypes.c_int32]),
    "cpfem_cell_stress": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(CpfemState), ctypes.POINTER(CpfemParams), vp]),
    "cpfem_update_state": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(CpfemState), ctypes.POINTER(CpfemParams),
                                          ctypes.POINTER(CpfemStateOut)]),
    "cpfem_solve": (ctypes.c_int, [vp, vp, vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(CpfemSolveInfo)]),
    "cpfem_solve_noalloc": (ctypes.c_int, [vp, vp, vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.POINTER(CpfemSolveInfo)]),
    "cpfem_spmv": (ctypes.c_int, [vp, vp, vp, vp, vp]),
    "cpfem_solve_direct": (ctypes.c_int, [vp, vp, vp, vp, vp, ctypes.c_int32, ctypes.POINTER(CpfemSolveInfo)]),
    "cpfem_band_order": (ctypes.c_int, [ctypes.c_int64, vp, vp, vp, vp]),
    "cpfem_box_numbering": (ctypes.c_int, [ctypes.c_int64, vp, vp, vp]),
    "cpfem_norm": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)]),
    "cpfem_axpby": (ctypes.c_int, [vp, vp, ctypes.c_double, vp, ctypes.c_double, vp]),
    "cpfem_to_csr": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
    "cpfem_local_vjp": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(CpfemState), ctypes.POINTER(CpfemParams),
                                       ctypes.POINTER(CpfemVjp)]),
    "cpfem_ugrad_transpose": (ctypes.c_int, [vp, vp, vp, vp]),
    "cpfem_grad_field": (ctypes.c_int, [vp, vp, vp, ctypes.c_double, vp]),
    "cpfem_transpose": (ctypes.c_int, [vp, vp, vp, vp]),
    "cpfem_profile": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32]),
    "cpfem_profile_read": (ctypes.c_int, [vp, ctypes.c_int32, vp, vp, vp]),
    "cpfem_copy": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_int64]),
    "cpfem_point_counts": (ctypes.c_int, [vp, vp, ctypes.c_int32, vp]),
    "cpfem_krylov_config": (ctypes.c_int, [vp, vp]),
    "cpfem_set_spmv_form": (ctypes.c_int, [vp, ctypes.c_int32]),
    "cpfem_set_box_columns": (ctypes.c_int, [vp, ctypes.c_int32]),
    "cpfem_set_point_park": (ctypes.c_int, [vp, ctypes.c_int32]),
}

PROF_CLASSES = ("element_qp", "element_K", "gather_matrix", "gather_residual", "krylov_spmv", "krylov_vector",
                "point_local", "copy", "krylov_replace", "element_qp_park", "point_tangent")

EXPORTED = tuple(_SIGS)
_lib = None


class CpfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cpfem error {code}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    """Load the shared library and declare every entry point (no device needed)."""
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build() (the HIP library is required)")
    L = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib():
    """The loaded library; requires a visible HIP device."""
    global _lib
    if _lib is None:
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("cpfem_amd needs a HIP device (MI355X); none is visible")
        _lib = load_library()
    return _lib


def check(rc):
    if rc != OK:
        raise CpfemError(rc, lib().cpfem_last_error().decode(errors="replace"))
    return rc
