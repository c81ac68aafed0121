"""ctypes binding of libgdx.so (the C ABI declared in include/gdx.h).

The library must have been built in-tree (``python gridapdistributed.jl_amd/build.py``);
there is no CPU fallback: every entry point fails loudly when the library or
the GPU is missing.
"""
import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("GDX_LIBRARY", os.path.join(_PKG_ROOT, "libgdx.so"))

GDX_OK = 0
ERRORS = {-1: "GDX_ERR_ARG", -2: "GDX_ERR_HIP", -3: "GDX_ERR_STATE", -4: "GDX_ERR_UNSUPPORTED", -5: "GDX_ERR_ALLOC"}
SUB_ASSEMBLED_ROWS, FULLY_ASSEMBLED_ROWS = 0, 1
PREC_FP64, PREC_FP32 = 0, 1
KERNEL_HBM_PATTERNS = 0x100  # gdx_problem_set_kernel flag: position patterns read from HBM in every tile
RHS_CONST, RHS_MANUFACTURED = 0, 1
OP_DIFFUSION, OP_FLUX, OP_ELASTICITY = 0, 1, 2
DIR_CLAMPED_X0 = 2
NEWTON_INIT_GIVEN, NEWTON_INIT_POISSON = 0, 1
SZ_NLDOFS, SZ_NOWN, SZ_NGHOST_ROWS, SZ_NGHOST_COLS, SZ_NNZ, SZ_COUNT = range(6)

# every function declared in include/gdx.h
EXPORTED = [
    "gdx_version", "gdx_last_error", "gdx_init", "gdx_finalize", "gdx_problem_create", "gdx_problem_destroy",
    "gdx_symbolic", "gdx_numeric", "gdx_num_parts", "gdx_num_global_dofs", "gdx_part_sizes", "gdx_export_fe",
    "gdx_export_rows", "gdx_export_cols", "gdx_export_csr", "gdx_export_rhs", "gdx_device_values", "gdx_stream",
    "gdx_sync", "gdx_phase_times", "gdx_owned_nnz", "gdx_spmv", "gdx_pcg", "gdx_set_state", "gdx_newton",
    "gdx_set_cell_coefficient", "gdx_cell_coefficient_grf", "gdx_export_cell_coefficient", "gdx_set_cell_lame",
    "gdx_set_vertex_coords", "gdx_set_reaction",
    "gdx_nn_create", "gdx_nn_destroy", "gdx_nn_num_params", "gdx_nn_save", "gdx_nn_load", "gdx_nn_forward",
    "gdx_nn_backward", "gdx_nn_attach", "gdx_nn_train_step", "gdx_nn_apply", "gdx_nn_apply_noalloc", "gdx_fpcg",
    "gdx_nn_set_options", "gdx_nn_coarse_operator", "gdx_nn_set_multigrid", "gdx_nn_level_operator",
    "gdx_batch_create", "gdx_batch_add", "gdx_batch_autotune", "gdx_batch_capture", "gdx_batch_launch",
    "gdx_batch_destroy", "gdx_problem_set_kernel", "gdx_problem_get_kernel",
    "gdx_pvector_create", "gdx_pvector_destroy", "gdx_pvector_length", "gdx_pvector_device", "gdx_pvector_set",
    "gdx_pvector_get", "gdx_pvector_fill", "gdx_pvector_copy_rhs", "gdx_pvector_consistent", "gdx_pvector_assemble",
    "gdx_pvector_axpy", "gdx_pvector_change_ghost", "gdx_psparse_mul", "gdx_spmv_timed",
]
PRANGE_ROWS, PRANGE_COLS = 0, 1


class GdxError(RuntimeError):
    def __init__(self, code, func, msg):
        super().__init__(f"{func} failed with {ERRORS.get(code, code)}: {msg}")
        self.code = code


class ProblemDesc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("order", ctypes.c_int32), ("qdeg", ctypes.c_int32), ("strategy", ctypes.c_int32),
        ("cells", ctypes.c_int64 * 3), ("parts", ctypes.c_int32 * 3), ("domain", ctypes.c_double * 6),
        ("coef_const", ctypes.c_double), ("nmodes", ctypes.c_int32), ("modes", ctypes.POINTER(ctypes.c_double)),
        ("rhs_kind", ctypes.c_int32), ("rhs_const", ctypes.c_double), ("dir_kind", ctypes.c_int32),
        ("dir_coef", ctypes.c_double * 4), ("op_kind", ctypes.c_int32), ("flux_p", ctypes.c_double),
        ("flux_alpha", ctypes.c_double), ("flux_beta", ctypes.c_double),
        ("lame_lambda", ctypes.c_double), ("lame_mu", ctypes.c_double), ("body_force", ctypes.c_double * 3),
    ]


class NewtonOpts(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("maxiter", ctypes.c_int32),
                ("lin_rtol", ctypes.c_double), ("lin_maxiter", ctypes.c_int32), ("init", ctypes.c_int32),
                ("precond", ctypes.c_void_p), ("jac_precision", ctypes.c_int32)]


class NewtonReport(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int32), ("lin_iters", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("res_ref", ctypes.c_double), ("res_final", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("ms_assembly", ctypes.c_double), ("ms_solve", ctypes.c_double), ("lin_fallbacks", ctypes.c_int32)]


_LIB = None


def lib():
    """Load libgdx.so; raises if it was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python gridapdistributed.jl_amd/build.py` "
                           "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    P = ctypes.POINTER
    sig = {
        "gdx_version": [P(i32), P(i32), P(i32)],
        "gdx_init": [i32, P(vp)],
        "gdx_finalize": [vp],
        "gdx_problem_create": [vp, P(ProblemDesc), P(vp)],
        "gdx_problem_destroy": [vp],
        "gdx_symbolic": [vp],
        "gdx_numeric": [vp, i32],
        "gdx_num_parts": [vp, P(i32)],
        "gdx_num_global_dofs": [vp, P(i64)],
        "gdx_part_sizes": [vp, i32, P(i64)],
        "gdx_export_fe": [vp, i32, vp, vp, vp],
        "gdx_export_rows": [vp, i32, vp, vp],
        "gdx_export_cols": [vp, i32, vp, vp],
        "gdx_export_csr": [vp, i32, vp, vp, vp],
        "gdx_export_rhs": [vp, i32, vp],
        "gdx_device_values": [vp, P(vp), P(i64)],
        "gdx_stream": [vp, P(vp)],
        "gdx_sync": [vp],
        "gdx_pvector_create": [vp, i32, P(vp)],
        "gdx_pvector_destroy": [vp],
        "gdx_pvector_length": [vp, P(i64)],
        "gdx_pvector_device": [vp, P(vp)],
        "gdx_pvector_set": [vp, vp],
        "gdx_pvector_get": [vp, vp],
        "gdx_pvector_fill": [vp, ctypes.c_double],
        "gdx_pvector_copy_rhs": [vp],
        "gdx_pvector_consistent": [vp],
        "gdx_pvector_assemble": [vp],
        "gdx_pvector_axpy": [ctypes.c_double, vp, vp],
        "gdx_pvector_change_ghost": [vp, vp, i32],
        "gdx_psparse_mul": [vp, vp, vp],
        "gdx_phase_times": [vp, P(ctypes.c_double), i32],
        "gdx_owned_nnz": [vp, P(i64)],
        "gdx_spmv": [vp, vp, vp],
        "gdx_spmv_timed": [vp, i32, P(ctypes.c_double), P(ctypes.c_double)],
        "gdx_pcg": [vp, ctypes.c_double, i32, vp, P(i32), P(ctypes.c_double), P(ctypes.c_double)],
        "gdx_set_state": [vp, vp],
        "gdx_set_cell_coefficient": [vp, vp],
        "gdx_cell_coefficient_grf": [vp, ctypes.c_uint64, ctypes.c_double, ctypes.c_double],
        "gdx_export_cell_coefficient": [vp, vp],
        "gdx_set_cell_lame": [vp, vp],
        "gdx_set_vertex_coords": [vp, vp],
        "gdx_set_reaction": [vp, ctypes.c_double, ctypes.c_double],
        "gdx_nn_create": [vp, i32, ctypes.c_uint64, P(vp)],
        "gdx_nn_destroy": [vp],
        "gdx_nn_num_params": [vp, P(i64)],
        "gdx_nn_save": [vp, vp],
        "gdx_nn_load": [vp, vp],
        "gdx_nn_forward": [vp, vp, vp],
        "gdx_nn_backward": [vp, vp, vp, vp],
        "gdx_nn_attach": [vp, vp],
        "gdx_nn_train_step": [vp, ctypes.c_uint64, ctypes.c_float, i32, P(ctypes.c_float)],
        "gdx_nn_apply": [vp, vp, vp],
        "gdx_nn_apply_noalloc": [vp, vp, vp],
        "gdx_nn_set_options": [vp, i32, i32],
        "gdx_nn_coarse_operator": [vp, vp],
        "gdx_nn_set_multigrid": [vp, i32, i32, vp],
        "gdx_nn_level_operator": [vp, i32, vp, vp],
        "gdx_batch_create": [vp, P(vp)],
        "gdx_batch_add": [vp, vp],
        "gdx_batch_autotune": [vp, i32, vp],
        "gdx_batch_capture": [vp],
        "gdx_batch_launch": [vp],
        "gdx_batch_destroy": [vp],
        "gdx_problem_set_kernel": [vp, i32, i32],
        "gdx_problem_get_kernel": [vp, P(i32), P(i32)],
        "gdx_fpcg": [vp, vp, vp, ctypes.c_double, i32, vp, P(i32), P(ctypes.c_double), P(ctypes.c_double), P(i32)],
        "gdx_newton": [vp, P(NewtonOpts), vp, P(NewtonReport)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.gdx_last_error.argtypes = []
    L.gdx_last_error.restype = ctypes.c_char_p
    _LIB = L
    return L


def check(rc, func):
    if rc != GDX_OK:
        msg = lib().gdx_last_error()
        raise GdxError(rc, func, msg.decode() if msg else "")


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)
    return rc
