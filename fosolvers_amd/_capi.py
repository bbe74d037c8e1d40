"""ctypes bindings of include/fosolvers_amd.h."""
import ctypes
import importlib.util
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfosolvers_amd.so")

_lib = None


class LibraryMissing(RuntimeError):
    pass


class PolyGridC(ctypes.Structure):
    _fields_ = [("nN", ctypes.c_int), ("nE", ctypes.c_int), ("maxNNE", ctypes.c_int),
                ("pN", ctypes.POINTER(ctypes.c_double)), ("sE", ctypes.POINTER(ctypes.c_int)),
                ("nNE", ctypes.POINTER(ctypes.c_int)), ("iNE", ctypes.POINTER(ctypes.c_int)),
                ("gid", ctypes.POINTER(ctypes.c_int))]


class FvGridC(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int) for k in ("nN", "nE", "nC", "nF", "nP", "nPi", "nPb", "maxNF", "maxNNE")] + [
        ("pN", ctypes.POINTER(ctypes.c_double)), ("sE", ctypes.POINTER(ctypes.c_int)),
        ("nNE", ctypes.POINTER(ctypes.c_int)), ("iNE", ctypes.POINTER(ctypes.c_int)),
        ("gid", ctypes.POINTER(ctypes.c_int)), ("p", ctypes.POINTER(ctypes.c_double)),
        ("v", ctypes.POINTER(ctypes.c_double)), ("iEP", ctypes.POINTER(ctypes.c_int)),
        ("aP", ctypes.POINTER(ctypes.c_double)), ("normP", ctypes.POINTER(ctypes.c_double)),
        ("pP", ctypes.POINTER(ctypes.c_double)), ("neib", ctypes.POINTER(ctypes.c_int))]


class CondTabC(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("maxPar", ctypes.c_int), ("gid", ctypes.POINTER(ctypes.c_int)),
                ("t", ctypes.POINTER(ctypes.c_int)), ("p", ctypes.POINTER(ctypes.c_double))]


VP = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
LL = ctypes.c_longlong


class BdfStatsC(ctypes.Structure):
    _fields_ = [(k, LL) for k in ("nst", "nfe", "nni", "nli", "npe", "nps", "njv", "nncfails", "nlinfails",
                                  "nfallback")] + [
        ("lastNewtonIters", I), ("lastLinIters", I), ("lastResidualNorm", D), ("nreorth", LL), ("netf", LL),
        ("lastOrder", I), ("nextOrder", I), ("lastStep", D), ("nextStep", D), ("lastResidualNormCvode", D)]


RHS_FN = ctypes.CFUNCTYPE(I, D, ctypes.POINTER(D), ctypes.POINTER(D), VP)
# double *(*)(void *nvector): the N_Vector data accessor of the *_nv callbacks
NV_DATA_FN = ctypes.CFUNCTYPE(VP, VP)

# name: (restype, argtypes)
SIGNATURES = {
    "fos_grid_box": (I, [I, I, I, D, D, D, D, ctypes.c_uint64, ctypes.POINTER(I), ctypes.POINTER(PolyGridC)]),
    "fos_grid_read_gmsh": (I, [ctypes.c_char_p, ctypes.POINTER(PolyGridC)]),
    "fos_grid_write_gmsh": (I, [ctypes.c_char_p, ctypes.POINTER(PolyGridC)]),
    "fos_grid_renumber": (I, [ctypes.POINTER(PolyGridC), I, ctypes.POINTER(I)]),
    "fos_grid_up": (I, [ctypes.POINTER(PolyGridC), ctypes.POINTER(FvGridC)]),
    "fos_grid_up_gpu": (I, [ctypes.POINTER(PolyGridC), ctypes.POINTER(FvGridC), VP]),
    "fos_grid_sort_pairs_reference": (I, [ctypes.POINTER(FvGridC)]),
    "fos_poly_grid_free": (None, [ctypes.POINTER(PolyGridC)]),
    "fos_fv_grid_free": (None, [ctypes.POINTER(FvGridC)]),
    "fos_read_cond_tab": (I, [ctypes.c_char_p, ctypes.POINTER(CondTabC)]),
    "fos_cond_tab_free": (None, [ctypes.POINTER(CondTabC)]),
    "evaluator_create": (VP, [ctypes.c_char_p]),
    "evaluator_destroy": (None, [VP]),
    "evaluator_evaluate": (D, [VP, I, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(D)]),
    "fos_read_udf_tab": (I, [ctypes.c_char_p, ctypes.POINTER(VP), I]),
    "fos_euler_create": (VP, [ctypes.POINTER(FvGridC), ctypes.POINTER(CondTabC), D, D, ctypes.c_uint, VP]),
    "fos_euler_destroy": (None, [VP]),
    "fos_euler_set_udf_bc": (I, [VP, ctypes.POINTER(D)]),
    "fos_euler_rhs": (I, [D, VP, VP, VP]),
    "fos_euler_jtimes": (I, [VP, VP, D, VP, VP, VP, VP]),
    "fos_euler_psetup": (I, [D, VP, VP, I, ctypes.POINTER(I), D, VP]),
    "fos_euler_psolve": (I, [D, VP, VP, VP, VP, D, D, I, VP]),
    "fos_euler_psetup_bjac": (I, [D, VP, VP, I, ctypes.POINTER(I), D, VP]),
    "fos_euler_psolve_bjac": (I, [D, VP, VP, VP, VP, D, D, I, VP]),
    "fos_euler_ilu_order": (I, [VP, VP]),
    "fos_euler_set_tolerances": (I, [VP, D, D]),
    "fos_find_euler": (I, [VP, VP, VP, VP, VP, D, VP, VP, VP]),
    "fos_find_euler_jac": (I, [VP, VP, VP]),
    "fos_euler_rhs_f32": (I, [VP, VP, VP]),
    "fos_euler_jtimes_f32": (I, [VP, VP, VP, VP, VP, D]),
    "fos_euler_set_stream": (I, [VP, VP]),
    "fos_euler_counters": (I, [VP, ctypes.POINTER(ctypes.c_longlong)]),
    "fos_euler_layout": (I, [VP]),
    "fos_bdf_create": (VP, [I, VP, I, VP, VP, VP]),
    "fos_bdf_destroy": (None, [VP]),
    "fos_bdf_set_jtimes": (I, [VP, VP]),
    "fos_bdf_set_newton_matvec": (I, [VP, VP]),
    "fos_bdf_set_options": (I, [VP, ctypes.c_uint, VP]),
    "fos_bdf_set_tol": (I, [VP, D, D]),
    "fos_bdf_set_max_steps": (I, [VP, I]),
    "fos_bdf_set_newton": (I, [VP, D, I, D, I, I]),
    "fos_bdf_set_lsetup_frequency": (I, [VP, I, I, D]),
    "fos_bdf_set_dt": (I, [VP, D]),
    "fos_bdf_set_adaptive": (I, [VP, I, D, D, D]),
    "fos_bdf_get_dky": (I, [VP, D, I, VP]),
    "fos_euler_sync_state": (I, [VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    "fos_euler_jtimes_analytic": (I, [VP, VP, D, VP, VP, VP, VP]),
    "fos_euler_newton_matvec": (I, [D, VP, VP, D, VP, VP, VP]),
    "fos_euler_newton_matvec_analytic": (I, [D, VP, VP, D, VP, VP, VP]),
    "fos_bdf_set_iv": (I, [VP, D, VP]),
    "fos_bdf_step": (I, [VP, D, ctypes.POINTER(D), VP]),
    "fos_bdf_get_dt": (I, [VP, ctypes.POINTER(D)]),
    "fos_bdf_get_stats": (I, [VP, ctypes.POINTER(BdfStatsC)]),
    "fos_bdf_set_fallback": (I, [VP, I, VP, VP]),
    "fos_bdf_last_triple": (I, [VP, VP, VP, VP]),
    "fos_learned_nparams": (I, []),
    "fos_learned_init": (I, [VP, ctypes.c_uint64]),
    "fos_learned_set_params": (I, [VP, ctypes.POINTER(ctypes.c_float)]),
    "fos_learned_get_params": (I, [VP, ctypes.POINTER(ctypes.c_float)]),
    "fos_learned_loss_grad": (I, [VP, VP, VP, VP, D, I, I, ctypes.POINTER(D), ctypes.POINTER(ctypes.c_float)]),
    "fos_learned_inputs": (I, [VP, VP, D, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "fos_learned_train_step": (I, [VP, VP, VP, VP, D, I, I, D, ctypes.POINTER(D)]),
    "fos_euler_psetup_learned": (I, [D, VP, VP, I, ctypes.POINTER(I), D, VP]),
    "fos_euler_psolve_learned": (I, [D, VP, VP, VP, VP, D, D, I, VP]),
    "fos_euler_jtimes_noalloc": (I, [VP, VP, D, VP, VP, VP, VP]),
    "fos_euler_nonfinite": (I, [VP, I]),
    "fos_euler_set_nvector_access": (I, [VP, VP]),
    "fos_euler_rhs_nv": (I, [D, VP, VP, VP]),
    "fos_euler_jtimes_nv": (I, [VP, VP, D, VP, VP, VP, VP]),
    "fos_euler_jtimes_analytic_nv": (I, [VP, VP, D, VP, VP, VP, VP]),
    "fos_euler_psetup_nv": (I, [D, VP, VP, I, ctypes.POINTER(I), D, VP]),
    "fos_euler_psolve_nv": (I, [D, VP, VP, VP, VP, D, D, I, VP]),
    "fos_euler_write_vtk": (I, [VP, ctypes.POINTER(FvGridC), VP, ctypes.c_char_p]),
    "fos_ilut_create": (VP, [I, I, I, ctypes.c_uint, VP]),
    "fos_ilut_destroy": (None, [VP]),
    "fos_ilut_set_coo": (I, [VP, VP, VP, VP, I]),
    "fos_ilut_set_csr": (I, [VP, VP, VP, VP]),
    "fos_ilut_mul_vec": (I, [VP, VP, VP]),
    "fos_ilut_fact": (I, [VP, D]),
    "fos_ilut_solve": (I, [VP, VP, VP]),
    "fos_ilut_get_msr": (I, [VP, VP, VP, VP]),
    "fos_ilut_levels": (I, [VP, ctypes.POINTER(I), ctypes.POINTER(I)]),
    "fos_last_error": (ctypes.c_char_p, []),
    "fos_version": (ctypes.c_char_p, []),
}


def lib():
    """Load libfosolvers_amd.so once; raise LibraryMissing when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} not built: run python -m fosolvers_amd.build")
        # PyTorch ships its own copy of the HIP runtime (same soname, libamdhip64.so.7).  Loaded
        # first, it also serves this library; loaded after the system copy, the process holds two
        # runtimes and the second to start finds no device.  So torch, when installed, comes first.
        if importlib.util.find_spec("torch") is not None:
            import torch  # noqa: F401
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def last_error():
    return lib().fos_last_error().decode(errors="replace")


def check(rc):
    if rc != 0:
        raise RuntimeError(f"fosolvers_amd error {rc}: {last_error()}")
    return rc
