"""ctypes binding of the C ABI in include/pzg.h (libpzg.so, built for gfx950).

There is no CPU fallback: if the HIP library is missing or a call fails, an
exception is raised.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libpzg.so")

PZG_OK, PZG_EINVAL, PZG_EHIP, PZG_ENOCONV, PZG_ESTATE, PZG_ENOMEM, PZG_EUNSUP = range(7)
SOLVER_PCG64, SOLVER_MIXED = 0, 1
PRECOND_JACOBI, PRECOND_BLOCK_JACOBI, PRECOND_LEARNED, PRECOND_MULTIGRID = 0, 1, 2, 3
INNER_AUTO, INNER_BSR, INNER_MATFREE = 0, 1, 2
MG_LEARNED = 4  # pzg_mg_apply: level-0 smoothing by the learned element smoother
LES_NH = 32  # hidden features of the learned smoother's network; its output layer is the last LES_NH weights
EXACT_POINTS, EXACT_CONST, EXACT_X, EXACT_SINSIN, EXACT_COSCOS = range(5)

EXPORTS = [
    "pzg_last_error", "pzg_create", "pzg_destroy", "pzg_mesh_set", "pzg_struct_create",
    "pzg_struct_get_pattern", "pzg_assemble", "pzg_get_system", "pzg_csr_set",
    "pzg_add_kel_batch", "pzg_multadd", "pzg_solve", "pzg_true_residual", "pzg_time_kernel",
    "pzg_get_sizes", "pzg_sync", "pzg_mesh_set_coef", "pzg_learned_setup", "pzg_learned_collect",
    "pzg_learned_train", "pzg_learned_weights", "pzg_learned_graph", "pzg_learned_samples",
    "pzg_learned_loss_grad", "pzg_scaled_apply", "pzg_mg_apply", "pzg_full_csr", "pzg_hp_tuning",
    "pzg_load_solution", "pzg_error_points", "pzg_evaluate_error", "pzg_sloan_renumber",
    "pzg_get_solution", "pzg_set_equation_filter", "pzg_learned_scales",
    "pzg_learned_apply", "pzg_true_residual_noalloc",
]

dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("material", C.c_int32), ("nstate", C.c_int32), ("porder", C.c_int32),
        ("nnodes", C.c_int64), ("xyz", dp),
        ("nvol", C.c_int64), ("vol_nodes", i64p), ("vol_conn", i64p), ("vol_coef", dp),
        ("nbc", C.c_int64), ("bc_nodes", i64p), ("bc_conn", i64p), ("bc_id", i32p),
        ("nbcid", C.c_int32), ("bc_type", i32p), ("bc_val", dp),
        ("nconn", C.c_int64), ("blk_pos", i64p), ("blk_size", i32p), ("neq", C.c_int64),
        ("scale", C.c_double), ("E", C.c_double), ("nu", C.c_double),
        ("rule_order_vol", C.c_int32), ("rule_order_bc", C.c_int32), ("force_kind", C.c_int32),
        ("force_const", C.c_double * 3),
        ("nbumps", C.c_int32), ("bumps", dp), ("bump_width", C.c_double),
        ("big_number", C.c_double),
        ("vol_rule", i32p), ("bc_rule", i32p), ("vol_etype", C.c_int32),
        ("dep_ptr", i64p), ("dep_conn", i64p), ("dep_coef", dp),
        ("bc_val1", dp),
        ("prestress", C.c_double * 3),
    ]


class SolveParams(C.Structure):
    _fields_ = [
        ("solver", C.c_int32), ("precond", C.c_int32), ("max_iter", C.c_int64),
        ("tol", C.c_double), ("inner_tol", C.c_double), ("max_outer", C.c_int32),
        ("check_every", C.c_int32), ("use_graph", C.c_int32), ("refine", C.c_int32),
        ("inner_op", C.c_int32), ("mg_degree", C.c_int32), ("mg_range", C.c_double),
    ]


class SolveInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64), ("outer", C.c_int32), ("converged", C.c_int32),
        ("achieved", C.c_double), ("true_residual", C.c_double), ("solve_ms", C.c_double),
        ("fallback", C.c_int32),
    ]


class Sizes(C.Structure):
    _fields_ = [
        ("neq", C.c_int64), ("nnz_upper", C.c_int64), ("nnz_full", C.c_int64),
        ("nblocks", C.c_int64), ("nnz_blocks", C.c_int64), ("nvol", C.c_int64),
        ("ncolors", C.c_int64), ("bs", C.c_int32),
    ]


class Exact(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("porder", C.c_int32), ("center", C.c_double * 3),
        ("constant", C.c_double), ("npts", C.c_int64), ("u", dp), ("du", dp),
    ]


class PzgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pzg error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Loads libpzg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is missing: run __graft_entry__.build() first "
                           "(the HIP path has no CPU fallback)")
    L = C.CDLL(LIBPATH)
    vp = C.c_void_p
    L.pzg_last_error.restype = C.c_char_p
    L.pzg_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.pzg_destroy.argtypes = [vp]
    L.pzg_destroy.restype = None
    L.pzg_mesh_set.argtypes = [vp, C.POINTER(MeshDesc)]
    L.pzg_struct_create.argtypes = [vp, i64p, i64p]
    L.pzg_struct_get_pattern.argtypes = [vp, i64p, i64p]
    L.pzg_assemble.argtypes = [vp]
    L.pzg_get_system.argtypes = [vp, dp, dp]
    L.pzg_csr_set.argtypes = [vp, C.c_int64, i64p, i64p, dp, dp]
    L.pzg_add_kel_batch.argtypes = [vp, C.c_int64, C.c_int32, i64p, dp]
    L.pzg_multadd.argtypes = [vp, dp, dp, dp, C.c_double, C.c_double]
    L.pzg_solve.argtypes = [vp, C.POINTER(SolveParams), dp, dp, C.POINTER(SolveInfo)]
    L.pzg_true_residual.argtypes = [vp, dp]
    L.pzg_true_residual_noalloc.argtypes = [vp, dp]
    L.pzg_time_kernel.argtypes = [vp, C.c_int32, C.c_int32, dp, dp]
    L.pzg_get_sizes.argtypes = [vp, C.POINTER(Sizes)]
    L.pzg_sync.argtypes = [vp]
    L.pzg_mesh_set_coef.argtypes = [vp, dp]
    L.pzg_learned_setup.argtypes = [vp, C.c_uint64]
    L.pzg_learned_collect.argtypes = [vp, C.c_int32, C.c_uint64]
    L.pzg_learned_train.argtypes = [vp, C.c_int32, C.c_double, dp, dp]
    L.pzg_learned_weights.argtypes = [vp, fp, fp, i32p]
    L.pzg_learned_graph.argtypes = [vp, i64p, i64p, i32p, fp, fp]
    L.pzg_learned_samples.argtypes = [vp, i32p, i32p, fp, fp, fp]
    L.pzg_learned_loss_grad.argtypes = [vp, dp, fp]
    L.pzg_learned_scales.argtypes = [vp, fp]
    L.pzg_learned_apply.argtypes = [vp, dp, dp]
    L.pzg_scaled_apply.argtypes = [vp, C.c_int32, dp, dp]
    L.pzg_mg_apply.argtypes = [vp, C.c_int32, C.c_int32, C.c_double, dp, dp, i32p, i32p, dp]
    L.pzg_full_csr.argtypes = [vp, i64p, i64p, dp]
    L.pzg_hp_tuning.argtypes = [vp, i32p, i32p, i32p, fp]
    L.pzg_load_solution.argtypes = [vp, dp]
    L.pzg_get_solution.argtypes = [vp, dp]
    L.pzg_error_points.argtypes = [vp, C.c_int32, i64p, dp]
    L.pzg_evaluate_error.argtypes = [vp, C.POINTER(Exact), dp, dp, i32p]
    L.pzg_sloan_renumber.argtypes = [C.c_int64, C.c_int64, i64p, i64p, i64p]
    L.pzg_set_equation_filter.argtypes = [vp, C.c_int64, i64p]
    _lib = L
    return L


def check(rc, allow=()):
    if rc != PZG_OK and rc not in allow:
        raise PzgError(rc, load().pzg_last_error().decode())
    return rc


def ptr(a, kind):
    if a is None:
        return None
    return a.ctypes.data_as(kind)


def make_mesh_desc(m):
    """Builds a MeshDesc from the dict of neopz_amd.mesh.build; returns (desc, keepalive)."""
    keep = []

    def arr(x, dt):
        if x is None:
            return None
        y = np.ascontiguousarray(x, dtype=dt)
        keep.append(y)
        return y

    d = MeshDesc()
    d.dim, d.material, d.nstate, d.porder = m["dim"], m["material"], m["nstate"], m["porder"]
    xyz = arr(m["xyz"], np.float64)
    d.nnodes = len(xyz) // 3
    d.xyz = ptr(xyz, dp)
    vn = arr(m["vol_nodes"], np.int64)
    vc = arr(m["vol_conn"], np.int64)
    d.nvol = vn.shape[0]
    d.vol_nodes = ptr(vn, i64p)
    d.vol_conn = ptr(vc, i64p)
    d.vol_coef = ptr(arr(m["vol_coef"], np.float64), dp)
    bn = arr(m["bc_nodes"], np.int64)
    d.nbc = bn.shape[0]
    d.bc_nodes = ptr(bn, i64p)
    d.bc_conn = ptr(arr(m["bc_conn"], np.int64), i64p)
    d.bc_id = ptr(arr(m["bc_id"], np.int32), i32p)
    bt = arr(m["bc_type"], np.int32)
    d.nbcid = len(bt)
    d.bc_type = ptr(bt, i32p)
    d.bc_val = ptr(arr(m["bc_val"], np.float64), dp)
    if m.get("bc_val1") is not None:
        d.bc_val1 = ptr(arr(m["bc_val1"], np.float64), dp)
    for k in range(3):
        d.prestress[k] = float(m.get("prestress", (0.0, 0.0, 0.0))[k])
    d.nconn = m["nconn"]
    d.blk_pos = ptr(arr(m["blk_pos"], np.int64), i64p)
    d.blk_size = ptr(arr(m["blk_size"], np.int32), i32p)
    d.neq = m["neq"]
    d.scale, d.E, d.nu = m["scale"], m["E"], m["nu"]
    d.rule_order_vol, d.rule_order_bc = m["rule_order_vol"], m["rule_order_bc"]
    d.force_kind = m["force_kind"]
    for k in range(3):
        d.force_const[k] = float(m["force_const"][k])
    b = arr(m["bumps"], np.float64)
    d.nbumps = 0 if b is None else b.shape[0]
    d.bumps = ptr(b, dp)
    d.bump_width = m["bump_width"]
    d.big_number = m["big_number"]
    d.vol_rule = ptr(arr(m.get("vol_rule"), np.int32), i32p)
    d.bc_rule = ptr(arr(m.get("bc_rule"), np.int32), i32p)
    d.vol_etype = int(m.get("vol_etype", 0))
    if m.get("dep_ptr") is not None:
        d.dep_ptr = ptr(arr(m["dep_ptr"], np.int64), i64p)
        d.dep_conn = ptr(arr(m["dep_conn"], np.int64), i64p)
        d.dep_coef = ptr(arr(m["dep_coef"], np.float64), dp)
    return d, keep


def hp_tuning(ctx):
    """Choices of the high-order assembly autotuner: per rule bucket the
    highest order, the kept variant and the milliseconds of every variant."""
    n = C.c_int32(64)
    order = np.zeros(64, np.int32)
    var = np.zeros(64, np.int32)
    ms = np.zeros(64 * 5, np.float32)
    check(ctx.lib.pzg_hp_tuning(ctx.h, C.byref(n), ptr(order, i32p), ptr(var, i32p), ptr(ms, fp)))
    k = n.value
    return [{"order": int(order[b]), "variant": int(var[b]),
             "ms_per_variant": [round(float(x), 4) for x in ms[5 * b:5 * b + 5]]} for b in range(k)]
