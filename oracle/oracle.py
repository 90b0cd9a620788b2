"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``, always as the checker.  The product
(``neopz_amd``) never imports it.

The C restatement follows NeoPZ's own algorithms (file:line citations in
``oracle/pz_oracle.c``).  The pins of each function against the reference's
golden vectors, closed forms and invariants live in ``tests/test_oracle.py``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

POISSON, DARCY, ELAST3D, MOCK1D = 0, 1, 2, 3
LINE, QUAD, HEX, TRI, TET = 1, 2, 3, 4, 5
MAXNE = 192  # largest element dof count (ORC_MAXNE)
FORCE_NONE, FORCE_CONST, FORCE_BUMPS = 0, 1, 2


class _Problem(C.Structure):
    _fields_ = [
        ("material", C.c_int), ("dim", C.c_int), ("p", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("scale", C.c_double), ("E", C.c_double), ("nu", C.c_double),
        ("el_coef", C.POINTER(C.c_double)),
        ("force_kind", C.c_int), ("force_porder", C.c_int),
        ("force_const", C.c_double * 3),
        ("nbumps", C.c_int), ("bumps", C.POINTER(C.c_double)),
        ("bump_width", C.c_double),
        ("bc_type", C.c_int * 6), ("bc_val", (C.c_double * 3) * 6),
        ("el_porder", C.POINTER(C.c_int)), ("simplex", C.c_int),
        ("bc_val1", (C.c_double * 9) * 6),
        ("prestress", C.c_double * 3),
    ]


EXACT_POINTS, EXACT_CONST, EXACT_X, EXACT_SINSIN, EXACT_COSCOS = range(5)


class _Exact(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("porder", C.c_int), ("center", C.c_double * 3),
        ("constant", C.c_double), ("u", C.POINTER(C.c_double)), ("du", C.POINTER(C.c_double)),
    ]


class _Mesh(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("p", C.c_int), ("nstate", C.c_int),
        ("nnodes", C.c_int64), ("xyz", C.POINTER(C.c_double)),
        ("nel", C.c_int64), ("nvol", C.c_int64),
        ("etype", C.POINTER(C.c_int)), ("bcface", C.POINTER(C.c_int)),
        ("enodes", C.POINTER(C.c_int64)), ("econn", C.POINTER(C.c_int64)),
        ("nconn", C.c_int64), ("conn_order", C.POINTER(C.c_int)),
        ("conn_nshape", C.POINTER(C.c_int)),
        ("conn_seq", C.POINTER(C.c_int64)), ("blk_pos", C.POINTER(C.c_int64)),
        ("blk_size", C.POINTER(C.c_int)), ("neq", C.c_int64),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P, M = C.POINTER(_Problem), C.POINTER(_Mesh)
        dp, i64p = C.POINTER(C.c_double), C.POINTER(C.c_int64)
        L.orc_gauss_legendre.argtypes = [C.c_int, dp, dp]
        L.orc_mesh_build.argtypes = [P, M]
        L.orc_mesh_free.argtypes = [M]
        L.orc_mesh_permute.argtypes = [M, i64p]
        L.orc_calcstiff.argtypes = [P, M, C.c_int64, dp, dp, i64p]
        L.orc_shape.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.orc_contribute.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, dp, dp,
                                     dp, C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_setup_pattern.argtypes = [M, C.POINTER(i64p), C.POINTER(i64p)]
        L.orc_setup_pattern.restype = C.c_int64
        L.orc_setup_pattern_full.argtypes = [M, C.POINTER(i64p), C.POINTER(i64p)]
        L.orc_setup_pattern_full.restype = C.c_int64
        L.orc_assemble_full.argtypes = [P, M, i64p, i64p, dp, dp]
        L.orc_assemble.argtypes = [P, M, i64p, i64p, dp, dp]
        L.orc_assemble_colored.argtypes = [P, M, i64p, i64p, dp, dp, C.c_int]
        L.orc_assemble_range.argtypes = [P, M, i64p, i64p, dp, dp, C.c_int64, C.c_int64]
        L.orc_addkel.argtypes = [i64p, i64p, dp, C.c_int, dp, i64p]
        L.orc_multadd.argtypes = [C.c_int64, i64p, i64p, dp, dp, dp, dp, C.c_double, C.c_double]
        L.orc_pcg.argtypes = [C.c_int64, i64p, i64p, dp, dp, dp, i64p, dp, C.c_int, dp]
        L.orc_pcg_block.argtypes = [C.c_int64, i64p, i64p, dp, dp, C.c_int64, C.POINTER(C.c_int32), dp,
                                    i64p, dp, dp]
        L.orc_simplex_rule.argtypes = [C.c_int, C.c_int, dp, dp]
        L.orc_collapsed_rule.argtypes = [C.c_int, C.c_int, dp, dp]
        L.orc_shape_h1.argtypes = [C.c_int, i64p, C.POINTER(C.c_int), dp, dp, dp]
        L.orc_error_points.argtypes = [P, M, C.c_int, dp]
        L.orc_error_points.restype = C.c_int64
        L.orc_evaluate_error.argtypes = [P, M, dp, C.POINTER(_Exact), dp, dp]
        ip = C.POINTER(C.c_int)
        L.orc_mesh_from_arrays.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, dp,
                                           C.c_int64, i64p, i64p, C.c_int64, i64p, i64p, ip,
                                           C.c_int64, i64p, ip, ip, C.c_int64, M]
        L.orc_setup_pattern_constrained.argtypes = [M, i64p, i64p, C.POINTER(i64p), C.POINTER(i64p)]
        L.orc_setup_pattern_constrained.restype = C.c_int64
        L.orc_assemble_constrained.argtypes = [P, M, i64p, i64p, dp, i64p, i64p, dp, dp]
        L.orc_setup_pattern_filtered.argtypes = [M, i64p, C.POINTER(i64p), C.POINTER(i64p)]
        L.orc_setup_pattern_filtered.restype = C.c_int64
        L.orc_assemble_filtered.argtypes = [P, M, i64p, i64p, i64p, dp, dp]
        L.orc_sloan.argtypes = [C.c_int64, i64p, i64p, i64p]
        L.orc_mesh_sloan.argtypes = [M, i64p]
        L.orc_elems_touching.argtypes = [M, C.c_void_p, i64p, C.c_int64]
        L.orc_elems_touching.restype = C.c_int64
        L.orc_assemble_rows.argtypes = [P, M, i64p, i64p, i64p, dp, dp, i64p, C.c_int64, C.c_int]
        L.orc_nnz_closed_form.argtypes = [C.c_int] * 6
        L.orc_nnz_closed_form.restype = C.c_int64
        L.free = C.CDLL(None).free
        L.free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def gauss_legendre(order):
    loc = np.zeros(64)
    w = np.zeros(64)
    n = lib().orc_gauss_legendre(order, _dp(loc), _dp(w))
    return loc[:n].copy(), w[:n].copy()


def simplex_rule(etype, order):
    pts = np.zeros(3 * 4096)
    w = np.zeros(4096)
    n = lib().orc_simplex_rule(etype, order, _dp(pts), _dp(w))
    if n < 0:
        raise ValueError("rule order not available")
    return pts[:3 * n].reshape(n, 3).copy(), w[:n].copy()


def collapsed_rule(etype, n):
    pts = np.zeros(3 * 4096)
    w = np.zeros(4096)
    q = lib().orc_collapsed_rule(etype, n, _dp(pts), _dp(w))
    if q < 0:
        raise ValueError("rule size not available")
    return pts[:3 * q].reshape(q, 3).copy(), w[:q].copy()


def shape_h1(etype, ids, orders, xi):
    """Hierarchical shapes of a triangle/tetrahedron with corner ids and
    connect orders (sides after the corners); returns (phi, dphi[dim, ns])."""
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    ids = np.ascontiguousarray(ids, dtype=np.int64)
    orders = np.ascontiguousarray(orders, dtype=np.int32)
    phi = np.zeros(256)
    dphi = np.zeros(3 * 256)
    ns = lib().orc_shape_h1(etype, _ip(ids), orders.ctypes.data_as(C.POINTER(C.c_int)),
                            _dp(np.r_[xi, np.zeros(3 - len(xi))]), _dp(phi), _dp(dphi))
    if ns < 0:
        raise ValueError("bad element type or order")
    dim = {TRI: 2, TET: 3}[etype]
    return phi[:ns].copy(), dphi[: dim * ns].reshape(dim, ns).copy()


def shape(etype, p, xi):
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    phi = np.zeros(27)
    dphi = np.zeros(81)
    ns = lib().orc_shape(etype, p, _dp(np.r_[xi, np.zeros(3 - len(xi))]), _dp(phi), _dp(dphi))
    dim = {LINE: 1, QUAD: 2, HEX: 3, TRI: 2, TET: 3}[etype]
    return phi[:ns].copy(), dphi[: dim * ns].reshape(dim, ns).copy()


def contribute(material, dim, weight, phi, dphix, f=(0.0, 0.0, 0.0), kappa=1.0, E=1.0, nu=0.3,
               ek=None, ef=None):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    dphix = np.ascontiguousarray(dphix, dtype=np.float64)
    nsh = phi.shape[0]
    ne = nsh * (3 if material == ELAST3D else 1)
    if ek is None:
        ek = np.zeros((ne, ne))
    if ef is None:
        ef = np.zeros(ne)
    fa = np.ascontiguousarray(f, dtype=np.float64)
    lib().orc_contribute(material, dim, nsh, ne, weight, _dp(phi), _dp(dphix), _dp(fa), kappa,
                         E, nu, _dp(ek), _dp(ef))
    return ek, ef


class Oracle:
    """A problem instance: structured mesh + first-touch numbering on the CPU.

    ``spec`` is a :class:`neopz_amd.problems.Problem`-like object (attribute
    access) describing material, mesh sizes, coefficients, loads and BCs.
    """

    def __init__(self, spec, mesh=None):
        """With ``mesh`` (a flattened description, e.g. of
        ``neopz_amd.mesh.build_refined2d``) the oracle takes that mesh's
        elements, connects, numbering and constraints instead of building the
        structured mesh of ``spec``; ``spec`` still gives the material, loads
        and boundary conditions (``el_coef`` per element of ``mesh``)."""
        self.spec = spec
        self._keep = []
        self.dep = None
        pb = _Problem()
        pb.material, pb.dim, pb.p = spec.material, spec.dim, spec.p
        pb.nx, pb.ny, pb.nz = spec.nx, spec.ny, spec.nz
        pb.scale, pb.E, pb.nu = spec.scale, spec.E, spec.nu
        coef = spec.el_coef
        if mesh is not None:  # coefficients per element of the given mesh
            coef = mesh.get("vol_coef")
        if coef is not None:
            c = np.ascontiguousarray(coef, dtype=np.float64)
            self._keep.append(c)
            pb.el_coef = _dp(c)
        pb.force_kind, pb.force_porder = spec.force_kind, spec.force_porder
        for k in range(3):
            pb.force_const[k] = spec.force_const[k]
        if spec.bumps is not None and len(spec.bumps):
            b = np.ascontiguousarray(spec.bumps, dtype=np.float64).reshape(-1, 6)
            self._keep.append(b)
            pb.nbumps = b.shape[0]
            pb.bumps = _dp(b)
        pb.bump_width = spec.bump_width
        for f in range(6):
            pb.bc_type[f] = spec.bc_type[f]
            for k in range(3):
                pb.bc_val[f][k] = spec.bc_val[f][k]
            v1 = getattr(spec, "bc_val1", None)
            for k in range(9):
                pb.bc_val1[f][k] = v1[f][k] if v1 is not None else 0.0
        pb.simplex = int(getattr(spec, "simplex", 0))
        for k in range(3):
            pb.prestress[k] = float(getattr(spec, "prestress", (0.0, 0.0, 0.0))[k])
        if getattr(spec, "el_porder", None) is not None:
            po = np.ascontiguousarray(spec.el_porder, dtype=np.int32)
            self._keep.append(po)
            pb.el_porder = po.ctypes.data_as(C.POINTER(C.c_int))
        self.pb = pb
        self.m = _Mesh()
        if mesh is not None:
            self._from_arrays(mesh)
        elif lib().orc_mesh_build(C.byref(pb), C.byref(self.m)) != 0:
            raise RuntimeError("oracle mesh build failed")
        self.neq = self.m.neq
        self.ia = self.ja = None

    def _from_arrays(self, md):
        def a(x, dt):
            y = np.ascontiguousarray(x, dtype=dt).ravel()
            self._keep.append(y)
            return y
        ip = C.POINTER(C.c_int)
        dim, ns = md["dim"], md["nstate"]
        vt = {2: QUAD, 3: HEX}[dim] if not md.get("vol_etype") else TET
        bt = {2: LINE, 3: QUAD}[dim] if vt != TET else TRI
        xyz = a(md["xyz"], np.float64)
        vn, vc = a(md["vol_nodes"], np.int64), a(md["vol_conn"], np.int64)
        bn, bc = a(md["bc_nodes"], np.int64), a(md["bc_conn"], np.int64)
        bf = a(md["bc_id"], np.int32)
        pos, size = a(md["blk_pos"], np.int64), a(md["blk_size"], np.int32)
        order = a(md["conn_order"], np.int32)
        nvol, nbc = len(md["vol_nodes"]), len(md["bc_nodes"])
        if lib().orc_mesh_from_arrays(dim, md["porder"], ns, vt, bt, len(xyz) // 3, _dp(xyz), nvol,
                                      _ip(vn), _ip(vc), nbc, _ip(bn), _ip(bc), bf.ctypes.data_as(ip),
                                      md["nconn"], _ip(pos), size.ctypes.data_as(ip),
                                      order.ctypes.data_as(ip), md["neq"], C.byref(self.m)) != 0:
            raise RuntimeError("oracle mesh from arrays failed")
        if md.get("dep_ptr") is not None:
            self.dep = (a(md["dep_ptr"], np.int64), a(md["dep_conn"], np.int64),
                        a(md["dep_coef"], np.float64))

    def __del__(self):
        try:
            lib().orc_mesh_free(C.byref(self.m))
        except Exception:
            pass

    # --- numbering ---------------------------------------------------------
    def permute(self, perm):
        """Renumbers the connects: new sequence number perm[old] (as after
        TPZCompMesh::Permute); block positions follow the new order."""
        pr = np.ascontiguousarray(perm, dtype=np.int64)
        if len(pr) != self.m.nconn or lib().orc_mesh_permute(C.byref(self.m), _ip(pr)) != 0:
            raise ValueError("not a permutation of the connects")
        self.ia = self.ja = None

    def connects(self):
        m = self.m
        econn = np.ctypeslib.as_array(m.econn, shape=(27 * m.nel,)).reshape(m.nel, 27).copy()
        blk_pos = np.ctypeslib.as_array(m.blk_pos, shape=(m.nconn,)).copy()
        blk_size = np.ctypeslib.as_array(m.blk_size, shape=(m.nconn,)).copy()
        return econn, blk_pos, blk_size

    def element(self, iel):
        """Element iel: type, corner node ids, corner coordinates and the
        orders of its side connects (sides after the corners)."""
        m = self.m
        t = m.etype[iel]
        nc = {LINE: 2, QUAD: 4, HEX: 8, TRI: 3, TET: 4}[t]
        ns = {LINE: 3, QUAD: 9, HEX: 27, TRI: 7, TET: 15}[t]
        ids = np.array([m.enodes[8 * iel + k] for k in range(nc)], dtype=np.int64)
        xyz = np.array([[m.xyz[3 * i + d] for d in range(3)] for i in ids])
        orders = np.array([m.conn_order[m.econn[27 * iel + s]] for s in range(nc, ns)],
                          dtype=np.int32)
        return {"etype": t, "ids": ids, "xyz": xyz, "orders": orders}

    def pattern(self):
        if self.ia is None:
            ia = C.POINTER(C.c_int64)()
            ja = C.POINTER(C.c_int64)()
            if self.dep is not None:
                nnz = lib().orc_setup_pattern_constrained(C.byref(self.m), _ip(self.dep[0]),
                                                          _ip(self.dep[1]), C.byref(ia), C.byref(ja))
                if nnz < 0:
                    raise RuntimeError("constrained connects must come after the others")
            else:
                nnz = lib().orc_setup_pattern(C.byref(self.m), C.byref(ia), C.byref(ja))
            self.ia = np.ctypeslib.as_array(ia, shape=(self.neq + 1,)).copy()
            self.ja = np.ctypeslib.as_array(ja, shape=(max(nnz, 1),))[:nnz].copy()
            lib().free(ia)
            lib().free(ja)
        return self.ia, self.ja

    def calcstiff(self, iel):
        ek = np.zeros(MAXNE * MAXNE)
        ef = np.zeros(MAXNE)
        dest = np.zeros(MAXNE, dtype=np.int64)
        ne = lib().orc_calcstiff(C.byref(self.pb), C.byref(self.m), iel, _dp(ek), _dp(ef),
                                 _ip(dest))
        if ne < 0:
            raise RuntimeError("oracle CalcStiff failed (element too large or order > 8)")
        return ek[: ne * ne].reshape(ne, ne).copy(), ef[:ne].copy(), dest[:ne].copy()

    def pattern_full(self):
        """TPZSpStructMatrix / TPZFYsmpMatrix pattern (both triangles)."""
        ia = C.POINTER(C.c_int64)()
        ja = C.POINTER(C.c_int64)()
        nnz = lib().orc_setup_pattern_full(C.byref(self.m), C.byref(ia), C.byref(ja))
        out_ia = np.ctypeslib.as_array(ia, shape=(self.neq + 1,)).copy()
        out_ja = np.ctypeslib.as_array(ja, shape=(max(nnz, 1),))[:nnz].copy()
        lib().free(ia)
        lib().free(ja)
        return out_ia, out_ja

    def assemble_full(self):
        """Assembly into the TPZFYsmpMatrix pattern (TPZFYsmpMatrix::AddKel)."""
        ia, ja = self.pattern_full()
        a = np.zeros(len(ja))
        rhs = np.zeros(self.neq)
        if lib().orc_assemble_full(C.byref(self.pb), C.byref(self.m), _ip(ia), _ip(ja), _dp(a),
                                   _dp(rhs)) != 0:
            raise RuntimeError("oracle full assembly failed")
        return ia, ja, a, rhs

    def assemble(self):
        ia, ja = self.pattern()
        a = np.zeros(len(ja))
        rhs = np.zeros(self.neq)
        if self.dep is not None:
            if lib().orc_assemble_constrained(C.byref(self.pb), C.byref(self.m), _ip(self.dep[0]),
                                              _ip(self.dep[1]), _dp(self.dep[2]), _ip(ia), _ip(ja),
                                              _dp(a), _dp(rhs)) != 0:
                raise RuntimeError("oracle constrained assembly failed")
        elif lib().orc_assemble(C.byref(self.pb), C.byref(self.m), _ip(ia), _ip(ja), _dp(a),
                                _dp(rhs)) != 0:
            raise RuntimeError("oracle assembly failed (boundary condition type outside the "
                               "material's ContributeBC, or element too large)")
        return a, rhs

    def filtered(self, active):
        """Pattern and assembly with an active TPZEquationFilter over the
        equations ``active`` (TPZEquationFilter::SetActiveEquations: sorted,
        duplicates removed): condensed (ia, ja, a) and the load in the mesh
        numbering (inactive entries 0)."""
        act = np.unique(np.asarray(active, dtype=np.int64))
        dest = np.full(self.neq, -1, dtype=np.int64)
        dest[act] = np.arange(len(act))
        ia = C.POINTER(C.c_int64)()
        ja = C.POINTER(C.c_int64)()
        nnz = lib().orc_setup_pattern_filtered(C.byref(self.m), _ip(dest), C.byref(ia), C.byref(ja))
        if nnz < 0:
            raise RuntimeError("filtered pattern failed")
        out_ia = np.ctypeslib.as_array(ia, shape=(len(act) + 1,)).copy()
        out_ja = np.ctypeslib.as_array(ja, shape=(max(nnz, 1),))[:nnz].copy()
        lib().free(ia)
        lib().free(ja)
        a = np.zeros(max(nnz, 1))
        rhs = np.zeros(self.neq)
        if lib().orc_assemble_filtered(C.byref(self.pb), C.byref(self.m), _ip(dest), _ip(out_ia),
                                       _ip(out_ja), _dp(a), _dp(rhs)) != 0:
            raise RuntimeError("filtered assembly failed")
        return out_ia, out_ja, a[:nnz].copy(), rhs

    def sloan(self):
        """TPZAnalysis::OptimizeBandwidth's permutation (TPZSloanRenumbering) of
        the current connect numbering: perm[seq] = new sequence number."""
        perm = np.zeros(self.m.nconn, dtype=np.int64)
        if lib().orc_mesh_sloan(C.byref(self.m), _ip(perm)) != 0:
            raise RuntimeError("oracle Sloan renumbering failed (graph not connected)")
        return perm

    # --- post-processing ---------------------------------------------------
    def error_points(self, porder):
        """Physical coordinates [npts, 3] of the error rule's points of every
        volume element (element order) for an exact solution of order porder."""
        n = lib().orc_error_points(C.byref(self.pb), C.byref(self.m), porder, None)
        if n < 0:
            raise RuntimeError("oracle error rule failed")
        xq = np.zeros((n, 3))
        lib().orc_error_points(C.byref(self.pb), C.byref(self.m), porder, _dp(xq))
        return xq

    def evaluate_error(self, u, kind=EXACT_POINTS, porder=0, center=(0.0, 0.0, 0.0), constant=0.0,
                       uq=None, duq=None):
        """TPZCompMesh::EvaluateError of solution u: (per-element errors
        [nvol, nerr], mesh totals [nerr]).  kind EXACT_POINTS takes the exact
        values uq [npts, nstate] and derivatives duq [npts, dim, nstate] at the
        points of error_points(porder)."""
        ex = _Exact()
        ex.kind, ex.porder = kind, porder
        for d in range(3):
            ex.center[d] = center[d]
        ex.constant = constant
        keep = []
        if kind == EXACT_POINTS:
            uq = np.ascontiguousarray(uq, dtype=np.float64)
            duq = np.ascontiguousarray(duq, dtype=np.float64)
            keep += [uq, duq]
            ex.u, ex.du = _dp(uq), _dp(duq)
        u = np.ascontiguousarray(u, dtype=np.float64)
        el = np.zeros((self.m.nvol, 6))
        tot = np.zeros(6)
        nerr = lib().orc_evaluate_error(C.byref(self.pb), C.byref(self.m), _dp(u), C.byref(ex),
                                        _dp(el), _dp(tot))
        if nerr < 0:
            raise RuntimeError("oracle EvaluateError failed")
        el = el.reshape(-1)[: self.m.nvol * nerr].reshape(self.m.nvol, nerr).copy()
        return el, tot[:nerr].copy()

    def assemble_colored(self, nthreads):
        """Coloured multi-threaded assembly (TPZStructMatrixOT's element
        loop): the same system up to the summation order."""
        ia, ja = self.pattern()
        a = np.zeros(len(ja))
        rhs = np.zeros(self.neq)
        if lib().orc_assemble_colored(C.byref(self.pb), C.byref(self.m), _ip(ia), _ip(ja), _dp(a),
                                      _dp(rhs), nthreads) != 0:
            raise RuntimeError("oracle coloured assembly failed")
        return a, rhs

    def assemble_range(self, e0, e1, a=None, rhs=None):
        ia, ja = self.pattern()
        if a is None:
            a = np.zeros(len(ja))
        if rhs is None:
            rhs = np.zeros(self.neq)
        lib().orc_assemble_range(C.byref(self.pb), C.byref(self.m), _ip(ia), _ip(ja), _dp(a),
                                 _dp(rhs), e0, e1)
        return a, rhs


def assemble_rows(o, rows, ia_s, ja_s, nthreads=8):
    """The oracle's assembled upper-CSR entries and load of the given rows
    (ascending, unique) over the pattern ia_s/ja_s of those rows: only the
    elements that touch them are integrated, and every entry receives the
    same terms in the same (element) order as in :meth:`Oracle.assemble`."""
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    mask = np.zeros(o.neq, dtype=np.uint8)
    mask[rows] = 1
    cap = o.m.nel
    elems = np.zeros(max(cap, 1), dtype=np.int64)
    ne = lib().orc_elems_touching(C.byref(o.m), mask.ctypes.data_as(C.c_void_p), _ip(elems), cap)
    if ne < 0:
        raise RuntimeError("element list overflow")
    rowmap = np.full(o.neq, -1, dtype=np.int64)
    rowmap[rows] = np.arange(len(rows))
    ia_s = np.ascontiguousarray(ia_s, dtype=np.int64)
    ja_s = np.ascontiguousarray(ja_s, dtype=np.int64)
    a_s = np.zeros(max(len(ja_s), 1))
    rhs_s = np.zeros(len(rows))
    rc = lib().orc_assemble_rows(C.byref(o.pb), C.byref(o.m), _ip(rowmap), _ip(ia_s), _ip(ja_s),
                                 _dp(a_s), _dp(rhs_s), _ip(elems), ne, nthreads)
    if rc == -1:
        raise ValueError("an element entry's column is missing from the given row pattern")
    if rc != 0:
        raise RuntimeError("oracle CalcStiff failed")
    return a_s[: len(ja_s)], rhs_s, int(ne)


def addkel(ia, ja, a, ek, dest):
    ek = np.ascontiguousarray(ek, dtype=np.float64)
    dest = np.ascontiguousarray(dest, dtype=np.int64)
    lib().orc_addkel(_ip(ia), _ip(ja), _dp(a), ek.shape[0], _dp(ek), _ip(dest))


def multadd(ia, ja, a, x, y=None, alpha=1.0, beta=0.0):
    n = len(ia) - 1
    x = np.ascontiguousarray(x, dtype=np.float64)
    z = np.zeros(n)
    yy = np.ascontiguousarray(y if y is not None else np.zeros(n), dtype=np.float64)
    lib().orc_multadd(n, _ip(ia), _ip(ja), _dp(a), _dp(x), _dp(yy), _dp(z), alpha, beta)
    return z


def pcg(ia, ja, a, b, maxit=100000, tol=1e-10, wasted_spmv=True, history=False):
    n = len(ia) - 1
    x = np.zeros(n)
    it = C.c_int64(maxit)
    t = C.c_double(tol)
    hist = np.zeros(maxit) if history else None
    b = np.ascontiguousarray(b, dtype=np.float64)
    ret = lib().orc_pcg(n, _ip(ia), _ip(ja), _dp(a), _dp(b), _dp(x), C.byref(it), C.byref(t),
                        1 if wasted_spmv else 0, _dp(hist) if history else None)
    out = {"x": x, "iters": it.value, "tol": t.value, "converged": ret == 0}
    if history:
        out["history"] = hist[: it.value].copy()
    return out


def pcg_block(ia, ja, a, b, bsize, maxit=100000, tol=1e-10, history=False):
    """CG with the EBlockJacobi(overlap) preconditioner: LU of the consecutive
    diagonal blocks of bsize[k] equations (orc_pcg_block)."""
    n = len(ia) - 1
    x = np.zeros(n)
    it = C.c_int64(maxit)
    t = C.c_double(tol)
    hist = np.zeros(maxit) if history else None
    b = np.ascontiguousarray(b, dtype=np.float64)
    bs = np.ascontiguousarray(bsize, dtype=np.int32)
    ret = lib().orc_pcg_block(n, _ip(ia), _ip(ja), _dp(a), _dp(b), len(bs),
                              bs.ctypes.data_as(C.POINTER(C.c_int32)), _dp(x), C.byref(it), C.byref(t),
                              _dp(hist) if history else None)
    if ret < 0:
        raise ValueError("pcg_block: block sizes do not sum to the number of equations")
    out = {"x": x, "iters": it.value, "tol": t.value, "converged": ret == 0}
    if history:
        out["history"] = hist[: it.value].copy()
    return out


def sloan(xadj, adj):
    """TPZSloanRenumbering::Resequence on a node graph (sorted adjacency, no
    self loops): perm[node] = new number."""
    xadj = np.ascontiguousarray(xadj, dtype=np.int64)
    adj = np.ascontiguousarray(adj, dtype=np.int64)
    n = len(xadj) - 1
    perm = np.zeros(max(n, 1), dtype=np.int64)
    if lib().orc_sloan(n, _ip(xadj), _ip(adj), _ip(perm)) != 0:
        raise RuntimeError("graph not connected")
    return perm[:n]


def nnz_closed_form(dim, p, nstate, nx, ny, nz):
    return lib().orc_nnz_closed_form(dim, p, nstate, nx, ny, nz)


def upper_to_dense(ia, ja, a):
    n = len(ia) - 1
    K = np.zeros((n, n))
    for r in range(n):
        for k in range(ia[r], ia[r + 1]):
            K[r, ja[k]] = a[k]
            K[ja[k], r] = a[k]
    return K
