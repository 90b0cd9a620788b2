"""Structured computational meshes in NeoPZ's layout (caller side of the C ABI).

In drop-in use the flattened mesh comes from ``TPZCompMesh`` (connect
sequence numbers, ``Block()`` positions, element connect lists).  Here it is
built directly for the unit square/cube, reproducing what
``TPZCompMesh::AutoBuild`` + ``TPZLinearAnalysis(cmesh, false)`` number:
connects are created in element order and, inside an element, in side order
(corners, edges, faces, interior) the first time a side is met
(``Mesh/TPZCompElH1.cpp:17-24``, ``Mesh/pzintel.cpp:652-703``,
``Mesh/pzcmesh.h:703-715``).  Volume elements come first (x fastest), then the
boundary elements face by face; boundary elements never create connects.

The numbering is computed in closed form: every side has a structured key
(node id, edge or face id by direction and lower corner, element id), and the
sequence number of a key is the rank of its first occurrence in the
element-major, side-minor walk.
"""
import numpy as np

from . import problems as P

QUAD, HEX, LINE = 2, 3, 1


def _first_touch(keys):
    """seq[k] for every entry of the walk = rank of the key's first occurrence."""
    uniq, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return rank[inv], len(uniq)


def build(pb):
    """Flattened mesh description of a :class:`neopz_amd.problems.Problem`."""
    if pb.dim == 2:
        return _build2d(pb)
    if pb.dim == 3 and getattr(pb, "simplex", 0):
        return _build3d_tet(pb)
    if pb.dim == 3:
        return _build3d(pb)
    raise ValueError("structured builder supports dim 2 and 3")


def _vol_rule(pb, p):
    """TPZMatSingleSpaceT::IntegrationRuleOrder (Material/TPZMatSingleSpace.cpp:
    60-72) of an element whose highest connect order is p; TPZElasticity3D
    uses 2p (Material/Elasticity/TPZElasticity3D.h:82-85)."""
    ff = pb.force_porder if pb.force_kind != P.FORCE_NONE else 0
    if pb.material == P.ELAST3D:
        return 2 * p
    return np.where(p < ff, ff + p, 2 * p)


# side kinds of a connect: corner, edge, quadrilateral face, hexahedron
# interior, triangle face, tetrahedron interior
K_CORNER, K_EDGE, K_QFACE, K_HEX, K_TFACE, K_TET = range(6)


def _nshape(kind, order):
    """TPZShape*::NConnectShapeF (Shape/pzshapequad.cpp:468-482,
    Shape/pzshapecube.cpp:573-584, Shape/pzshapetriang.cpp:452-470,
    Shape/pzshapetetra.cpp:451-470) per connect."""
    q = np.maximum(order - 1, 0)
    return np.select(
        [kind == K_CORNER, kind == K_EDGE, kind == K_QFACE, kind == K_HEX, kind == K_TFACE],
        [np.ones_like(q), q, q * q, q * q * q, np.maximum((order - 1) * (order - 2) // 2, 0)],
        np.maximum((order - 1) * (order - 2) * (order - 3) // 6, 0)).astype(np.int64)


def _common(pb, econn_vol, econn_bc, nconn, corner_flags, xyz, vol_nodes, bc_nodes, bc_id,
            side_kind=None, vol_etype=0):
    ns = pb.nstate
    # connect orders: corners 1; other sides the minimum preferred order of the
    # elements sharing them (TPZInterpolatedElement::IdentifySideOrder /
    # ComputeSideOrder, Mesh/pzintel.cpp:214-290, 1370-1390).  Boundary
    # elements prefer the default order p, which never lowers the minimum.
    pe = (np.full(econn_vol.shape[0], pb.p, dtype=np.int64) if pb.el_porder is None
          else np.asarray(pb.el_porder, dtype=np.int64))
    if pe.shape[0] != econn_vol.shape[0] or pe.min() < 1 or pe.max() > pb.p:
        raise ValueError("el_porder must hold one order in 1..p per volume element")
    order = np.full(nconn, pb.p, dtype=np.int64)
    np.minimum.at(order, econn_vol.ravel(), np.repeat(pe, econn_vol.shape[1]))
    order[corner_flags] = 1
    if side_kind is None:
        # quadrilaterals / hexahedra, p <= 2: corners 1, other sides 1 from order 2 on
        if pb.p > 2:
            raise ValueError("quadrilateral/hexahedral meshes support p <= 2")
        nshape = np.where(corner_flags, 1, (order >= 2).astype(np.int64))
    else:
        nshape = _nshape(side_kind, order)
    blk_size = (nshape * ns).astype(np.int32)
    # element integration rules from the highest connect order (MaxOrder,
    # Mesh/pzinterpolationspace.cpp:53-83)
    vol_rule = _vol_rule(pb, order[econn_vol].max(axis=1)).astype(np.int32)
    bc_rule = (2 * order[econn_bc].max(axis=1)).astype(np.int32) if len(econn_bc) else \
        np.zeros(0, np.int32)
    blk_pos = np.zeros(nconn, dtype=np.int64)
    if nconn > 1:
        blk_pos[1:] = np.cumsum(blk_size[:-1].astype(np.int64))
    neq = int(blk_size.astype(np.int64).sum())
    ff = pb.force_porder if pb.force_kind != P.FORCE_NONE else 0
    if pb.material == P.ELAST3D:
        rule_vol = 2 * pb.p
    else:
        rule_vol = ff + pb.p if pb.p < ff else 2 * pb.p
    big = 1.0e12 if pb.material == P.ELAST3D else 10.0 ** 17 * 2 / 3
    bc_ids_used = [f for f in range(6) if pb.bc_type[f] >= 0]
    hp = pb.el_porder is not None
    return dict(
        vol_rule=vol_rule if hp else None, bc_rule=bc_rule if hp else None,
        dim=pb.dim, material=pb.material, nstate=ns, porder=pb.p, xyz=xyz,
        vol_nodes=vol_nodes, vol_conn=econn_vol,
        vol_coef=None if pb.el_coef is None else np.ascontiguousarray(pb.el_coef, dtype=np.float64),
        bc_nodes=bc_nodes, bc_conn=econn_bc, bc_id=bc_id,
        bc_type=np.array([pb.bc_type[f] for f in range(6)], dtype=np.int32),
        bc_val=np.array([pb.bc_val[f] for f in range(6)], dtype=np.float64).ravel(),
        prestress=np.array(getattr(pb, "prestress", (0.0, 0.0, 0.0)), dtype=np.float64),
        bc_val1=np.array([getattr(pb, "bc_val1", [[0.0] * 9] * 6)[f] for f in range(6)],
                         dtype=np.float64).ravel(),
        nconn=nconn, blk_pos=blk_pos, blk_size=blk_size, neq=neq,
        scale=pb.scale, E=pb.E, nu=pb.nu, rule_order_vol=rule_vol, rule_order_bc=2 * pb.p,
        force_kind=pb.force_kind, force_const=np.array(pb.force_const, dtype=np.float64),
        bumps=None if pb.bumps is None else np.ascontiguousarray(pb.bumps, dtype=np.float64).reshape(-1, 6),
        bump_width=pb.bump_width, big_number=big, bc_faces=bc_ids_used, vol_etype=vol_etype,
    )


def _build2d(pb):
    nx, ny = pb.nx, pb.ny
    NX = nx + 1
    nv = NX * (ny + 1)
    nxe = nx * (ny + 1)
    nye = NX * ny
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ix = ix.ravel().astype(np.int64)
    iy = iy.ravel().astype(np.int64)
    node = lambda i, j: i + NX * j
    xe = lambda i, j: nv + i + nx * j
    ye = lambda i, j: nv + nxe + i + NX * j
    face = lambda e: nv + nxe + nye + e
    e = np.arange(nx * ny, dtype=np.int64)
    vol_nodes = np.stack([node(ix, iy), node(ix + 1, iy), node(ix + 1, iy + 1), node(ix, iy + 1)], 1)
    keys = np.stack([node(ix, iy), node(ix + 1, iy), node(ix + 1, iy + 1), node(ix, iy + 1),
                     xe(ix, iy), ye(ix + 1, iy), xe(ix, iy + 1), ye(ix, iy), face(e)], 1)
    ktot = nv + nxe + nye + nx * ny
    seq_flat, nconn = _first_touch(keys.ravel())
    key2seq = np.full(ktot, -1, dtype=np.int64)
    key2seq[keys.ravel()] = seq_flat
    econn = seq_flat.reshape(-1, 9)
    corner = np.zeros(nconn, dtype=bool)
    corner[key2seq[:nv][key2seq[:nv] >= 0]] = True
    xs = np.arange(NX) / nx
    ys = np.arange(ny + 1) / ny
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    xyz = np.stack([X.ravel(), Y.ravel(), np.zeros(nv)], 1).ravel()
    bn, bk, bid = [], [], []
    for f in range(4):
        if pb.bc_type[f] < 0:
            continue
        ax, side = f // 2, f % 2
        if ax == 0:
            c = nx if side else 0
            a = np.arange(ny, dtype=np.int64)
            bn.append(np.stack([node(c, a), node(c, a + 1)], 1))
            bk.append(np.stack([node(c, a), node(c, a + 1), ye(c, a)], 1))
        else:
            c = ny if side else 0
            a = np.arange(nx, dtype=np.int64)
            bn.append(np.stack([node(a, c), node(a + 1, c)], 1))
            bk.append(np.stack([node(a, c), node(a + 1, c), xe(a, c)], 1))
        bid.append(np.full(len(a), f, dtype=np.int32))
    if bn:
        bc_nodes = np.concatenate(bn).astype(np.int64)
        bc_conn = key2seq[np.concatenate(bk)]
        bc_id = np.concatenate(bid)
    else:
        bc_nodes = np.zeros((0, 2), np.int64)
        bc_conn = np.zeros((0, 3), np.int64)
        bc_id = np.zeros(0, np.int32)
    return _common(pb, econn.astype(np.int64), bc_conn.astype(np.int64), nconn, corner, xyz,
                   vol_nodes.astype(np.int64), bc_nodes, bc_id)


def _build3d(pb):
    nx, ny, nz = pb.nx, pb.ny, pb.nz
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    nv = NX * NY * NZ
    n_xe, n_ye, n_ze = nx * NY * NZ, NX * ny * NZ, NX * NY * nz
    n_xf, n_yf, n_zf = NX * ny * nz, nx * NY * nz, nx * ny * NZ
    o_xe = nv
    o_ye = o_xe + n_xe
    o_ze = o_ye + n_ye
    o_xf = o_ze + n_ze
    o_yf = o_xf + n_xf
    o_zf = o_yf + n_yf
    o_v = o_zf + n_zf
    ktot = o_v + nx * ny * nz
    node = lambda i, j, k: i + NX * (j + NY * k)
    xe = lambda i, j, k: o_xe + i + nx * (j + NY * k)
    ye = lambda i, j, k: o_ye + i + NX * (j + ny * k)
    ze = lambda i, j, k: o_ze + i + NX * (j + NY * k)
    xf = lambda i, j, k: o_xf + i + NX * (j + ny * k)
    yf = lambda i, j, k: o_yf + i + nx * (j + NY * k)
    zf = lambda i, j, k: o_zf + i + nx * (j + ny * k)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ix = ix.ravel().astype(np.int64)
    iy = iy.ravel().astype(np.int64)
    iz = iz.ravel().astype(np.int64)
    e = np.arange(nx * ny * nz, dtype=np.int64)
    c = [node(ix, iy, iz), node(ix + 1, iy, iz), node(ix + 1, iy + 1, iz), node(ix, iy + 1, iz),
         node(ix, iy, iz + 1), node(ix + 1, iy, iz + 1), node(ix + 1, iy + 1, iz + 1),
         node(ix, iy + 1, iz + 1)]
    vol_nodes = np.stack(c, 1)
    keys = np.stack(c + [
        xe(ix, iy, iz), ye(ix + 1, iy, iz), xe(ix, iy + 1, iz), ye(ix, iy, iz),
        ze(ix, iy, iz), ze(ix + 1, iy, iz), ze(ix + 1, iy + 1, iz), ze(ix, iy + 1, iz),
        xe(ix, iy, iz + 1), ye(ix + 1, iy, iz + 1), xe(ix, iy + 1, iz + 1), ye(ix, iy, iz + 1),
        zf(ix, iy, iz), yf(ix, iy, iz), xf(ix + 1, iy, iz), yf(ix, iy + 1, iz), xf(ix, iy, iz),
        zf(ix, iy, iz + 1), o_v + e], 1)
    seq_flat, nconn = _first_touch(keys.ravel())
    key2seq = np.full(ktot, -1, dtype=np.int64)
    key2seq[keys.ravel()] = seq_flat
    econn = seq_flat.reshape(-1, 27)
    corner = np.zeros(nconn, dtype=bool)
    corner[key2seq[:nv][key2seq[:nv] >= 0]] = True
    gz, gy, gx = np.meshgrid(np.arange(NZ) / nz, np.arange(NY) / ny, np.arange(NX) / nx, indexing="ij")
    xyz = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1).ravel()
    bn, bk, bid = [], [], []
    for f in range(6):
        if pb.bc_type[f] < 0:
            continue
        ax, side = f // 2, f % 2
        nu_ = ny if ax == 0 else nx
        nv_ = ny if ax == 2 else nz
        cc = (nx if ax == 0 else ny if ax == 1 else nz) if side else 0
        b, a = np.meshgrid(np.arange(nv_), np.arange(nu_), indexing="ij")
        a = a.ravel().astype(np.int64)
        b = b.ravel().astype(np.int64)
        if ax == 0:
            nodes = [node(cc, a, b), node(cc, a + 1, b), node(cc, a + 1, b + 1), node(cc, a, b + 1)]
            sides = [ye(cc, a, b), ze(cc, a + 1, b), ye(cc, a, b + 1), ze(cc, a, b), xf(cc, a, b)]
        elif ax == 1:
            nodes = [node(a, cc, b), node(a + 1, cc, b), node(a + 1, cc, b + 1), node(a, cc, b + 1)]
            sides = [xe(a, cc, b), ze(a + 1, cc, b), xe(a, cc, b + 1), ze(a, cc, b), yf(a, cc, b)]
        else:
            nodes = [node(a, b, cc), node(a + 1, b, cc), node(a + 1, b + 1, cc), node(a, b + 1, cc)]
            sides = [xe(a, b, cc), ye(a + 1, b, cc), xe(a, b + 1, cc), ye(a, b, cc), zf(a, b, cc)]
        bn.append(np.stack(nodes, 1))
        bk.append(np.stack(nodes + sides, 1))
        bid.append(np.full(len(a), f, dtype=np.int32))
    if bn:
        bc_nodes = np.concatenate(bn).astype(np.int64)
        bc_conn = key2seq[np.concatenate(bk)]
        bc_id = np.concatenate(bid)
    else:
        bc_nodes = np.zeros((0, 4), np.int64)
        bc_conn = np.zeros((0, 9), np.int64)
        bc_id = np.zeros(0, np.int32)
    return _common(pb, econn.astype(np.int64), bc_conn.astype(np.int64), nconn, corner, xyz,
                   vol_nodes.astype(np.int64), bc_nodes, bc_id)


def permute(m, perm):
    """The mesh description after a connect renumbering (new sequence number
    perm[old], as TPZCompMesh::Permute leaves it after a bandwidth
    optimisation): element connect lists relabelled, Block() sizes and
    positions in the new sequence order."""
    perm = np.asarray(perm, dtype=np.int64)
    n = m["nconn"]
    if perm.shape != (n,) or not np.array_equal(np.sort(perm), np.arange(n)):
        raise ValueError("not a permutation of the connects")
    out = dict(m)
    out["vol_conn"] = perm[m["vol_conn"]]
    out["bc_conn"] = perm[m["bc_conn"]]
    size = np.empty(n, dtype=np.int32)
    size[perm] = m["blk_size"]
    out["blk_size"] = size
    pos = np.zeros(n, dtype=np.int64)
    pos[1:] = np.cumsum(size[:-1].astype(np.int64))
    out["blk_pos"] = pos
    return out


# Kuhn split of every cube into 6 tetrahedra around its diagonal 0 -> 6, one
# per ordering of the axes (xyz, xzy, yxz, yzx, zxy, zyx), cube corners in the
# hexahedron order; sides of a tetrahedron: 4 corners, 6 edges, 4 faces, the
# interior (Topology/tpztetrahedron.h:280-283).
_KUHN = np.array([[0, 1, 2, 6], [0, 1, 5, 6], [0, 3, 2, 6], [0, 3, 7, 6], [0, 4, 5, 6], [0, 4, 7, 6]])
_TET_SIDES = [(0,), (1,), (2,), (3,), (0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3),
              (0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3), (0, 1, 2, 3)]
_TRI_SIDES = [(0,), (1,), (2,), (0, 1), (1, 2), (2, 0), (0, 1, 2)]


def _side_keys(nodes, sides, interior_base=None):
    """(n_elem, n_sides, 4) keys: sorted corner nodes of every side, padded
    with -1; element interiors get (-2, element, -1, -1)."""
    ne = nodes.shape[0]
    keys = np.full((ne, len(sides), 4), -1, dtype=np.int64)
    for s, loc in enumerate(sides):
        if interior_base is not None and len(loc) == 4:
            keys[:, s, 0] = -2
            keys[:, s, 1] = interior_base + np.arange(ne)
            continue
        keys[:, s, :len(loc)] = np.sort(nodes[:, list(loc)], axis=1)
    return keys


def _build3d_tet(pb):
    nx, ny, nz = pb.nx, pb.ny, pb.nz
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    node = lambda i, j, k: i + NX * (j + NY * k)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ix, iy, iz = (a.ravel().astype(np.int64) for a in (ix, iy, iz))
    cube = np.stack([node(ix, iy, iz), node(ix + 1, iy, iz), node(ix + 1, iy + 1, iz),
                     node(ix, iy + 1, iz), node(ix, iy, iz + 1), node(ix + 1, iy, iz + 1),
                     node(ix + 1, iy + 1, iz + 1), node(ix, iy + 1, iz + 1)], 1)
    tets = cube[:, _KUHN].reshape(-1, 4)                      # cubes x fastest, 6 per cube
    gz, gy, gx = np.meshgrid(np.arange(NZ) / nz, np.arange(NY) / ny, np.arange(NX) / nx, indexing="ij")
    xyz = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    # boundary triangles: faces 0..5, tetrahedra in order, their faces in side order
    nid = np.stack([np.arange(NX * NY * NZ) % NX, (np.arange(NX * NY * NZ) // NX) % NY,
                    np.arange(NX * NY * NZ) // (NX * NY)], 1)
    bn, bid = [], []
    faces = np.array(_TET_SIDES[10:14])
    for f in range(6):
        if pb.bc_type[f] < 0:
            continue
        ax, side = f // 2, f % 2
        plane = (nx, ny, nz)[ax] if side else 0
        fn = tets[:, faces]                                     # (ntet, 4, 3)
        on = (nid[fn][..., ax] == plane).all(axis=2)            # (ntet, 4)
        t_idx, f_idx = np.nonzero(on)                           # row-major: tet order, face order
        bn.append(fn[t_idx, f_idx])
        bid.append(np.full(len(t_idx), f, dtype=np.int32))
    bc_nodes = np.concatenate(bn).astype(np.int64) if bn else np.zeros((0, 3), np.int64)
    bc_id = np.concatenate(bid) if bid else np.zeros(0, np.int32)
    # first-touch numbering over the side keys (volume elements, then boundary)
    kv = _side_keys(tets, _TET_SIDES, interior_base=0)
    kb = _side_keys(bc_nodes, _TRI_SIDES)
    allk = np.concatenate([kv.reshape(-1, 4), kb.reshape(-1, 4)])
    uniq, first, inv = np.unique(allk, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    seq = rank[inv.ravel()]
    nconn = len(uniq)
    nv = kv.shape[0] * kv.shape[1]
    econn = seq[:nv].reshape(kv.shape[0], kv.shape[1])
    bc_conn = seq[nv:].reshape(kb.shape[0], kb.shape[1]) if len(kb) else np.zeros((0, 7), np.int64)
    corner = np.zeros(nconn, dtype=bool)
    corner[econn[:, :4].ravel()] = True
    kind = np.empty(nconn, dtype=np.int64)
    kind[econn[:, :4].ravel()] = K_CORNER
    kind[econn[:, 4:10].ravel()] = K_EDGE
    kind[econn[:, 10:14].ravel()] = K_TFACE
    kind[econn[:, 14].ravel()] = K_TET
    return _common(pb, econn.astype(np.int64), bc_conn.astype(np.int64), nconn, corner,
                   xyz.ravel(), tets.astype(np.int64), bc_nodes, bc_id, side_kind=kind,
                   vol_etype=5)


# Flat little-endian file of a mesh description, read by the C++ host mirror
# (neopz_amd/host/pzhip_run.cpp).  Fields in this fixed order; arrays are
# preceded by their element count (int64).
_FILE_MAGIC = b"PZGMESH5"


def save(m, path):
    def i32(f, v):
        f.write(np.int32(v).tobytes())

    def i64(f, v):
        f.write(np.int64(v).tobytes())

    def f64(f, v):
        f.write(np.float64(v).tobytes())

    def arr(f, x, dt):
        a = np.zeros(0, dt) if x is None else np.ascontiguousarray(x, dtype=dt).ravel()
        i64(f, a.size)
        f.write(a.tobytes())

    with open(path, "wb") as f:
        f.write(_FILE_MAGIC)
        for k in ("dim", "material", "nstate", "porder"):
            i32(f, m[k])
        arr(f, m["xyz"], np.float64)
        arr(f, m["vol_nodes"], np.int64)
        arr(f, m["vol_conn"], np.int64)
        arr(f, m["vol_coef"], np.float64)
        arr(f, m["bc_nodes"], np.int64)
        arr(f, m["bc_conn"], np.int64)
        arr(f, m["bc_id"], np.int32)
        arr(f, m["bc_type"], np.int32)
        arr(f, m["bc_val"], np.float64)
        arr(f, m["blk_pos"], np.int64)
        arr(f, m["blk_size"], np.int32)
        i64(f, m["neq"])
        for k in ("scale", "E", "nu"):
            f64(f, m[k])
        for k in ("rule_order_vol", "rule_order_bc", "force_kind"):
            i32(f, m[k])
        arr(f, m["force_const"], np.float64)
        arr(f, m["bumps"], np.float64)
        for k in ("bump_width", "big_number"):
            f64(f, m[k])
        arr(f, m.get("vol_rule"), np.int32)
        arr(f, m.get("bc_rule"), np.int32)
        i32(f, m.get("vol_etype", 0))
        arr(f, m.get("dep_ptr"), np.int64)
        arr(f, m.get("dep_conn"), np.int64)
        arr(f, m.get("dep_coef"), np.float64)
        arr(f, m.get("bc_val1"), np.float64)


def build_refined2d(pb, refine):
    """A quadrilateral mesh of the unit square with hanging sides: the coarse
    nx x ny elements flagged in ``refine`` are divided into four (children in
    corner order, TPZRefQuad), the others stay.  Returns the flattened mesh
    (volume elements in coarse order, children in place of their parent) with
    the constraints of the hanging sides:

    * the middle node M of a coarse edge (A, B) next to a divided element
      depends on the coarse edge's connects A, B and, for p = 2, its edge
      connect E with coefficients 1/2, 1/2, 1 (the coarse side's functions at
      M);
    * for p = 2 each half edge's connect depends on A, B, E with 0, 0, 1/4
      (the coarse edge function restricted to the half edge is 1/4 of the
      half edge's function plus a linear part).

    These are the restraints TPZInterpolatedElement::RestrainSide computes by
    projection (Mesh/pzintel.cpp:847-1010): only the side's own connect is
    restrained, against every connect of the large side that has shapes, the
    block of a negligible coefficient zeroed but kept.  Connects are numbered
    by first touch over the elements (sides in order); the constrained ones
    then move behind the others, keeping their order, so that their blocks
    hold no equations.  The element coefficients of ``pb.el_coef`` (one per
    coarse element) pass to the children."""
    if pb.dim != 2 or pb.p > 2 or pb.el_porder is not None:
        raise ValueError("build_refined2d: quadrilaterals, uniform p <= 2")
    nx, ny = pb.nx, pb.ny
    refine = np.asarray(refine, dtype=bool).reshape(ny, nx)
    NX = 2 * nx + 1
    lat = lambda I, J: I + NX * J
    els, parent = [], []
    for j in range(ny):
        for i in range(nx):
            I, J = 2 * i, 2 * j
            if refine[j, i]:
                for (a, b) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                    I0, J0 = I + a, J + b
                    els.append([lat(I0, J0), lat(I0 + 1, J0), lat(I0 + 1, J0 + 1), lat(I0, J0 + 1)])
                    parent.append(i + nx * j)
            else:
                els.append([lat(I, J), lat(I + 2, J), lat(I + 2, J + 2), lat(I, J + 2)])
                parent.append(i + nx * j)
    vol_nodes = np.array(els, dtype=np.int64)
    nel = len(els)
    # side keys: corner = lattice node; edge = ('e', min, max); interior = ('i', e)
    keymap, seq_of = {}, []
    econn = np.zeros((nel, 9), dtype=np.int64)
    for e, cn in enumerate(els):
        keys = [("n", c) for c in cn]
        keys += [("e", min(cn[k], cn[(k + 1) % 4]), max(cn[k], cn[(k + 1) % 4])) for k in range(4)]
        keys.append(("i", e))
        for s, k in enumerate(keys):
            if k not in keymap:
                keymap[k] = len(keymap)
                seq_of.append(k)
            econn[e, s] = keymap[k]
    nconn = len(keymap)
    # hanging sides: a coarse edge (A, B) whose middle node is a corner of
    # some element; the coarse element holding the edge is the large side
    dep = {}
    for e, cn in enumerate(els):
        for k in range(4):
            A, B = cn[k], cn[(k + 1) % 4]
            mid = (A + B) // 2
            if (A + B) % 2 or ("n", mid) not in keymap:
                continue
            if abs(A - B) not in (2, 2 * NX):
                continue  # not a coarse edge
            cA, cB, cE = keymap[("n", A)], keymap[("n", B)], econn[e, 4 + k]
            large = [cA, cB] + ([cE] if pb.p == 2 else [])
            dep[keymap[("n", mid)]] = (large, [0.5, 0.5] + ([1.0] if pb.p == 2 else []))
            if pb.p == 2:
                for h in ((min(A, mid), max(A, mid)), (min(mid, B), max(mid, B))):
                    dep[keymap[("e",) + h]] = (large, [0.0, 0.0, 0.25])
    # boundary elements: boundary edges of the elements, face by face
    xs = lambda I: I / (2.0 * nx)
    ys = lambda J: J / (2.0 * ny)
    bn, bk, bid = [], [], []
    for f in range(4):
        if pb.bc_type[f] < 0:
            continue
        for e, cn in enumerate(els):
            for k in range(4):
                A, B = cn[k], cn[(k + 1) % 4]
                IA, JA, IB, JB = A % NX, A // NX, B % NX, B // NX
                on = {0: IA == 0 and IB == 0, 1: IA == 2 * nx and IB == 2 * nx,
                      2: JA == 0 and JB == 0, 3: JA == 2 * ny and JB == 2 * ny}[f]
                if on:
                    lo, hi = (A, B) if (IA, JA) < (IB, JB) else (B, A)
                    bn.append((lo, hi))
                    bk.append((keymap[("n", lo)], keymap[("n", hi)], econn[e, 4 + k]))
                    bid.append(f)
    bc_nodes = np.array(bn, dtype=np.int64).reshape(-1, 2)
    bc_conn = np.array(bk, dtype=np.int64).reshape(-1, 3)
    bc_id = np.array(bid, dtype=np.int32)
    # constrained connects behind the others (relative order kept)
    is_dep = np.zeros(nconn, dtype=bool)
    is_dep[list(dep.keys())] = True
    order = np.concatenate([np.flatnonzero(~is_dep), np.flatnonzero(is_dep)])
    perm = np.empty(nconn, dtype=np.int64)
    perm[order] = np.arange(nconn)
    econn = perm[econn]
    bc_conn = perm[bc_conn] if len(bc_conn) else bc_conn
    kinds = np.array([{"n": K_CORNER, "e": K_EDGE, "i": K_QFACE}[k[0]] for k in seq_of])[order]
    ns = pb.nstate
    nshape = np.where(kinds == K_CORNER, 1, 1 if pb.p == 2 else 0)
    blk_size = (nshape * ns).astype(np.int32)
    blk_pos = np.zeros(nconn, dtype=np.int64)
    blk_pos[1:] = np.cumsum(blk_size[:-1].astype(np.int64))
    ndep = int(is_dep.sum())
    neq = int(blk_size[: nconn - ndep].astype(np.int64).sum())
    dep_ptr = np.zeros(nconn + 1, dtype=np.int64)
    dc, dv = [], []
    for newc in range(nconn):
        oldc = order[newc]
        if oldc in dep:
            large, coef = dep[oldc]
            dc += [int(perm[c]) for c in large]
            dv += coef
        dep_ptr[newc + 1] = len(dc)
    xyz = np.zeros((NX * (2 * ny + 1), 3))
    xyz[:, 0] = xs(np.arange(NX * (2 * ny + 1)) % NX)
    xyz[:, 1] = ys(np.arange(NX * (2 * ny + 1)) // NX)
    coef = None if pb.el_coef is None else np.asarray(pb.el_coef, dtype=np.float64)[parent]
    rule = 2 * pb.p
    big = 1.0e12 if pb.material == P.ELAST3D else 10.0 ** 17 * 2 / 3
    return dict(
        vol_rule=None, bc_rule=None, dim=2, material=pb.material, nstate=ns, porder=pb.p,
        xyz=xyz.ravel(), vol_nodes=vol_nodes, vol_conn=econn, vol_coef=coef,
        bc_nodes=bc_nodes, bc_conn=bc_conn, bc_id=bc_id,
        bc_type=np.array([pb.bc_type[f] for f in range(6)], dtype=np.int32),
        bc_val=np.array([pb.bc_val[f] for f in range(6)], dtype=np.float64).ravel(),
        prestress=np.array(getattr(pb, "prestress", (0.0, 0.0, 0.0)), dtype=np.float64),
        bc_val1=np.array([getattr(pb, "bc_val1", [[0.0] * 9] * 6)[f] for f in range(6)],
                         dtype=np.float64).ravel(),
        nconn=nconn, blk_pos=blk_pos, blk_size=blk_size, neq=neq,
        scale=pb.scale, E=pb.E, nu=pb.nu, rule_order_vol=rule, rule_order_bc=2 * pb.p,
        force_kind=pb.force_kind, force_const=np.array(pb.force_const, dtype=np.float64),
        bumps=None if pb.bumps is None else np.ascontiguousarray(pb.bumps, dtype=np.float64).reshape(-1, 6),
        bump_width=pb.bump_width, big_number=big, bc_faces=[f for f in range(6) if pb.bc_type[f] >= 0],
        vol_etype=0, dep_ptr=dep_ptr, dep_conn=np.array(dc, dtype=np.int64),
        dep_coef=np.array(dv, dtype=np.float64), conn_order=np.where(kinds == K_CORNER, 1, pb.p),
        leaf_parent=np.array(parent, dtype=np.int64),
    )


_HEX_EDGE = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 5), (2, 6), (3, 7), (4, 5), (5, 6), (6, 7), (7, 4)]
_HEX_FACE = [(0, 1, 2, 3), (0, 1, 5, 4), (1, 2, 6, 5), (3, 2, 6, 7), (0, 3, 7, 4), (4, 5, 6, 7)]
_HEX_BITS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
_HEX_FACE_BC = {4: 0, 2: 1, 1: 2, 3: 3, 0: 4, 5: 5}  # hexahedron face -> boundary face id


def build_refined3d(pb, refine):
    """A hexahedral mesh of the unit cube with hanging sides: the coarse
    elements flagged in ``refine`` are divided into eight (children in corner
    order, TPZRefCube); boundary elements are the boundary faces of the
    elements (a divided element's children's faces, which may hold
    constrained connects on a coarse boundary edge).  As in
    :func:`build_refined2d`, every side of a divided element that lies inside a
    side of an undivided one (the smallest such large side: an edge, else a
    face) has its own connect restrained against the large side's connects
    with shapes, with the coefficients of the large side's functions in the
    small side's basis (TPZInterpolatedElement::RestrainSide,
    Mesh/pzintel.cpp:847-1010).  For p <= 2 the large side's functions are
    products of 1-t, t and 4t(1-t) along its directions and the coefficients
    follow from values at points: a node takes phi(P); a half edge
    phi(mid) - (phi(P0) + phi(P1))/2; a quarter face phi(centre) -
    sum(phi(edge middles))/2 + sum(phi(corners))/4."""
    if pb.dim != 3 or pb.p > 2 or pb.el_porder is not None or getattr(pb, "simplex", 0):
        raise ValueError("build_refined3d: hexahedra, uniform p <= 2")
    nx, ny, nz = pb.nx, pb.ny, pb.nz
    refine = np.asarray(refine, dtype=bool).reshape(nz, ny, nx)
    NX, NY, NZ = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
    lat = lambda I, J, K: I + NX * (J + NY * K)
    coords = lambda n: np.array([n % NX, (n // NX) % NY, n // (NX * NY)], dtype=np.int64)
    els, parent, size = [], [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                e = i + nx * (j + ny * k)
                b = (2 * i, 2 * j, 2 * k)
                if refine[k, j, i]:
                    for c in range(8):
                        o = [b[d] + _HEX_BITS[c][d] for d in range(3)]
                        els.append([lat(o[0] + q[0], o[1] + q[1], o[2] + q[2]) for q in _HEX_BITS])
                        parent.append(e)
                        size.append(1)
                else:
                    els.append([lat(b[0] + 2 * q[0], b[1] + 2 * q[1], b[2] + 2 * q[2]) for q in _HEX_BITS])
                    parent.append(e)
                    size.append(2)
    nel = len(els)

    def side_keys(cn, e):
        keys = [("n", c) for c in cn]
        keys += [("e",) + tuple(sorted((cn[a], cn[b]))) for a, b in _HEX_EDGE]
        keys += [("f",) + tuple(sorted(cn[q] for q in f)) for f in _HEX_FACE]
        keys.append(("i", e))
        return keys

    keymap, seq_of = {}, []
    econn = np.zeros((nel, 27), dtype=np.int64)
    for e, cn in enumerate(els):
        for s, key in enumerate(side_keys(cn, e)):
            if key not in keymap:
                keymap[key] = len(keymap)
                seq_of.append(key)
            econn[e, s] = keymap[key]
    nconn = len(keymap)
    used = {k[1] for k in keymap if k[0] == "n"}
    has = lambda key: key in keymap
    ekey = lambda a, b: ("e",) + tuple(sorted((a, b)))
    fkey = lambda ns: ("f",) + tuple(sorted(ns))

    def mid(a, b):  # lattice node halfway between two lattice nodes (coarse sides)
        return int(lat(*((coords(a) + coords(b)) // 2)))

    X = lambda n: coords(n).astype(np.float64)  # lattice coordinates of a node
    half = lambda a, b: (X(a) + X(b)) / 2.0      # a point halfway (may be off the lattice)

    dep = {}

    def restrain(small_keys, large_conn, phi):
        """small_keys: (key, kind, points) of small sides; phi(P) -> values of
        the large side's functions at lattice point P."""
        for key, kind, pts in small_keys:
            if not has(key):
                continue
            if kind == "n":
                v = phi(pts[0])
            elif kind == "e":
                v = phi(pts[2]) - (phi(pts[0]) + phi(pts[1])) / 2
            else:  # quarter face: corners pts[0:4], edge middles pts[4:8], centre pts[8]
                v = phi(pts[8]) - sum(phi(q) for q in pts[4:8]) / 2 + sum(phi(q) for q in pts[0:4]) / 4
            v = np.where(np.abs(v) < 1e-14, 0.0, v)
            dep[keymap[key]] = (list(large_conn), [float(x) for x in v])

    for e, cn in enumerate(els):
        if size[e] != 2:
            continue
        # hanging edges
        for s, (a, b) in enumerate(_HEX_EDGE):
            A, B = cn[a], cn[b]
            M = mid(A, B)
            if M not in used:
                continue
            cA, cB, cE = keymap[("n", A)], keymap[("n", B)], econn[e, 8 + s]
            large = [cA, cB] + ([cE] if pb.p == 2 else [])
            pa, pb_ = coords(A), coords(B)

            def phi_e(P, pa=pa, pb_=pb_):
                t = float(np.dot(P - pa, pb_ - pa)) / float(np.dot(pb_ - pa, pb_ - pa))
                f = [1 - t, t] + ([4 * t * (1 - t)] if pb.p == 2 else [])
                return np.array(f)

            small = [(("n", M), "n", [X(M)])]
            if pb.p == 2:
                small += [(ekey(A, M), "e", [X(A), X(M), half(A, M)]),
                          (ekey(M, B), "e", [X(M), X(B), half(M, B)])]
            restrain(small, large, phi_e)
        # hanging faces
        for s, f in enumerate(_HEX_FACE):
            C = [cn[q] for q in f]
            O = int(lat(*((coords(C[0]) + coords(C[2])) // 2)))
            if O not in used:
                continue
            large = [keymap[("n", c)] for c in C]
            if pb.p == 2:
                large += [keymap[ekey(C[q], C[(q + 1) % 4])] for q in range(4)] + [econn[e, 20 + s]]
            c0, c1, c3 = coords(C[0]), coords(C[1]), coords(C[3])

            def phi_f(P, c0=c0, c1=c1, c3=c3):
                r = P - c0
                u = float(np.dot(r, c1 - c0)) / float(np.dot(c1 - c0, c1 - c0))
                v = float(np.dot(r, c3 - c0)) / float(np.dot(c3 - c0, c3 - c0))
                vals = [(1 - u) * (1 - v), u * (1 - v), u * v, (1 - u) * v]
                if pb.p == 2:
                    vals += [4 * u * (1 - u) * (1 - v), 4 * v * (1 - v) * u, 4 * u * (1 - u) * v,
                             4 * v * (1 - v) * (1 - u), 16 * u * (1 - u) * v * (1 - v)]
                return np.array(vals)

            Ms = [mid(C[q], C[(q + 1) % 4]) for q in range(4)]
            small = [(("n", O), "n", [X(O)])]
            if pb.p == 2:
                for q in range(4):
                    small.append((ekey(O, Ms[q]), "e", [X(O), X(Ms[q]), half(O, Ms[q])]))
                for q in range(4):  # quarter face at corner C[q]
                    qc = [C[q], Ms[q], O, Ms[(q + 3) % 4]]
                    qm = [half(qc[r], qc[(r + 1) % 4]) for r in range(4)]
                    small.append((fkey(qc), "f", [X(n) for n in qc] + qm + [half(qc[0], qc[2])]))
            restrain(small, large, phi_f)

    # boundary elements: boundary faces of the (undivided) elements
    bn, bk, bid = [], [], []
    for f_id in range(6):
        if pb.bc_type[f_id] < 0:
            continue
        for e, cn in enumerate(els):
            for s, f in enumerate(_HEX_FACE):
                if _HEX_FACE_BC[s] != f_id:
                    continue
                C = [cn[q] for q in f]
                cc = np.array([coords(c) for c in C])
                ax, lo = {0: (0, 0), 1: (0, 2 * nx), 2: (1, 0), 3: (1, 2 * ny), 4: (2, 0), 5: (2, 2 * nz)}[f_id]
                if not np.all(cc[:, ax] == lo):
                    continue
                bn.append(C)
                conn = [keymap[("n", c)] for c in C] + \
                       [keymap[ekey(C[q], C[(q + 1) % 4])] for q in range(4)] + [econn[e, 20 + s]]
                bk.append(conn)
                bid.append(f_id)
    bc_nodes = np.array(bn, dtype=np.int64).reshape(-1, 4)
    bc_conn = np.array(bk, dtype=np.int64).reshape(-1, 9)
    bc_id = np.array(bid, dtype=np.int32)
    # constrained connects behind the others (relative order kept)
    is_dep = np.zeros(nconn, dtype=bool)
    is_dep[list(dep.keys())] = True
    order = np.concatenate([np.flatnonzero(~is_dep), np.flatnonzero(is_dep)])
    perm = np.empty(nconn, dtype=np.int64)
    perm[order] = np.arange(nconn)
    econn = perm[econn]
    bc_conn = perm[bc_conn] if len(bc_conn) else bc_conn
    kinds = np.array([{"n": K_CORNER, "e": K_EDGE, "f": K_QFACE, "i": K_HEX}[k[0]] for k in seq_of])[order]
    ns = pb.nstate
    nshape = np.where(kinds == K_CORNER, 1, 1 if pb.p == 2 else 0)
    blk_size = (nshape * ns).astype(np.int32)
    blk_pos = np.zeros(nconn, dtype=np.int64)
    blk_pos[1:] = np.cumsum(blk_size[:-1].astype(np.int64))
    ndep = int(is_dep.sum())
    neq = int(blk_size[: nconn - ndep].astype(np.int64).sum())
    dep_ptr = np.zeros(nconn + 1, dtype=np.int64)
    dc, dv = [], []
    for newc in range(nconn):
        oldc = order[newc]
        if oldc in dep:
            large, coef = dep[oldc]
            dc += [int(perm[c]) for c in large]
            dv += coef
        dep_ptr[newc + 1] = len(dc)
    nn = NX * NY * NZ
    xyz = np.zeros((nn, 3))
    idx = np.arange(nn)
    xyz[:, 0] = (idx % NX) / (2.0 * nx)
    xyz[:, 1] = ((idx // NX) % NY) / (2.0 * ny)
    xyz[:, 2] = (idx // (NX * NY)) / (2.0 * nz)
    coef = None if pb.el_coef is None else np.asarray(pb.el_coef, dtype=np.float64)[parent]
    ff = pb.force_porder if pb.force_kind != P.FORCE_NONE else 0
    rule = 2 * pb.p if (pb.material == P.ELAST3D or pb.p >= ff) else ff + pb.p
    big = 1.0e12 if pb.material == P.ELAST3D else 10.0 ** 17 * 2 / 3
    return dict(
        vol_rule=None, bc_rule=None, dim=3, material=pb.material, nstate=ns, porder=pb.p,
        xyz=xyz.ravel(), vol_nodes=np.array(els, dtype=np.int64), vol_conn=econn, vol_coef=coef,
        bc_nodes=bc_nodes, bc_conn=bc_conn, bc_id=bc_id,
        bc_type=np.array([pb.bc_type[f] for f in range(6)], dtype=np.int32),
        bc_val=np.array([pb.bc_val[f] for f in range(6)], dtype=np.float64).ravel(),
        prestress=np.array(getattr(pb, "prestress", (0.0, 0.0, 0.0)), dtype=np.float64),
        bc_val1=np.array([getattr(pb, "bc_val1", [[0.0] * 9] * 6)[f] for f in range(6)],
                         dtype=np.float64).ravel(),
        nconn=nconn, blk_pos=blk_pos, blk_size=blk_size, neq=neq,
        scale=pb.scale, E=pb.E, nu=pb.nu, rule_order_vol=rule, rule_order_bc=2 * pb.p,
        force_kind=pb.force_kind, force_const=np.array(pb.force_const, dtype=np.float64),
        bumps=None if pb.bumps is None else np.ascontiguousarray(pb.bumps, dtype=np.float64).reshape(-1, 6),
        bump_width=pb.bump_width, big_number=big, bc_faces=[f for f in range(6) if pb.bc_type[f] >= 0],
        vol_etype=0, dep_ptr=dep_ptr, dep_conn=np.array(dc, dtype=np.int64),
        dep_coef=np.array(dv, dtype=np.float64), conn_order=np.where(kinds == K_CORNER, 1, pb.p),
        leaf_parent=np.array(parent, dtype=np.int64),
    )


def volume_dofs(m):
    """Equation of every dof of every volume element, shape-major
    (index = nstate * shape + state), the order of the element matrix: the
    shapes are the sides whose connect carries equations (uniform order
    p <= 2: all sides for p = 2, the corners for p = 1).  -1 where a shape's
    connect has none in this element."""
    vc = np.asarray(m["vol_conn"]).reshape(len(m["vol_nodes"]), -1)
    pos = np.asarray(m["blk_pos"])
    size = np.asarray(m["blk_size"])
    ns = m["nstate"]
    has = (size[vc] > 0).any(axis=0)  # sides that carry shapes in some element
    sides = np.flatnonzero(has)
    out = np.full((vc.shape[0], len(sides) * ns), -1, dtype=np.int64)
    for j, sd in enumerate(sides):
        c = vc[:, sd]
        ok = size[c] > 0
        for k in range(ns):
            out[ok, j * ns + k] = pos[c[ok]] + k
    return out
