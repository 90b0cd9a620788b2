"""FP64 numpy statement of the learned element smoother (test
infrastructure: the checker of the HIP kernels in
neopz_amd/csrc/pzg_learned.hip; the product never imports it).

Parameters theta (flat): G (ne x ne, row-major), W0 (4 x 32), then A_l, B_l
(32 x 32) for l = 1..3, then w (32); matrices stored [in][out].
Network on the element graph (cia, cja; mean over the neighbours):
    H0 = X W0;  P_l = H_{l-1} A_l + mean_nb(H_{l-1}) B_l;  H_l = lrelu(P_l);
    z = H3 w;   s = exp(clip(z, -4, 4))          (lrelu slope 0.1)
Loss on pairs R, E ([element][pair][ne]) with the dof mask Mk ([element][ne]):
    L = sum_{e,k} |Mk_e * (s_e G r_ek - e_ek)|^2 / sum |e_ek|^2
Smoother: z = sum_e s_e P_e^T G (mask r)_e on unpenalized equations, r on
penalized ones.
"""
import numpy as np

NF, NH, NL, SLOPE = 4, 32, 3, 0.1


def offsets(ne):
    o = {"g": 0}
    k = ne * ne
    o["w0"] = k
    k += NF * NH
    o["a"], o["b"] = [], []
    for _ in range(NL):
        o["a"].append(k)
        k += NH * NH
        o["b"].append(k)
        k += NH * NH
    o["w"] = k
    k += NH
    o["n"] = k
    return o


def unpack(theta, ne):
    o = offsets(ne)
    t = np.asarray(theta, dtype=np.float64)
    G = t[o["g"]:o["g"] + ne * ne].reshape(ne, ne)
    W0 = t[o["w0"]:o["w0"] + NF * NH].reshape(NF, NH)
    A = [t[a:a + NH * NH].reshape(NH, NH) for a in o["a"]]
    B = [t[b:b + NH * NH].reshape(NH, NH) for b in o["b"]]
    w = t[o["w"]:o["w"] + NH]
    return G, W0, A, B, w


def _mean_nb(H, cia, cja):
    out = np.zeros_like(H)
    for e in range(len(cia) - 1):
        j0, j1 = cia[e], cia[e + 1]
        if j1 > j0:
            out[e] = H[cja[j0:j1]].mean(axis=0)
    return out


def _mean_nb_t(Y, cia, cja):
    """transpose of the neighbour mean: out[j] += Y[e] / deg(e) for j in N(e)"""
    out = np.zeros_like(Y)
    for e in range(len(cia) - 1):
        j0, j1 = cia[e], cia[e + 1]
        if j1 > j0:
            out[cja[j0:j1]] += Y[e] / (j1 - j0)
    return out


def forward(theta, ne, X, cia, cja):
    G, W0, A, B, w = unpack(theta, ne)
    H = [np.asarray(X, np.float64) @ W0]
    Ms, Ps = [], []
    for l in range(NL):
        Mm = _mean_nb(H[-1], cia, cja)
        P = H[-1] @ A[l] + Mm @ B[l]
        Ms.append(Mm)
        Ps.append(P)
        H.append(np.where(P > 0, P, SLOPE * P))
    z = H[-1] @ w
    s = np.exp(np.clip(z, -4.0, 4.0))
    return z, s, (H, Ms, Ps)


def loss_grad(theta, ne, X, cia, cja, R, E, Mk):
    G, W0, A, B, w = unpack(theta, ne)
    z, s, (H, Ms, Ps) = forward(theta, ne, X, cia, cja)
    R = np.asarray(R, np.float64)
    E = np.asarray(E, np.float64)
    Mk = np.asarray(Mk, np.float64)
    den = float(np.sum(E * E))
    Y = np.einsum("ij,ekj->eki", G, R)
    D = Mk[:, None, :] * (s[:, None, None] * Y - E)
    loss = float(np.sum(D * D)) / den
    gG = 2.0 * np.einsum("e,eki,ekj->ij", s, D, R) / den
    gG = 0.5 * (gG + gG.T)
    ds = 2.0 * np.einsum("eki,eki->e", D, Y) / den
    dz = np.where((z > -4) & (z < 4), s * ds, 0.0)
    o = offsets(ne)
    grad = np.zeros(o["n"])
    grad[o["g"]:o["g"] + ne * ne] = gG.ravel()
    grad[o["w"]:o["w"] + NH] = H[-1].T @ dz
    dH = np.outer(dz, w)
    for l in range(NL - 1, -1, -1):
        Gl = dH * np.where(Ps[l] > 0, 1.0, SLOPE)
        grad[o["a"][l]:o["a"][l] + NH * NH] = (H[l].T @ Gl).ravel()
        grad[o["b"][l]:o["b"][l] + NH * NH] = (Ms[l].T @ Gl).ravel()
        dH = Gl @ A[l].T + _mean_nb_t(Gl @ B[l].T, cia, cja)
    grad[o["w0"]:o["w0"] + NF * NH] = (np.asarray(X, np.float64).T @ dH).ravel()
    return loss, grad


def smoother_apply(G, s, dests, pen, r):
    """z = sum_e s_e P_e^T G (mask r)_e, r on penalized equations; dests
    [element][ne] equation numbers (-1 absent), pen [neq] 0/1."""
    r = np.asarray(r, np.float64)
    keep = (dests >= 0)
    d = np.where(keep, dests, 0)
    m = keep & (np.asarray(pen)[d] == 0)
    Re = np.where(m, r[d], 0.0)
    Ze = s[:, None] * (Re @ np.asarray(G, np.float64).T)
    z = np.zeros_like(r)
    np.add.at(z, d[keep], Ze[keep])
    z[np.asarray(pen) != 0] = r[np.asarray(pen) != 0]
    return z
