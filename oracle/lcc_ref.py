"""FP64 numpy reference of the learned coarse-correction network (checker only).

TEST INFRASTRUCTURE: used by tests/ to check the HIP forward/backward kernels of
the learned preconditioner (neopz_amd/csrc/pzg_learned.hip).  NeoPZ has no
learned component, so this file restates the network defined by the product,
not a reference algorithm; its own backward pass is pinned by central finite
differences in tests/test_learned_oracle.py.

Network on a coarse graph (CSR cia/cja, no self loops, deg_I = row length):
    H0 = X W0                                  (nc, F)
    M_l = Ahat H_{l-1},  Ahat = D^-1 A         (mean over neighbours)
    P_l = H_{l-1} A_l + M_l B_l,  H_l = lrelu(P_l)     l = 1..L
    Z  = H_L Wout                              (nc, bs)
Loss over samples s:  sum_s  sum_I w_I |Z_sI - T_sI|^2 / sum_I w_I |T_sI|^2
"""
import numpy as np

SLOPE = 0.1


def lrelu(x):
    return np.where(x > 0, x, SLOPE * x)


def dlrelu(x):
    return np.where(x > 0, 1.0, SLOPE)


def mean_agg(cia, cja, H):
    out = np.zeros_like(H)
    for i in range(len(cia) - 1):
        nb = cja[cia[i]:cia[i + 1]]
        if len(nb):
            out[..., i, :] = H[..., nb, :].sum(axis=-2) / len(nb)
    return out


def mean_agg_t(cia, cja, Y):
    deg = np.maximum(np.diff(cia), 1)
    out = np.zeros_like(Y)
    for j in range(len(cia) - 1):
        nb = cja[cia[j]:cia[j + 1]]
        if len(nb):
            out[..., j, :] = (Y[..., nb, :] / deg[nb][:, None]).sum(axis=-2)
    return out


def unpack(theta, bs, F, L):
    k = 0
    W0 = theta[k:k + bs * F].reshape(bs, F); k += bs * F
    A, B = [], []
    for _ in range(L):
        A.append(theta[k:k + F * F].reshape(F, F)); k += F * F
        B.append(theta[k:k + F * F].reshape(F, F)); k += F * F
    Wout = theta[k:k + F * bs].reshape(F, bs); k += F * bs
    return W0, A, B, Wout


def nparams(bs, F, L):
    return bs * F + 2 * L * F * F + F * bs


def forward(theta, X, cia, cja, bs, F, L):
    W0, A, B, Wout = unpack(theta, bs, F, L)
    H = [X @ W0]
    Ms, Ps = [], []
    for l in range(L):
        M = mean_agg(cia, cja, H[-1])
        P = H[-1] @ A[l] + M @ B[l]
        Ms.append(M)
        Ps.append(P)
        H.append(lrelu(P))
    Z = H[-1] @ Wout
    return Z, (H, Ms, Ps)


def loss_and_grad(theta, X, T, w, cia, cja, bs, F, L):
    """X, T: (nsamp, nc, bs); w: (nc,). Returns loss, grad (same layout as theta)."""
    W0, A, B, Wout = unpack(theta, bs, F, L)
    Z, (H, Ms, Ps) = forward(theta, X, cia, cja, bs, F, L)
    den = np.einsum("i,sik->s", w, T * T)
    den = np.maximum(den, 1e-300)
    diff = Z - T
    loss = float(np.sum(np.einsum("i,sik->s", w, diff * diff) / den))
    dZ = 2.0 * w[None, :, None] * diff / den[:, None, None]
    gWout = np.einsum("sif,sik->fk", H[-1], dZ)
    dH = dZ @ Wout.T
    gA, gB = [None] * L, [None] * L
    for l in reversed(range(L)):
        G = dH * dlrelu(Ps[l])
        gA[l] = np.einsum("sif,sig->fg", H[l], G)
        gB[l] = np.einsum("sif,sig->fg", Ms[l], G)
        dH = G @ A[l].T + mean_agg_t(cia, cja, G @ B[l].T)
    gW0 = np.einsum("sik,sif->kf", X, dH)
    g = [gW0.ravel()]
    for l in range(L):
        g += [gA[l].ravel(), gB[l].ravel()]
    g.append(gWout.ravel())
    return loss, np.concatenate(g)
