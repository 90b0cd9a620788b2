"""Learned element smoother of the mixed solver's multigrid cycle (host
driver over the C ABI).

The operator, the network and its training kernels live in
``neopz_amd/csrc/pzg_learned.hip``: the smoothing of the p level uses
``z = sum_e s_e P_e^T G P_e r`` with one learned symmetric element matrix G
and per-element scales s_e from a message-passing network on the element
graph.  Training pairs are the (residual, error) pairs of every element's
local problem in seeded training systems.  NeoPZ has no learned component;
the preconditioner is an option of the mixed solver (``PZG_PRECOND_LEARNED``)
whose fallback is plain FP64 Jacobi-PCG.
"""
import ctypes as C

import numpy as np

from . import _lib as L

NFEAT = 4    # element features
NHID = 32    # hidden features
NLAYER = 3   # message-passing layers


class LearnedSmoother:
    def __init__(self, ctx):
        self.ctx = ctx
        self.lib = ctx.lib

    def setup(self, seed=0):
        L.check(self.lib.pzg_learned_setup(self.ctx.h, seed))

    def collect(self, nrhs=4, seed=1):
        L.check(self.lib.pzg_learned_collect(self.ctx.h, nrhs, seed))

    def train(self, steps=200, lr=2e-3):
        l0, l1 = C.c_double(), C.c_double()
        L.check(self.lib.pzg_learned_train(self.ctx.h, steps, lr, C.byref(l0), C.byref(l1)))
        return l0.value, l1.value

    def nparams(self):
        n = C.c_int32()
        L.check(self.lib.pzg_learned_weights(self.ctx.h, None, None, C.byref(n)))
        return n.value

    def get_weights(self):
        w = np.zeros(self.nparams(), np.float32)
        L.check(self.lib.pzg_learned_weights(self.ctx.h, L.ptr(w, L.fp), None, None))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert len(w) == self.nparams()
        L.check(self.lib.pzg_learned_weights(self.ctx.h, None, L.ptr(w, L.fp), None))

    def graph(self, neq):
        """(cia, cja, features [nel, 4], penalized mask [neq]) of the current
        assembly."""
        nel = C.c_int64()
        L.check(self.lib.pzg_learned_graph(self.ctx.h, C.byref(nel), None, None, None, None))
        cia = np.zeros(nel.value + 1, np.int64)
        L.check(self.lib.pzg_learned_graph(self.ctx.h, None, L.ptr(cia, L.i64p), None, None, None))
        cja = np.zeros(max(int(cia[-1]), 1), np.int32)
        X = np.zeros((nel.value, NFEAT), np.float32)
        pen = np.zeros(neq, np.float32)
        L.check(self.lib.pzg_learned_graph(self.ctx.h, None, None, L.ptr(cja, L.i32p), L.ptr(X, L.fp),
                                           L.ptr(pen, L.fp)))
        return cia, cja[: cia[-1]].astype(np.int64), X, pen

    def samples(self, nel):
        ne, nr = C.c_int32(), C.c_int32()
        L.check(self.lib.pzg_learned_samples(self.ctx.h, C.byref(ne), C.byref(nr), None, None, None))
        R = np.zeros((nel, nr.value, ne.value), np.float32)
        E = np.zeros_like(R)
        Mk = np.zeros((nel, ne.value), np.float32)
        L.check(self.lib.pzg_learned_samples(self.ctx.h, None, None, L.ptr(R, L.fp), L.ptr(E, L.fp),
                                             L.ptr(Mk, L.fp)))
        return R, E, Mk

    def loss_grad(self):
        g = np.zeros(self.nparams(), np.float32)
        loss = C.c_double()
        L.check(self.lib.pzg_learned_loss_grad(self.ctx.h, C.byref(loss), L.ptr(g, L.fp)))
        return loss.value, g

    def apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        L.check(self.lib.pzg_learned_apply(self.ctx.h, L.ptr(r, L.dp), L.ptr(z, L.dp)))
        return z

    def scales(self, nel):
        s = np.zeros(nel, np.float32)
        L.check(self.lib.pzg_learned_scales(self.ctx.h, L.ptr(s, L.fp)))
        return s


def train_on_seeds(ctx, strmat, K, make_coef, seeds, nrhs=4, steps=100, lr=2e-3, setup_seed=0,
                   rounds=1, progress=None):
    """Trains the learned smoother on seeded coefficient fields of one mesh:
    for each seed, pzg_mesh_set_coef (skipped when make_coef returns None:
    the mesh's own coefficients) + assembly + pairs of the local problems,
    then the least-squares fit of G on every pair so far and ``steps`` Adam
    steps of the network on the seed's pairs (``rounds`` passes over the
    seeds; ``progress(k, loss)``, if given, after every 100th seed).  Returns
    (LearnedSmoother, loss of the initial parameters on the first seed's
    pairs, loss after the last step)."""
    ls = LearnedSmoother(ctx)
    ls.setup(seed=setup_seed)
    first = last = None
    k = 0
    for _ in range(rounds):
        for s in seeds:
            coef = make_coef(s)
            if coef is not None:  # (a family without per-element coefficients keeps its mesh's)
                coef = np.ascontiguousarray(coef, np.float64)
                if coef.shape != (ctx.sizes()["nvol"],):
                    raise ValueError("make_coef must return one coefficient per volume element")
                L.check(ctx.lib.pzg_mesh_set_coef(ctx.h, L.ptr(coef, L.dp)))
            strmat.Assemble(K)
            ls.collect(nrhs=nrhs, seed=1000 + s)
            if first is None:  # the loss of the initial parameters
                first = ls.loss_grad()[0]
            _, last = ls.train(steps=steps, lr=lr)
            k += 1
            if progress is not None and k % 100 == 0:
                progress(k, last)
    return ls, first, last
