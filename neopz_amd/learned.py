"""Learned coarse-correction preconditioner of the mixed solver (host driver).

The network, its training kernels and its use inside FCG live in
``neopz_amd/csrc/pzg_learned.hip``; this module drives the C ABI:
aggregation set-up, pair generation on seeded problem instances, training,
and weight transfer.  NeoPZ has no learned component; the preconditioner is
an option of the mixed solver (``PZG_PRECOND_LEARNED``) whose fallback is
block Jacobi.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class LearnedPreconditioner:
    def __init__(self, ctx):
        self.ctx = ctx
        self.lib = ctx.lib

    def setup(self, coarsen=4, seed=0, max_samples=512):
        L.check(self.lib.pzg_learned_setup(self.ctx.h, coarsen, seed, max_samples))

    def collect(self, nrhs=2, record=(2, 5, 10, 20, 40, 80), seed=1):
        rec = np.ascontiguousarray(record, dtype=np.int32)
        L.check(self.lib.pzg_learned_collect(self.ctx.h, nrhs, L.ptr(rec, L.i32p), len(rec), seed))

    def train(self, steps=300, lr=3e-3):
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

    def graph(self, nb):
        nc = C.c_int64()
        ns = C.c_int64()
        L.check(self.lib.pzg_learned_graph(self.ctx.h, C.byref(nc), None, None, None, C.byref(ns)))
        cia = np.zeros(nc.value + 1, np.int64)
        L.check(self.lib.pzg_learned_graph(self.ctx.h, None, L.ptr(cia, L.i64p), None, None, None))
        cja = np.zeros(max(int(cia[-1]), 1), np.int32)
        agg = np.zeros(nb, np.int32)
        L.check(self.lib.pzg_learned_graph(self.ctx.h, None, None, L.ptr(cja, L.i32p),
                                           L.ptr(agg, L.i32p), None))
        return cia, cja[: cia[-1]].astype(np.int64), agg, ns.value

    def samples(self, nc, bs, ns):
        X = np.zeros(ns * nc * bs, np.float32)
        T = np.zeros(ns * nc * bs, np.float32)
        L.check(self.lib.pzg_learned_samples(self.ctx.h, L.ptr(X, L.fp), L.ptr(T, L.fp)))
        return X.reshape(ns, nc, bs), T.reshape(ns, nc, bs)

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.nparams(), np.float32)
        loss = C.c_double()
        L.check(self.lib.pzg_learned_loss_grad(self.ctx.h, X.shape[0], L.ptr(X, L.fp), L.ptr(T, L.fp),
                                               C.byref(loss), L.ptr(g, L.fp)))
        return loss.value, g


def train_on_seeds(ctx, strmat, K, make_coef, seeds, coarsen=4, nrhs=2,
                   record=(2, 5, 10, 20, 40, 80), steps=400, lr=3e-3, setup_seed=0):
    """Generates pairs on every seeded coefficient field of the same mesh
    (pzg_mesh_set_coef + pzg_assemble + pzg_learned_collect), then trains.
    Returns (LearnedPreconditioner, loss_first, loss_last)."""
    lp = LearnedPreconditioner(ctx)
    lp.setup(coarsen=coarsen, seed=setup_seed, max_samples=len(seeds) * nrhs * len(record))
    for s in seeds:
        coef = np.ascontiguousarray(make_coef(s), np.float64)
        L.check(ctx.lib.pzg_mesh_set_coef(ctx.h, L.ptr(coef, L.dp)))
        strmat.Assemble(K)
        lp.collect(nrhs=nrhs, record=record, seed=1000 + s)
    l0, l1 = lp.train(steps=steps, lr=lr)
    return lp, l0, l1
