"""Seeded problem descriptions for the linear FEM path (synthetic inputs).

A :class:`Problem` is what a NeoPZ user would set up before
``TPZLinearAnalysis::Assemble()``: a structured geometric mesh of the unit
square/cube, an H1 approximation order, one volume material
(``TPZMatPoisson``, ``TPZDarcyFlow`` or ``TPZElasticity3D``), per-element
coefficients and boundary conditions on the six (four) boundary faces.

NeoPZ's own random generators are ``clock()``-seeded
(``Random/TPZUniformRandom.h:20``), so every random field here comes from
numpy's seeded ``default_rng`` and is handed to both the HIP path and the CPU
oracle as data.

Boundary face numbering: 0: x=0, 1: x=1, 2: y=0, 3: y=1, 4: z=0, 5: z=1.
BC type -1 = no boundary element, else the TPZBndCondT type of the material's
ContributeBC (0 Dirichlet penalty, 1 Neumann, 2 Robin/mixed; TPZElasticity3D
also 3..8), with Val2 = ``bc_val`` and Val1 = ``bc_val1``.
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

POISSON, DARCY, ELAST3D, MOCK1D = 0, 1, 2, 3
FORCE_NONE, FORCE_CONST, FORCE_BUMPS = 0, 1, 2


@dataclass
class Problem:
    material: int
    dim: int
    p: int
    nx: int
    ny: int = 1
    nz: int = 1
    scale: float = 1.0
    E: float = 1.0
    nu: float = 0.3
    el_coef: Optional[np.ndarray] = None
    force_kind: int = FORCE_NONE
    force_porder: int = 0
    force_const: tuple = (0.0, 0.0, 0.0)
    bumps: Optional[np.ndarray] = None
    bump_width: float = 0.1
    bc_type: list = field(default_factory=lambda: [-1] * 6)
    bc_val: list = field(default_factory=lambda: [[0.0, 0.0, 0.0] for _ in range(6)])
    # TPZBndCondT::Val1 per face, row-major 3x3 (Val1(0,0) for scalar materials)
    bc_val1: list = field(default_factory=lambda: [[0.0] * 9 for _ in range(6)])
    # TPZElasticity3D::fPreStress (XX, YY, ZZ)
    prestress: tuple = (0.0, 0.0, 0.0)
    name: str = ""
    # hp: preferred order per volume element (1..p); None = p everywhere.
    # Boundary elements keep the default order p.
    el_porder: Optional[np.ndarray] = None
    # dim 3: 1 = six tetrahedra per cube (Kuhn split), boundary triangles
    simplex: int = 0

    @property
    def nstate(self):
        return 3 if self.material == ELAST3D else 1

    @property
    def nvol(self):
        n = self.nx * (self.ny if self.dim >= 2 else 1) * (self.nz if self.dim >= 3 else 1)
        return 6 * n if self.simplex else n

    def n_equations(self):
        n = 1
        for d, k in zip(range(self.dim), (self.nx, self.ny, self.nz)):
            n *= self.p * k + 1
        return n * self.nstate


def gaussian_random_field(shape, corr_len, sigma, rng):
    """Stationary Gaussian field with covariance sigma^2 exp(-r^2/(2 l^2)) at the
    element centres of a unit-box grid of ``shape`` cells (circulant embedding
    on the doubled periodic grid; negative eigenvalues of the embedding are
    clipped to zero)."""
    shape = tuple(shape)
    big = tuple(2 * s for s in shape)
    axes = []
    for s, S in zip(shape, big):
        h = 1.0 / s
        k = np.arange(S)
        d = np.minimum(k, S - k) * h
        axes.append(d)
    grids = np.meshgrid(*axes, indexing="ij")
    r2 = sum(g * g for g in grids)
    cov = sigma * sigma * np.exp(-r2 / (2.0 * corr_len * corr_len))
    lam = np.fft.fftn(cov).real
    lam = np.maximum(lam, 0.0)
    noise = rng.standard_normal(big) + 1j * rng.standard_normal(big)
    field_ = np.fft.ifftn(np.sqrt(lam) * noise) * np.sqrt(np.prod(big))
    sl = tuple(slice(0, s) for s in shape)
    out = field_.real[sl]
    # element index is x fastest: transpose (x,y,z) -> (z,y,x) before ravel
    return np.ascontiguousarray(out.transpose(tuple(range(len(shape)))[::-1])).ravel()


def poisson2d(n=64, p=2, seed=0, nbumps=8, width=0.1):
    """2D Poisson on [0,1]^2, f = sum of seeded Gaussian bumps, u=0 on all sides."""
    rng = np.random.default_rng(seed)
    b = np.zeros((nbumps, 6))
    b[:, 0:2] = rng.uniform(0.0, 1.0, size=(nbumps, 2))
    b[:, 3] = rng.standard_normal(nbumps)
    return Problem(material=POISSON, dim=2, p=p, nx=n, ny=n, force_kind=FORCE_BUMPS,
                   force_porder=2, bumps=b, bump_width=width, bc_type=[0, 0, 0, 0, -1, -1],
                   name=f"poisson2d_q{n}x{n}_p{p}")


def elasticity3d(n=48, p=2, seed=0, sigma=1.0, nu=0.3, field_kind="iid", corr_len=0.1,
                 load="traction"):
    """3D linear elasticity on [0,1]^3 with per-element Young's modulus.

    load="traction": clamped at x=0, unit traction (1,0,0) at x=1.
    load="body": clamped base z=0, seeded smooth body load (Gaussian bumps)."""
    rng = np.random.default_rng(seed)
    if field_kind == "iid":
        E = np.exp(sigma * rng.standard_normal(n * n * n))
    else:
        E = np.exp(gaussian_random_field((n, n, n), corr_len, sigma, rng))
    pb = Problem(material=ELAST3D, dim=3, p=p, nx=n, ny=n, nz=n, E=1.0, nu=nu, el_coef=E,
                 name=f"elasticity3d_h{n}_p{p}_{field_kind}_{load}")
    if load == "traction":
        pb.bc_type = [0, 1, -1, -1, -1, -1]
        pb.bc_val[1] = [1.0, 0.0, 0.0]
    else:
        pb.bc_type = [-1, -1, -1, -1, 0, -1]
        nb = 4
        bumps = np.zeros((nb, 6))
        bumps[:, 0:3] = rng.uniform(0.0, 1.0, size=(nb, 3))
        bumps[:, 3:6] = rng.standard_normal((nb, 3))
        pb.force_kind = FORCE_BUMPS
        pb.force_porder = 2
        pb.bumps = bumps
        pb.bump_width = 0.25
    return pb


def darcy2d(n=256, p=1, seed=0, sigma=1.5, corr_len=0.05):
    """2D Darcy-type diffusion, kappa = exp(GRF) per element, f = 1, u = 0 on the boundary."""
    rng = np.random.default_rng(seed)
    k = np.exp(gaussian_random_field((n, n), corr_len, sigma, rng))
    return Problem(material=DARCY, dim=2, p=p, nx=n, ny=n, el_coef=k, force_kind=FORCE_CONST,
                   force_const=(1.0, 0.0, 0.0), force_porder=0, bc_type=[0, 0, 0, 0, -1, -1],
                   name=f"darcy2d_q{n}_p{p}_s{seed}")


def darcy3d_tet(n=32, pmax=2, seed=0, sigma=1.0, hp=True):
    """hp Darcy on the Kuhn tetrahedral mesh of the unit cube (6 tetrahedra per
    cube of an n^3 grid): kappa_e = exp(N(0, sigma^2)) per element, f = 1,
    u = 0 on the boundary; per-element order drawn uniformly from 1..pmax by
    seed (pmax <= 5)."""
    rng = np.random.default_rng(seed)
    nvol = 6 * n ** 3
    k = np.exp(sigma * rng.standard_normal(nvol))
    po = rng.integers(1, pmax + 1, nvol).astype(np.int32) if hp else None
    return Problem(material=DARCY, dim=3, p=pmax, nx=n, ny=n, nz=n, el_coef=k,
                   force_kind=FORCE_CONST, force_const=(1.0, 0.0, 0.0), force_porder=0,
                   bc_type=[0] * 6, simplex=1, el_porder=po, name=f"darcy_tet{n}_p{pmax}_s{seed}")


def mock1d(n=20):
    """1D mesh with the constant mock element matrix [[2,1],[1,2]]
    (UnitTest_PZ/TestStruct/StructMatrixUnitTest.cpp:333-362)."""
    return Problem(material=MOCK1D, dim=1, p=1, nx=n, name=f"mock1d_{n}")


CONFIGS = {
    # name: (factory, kwargs) -- the workloads quoted in the benchmark table
    "C1": (poisson2d, dict(n=64, p=2)),
    "C2": (elasticity3d, dict(n=48, p=2, field_kind="iid", load="traction")),
    "C3": (darcy3d_tet, dict(n=32, pmax=5)),
    "C4p1": (darcy2d, dict(n=256, p=1)),
    "C4p2": (darcy2d, dict(n=256, p=2)),
    "C5": (elasticity3d, dict(n=64, p=2, field_kind="grf", corr_len=0.1, load="body")),
}


def config(name, seed=0, **over):
    f, kw = CONFIGS[name]
    kw = dict(kw)
    kw.update(over)
    return f(seed=seed, **kw)
