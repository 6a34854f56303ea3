"""Problem description for residual/tangent assembly of C1 quadratic B-spline
gradient elasticity on a structured (tensor-product) patch.

Mirrors what the reference's structs carry for this path:
  Nurbs (include/nurbs.h:4-24): porder 2, nelem[3], nbasis[3], nknot[3],
      knotVector[3], nboun/boun
  Phys  (include/phys.h:11-26): ndof 3, torder 0 (ge) / 2 (getime), par_mat[9]
  BC    (include/bc.h:4-11): bc_type[36], par_periodic[9], par_dirichlet[9],
      par_neumann
plus what the device path adds: nquad (the reference hard-codes 4,
src/snesiga_Residual.c:99) and optional per-element material coefficients.

Material coefficients: W is linear in coef = (A36, A37, A38, e0, e345, le^2);
the reference computes the first three from par_mat r, curv
(phys/ge/residual.c:45-47).
"""
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

PORDER = 2
NDOF = 3
GE = 0
GETIME = 1

# r, curv, e0, e345, le, rho, cii, dt, t  (phys/ge/init.c:18, phys/getime/init.c:18)
PAR_MAT_REFERENCE = np.array([0.25, 0.50, 500., 250., 0.025, 1.0, 1.0, 1.e-3 * (1. / 2.), 0.0])


def uniform_knots(nelem):
    """Open uniform knots on [0,1] with p = 2 (nurbs_genuniform.c:34-38), any nelem."""
    k = np.empty(nelem + 1 + 2 * PORDER)
    k[:PORDER + 1] = 0.0
    for i in range(PORDER + 1, nelem + PORDER):
        k[i] = (i - PORDER) * (1.0 / nelem)
    k[nelem + PORDER:] = 1.0
    return k


def coef_from_par(par):
    """(A36, A37, A38, e0, e345, le^2) from the reference par_mat."""
    r, c = par[0], par[1]
    return np.array([4.0 * (1.0 - c) / (r * r * r * (1.0 + 2.0 * c)),
                     -6.0 * c / (r * r * (1.0 + 2.0 * c)),
                     3.0 / (r * r * r * r * (1.0 + 2.0 * c)),
                     par[2], par[3], par[4] * par[4]])


def coef_from_lame(lam, mu, ell):
    """St-Venant-Kirchhoff strain-gradient variant: W = (3 lam/2 + mu) e1^2 +
    mu (e2^2 + e3^2) + 2 mu (E01^2 + E02^2 + E12^2) + l^2 |grad dev|^2.
    Vectorised over arrays of lam, mu, ell.  (Not reference physics: the
    reference only has the par_mat double-well.)"""
    lam, mu, ell = np.broadcast_arrays(np.asarray(lam, float), np.asarray(mu, float), np.asarray(ell, float))
    z = np.zeros_like(lam)
    return np.stack([z, mu, z, 1.5 * lam + mu, 2.0 * mu, ell * ell], axis=-1)


def splitmix64(seed, n):
    """n uint64 draws of splitmix64 starting from state `seed`."""
    M = (1 << 64) - 1
    out = np.empty(n, np.uint64)
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = (np.uint64(seed & M) + idx * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        out[:] = z ^ (z >> np.uint64(31))
    return out


def uniform01(seed, n):
    """Uniform doubles in [0,1) from the top 53 bits of splitmix64."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


@dataclass
class Problem:
    nelem: tuple
    knots: list
    physics: int = GE
    nquad: int = 4
    par_mat: np.ndarray = field(default_factory=lambda: PAR_MAT_REFERENCE.copy())
    elem_coef: Optional[np.ndarray] = None
    boun: tuple = (0, 1, 2, 3, 4, 5)
    bc_type: np.ndarray = field(default_factory=lambda: np.ones(36, np.int32))
    par_dirichlet: np.ndarray = field(default_factory=lambda: np.zeros(9))
    par_periodic: np.ndarray = field(default_factory=lambda: np.zeros(9))
    par_neumann: Optional[np.ndarray] = None
    neumann: Optional[Callable] = None

    @property
    def torder(self):
        return 2 if self.physics == GETIME else 0

    @property
    def nhist(self):
        return self.torder + 1

    @property
    def nbasis(self):
        return tuple(n + PORDER for n in self.nelem)

    @property
    def n_elem(self):
        return int(self.nelem[0]) * int(self.nelem[1]) * int(self.nelem[2])

    @property
    def ndof(self):
        nb = self.nbasis
        return NDOF * nb[0] * nb[1] * nb[2]

    @property
    def nnz(self):
        def s(nb):
            return sum(min(a + PORDER, nb - 1) - max(a - PORDER, 0) + 1 for a in range(nb))
        nb = self.nbasis
        return NDOF * NDOF * s(nb[0]) * s(nb[1]) * s(nb[2])

    def coef(self):
        """Per-element coefficients (n_elem, 6); the reference's global ones if unset."""
        if self.elem_coef is not None:
            return np.ascontiguousarray(self.elem_coef, dtype=np.float64)
        return np.tile(coef_from_par(self.par_mat), (self.n_elem, 1))

    def greville(self):
        """Greville abscissae per direction (control-point positions of x)."""
        return [0.5 * (k[1:-2] + k[2:-1]) for k in self.knots]


def bc_ex1():
    """example/ex1/main.c:73-95: all faces u = 0 (standard Dirichlet),
    high-order Neumann with zero traction."""
    bc = np.zeros(36, np.int32)
    for f in range(6):
        bc[6 * f + 3:6 * f + 6] = 1
    return bc


def bc_clamp_stretch():
    """Face 0 clamped (standard and high-order Dirichlet), face 1 u_x
    prescribed with free tangential motion and high-order Neumann, faces 2-5
    Neumann (zero traction)."""
    bc = np.ones(36, np.int32)
    bc[0:6] = 0             # face 0: ux, uy, uz and their normal gradients
    bc[6 + 0] = 0           # face 1: ux Dirichlet; uy, uz and high order Neumann
    return bc


def bc_all_clamped():
    return np.zeros(36, np.int32)


def reference_example(n, physics=GE, nquad=4):
    """The reference's example (ex1): unit cube, n^3 elements, all faces
    standard-Dirichlet zero, high-order Neumann zero, reference par_mat."""
    return Problem(nelem=(n, n, n), knots=[uniform_knots(n) for _ in range(3)], physics=physics,
                   nquad=nquad, bc_type=bc_ex1())


def stretch_problem(n, a=0.01, lam=1.0, mu=0.5, ell=0.1, noise=0.0, seed=0x5EED0002, nquad=4,
                    material="lame"):
    """Clamped at x=0, u_x = a X imposed at x=1 (par_dirichlet[0] = a), the
    rest traction free.  material='lame' maps (lam, mu, ell) per element with
    optional +-noise relative perturbation; material='reference' uses par_mat."""
    n3 = (n, n, n) if np.isscalar(n) else tuple(n)
    pb = Problem(nelem=n3, knots=[uniform_knots(m) for m in n3], nquad=nquad, bc_type=bc_clamp_stretch())
    pb.par_dirichlet = np.zeros(9)
    pb.par_dirichlet[0] = a
    if material == "lame":
        ne = pb.n_elem
        if noise:
            r = uniform01(seed, 2 * ne).reshape(2, ne)
            lam_e = lam * (1.0 + noise * (2.0 * r[0] - 1.0))
            mu_e = mu * (1.0 + noise * (2.0 * r[1] - 1.0))
        else:
            lam_e = np.full(ne, lam)
            mu_e = np.full(ne, mu)
        pb.elem_coef = coef_from_lame(lam_e, mu_e, np.full(ne, ell))
    return pb


def gaussian01(seed, n):
    """n standard normal draws (Box-Muller on splitmix64 uniforms)."""
    u = uniform01(seed, 2 * n).reshape(n, 2)
    u1 = np.maximum(u[:, 0], 1.0 / (1 << 53))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[:, 1])


def gaussian_random_field(coords, corr_len, seed, std=1.0):
    """Zero-mean Gaussian random field on the tensor grid coords (3 1D
    arrays): white noise (gaussian01) smoothed by a separable Gaussian kernel
    exp(-d^2 / corr_len^2) per direction, so the covariance is proportional
    to exp(-d^2 / (2 corr_len^2)); then shifted and scaled to sample mean 0
    and sample standard deviation std."""
    shape = tuple(len(c) for c in coords)
    f = gaussian01(seed, int(np.prod(shape))).reshape(shape)
    for d, c in enumerate(coords):
        c = np.asarray(c, float)
        Kd = np.exp(-((c[:, None] - c[None, :]) / corr_len) ** 2)
        f = np.moveaxis(np.tensordot(Kd, np.moveaxis(f, d, 0), axes=(1, 0)), 0, d)
    f = f - f.mean()
    return f * (std / f.std())


def element_centres(pb):
    return [0.5 * (k[2:-3] + k[3:-2]) for k in pb.knots]


def young_to_coef(E, nu, ell):
    """Per-element coefficients of the Lame mapping for Young's modulus E and
    Poisson ratio nu: lam = E nu / ((1 + nu)(1 - 2 nu)), mu = E / (2 (1 + nu))."""
    E = np.asarray(E, float)
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    return coef_from_lame(lam, mu, np.full(E.shape, ell))


def dcoef_dyoung(nu):
    """d(coef)/dE of young_to_coef (constant: the mapping is linear in E)."""
    dl = nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    dm = 1.0 / (2.0 * (1.0 + nu))
    return np.array([0.0, dm, 0.0, 1.5 * dl + dm, 2.0 * dm, 0.0])


def lognormal_young(pb, sigma=0.3, corr_len=0.25, seed=0x5EED0005):
    """E_e = exp(-sigma^2/2 + sigma Z_e) (mean 1), Z a unit Gaussian random
    field at the element centres (elements in reference order)."""
    Z = gaussian_random_field(element_centres(pb), corr_len, seed).reshape(-1)
    return np.exp(-0.5 * sigma * sigma + sigma * Z)


def adjoint_problem(n=32, sigma=0.3, corr_len=0.25, nu=0.3, ell=0.1, a=0.01, seed=0x5EED0005,
                    target_seed=0x5EED0006, nquad=4):
    """Material calibration problem on n^3 elements: clamp/stretch loading,
    per-element Young's modulus from a log-normal field (seed), Poisson ratio
    nu, length scale ell.  Returns (pb, E, E_target); the target displacement
    is the solution for E_target (second field, target_seed)."""
    pb = stretch_problem(n, a=a, nquad=nquad)
    E = lognormal_young(pb, sigma, corr_len, seed)
    Et = lognormal_young(pb, sigma, corr_len, target_seed)
    pb.elem_coef = young_to_coef(E, nu, ell)
    return pb, E, Et


def transient_problem(n=48, rho=1.0, damping=0.0, dt=1e-3, lam=1.0, mu=0.5, ell=0.05, nquad=4):
    """phys/getime on n^3 elements, all faces clamped (standard and high-order
    Dirichlet zero), uniform Lame coefficients; par_mat[5..7] = (rho, c, dt)
    as the getime kernels read them (phys/getime/residual.c:4127)."""
    pb = reference_example(n, physics=GETIME, nquad=nquad)
    pb.bc_type = bc_all_clamped()
    pb.par_mat = np.array(pb.par_mat, dtype=np.float64)
    pb.par_mat[5], pb.par_mat[6], pb.par_mat[7] = rho, damping, dt
    pb.elem_coef = coef_from_lame(np.full(pb.n_elem, lam), np.full(pb.n_elem, mu), np.full(pb.n_elem, ell))
    return pb


def clamped_mask(pb, layers=2):
    """True on the control points of the first/last `layers` layers of every
    face (the rows an all-faces clamp fixes), per dof."""
    nb = pb.nbasis
    m = np.zeros(tuple(nb), bool)
    m[:layers], m[-layers:] = True, True
    m[:, :layers], m[:, -layers:] = True, True
    m[:, :, :layers], m[:, :, -layers:] = True, True
    return np.repeat(m.reshape(-1), 3)


def initial_velocity(pb, std=1e-3, corr_len=0.2, seed=0x5EED0003):
    """Velocity at the control points: one Gaussian random field per
    component (seeds seed, seed+1, seed+2; zero mean, standard deviation std,
    correlation length corr_len), zero on the clamped layers."""
    g = pb.greville()
    comps = [gaussian_random_field(g, corr_len, seed + i, std=std).reshape(-1) for i in range(3)]
    v = np.stack(comps, axis=-1).reshape(-1)
    v[clamped_mask(pb)] = 0.0
    return v


def affine_guess(pb, nhist=None):
    """Control values of the affine field u_i = sum_J a_iJ X_J with
    a_iJ = par_dirichlet[3 i + J] (the field src/iga_dirichlet.c:42-44
    imposes on the Dirichlet faces) at every control point: a smooth initial
    guess that meets the Dirichlet rows of faces whose data it is (quadratic
    B-splines reproduce affine fields with control values at the Greville
    abscissae).  History slices (getime) are copies."""
    g = pb.greville()
    A = np.asarray(pb.par_dirichlet, dtype=np.float64).reshape(3, 3)
    X = np.stack(np.meshgrid(*g, indexing="ij"), axis=-1)       # (nb0, nb1, nb2, 3)
    U = np.einsum("abcJ,iJ->abci", X, A).reshape(-1)
    nh = pb.nhist if nhist is None else nhist
    return np.tile(U, nh)


def smooth_state(pb, amp=1e-3, noise=1e-6, seed=0x5EED0012, nhist=None):
    """Non-equilibrium displacement at the control points:
    U_i = amp sin(pi X) sin(pi Y) sin(pi Z) (1, .5, .25)_i + noise U(-1,1),
    in reference layout (dof fastest, then z, y, x).  History slices (getime)
    are perturbed copies."""
    g = pb.greville()
    sx, sy, sz = (np.sin(np.pi * gi) for gi in g)
    base = amp * (sx[:, None, None] * sy[None, :, None] * sz[None, None, :])
    U = base[..., None] * np.array([1.0, 0.5, 0.25])
    U = U.reshape(-1)
    nh = pb.nhist if nhist is None else nhist
    out = np.empty(nh * U.size)
    r = uniform01(seed, nh * U.size)
    for h in range(nh):
        out[h * U.size:(h + 1) * U.size] = U * (1.0 - 0.1 * h) + noise * (2.0 * r[h * U.size:(h + 1) * U.size] - 1.0)
    return out
