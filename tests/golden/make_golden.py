"""Writes the golden vectors under tests/golden/ from the reference's own
libm-only C sources, compiled by oracle/Makefile into oracle/_ref/ (run in a
container that has the reference tree; the vectors are then committed).

    make -C oracle ref && python tests/golden/make_golden.py

Covers (reference file:line):
  pointwise_ge.npz      phys/ge/residual.c:6-78, tangent.c:6-948, density.c:49
  pointwise_getime.npz  phys/getime/residual.c:6-4157, tangent.c:6-5027
  basis.npz             mesh/nurbs/eval_Bspline.c:9-93, evalN.c:7-62,
                        src/quadrature.c:381-429
  dirichlet.npz         src/iga_dirichlet.c:7-80
  density.npz           phys/ge/density.c:6-84, phys/getime/density.c:6-62
Inputs are seeded; the reference parameters are phys/ge/init.c:18.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

PAR = np.array([0.25, 0.50, 500., 250., 0.025, 1.0, 1.0, 1.e-3 * (1. / 2.), 0.0])


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libigap4ref.so missing: make -C oracle ref")
    rng = np.random.default_rng(0x601DE)
    amps = np.repeat([1e-3, 1e-2, 5e-2, 1e-1], 16)
    # ---- ge
    U = rng.uniform(-1, 1, (amps.size, 30)) * amps[:, None]
    R = np.array([O.ref_ge_residual(u, PAR) for u in U])
    D = np.array([O.ref_ge_tangent(u, PAR).reshape(-1) for u in U])
    W = np.array([O.ref_ge_energy(u, PAR) for u in U])
    np.savez_compressed(os.path.join(HERE, "pointwise_ge.npz"), par=PAR, u=U, r=R, d=D, w=W)
    # ---- getime: u1, u0 around a common state, u a small step away
    base = rng.uniform(-1, 1, (amps.size, 30)) * amps[:, None]
    U90 = np.concatenate([base + rng.uniform(-1, 1, (amps.size, 30)) * amps[:, None] * 0.3,
                          base + rng.uniform(-1, 1, (amps.size, 30)) * amps[:, None] * 0.1,
                          base], axis=1)
    R = np.array([O.ref_getime_residual(u, PAR) for u in U90])
    D = np.array([O.ref_getime_tangent(u, PAR).reshape(-1) for u in U90])
    np.savez_compressed(os.path.join(HERE, "pointwise_getime.npz"), par=PAR, u=U90, r=R, d=D)
    # ---- basis on uniform and non-uniform windows (open ends included)
    wins, xqs, Ns = [], [], []
    for t in range(48):
        kw = []
        for d in range(3):
            kind = rng.integers(3)
            if kind == 0:      # first element of an open knot vector
                h = rng.uniform(0.05, 0.3)
                k = np.array([0, 0, 0, h, 2 * h, 3 * h])
            elif kind == 1:    # last element
                h = rng.uniform(0.05, 0.3)
                k = np.array([1 - 3 * h, 1 - 2 * h, 1 - h, 1, 1, 1])
            else:              # interior, non-uniform
                k = np.cumsum(rng.uniform(0.05, 0.3, 6))
            kw.append(k)
        xq = np.array([kw[d][2] + rng.uniform(0.02, 0.98) * (kw[d][3] - kw[d][2]) for d in range(3)])
        wins.append(np.stack(kw))
        xqs.append(xq)
        Ns.append(O.ref_eval_bspline(kw, xq))
    kv = np.array([0, 0, 0, .125, .3, .5, .75, .8, 1, 1, 1.])
    ev = []
    for k in range(3):
        for i in range(len(kv) - 3):
            for xi in np.concatenate([np.linspace(0, 1, 21), [1.0 - 1e-15]]):
                ev.append((k, i, xi, O.ref_evalN(kv, k, i, 2, xi)))
    g = {n: O.ref_gauss(n) for n in (1, 2, 3, 4)}
    np.savez_compressed(os.path.join(HERE, "basis.npz"), windows=np.array(wins), xq=np.array(xqs),
                        n=np.array(Ns), evalN_kv=kv, evalN=np.array(ev),
                        **{"gauss_x%d" % n: g[n][0] for n in g}, **{"gauss_w%d" % n: g[n][1] for n in g})
    # ---- iga_dirichlet on two meshes, every face/order/dof
    pd = np.array([0.01, 0.002, -0.003, 0.004, 0.005, -0.006, 0.007, 0.008, 0.009])
    recs = []
    for n in (4, 7):
        kvn = [np.concatenate([[0, 0], np.linspace(0, 1, n + 1), [1, 1]])] * 3
        nb = [n + 2] * 3
        for f in range(6):
            for o in range(2):
                for i in range(3):
                    recs.append(np.concatenate([[n, f, o, i], O.ref_iga_dirichlet(pd, f, o, i, kvn, nb)]))
    width = max(len(r) for r in recs)
    arr = np.full((len(recs), width), np.nan)
    for j, r in enumerate(recs):
        arr[j, :len(r)] = r
    np.savez_compressed(os.path.join(HERE, "dirichlet.npz"), par=pd, rec=arr)
    # ---- every density output of ge and getime, value slots included, and
    # the boundary densities (own generator: the draws above are unchanged)
    rd = np.random.default_rng(0xDE75)
    amps = np.repeat([1e-3, 1e-2, 5e-2, 1e-1, 3e-1], 12)
    Uge = rd.uniform(-1, 1, (amps.size, 30)) * amps[:, None]
    Hge = np.array([O.ref_ge_density(u, PAR) for u in Uge])
    Ugt = rd.uniform(-1, 1, (amps.size, 90)) * amps[:, None]
    Hgt = np.array([O.ref_getime_density(u, PAR) for u in Ugt])
    nb = 24
    UB = rd.uniform(-1, 1, (nb, 3))
    T = rd.uniform(-2, 2, nb)
    ORD = rd.integers(0, 2, nb)
    HBge = np.array([O.ref_densityb(0, UB[k, :1], T[k], PAR, int(ORD[k])) for k in range(nb)])
    HBgt = np.array([O.ref_densityb(1, UB[k], T[k], PAR, int(ORD[k])) for k in range(nb)])
    np.savez_compressed(os.path.join(HERE, "density.npz"), par=PAR, u_ge=Uge, hh_ge=Hge, u_getime=Ugt,
                        hh_getime=Hgt, ub=UB, traction=T, order=ORD, hhb_ge=HBge, hhb_getime=HBgt)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
