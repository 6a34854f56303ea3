"""Pins the oracle's numerics and pointwise physics against the reference's
golden vectors (tests/golden, made from the reference's compiled C sources)
and against what the mathematics fixes (FD, symmetry, invariances)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PAR = np.array([0.25, 0.50, 500., 250., 0.025, 1.0, 1.0, 1.e-3 * (1. / 2.), 0.0])
GRAD = np.array([s % 10 != 0 for s in range(30)])


def load(name):
    return np.load(os.path.join(GOLD, name))


def test_gauss_closed_form_vs_reference(oracle):
    g = load("basis.npz")
    for n in (1, 2, 3, 4):
        x, w = oracle.gauss(n)
        np.testing.assert_allclose(x, g["gauss_x%d" % n], rtol=0, atol=4e-16)
        np.testing.assert_allclose(w, g["gauss_w%d" % n], rtol=0, atol=4e-16)
        # exact for polynomials of degree <= 2n-1 on [0,1]
        for k in range(2 * n):
            assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-15


def test_bspline_vs_reference(oracle):
    g = load("basis.npz")
    for win, xq, n in zip(g["windows"], g["xq"], g["n"]):
        np.testing.assert_array_equal(oracle.eval_bspline(list(win), xq), n)


def test_bspline_invariants(oracle):
    rng = np.random.default_rng(3)
    k = np.array([0.0, 0.0, 0.0, 0.125, 0.25, 0.375])
    kint = np.array([0.125, 0.25, 0.375, 0.5, 0.625, 0.75])
    for win in ([k, kint, k], [kint, kint, kint]):
        for _ in range(5):
            xq = np.array([w[2] + rng.uniform(0.01, 0.99) * (w[3] - w[2]) for w in win])
            N = oracle.eval_bspline(win, xq)
            # partition of unity; derivatives of 1 vanish
            assert abs(N[:, 0].sum() - 1) < 1e-13
            assert np.abs(N[:, 1:].sum(axis=0)).max() < 1e-10
            # Marsden: sum_A greville_A N_A = x (per direction) and
            # sum_A ga gb N_A = x y for tensor products of Greville abscissae
            gre = [0.5 * (w[1:4] + w[2:5]) for w in win]
            G = np.array([[gre[0][a], gre[1][b], gre[2][c]] for a in range(3) for b in range(3) for c in range(3)])
            for d in range(3):
                assert abs(G[:, d] @ N[:, 0] - xq[d]) < 1e-13
                assert abs(G[:, d] @ N[:, 1 + d] - 1.0) < 1e-11
            assert abs((G[:, 0] * G[:, 1]) @ N[:, 0] - xq[0] * xq[1]) < 1e-13
            assert abs((G[:, 0] * G[:, 1]) @ N[:, 5] - 1.0) < 1e-10
            # FD of the value slot in x reproduces slot 1
            h = 1e-7
            e = np.array([h, 0, 0])
            fd = (oracle.eval_bspline(win, xq + e)[:, 0] - oracle.eval_bspline(win, xq - e)[:, 0]) / (2 * h)
            assert np.abs(fd - N[:, 1]).max() < 1e-6 * np.abs(N[:, 1]).max()


def test_evalN_vs_reference(oracle):
    g = load("basis.npz")
    kv = g["evalN_kv"]
    for k, i, xi, v in g["evalN"]:
        assert oracle.evalN(kv, int(k), int(i), 2, xi) == v


def test_iga_dirichlet_vs_reference_and_affine(oracle):
    g = load("dirichlet.npz")
    pd = g["par"]
    for rec in g["rec"]:
        n, f, o, i = (int(v) for v in rec[:4])
        kvn = [np.concatenate([[0, 0], np.linspace(0, 1, n + 1), [1, 1]])] * 3
        out = oracle.iga_dirichlet(pd, f, o, i, kvn, [n + 2] * 3)
        np.testing.assert_array_equal(out, rec[4:4 + out.size])
        # affine reproduction: g equals a . (Greville abscissae, face coordinate)
        gre = 0.5 * (kvn[0][1:-2] + kvn[0][2:-1])
        perm = {0: (1, 2, 0), 1: (0, 2, 1), 2: (0, 1, 2)}[f // 2]
        a = pd[3 * i:3 * i + 3][list(perm)]
        Z = 0.0 if f % 2 == 0 else 1.0 - 1e-15
        expect = (a[0] * gre[:, None] + a[1] * gre[None, :] + a[2] * Z).reshape(-1)
        assert np.abs(out - expect).max() < 1e-15


def test_ge_pointwise_vs_reference(oracle):
    g = load("pointwise_ge.npz")
    coef = oracle.coef_from_par(g["par"])
    for u, r, d, w in zip(g["u"], g["r"], g["d"], g["w"]):
        rr = oracle.ge_residual(u, coef)
        assert np.abs(rr - r).max() <= 1e-12 * np.abs(r).max() + 1e-14
        dd = oracle.ge_tangent(u, coef).reshape(-1)
        assert np.abs(dd - d).max() <= 1e-13 * np.abs(d).max()
        assert abs(oracle.ge_energy(u, coef) - w) <= 1e-12 * abs(w) + 1e-16


def test_getime_pointwise_vs_reference(oracle):
    g = load("pointwise_getime.npz")
    par = g["par"]
    coef, tp = oracle.coef_from_par(par), par[5:8]
    for u, r, d in zip(g["u"], g["r"], g["d"]):
        rr = oracle.getime_residual(u, coef, tp)
        # value slots (inertia) and gradient slots (averaged stress) separately
        assert np.abs(rr[~GRAD] - r[~GRAD]).max() <= 1e-15 * np.abs(r[~GRAD]).max()
        assert np.abs(rr[GRAD] - r[GRAD]).max() <= 1e-12 * np.abs(r[GRAD]).max() + 1e-14
        dd = oracle.getime_tangent(u, coef, tp)
        dg = d.reshape(30, 30)
        assert np.abs(dd[np.ix_(GRAD, GRAD)] - dg[np.ix_(GRAD, GRAD)]).max() <= 1e-13 * np.abs(dg[np.ix_(GRAD, GRAD)]).max()
        assert np.abs(dd - dg).max() <= 1e-15 * np.abs(dg).max()


def test_ge_residual_is_energy_gradient(oracle):
    coef = oracle.coef_from_par(PAR)
    rng = np.random.default_rng(11)
    for amp in (1e-2, 1e-1):
        u = rng.uniform(-1, 1, 30) * amp
        h = 1e-6
        fd = np.array([(oracle.ge_energy(u + h * e, coef) - oracle.ge_energy(u - h * e, coef)) / (2 * h)
                       for e in np.eye(30)])
        r = oracle.ge_residual(u, coef)
        assert np.abs(fd - r).max() < 1e-7 * np.abs(r).max()
        J = np.array([(oracle.ge_residual(u + h * e, coef) - oracle.ge_residual(u - h * e, coef)) / (2 * h)
                      for e in np.eye(30)]).T
        D = oracle.ge_tangent(u, coef)
        assert np.abs(J - D).max() < 1e-7 * np.abs(D).max()
        assert np.array_equal(D, D.T)
        assert not D[~GRAD].any() and not D[:, ~GRAD].any()
    assert not oracle.ge_residual(np.zeros(30), coef).any()


def test_ge_frame_indifference(oracle):
    """W(Q F, Q grad F) = W(F, grad F) for a rotation Q."""
    coef = oracle.coef_from_par(PAR)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, 30) * 0.05
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    F = np.eye(3) + np.array([[u[10 * a + 1 + J] for J in range(3)] for a in range(3)])
    Hs = np.array([[u[10 * a + 4 + h] for h in range(6)] for a in range(3)])
    F2, H2 = Q @ F, Q @ Hs
    v = np.zeros(30)
    for a in range(3):
        v[10 * a] = rng.normal()
        v[10 * a + 1:10 * a + 4] = F2[a] - (np.arange(3) == a)
        v[10 * a + 4:10 * a + 10] = H2[a]
    assert abs(oracle.ge_energy(v, coef) - oracle.ge_energy(u, coef)) < 1e-11 * abs(oracle.ge_energy(u, coef))


def test_getime_reduces_to_ge(oracle):
    """delta = 0 (u = u0): averaged stress = ge stress at (u0+u1)/2, gradient
    block of the tangent = ge tangent / 4; value rows closed form."""
    coef = oracle.coef_from_par(PAR)
    tp = PAR[5:8]
    rng = np.random.default_rng(7)
    u1, u0 = rng.uniform(-1, 1, 30) * 0.05, rng.uniform(-1, 1, 30) * 0.05
    ub = (u0 + u1) / 2
    r = oracle.getime_residual(np.concatenate([u0, u1, u0]), coef, tp)
    assert np.abs(r[GRAD] - oracle.ge_residual(ub, coef)[GRAD]).max() < 1e-13 * np.abs(r[GRAD]).max()
    for s in (0, 10, 20):
        assert abs(r[s] - tp[0] * (2 * u0[s] - 2 * u1[s]) / tp[2] ** 2) < 1e-9 * abs(r[s])
    D = oracle.getime_tangent(np.concatenate([u0, u1, u0]), coef, tp)
    np.testing.assert_allclose(D[np.ix_(GRAD, GRAD)], oracle.ge_tangent(ub, coef)[np.ix_(GRAD, GRAD)] / 4,
                               rtol=0, atol=1e-13 * np.abs(D[np.ix_(GRAD, GRAD)]).max())
    # FD Jacobian of the residual in the current slice
    u = u0 + rng.uniform(-1, 1, 30) * 0.02
    h = 1e-7
    J = np.array([(oracle.getime_residual(np.concatenate([u + h * e, u1, u0]), coef, tp)
                   - oracle.getime_residual(np.concatenate([u - h * e, u1, u0]), coef, tp)) / (2 * h)
                  for e in np.eye(30)]).T
    D = oracle.getime_tangent(np.concatenate([u, u1, u0]), coef, tp)
    assert np.abs(J - D)[np.ix_(GRAD, GRAD)].max() < 1e-6 * np.abs(D[np.ix_(GRAD, GRAD)]).max()


@pytest.mark.skipif(not __import__("oracle.oracle", fromlist=["x"]).ref_available(),
                    reason="oracle/_ref not built (needs the reference tree)")
def test_live_reference_random(oracle):
    """Fresh random inputs against the reference build itself."""
    coef = oracle.coef_from_par(PAR)
    rng = np.random.default_rng(99)
    for _ in range(20):
        u = rng.uniform(-1, 1, 30) * 0.08
        r = oracle.ref_ge_residual(u, PAR)
        assert np.abs(oracle.ge_residual(u, coef) - r).max() <= 1e-12 * np.abs(r).max()
        u90 = np.concatenate([u, u * 0.9, u * 0.8])
        r = oracle.ref_getime_residual(u90, PAR)
        assert np.abs(oracle.getime_residual(u90, coef, PAR[5:8]) - r)[GRAD].max() <= 1e-12 * np.abs(r[GRAD]).max()


def test_getime_energy_matches_reference_density(oracle):
    """phys/getime/density.c: W of the ge energy at (u0 + u1) / 2, against the
    compiled reference (live) at random states."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    import numpy as np
    par = np.array([0.25, 0.5, 500.0, 250.0, 0.025, 1.0, 1.0, 5e-4, 0.0])
    coef = oracle.coef_from_par(par)
    rng = np.random.default_rng(21)
    for amp in (1e-3, 1e-2, 1e-1):
        u90 = rng.uniform(-1, 1, 90) * amp
        for i in range(3):
            for blk in range(3):
                u90[30 * blk + 10 * i] = 0.0
        ub = 0.5 * (u90[30:60] + u90[60:90])
        e_or = oracle.ge_energy(ub, coef)
        e_ref = oracle.ref_getime_energy(u90, par)
        assert abs(e_or - e_ref) <= 1e-12 * max(1.0, abs(e_ref)), (amp, e_or, e_ref)


def test_density_outputs_match_golden(oracle):
    """All 13 ge density outputs (phys/ge/density.c:6-66: W, the stress-like
    outputs, the variant indicators) and both getime outputs (W at the mean
    state plus the kinetic energy of the value slots) against the compiled
    reference's vectors (tests/golden/density.npz)."""
    import numpy as np
    g = np.load(os.path.join(GOLD, "density.npz"))
    par = g["par"]
    coef = oracle.coef_from_par(par)
    for u, h in zip(g["u_ge"], g["hh_ge"]):
        d = oracle.density(0, u, coef)
        assert np.abs(d[:7] - h[:7]).max() <= 1e-12 * np.abs(h[:7]).max()
        assert np.array_equal(d[7:], h[7:])
    assert g["hh_ge"][:, 7:].any()           # the indicators are exercised
    for u, h in zip(g["u_getime"], g["hh_getime"]):
        d = oracle.density(1, u, coef, par[5:8])
        assert abs(d[0] - h[0]) <= 1e-12 * abs(h[0]) and d[1] == h[1] == 0.0
    # the kinetic term is present: value slots differ between the slices
    assert np.abs(g["u_getime"][:, 30] - g["u_getime"][:, 60]).max() > 0


def test_boundary_density_matches_golden():
    """densityb as or_intdensity applies it: ge hh[0] = u t, the rest 0
    (phys/ge/density.c:68-84); getime hh[0] = (u1 + u0) / 2 t (slices 1 and
    2), hh[1] = 1 (phys/getime/density.c:57-62)."""
    import numpy as np
    g = np.load(os.path.join(GOLD, "density.npz"))
    ub, t = g["ub"], g["traction"]
    assert np.array_equal(g["hhb_ge"][:, 0], ub[:, 0] * t) and not g["hhb_ge"][:, 1:].any()
    assert np.array_equal(g["hhb_getime"][:, 0], (ub[:, 1] + ub[:, 2]) / 2 * t)
    assert np.all(g["hhb_getime"][:, 1] == 1.0)
