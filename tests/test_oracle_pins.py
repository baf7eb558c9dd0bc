"""Pins of the CPU oracle (oracle/phoracle.c).

1. Bitwise against the reference's own compiled Fortran: committed golden
   vectors from oracle/_ref (tests/golden/make_golden.py) and, where
   oracle/_ref is built, live calls of ElmGMRs, e3, i3LU, SparseAp,
   Spsi3pre, bc3per and itrPredict/itrCorrect/itrUpdate/itrBC.
2. What the mathematics fixes for the parts oracle/_ref does not build
   (SolGMRs' GMRES loop) and as independent checks of the rest: quadrature
   rules, a zero residual and an exact tangent (= finite-difference
   Jacobian) at a uniform state on a fully periodic box, the block-LU
   factor identity, GMRES = exact solve on the full Krylov space and =
   Krylov least squares below it, the minIters/etol stopping rule.
"""
import os

import numpy as np
import pytest

import pyoracle
import pyref
from phasta_amd import config, mesh

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# alpha factors the itrUpdate fixtures were made with (tests/golden/make_golden.py)
UPDATE_ALPHA = dict(almi=0.75, alfi=0.6, gami=0.65)


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


# ---------------------------------------------------------------- golden
@pytest.mark.parametrize("N", [2, 3])
def test_oracle_bitwise_equals_reference_golden(N):
    g = np.load(os.path.join(GOLD, f"ref_box{N}.npz"))
    case = mesh.box_case(N, mach=0.3, kspace=10)
    o = pyoracle.Oracle(case)
    y, ac = o.itr_predict_bc(case.yold)
    assert np.array_equal(y, g["y"])
    res, rmes, bd, lk = o.elmgmrs(g["y"], g["ac"])
    assert np.array_equal(res, g["res"])
    assert np.array_equal(bd, g["BDiag"])
    assert np.array_equal(lk, g["lhsK"])
    assert not rmes.any()
    lu = bd.copy(order="F")
    o.i3lu(lu, None, "LU_Fact ")
    assert np.array_equal(lu, g["BDiag_lu"])
    ap = g["p"].copy(order="F")
    o.sparse_ap(np.asfortranarray(g["lhsK"]), ap)
    assert np.array_equal(ap, g["Ap"])
    ls = np.array(g["lhsK"], order="F")
    o.spsi3pre(np.asfortranarray(g["BDiag_lu"]), ls)
    assert np.array_equal(ls, g["lhsK_scaled"])
    y1, ac1 = o.itr_correct_bc(g["y"], g["ac"], case.yold, np.zeros_like(g["ac"]),
                               np.asfortranarray(g["dy"]))
    assert np.array_equal(y1, g["y_corr"])
    assert np.array_equal(ac1, g["ac_corr"])
    ou = pyoracle.Oracle(case, params=dict(case.params, **UPDATE_ALPHA))
    yo, aco = ou.itr_update_bc(case.yold, np.asfortranarray(g["acold"]), g["y_corr"],
                               g["ac_corr"])
    assert np.array_equal(yo, g["yold_upd"])
    assert np.array_equal(aco, g["acold_upd"])


# ------------------------------------------------- generalized alpha
def test_time_integration_equals_reference_itrsetup():
    """config.time_integration against the reference's compiled itrSetup
    (fixture tests/golden/ref_alpha_box3.npz): the factors bitwise, the
    predictor it settles on, and a zero ipred = 4 start-up shift."""
    g = np.load(os.path.join(GOLD, "ref_alpha_box3.npz"))
    for r, fac, ipo in zip(g["rhoinf"], g["factors"], g["ipred"]):
        for k, ip in enumerate((1, 4)):
            t = config.time_integration(float(r), ip)
            assert (t["almi"], t["alfi"], t["gami"]) == tuple(fac[k]), (r, ip)
            assert t["ipred"] == ipo[k], (r, ip)
    assert not g["shift"].any()


def _amplification(rhoinf, lam_dt):
    """Amplification matrix of the generalized-alpha scheme on y' = lam y
    in the variables (y, y') (Jansen, Whiting and Hulbert, CMAME 190, 2000):
    ydot_{n+am} = lam y_{n+af},  y_{n+1} = y_n + dt ydot_n + gam dt (ydot_{n+1} - ydot_n)."""
    t = config.time_integration(rhoinf)
    am, af, gm = t["almi"], t["alfi"], t["gami"]
    # unknowns (y1, v1) with v = dt y'; rows: the two relations above
    A = np.array([[-lam_dt * af, am], [1.0, -gm]])
    B = np.array([[lam_dt * (1 - af), -(1 - am)], [1.0, 1.0 - gm]])
    return np.linalg.solve(A, B)


@pytest.mark.parametrize("rhoinf", [0.0, 0.25, 0.5, 1.0])
def test_generalized_alpha_spectral_radius_and_order(rhoinf):
    """What the mathematics fixes for the factors: the spectral radius of
    the amplification matrix at lam dt -> -inf is rho_inf, the scheme is
    second-order accurate (one-step error against exp(lam dt) falls 8x per
    halving of dt) and A-stable on the negative real axis."""
    # (at rho_inf = 0 the double eigenvalue 0 is a Jordan block: the radius
    # falls as |lam dt|^-1/2, 3e-5 here)
    G = _amplification(rhoinf, -1e9)
    assert abs(max(abs(np.linalg.eigvals(G))) - rhoinf) < 1e-4
    errs = []
    for h in (0.1, 0.05, 0.025):
        G = _amplification(rhoinf, -h)
        errs.append(min(abs(np.linalg.eigvals(G) - np.exp(-h))))
    assert 7.0 < errs[0] / errs[1] < 9.0 and 7.0 < errs[1] / errs[2] < 9.0
    for ld in (-1e-2, -1.0, -1e2, -1e4):
        assert max(abs(np.linalg.eigvals(_amplification(rhoinf, ld)))) <= 1.0 + 1e-12


def test_generalized_alpha_step_equals_reference_golden():
    """itrPredict + itrBC and itrCorrect + itrBC at rho_inf = 0.5 from a
    nonzero acold, bitwise against the reference's compiled routines; the
    oracle refuses the predictors the reference leaves undefined."""
    g = np.load(os.path.join(GOLD, "ref_alpha_box3.npz"))
    case = mesh.box_case(3, mach=0.3, kspace=10)
    o = pyoracle.Oracle(case, params=dict(case.params, **config.time_integration(0.5)))
    acold = np.asfortranarray(g["acold"])
    y, ac = o.itr_predict_bc(case.yold, acold)
    assert np.array_equal(y, g["y"]) and np.array_equal(ac, g["ac"])
    y1, ac1 = o.itr_correct_bc(y, ac, case.yold, acold, np.asfortranarray(g["dy"]))
    assert np.array_equal(y1, g["y_corr"]) and np.array_equal(ac1, g["ac_corr"])
    for ip in (2, 3, 4):
        with pytest.raises(RuntimeError):
            pyoracle.Oracle(case, params=dict(case.params, ipred=ip)).itr_predict_bc(case.yold)


# ------------------------------ global reconstruction of the diffusive flux
def test_idiff_elmgmrs_equals_reference_golden():
    """ElmGMRs with idiff 1 and 3 (AsIq, e3q, qpbc and the divergence term
    of e3ivar/e3LS) bitwise against the reference's compiled routines
    (fixture tests/golden/ref_idiff_box3.npz); the tangent does not change."""
    g = np.load(os.path.join(GOLD, "ref_idiff_box3.npz"))
    case = mesh.box_case(3, mach=0.5, kspace=10, jitter=0.1)
    y, ac = np.asfortranarray(g["y"]), np.asfortranarray(g["ac"])
    for idiff in (1, 3):
        o = pyoracle.Oracle(case, params=dict(case.params, idiff=idiff))
        res, rmes, bd, lk = o.elmgmrs(y, ac)
        assert np.array_equal(res, g[f"res{idiff}"]), idiff
        assert np.array_equal(bd, g["BDiag"]) and np.array_equal(lk, g["lhsK"])
    res0 = pyoracle.Oracle(case).elmgmrs(y, ac)[0]
    assert rel(g["res1"], res0) > 1e-9      # the term is there
    with pytest.raises(RuntimeError):
        pyoracle.Oracle(case, params=dict(case.params, idiff=2)).elmgmrs(y, ac)


@pytest.mark.skipif(not pyref.available(), reason="oracle/_ref not built")
def test_idiff_elmgmrs_bitwise_equals_live_reference():
    case = mesh.box_case(4, mach=0.3, kspace=10, jitter=0.1, grade=1.1)
    rng = np.random.default_rng(8)
    o0 = pyoracle.Oracle(case)
    y, ac = o0.itr_predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.02 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    for idiff in (1, 3):
        prm = dict(case.params, idiff=idiff)
        for a, b in zip(pyoracle.Oracle(case, params=prm).elmgmrs(y, ac),
                        pyref.Reference(case, params=prm).elmgmrs(y, ac)):
            assert np.array_equal(a, b), idiff


# ------------------------------------------------ discontinuity capturing
def dc_case():
    return mesh.box_case(3, mach=0.5, kspace=10, jitter=0.1)


def test_dc_elmgmrs_equals_reference_golden():
    """ElmGMRs with iDC 1, 2, 3 (e3dc.f, with e3mtrx.f's A0DC/A0inv/dVdY and
    e3tau.f's rTLS/raLS/giju) bitwise against the reference's compiled
    routines (fixture tests/golden/ref_dc_box3.npz); DC changes the residual
    and the tangent, and the three variants differ."""
    g = np.load(os.path.join(GOLD, "ref_dc_box3.npz"))
    case = dc_case()
    y, ac = np.asfortranarray(g["y"]), np.asfortranarray(g["ac"])
    base = pyoracle.Oracle(case).elmgmrs(y, ac)
    for idc in (1, 2, 3):
        res, rmes, bd, lk = pyoracle.Oracle(case, params=dict(case.params, idc=idc)).elmgmrs(y, ac)
        assert np.array_equal(res, g[f"res{idc}"]), idc
        assert np.array_equal(bd, g[f"BDiag{idc}"]), idc
        assert np.array_equal(lk, g[f"lhsK{idc}"]), idc
        assert rel(res, base[0]) > 1e-6 and rel(lk, base[3]) > 1e-6, idc
    assert rel(g["res1"], g["res2"]) > 1e-9 and rel(g["res1"], g["res3"]) > 1e-9
    with pytest.raises(RuntimeError):
        pyoracle.Oracle(case, params=dict(case.params, idc=4)).elmgmrs(y, ac)


@pytest.mark.skipif(not pyref.available(), reason="oracle/_ref not built")
def test_dc_elmgmrs_bitwise_equals_live_reference():
    case = mesh.box_case(4, mach=0.3, kspace=10, jitter=0.1, grade=1.1)
    rng = np.random.default_rng(9)
    y, ac = pyoracle.Oracle(case).itr_predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.05 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    for idc in (1, 2, 3):
        prm = dict(case.params, idc=idc)
        for a, b in zip(pyoracle.Oracle(case, params=prm).elmgmrs(y, ac),
                        pyref.Reference(case, params=prm).elmgmrs(y, ac)):
            assert np.array_equal(a, b), idc


def test_asiq_reconstructs_linear_fields_exactly():
    """What the mathematics fixes for AsIq + qpbc: on a uniform box (every
    node patch point-symmetric) the lumped L2 projection of a piecewise
    linear flux is exact at the nodes off the walls, so the shear flow
    u = a y + b, T = T0 + c y gives tau_12 = tau_21 = mu a and the y
    energy flux u tau_21 + k T,y = (a y + b) mu a + k c, every other
    component zero; the lumped masses sum to the box volume."""
    case = mesh.box_case(4, jitter=0.0)
    a, b, c = 37.0, 11.0, 25.0
    x = case.x
    y = np.zeros((case.nshg, 5), order="F")
    y[:, 0] = a * x[:, 1] + b
    y[:, 3] = 1e5
    y[:, 4] = 300.0 + c * x[:, 1]
    for idiff in (1, 3):
        o = pyoracle.Oracle(case, params=dict(case.params, idiff=idiff))
        qres, rmass = o.asiq(y)
        mu = config.MU
        kap = mu * (config.RGAS * config.GAMMA / config.GAMMA1) / config.PRANDTL
        exact = np.zeros_like(qres)
        exact[:, 1] = exact[:, 4] = mu * a
        exact[:, 7] = (a * x[:, 1] + b) * mu * a + kap * c
        inner = (x[:, 1] > 1e-9) & (x[:, 1] < 1 - 1e-9)
        err = np.abs(qres[inner] - exact[inner]).max() / np.abs(exact).max()
        assert err < 1e-12, (idiff, err)
        slave = (case.iBC & 1024) != 0
        assert abs(np.sum(1.0 / rmass[~slave]) - 1.0) < 1e-12


@pytest.mark.skipif(not pyref.available(), reason="oracle/_ref not built")
def test_oracle_bitwise_equals_live_reference():
    case = mesh.box_case(4, mach=0.5, kspace=10)
    o, r = pyoracle.Oracle(case), pyref.Reference(case)
    y, ac = o.itr_predict_bc(case.yold)
    y2, ac2 = r.predict_bc(case.yold)
    assert np.array_equal(y, y2) and np.array_equal(ac, ac2)
    rng = np.random.default_rng(5)
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    for a, b in zip(o.elmgmrs(y, ac), r.elmgmrs(y, ac)):
        assert np.array_equal(a, b)
    acold = np.asfortranarray(1e2 * rng.standard_normal(ac.shape))
    for a, b in zip(o.itr_update_bc(case.yold, acold, y, ac),
                    r.update_bc(case.yold, acold, y, ac)):
        assert np.array_equal(a, b)
    # element kernel on perturbed, non-periodic-consistent states
    for e in rng.choice(case.nel, 8, replace=False):
        nd = case.ien[e] - 1
        yl = np.array([[y[g, 3], y[g, 0], y[g, 1], y[g, 2], y[g, 4]] for g in nd])
        yl *= 1 + 0.05 * rng.standard_normal(yl.shape)
        acl = 1e3 * rng.standard_normal((4, 5))
        rl1, eg1 = o.e3(yl, acl, case.x[nd])
        rl2, eg2 = r.e3(yl, acl, case.x[nd])
        assert np.array_equal(rl1, rl2) and np.array_equal(eg1, eg2)


# ----------------------------------------------------------- quadrature
def test_quadrature_rules():
    import ctypes
    shp = np.zeros((4, 4)); shgl = np.zeros((4, 3, 4)); qwt = np.zeros(4)
    shpb = np.zeros((3, 4)); shglb = np.zeros((3, 3, 4)); qwtb = np.zeros(3)
    dp = ctypes.POINTER(ctypes.c_double)
    pyoracle.lib().orc_quadrature(*(a.ctypes.data_as(dp) for a in
                                    (shp, shgl, qwt, shpb, shglb, qwtb)))
    assert abs(qwt.sum() - 4.0 / 3.0) < 1e-15          # reference tet volume 1/6 * 8
    assert np.allclose(shp.sum(1), 1.0, atol=1e-15)
    assert np.allclose(shgl.sum(2), 0.0)
    # the 4-point rule is exact for quadratics: int N_a N_b = (1+d_ab)/20 * V
    M = (shp.T * qwt) @ shp / qwt.sum()
    assert np.allclose(M, (np.ones((4, 4)) + np.eye(4)) / 20, atol=1e-15)
    # wall rule: 3 points on the face (4th shape function ~0), weights 2/3
    assert np.allclose(shpb[:, 3], 0.0, atol=1e-14)
    assert np.allclose(qwtb.sum(), 2.0, atol=1e-14)


# ---------------------------------------------------- fully periodic box
def periodic_box(N):
    c = mesh.box_case(N, jitter=0.1)
    n = np.arange(c.nshg)
    i, j, k = n % (N + 1), (n // (N + 1)) % (N + 1), n // (N + 1) ** 2
    slave = (i == N) | (j == N) | (k == N)
    c.iBC = np.where(slave, 1024, 0).astype(np.int32)
    c.iper = ((i % N) + (N + 1) * ((j % N) + (N + 1) * (k % N)) + 1).astype(np.int32)
    c.ienb = np.zeros((0, 4), dtype=np.int32, order="F")
    return c


def uniform_state(c):
    y = np.zeros((c.nshg, 5), order="F")
    y[:, 0], y[:, 1], y[:, 2] = 0.3 * config.C0, 0.05 * config.C0, -0.02 * config.C0
    y[:, 3], y[:, 4] = 1.0e5, 300.0
    return y


def dense_folded(c, lk):
    """P^T K P restricted to periodic masters, rows/cols in local dof order."""
    n = c.nshg
    K = np.zeros((5 * n, 5 * n))
    for r in range(n):
        for k in range(c.colm[r] - 1, c.colm[r + 1] - 1):
            col = c.rowp[k] - 1
            K[5 * r:5 * r + 5, 5 * col:5 * col + 5] = lk[:, k].reshape(5, 5, order="F")
    m = c.iper - 1
    masters = np.unique(m)
    idx = {mm: t for t, mm in enumerate(masters)}
    P = np.zeros((5 * n, 5 * len(masters)))
    for r in range(n):
        P[5 * r:5 * r + 5, 5 * idx[m[r]]:5 * idx[m[r]] + 5] = np.eye(5)
    return P.T @ K @ P, masters


def test_uniform_state_residual_vanishes_on_periodic_box():
    c = periodic_box(3)
    o = pyoracle.Oracle(c)
    y = uniform_state(c)
    res, _, _, _ = o.elmgmrs(y, np.zeros_like(y))
    scale = 1e5 * (1.0 / 3) ** 2  # pressure x face area
    assert np.abs(res).max() < 1e-9 * scale


def test_tangent_is_fd_jacobian_at_uniform_state():
    """At a uniform state every term the reference freezes (dA/dY, dtau/dY,
    the non-conservative Galerkin form) vanishes or cancels over the
    periodic box, so the assembled tangent must be the Jacobian of the
    assembled residual with ac = c (y - yold)."""
    c = periodic_box(2)
    prm = dict(c.params)
    o = pyoracle.Oracle(c, prm)
    y0 = uniform_state(c)
    cf = prm["almi"] * prm["Dtgl"] / prm["gami"] / prm["alfi"]
    _, _, _, lk = o.elmgmrs(y0, np.zeros_like(y0))
    Km, masters = dense_folded(c, lk)
    loc2glob = [3, 0, 1, 2, 4]
    hstep = {0: 1.0, 1: 1e-3, 2: 1e-3, 3: 1e-3, 4: 3e-3}   # per local dof
    m = c.iper - 1
    J = np.zeros_like(Km)
    for t, mm in enumerate(masters):
        for g in range(5):
            cols = []
            for sgn in (1, -1):
                y = y0.copy(order="F")
                y[m == mm, loc2glob[g]] += sgn * hstep[g]
                res, _, _, _ = o.elmgmrs(y, np.asfortranarray(cf * (y - y0)))
                cols.append(res[masters].reshape(-1))
            J[:, 5 * t + g] = (cols[0] - cols[1]) / (2 * hstep[g])
    assert rel(J, Km) < 1e-8


# ------------------------------------------------------------- block LU
def test_block_lu_identity_with_unreduced_u45():
    """i3LU (i3lu.f:41-97): unit-lower L and U with inverted diagonal, but
    U(4,5) is never reduced, so (L U)(4,5) = A(4,5) + sum_k<4 l4k u5k and
    every other entry reproduces A."""
    rng = np.random.default_rng(7)
    case = mesh.box_case(2)
    n = case.nshg
    A = rng.standard_normal((n, 5, 5)) + 6 * np.eye(5)
    o = pyoracle.Oracle(case)
    F = np.asfortranarray(A.copy())
    o.i3lu(F, None, "LU_Fact ")
    for node in range(n):
        L = np.tril(F[node], -1) + np.eye(5)
        U = np.triu(F[node], 1) + np.diag(1.0 / np.diag(F[node]))
        LU = L @ U
        extra = L[3, :3] @ U[:3, 4]
        expect = A[node].copy()
        expect[3, 4] += extra
        assert np.allclose(LU, expect, rtol=1e-13, atol=1e-13)
    # forward then backward applies (L U)^{-1}
    r = np.asfortranarray(rng.standard_normal((n, 5)))
    x = r.copy(order="F")
    o.i3lu(F, x, "forward ")
    o.i3lu(F, x, "backward")
    for node in range(n):
        L = np.tril(F[node], -1) + np.eye(5)
        U = np.triu(F[node], 1) + np.diag(1.0 / np.diag(F[node]))
        assert np.allclose(L @ U @ x[node], r[node], rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- GMRES
def preconditioned_operator(o, c, y, ac):
    """Dense A~ (scaled, folded), b~ = L^-1 r, restricted to master dofs."""
    res, _, bd, lk = o.elmgmrs(y, ac)
    lu = bd.copy(order="F")
    o.i3lu(lu, None, "LU_Fact ")
    b = res.copy(order="F")
    o.i3lu(lu, b, "forward ")
    ls = lk.copy(order="F")
    o.spsi3pre(lu, ls)
    masters = np.unique(c.iper - 1)
    cols = []
    for mm in masters:
        for f in range(5):
            e = np.zeros((c.nshg, 5), order="F")
            e[mm, f] = 1.0
            o.sparse_ap(ls, e)
            o.bc3per(e)
            cols.append(e[masters].reshape(-1))
    return np.array(cols).T, b[masters].reshape(-1), lu, masters, res


def test_gmres_full_krylov_space_is_the_exact_solve():
    c = mesh.box_case(2, mach=0.3)
    o = pyoracle.Oracle(c)
    y, ac = o.itr_predict_bc(c.yold)
    At, bt, lu, masters, res = preconditioned_operator(o, c, y, ac)
    nd = At.shape[0]
    out = o.solgmrs(y, ac, Kspace=nd, etol=1e-13)
    _, _, _, lk = o.elmgmrs(y, ac)
    Km, _ = dense_folded(c, lk)
    x = np.linalg.solve(Km, res[masters].reshape(-1))
    assert rel(out["Dy"][masters].reshape(-1), x) < 1e-9
    assert not out["Dy"][np.setdiff1d(np.arange(c.nshg), masters)].any()


@pytest.mark.parametrize("k", [3, 7])
def test_gmres_iterate_is_krylov_least_squares(k):
    c = mesh.box_case(3, mach=0.3)
    o = pyoracle.Oracle(c)
    y, ac = o.itr_predict_bc(c.yold)
    At, bt, lu, masters, _ = preconditioned_operator(o, c, y, ac)
    V = np.empty((bt.size, k))
    v = bt / np.linalg.norm(bt)
    for j in range(k):                      # orthonormal Krylov basis
        V[:, j] = v
        v = At @ v
        v -= V[:, :j + 1] @ (V[:, :j + 1].T @ v)
        v -= V[:, :j + 1] @ (V[:, :j + 1].T @ v)
        v /= np.linalg.norm(v)
    coef, *_ = np.linalg.lstsq(At @ V, bt, rcond=None)
    z = np.zeros((c.nshg, 5), order="F")
    z[masters] = (V @ coef).reshape(-1, 5)
    o.i3lu(lu, z, "backward")
    out = o.solgmrs(y, ac, Kspace=k, etol=0.0)
    assert out["iKs"] == k
    assert rel(out["Dy"], z) < 1e-9


def test_stopping_rule_and_min_iters():
    c = mesh.box_case(3, mach=0.3)
    o = pyoracle.Oracle(c)
    y, ac = o.itr_predict_bc(c.yold)
    a = o.solgmrs(y, ac, Kspace=20, etol=1e-1, minIters=0)
    assert a["echeck"] <= a["epsnrm"] and a["iKs"] < 10
    b = o.solgmrs(y, ac, Kspace=20, etol=1e-1, minIters=10)
    assert b["iKs"] == 10                 # solgmr.f:669: converged but iKs < minIters
    assert b["echeck"] <= a["echeck"]


def test_rstat_norms_and_abort_rules(box4):
    """rstat.f:95-108 norms, rstatCheck.f:36-79 divergence and tnanq.f:10-28
    NaN aborts of the oracle's SolGMRs."""
    o = pyoracle.Oracle(box4)
    y, ac = o.itr_predict_bc(box4.yold)
    out = o.solgmrs(y, ac)
    n = box4.nshg
    assert out["totres"][0] == pytest.approx(np.linalg.norm(out["res"]) / n, rel=1e-14)
    assert out["totres"][1] == pytest.approx(np.linalg.norm(out["rmes"]) / n, rel=1e-14)
    # from the second Newton iteration of a time step on, a residual 1e4
    # times above the last one aborts (code 11); the first is not tested
    o.reslast = 1e-4 * np.sqrt(np.sum(out["res"] ** 2) / n) * 0.999
    with pytest.raises(RuntimeError, match="rc=11"):
        o.solgmrs(y, ac, iter=2)
    o.solgmrs(y, ac, iter=1)
    o.reslast = 1e-4 * np.sqrt(np.sum(out["res"] ** 2) / n) * 1.001
    o.solgmrs(y, ac, iter=2)  # just below the bound
    o.reslast = 0.0
    bad = y.copy(order="F")
    bad[3, 2] = np.nan
    with pytest.raises(RuntimeError, match="rc=10"):
        o.solgmrs(bad, ac)


def test_divergence_check_skips_the_first_iteration_of_a_time_step(box4):
    """rstatCheck.f:73 tests iter > 1 and itrdrv.f:409 resets iter at each
    time step: a converged step (here a resting uniform state, residual at
    rounding level) followed by a new step whose first residual is far more
    than 1e4 times larger does not abort; the same residual as the step's
    second Newton iteration does."""
    o = pyoracle.Oracle(box4)
    rest = np.zeros((box4.nshg, 5), order="F")
    rest[:, 3], rest[:, 4] = 1.0e5, 300.0
    acold = np.zeros((box4.nshg, 5), order="F")
    y0, ac0 = o.itr_predict_bc(rest)
    o.solgmrs(y0, ac0, iter=1)
    quiet = o.reslast
    y, ac = o.itr_predict_bc(box4.yold, acold)
    out = o.solgmrs(y, ac, iter=1)
    assert o.reslast > 1e4 * quiet
    assert out["totres"][0] > 0.0
    o.reslast = quiet
    with pytest.raises(RuntimeError, match="rc=11"):
        o.solgmrs(y, ac, iter=2)
    # the time-step driver numbers the Newton iterations from 1 per step
    o2 = pyoracle.Oracle(box4)
    o2.reslast = quiet
    yo, aco, hist = o2.time_steps(box4.yold, acold, steps=1, newton=2)
    assert len(hist) == 2


# ------------------------------------------------ matrix tau (itau = 10)
def test_matrix_tau_elmgmrs_equals_reference_golden():
    """ElmGMRs with the matrix tau of e3tau_nd (e3tau.f:284-744 with
    e3eig1/e3eig2 and e3visc's compK), alone and with iDC 1, bitwise against
    the reference's compiled routines (fixture tests/golden/ref_mtau_box3.npz);
    the matrix tau changes the residual and the tangent; itau values the
    build does not implement are refused."""
    g = np.load(os.path.join(GOLD, "ref_mtau_box3.npz"))
    case = mesh.box_case(3, mach=0.5, kspace=10, jitter=0.1)
    y, ac = np.asfortranarray(g["y"]), np.asfortranarray(g["ac"])
    base = pyoracle.Oracle(case).elmgmrs(y, ac)
    for tag, extra in (("", {}), ("_dc", {"idc": 1})):
        prm = dict(case.params, itau=10, **extra)
        res, rmes, bd, lk = pyoracle.Oracle(case, params=prm).elmgmrs(y, ac)
        assert np.array_equal(res, g[f"res{tag}"]), tag
        assert np.array_equal(bd, g[f"BDiag{tag}"]), tag
        assert np.array_equal(lk, g[f"lhsK{tag}"]), tag
    assert rel(g["res"], base[0]) > 1e-2 and rel(g["lhsK"], base[3]) > 1e-2
    # itau 2 and 3 have no branch in the compressible e3tau (tau left unset)
    for itau in (2, 3):
        with pytest.raises(RuntimeError):
            pyoracle.Oracle(case, params=dict(case.params, itau=itau)).elmgmrs(y, ac)


# ------------------------------------------------ Franca's tau (itau = 1), itau = 11
def test_franca_tau_equals_reference_golden():
    """ElmGMRs with Franca's diagonal tau (e3tau.f:54-139) alone, with iDC 1
    and with idiff 1, and ElmMFG with it, bitwise against the reference's
    compiled routines (fixture tests/golden/ref_franca_box3.npz); the tau
    changes res and the tangent by O(1) against Shakib's."""
    g = np.load(os.path.join(GOLD, "ref_franca_box3.npz"))
    case = mesh.box_case(3, mach=0.5, kspace=10, jitter=0.1)
    y, ac = np.asfortranarray(g["y"]), np.asfortranarray(g["ac"])
    for tag, extra in (("", {}), ("_dc", {"idc": 1}), ("_idiff", {"idiff": 1})):
        prm = dict(case.params, itau=1, **extra)
        res, rmes, bd, lk = pyoracle.Oracle(case, params=prm).elmgmrs(y, ac)
        assert np.array_equal(res, g[f"res{tag}"]), tag
        assert np.array_equal(bd, g[f"BDiag{tag}"]), tag
        assert np.array_equal(lk, g[f"lhsK{tag}"]), tag
    res, rmes, bd = pyoracle.Oracle(case, params=dict(case.params, itau=1, lhs=0,
                                                      iprec=1)).elmmfg(y, ac)
    assert np.array_equal(res, g["mfg_res"]) and np.array_equal(rmes, g["mfg_rmes"])
    assert np.array_equal(bd, g["mfg_BDiag"])
    base = pyoracle.Oracle(case).elmgmrs(y, ac)
    assert rel(g["res"], base[0]) > 1e-1 and rel(g["lhsK"], base[3]) > 1e-1


def test_modal_matrix_tau_is_the_mallet_one():
    """itau = 11 (Matrix-Modal) reaches the same e3tau_nd as 10: e3ls.f:188
    branches on itau < 10 only and e3tau_nd reads no itau."""
    case = mesh.box_case(4, mach=0.3, kspace=10, jitter=0.1, grade=1.1)
    y, ac = pyoracle.Oracle(case).itr_predict_bc(case.yold)
    a = pyoracle.Oracle(case, params=dict(case.params, itau=10)).elmgmrs(y, ac)
    b = pyoracle.Oracle(case, params=dict(case.params, itau=11)).elmgmrs(y, ac)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


@pytest.mark.skipif(not pyref.available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("itau,extra", [(1, {}), (1, {"idc": 2}), (1, {"idiff": 3}),
                                        (11, {}), (11, {"idc": 1})], ids=str)
def test_tau_variants_bitwise_equal_live_reference(itau, extra):
    case = mesh.box_case(4, mach=0.3, kspace=10, jitter=0.1, grade=1.1)
    rng = np.random.default_rng(31)
    y, ac = pyoracle.Oracle(case).itr_predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.05 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    prm = dict(case.params, itau=itau, **extra)
    a = pyoracle.Oracle(case, params=prm).elmgmrs(y, ac)
    b = pyref.Reference(case, params=prm).elmgmrs(y, ac)
    for k in (0, 2, 3):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.skipif(not pyref.available(), reason="oracle/_ref not built")
def test_matrix_tau_elmgmrs_bitwise_equals_live_reference():
    case = mesh.box_case(4, mach=0.3, kspace=10, jitter=0.1, grade=1.1)
    rng = np.random.default_rng(21)
    y, ac = pyoracle.Oracle(case).itr_predict_bc(case.yold)
    y = np.asfortranarray(y * (1 + 0.05 * rng.standard_normal(y.shape)))
    ac = np.asfortranarray(ac + 1e3 * rng.standard_normal(ac.shape))
    for extra in ({}, {"idc": 2}, {"idiff": 1}):
        prm = dict(case.params, itau=10, **extra)
        for a, b in zip(pyoracle.Oracle(case, params=prm).elmgmrs(y, ac),
                        pyref.Reference(case, params=prm).elmgmrs(y, ac)):
            assert np.array_equal(a, b), extra
