"""Pins of the CPU oracle (oracle/oracle.c) and of the library's host-side
operator setup: reference-held constants, probe values printed by the compiled
reference, closed forms and invariants that a dropped term, a wrong sign or a
transposed operand would break.  CPU only."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg
from scipy import special

import bfam_amd as bx
import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_constants.json")))
ULP = np.finfo(np.float64).eps


def legendre_P(N, x):
    c = np.zeros(N + 1)
    c[N] = 1.0
    return npleg.legval(x, c)


# ------------------------------------------------------------------ A1 Jacobi / LGL

@pytest.mark.parametrize("N", range(1, 13))
def test_lgl_closed_forms(N):
    r, w = oracle.gauss_lobatto(N)
    assert r[0] == -1.0 and r[-1] == 1.0
    assert np.all(np.diff(r) > 0)
    # interior nodes are the roots of P_N'
    c = np.zeros(N + 1)
    c[N] = 1.0
    roots = np.sort(np.real(npleg.legroots(npleg.legder(c)))) if N > 1 else np.array([])
    assert np.allclose(r[1:-1], roots, atol=1e-14, rtol=0)
    # w_i = 2 / (N (N+1) P_N(r_i)^2)
    wc = 2.0 / (N * (N + 1) * legendre_P(N, r) ** 2)
    assert np.allclose(w, wc, rtol=1e-14, atol=0)
    assert np.allclose(r, -r[::-1], atol=1e-15) and np.allclose(w, w[::-1], rtol=1e-14)
    assert abs(w.sum() - 2.0) < 1e-14
    # exact on polynomials of degree 2N-1
    for k in range(2 * N):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, r ** k) - exact) < 1e-14


def test_lgl_known_nodes():
    r, w = oracle.gauss_lobatto(4)
    assert np.allclose(r, [-1, -math.sqrt(3 / 7), 0, math.sqrt(3 / 7), 1], atol=1e-16)
    assert np.allclose(w, [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], rtol=1e-15)


def test_lgl_reference_probe():
    g = GOLD["lgl_probe"]
    r3, w3 = oracle.gauss_lobatto(3)
    assert abs(r3[1] - g["N3_r1"]) <= 2 * ULP
    assert abs(w3[0] - g["N3_w0"]) <= 2 * ULP
    r7, w7, _, D7 = oracle.dgx_ops(7)
    assert abs(w7[0] - g["N7_w0"]) <= 4 * ULP * g["N7_w0"] + 1e-17
    assert abs(D7[0, 0] - g["N7_D00"]) <= 16 * ULP * 14
    assert np.abs(D7 @ np.ones(8)).max() < 10 * g["N7_max_abs_D_times_ones"]
    assert abs(w7.sum() - 2.0) < 1e-14


@pytest.mark.parametrize("ab", [(0.0, 0.0), (1.0, 1.0), (0.5, -0.5), (2.0, 0.0)])
@pytest.mark.parametrize("N", [0, 1, 3, 6, 9])
def test_jacobi_p_orthonormal_against_scipy(ab, N):
    a, b = ab
    x = np.linspace(-0.99, 0.99, 23)
    got = oracle.jacobi_p(a, b, N, x)
    # orthonormal scaling: h_N = 2^(a+b+1) G(N+a+1) G(N+b+1) / ((2N+a+b+1) N! G(N+a+b+1))
    lg = ((a + b + 1) * math.log(2) + math.lgamma(N + a + 1) + math.lgamma(N + b + 1)
          - math.log(2 * N + a + b + 1) - math.lgamma(N + 1) - math.lgamma(N + a + b + 1))
    want = special.eval_jacobi(N, a, b, x) / math.exp(0.5 * lg)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    if N > 0:
        # d/dx P_N^{(a,b)} = (N+a+b+1)/2 P_{N-1}^{(a+1,b+1)} (unnormalised)
        gd = oracle.grad_jacobi_p(a, b, N, x)
        wd = 0.5 * (N + a + b + 1) * special.eval_jacobi(N - 1, a + 1, b + 1, x) / math.exp(0.5 * lg)
        assert np.allclose(gd, wd, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("ab", [(0.0, 0.0), (1.0, 1.0), (0.5, 1.5)])
@pytest.mark.parametrize("N", [0, 2, 5, 11])
def test_gauss_against_scipy(ab, N):
    a, b = ab
    r, w = oracle.gauss(N, a, b)
    xr, wr = special.roots_jacobi(N + 1, a, b)
    assert np.allclose(r, xr, atol=1e-14)
    assert np.allclose(w, wr, rtol=1e-12)


# ------------------------------------------------------------- A2/A3 operators

@pytest.mark.parametrize("N", range(1, 11))
def test_differentiation_closed_form(N):
    r, w, wi, D = oracle.dgx_ops(N)
    P = legendre_P(N, r)
    Dc = np.zeros((N + 1, N + 1))
    for i in range(N + 1):
        for j in range(N + 1):
            if i != j:
                Dc[i, j] = P[i] / (P[j] * (r[i] - r[j]))
    Dc[0, 0] = -N * (N + 1) / 4.0
    Dc[N, N] = N * (N + 1) / 4.0
    assert np.allclose(D, Dc, atol=2e-12 * N * N, rtol=0)
    # exact derivative of monomials up to degree N
    for k in range(N + 1):
        want = k * r ** (k - 1) if k else np.zeros_like(r)
        assert np.allclose(D @ r ** k, want, atol=1e-11 * N * N)
    assert np.allclose(wi, 1.0 / w, rtol=1e-15)


@pytest.mark.parametrize("N", [1, 3, 4, 7])
def test_vandermonde_mass_interp(N):
    r, w = oracle.gauss_lobatto(N)
    V = oracle.vandermonde(N, r)
    # exact mass matrix of the Lagrange basis on the LGL nodes
    M = oracle.mass(N)
    xq, wq = npleg.leggauss(N + 2)
    L = np.stack([bx.lagrange_eval(N, r, x) for x in xq])  # (nq, N+1)
    Mex = (L * wq[:, None]).T @ L
    assert np.allclose(M, Mex, atol=1e-13)
    assert np.allclose(M, np.linalg.inv(V @ V.T), atol=1e-12)
    # LGL lumping is exact on the diagonal only for N=1 row sums
    assert np.allclose(M.sum(axis=1), w, atol=1e-13)


def test_host_operators_match_oracle():
    for N in range(1, 11):
        r, w = bx.lgl(N)
        ro, wo, _, Do = oracle.dgx_ops(N)
        # the library uses Kopriva's Newton iteration on L_N, the oracle the
        # reference's Jacobi-polynomial route: both within a few ulps
        assert np.abs(r - ro).max() <= 4 * ULP
        assert np.abs(w - wo).max() <= 1e-14 * w.max()
        D = bx.diff_matrix(N, r)
        assert np.abs(D - Do).max() <= 1e-13 * N * N


# ------------------------------------------------------------- A8 mortar ops

@pytest.mark.parametrize("Na,Nb", [(3, 3), (3, 5), (5, 7), (4, 7), (7, 7), (2, 6)])
def test_mortar_identities(Na, Nb):
    I = oracle.interpolator(Na, Nb)   # a -> b
    Pj = oracle.interpolator(Nb, Na)  # b -> a
    ra, wa = oracle.gauss_lobatto(Na)
    rb, wb = oracle.gauss_lobatto(Nb)
    Ma, Mb = oracle.mass(Na), oracle.mass(Nb)
    Iab = I["prj"][0] if I["prj"][0] is not None else np.eye(Nb + 1)
    Pba = Pj["prj"][0] if Pj["prj"][0] is not None else np.eye(Na + 1)
    # interpolation exact on degree <= Na
    for k in range(Na + 1):
        assert np.allclose(Iab @ ra ** k, rb ** k, atol=1e-13)
    # L2 projection after interpolation is the identity (Nb >= Na)
    assert np.allclose(Pba @ Iab, np.eye(Na + 1), atol=1e-13)
    # exact-mass projection identity M_a P = I^T M_b
    assert np.allclose(Ma @ Pba, Iab.T @ Mb, atol=1e-13)
    # hanging pair: P_bottom I_bottom + P_top I_top = I (each half carries M/2)
    HI = oracle.interpolator(Na, Nb)["prj"]
    HP = oracle.interpolator(Nb, Na)["prj"]
    S = HP[1] @ HI[3] + HP[2] @ HI[4]
    assert np.allclose(S, np.eye(Na + 1), atol=1e-13)
    # children at (r -+ 1)/2: refine-to-bottom interpolates onto the lower half
    for k in range(Na + 1):
        assert np.allclose(HI[3] @ ra ** k, ((rb - 1) / 2) ** k, atol=1e-13)
        assert np.allclose(HI[4] @ ra ** k, ((rb + 1) / 2) ** k, atol=1e-13)
    # projection conserves the integral
    f = np.cos(1.3 * rb)
    assert abs(wa @ (Pba @ f) - (Mb @ f).sum()) < 1e-13


@pytest.mark.parametrize("Na,Nb", [(3, 5), (5, 5), (7, 4), (4, 7)])
def test_host_mortar_ops_match_oracle(Na, Nb):
    h = bx.mortar_ops(Na, Nb)
    o = oracle.interpolator(Na, Nb)
    for k in range(5):
        if o["prj"][k] is None:
            assert h["prj0_identity"]
            continue
        assert np.allclose(h["prj"][k], o["prj"][k], atol=1e-13)
        assert np.allclose(h["mass_prj"][k], o["mass_prj"][k], atol=1e-13)
        assert np.allclose(h["wi_mass_prj"][k], o["wi_mass_prj"][k], atol=1e-12)


# ------------------------------------------------------------------ A10 LSRK

METHODS = {"KC54": oracle.LSRK_KC54, "W33": oracle.LSRK_W33, "HEUN": oracle.LSRK_HEUN,
           "FE": oracle.LSRK_FE}
ORDER = {"KC54": 4, "W33": 3, "HEUN": 2, "FE": 1}


def rational(pair):
    return Fraction(int(pair[0]), int(pair[1]))


@pytest.mark.parametrize("name", list(METHODS))
def test_lsrk_tables_match_reference_rationals(name):
    A, B, Cc = oracle.lsrk_table(METHODS[name])
    g = GOLD["lsrk"][name]
    for got, want in ((A, g["A"]), (B, g["B"]), (Cc, g["C"])):
        assert len(got) == len(want)
        for x, p in zip(got, want):
            # long-double quotient rounded to double: within one ulp of the
            # correctly rounded rational
            exact = float(rational(p))
            assert abs(x - exact) <= ULP * abs(exact)
    hA, hB, hC = bx.lsrk_init({"KC54": bx.LSRK_KC54, "W33": bx.LSRK_W33,
                               "HEUN": bx.LSRK_HEUN, "FE": bx.LSRK_FE}[name])
    assert np.array_equal(hA, A) and np.array_equal(hB, B) and np.array_equal(hC, Cc)


@pytest.mark.parametrize("name", list(METHODS))
def test_lsrk_order_conditions(name):
    """The 2N-storage recursion applied to y' = z y gives a stability
    polynomial whose Taylor coefficients match e^z to the method's order
    (computed in exact rationals from the reference-held tables)."""
    g = GOLD["lsrk"][name]
    A = [rational(p) for p in g["A"]]
    B = [rational(p) for p in g["B"]]
    ns = len(A)
    # polynomials in z as coefficient lists
    y, dq = [Fraction(1)], [Fraction(0)]

    def add(a, b):
        n = max(len(a), len(b))
        return [(a[i] if i < len(a) else 0) + (b[i] if i < len(b) else 0) for i in range(n)]

    def scale(a, s):
        return [s * x for x in a]

    def shift(a):
        return [Fraction(0)] + a

    for s in range(ns):
        dq = add(scale(dq, A[s]), shift(y))
        y = add(y, scale(dq, B[s]))
    p = ORDER[name]
    for k in range(p + 1):
        assert y[k] == Fraction(1, math.factorial(k)) or \
            abs(float(y[k]) - 1 / math.factorial(k)) < 1e-12, (k, y[k])
    # stage times: C[s] is the time reached by the stage-s input for y' = 1
    Cc = [rational(p) for p in g["C"]]
    t, dqt = Fraction(0), Fraction(0)
    for s in range(ns):
        assert abs(float(t - Cc[s])) < 1e-12
        dqt = A[s] * dqt + 1
        t = t + B[s] * dqt
    assert abs(float(t - 1)) < 1e-12


def test_lsrk_convergence_order_oracle():
    """KC54 converges at 4th order on a periodic acoustic plane wave when the
    space error is negligible (N=8, 6x6)."""
    from util import oracle_problem
    import bfam_amd.workloads as W
    cfg = W.acoustic2d_plane_wave(n=6, N=8)
    P = oracle_problem(cfg)
    T = 0.05
    errs = []
    for nsteps in (10, 20):
        dt = T / nsteps
        q = P.run(cfg["q0"], dt, nsteps)
        ex = W.acoustic2d_exact(cfg, T)
        errs.append(np.sqrt(((q - ex) ** 2).sum() / (ex ** 2).sum()))
    rate = math.log2(errs[0] / errs[1])
    assert rate > 3.6, errs


# ------------------------------------------------------------------ PCG32

def test_pcg32_known_answer():
    """O'Neill's pcg32 demo: seed (42, 54) -> first six outputs."""
    want = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    o = oracle.PCG32(42, 54)
    h = bx.PCG32(42, 54)
    assert [o.next() for _ in range(6)] == want
    assert [h.next() for _ in range(6)] == want
    o, h = oracle.PCG32(7, 3), bx.PCG32(7, 3)
    for b in (1, 3, 10, 1000, 2 ** 31 + 5):
        for _ in range(50):
            x, y = o.bounded(b), h.bounded(b)
            assert x == y and 0 <= x < b


def test_ricker():
    f0, t0 = 3.0, 0.4
    for t in (0.0, 0.3, 0.4, 0.55):
        a = (math.pi * f0 * (t - t0)) ** 2
        assert abs(oracle.ricker(f0, t0, t) - (1 - 2 * a) * math.exp(-a)) < 1e-15
