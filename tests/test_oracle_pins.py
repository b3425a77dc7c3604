"""Pins of the CPU oracle: each check is a closed form, an invariant, a known answer from the
reference's own comments, or brute force, chosen so that a plausible mistake (dropped term, wrong
sign or index, transposed operand) fails one of them. CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as PO

p_ = PO.p
dp = C.POINTER(C.c_double)


def phys(flux=0, mu0=1e-3, gamma=1.4):
    o = PO.OPhys()
    o.gamma, o.molecular_weight, o.prandtl, o.mu0, o.T0, o.S = gamma, 8.314462618, 0.72, mu0, 1.0, 0.4
    o.inv_surf_flux = flux
    o.threshold_factor, o.threshold_exp_factor, o.sharpness_factor = 0.5, 1.8, 9.21024
    o.alpha_min, o.alpha_max, o.wall_blender_limit = 1e-3, 0.5, 0.5
    return o


def basis(N):
    lib = PO.load()
    x, w, D = np.zeros(N + 1), np.zeros(N + 1), np.zeros((N + 1) * (N + 1))
    lib.oracle_basis(N, x.ctypes.data_as(dp), w.ctypes.data_as(dp), D.ctypes.data_as(dp))
    return x, w, D.reshape(N + 1, N + 1)


# ------------------------------------------------------------------ GLL basis on [0,1]
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5])
def test_gll_quadrature_exactness(N):
    x, w, _ = basis(N)
    assert x[0] == 0.0 and abs(x[-1] - 1.0) < 1e-15
    assert abs(w.sum() - 1.0) < 1e-14
    for k in range(2 * N):  # exact for degree <= 2N-1
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-14, k
    # and not for 2N (Lobatto rule has degree 2N-1): guards against a Gauss rule in its place
    assert abs(np.dot(w, x ** (2 * N)) - 1.0 / (2 * N + 1)) > 1e-8


def test_gll_closed_forms():
    x, w, _ = basis(3)
    assert np.allclose(x, [0, 0.5 - math.sqrt(5) / 10, 0.5 + math.sqrt(5) / 10, 1], atol=1e-15)
    assert np.allclose(w, [1 / 12, 5 / 12, 5 / 12, 1 / 12], atol=1e-15)
    x, w, _ = basis(4)
    assert np.allclose(x, [0, 0.5 - math.sqrt(21) / 14, 0.5, 0.5 + math.sqrt(21) / 14, 1], atol=1e-15)
    assert np.allclose(w, [0.05, 49 / 180, 32 / 90, 49 / 180, 0.05], atol=1e-15)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5])
def test_differentiation_matrix(N):
    x, w, D = basis(N)
    assert np.abs(D.sum(axis=1)).max() < 1e-12  # D 1 = 0
    for k in range(1, N + 1):  # exact on polynomials of degree <= N
        assert np.abs(D @ x ** k - k * x ** (k - 1)).max() < 1e-11
    Q = np.diag(w) @ D  # SBP: Q + Q^T = B
    B = np.zeros((N + 1, N + 1))
    B[0, 0], B[-1, -1] = -1, 1
    assert np.abs(Q + Q.T - B).max() < 1e-13


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5])
def test_product_basis_matches_oracle(N):
    from plens_amd import _lib as L
    lib = L.load()
    x, w, D, C1 = (np.zeros(N + 1), np.zeros(N + 1), np.zeros((N + 1) ** 2), np.zeros((N + 1) ** 2))
    assert lib.plens_basis_1d(N, *(a.ctypes.data_as(dp) for a in (x, w, D, C1))) == 0
    xo, wo, Do = basis(N)
    assert np.abs(x - xo).max() < 2e-16 and np.abs(w - wo).max() < 2e-16
    assert np.abs(D.reshape(N + 1, N + 1) - Do).max() < 1e-13


# ------------------------------------------------------------------ change of basis (blender)
def legendre01(k, x):
    """orthonormal Legendre on [0,1] from numpy's Legendre class"""
    c = np.zeros(k + 1)
    c[k] = 1
    return math.sqrt(2 * k + 1) * np.polynomial.legendre.legval(2 * x - 1, c)


def cob(N):
    n = (N + 1) ** 3
    Cm = np.zeros(n * n)
    PO.load().oracle_cob_matrix(N, Cm.ctypes.data_as(dp))
    return Cm.reshape(n, n)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_change_of_basis_modes(N):
    x, _, _ = basis(N)
    Cm = cob(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    lex = lambda A: A.transpose(2, 1, 0).reshape(-1)  # node id i + j(N+1) + k(N+1)^2
    # a single tensor Legendre mode comes back as exactly that mode
    rng = np.random.default_rng(0)
    for _ in range(5):
        a, b, c = rng.integers(0, N + 1, size=3)
        f = legendre01(a, X) * legendre01(b, Y) * legendre01(c, Z)
        m = Cm @ lex(f)
        e = np.zeros_like(m)
        e[a + (N + 1) * b + (N + 1) ** 2 * c] = 1
        assert np.abs(m - e).max() < 1e-12
    # Parseval for a random polynomial of degree <= N per direction
    coef = rng.standard_normal((N + 1,) * 3)
    f = sum(coef[a, b, c] * legendre01(a, X) * legendre01(b, Y) * legendre01(c, Z)
            for a in range(N + 1) for b in range(N + 1) for c in range(N + 1))
    m = Cm @ lex(f)
    assert abs((m ** 2).sum() - (coef ** 2).sum()) < 1e-11 * (coef ** 2).sum()


@pytest.mark.parametrize("N", [2, 3, 4])
def test_product_cob_is_tensor_of_oracle(N):
    from plens_amd import _lib as L
    x, w, D, C1 = (np.zeros(N + 1), np.zeros(N + 1), np.zeros((N + 1) ** 2), np.zeros((N + 1) ** 2))
    L.load().plens_basis_1d(N, *(a.ctypes.data_as(dp) for a in (x, w, D, C1)))
    C1 = C1.reshape(N + 1, N + 1)
    full = np.kron(np.kron(C1, C1), C1)  # lexicographic x-fastest: kron(z, y, x)
    assert np.abs(full - cob(N)).max() < 1e-13


# ------------------------------------------------------------------ blender (Hennemann)
@pytest.mark.parametrize("N", [3, 4])
def test_blender_closed_form(N):
    x, _, _ = basis(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    lex = lambda A: A.transpose(2, 1, 0).reshape(-1)
    ph = phys()
    T = 0.5 * 10 ** (-1.8 * (N + 1) ** 0.25)
    lib = PO.load()
    for amp in [1e-4, 3e-3, 1e-2, 0.05, 0.2]:
        f = 1.0 + amp * legendre01(N, X)  # energies: tot 1+a^2, N-1: 1, N-2: 1
        trouble = amp * amp / (1 + amp * amp)
        at = 1 / (1 + math.exp(-9.21024 * (trouble / T - 1)))
        expect = 0.0 if at < 1e-3 else min(at, 0.5)
        got = lib.oracle_blender_value(C.byref(ph), N, p_(lex(f)))
        assert abs(got - expect) < 1e-9, (amp, got, expect)
    # second-highest mode triggers the Hennemann term 1 - E_{N-2}/E_{N-1}
    f = 1.0 + 0.05 * legendre01(N - 1, Y)
    trouble = 0.0025 / 1.0025
    at = 1 / (1 + math.exp(-9.21024 * (trouble / T - 1)))
    got = lib.oracle_blender_value(C.byref(ph), N, p_(lex(f)))
    assert abs(got - min(at, 0.5) * (at >= 1e-3)) < 1e-9
    # a trilinear field has no energy in the top modes -> alpha = 0
    f = 1.0 + 0.1 * X + 0.2 * Y * Z
    assert lib.oracle_blender_value(C.byref(ph), N, p_(lex(f))) == 0.0


# ------------------------------------------------------------------ log mean
def test_log_avg():
    lib = PO.load()
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b = np.exp(rng.uniform(-3, 3, 2))
        exact = (a - b) / (math.log(a) - math.log(b)) if a != b else a
        f = (a / b - 1) / (a / b + 1)
        tol = 1e-14 if f * f >= 1e-2 else 1e-9  # series truncation of the reference below u=1e-2
        assert abs(lib.oracle_log_avg(a, b) - exact) <= tol * exact
    assert lib.oracle_log_avg(2.5, 2.5) == 2.5
    # both branches close to the switch agree with the exact mean
    for r in [1.2, 1.222, 1.23, 0.82]:
        exact = (r - 1) / math.log(r)
        assert abs(lib.oracle_log_avg(r, 1.0) - exact) < 2e-9


# ------------------------------------------------------------------ fluxes
def rand_state(rng, g=1.4, mach=2.0):
    rho = rng.uniform(0.5, 3.0)
    v = rng.uniform(-1, 1, 3) * mach
    p = rng.uniform(0.5, 3.0)
    return np.array([rho, rho * v[0], rho * v[1], rho * v[2], p / (g - 1) + 0.5 * rho * v @ v])


def euler_flux(u, n, g=1.4):
    rho, m, E = u[0], u[1:4], u[4]
    v = m / rho
    p = (g - 1) * (E - 0.5 * rho * v @ v)
    vn = v @ n
    return np.array([rho * vn, *(m * vn + p * n), (E + p) * vn])


def surf(ph, scheme, a, ul, ur, n):
    f = np.zeros(5)
    PO.load().oracle_inv_surf_flux(C.byref(ph), scheme, a, p_(ul), p_(ur), p_(n), f.ctypes.data_as(dp))
    return f


def unit(rng):
    n = rng.standard_normal(3)
    return n / np.linalg.norm(n)


ALL_SCHEMES = list(range(10))  # HLLC ... Ismail-Roe, as PLENS_FLUX_*


@pytest.mark.parametrize("scheme", ALL_SCHEMES)
def test_surface_flux_consistency_and_conservation(scheme):
    rng = np.random.default_rng(2 + scheme)
    ph = phys(flux=scheme)
    for _ in range(100):
        u, ur, n = rand_state(rng), rand_state(rng), unit(rng)
        a = rng.uniform(0, 1)
        f = surf(ph, scheme, a, u, u, n)
        assert np.abs(f - euler_flux(u, n)).max() <= 1e-12 * np.abs(euler_flux(u, n)).max() + 1e-12
        # F(uL,uR,n) = -F(uR,uL,-n)
        f1, f2 = surf(ph, scheme, a, u, ur, n), surf(ph, scheme, a, ur, u, -n)
        assert np.abs(f1 + f2).max() <= 1e-11 * (1 + np.abs(f1).max())


@pytest.mark.parametrize("scheme", ALL_SCHEMES)
def test_surface_flux_rotational_invariance(scheme):
    rng = np.random.default_rng(10 + scheme)
    ph = phys(flux=scheme)
    for _ in range(50):
        ul, ur, n = rand_state(rng), rand_state(rng), unit(rng)
        A = np.linalg.qr(rng.standard_normal((3, 3)))[0]
        rot = lambda u: np.array([u[0], *(A @ u[1:4]), u[4]])
        f = surf(ph, scheme, 0.3, ul, ur, n)
        fr = surf(ph, scheme, 0.3, rot(ul), rot(ur), A @ n)
        assert np.abs(rot(f) - fr).max() <= 1e-11 * (1 + np.abs(f).max())


def test_supersonic_upwinding():
    # both states supersonic along +n: every scheme except the central ones returns f(uL)
    rng = np.random.default_rng(3)
    for scheme in (0, 2):
        ph = phys(flux=scheme)
        for _ in range(20):
            n = unit(rng)
            rho, p = rng.uniform(0.5, 2, 2)
            a = math.sqrt(1.4 * p / rho)
            v = n * a * rng.uniform(2.5, 4)
            ul = np.array([rho, *(rho * v), p / 0.4 + 0.5 * rho * v @ v])
            v2 = v * 1.1
            ur = np.array([rho * 0.9, *(rho * 0.9 * v2), p * 1.2 / 0.4 + 0.5 * rho * 0.9 * v2 @ v2])
            f = surf(ph, scheme, 0.0, ul, ur, n)
            assert np.abs(f - euler_flux(ul, n)).max() <= 1e-11 * np.abs(f).max()


def test_hllc_contact_exact():
    # isolated contact (same u, p, different rho) moving with +u: HLLC returns the upwind flux
    ph = phys(flux=0)
    n = np.array([1.0, 0, 0])
    for u0 in [0.3, -0.4]:
        ul = np.array([1.0, u0, 0.2, -0.1, 1 / 0.4 + 0.5 * (u0 ** 2 + 0.05)])
        ur = np.array([2.0, 2 * u0, 2 * 0.2, 2 * -0.1, 1 / 0.4 + 0.5 * 2 * (u0 ** 2 + 0.05)])
        f = surf(ph, 0, 0.0, ul, ur, n)
        up = ul if u0 > 0 else ur
        assert np.abs(f - euler_flux(up, n)).max() < 1e-12


def test_rusanov_closed_form():
    ph = phys(flux=1)
    rng = np.random.default_rng(5)
    n = np.array([1.0, 0, 0])
    for _ in range(20):
        ul, ur = rand_state(rng), rand_state(rng)
        c = lambda u: math.sqrt(1.4 * 0.4 * (u[4] - 0.5 * u[1:4] @ u[1:4] / u[0]) / u[0]) + abs(u[1] / u[0])
        S = max(c(ul), c(ur))
        expect = 0.5 * (euler_flux(ul, n) + euler_flux(ur, n)) - 0.5 * S * (ur - ul)
        assert np.abs(surf(ph, 1, 0, ul, ur, n) - expect).max() < 1e-12 * np.abs(expect).max()


def entropy_vars(u, g=1.4):
    rho, v = u[0], u[1:4] / u[0]
    p = (g - 1) * (u[4] - 0.5 * rho * v @ v)
    s = math.log(p) - g * math.log(rho)
    beta = 0.5 * rho / p
    return np.array([(g - s) / (g - 1) - beta * v @ v, *(2 * beta * v), -2 * beta])


def test_volume_flux_entropy_conservative():
    """Tadmor's condition (w2-w1).F#(u1,u2) = psi2 - psi1 with psi = rho u.n for Chandrashekhar's
    flux, plus symmetry and consistency."""
    ph = phys()
    lib = PO.load()
    rng = np.random.default_rng(6)
    for _ in range(100):
        u1, u2, n = rand_state(rng), rand_state(rng), unit(rng)
        f, fs = np.zeros(5), np.zeros(5)
        lib.oracle_vol_flux(C.byref(ph), p_(u1), p_(u2), p_(n), f.ctypes.data_as(dp))
        lib.oracle_vol_flux(C.byref(ph), p_(u2), p_(u1), p_(n), fs.ctypes.data_as(dp))
        assert np.abs(f - fs).max() < 1e-12 * np.abs(f).max()
        lhs = (entropy_vars(u2) - entropy_vars(u1)) @ f
        rhs = u2[1:4] @ n - u1[1:4] @ n
        # exact up to the reference's log-mean series truncation (~1e-9 relative below u=1e-2)
        assert abs(lhs - rhs) < 1e-8 * (1 + abs(rhs))
        lib.oracle_vol_flux(C.byref(ph), p_(u1), p_(u1), p_(n), f.ctypes.data_as(dp))
        assert np.abs(f - euler_flux(u1, n)).max() < 1e-12 * np.abs(f).max()


def test_chandrashekhar_surface_dissipates_entropy():
    # the matrix dissipation term removes entropy for alpha in [0,1]: (w_R-w_L).(F* - F#) <= 0
    lib = PO.load()
    rng = np.random.default_rng(8)
    ph = phys(flux=3)
    n = np.array([1.0, 0, 0])
    for _ in range(100):
        ul, ur = rand_state(rng, mach=0.5), rand_state(rng, mach=0.5)
        fec = np.zeros(5)
        lib.oracle_vol_flux(C.byref(ph), p_(ul), p_(ur), p_(n), fec.ctypes.data_as(dp))
        for a in (0.0, 0.5, 1.0):
            fs = surf(ph, 3, a, ul, ur, n)
            assert (entropy_vars(ur) - entropy_vars(ul)) @ (fs - fec) <= 1e-12


def vol(ph, u1, u2, n):
    f = np.zeros(5)
    PO.load().oracle_vol_flux(C.byref(ph), p_(u1), p_(u2), p_(n), f.ctypes.data_as(dp))
    return f


def ir_phys():
    ph = phys(flux=9)
    ph.inv_vol_flux = 2
    return ph


def test_ismail_roe_volume_flux_entropy_conservative():
    """Tadmor's condition (w2-w1).F#(u1,u2) = psi2 - psi1, psi = rho u.n, for the Ismail-Roe flux
    (its entropy is the same -rho s/(gamma-1)), plus symmetry and consistency."""
    ph = ir_phys()
    rng = np.random.default_rng(16)
    for _ in range(100):
        u1, u2, n = rand_state(rng), rand_state(rng), unit(rng)
        f, fs = vol(ph, u1, u2, n), vol(ph, u2, u1, n)
        assert np.abs(f - fs).max() < 1e-12 * np.abs(f).max()
        lhs = (entropy_vars(u2) - entropy_vars(u1)) @ f
        rhs = u2[1:4] @ n - u1[1:4] @ n
        assert abs(lhs - rhs) < 1e-8 * (1 + abs(rhs))
        assert np.abs(vol(ph, u1, u1, n) - euler_flux(u1, n)).max() < 1e-12 * np.abs(euler_flux(u1, n)).max()
    # a dropped term fails it: Chandrashekhar's flux is entropy conservative too, KG's is not
    ph.inv_vol_flux = 1
    u1, u2, n = rand_state(rng), rand_state(rng), unit(rng)
    lhs = (entropy_vars(u2) - entropy_vars(u1)) @ vol(ph, u1, u2, n)
    assert abs(lhs - (u2[1:4] @ n - u1[1:4] @ n)) > 1e-6


@pytest.mark.parametrize("N", [2, 4])
def test_kennedy_gruber_split_form(N):
    """With any D whose rows sum to zero, sum_j 2 D_ij F#(u_i,u_j) of a flux built from products of
    arithmetic means equals the split form of Kennedy & Gruber (Gassner et al. 2016, eq. 3.10):
    (a b c): 1/4 [D(abc) + a D(bc) + b D(ac) + c D(ab) + bc Da + ac Db + ab Dc], (a b):
    1/2 [D(ab) + a Db + b Da] — with a, b, c among rho, v.n, v_d, E/rho and p."""
    ph = phys()
    ph.inv_vol_flux = 1
    _, _, D = basis(N)
    rng = np.random.default_rng(17)
    u = [rand_state(rng) for _ in range(N + 1)]
    n = unit(rng)
    rho = np.array([q[0] for q in u])
    v = np.array([q[1:4] / q[0] for q in u])
    vn = v @ n
    eps = np.array([q[4] / q[0] for q in u])
    p = 0.4 * (np.array([q[4] for q in u]) - 0.5 * rho * (v * v).sum(1))
    d = lambda x: D @ x

    def tri(a, b, c):
        return 0.25 * (d(a * b * c) + a * d(b * c) + b * d(a * c) + c * d(a * b) + b * c * d(a) + a * c * d(b)
                       + a * b * d(c))

    def duo(a, b):
        return 0.5 * (d(a * b) + a * d(b) + b * d(a))

    expect = np.zeros((N + 1, 5))
    expect[:, 0] = duo(rho, vn)
    for k in range(3):
        expect[:, 1 + k] = tri(rho, vn, v[:, k]) + d(p) * n[k]
    expect[:, 4] = tri(rho, vn, eps) + duo(vn, p)
    got = np.zeros((N + 1, 5))
    for i in range(N + 1):
        for j in range(N + 1):
            got[i] += 2 * D[i, j] * vol(ph, u[i], u[j], n)
    assert np.abs(got - expect).max() < 1e-11 * np.abs(expect).max()


def test_kennedy_gruber_surface_stabilisation():
    # KG surface flux = KG two-point flux along x minus 1/2 max(|v_x| + a) (u_R - u_L)
    ph = phys(flux=7)
    ph.inv_vol_flux = 1
    rng = np.random.default_rng(18)
    x = np.array([1.0, 0, 0])
    for _ in range(20):
        ul, ur = rand_state(rng), rand_state(rng)
        c = lambda u: math.sqrt(1.4 * 0.4 * (u[4] - 0.5 * u[1:4] @ u[1:4] / u[0]) / u[0]) + abs(u[1] / u[0])
        expect = vol(ph, ul, ur, x) - 0.5 * max(c(ul), c(ur)) * (ur - ul)
        assert np.abs(surf(ph, 7, 0, ul, ur, x) - expect).max() < 1e-12 * np.abs(expect).max()


def cons_from_entropy(w, g=1.4):
    # inverse of entropy_vars: -rho/p = w4, v = -w_{1..3}/w4, s from w0
    rp = -w[4]
    v = w[1:4] / rp
    s = g - (g - 1) * (w[0] + 0.5 * rp * v @ v)
    # p rho^-g = e^s and rho/p = rp  =>  rho^(1-g) = rp e^s
    rho = (rp * math.exp(s)) ** (1 / (1 - g))
    p = rho / rp
    return np.array([rho, *(rho * v), p / (g - 1) + 0.5 * rho * v @ v])


def test_ismail_roe_surface_dissipation_is_entropy_jacobian():
    """IR surface flux = F#_x - 1/2 lambda du/dw (w_R - w_L), with du/dw the Jacobian of the
    conservative state in entropy variables at the IR mean state (rho^, v^, p2^) and
    lambda = |v^_x| + sqrt(gamma p2^/rho^). The mean state is read back from the two-point flux in
    three directions; du/dw comes from central differences of the inverse entropy map."""
    ph = ir_phys()
    rng = np.random.default_rng(19)
    E = np.eye(3)
    for _ in range(20):
        ul, ur = rand_state(rng), rand_state(rng)
        fx, fy = vol(ph, ul, ur, E[0]), vol(ph, ul, ur, E[1])
        fz = vol(ph, ul, ur, E[2])
        m = np.array([fx[0], fy[0], fz[0]])          # rho^ v^
        rh = m[0] * m[1] / fx[2]                     # rho^ v0 v1 = f_x[2]
        vh = m / rh
        hh = fx[4] / m[0]
        p2 = (hh - 0.5 * vh @ vh) * rh * 0.4 / 1.4
        uh = np.array([rh, *(rh * vh), p2 / 0.4 + 0.5 * rh * vh @ vh])
        wh = entropy_vars(uh)
        J = np.zeros((5, 5))
        for k in range(5):
            h = 1e-6 * max(1.0, abs(wh[k]))
            dw = np.zeros(5)
            dw[k] = h
            J[:, k] = (cons_from_entropy(wh + dw) - cons_from_entropy(wh - dw)) / (2 * h)
        lam = abs(vh[0]) + math.sqrt(1.4 * p2 / rh)
        expect = fx - 0.5 * lam * J @ (entropy_vars(ur) - entropy_vars(ul))
        got = surf(ph, 9, 0, ul, ur, E[0])
        assert np.abs(got - expect).max() < 1e-6 * np.abs(expect).max()
        # the dissipation removes entropy
        assert (entropy_vars(ur) - entropy_vars(ul)) @ (got - fx) <= 1e-12


def test_modified_steger_warming_matrix_functions():
    """Modified Steger-Warming: F = A+(u+) u_L + A-(u-) u_R with A+- = g+-(df/du) the matrix
    functions of the Euler flux Jacobian (numerical: central differences, numpy eigensolver)
    with g+-(x) = (x +- sqrt(x^2 + (0.3 a)^2))/2, at the pressure-weighted states of Druguet,
    Candler & Nompelis (2005): w = 1/2 / (1 + (1/2 dp / min p)^2), u+ = (1-w) L + w R,
    u- = (1-w) R + w L in rho, v, p."""
    ph = phys(flux=6)
    rng = np.random.default_rng(20)
    x = np.array([1.0, 0, 0])

    def prim(u):
        v = u[1:4] / u[0]
        return u[0], v, 0.4 * (u[4] - 0.5 * u[0] * v @ v)

    def matfun(u, sgn):
        J = np.zeros((5, 5))
        for k in range(5):
            h = 1e-7 * max(1.0, abs(u[k]))
            du = np.zeros(5)
            du[k] = h
            J[:, k] = (euler_flux(u + du, x) - euler_flux(u - du, x)) / (2 * h)
        lam, V = np.linalg.eig(J)
        lam = lam.real
        rho, v, p = prim(u)
        e = 0.3 * math.sqrt(1.4 * p / rho)
        g = 0.5 * (lam + sgn * np.sqrt(lam * lam + e * e))
        return (V @ np.diag(g) @ np.linalg.inv(V)).real

    for _ in range(20):
        ul, ur = rand_state(rng), rand_state(rng)
        (rl, vl, pl), (rr, vr, pr) = prim(ul), prim(ur)
        w = 0.5 / (1 + (0.5 * (pr - pl) / min(pl, pr)) ** 2)

        def mix(a, b):
            rho, v, p = (1 - w) * a[0] + w * b[0], (1 - w) * a[1] + w * b[1], (1 - w) * a[2] + w * b[2]
            return np.array([rho, *(rho * v), p / 0.4 + 0.5 * rho * v @ v])

        up, un = mix((rl, vl, pl), (rr, vr, pr)), mix((rr, vr, pr), (rl, vl, pl))
        expect = matfun(up, 1) @ ul + matfun(un, -1) @ ur
        got = surf(ph, 6, 0, ul, ur, x)
        assert np.abs(got - expect).max() < 1e-6 * np.abs(expect).max()


@pytest.mark.parametrize("scheme,other", [(4, 0), (5, 2), (8, 7)])
def test_rusanov_blends(scheme, other):
    # Rusanov gets the flux-blender weight, the other scheme its complement
    rng = np.random.default_rng(21)
    ph = phys()
    for _ in range(10):
        ul, ur, n = rand_state(rng), rand_state(rng), unit(rng)
        a = rng.uniform(0, 1)
        expect = a * surf(ph, 1, a, ul, ur, n) + (1 - a) * surf(ph, other, a, ul, ur, n)
        assert np.abs(surf(ph, scheme, a, ul, ur, n) - expect).max() < 1e-13 * np.abs(expect).max()


def test_dif_flux_closed_form():
    rng = np.random.default_rng(9)
    for _ in range(20):
        u, av, n = rand_state(rng), rng.standard_normal(9), unit(rng)
        tau = np.array([[av[0], av[1], av[2]], [av[1], av[3], av[4]], [av[2], av[4], av[5]]])
        q = av[6:9]
        v = u[1:4] / u[0]
        expect = np.array([0, *(tau @ n), -(q @ n) + v @ (tau @ n)])
        f = np.zeros(5)
        PO.load().oracle_dif_flux(p_(u), p_(av), p_(n), f.ctypes.data_as(dp))
        assert np.abs(f - expect).max() < 1e-13 * (1 + np.abs(expect).max())


# ------------------------------------------------------------------ RK coefficients
def test_rk_order_conditions():
    ao, ai, b = np.zeros(5), np.zeros(4), np.zeros(5)
    PO.load().oracle_rk4_coeffs(ao.ctypes.data_as(dp), ai.ctypes.data_as(dp), b.ctypes.data_as(dp))
    A = np.zeros((5, 5))  # Butcher tableau of the 3R+ scheme (dgsem/rk4_stage5_register3.h:34-53)
    A[1, 0] = ao[0]
    A[2, 0], A[2, 1] = ai[0], ao[1]
    A[3, 0], A[3, 1], A[3, 2] = b[0], ai[1], ao[2]
    A[4, 0], A[4, 1], A[4, 2], A[4, 3] = b[0], b[1], ai[2], ao[3]
    c = A.sum(axis=1)
    conds = [(b.sum(), 1), (b @ c, 1 / 2), (b @ c ** 2, 1 / 3), (b @ A @ c, 1 / 6), (b @ c ** 3, 1 / 4),
             (b @ (c * (A @ c)), 1 / 8), (b @ A @ c ** 2, 1 / 12), (b @ A @ A @ c, 1 / 24)]
    for got, exp in conds:
        assert abs(got - exp) < 1e-14
    assert ao[4] == b[4] and ai[3] == b[3]


def _butcher():
    ao, ai, b = np.zeros(5), np.zeros(4), np.zeros(5)
    PO.load().oracle_rk4_coeffs(ao.ctypes.data_as(dp), ai.ctypes.data_as(dp), b.ctypes.data_as(dp))
    A = np.zeros((5, 5))
    A[1, 0] = ao[0]
    A[2, 0], A[2, 1] = ai[0], ao[1]
    A[3, 0], A[3, 1], A[3, 2] = b[0], ai[1], ao[2]
    A[4, 0], A[4, 1], A[4, 2], A[4, 3] = b[0], b[1], ai[2], ao[3]
    return A, b


@pytest.mark.parametrize("which", ["periodic_cube", "shock_tube"])
def test_register_step_equals_butcher_form(which):
    """One step of the oracle's three-register update (dgsem/plens.cc:4334-4403) equals the Butcher
    form y1 = y0 + dt sum b_i k_i, k_i = f(y0 + dt sum_j a_ij k_j) built here from the oracle's
    right-hand side: a wrong register, stage coefficient or stage order fails. For y' = lambda y the
    Butcher form is the stability polynomial R(z) = sum_k z^k b^T A^(k-1) 1, whose first five
    coefficients are 1/k! (fourth order)."""
    from plens_amd import cases as CS
    if which == "periodic_cube":
        case, dt = CS.periodic_cube(n=2, degree=3), 0.004
    else:
        case, dt = CS.shock_tube(degree=2, nx=6, ny=2, nz=2), 0.002
    A, b = _butcher()
    o = PO.OraclePlens(case)
    y0 = case.Q0.copy()
    ks = []
    for i in range(5):
        o.set_state(y0 + dt * sum(A[i, j] * ks[j] for j in range(i)) if i else y0)
        ks.append(o.calc_rhs()[0])
    y1 = y0 + dt * sum(b[i] * ks[i] for i in range(5))
    o.set_state(y0)
    o.update_fixed_dt(dt)
    got = o.get_state()
    # relative to the step's increment: about 1e-13 from rounding, 2e-3 with two weights swapped
    assert np.linalg.norm(got - y1) <= 1e-12 * np.linalg.norm(y1 - y0)
    one = np.ones(5)
    coeff = [b @ np.linalg.matrix_power(A, k - 1) @ one for k in range(1, 5)]
    assert np.allclose(coeff, [1, 1 / 2, 1 / 6, 1 / 24], rtol=0, atol=1e-15)


# ------------------------------------------------------------------ checker hook: given blender values
def test_given_alpha_hook():
    """the oracle's own blender values handed back reproduce its RHS bit for bit; other values change
    only the blended terms (a cell with alpha = 0 everywhere gives the pure high-order RHS)"""
    from plens_amd import shock_tube
    case = shock_tube(nx=16, ny=2, nz=2)
    o = PO.OraclePlens(case)
    r0, a0, ad0 = o.calc_rhs()
    assert (a0 > 0).any()
    o.set_alpha(a0, ad0)
    r1, a1, ad1 = o.calc_rhs()
    assert np.array_equal(r0, r1) and np.array_equal(a0, a1) and np.array_equal(ad0, ad1)
    z = np.zeros_like(a0)
    o.set_alpha(z, z)
    r2, a2, _ = o.calc_rhs()
    assert np.array_equal(a2, z) and not np.array_equal(r2, r0)
    o.set_alpha()
    r3, a3, _ = o.calc_rhs()
    assert np.array_equal(r3, r0) and np.array_equal(a3, a0)
