"""The oracle's HLLC and AUSM+-up fluxes on Toro's five Riemann test states (Toro, Riemann Solvers
and Numerical Methods for Fluid Dynamics, 3rd ed., Table 4.1), against textbook code written here
from the publications, not from the reference:

- HLLC: Toro sec. 10.4 (star states, eq. 10.39) with the adaptive noniterative wave-speed estimate
  of secs. 9.5.2 and 10.6 (PVRS inside [p_min, p_max] when p_max/p_min < 2, else TRRS below p_min,
  else TSRS from max(0, p_PVRS)), which is the estimate the reference selects
  (modelling/navier_stokes.cc:956-986, tested at navier_stokes.cc:1713-1724).
- AUSM+-up: Liou, J. Comput. Phys. 214 (2006) 137-170, eqs. 18-24 and 73-75 in their expanded
  polynomial forms, with K_p = 0.25, K_u = 0.75, sigma = 1, beta = 1/8, M_inf = 0.3 and
  a_1/2 = (a_L + a_R)/2 (modelling/ausm_functions.h, navier_stokes.cc:1070-1121). The reference
  writes alpha = 3/16*(-4 + 5 f_a^2) in integer arithmetic, which is 0; the textbook code takes
  alpha as an argument, and the test checks that the paper's alpha would give a different flux, so
  the comparison would see that term.

Each state pair is also shifted by Galilean velocities so that every HLLC branch (left, left star,
right star, right) and both AUSM mass-flux signs are reached, given tangential velocities, and
rotated onto random normals. CPU only."""
import math

import numpy as np
import pytest

from tests.test_oracle_pins import phys, surf

G = 1.4
# (rho, u, p) left | right, Toro Table 4.1
TORO = [
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1)),
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4)),
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01)),
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0)),
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950)),
]


def cons(rho, u, v, w, p, g=G):
    return np.array([rho, rho * u, rho * v, rho * w, p / (g - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def xflux(U, g=G):
    rho, u, v, w = U[0], U[1] / U[0], U[2] / U[0], U[3] / U[0]
    p = (g - 1) * (U[4] - 0.5 * rho * (u * u + v * v + w * w))
    return np.array([rho * u, rho * u * u + p, rho * u * v, rho * u * w, u * (U[4] + p)])


def prim(U, g=G):
    rho = U[0]
    u, v, w = U[1:4] / rho
    p = (g - 1) * (U[4] - 0.5 * rho * (u * u + v * v + w * w))
    return rho, u, v, w, p


def toro_hllc(UL, UR, g=G):
    """HLLC flux across x (Toro 10.4-10.6); returns the flux and the branch taken (0..3)."""
    rl, ul, vl, wl, pl = prim(UL, g)
    rr, ur, vr, wr, pr = prim(UR, g)
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    pmin, pmax = min(pl, pr), max(pl, pr)
    ppv = 0.5 * (pl + pr) - 0.5 * (ur - ul) * (0.5 * (rl + rr)) * (0.5 * (al + ar))  # Toro eq. 9.20
    if pmax / pmin < 2.0 and pmin <= ppv <= pmax:
        ps = ppv
    elif ppv < pmin:  # two-rarefaction, Toro eq. 9.32
        z = (g - 1) / (2 * g)
        ps = ((al + ar - 0.5 * (g - 1) * (ur - ul)) / (al / pl ** z + ar / pr ** z)) ** (1 / z)
    else:  # two-shock, Toro eqs. 9.42-9.43
        p0 = max(0.0, ppv)
        gk = lambda r, p: math.sqrt((2 / ((g + 1) * r)) / (p0 + (g - 1) / (g + 1) * p))
        gl, gr = gk(rl, pl), gk(rr, pr)
        ps = (gl * pl + gr * pr - (ur - ul)) / (gl + gr)
    q = lambda p: 1.0 if ps <= p else math.sqrt(1 + (g + 1) / (2 * g) * (ps / p - 1))  # Toro eq. 10.59
    SL, SR = ul - al * q(pl), ur + ar * q(pr)
    Ss = (pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))  # eq. 10.37

    def star(U, r, u, v, w, p, S):  # Toro eq. 10.39
        c = r * (S - u) / (S - Ss)
        return c * np.array([1.0, Ss, v, w, U[4] / r + (Ss - u) * (Ss + p / (r * (S - u)))])

    if SL > 0:
        return xflux(UL, g), 0
    if Ss > 0:
        return xflux(UL, g) + SL * (star(UL, rl, ul, vl, wl, pl, SL) - UL), 1
    if SR > 0:
        return xflux(UR, g) + SR * (star(UR, rr, ur, vr, wr, pr, SR) - UR), 2
    return xflux(UR, g), 3


def liou_ausm_plus_up(UL, UR, alpha_paper=False, g=G):
    """AUSM+-up flux across x (Liou 2006); returns the flux and the sign of the interface Mach number."""
    Kp, Ku, sigma, beta, Minf = 0.25, 0.75, 1.0, 0.125, 0.3
    rl, ul, vl, wl, pl = prim(UL, g)
    rr, ur, vr, wr, pr = prim(UR, g)
    a = 0.5 * (math.sqrt(g * pl / rl) + math.sqrt(g * pr / rr))
    Ml, Mr = ul / a, ur / a
    Mb2 = 0.5 * (Ml * Ml + Mr * Mr)  # eq. 70
    Mo = math.sqrt(min(1.0, max(Mb2, Minf * Minf)))  # eq. 71
    fa = Mo * (2 - Mo)  # eq. 72
    alpha = 3.0 / 16.0 * (-4 + 5 * fa * fa) if alpha_paper else 0.0

    def M4(M, s):  # eq. 20, expanded: +-(M +- 1)^2/4 +- beta (M^2 - 1)^2
        if abs(M) >= 1:
            return 0.5 * (M + s * abs(M))
        return s * 0.25 * (M + s) ** 2 + s * beta * (M * M - 1) ** 2

    def P5(M, s):  # eq. 24, expanded: (M +- 1)^2 (2 -+ M)/4 +- alpha M (M^2 - 1)^2
        if abs(M) >= 1:
            return 0.5 * (M + s * abs(M)) / M
        return 0.25 * (M + s) ** 2 * (2 - s * M) + s * alpha * M * (M * M - 1) ** 2

    rho_h = 0.5 * (rl + rr)
    Mh = M4(Ml, 1) + M4(Mr, -1) - Kp / fa * max(1 - sigma * Mb2, 0.0) * (pr - pl) / (rho_h * a * a)  # eq. 73
    Pp, Pm = P5(Ml, 1), P5(Mr, -1)
    ph = Pp * pl + Pm * pr - Ku * Pp * Pm * (rl + rr) * (fa * a) * (ur - ul)  # eq. 75
    if Mh > 0:
        md, psi = a * Mh * rl, np.array([1.0, ul, vl, wl, (UL[4] + pl) / rl])
    else:
        md, psi = a * Mh * rr, np.array([1.0, ur, vr, wr, (UR[4] + pr) / rr])
    return md * psi + ph * np.array([0.0, 1.0, 0.0, 0.0, 0.0]), 1 if Mh > 0 else -1


def _cases():
    rng = np.random.default_rng(11)
    out = []
    for (L, R) in TORO:
        aL = math.sqrt(G * L[2] / L[0])
        for shift in (0.0, 3.0 * aL, -3.0 * aL, 25.0, -25.0):
            for swap in (False, True):
                l, r = (R, L) if swap else (L, R)
                sgn = -1.0 if swap else 1.0  # mirror the velocities with the states
                t = rng.uniform(-1, 1, 4) * aL
                UL = cons(l[0], sgn * l[1] + shift, t[0], t[1], l[2])
                UR = cons(r[0], sgn * r[1] + shift, t[2], t[3], r[2])
                out.append((UL, UR))
    return out


def _rot(n):
    """Orthogonal matrix whose first column is n."""
    A = np.linalg.qr(np.column_stack([n, np.eye(3)[:, :2]]))[0]
    return A * np.sign(A[:, 0] @ n)


def _rotated(U, A):
    return np.array([U[0], *(A @ U[1:4]), U[4]])


def test_hllc_on_toro_states():
    ph = phys(flux=0)
    branches = set()
    rng = np.random.default_rng(12)
    for UL, UR in _cases():
        ft, br = toro_hllc(UL, UR)
        branches.add(br)
        f = surf(ph, 0, 0.0, UL, UR, np.array([1.0, 0.0, 0.0]))
        scale = np.abs(ft).max()
        assert np.abs(f - ft).max() <= 1e-13 * scale, (UL, UR, f, ft)
        # the same Riemann problem across a random normal
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        A = _rot(n)
        fr = surf(ph, 0, 0.0, _rotated(UL, A), _rotated(UR, A), n)
        assert np.abs(fr - _rotated(ft, A)).max() <= 1e-12 * scale
    assert branches == {0, 1, 2, 3}


def test_ausm_plus_up_on_toro_states():
    ph = phys(flux=2)
    signs, differs = set(), 0
    rng = np.random.default_rng(13)
    for UL, UR in _cases():
        ft, s = liou_ausm_plus_up(UL, UR)
        signs.add(s)
        f = surf(ph, 2, 0.0, UL, UR, np.array([1.0, 0.0, 0.0]))
        scale = np.abs(ft).max()
        assert np.abs(f - ft).max() <= 1e-13 * scale, (UL, UR, f, ft)
        fp, _ = liou_ausm_plus_up(UL, UR, alpha_paper=True)
        differs += np.abs(fp - ft).max() > 1e-10 * scale
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        A = _rot(n)
        fr = surf(ph, 2, 0.0, _rotated(UL, A), _rotated(UR, A), n)
        assert np.abs(fr - _rotated(ft, A)).max() <= 1e-12 * scale
    assert signs == {-1, 1}
    assert differs > 10  # the alpha term is visible on these states


@pytest.mark.parametrize("scheme,blend", [(4, 0.0), (4, 0.37), (4, 1.0), (5, 0.0), (5, 0.61), (5, 1.0)])
def test_rusanov_blends_on_toro_states(scheme, blend):
    """Rusanov-HLLC and Rusanov-AUSM+-up: the flux-blender weight on Rusanov, the rest on the
    textbook flux (navier_stokes.cc:1132-1159)."""
    ph = phys(flux=scheme)
    for UL, UR in _cases():
        fb = toro_hllc(UL, UR)[0] if scheme == 4 else liou_ausm_plus_up(UL, UR)[0]
        al = math.sqrt(G * prim(UL)[4] / UL[0]) + abs(UL[1] / UL[0])
        ar = math.sqrt(G * prim(UR)[4] / UR[0]) + abs(UR[1] / UR[0])
        fr = 0.5 * (xflux(UL) + xflux(UR)) - 0.5 * max(al, ar) * (UR - UL)
        expect = blend * fr + (1 - blend) * fb
        f = surf(ph, scheme, blend, UL, UR, np.array([1.0, 0.0, 0.0]))
        assert np.abs(f - expect).max() <= 1e-13 * max(np.abs(fb).max(), np.abs(fr).max())
