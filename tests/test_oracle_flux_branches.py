"""Pin the oracle's Roe entropy fix and HLLC star states independently of
their closed forms in fluxes.py:

* Roe (fluxes.py:54-98): F = (F- + F+)/2 - |A|~ (q+ - q-)/2 with |A|~ built
  from a numerical eigendecomposition of the normal-flux Jacobian at the
  Roe-averaged state (Jacobian by complex-step differentiation of a
  separately written flux), Harten's fix applied to the smallest and largest
  eigenvalues with delta = 0.1 (|mid| + half-spread) of the spectrum
  (fluxes.py:21,73-77); states chosen so the fix is active on u_n - c or
  on u_n + c (transonic rarefactions) and inactive. The Roe property
  A(q~)(q+ - q-) = F(q+) - F(q-) pins the average itself.
* HLLC (fluxes.py:101-143): Toro's one-dimensional HLLC (Riemann Solvers and
  Numerical Methods for Fluid Dynamics, 3rd ed., eqs. 10.37-10.39) in the
  face frame (normal, tangential), rotated back, with the reference's
  Einfeldt/Roe wave-speed estimates; both star branches with a contact speed
  S* different from the side's normal velocity, all four components. Both
  star states move with normal velocity S* and satisfy the integral
  consistency condition of the HLLC fan (Toro eq. 10.29).
"""

import numpy as np
import pytest

from oracle.physics import hllc, primitive_to_conserved, roe

GAMMA = 1.4


def cons(p):
    return np.asarray(primitive_to_conserved(*p, GAMMA), dtype=np.float64)


def flux_n(q, nx, ny):
    """F(q).n written out for complex arguments."""
    rho, mx, my, e = q
    u, v = mx / rho, my / rho
    p = (GAMMA - 1.0) * (e - 0.5 * (mx * u + my * v))
    un = u * nx + v * ny
    return np.array([rho * un, mx * un + p * nx, my * un + p * ny, (e + p) * un])


def jacobian(q, nx, ny, h=1e-30):
    a = np.empty((4, 4))
    for k in range(4):
        z = q.astype(complex)
        z[k] += 1j * h
        a[:, k] = flux_n(z, nx, ny).imag / h
    return a


def roe_state(qm, qp):
    """Conserved state at the Roe average (sqrt(rho)-weighted u, v, H)."""
    def prim(q):
        rho, u, v = q[0], q[1] / q[0], q[2] / q[0]
        p = (GAMMA - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v))
        return rho, u, v, (q[3] + p) / rho
    rm, um, vm, hm = prim(qm)
    rp, up, vp, hp = prim(qp)
    wm, wp = np.sqrt(rm), np.sqrt(rp)
    u, v, hh = [(wm * a + wp * b) / (wm + wp) for a, b in ((um, up), (vm, vp), (hm, hp))]
    rho = wm * wp
    p = (GAMMA - 1.0) / GAMMA * rho * (hh - 0.5 * (u * u + v * v))
    return np.array([rho, rho * u, rho * v, rho * hh - p])


def roe_by_eigendecomposition(qm, qp, nx, ny, delta_factor=0.1):
    """|A| by Sylvester's formula over the three distinct eigenvalues
    u_n - c, u_n (double: entropy and shear waves), u_n + c of the Roe
    matrix: f(A) = sum_i f(l_i) prod_{j != i} (A - l_j) / (l_i - l_j). It
    needs no eigenvectors, which are ill-conditioned at the double root."""
    a = jacobian(roe_state(qm, qp), nx, ny)
    lam = np.sort(np.linalg.eigvals(a).real)
    distinct = np.array([lam[0], 0.5 * (lam[1] + lam[2]), lam[3]])
    mid, half = distinct[1], 0.5 * (lam[3] - lam[0])  # u_n and c of the average
    delta = delta_factor * (abs(mid) + half)
    mag = np.abs(distinct)
    for k in (0, 2):  # acoustic waves only
        if mag[k] < delta:
            mag[k] = (distinct[k] ** 2 + delta ** 2) / (2.0 * delta)
    absa = np.zeros((4, 4))
    eye = np.eye(4)
    for i in range(3):
        proj = eye.copy()
        for j in range(3):
            if j != i:
                proj = proj @ (a - distinct[j] * eye) / (distinct[i] - distinct[j])
        absa += mag[i] * proj
    return 0.5 * (flux_n(qm, nx, ny) + flux_n(qp, nx, ny)) - 0.5 * absa @ (qp - qm), lam, delta


# (left, right, normal, which acoustic wave the fix acts on)
ROE_CASES = [
    ((1.0, 0.9, 0.1, 1.0), (0.6, 1.25, 0.05, 0.5), (1.0, 0.0), "u-c"),
    ((1.0, 0.9, 0.3, 1.0), (0.7, 1.1, 0.2, 0.6), (0.8, 0.6), "u-c"),
    ((0.6, -1.25, 0.2, 0.5), (1.0, -0.95, 0.1, 1.0), (1.0, 0.0), "u+c"),
    ((0.6, 0.2, -1.25, 0.5), (1.0, 0.1, -0.95, 1.0), (0.0, 1.0), "u+c"),
    ((1.3, 0.7, -0.4, 2.1), (0.4, -1.2, 0.9, 0.3), (0.6, -0.8), "none"),
    ((1.0, 0.0, 0.0, 1.0), (0.125, 0.0, 0.0, 0.1), (1.0, 0.0), "none"),
]


@pytest.mark.parametrize("left,right,normal,wave", ROE_CASES)
def test_roe_equals_eigendecomposition_with_harten_fix(left, right, normal, wave):
    qm, qp = cons(left), cons(right)
    nx, ny = normal
    ref, lam, delta = roe_by_eigendecomposition(qm, qp, nx, ny)
    active = [abs(lam[0]) < delta, abs(lam[3]) < delta]
    assert active == [wave == "u-c", wave == "u+c"]
    got = roe(qm, qp, nx, ny, GAMMA)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-12 * scale
    if wave != "none":
        # The fix moves the flux: a different delta would be seen.
        for wrong in (0.05, 0.2):
            other, _, _ = roe_by_eigendecomposition(qm, qp, nx, ny, delta_factor=wrong)
            assert np.abs(other - ref).max() > 1e-6 * scale


@pytest.mark.parametrize("left,right,normal,wave", ROE_CASES)
def test_roe_average_has_the_roe_property(left, right, normal, wave):
    qm, qp = cons(left), cons(right)
    nx, ny = normal
    a = jacobian(roe_state(qm, qp), nx, ny)
    df = flux_n(qp, nx, ny) - flux_n(qm, nx, ny)
    assert np.allclose(a @ (qp - qm), df, rtol=0, atol=1e-13 * np.abs(df).max())


def hllc_toro(qm, qp, nx, ny):
    """Toro's HLLC in the face frame with the reference's wave speeds."""
    def rotate(q):
        rho, mx, my, e = q
        return np.array([rho, mx * nx + my * ny, -mx * ny + my * nx, e])

    def prim(w):
        rho, un, ut = w[0], w[1] / w[0], w[2] / w[0]
        p = (GAMMA - 1.0) * (w[3] - 0.5 * rho * (un * un + ut * ut))
        return rho, un, ut, p

    def f1d(w):
        rho, un, ut, p = prim(w)
        return np.array([rho * un, rho * un * un + p, rho * un * ut, (w[3] + p) * un])

    wl, wr = rotate(qm), rotate(qp)
    rl, ul, tl, pl = prim(wl)
    rr, ur, tr, pr = prim(wr)
    cl, cr = np.sqrt(GAMMA * pl / rl), np.sqrt(GAMMA * pr / rr)
    sqrl, sqrr = np.sqrt(rl), np.sqrt(rr)
    hl, hr = (wl[3] + pl) / rl, (wr[3] + pr) / rr
    ub = (sqrl * ul + sqrr * ur) / (sqrl + sqrr)
    tb = (sqrl * tl + sqrr * tr) / (sqrl + sqrr)
    hb = (sqrl * hl + sqrr * hr) / (sqrl + sqrr)
    cb = np.sqrt((GAMMA - 1.0) * (hb - 0.5 * (ub * ub + tb * tb)))
    sl, sr = min(ul - cl, ub - cb), max(ur + cr, ub + cb)
    ss = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur))  # 10.37

    def star(w, rho, un, ut, p, s):  # 10.39
        k = rho * (s - un) / (s - ss)
        return k * np.array([1.0, ss, ut, w[3] / rho + (ss - un) * (ss + p / (rho * (s - un)))])

    ustar_l, ustar_r = star(wl, rl, ul, tl, pl, sl), star(wr, rr, ur, tr, pr, sr)
    if sl >= 0:
        f, branch = f1d(wl), "left"
    elif ss >= 0:
        f, branch = f1d(wl) + sl * (ustar_l - wl), "star-left"
    elif sr > 0:
        f, branch = f1d(wr) + sr * (ustar_r - wr), "star-right"
    else:
        f, branch = f1d(wr), "right"
    out = np.array([f[0], f[1] * nx - f[2] * ny, f[1] * ny + f[2] * nx, f[3]])
    return out, branch, (ustar_l, ustar_r, ss, ul, ur)


HLLC_CASES = [
    ((1.0, 0.3, 0.2, 1.0), (0.5, 0.1, -0.4, 0.4), (1.0, 0.0), "star-left"),
    ((1.0, 0.3, 0.2, 1.0), (0.5, 0.1, -0.4, 0.4), (0.6, -0.8), "star-left"),
    ((0.5, -0.4, 0.3, 0.4), (1.0, -0.6, 0.2, 1.0), (0.8, 0.6), "star-right"),
    ((0.4, -0.9, 0.5, 0.3), (1.3, -0.2, -0.7, 2.1), (0.0, 1.0), "star-right"),
    ((1.3, 0.7, -0.4, 2.1), (0.4, -1.2, 0.9, 0.3), (0.6, -0.8), None),
    ((1.0, 3.0, 0.5, 1.0), (0.8, 2.5, 0.0, 1.2), (1.0, 0.0), "left"),
    ((0.8, -2.5, 0.0, 1.2), (1.0, -3.0, 0.5, 1.0), (1.0, 0.0), "right"),
]


@pytest.mark.parametrize("left,right,normal,branch", HLLC_CASES)
def test_hllc_equals_toro_in_the_face_frame(left, right, normal, branch):
    qm, qp = cons(left), cons(right)
    nx, ny = normal
    ref, got_branch, (usl, usr, ss, ul, ur) = hllc_toro(qm, qp, nx, ny)
    if branch is not None:
        assert got_branch == branch
    if got_branch.startswith("star"):
        # a real shift: the star state differs from the side's normal velocity
        un = ul if got_branch == "star-left" else ur
        assert abs(ss - un) > 0.05
    out = hllc(qm, qp, nx, ny, GAMMA)
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("left,right,normal,branch", HLLC_CASES[:4])
def test_hllc_star_states_are_consistent(left, right, normal, branch):
    """Both star states move with normal velocity S*, and they satisfy the
    integral consistency condition of the HLLC fan (Toro eq. 10.29):
    (S* - S_L) U*_L + (S_R - S*) U*_R = S_R U_R - S_L U_L - (F_R - F_L)."""
    qm, qp = cons(left), cons(right)
    nx, ny = normal
    _, _, (usl, usr, ss, _, _) = hllc_toro(qm, qp, nx, ny)
    assert usl[1] / usl[0] == pytest.approx(ss, rel=1e-13)
    assert usr[1] / usr[0] == pytest.approx(ss, rel=1e-13)
    wl = np.array([qm[0], qm[1] * nx + qm[2] * ny, -qm[1] * ny + qm[2] * nx, qm[3]])
    wr = np.array([qp[0], qp[1] * nx + qp[2] * ny, -qp[1] * ny + qp[2] * nx, qp[3]])

    def f1d(w):
        un, ut = w[1] / w[0], w[2] / w[0]
        p = (GAMMA - 1.0) * (w[3] - 0.5 * w[0] * (un * un + ut * ut))
        return np.array([w[1], w[1] * un + p, w[2] * un, (w[3] + p) * un])

    # recover S_L, S_R from the star states' definition U*_K = rho_K (S_K - u_K)/(S_K - S*) (1, ...)
    def speed(w, ustar):
        k, un = ustar[0] / w[0], w[1] / w[0]  # k = (S - u)/(S - S*)
        return (k * ss - un) / (k - 1.0)
    sl, sr = speed(wl, usl), speed(wr, usr)
    lhs = (ss - sl) * usl + (sr - ss) * usr
    rhs = sr * wr - sl * wl - (f1d(wr) - f1d(wl))
    assert np.allclose(lhs, rhs, rtol=0, atol=1e-13 * np.abs(rhs).max())
