"""Mie series for a homogeneous sphere in a homogeneous medium (Bohren & Huffman, "Absorption
and Scattering of Light by Small Particles", 1983, ch. 4): the scattered field outside the
sphere for an x-polarised plane wave travelling along +z, E_i = x exp(i k z), time factor
exp(-i w t).  H is returned in the reference's units, Z0 H (a plane wave in a medium of index
n0 has Z0 H = n0 z x E).  Test infrastructure: an independent physics check of the
non-periodic solve, not a restatement of the reference."""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def _coeffs(x, m, nmax):
    n = np.arange(1, nmax + 1)
    mx = m * x

    def psi(nn, z):
        return z * spherical_jn(nn, z)

    def dpsi(nn, z):
        return spherical_jn(nn, z) + z * spherical_jn(nn, z, derivative=True)

    def xi(nn, z):
        return z * (spherical_jn(nn, z) + 1j * spherical_yn(nn, z))

    def dxi(nn, z):
        h = spherical_jn(nn, z) + 1j * spherical_yn(nn, z)
        dh = spherical_jn(nn, z, derivative=True) + 1j * spherical_yn(nn, z, derivative=True)
        return h + z * dh

    a = (m * psi(n, mx) * dpsi(n, x) - psi(n, x) * dpsi(n, mx)) / (m * psi(n, mx) * dxi(n, x) - xi(n, x) * dpsi(n, mx))
    b = (psi(n, mx) * dpsi(n, x) - m * psi(n, x) * dpsi(n, mx)) / (psi(n, mx) * dxi(n, x) - m * xi(n, x) * dpsi(n, mx))
    return n, a, b


def scattered(r, radius, n_sphere, n0, lam0, nmax=None):
    """Scattered E and Z0 H at a point r outside the sphere (centre at the origin)."""
    k = 2.0 * np.pi * n0 / lam0
    x = k * radius
    m = n_sphere / n0
    if nmax is None:
        nmax = int(x + 4.05 * x ** (1.0 / 3.0) + 8)
    n, a, b = _coeffs(x, m, nmax)
    r = np.asarray(r, float)
    R = np.linalg.norm(r)
    th = np.arccos(r[2] / R)
    ph = np.arctan2(r[1], r[0])
    mu, st = np.cos(th), np.sin(th)
    pi = np.zeros(nmax + 1)
    tau = np.zeros(nmax + 1)
    pi[1] = 1.0
    tau[1] = mu
    for j in range(2, nmax + 1):
        pi[j] = (2 * j - 1) / (j - 1) * mu * pi[j - 1] - j / (j - 1) * pi[j - 2]
        tau[j] = j * mu * pi[j] - (j + 1) * pi[j - 1]
    pi, tau = pi[1:], tau[1:]
    rho = k * R
    h = spherical_jn(n, rho) + 1j * spherical_yn(n, rho)
    dh = spherical_jn(n, rho, derivative=True) + 1j * spherical_yn(n, rho, derivative=True)
    drz = (h + rho * dh) / rho  # [rho z_n(rho)]' / rho
    En = (1j ** n) * (2 * n + 1) / (n * (n + 1))
    cp, sp = np.cos(ph), np.sin(ph)
    # N_e1n, M_o1n, N_o1n, M_e1n in (r, theta, phi) components
    Ne = np.array([cp * n * (n + 1) * st * pi * h / rho, cp * tau * drz, -sp * pi * drz])
    Mo = np.array([0 * n, cp * pi * h, -sp * tau * h])
    No = np.array([sp * n * (n + 1) * st * pi * h / rho, sp * tau * drz, cp * pi * drz])
    Me = np.array([0 * n, -sp * pi * h, -cp * tau * h])
    Es = (En * (1j * a * Ne - b * Mo)).sum(1)
    Hs = n0 * (En * (1j * b * No + a * Me)).sum(1)
    er = np.array([st * cp, st * sp, mu])
    et = np.array([mu * cp, mu * sp, -st])
    ep = np.array([-sp, cp, 0.0])
    T = np.stack([er, et, ep], 1)
    return T @ Es, T @ Hs


def incident(r, n0, lam0):
    k = 2.0 * np.pi * n0 / lam0
    e = np.exp(1j * k * r[2])
    return np.array([e, 0, 0], complex), n0 * np.array([0, e, 0], complex)
