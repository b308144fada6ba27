"""Internal field of a homogeneous sphere under a plane wave (Mie series).

Bohren & Huffman, "Absorption and Scattering of Light by Small Particles",
eqs. 4.40 and 4.52 (internal coefficients c_n, d_n, mu = mu_1 = 1), with the
exp(-i omega t) convention and incident field x_hat exp(i k z).  The VIE code
uses exp(+j omega t), incident x_hat exp(-j k0 z) and complex relative
permittivity eps_r - j sigma/(omega e0); its field is the complex conjugate of
the B&H field computed with the conjugate refractive index.
Used only as a physics sanity check of the assembled operator.
"""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def _jn(n, z, d=False):
    return spherical_jn(n, z, derivative=d)


def _hn(n, z, d=False):
    return spherical_jn(n, z, derivative=d) + 1j * spherical_yn(n, z, derivative=d)


def internal_field_bh(points, k, a, m, nmax=None):
    """E inside the sphere (B&H convention), points (N, 3) with |r| < a."""
    x = k * a
    if nmax is None:
        nmax = int(np.ceil(abs(x * m) + 4 * abs(x * m) ** (1 / 3) + 8))
    r = np.linalg.norm(points, axis=1)
    cost = points[:, 2] / r
    sint = np.sqrt(np.maximum(0.0, 1 - cost ** 2))
    phi = np.arctan2(points[:, 1], points[:, 0])
    cphi, sphi = np.cos(phi), np.sin(phi)
    rho = m * k * r
    Er = np.zeros(len(r), complex)
    Et = np.zeros(len(r), complex)
    Ep = np.zeros(len(r), complex)
    pi_prev, pi_cur = np.zeros_like(cost), np.ones_like(cost)   # pi_0, pi_1
    for n in range(1, nmax + 1):
        if n > 1:
            pi_new = ((2 * n - 1) / (n - 1)) * cost * pi_cur - (n / (n - 1)) * pi_prev
            pi_prev, pi_cur = pi_cur, pi_new
        pin = pi_cur
        taun = n * cost * pin - (n + 1) * pi_prev
        jx, jxd = _jn(n, x), _jn(n, x, True)
        hx, hxd = _hn(n, x), _hn(n, x, True)
        jm, jmd = _jn(n, m * x), _jn(n, m * x, True)
        xjx_d = jx + x * jxd                 # [x j_n(x)]'
        xhx_d = hx + x * hxd                 # [x h_n(x)]'
        mxjm_d = jm + m * x * jmd            # [mx j_n(mx)]'
        cn = (jx * xhx_d - hx * xjx_d) / (jm * xhx_d - hx * mxjm_d)
        dn = (m * jx * xhx_d - m * hx * xjx_d) / (m ** 2 * jm * xhx_d - hx * mxjm_d)
        En = (1j ** n) * (2 * n + 1) / (n * (n + 1))
        z = _jn(n, rho)
        dz = (z + rho * _jn(n, rho, True)) / rho          # [rho z]' / rho
        # M_o1n and N_e1n (B&H 4.50)
        Mt, Mp = cphi * pin * z, -sphi * taun * z
        Nr = cphi * n * (n + 1) * sint * pin * z / rho
        Nt, Np = cphi * taun * dz, -sphi * pin * dz
        Er += En * (-1j * dn * Nr)
        Et += En * (cn * Mt - 1j * dn * Nt)
        Ep += En * (cn * Mp - 1j * dn * Np)
    ex = Er * sint * cphi + Et * cost * cphi - Ep * sphi
    ey = Er * sint * sphi + Et * cost * sphi + Ep * cphi
    ez = Er * cost - Et * sint
    return np.stack([ex, ey, ez], axis=1)


def internal_field(points, k0, a, eps_r, sigma, omega, e0):
    """E inside the sphere in the VIE convention (exp(+j omega t))."""
    eps_c = eps_r - 1j * sigma / (omega * e0)
    m = np.sqrt(eps_c)
    return np.conj(internal_field_bh(points, k0, a, np.conj(m)))
