"""Mie series for the bistatic RCS of a PEC sphere (textbook closed form,
Bohren & Huffman, "Absorption and Scattering of Light by Small Particles",
ch. 4, in the limit of infinite refractive index).

Used to pin the whole oracle chain (fill -> solve -> RCS) against physics.
Conventions follow the reference's incidence (EMSolver/EM.cpp:33-34): the
plane wave travels along -z (comes from +z) with E along -x for pole=0.
Scattering direction s = (sin t cos p, sin t sin p, cos t); t=0 is back
scatter, so the scattering angle from the propagation direction is pi - t.
"""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def pec_sphere_rcs(radius, k, theta_deg, phi_deg, nmax=None):
    """Bistatic sigma (m^2) at scattering angles (theta, phi) in degrees."""
    x = k * radius
    if nmax is None:
        nmax = int(x + 4.05 * x ** (1.0 / 3.0) + 10)
    n = np.arange(1, nmax + 1)
    jn = spherical_jn(n, x)
    yn = spherical_yn(n, x)
    jnp = spherical_jn(n, x, derivative=True)
    ynp = spherical_yn(n, x, derivative=True)
    psi = x * jn
    dpsi = jn + x * jnp
    xi = x * (jn + 1j * yn)
    dxi = (jn + 1j * yn) + x * (jnp + 1j * ynp)
    a = dpsi / dxi           # PEC limit of B&H a_n
    b = psi / xi             # PEC limit of B&H b_n

    theta = np.radians(np.atleast_1d(np.asarray(theta_deg, dtype=float)))
    phi = np.radians(np.atleast_1d(np.asarray(phi_deg, dtype=float)))
    theta, phi = np.broadcast_arrays(theta, phi)
    mu = np.cos(np.pi - theta)
    S1 = np.zeros(mu.shape, dtype=complex)
    S2 = np.zeros(mu.shape, dtype=complex)
    pi_prev = np.zeros_like(mu)      # pi_0
    pi_cur = np.ones_like(mu)        # pi_1
    for i, nn in enumerate(n):
        if nn > 1:
            pi_new = ((2 * nn - 1) / (nn - 1)) * mu * pi_cur - (nn / (nn - 1)) * pi_prev
            pi_prev, pi_cur = pi_cur, pi_new
        tau = nn * mu * pi_cur - (nn + 1) * pi_prev
        fac = (2 * nn + 1) / (nn * (nn + 1))
        S1 += fac * (a[i] * pi_cur + b[i] * tau)
        S2 += fac * (a[i] * tau + b[i] * pi_cur)
    amp2 = np.abs(S2) ** 2 * np.cos(phi) ** 2 + np.abs(S1) ** 2 * np.sin(phi) ** 2
    return 4.0 * np.pi * amp2 / (k * k)


def pec_sphere_rcs_db(radius, k, theta_deg, phi_deg):
    return 10.0 * np.log10(pec_sphere_rcs(radius, k, theta_deg, phi_deg))
