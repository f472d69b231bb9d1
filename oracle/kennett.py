"""CPU oracle of the Kennett layer reflectivity and the ocean-bottom dressing
-- TEST INFRASTRUCTURE ONLY.

Restates, in closed form, the modified (sqrt(eta rho)-normalised) interface
coefficients of cubic_scattering/kennett_layers.py -- solid-solid 290-386
(Aki & Richards' expressions E, F, G, H, Det and the Q..V combinations),
fluid-solid 389-474, SH 477-506 -- the upward recursion 549-736 and the
ocean-bottom composition of cubic_scattering/ocean_bottom.py:116-156 and
159-368 (Kennett background of [water | sediment | half-space], sub-ocean
reflection of [sediment | half-space], sediment two-way phase, the water
step R = Rd + Tu M (1 - Ru M)^{-1} Td, water phase, optional free surface,
Ricker spectrum seismic_survey.py:217-234, Hermitian inverse FFT with
damping compensation).  Layers are tuples (alpha, beta, rho, thickness) with
beta = 0 for a fluid.

Pins (tests/test_kennett.py): unitarity and reciprocity of the coefficients,
identical media give no reflection, the normal-incidence impedance formula,
the Airy single-layer formula, and agreement with the package's independent
boundary-condition derivation.
"""

from __future__ import annotations

import numpy as np


def slowness(v, Q=np.inf):
    if np.isinf(Q):
        return complex(1.0 / v)
    q = 2.0 * Q
    den = (1.0 + q * q) * v
    return complex(q * q / den, q / den)


def vslow(s, p):
    e = np.sqrt(complex((s + p) * (s - p)))
    if e.imag < 0.0 or (e.imag == 0.0 and e.real < 0.0):
        e = -e
    return e


def solid_solid(p, eta1, neta1, rho1, beta1, eta2, neta2, rho2, beta2):
    """(Rd, Ru, Td, Tu) 2x2 each (kennett_layers.py:290-386)."""
    za1, za2 = np.sqrt(eta1 * rho1), np.sqrt(eta2 * rho2)
    zb1, zb2 = np.sqrt(neta1 * rho1), np.sqrt(neta2 * rho2)
    p2 = complex(p * p)
    d = 2.0 * (rho2 * beta2 ** 2 - rho1 * beta1 ** 2)
    a = (rho2 - rho1) - p2 * d
    b = rho2 - p2 * d
    c = rho1 + p2 * d
    E = b * eta1 + c * eta2
    F = b * neta1 + c * neta2
    G = a - d * eta1 * neta2
    H = a - d * eta2 * neta1
    D = E * F + G * H * p2
    E, F, G, H = E / D, F / D, G / D, H / D
    q = (b * eta1 - c * eta2) * F
    r = (a + d * eta1 * neta2) * H * p2
    s = (a * b + c * d * eta2 * neta2) * p / D
    t = (b * neta1 - c * neta2) * E
    u = (a + d * eta2 * neta1) * G * p2
    v = (a * c + b * d * eta1 * neta1) * p / D
    k = -2j
    Rd = np.array([[q - r, k * np.sqrt(eta1) * np.sqrt(neta1) * s],
                   [k * np.sqrt(eta1) * np.sqrt(neta1) * s, t - u]])
    Td = np.array([[2 * za1 * za2 * F, k * zb1 * za2 * G * p],
                   [k * za1 * zb2 * H * p, 2 * zb1 * zb2 * E]])
    Ru = np.array([[-(q + u), k * np.sqrt(eta2) * np.sqrt(neta2) * v],
                   [k * np.sqrt(eta2) * np.sqrt(neta2) * v, -(t + r)]])
    return Rd, Ru, Td, Td.T.copy()


def fluid_solid(p, eta1, rho1, eta2, neta2, rho2, beta2):
    """(Rd, Ru, Td, Tu) of a fluid over a solid (kennett_layers.py:389-474)."""
    za1, za2, zb2 = np.sqrt(rho1 * eta1), np.sqrt(rho2 * eta2), np.sqrt(rho2 * neta2)
    p2 = complex(p * p)
    d = 2.0 * rho2 * beta2 ** 2
    a = (rho2 - rho1) - p2 * d
    b = rho2 - p2 * d
    c = rho1 + p2 * d
    E = b * eta1 + c * eta2
    F = b
    G = a - d * eta1 * neta2
    H = -d * eta2
    D = E * F + G * H * p2
    En, Fn, Gn, Hn = E / D, F / D, G / D, H / D
    t1 = (b * eta1 - c * eta2) * Fn
    t2 = (a + d * eta1 * neta2) * Hn * p2
    t4 = b * En
    t5 = d * eta2 * Gn * p2
    t6 = b * d * eta1 * p / D
    z = 0j
    pp_t = 2.0 * za1 * za2 * Fn
    ps_t = -2j * za1 * zb2 * Hn * p
    up_ps = -2j * np.sqrt(eta2) * np.sqrt(neta2) * t6
    Rd = np.array([[t1 - t2, z], [z, z]])
    Td = np.array([[pp_t, z], [ps_t, z]])
    Tu = np.array([[pp_t, ps_t], [z, z]])
    Ru = np.array([[-(t1 + t5), up_ps], [up_ps, -(t2 + t4)]])
    return Rd, Ru, Td, Tu


def sh(neta1, rho1, beta1, neta2, rho2, beta2):
    z1 = rho1 * beta1 ** 2 * neta1
    z2 = rho2 * beta2 ** 2 * neta2
    rd = (z1 - z2) / (z1 + z2)
    td = 2.0 * np.sqrt(z1 * z2) / (z1 + z2)
    return rd, -rd, td, td


def _media(layers, p):
    out = []
    for al, be, rho, h in layers:
        sp = slowness(al)
        eta = vslow(sp, p)
        if be == 0.0:
            out.append((eta, 0j, 0j))
        else:
            ss = slowness(be)
            out.append((eta, vslow(ss, p), 1.0 / ss))
    return out


def reflectivity(layers, p, omega):
    """(RD_psv (nf,2,2), RD_sh (nf,)) of the stack (kennett_layers.py:549-736)."""
    omega = np.atleast_1d(np.asarray(omega))
    med = _media(layers, p)
    nf = omega.shape[0]
    R = np.zeros((nf, 2, 2), complex)
    Rs = np.zeros(nf, complex)
    for i in range(len(layers) - 2, -1, -1):
        a1, b1, r1, _ = layers[i]
        a2, b2, r2, h2 = layers[i + 1]
        e1, n1, c1 = med[i]
        e2, n2, c2 = med[i + 1]
        if b1 == 0.0:
            Rd, Ru, Td, Tu = fluid_solid(p, e1, r1, e2, n2, r2, c2)
        else:
            Rd, Ru, Td, Tu = solid_solid(p, e1, n1, r1, c1, e2, n2, r2, c2)
        if np.isinf(h2):
            ea = eb = np.ones(nf, complex)
        else:
            ea = np.exp(1j * omega * e2 * h2)
            eb = np.exp(1j * omega * n2 * h2)
        for w in range(nf):
            Ew = np.diag([ea[w], eb[w]])
            M = Ew @ R[w] @ Ew
            R[w] = Rd + Tu @ M @ np.linalg.inv(np.eye(2) - Ru @ M) @ Td
            if b1 != 0.0:
                rd, ru, td, tu = sh(n1, r1, c1, n2, r2, c2)
                m = eb[w] ** 2 * Rs[w]
                Rs[w] = rd + tu * m * td / (1.0 - ru * m)
    return R, Rs


def ricker(omega, f_peak):
    x = (omega / (2.0 * np.pi * f_peak)) ** 2
    return x * np.exp(-x)


def ocean_bottom(water, sed, hs, H, p, omega_real, gamma, active, R_slab, free_surface, nw, T,
                 f_peak):
    """R_bg, R_total, trace_total, trace_homogeneous (ocean_bottom.py:234-368).

    water = (alpha, rho, depth); sed = (alpha, beta, rho); hs = (alpha, beta, rho);
    R_slab has one entry per frequency (zero outside the active band)."""
    wd = omega_real + 1j * gamma
    wa = wd[active]
    full = [(water[0], 0.0, water[1], water[2]), (sed[0], sed[1], sed[2], H),
            (hs[0], hs[1], hs[2], np.inf)]
    R_bg = np.zeros(omega_real.shape, complex)
    R_bg[active] = reflectivity(full, p, wa)[0][:, 0, 0]
    sub = [(sed[0], sed[1], sed[2], H), (hs[0], hs[1], hs[2], np.inf)]
    r_sh = reflectivity(sub, p, wa)[0][0, 0, 0]
    eta_w = vslow(slowness(water[0]), p)
    eta_s = vslow(slowness(sed[0]), p)
    neta_s = vslow(slowness(sed[1]), p)
    MT = np.exp(2j * wd * eta_s * H) * r_sh + R_slab
    Rd, Ru, Td, Tu = fluid_solid(p, eta_w, water[1], eta_s, neta_s, sed[2], 1.0 / slowness(sed[1]))
    R_tot = np.zeros(omega_real.shape, complex)
    m = MT[active]
    R_tot[active] = Rd[0, 0] + Tu[0, 0] * m * Td[0, 0] / (1.0 - Ru[0, 0] * m)
    ph = np.exp(2j * wd * eta_w * water[2])

    def water_step(R):
        x = ph * R
        return x / (1.0 + x) if free_surface else x

    nt = 2 * nw
    t = np.arange(nt) * (T / nt)
    S = ricker(omega_real, f_peak)

    def trace(R):
        U = S * R
        full_spec = np.zeros(nt, complex)
        full_spec[1:nw] = U
        full_spec[nw + 1:] = np.conj(U[::-1])
        return np.real(np.fft.fft(full_spec)) * np.exp(-gamma * t)

    return R_bg, R_tot, trace(water_step(R_tot)), trace(water_step(R_bg)), t
