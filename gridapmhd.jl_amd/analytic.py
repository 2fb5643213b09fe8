"""Closed-form fully developed duct solutions and pressure-drop correlations.

* ``analytical_hunt_u`` / ``analytical_hunt_j``: Hunt's flow (perfectly
  conducting Hartmann walls, insulating side walls) as the truncated Fourier
  series of src/Applications/hunt.jl:372-457.
* ``analytical_general_hunt_u``: the same duct with thin conducting Hartmann
  walls of conductance ratio d_b (d_b = 0: Shercliff's insulating duct),
  src/Applications/FullyDevelopedFlow.jl:284-316.
* ``kp_*``: the pressure-gradient correlations of src/utils.jl:69-95.

All functions are vectorised over points x of shape (N, 3); the velocity
functions also return the exact gradient of the truncated series (the
reference differentiates these functions with automatic differentiation when
it forms grad(u - u_h) for the H1 error).  Gradients follow Gridap's layout
G[n, i, k] = d u_k / d x_i.
"""
from __future__ import annotations

import math

import numpy as np


def _modes(l, Ha, k):
    a = (k + 0.5) * math.pi / l
    N = math.sqrt(Ha * Ha + 4.0 * a * a)
    return a, N, 0.5 * (Ha + N), 0.5 * (-Ha + N)


def analytical_hunt_u(a, b, mu, grad_pz, Ha, n, x, grad=False):
    """u = (0, 0, u_z(x, y)) of Hunt's flow with n + 1 Fourier modes."""
    x = np.asarray(x, dtype=np.float64)
    l = b / a
    xi = x[:, 0] / a
    eta = x[:, 1] / a
    inside = (np.abs(xi) <= 1.0) & (np.abs(eta) <= 1.0)
    V = np.zeros(x.shape[0])
    dVx = np.zeros(x.shape[0])
    dVy = np.zeros(x.shape[0])
    for k in range(n + 1):
        ak, N, r1, r2 = _modes(l, Ha, k)
        e1m, e1p = np.exp(-r1 * (1 - eta)), np.exp(-r1 * (1 + eta))
        e2m, e2p = np.exp(-r2 * (1 - eta)), np.exp(-r2 * (1 + eta))
        d1, d2 = 1 + math.exp(-2 * r1), 1 + math.exp(-2 * r2)
        V2 = (r2 / N) * (e1m + e1p) / d1
        V3 = (r1 / N) * (e2m + e2p) / d2
        dV2 = (r2 / N) * r1 * (e1m - e1p) / d1
        dV3 = (r1 / N) * r2 * (e2m - e2p) / d2
        c = 2 * (-1) ** k / (l * ak ** 3)
        V += c * np.cos(ak * xi) * (1 - V2 - V3)
        if grad:
            dVx += -c * ak * np.sin(ak * xi) * (1 - V2 - V3)
            dVy += -c * np.cos(ak * xi) * (dV2 + dV3)
    s = (-grad_pz) * a * a / mu
    u = np.zeros((x.shape[0], 3))
    u[:, 2] = np.where(inside, V * s, 0.0)
    if not grad:
        return u
    G = np.zeros((x.shape[0], 3, 3))
    G[:, 0, 2] = np.where(inside, dVx * s / a, 0.0)
    G[:, 1, 2] = np.where(inside, dVy * s / a, 0.0)
    return u, G


def analytical_hunt_j(a, b, sigma, mu, grad_pz, Ha, n, x):
    """j = (j_x, j_y, 0) of Hunt's flow with n + 1 Fourier modes."""
    x = np.asarray(x, dtype=np.float64)
    l = b / a
    xi = x[:, 0] / a
    eta = x[:, 1] / a
    inside = (np.abs(xi) <= 1.0) & (np.abs(eta) <= 1.0)
    Hdx = np.zeros(x.shape[0])
    Hdy = np.zeros(x.shape[0])
    for k in range(n + 1):
        ak, N, r1, r2 = _modes(l, Ha, k)
        e1m, e1p = np.exp(-r1 * (1 - eta)), np.exp(-r1 * (1 + eta))
        e2m, e2p = np.exp(-r2 * (1 - eta)), np.exp(-r2 * (1 + eta))
        d1, d2 = 1 + math.exp(-2 * r1), 1 + math.exp(-2 * r2)
        H2 = (r2 / N) * (e1m - e1p) / d1
        H2y = (r2 / N) * (e1m * (r1 / a) + e1p * (r1 / a)) / d1
        H3 = (r1 / N) * (e2m - e2p) / d2
        H3y = (r1 / N) * (e2m * (r2 / a) + e2p * (r2 / a)) / d2
        Hdx += -2 * (-1) ** k * np.sin(ak * xi) / (a * l * ak ** 2) * (H2 - H3)
        Hdy += 2 * (-1) ** k * np.cos(ak * xi) / (l * ak ** 3) * (H2y - H3y)
    s = a * a * math.sqrt(sigma) / math.sqrt(mu) * (-grad_pz)
    j = np.zeros((x.shape[0], 3))
    j[:, 0] = np.where(inside, s * Hdy, 0.0)
    j[:, 1] = np.where(inside, -s * Hdx, 0.0)
    return j


def analytical_general_hunt_u(l, d_b, grad_pz, Ha, n, x, grad=False):
    """u = (0, 0, u_z) of the duct with thin conducting Hartmann walls
    (conductance ratio d_b) and insulating side walls, in the Hartmann
    scaling of the FullyDeveloped application (x normalised by the
    Hartmann-wall half width; side walls at x = +-l)."""
    x = np.asarray(x, dtype=np.float64)
    X, Y = x[:, 0], x[:, 1]
    V = np.zeros(x.shape[0])
    dVx = np.zeros(x.shape[0])
    dVy = np.zeros(x.shape[0])
    for k in range(n + 1):
        ak, N, r1, r2 = _modes(l, Ha, k)
        ep1, em1 = 1 + math.exp(-2 * r1), 1 - math.exp(-2 * r1)
        ep2, em2 = 1 + math.exp(-2 * r2), 1 - math.exp(-2 * r2)
        ep = 1 + math.exp(-2 * (r1 + r2))
        ex1 = 0.5 * (np.exp(-r1 * (1 - Y)) + np.exp(-r1 * (1 + Y)))
        ex2 = 0.5 * (np.exp(-r2 * (1 - Y)) + np.exp(-r2 * (1 + Y)))
        dex1 = 0.5 * r1 * (np.exp(-r1 * (1 - Y)) - np.exp(-r1 * (1 + Y)))
        dex2 = 0.5 * r2 * (np.exp(-r2 * (1 - Y)) - np.exp(-r2 * (1 + Y)))
        c2 = (d_b * r2 + em2 / ep2) / (0.5 * N * d_b * ep1 + ep / ep2)
        c3 = (d_b * r1 + em1 / ep1) / (0.5 * N * d_b * ep2 + ep / ep1)
        c = 2 * (-1) ** k / (l * ak ** 3)
        V += c * np.cos(ak * X) * (1 - c2 * ex1 - c3 * ex2)
        if grad:
            dVx += -c * ak * np.sin(ak * X) * (1 - c2 * ex1 - c3 * ex2)
            dVy += -c * np.cos(ak * X) * (c2 * dex1 + c3 * dex2)
    s = Ha * Ha * (-grad_pz)
    u = np.zeros((x.shape[0], 3))
    u[:, 2] = V * s
    if not grad:
        return u
    G = np.zeros((x.shape[0], 3, 3))
    G[:, 0, 2] = dVx * s
    G[:, 1, 2] = dVy * s
    return u, G


def kp_shercliff_cartesian(b, Ha):
    return 1.0 / (Ha * (1.0 - 0.852 * Ha ** -0.5 / b - 1.0 / Ha))


def kp_shercliff_cylinder(Ha):
    return (3.0 / 8.0) * math.pi / (Ha - 1.5 * math.pi)


def kp_hunt(b, Ha):
    return 1.0 / (Ha * (1.0 - 0.956 * Ha ** -0.5 / b - 1.0 / Ha))


def kp_tillac(b, Ha, cw_s, cw_Ha):
    k_s = (1.0 / (3.0 * b)) * (Ha ** 0.5 / (1.0 + cw_s * Ha ** 0.5))
    k_Ha = (1.0 + cw_Ha) / (1.0 / Ha + cw_Ha)
    return 1.0 / (k_s + k_Ha)


def kp_glukhih(Ha, cw):
    return (3.0 / 8.0) * math.pi * (1.0 + 0.833 * cw * Ha - 0.019 * (cw * Ha) ** 2) / Ha
