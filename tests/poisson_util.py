"""The pressure operator of the reference restated in numpy, for the converged-solve tests.

PressureSolver<cSaturnModel>::run (planet/PressureSolver.h:350-353, 476-491, no obstacle): at an
interior cell (i 1..im-2, j 1..jm-2, k 0..km-2, the phi neighbours round the ring)
    Lc p = num1 (p[i+1] + p[i-1]) + num2 (p[j+1] + p[j-1]) + num3 (p[k+1] + p[k-1]) - den p,
num1 = 1/dr^2, num2 = 1/(rm^2 dtheta^2), num3 = 1/(rm^2 sin^2 dphi^2), den = 2 (num1 + num2 + num3);
the boundary pass :718-740: p0 = 4/3 p1 - 1/3 p2 on the radial planes, then on the polar rows,
then the seam copy k = km-1 <- 0.
"""
import math

import numpy as np


def geometry(g, c):
    dr = g.dr
    dthe = g.dthe_deg / (180.0 / math.pi)
    dphi = g.dphi_deg / (180.0 / math.pi)
    radz = np.empty(g.im)
    radz[0] = 1.0
    for l in range(1, g.im):                       # lib/Array_1D.cpp:54-59, repeated addition
        radz[l] = radz[l - 1] + dr
    thez = np.empty(g.jm)
    thez[0] = 0.0
    for l in range(1, g.jm):
        thez[l] = thez[l - 1] + dthe
    rm = radz + (c.metric_radius_km / c.L_atm - 1.0) if c.metric_radius_km > 0 else radz
    sp = np.sin(thez)
    sp = np.where(sp < 0.0, 0.0, sp)
    sp = np.where(sp == 0.0, 1.0e-5, sp)
    inv_rm2 = 1.0 / (rm * rm)
    num1 = 1.0 / (dr * dr)
    num2 = inv_rm2 * (1.0 / (dthe * dthe))                      # [im]
    num3 = (inv_rm2[:, None] / (sp * sp)[None, :]) * (1.0 / (dphi * dphi))   # [im, jm]
    return num1, num2, num3


def boundary_pass(p):
    """The solver's boundary pass, in place; p is [..., im, jm, km]."""
    c43, c13 = 4.0 / 3.0, 1.0 / 3.0
    p[..., 0, :, :] = c43 * p[..., 1, :, :] - c13 * p[..., 2, :, :]
    p[..., -1, :, :] = c43 * p[..., -2, :, :] - c13 * p[..., -3, :, :]
    p[..., :, 0, :] = c43 * p[..., :, 1, :] - c13 * p[..., :, 2, :]
    p[..., :, -1, :] = c43 * p[..., :, -2, :] - c13 * p[..., :, -3, :]
    p[..., :, :, -1] = p[..., :, :, 0]
    return p


def apply(p, g, c):
    """Lc p on the interior cells (zero elsewhere)."""
    num1, num2, num3 = geometry(g, c)
    out = np.zeros_like(p)
    n = g.km - 1
    P = p[..., :, :, :n]
    pe = np.roll(P, -1, axis=-1)
    pw = np.roll(P, 1, axis=-1)
    I = slice(1, g.im - 1)
    J = slice(1, g.jm - 1)
    n2 = num2[I][:, None, None]
    n3 = num3[I, J][:, :, None]
    den = 2.0 * num1 + 2.0 * n2 + 2.0 * n3
    acc = (P[..., 2:, J, :] + P[..., :-2, J, :]) * num1 + (P[..., I, 2:, :] + P[..., I, :-2, :]) * n2 \
        + (pe[..., I, J, :] + pw[..., I, J, :]) * n3
    out[..., I, J, :n] = acc - den * P[..., I, J, :]
    return out


def weights(g):
    y_i = np.ones(g.im - 2)
    y_i[[0, -1]] = 1.5
    y_j = np.ones(g.jm - 2)
    y_j[[0, -1]] = 1.5
    return y_i[:, None, None] * y_j[None, :, None] * np.ones(g.km - 1)[None, None, :]


def interior(a, g):
    return a[..., 1:g.im - 1, 1:g.jm - 1, :g.km - 1]


def project(src, g):
    """src minus its component along the operator's left null vector (interior cells)."""
    w = weights(g)
    s = src.copy()
    I = interior(s, g)
    c = np.sum(w * I, axis=(-3, -2, -1), keepdims=True) / np.sum(w)
    I -= c
    return s


def rel_residual(p, src, g, c):
    r = interior(src, g) - interior(apply(p, g, c), g)
    s = interior(src, g)
    ax = (-3, -2, -1)
    return np.sqrt(np.sum(r * r, axis=ax)) / np.sqrt(np.sum(s * s, axis=ax))


def wmean(p, g):
    w = weights(g)
    return np.sum(w * interior(p, g), axis=(-3, -2, -1)) / np.sum(w)
