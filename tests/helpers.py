"""Independent host-side checks used by the parity tests.

These restate the quantities of the reference's own test helpers with a
different formulation (direct trigonometric evaluation, atan2, brute force over
all pairs) so that they can catch mistakes of the oracle and of the device
path.
"""
import math

import numpy as np


def contour_direct(lab, R0, K, psi):
    """R_j(ψ) = R0 [a0 + Σ αn cos nψ + βn sin nψ] by direct cos/sin (no
    recurrence), src/geometry.jl:124-156 restated via test_contact.jl:149."""
    psi = np.asarray(psi, dtype=np.float64)
    r = np.full_like(psi, lab[0])
    for m in range(1, K + 1):
        n = m + 1
        r = r + lab[m] * np.cos(n * psi) + lab[K + m] * np.sin(n * psi)
    return R0 * r


def rdirect_body(coef, theta, R0, phi):
    """Body contour from body-frame coefficients, test/test_geometry.jl:35."""
    K = (coef.shape[0] - 1) // 2
    phi = np.asarray(phi, dtype=np.float64)
    r = np.full_like(phi, coef[0])
    for m in range(1, K + 1):
        n = m + 1
        r = r + coef[m] * np.cos(n * (phi - theta)) + coef[K + m] * np.sin(n * (phi - theta))
    return R0 * r


def min_image(d, L):
    return d - L * np.round(d / L)


def brute_force_pressure(b, R, lab, box, terms_fn, nth):
    """brute_force_pressure (test/test_contact.jl:4-22): all pairs, atan2."""
    N = len(b.x)
    K = (lab.shape[0] - 1) // 2
    P = np.zeros((nth, N))
    phi = 2 * math.pi * np.arange(nth) / nth
    for i in range(N):
        ux = b.x[i] + R[:, i] * np.cos(phi)
        uy = b.y[i] + R[:, i] * np.sin(phi)
        for j in range(N):
            if j == i:
                continue
            dx = min_image(ux - b.x[j], box.Lx)
            dy = min_image(uy - b.y[j], box.Ly)
            d = contour_direct(lab[:, j], b.R0[j], K, np.arctan2(dy, dx)) - np.hypot(dx, dy)
            P[:, i] += terms_fn(d, b.type[i], b.type[j], b.parent[i] == b.parent[j])
    return P


def field_energy(b, R, lab, nth, terms_fn, nr=4000):
    """field_energy (test/test_contact.jl:41-57): radial midpoint quadrature of
    body 1 (index 0) in the field of body 2 (index 1)."""
    K = (lab.shape[0] - 1) // 2
    E = 0.0
    t = (np.arange(1, nr + 1) - 0.5) / nr
    for k in range(nth):
        phi = 2 * math.pi * k / nth
        s = R[k, 0] * t
        dx = b.x[0] + s * math.cos(phi) - b.x[1]
        dy = b.y[0] + s * math.sin(phi) - b.y[1]
        d = contour_direct(lab[:, 1], b.R0[1], K, np.arctan2(dy, dx)) - np.hypot(dx, dy)
        E += np.sum(terms_fn(d, b.type[0], b.type[1], False) * s * R[k, 0] / nr)
    return E * 2 * math.pi / nth


def read_reference(path):
    """read_reference (test/test_dynamics.jl:5-20): initial rows and the
    snapshots after T+1 steps."""
    init, snaps, current = [], {}, -1
    with open(path) as f:
        for line in f:
            if line.startswith("#init"):
                init.append([float(v) for v in line.split()[1:]])
            elif line.startswith("#T="):
                current = int(line[3:].split()[0]) + 1
                snaps[current] = []
            elif line.startswith("#cell"):
                snaps[current].append([float(v) for v in line.split()[1:]])
    return np.array(init).T, {k: np.array(v).T for k, v in snaps.items()}


def rel_err(a, ref, scale):
    return np.abs(a - ref) / np.maximum(np.maximum(np.abs(ref), scale), 1e-300)


def force_scales(sim, g):
    """Σk |Pk| Δφ |nk| and Σk |Pk| Δφ |Rk R'k| per body from the oracle."""
    geo = sim.geom
    dphi = 2 * math.pi / sim.ntheta
    phi = dphi * np.arange(sim.ntheta)[:, None]
    R, dR = geo.R, geo.dR
    nx = R * np.cos(phi) + dR * np.sin(phi)
    ny = R * np.sin(phi) - dR * np.cos(phi)
    w = np.abs(g.P) * dphi
    return np.sum(w * np.hypot(nx, ny), axis=0), np.sum(w * np.abs(R * dR), axis=0)


def plain_force_errors(dev, ref, ok):
    """The unnormalised force metrics beside the scale-normalised one, on the
    bodies without a flipped node: per-body |ΔF| / |F| (median, 99th
    percentile, maximum) and the field error ‖ΔF‖ / ‖F‖ over all of them."""
    dF = np.hypot(dev.gx - ref.gx, dev.gy - ref.gy)[ok]
    Fn = np.hypot(ref.gx, ref.gy)[ok]
    per = dF / np.maximum(Fn, 1e-300)
    return np.median(per), np.quantile(per, 0.99), per.max(), np.linalg.norm(dF) / np.linalg.norm(Fn)
