"""Comparison helpers shared by the parity tests."""
import numpy as np


def read_set_mask(nx, ny, nz, ng):
    """Cells of the ghosted primitive array that the update reads: the interior
    plus, on each active axis, the ghost layers of rows whose transverse indices
    are interior (corners and inactive-axis ghosts are never read)."""
    shape = (nz + 2 * ng + 1, ny + 2 * ng + 1, nx + 2 * ng + 1)
    m = np.zeros(shape, bool)
    I = slice(ng + 1, ng + 1 + nx)
    J = slice(ng + 1, ng + 1 + ny)
    K = slice(ng + 1, ng + 1 + nz)
    m[K, J, :] = True  # x rows incl. ghosts (x always has ghost layers)
    if ny > 1:
        m[K, :, I] = True
    if nz > 1:
        m[:, J, I] = True
    return m


def rel_errors(a, b):
    """Per-variable relative L1 and Linf of a against b (last axis = variable):
    L1 = sum|a-b| / sum|b|,  Linf = max|a-b| / max|b|."""
    a = np.asarray(a, float).reshape(-1, a.shape[-1])
    b = np.asarray(b, float).reshape(-1, b.shape[-1])
    d = np.abs(a - b)
    s1 = np.abs(b).sum(axis=0)
    sm = np.abs(b).max(axis=0)
    l1 = np.where(s1 > 0, d.sum(axis=0) / np.where(s1 > 0, s1, 1), d.sum(axis=0))
    li = np.where(sm > 0, d.max(axis=0) / np.where(sm > 0, sm, 1), d.max(axis=0))
    return l1, li


def seeded_noise(p, nx, ny, nz, ng, kind, amp, seed):
    """Multiplicative noise on the interior (kind 1 pressure, 2 density with
    p scaled by f^Gamma is left to the caller); numpy generator, for feeding the
    same perturbed state to the oracle and the device."""
    rng = np.random.default_rng(seed)
    q = p.copy()
    sl = (slice(ng + 1, ng + 1 + nz), slice(ng + 1, ng + 1 + ny), slice(ng + 1, ng + 1 + nx))
    rho = q[sl + (0,)]
    mask = rho > 0.05 * rho.max()
    f = 1.0 + amp * (2.0 * rng.random(rho.shape) - 1.0)
    f = np.where(mask, f, 1.0)
    if kind == 1:
        q[sl + (1,)] *= f
    else:
        q[sl + (0,)] *= f
        q[sl + (1,)] *= f ** 1.3333333333333333
    return q


def michel_norms(p, exact, nx, ny, ng):
    """Density error norms of the Michel problem against the analytic
    solution, as calc_convergence_norms computes them (output.f90:1189-1213,
    1223-1224): cells i0 = max(1, int(0.10 nx)) .. i1 = min(nx, max(i0,
    int(0.95 nx))) of row j = max(1, ny // 2), k = 1; returns (L1, L2, Linf)
    with L1 and L2 divided by the cell count."""
    i0 = max(1, int(0.10 * nx))
    i1 = min(nx, max(i0, int(0.95 * nx)))
    j = max(1, ny // 2)
    rows = slice(ng + i0, ng + i1 + 1)
    d = np.abs(p[ng + 1, ng + j, rows, 0] - exact[ng + 1, ng + j, rows, 0])
    n = i1 - i0 + 1
    return np.array([d.sum() / n, np.sqrt((d ** 2).sum() / n), d.max()])


def dust_exact_density(r, bh_mass, accretion_constant):
    """Density of the stationary dust solution (evaluate_dust_accretion_state,
    conditions.f90:249-278): rho = -C / (r^2 sqrt(2 M / r))."""
    return -accretion_constant / (r ** 2 * np.sqrt(2.0 * bh_mass / r))


def dust_norms(p, rho_exact, nx, ny, ng):
    """The dust branch of calc_convergence_norms (output.f90:1214-1235): cells
    max(1, int(0.10 nx)) .. min(nx, max(i0, int(0.95 nx))) of row
    j = max(1, ny // 2), k = 1, against rho_exact[i - 1]; (L1, L2, Linf)."""
    i0 = max(1, int(0.10 * nx))
    i1 = min(nx, max(i0, int(0.95 * nx)))
    j = max(1, ny // 2)
    d = np.abs(p[ng + 1, ng + j, ng + i0:ng + i1 + 1, 0] - rho_exact[i0 - 1:i1])
    n = i1 - i0 + 1
    return np.array([d.sum() / n, np.sqrt((d ** 2).sum() / n), d.max()])
