"""Shared input builders for the tests (seeded, deterministic)."""
import numpy as np


def bumpy_sdf(shape, h, seed, r=0.2, noise=0.01):
    """Signed distance to a sphere/circle centred in the padded box plus seeded noise."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    d2 = np.zeros(shape)
    for a, (gr, s) in enumerate(zip(grids, shape)):
        d2 += ((gr - (s - 1) / 2.0) * h[a]) ** 2
    return np.sqrt(d2) - r + noise * rng.standard_normal(shape)


def rand_vel(shape, seed, scale=1.0, n=3):
    rng = np.random.default_rng(seed + 1000)
    return [scale * rng.standard_normal(shape) for _ in range(n)]


def rel_linf(a, b, sl=None):
    if sl is not None:
        a, b = a[sl], b[sl]
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def crossing_positions(G, axis, h):
    """Zero-level-set positions on grid edges along `axis`: sign pattern of each
    edge (True where G changes sign) and the linearly interpolated crossing
    offset in units of length.  NaN where there is no crossing."""
    a = np.moveaxis(G, axis, -1)
    g0, g1 = a[..., :-1], a[..., 1:]
    cross = (g0 * g1) < 0
    with np.errstate(divide="ignore", invalid="ignore"):
        off = np.where(cross, h * g0 / (g0 - g1), np.nan)
    return cross, off


def front_parity(got, ref, h, rel_tol=1e-12, interior=None):
    """Zero-level-set parity: along every axis the interpolated crossing
    positions x_i + h G_i/(G_i - G_{i+1}) must agree within rel_tol * L per edge
    (L = h * cells along the axis).  Edges whose sign pattern differs are
    accepted only where the front sits on a node within rounding of zero in
    both fields (|G| <= rel_tol * max|G|).  Returns the worst position error
    relative to L."""
    if interior is not None:
        got, ref = got[interior], ref[interior]
    scale = max(float(np.abs(ref).max()), 1e-300)
    worst = 0.0
    for ax in range(got.ndim):
        hh = h[ax]
        L = hh * got.shape[ax]
        cg, og = crossing_positions(got, ax, hh)
        cr, orr = crossing_positions(ref, ax, hh)
        both = cg & cr
        if both.any():
            worst = max(worst, float(np.abs(og[both] - orr[both]).max()) / L)
        diff = cg ^ cr
        if diff.any():
            a = np.moveaxis(got, ax, -1)
            b = np.moveaxis(ref, ax, -1)
            near = (np.minimum(np.abs(a[..., :-1]), np.abs(a[..., 1:])) <= rel_tol * scale) & \
                   (np.minimum(np.abs(b[..., :-1]), np.abs(b[..., 1:])) <= rel_tol * scale)
            assert near[diff].all(), "crossing pattern differs away from a node on axis %d" % ax
    assert worst <= rel_tol, "front position differs by %.3e L" % worst
    return worst


def front_parity_torch(got, ref, h, rel_tol=1e-12):
    """front_parity on device tensors (interior views), for grids too large to
    copy to the host every step: the same checks along every axis.  Returns
    (worst position error relative to L, count of edges whose sign pattern
    differs only at nodes within rounding of zero)."""
    import torch
    scale = max(float(ref.abs().max()), 1e-300)
    worst, near_only = 0.0, 0
    for ax in range(got.dim()):
        n = got.shape[ax]
        L = h[ax] * n
        a0, a1 = got.narrow(ax, 0, n - 1), got.narrow(ax, 1, n - 1)
        b0, b1 = ref.narrow(ax, 0, n - 1), ref.narrow(ax, 1, n - 1)
        cg, cr = (a0 * a1) < 0, (b0 * b1) < 0
        both = cg & cr
        if bool(both.any()):
            og = h[ax] * a0[both] / (a0[both] - a1[both])
            orr = h[ax] * b0[both] / (b0[both] - b1[both])
            worst = max(worst, float((og - orr).abs().max()) / L)
            del og, orr
        diff = cg ^ cr
        if bool(diff.any()):
            near = (torch.minimum(a0.abs(), a1.abs()) <= rel_tol * scale) & \
                   (torch.minimum(b0.abs(), b1.abs()) <= rel_tol * scale)
            assert bool(near[diff].all()), "crossing pattern differs away from a node on axis %d" % ax
            near_only += int(diff.sum())
        del cg, cr, both, diff
    assert worst <= rel_tol, "front position differs by %.3e L" % worst
    return worst, near_only
