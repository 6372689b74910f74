"""Pins the oracle's libgpw folded-table path (orc_fold_tables,
orc_collocate_folded, orc_integrate_folded; reference collocate.cc:32-57,
276-331, 626-745, polynomials.hpp:5-86, utils.hpp:20-181).

  * fold: the folded table of a Gaussian x polynomial sampled on a window
    wider than the period equals the periodic image sum evaluated directly
    at the image coordinates (a wrong centring, sign or period fails it);
  * fold + collocate over the full period equal a brute-force 3D image sum
    of the separable Gaussian product sum_{a,g,b} co X^a Z^g Y^b exp(...)
    evaluated point by point, image by image;
  * the mask is the support of the folded table, all ones when the window
    covers the period;
  * a local window is the matching slice of the full-period result, and
    points whose mask is clear are left untouched;
  * integrate is the adjoint of collocate on the same masked window.
"""
import numpy as np
import pytest

from helpers import rel_err

H = 0.3  # grid spacing (bohr)


def tables(L, zeta, xc, center, half, n_extra=0):
    """unfolded tables (L, 2*half+1+n_extra) of (x - xc)^a exp(-zeta (x - xc)^2)
    on the grid coordinates center - half + i (times H)"""
    length = 2 * half + 1 + n_extra
    x = (center - (length - 1) // 2 + np.arange(length)) * H - xc
    e = np.exp(-zeta * x * x)
    return np.stack([x ** a * e for a in range(L)]), length


def image_sum(L, zeta, xc, period, lo, hi):
    """F[a][g] = sum over images n of (X - xc)^a exp(-zeta (X - xc)^2),
    X = (g + n period) H, for all image coordinates g + n period in [lo, hi)"""
    out = np.zeros((L, period))
    for g in range(period):
        for n in range(-50, 51):
            c = g + n * period
            if lo <= c < hi:
                x = c * H - xc
                for a in range(L):
                    out[a, g] += x ** a * np.exp(-zeta * x * x)
    return out


@pytest.mark.parametrize("period,half,n_extra", [(12, 30, 0), (16, 5, 1), (9, 40, 3), (20, 9, 0)])
def test_fold_is_periodic_image_sum(oracle, period, half, n_extra):
    L = 4
    rng = np.random.default_rng(period + half)
    unf = []
    meta = []
    for d in range(3):
        center = int(rng.integers(-3, period + 3))
        xc = (center + rng.uniform(0, 1)) * H
        zeta = rng.uniform(0.05, 0.5)
        t, length = tables(L, zeta, xc, center, half, n_extra)
        unf.append(t)
        meta.append((center, xc, zeta, length))
    lengths = [m[3] for m in meta]
    U = np.zeros((3, L, max(lengths)))
    for d in range(3):
        U[d, :, :lengths[d]] = unf[d]
    folded, mask = oracle.fold_tables(U, lengths, [m[0] for m in meta], [period] * 3, ldp=period + 2)
    for d in range(3):
        center, xc, zeta, length = meta[d]
        lo = center - (length - 1) // 2
        ref = image_sum(L, zeta, xc, period, lo, lo + length)
        assert rel_err(folded[d, :, :period], ref) < 1e-14, d
        assert not folded[d, :, period:].any()
        # the mask is the support (Gaussians have no zeros inside the window)
        support = np.abs(ref).sum(axis=0) > 0
        assert np.array_equal(mask[d, :period].astype(bool), support), d
        if length >= period:
            assert mask[d, :period].all()
        assert not mask[d, period:].any()


def brute_density(co, zeta, xc, period, images=3):
    """sum over images of sum_{a,g,b} co[a][g][b] X^a Z^g Y^b exp(-zeta r^2),
    X = x - xc_x + n_x Lx etc., point by point"""
    L = co.shape[0]
    nz, ny, nx = period[0], period[1], period[2]
    out = np.zeros((nz, ny, nx))
    ns = range(-images, images + 1)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                v = 0.0
                for mz in ns:
                    Z = (z + mz * nz) * H - xc[0]
                    for my in ns:
                        Y = (y + my * ny) * H - xc[1]
                        for mx in ns:
                            X = (x + mx * nx) * H - xc[2]
                            e = np.exp(-zeta * (X * X + Y * Y + Z * Z))
                            if e < 1e-300:
                                continue
                            p = 0.0
                            for a in range(L):
                                for g in range(L):
                                    for b in range(L):
                                        p += co[a, g, b] * X ** a * Z ** g * Y ** b
                            v += p * e
                out[z, y, x] = v
    return out


def test_fold_collocate_is_brute_force_image_sum(oracle):
    """Direction order of the folded tables: 0 = z (slowest axis of the
    grid), 1 = y, 2 = x; co[a][g][b] has a the x power, g the z power, b the
    y power (collocate_core, collocate.cc:543-623)."""
    rng = np.random.default_rng(5)
    L, zeta = 3, 1.2
    period = (7, 8, 9)          # (z, y, x)
    half = 14                   # 29-point windows: wider than every period
    centers, xcs, unf = [], [], []
    for d in range(3):
        c = int(rng.integers(0, period[d]))
        xc = (c + rng.uniform(0, 1)) * H
        t, length = tables(L, zeta, xc, c, half)
        centers.append(c); xcs.append(xc); unf.append(t)
    U = np.stack(unf)
    folded, mask = oracle.fold_tables(U, [U.shape[2]] * 3, centers, list(period), ldp=max(period))
    co = rng.uniform(-1, 1, (L, L, L))
    dens = oracle.collocate_folded(co, folded, mask, period, (0, 0, 0))
    ref = brute_density(co, zeta, xcs, period, images=2)
    # exactly: the brute force restricted to the window's coordinates (the
    # tables reach 14 points = 4.2 bohr from the centre on each side)
    ref_w = np.zeros(period)
    lo = [centers[d] - (U.shape[2] - 1) // 2 for d in range(3)]
    for z in range(period[0]):
        for y in range(period[1]):
            for x in range(period[2]):
                v = 0.0
                for cz in range(lo[0], lo[0] + U.shape[2]):
                    if (cz - z) % period[0]:
                        continue
                    Z = cz * H - xcs[0]
                    for cy in range(lo[1], lo[1] + U.shape[2]):
                        if (cy - y) % period[1]:
                            continue
                        Y = cy * H - xcs[1]
                        for cx in range(lo[2], lo[2] + U.shape[2]):
                            if (cx - x) % period[2]:
                                continue
                            X = cx * H - xcs[2]
                            e = np.exp(-zeta * (X * X + Y * Y + Z * Z))
                            p = sum(co[a, g, b] * X ** a * Z ** g * Y ** b
                                    for a in range(L) for g in range(L) for b in range(L))
                            ref_w[z, y, x] += p * e
    assert rel_err(dens, ref_w) < 1e-13
    # and the untruncated image sum up to the cut, exp(-1.2 * 4.2^2) ~ 7e-10
    assert rel_err(dens, ref) < 1e-8


def _random_folded(rng, L, period, ldp):
    P = rng.uniform(-1, 1, (3, L, ldp))
    mask = (rng.uniform(0, 1, (3, ldp)) < 0.75).astype(np.int32)
    return P, mask


def test_folded_window_is_slice_and_masked_points_untouched(oracle):
    rng = np.random.default_rng(9)
    L, period = 4, (10, 11, 12)
    P, mask = _random_folded(rng, L, period, 12)
    co = rng.uniform(-1, 1, (L, L, L))
    full_mask = np.ones_like(mask)
    full = oracle.collocate_folded(co, P, full_mask, period, (0, 0, 0))
    # all masks set: collocate_core on the tables of the period
    core = oracle.collocate_core(period, co, P)
    assert rel_err(full, core) < 1e-14
    lc, gs = (2, 3, 1), (6, 5, 9)
    base = rng.uniform(-1, 1, gs)
    win = oracle.collocate_folded(co, P, mask, gs, lc, density=base.copy())
    sl = full[lc[0]:lc[0] + gs[0], lc[1]:lc[1] + gs[1], lc[2]:lc[2] + gs[2]]
    m3 = (mask[0, lc[0]:lc[0] + gs[0]][:, None, None] * mask[1, lc[1]:lc[1] + gs[1]][None, :, None] *
          mask[2, lc[2]:lc[2] + gs[2]][None, None, :]).astype(bool)
    assert m3.any() and (~m3).any()
    assert np.array_equal(win[~m3], base[~m3])
    assert rel_err(win[m3] - base[m3], sl[m3]) < 1e-14


def test_folded_integrate_is_adjoint(oracle):
    rng = np.random.default_rng(10)
    L = 5
    P, mask = _random_folded(rng, L, (9, 10, 11), 13)
    lc, gs = (1, 2, 0), (8, 7, 11)
    co = rng.uniform(-1, 1, (L, L, L))
    W = rng.uniform(-1, 1, gs)
    lhs = float((oracle.collocate_folded(co, P, mask, gs, lc) * W).sum())
    I = oracle.integrate_folded(P, mask, gs, lc, W)          # [g][b][a]
    rhs = float(np.einsum("agb,gba->", co, I))
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))
    # nothing set: zero block
    assert not oracle.integrate_folded(P, np.zeros_like(mask), gs, lc, W).any()
