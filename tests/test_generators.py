"""Host input generators of the coupled path (no GPU): Rosenthal frames and the
Voronoi substrate, against direct numpy evaluations."""
import numpy as np
import pytest

import remelt_cases as rc


def test_rosenthal_frames_match_formula(amca):
    from amca3d_amd import remelt
    grid = dict(nx=30, ny=12, nz=10, h=2.5e-6, x0=0.0, y0=0.0, z0=0.0)
    fe = rc.fe_mesh(grid, 4)
    times = 1e-5 * np.arange(5)
    kw = dict(P=200.0, eta=0.35, v=1.0, k=20.0, alpha=5e-6, T0=300.0, x_start=1e-5,
              y_c=15e-6, z_top=25e-6, r_min=2.5e-6)
    got = remelt.rosenthal_frames(fe["xyz"], times, **kw)
    ref = np.stack([rc.rosenthal(fe["xyz"], t, P=kw["P"], eta=kw["eta"], v=kw["v"], k=kw["k"],
                                 alpha=kw["alpha"], T0=kw["T0"], x_start=kw["x_start"],
                                 y_c=kw["y_c"], z_top=kw["z_top"], r_min=kw["r_min"])
                    for t in times])
    np.testing.assert_allclose(got, ref, rtol=1e-12)
    assert got.max() > 1600.0  # the pool is hotter than the melting point


@pytest.mark.parametrize("n_seeds", [1, 7, 300])
def test_voronoi_labels_are_nearest_seed(amca, n_seeds):
    from amca3d_amd import remelt
    grid = dict(nx=23, ny=17, nz=9, h=1e-6, x0=1e-6, y0=-2e-6, z0=0.0)
    lab, eu = remelt.voronoi_substrate(grid, n_seeds, 12345)
    nx, ny, nz, h = grid["nx"], grid["ny"], grid["nz"], grid["h"]
    nxG, nyG = nx + 2, ny + 2
    L = lab.reshape(nz + 2, nyG, nxG)
    assert (L[0] == -1).all() and (L[-1] == -1).all() and (L[:, 0] == -1).all()
    assert (L[:, :, 0] == -1).all() and (L[:, :, -1] == -1).all()
    inner = L[1:-1, 1:-1, 1:-1]
    assert inner.min() >= 0 and inner.max() < n_seeds
    # the seed points, re-drawn from the same splitmix64 stream: every cell's
    # label is its nearest seed by brute force
    seeds = _splitmix_seeds(grid, n_seeds, 12345)
    k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1),
                          indexing="ij")
    cen = np.stack([grid["x0"] + h * (i - 0.5), grid["y0"] + h * (j - 0.5),
                    grid["z0"] + h * (k - 0.5)], axis=-1).reshape(-1, 3)
    d = ((cen[:, None, :] - seeds[None, :, :]) ** 2).sum(-1)
    np.testing.assert_array_equal(inner.ravel(), d.argmin(1))
    # Euler angles in range
    assert (eu[:, 0] >= 0).all() and (eu[:, 0] < 2 * np.pi).all()
    assert (eu[:, 1] >= 0).all() and (eu[:, 1] <= np.pi).all()
    assert (eu[:, 2] >= 0).all() and (eu[:, 2] < 0.5 * np.pi).all()


def _splitmix_seeds(grid, n, seed):
    """The seed points of the generator: splitmix64 doubles in [0, 1), three per
    seed, scaled to the box."""
    M = (1 << 64) - 1
    s = seed
    out = []
    for _ in range(3 * n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out.append((z >> 11) / 9007199254740992.0)
    u = np.array(out).reshape(n, 3)
    lo = np.array([grid["x0"], grid["y0"], grid["z0"]])
    L = np.array([grid["nx"], grid["ny"], grid["nz"]]) * grid["h"]
    return lo + u * L
