"""Coefficient fields of the benchmark workloads, pinned on the CPU.

The L-shaped waveguide: the product's bump field (which sums only the
lattice cells around each node) against the oracle's brute-force sum over
the whole lattice, and the oracle's own lattice geometry against closed
forms at every lattice centre.  The Fourier field: band limit, mean and
bound from its definition."""
import math

import numpy as np
import torch

from oracle import hps_oracle as O
from slablu_amd.src_disc import built_in_funcs as bif


def test_waveguide_lattice_geometry_closed_form():
    """At a lattice centre the field is amp for a kept bump and 0 for a
    removed one (the next bump is 0.04 away: exp(-(0.04/0.005)^2) = 1.6e-28);
    the removed bumps are the half row j = 12, i >= 12 and the half column
    i = 12, j >= 12 of the 25 x 25 lattice (25 of 625)."""
    n, h = 25, 0.04
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    xx = np.stack(((ii.ravel() + 0.5) * h, (jj.ravel() + 0.5) * h), axis=1)
    v = O.lshaped_waveguide_bumps(xx)
    removed = ((jj.ravel() == 12) & (ii.ravel() >= 12)) | ((ii.ravel() == 12) & (jj.ravel() >= 12))
    assert removed.sum() == 25
    assert np.abs(v[~removed] - 1.0).max() < 1e-15
    assert np.abs(v[removed]).max() < 1e-27
    # the removed half row and half column meet at the lattice centre (12, 12)
    assert removed[12 * n + 12] and removed[24 * n + 12] and removed[12 * n + 24]
    assert not removed[11 * n + 12] and not removed[12 * n + 11]


def test_waveguide_field_matches_brute_force():
    """Product field (neighbouring cells only) = brute-force lattice sum, on
    random points, on points near the channel and on the leaf nodes of the
    n=200, p=22 grid (10 x 10 leaves)."""
    rng = np.random.default_rng(5)
    pts = [rng.uniform(0, 1, (4000, 2))]
    # near the removed half row / half column and their corner
    pts.append(np.stack((rng.uniform(0.44, 1.0, 1500), rng.uniform(0.47, 0.53, 1500)), axis=1))
    pts.append(np.stack((rng.uniform(0.47, 0.53, 1500), rng.uniform(0.44, 1.0, 1500)), axis=1))
    hps = O.HPSOracle(O.helmholtz_pdo(1.0), np.array([[0, 1.0], [0, 1.0]]), 1 / (2 * 200 / 20), 22)
    pts.append(hps.xx_tot)
    xx = np.concatenate(pts)
    ref = O.lshaped_waveguide_bumps(xx)
    got = bif.lshaped_waveguide_bumps(torch.as_tensor(xx)).numpy()
    assert np.abs(got - ref).max() < 1e-14
    kh = 253.3
    c = bif.bfield_lshaped_waveguide(torch.as_tensor(xx), kh).numpy()
    assert np.abs(c - O.bfield_lshaped_waveguide(xx, kh)).max() < 1e-14 * kh ** 2


def test_fourier_field_definition():
    """1 + 0.1 * (sum of 32 modes cos(2 pi k.x + phase), |k| <= 8): on a
    64 x 64 periodic grid the DFT of the field has mean 1 exactly and no
    energy beyond |k| = 8; |b - 1| <= 0.1 everywhere; seeded (same values
    for the same seed, different for another)."""
    m = 64
    g = (np.arange(m) + 0.0) / m
    X, Y = np.meshgrid(g, g, indexing="ij")
    xx = torch.as_tensor(np.stack((X.ravel(), Y.ravel()), axis=1))
    b = bif.fourier_field(xx).numpy().reshape(m, m)
    F = np.fft.fft2(b - 1.0) / m ** 2
    kx = np.fft.fftfreq(m, 1.0 / m)
    KX, KY = np.meshgrid(kx, kx, indexing="ij")
    outside = np.hypot(KX, KY) > 8 + 1e-9
    assert np.abs(F[outside]).max() < 1e-14
    assert abs(F[0, 0]) < 1e-15
    assert np.count_nonzero(np.abs(F) > 1e-12) <= 2 * 32
    assert np.abs(b - 1.0).max() <= 0.1 + 1e-15
    assert torch.equal(bif.fourier_field(xx), bif.fourier_field(xx))
    assert not torch.equal(bif.fourier_field(xx), bif.fourier_field(xx, seed=1))
    assert math.isclose(float(b.mean()), 1.0, abs_tol=1e-15)
