"""Curved-domain parameter maps (src_disc/pdo.py:133-334): the maps invert
each other and the transformed coefficients equal the metric of the inverse
map computed by finite differences (c11 = |grad y1|^2, c22 = |grad y2|^2,
c12 = grad y1 . grad y2, c1 = -Lap y1, c2 = -Lap y2)."""
import math

import pytest
import torch

from slablu_amd.src_disc import pdo


@pytest.mark.parametrize("geom,xmax", [("sinusoidal", 1.0), ("annulus", 1.0), ("curvy_annulus", 3.0)])
def test_param_map_coefficients(geom, xmax):
    op, z, y = pdo.get_param_map_and_pdo(geom, lambda zz, kh: -kh ** 2 * torch.ones(zz.shape[0], dtype=zz.dtype),
                                         10.0)
    g = torch.Generator().manual_seed(3)
    xx = torch.rand(300, 2, generator=g, dtype=torch.float64)
    xx[:, 0] *= xmax                         # curvy annulus: stay on one branch of atan2
    zz = z(xx)
    assert (y(zz) - xx).abs().max() < 1e-12
    h = 1e-6
    J = torch.zeros(300, 2, 2, dtype=torch.float64)
    for k in range(2):
        e = torch.zeros(2, dtype=torch.float64)
        e[k] = h
        J[:, :, k] = (y(zz + e) - y(zz - e)) / (2 * h)
    hh = 1e-4
    lap = torch.zeros(300, 2, dtype=torch.float64)
    for k in range(2):
        e = torch.zeros(2, dtype=torch.float64)
        e[k] = hh
        lap += (y(zz + e) - 2 * y(zz) + y(zz - e)) / hh ** 2
    c11 = J[:, 0, 0] ** 2 + J[:, 0, 1] ** 2
    c22 = J[:, 1, 0] ** 2 + J[:, 1, 1] ** 2
    c12 = J[:, 0, 0] * J[:, 1, 0] + J[:, 0, 1] * J[:, 1, 1]
    ev = lambda f: f(xx) if f is not None else torch.zeros(300, dtype=torch.float64)
    assert (ev(op.c11) - c11).abs().max() < 1e-8
    assert (ev(op.c22) - c22).abs().max() < 1e-8
    assert (ev(op.c12) - c12).abs().max() < 1e-8
    assert (ev(op.c1) + lap[:, 0]).abs().max() < 1e-5
    assert (ev(op.c2) + lap[:, 1]).abs().max() < 1e-5
    assert (op.c(xx) + 100.0).abs().max() < 1e-12


def test_unknown_geometry_raises():
    with pytest.raises(ValueError):
        pdo.get_param_map_and_pdo("corner", None, 1.0)
