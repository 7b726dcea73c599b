"""Dirichlet data and coefficient fields (src_disc/built_in_funcs.py of the
reference for bfield_constant and uu_dir_func_greens; the variable fields of
the benchmark workloads are defined here)."""
import math

import numpy as np
import torch
from scipy.special import hankel1


def uu_dir_func_greens(xx, kh, center=(-0.1, 0.5)):
    """Re[(i/4) H0^(1)(kh |x - center|)]  (kh != 0) or (1/pi) log |x-c|^2.

    Same values as src_disc/built_in_funcs.py:59-70 (scipy.special.hankel1
    on the host); returns an (M, 1) tensor on the device of xx."""
    x = xx.detach().to("cpu", torch.float64).numpy()
    c = np.asarray(center.cpu() if torch.is_tensor(center) else center, dtype=np.float64)
    dd0 = x[:, 0] - c[0]
    dd1 = x[:, 1] - c[1]
    ddsq = dd0 * dd0 + dd1 * dd1
    if kh == 0:
        u = (1 / np.pi) * np.log(ddsq)
    else:
        u = np.real((1j / 4) * hankel1(0, kh * np.sqrt(ddsq)))
    return torch.from_numpy(np.ascontiguousarray(u)).to(xx.device).unsqueeze(-1)


def uu_dir_func_greens_multi(xx, kh, centers):
    """One column of Green's-function data per source centre: (M, len(centers))."""
    cols = [uu_dir_func_greens(xx, kh, center=c) for c in centers]
    return torch.cat(cols, dim=1)


def uu_dir_func_mms(xx, Lx, Ly):
    """sin(Lx x) exp(Ly y), (M, 1): the manufactured solution of the Poisson
    test (src_disc/built_in_funcs.py:37-41)."""
    xx = torch.as_tensor(xx)
    return (torch.sin(Lx * xx[:, 0]) * torch.exp(Ly * xx[:, 1])).unsqueeze(-1)


def ff_body_func_mms(xx, Lx, Ly):
    """-Lap of uu_dir_func_mms: (Lx^2 - Ly^2) sin(Lx x) exp(Ly y)
    (src_disc/built_in_funcs.py:43-48)."""
    return (Lx ** 2 - Ly ** 2) * uu_dir_func_mms(xx, Lx, Ly)


def uu_dir_pulse(xx, kh):
    """A Gaussian pulse exp(-200 (y - 1/2)^2) on the left boundary x = 0,
    zero elsewhere (src_disc/built_in_funcs.py:50-58)."""
    xx = torch.as_tensor(xx)
    u = torch.zeros(xx.shape[0], 1, dtype=torch.float64, device=xx.device)
    left = xx[:, 0] == 0
    u[left, 0] = torch.exp(-200.0 * (xx[left, 1] - 0.5) ** 2)
    return u


def annulus_sources(count=64, seed=0, inner=1.2, outer=2.0, center=(0.5, 0.5)):
    """`count` source points drawn uniformly (by area, seeded) from the
    annulus inner < |x - center| < outer: a list of (x, y) tuples, for the
    multi-right-hand-side workload (one Green's-function column each)."""
    g = torch.Generator().manual_seed(seed)
    u = torch.rand(count, generator=g, dtype=torch.float64)
    th = torch.rand(count, generator=g, dtype=torch.float64) * (2 * math.pi)
    r = torch.sqrt(inner ** 2 + u * (outer ** 2 - inner ** 2))
    return [(float(center[0] + r[i] * math.cos(th[i])), float(center[1] + r[i] * math.sin(th[i])))
            for i in range(count)]


def bfield_constant(xx, kh):
    """c(x) = -kh^2 (src_disc/built_in_funcs.py:79-80)."""
    return -(kh ** 2) * torch.ones(xx.shape[0], device=xx.device, dtype=xx.dtype)


def lshaped_waveguide_bumps(xx, amp=1.0, width=0.005, spacing=0.04, removed_row=None, removed_col=None):
    """sum of Gaussian bumps exp(-|x - x_k|^2 / width^2) * amp on the square
    lattice x_k = spacing*(i + 1/2), with one lattice row and one lattice
    column removed so that the missing bumps form an L-shaped channel."""
    n = int(round(1.0 / spacing))
    rr = n // 2 if removed_row is None else removed_row
    cc = n // 2 if removed_col is None else removed_col
    b = torch.zeros(xx.shape[0], device=xx.device, dtype=xx.dtype)
    centres = (torch.arange(n, device=xx.device, dtype=xx.dtype) + 0.5) * spacing
    # each node only sees the bumps of its own and the neighbouring lattice
    # cells (exp(-(0.04/0.005)^2) = 1.6e-28 beyond that)
    ix = torch.clamp((xx[:, 0] / spacing).floor().long(), 0, n - 1)
    iy = torch.clamp((xx[:, 1] / spacing).floor().long(), 0, n - 1)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            jx, jy = ix + dx, iy + dy
            ok = (jx >= 0) & (jx < n) & (jy >= 0) & (jy < n)
            ok = ok & ~((jy == rr) & (jx >= cc)) & ~((jx == cc) & (jy >= rr))
            jxc, jyc = jx.clamp(0, n - 1), jy.clamp(0, n - 1)
            d2 = (xx[:, 0] - centres[jxc]) ** 2 + (xx[:, 1] - centres[jyc]) ** 2
            b = b + torch.where(ok, amp * torch.exp(-d2 / (width * width)), torch.zeros_like(d2))
    return b


def bfield_lshaped_waveguide(xx, kh):
    """c(x) = -kh^2 (1 - b(x)) with b the L-shaped bump lattice."""
    return -(kh ** 2) * (1.0 - lshaped_waveguide_bumps(xx))


def fourier_field(xx, nmodes=32, kmax=8, seed=0, amp=0.1):
    """1 + amp * (sum of nmodes seeded random Fourier modes with |k| <= kmax),
    normalised so the mode sum has unit maximum coefficient magnitude."""
    g = torch.Generator().manual_seed(seed)
    ks = []
    while len(ks) < nmodes:
        k = torch.randint(-kmax, kmax + 1, (2,), generator=g)
        if 0 < float(k.double().norm()) <= kmax:
            ks.append(k)
    K = torch.stack(ks).to(xx.device, xx.dtype)
    ph = torch.rand(nmodes, generator=g, dtype=torch.float64).to(xx.device, xx.dtype) * 2 * math.pi
    w = (torch.rand(nmodes, generator=g, dtype=torch.float64) * 2 - 1).to(xx.device, xx.dtype) / nmodes
    arg = 2 * math.pi * (xx @ K.T) + ph
    return 1.0 + amp * (torch.cos(arg) @ w)
