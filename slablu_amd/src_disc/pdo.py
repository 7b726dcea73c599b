"""Second-order PDE operator -c11 u_xx - c22 u_yy - 2 c12 u_xy + c1 u_x
+ c2 u_y + c u (same interface as src_disc/pdo.py:6-31 of the reference).

Coefficients are callables xx (M, 2) -> (M,) or (M, 1) torch tensors on the
device of xx; None means the term is absent (c11 and c22 are required)."""
import torch


class PDO_2d:
    def __init__(self, c11, c22, c12=None, c1=None, c2=None, c=None):
        self.c11, self.c22 = c11, c22
        self.c12 = c12
        self.c1, self.c2 = c1, c2
        self.c = c


def ones(xxloc):
    return torch.ones(xxloc.shape[0], device=xxloc.device, dtype=xxloc.dtype)


def const(c=1):
    return lambda xxloc: c * torch.ones(xxloc.shape[0], device=xxloc.device, dtype=xxloc.dtype)


class _Geometry:
    """A curved domain Psi given as the image of a reference rectangle Omega:
    z(x) maps Omega -> Psi and y(z) is its inverse, with the first and
    second derivatives of y = (y1, y2) with respect to z = (z1, z2)
    (each None when identically zero).  Composing u(z) = v(y(z)) turns
    -Lap_z u + c u into an operator on Omega with
        c11 = |grad y1|^2, c22 = |grad y2|^2, c12 = grad y1 . grad y2,
        c1 = -Lap y1,      c2 = -Lap y2
    evaluated at z(x) (the convention of src_disc/pdo.py:36-131)."""

    def __init__(self, z, y, dy, ddy):
        self.z, self.y = z, y
        self.dy = dy      # {(i, k): callable} for d y_i / d z_k, i, k in (1, 2)
        self.ddy = ddy    # {(i, k): callable} for d^2 y_i / d z_k^2

    def _grad(self, i, zz):
        return [self.dy.get((i, k)) for k in (1, 2)]

    def pdo(self, bfield, kh):
        z = self.z
        dy, ddy = self.dy, self.ddy

        def sq(i):
            fs = [dy.get((i, k)) for k in (1, 2) if dy.get((i, k)) is not None]
            if not fs:
                return None
            return lambda xx: sum(f(z(xx)) ** 2 for f in fs)

        def cross():
            pairs = [(dy.get((1, k)), dy.get((2, k))) for k in (1, 2)
                     if dy.get((1, k)) is not None and dy.get((2, k)) is not None]
            if not pairs:
                return None
            return lambda xx: sum(a(z(xx)) * b(z(xx)) for a, b in pairs)

        def neglap(i):
            fs = [ddy.get((i, k)) for k in (1, 2) if ddy.get((i, k)) is not None]
            if not fs:
                return None
            return lambda xx: -sum(f(z(xx)) for f in fs)

        c = lambda xx: bfield(z(xx), kh)
        return PDO_2d(sq(1), sq(2), c12=cross(), c1=neglap(1), c2=neglap(2), c=c)


def _stack(a, b):
    return torch.stack((a, b), dim=1)


def _sinusoidal():
    """Strip with a sinusoidal top: y = (z1, z2 psi(z1)), psi = 1 - sin(6 z1)/4
    (src_disc/pdo.py:166-189)."""
    psi = lambda t: 1 - 0.25 * torch.sin(6 * t)
    dpsi = lambda t: -1.5 * torch.cos(6 * t)
    ddpsi = lambda t: 9.0 * torch.sin(6 * t)
    z = lambda xx: _stack(xx[:, 0], xx[:, 1] / psi(xx[:, 0]))
    y = lambda zz: _stack(zz[:, 0], zz[:, 1] * psi(zz[:, 0]))
    dy = {(1, 1): lambda zz: torch.ones_like(zz[:, 0]),
          (2, 1): lambda zz: zz[:, 1] * dpsi(zz[:, 0]),
          (2, 2): lambda zz: psi(zz[:, 0])}
    ddy = {(2, 1): lambda zz: zz[:, 1] * ddpsi(zz[:, 0])}
    return _Geometry(z, y, dy, ddy)


def _annulus(amp=0.0, phase=5):
    """Annular sector (amp = 0, src_disc/pdo.py:192-216) or the curvy annulus
    (amp = 0.2, phase 5, src_disc/pdo.py:218-276): the reference coordinate
    x1 is the angle scaled by 3/pi and x2 the radial offset,
        z = (1 + amp cos(phase theta) + x2) (cos theta, sin theta),
        theta = x1 pi / 3,  y1 = 3 theta / pi,  y2 = r - amp cos(phase theta) - 1."""
    s = 3.0 / torch.pi if hasattr(torch, "pi") else 3.0 / 3.141592653589793
    r = lambda zz: torch.sqrt(zz[:, 0] ** 2 + zz[:, 1] ** 2)
    th = lambda zz: torch.atan2(zz[:, 1], zz[:, 0])

    def z(xx):
        t = xx[:, 0] / s
        rad = 1 + amp * torch.cos(phase * t) + xx[:, 1]
        return _stack(rad * torch.cos(t), rad * torch.sin(t))

    y = lambda zz: _stack(s * th(zz), r(zz) - amp * torch.cos(phase * th(zz)) - 1)
    dy = {(1, 1): lambda zz: -s * zz[:, 1] / r(zz) ** 2,
          (1, 2): lambda zz: s * zz[:, 0] / r(zz) ** 2,
          (2, 1): lambda zz: zz[:, 0] / r(zz) - amp * phase * zz[:, 1] * torch.sin(phase * th(zz)) / r(zz) ** 2,
          (2, 2): lambda zz: zz[:, 1] / r(zz) + amp * phase * zz[:, 0] * torch.sin(phase * th(zz)) / r(zz) ** 2}
    ddy = {(2, 1): lambda zz: zz[:, 1] ** 2 / r(zz) ** 3 + amp * phase * zz[:, 1]
                   * (2 * zz[:, 0] * torch.sin(phase * th(zz)) + phase * zz[:, 1] * torch.cos(phase * th(zz)))
                   / r(zz) ** 4,
           (2, 2): lambda zz: zz[:, 0] ** 2 / r(zz) ** 3 + amp * phase * zz[:, 0]
                   * (phase * zz[:, 0] * torch.cos(phase * th(zz)) - 2 * zz[:, 1] * torch.sin(phase * th(zz)))
                   / r(zz) ** 4}
    if amp != 0.0:
        # the angle's Laplacian vanishes; the curvy map keeps both (cancelling)
        # second derivatives, as src_disc/pdo.py:231-232 does
        ddy[(1, 1)] = lambda zz: 2 * s * zz[:, 0] * zz[:, 1] / r(zz) ** 4
        ddy[(1, 2)] = lambda zz: -2 * s * zz[:, 0] * zz[:, 1] / r(zz) ** 4
    return _Geometry(z, y, dy, ddy)


def get_param_map_and_pdo(geom, bfield, kh):
    """(PDO on the reference rectangle, parameter map Omega -> Psi, inverse
    map) for geom in 'square', 'sinusoidal' ('curved'), 'annulus',
    'curvy_annulus' (src_disc/pdo.py:164-334)."""
    if geom == "square":
        param_map = lambda xx: xx.clone()
        inv_param_map = lambda xx: xx.clone()

        def c_func(xx):
            return bfield(xx, kh)
        return PDO_2d(const(1), const(1), c=c_func), param_map, inv_param_map
    if geom in ("sinusoidal", "curved"):
        g = _sinusoidal()
    elif geom == "annulus":
        g = _annulus(0.0)
    elif geom == "curvy_annulus":
        g = _annulus(0.2)
    else:
        raise ValueError("geometry %r is not available" % (geom,))
    return g.pdo(bfield, kh), g.z, g.y
