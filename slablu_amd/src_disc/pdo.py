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


def get_param_map_and_pdo(geom, bfield, kh):
    """Square geometry: identity parameter maps and the variable-coefficient
    Helmholtz operator (src_disc/pdo.py:321-334).  Curved geometries need the
    c12 cross term and are not provided by this package."""
    if geom == "square":
        param_map = lambda xx: xx.clone()
        inv_param_map = lambda xx: xx.clone()

        def c_func(xx):
            return bfield(xx, kh)
        return PDO_2d(const(1), const(1), c=c_func), param_map, inv_param_map
    raise NotImplementedError("geometry %r: only 'square' is provided" % (geom,))
