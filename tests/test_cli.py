"""The command-line driver's flags and parameter derivation (the reference's
argparse_driver.py:17-38, 86-163) on the CPU: the wavenumber rule, the HPS
leaf size a, the manufactured and pulse data, and the errors of unsupported
choices (they exit before any GPU work)."""
import math

import pytest
import torch

from slablu_amd import argparse_driver as drv
from slablu_amd.src_disc import built_in_funcs as bif

REF_FLAGS = ["--n", "400", "--pde", "bfield_constant", "--bc", "free_space", "--pickle", "tmp_test_file",
             "--disc", "hps", "--p", "22", "--domain", "square", "--ppw", "40", "--solver", "slabLU",
             "--sparse_assembly", "reduced_cpu", "--box_xlim", "1.000000", "--box_ylim", "1.000000",
             "--disable_cuda"]


def test_reference_flags_parse():
    a = drv.parse(REF_FLAGS)
    assert (a.n, a.p, a.ppw, a.pde, a.bc, a.domain, a.solver) == (400, 22, 40, "bfield_constant", "free_space",
                                                                    "square", "slabLU")
    assert a.disable_cuda and not a.periodic_bc and not a.store_sol
    assert a.box_xlim == 1.0 and a.sparse_assembly == "reduced_cpu" and a.precision == "fp64"
    b = drv.parse(REF_FLAGS[:-1] + ["--periodic_bc", "--store_sol", "--precision", "mixed"])
    assert b.periodic_bc and b.store_sol and b.precision == "mixed"


def test_wavenumber_rule():
    # the reference's figure for n = 400, ppw = 40 (tests/test_2d.py's case)
    assert drv.wavenumber(400, 40, None) == pytest.approx(64.82034863101124, rel=1e-15)
    assert drv.wavenumber(1000, 10, None) == pytest.approx((100 + 0.03) * 2 * math.pi + 1.8, rel=1e-15)
    assert drv.wavenumber(1000, None, 2.5) == pytest.approx(2.53 * 2 * math.pi + 1.8, rel=1e-15)
    with pytest.raises(ValueError):
        drv.wavenumber(400, None, None)
    with pytest.raises(ValueError):
        drv.wavenumber(400, 40, 3.0)


@pytest.mark.parametrize("extra,msg", [(["--disc", "fd"], "HPS"), (["--solver", "superLU"], "slabLU"),
                                       (["--pde", "bfield_crystal"], "not provided")])
def test_unsupported_choices_exit(extra, msg, monkeypatch):
    monkeypatch.setattr(torch.cuda, "is_available", lambda: True)
    argv = list(REF_FLAGS)
    for i in range(0, len(extra), 2):
        argv[argv.index(extra[i]) + 1] = extra[i + 1]
    with pytest.raises(SystemExit, match=msg):
        drv.main(argv)


def test_no_gpu_exits():
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(SystemExit, match="GPU"):
        drv.main(REF_FLAGS)


def test_mms_and_pulse_data():
    """u = sin(Lx x) e^(Ly y) and f = -Lap u by central differences; the
    pulse lives on x = 0 only."""
    Lx, Ly, h = 4 * math.pi, 1.0, 1e-4
    xx = torch.tensor([[0.3, 0.7], [0.11, 0.25]], dtype=torch.float64)
    u = lambda x: bif.uu_dir_func_mms(x, Lx, Ly)[:, 0]
    ex, ey = torch.tensor([h, 0.0], dtype=torch.float64), torch.tensor([0.0, h], dtype=torch.float64)
    lap = (u(xx + ex) + u(xx - ex) + u(xx + ey) + u(xx - ey) - 4 * u(xx)) / h ** 2
    assert torch.allclose(bif.ff_body_func_mms(xx, Lx, Ly)[:, 0], -lap, rtol=1e-5)
    pts = torch.tensor([[0.0, 0.5], [0.0, 0.6], [0.5, 0.5]], dtype=torch.float64)
    pu = bif.uu_dir_pulse(pts, 10.0)[:, 0]
    assert pu[0] == 1.0 and pu[1] == pytest.approx(math.exp(-2.0)) and pu[2] == 0.0
