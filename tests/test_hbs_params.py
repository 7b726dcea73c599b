"""HBS tree parameters (host only): leaf block size = the divisor of n + n_pad
nearest above 2k, leaf count padded to a power of two with identity rows
(src_solver/hbs_reconstruction.py:27-63 of the reference)."""
from slablu_amd.src_solver.hbs_reconstruction import HBS_matrix, get_nearest_div


def test_nearest_divisor_rounds_up():
    assert get_nearest_div(1000, 206) == 250
    assert get_nearest_div(256, 48) == 64
    assert get_nearest_div(97, 10) == 97
    assert get_nearest_div(12, 20) == 12


def test_tree_shapes():
    h = HBS_matrix(1, 256, 200)
    assert (h.bs, h.m, h.n_pad) == (256, 1, 0)
    h = HBS_matrix(4, 1000, 103)
    assert (h.bs, h.m, h.n_pad, h.N) == (250, 4, 0, 1000)
    h = HBS_matrix(1, 200, 8)
    assert (h.bs, h.m, h.n_pad, h.N) == (20, 16, 120, 320)
    h = HBS_matrix(2, 256, 32)
    assert (h.bs, h.m, h.n_pad) == (64, 4, 0)
