"""Host-side logic of the tile pipeline (no GPU): the exactness rule for
sharing raster-wide coarse means between topousm_fast windows."""
import pytest

from fujishadergpu_amd.params import coarse_radius, shared_coarse_ok


def test_coarse_radius_rounds_like_the_reference():
    # max(1, round(r / ds)) with Python's round (half to even)
    assert coarse_radius(128, 4) == 32
    assert coarse_radius(2048, 16) == 128
    assert coarse_radius(300, 8) == 38      # 37.5 -> 38
    assert coarse_radius(20, 8) == 2        # 2.5 -> 2
    assert coarse_radius(3, 16) == 1


def test_stratified_prepass_window_of_the_16384_job_shares():
    # 8256^2 window at (4064, 4064), trimmed by 2064: the radius-2048 box
    # (ds 16, 128 coarse cells) reaches coarse rows 0 .. 515 exactly
    reg = (2064, 2064, 4128, 4128)
    assert shared_coarse_ok((4064, 4064), (8256, 8256), reg, 16, 128, (16384, 16384))
    assert shared_coarse_ok((4064, 4064), (8256, 8256), reg, 4, 32, (16384, 16384))
    # one coarse cell more of radius would reflect at the window edge
    assert not shared_coarse_ok((4064, 4064), (8256, 8256), reg, 16, 129, (16384, 16384))


def test_main_tiles_of_the_16384_job_share():
    # interior tile window (4096 + 2 * 160) and the raster-corner tile
    assert shared_coarse_ok((3936, 7968), (4416, 4416), (160, 160, 4096, 4096), 4, 32,
                            (16384, 16384))
    assert shared_coarse_ok((0, 0), (4256, 4256), (0, 0, 4096, 4096), 4, 32, (16384, 16384))


def test_misaligned_window_does_not_share():
    reg = (160, 160, 4096, 4096)
    assert not shared_coarse_ok((3937, 7968), (4416, 4416), reg, 4, 32, (16384, 16384))
    assert not shared_coarse_ok((3936, 7968), (4417, 4416), reg, 4, 32, (16384, 16384))


def test_region_near_an_inner_window_edge_does_not_share():
    # region starting at the window's first row, window edge inside the raster
    assert not shared_coarse_ok((4000, 4000), (1024, 1024), (0, 0, 512, 512), 4, 8,
                                (8192, 8192))
    # the same region on the raster's own edge reflects alike on both grids
    assert shared_coarse_ok((0, 4000), (1024, 1024), (0, 200, 512, 512), 4, 8, (8192, 8192))


def test_window_spanning_the_raster_always_shares():
    assert shared_coarse_ok((0, 0), (96, 80), (0, 0, 96, 80), 8, 64, (96, 80))


def test_gaussian_coarse_radius_uses_its_truncation_margin():
    # r <= 1 is the sigma=1 Gaussian truncated at 4 cells
    # ds-grid 32 x 32, zoom 31/511: region row 80 -> cell 4, row 431 -> cells 26..27
    assert shared_coarse_ok((64, 64), (512, 512), (80, 80, 352, 352), 16, 1, (1024, 1024))
    # row 64 -> cell 3: the truncated Gaussian would reflect at the window edge
    assert not shared_coarse_ok((64, 64), (512, 512), (64, 64, 384, 384), 16, 1, (1024, 1024))
    # a box of radius 2 needs less margin than the Gaussian
    assert shared_coarse_ok((64, 64), (512, 512), (64, 64, 384, 384), 16, 2, (1024, 1024))
