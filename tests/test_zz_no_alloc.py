"""fsg_topousm_noalloc (the entry point that allocates nothing) against
fsg_topousm with a caller workspace, through the C ABI, on two calls in a
row -- the second on a smaller window, so it runs on the scratch array the
first one left behind."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dem(h, w, seed, nan_discs=0):
    from fujishadergpu_amd import synth
    d = synth.fbm(h, w, 0.75, 3776.0, seed=seed, device="cpu")
    if nan_discs:
        discs = synth.nodata_discs(h, w, 0.01, seed=seed + 1)[:nan_discs]
        synth.apply_discs(d, discs, float("nan"))
    return d


def _params(radii, weights, pixel_size):
    from fujishadergpu_amd.algorithms import TopoUSMFastAlgorithm
    return TopoUSMFastAlgorithm().params_struct(
        {"radii": radii, "weights": weights, "pixel_size": pixel_size})


def test_noalloc_topousm_equals_the_workspace_entry_point(cuda):
    from fujishadergpu_amd import _native as N
    from fujishadergpu_amd import ops
    from fujishadergpu_amd.encoding import quantize_params
    qp = quantize_params(-1.176, 1.176, "uint8")
    cases = [
        # 2300 x 2500 window, r = 128 decimated by 4 (its grid in the scratch
        # array), NaN discs (void fill), region inside a halo of 160
        (2300, 2500, 11, 40, [2, 8, 32, 128], [8 / 15, 4 / 15, 2 / 15, 1 / 15], 5.0, 160),
        # then a smaller window on what the first call left in the array
        (1100, 1300, 12, 6, [3, 12, 128], [0.5, 0.3, 0.2], 5.0, 144),
    ]
    for h, w, seed, discs, radii, weights, px, halo in cases:
        print("case", (h, w), "radii", radii, "discs", discs, flush=True)
        g = _dem(h, w, seed, nan_discs=discs).to(cuda)
        prm = _params(radii, weights, px)
        assert any(int(prm.ds[i]) > 1 for i in range(prm.n_radii)) == (128 in radii)
        reg = ops.region(halo, halo, h - 2 * halo, w - 2 * halo)
        stats = (23.5, 0.0)
        outs = []
        for entry in ("workspace", "noalloc"):
            f32 = torch.full((reg.h, reg.w), -1.0, dtype=torch.float32, device=cuda)
            u8 = torch.zeros((reg.h, reg.w), dtype=torch.uint8, device=cuda)
            o1 = ops.output(f32, post=N.POST_DIV, host_stats=stats)
            o2 = ops.output(u8, qp=qp, post=N.POST_DIV, host_stats=stats)
            if entry == "workspace":
                ops.topousm(g, reg, prm, o1, ops.Workspace(), out2=o2)
            else:
                ops.topousm_noalloc(g, reg, prm, o1, out2=o2)
            torch.cuda.synchronize()
            outs.append((f32, u8))
        (a32, a8), (b32, b8) = outs
        assert bool(torch.isnan(a32).any()) == bool(discs)
        assert torch.equal(torch.isnan(a32), torch.isnan(b32))
        assert torch.equal(torch.nan_to_num(a32), torch.nan_to_num(b32))
        assert torch.equal(a8, b8)


def test_noalloc_topousm_rejects_windows_above_the_maximum(cuda):
    from fujishadergpu_amd import _native as N
    from fujishadergpu_amd import ops
    L = N.lib()
    # 8192^2 with r = 128 at ds 4: six 2048^2 float grids, about 100 MB
    g = torch.zeros((8192, 8192), dtype=torch.float32, device=cuda)
    prm = _params([2, 8, 32, 128], [0.4, 0.3, 0.2, 0.1], 5.0)
    w = ops.window(g)
    need = L.fsg_topousm_workspace_size(C.byref(w), C.byref(prm))
    assert need > (64 << 20)
    out = torch.empty((16, 16), dtype=torch.float32, device=cuda)
    rc = L.fsg_topousm_noalloc(C.byref(w), C.byref(ops.region(0, 0, 16, 16)), C.byref(prm),
                               C.byref(ops.output(out)), None, None)
    assert rc == N.FSG_E_WORKSPACE
    assert b"fixed maximum" in L.fsg_last_error()
