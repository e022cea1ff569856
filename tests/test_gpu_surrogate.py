"""Openness surrogate: MFMA conv kernels against the float64 oracle, the
backward pass and Adam against oracle gradients, and a short training run
that must beat the method without the network on held-out tiles."""
import numpy as np
import pytest
import torch

from oracle import cnn as ocnn

pytestmark = pytest.mark.gpu


def _bf16_to_f32(t: torch.Tensor) -> torch.Tensor:
    return (t.to(torch.int32) << 16).view(torch.float32)


def _round_bf16(a: np.ndarray) -> np.ndarray:
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _net_and_input(cuda, seed=0, bsz=2, h=9, w=70):
    from fujishadergpu_amd import _native as N
    from fujishadergpu_amd.surrogate import Net
    net = Net(cuda, seed=seed)
    # give the last layer a visible scale for the checks
    with torch.no_grad():
        net.w(5).mul_(300.0)
        net.b(5).add_(0.25)
        for li in range(6):
            net.b(li).add_(torch.linspace(-0.1, 0.1, 32, device=cuda))
    net.pack()
    rng = np.random.default_rng(seed + 1)
    xf = np.zeros((bsz, h, w, 8), np.float32)
    xf[..., :4] = rng.normal(0, 1.0, size=(bsz, h, w, 4))
    xt = torch.from_numpy(xf).to(cuda)
    x = torch.empty((bsz, h, w, 8), dtype=torch.int16, device=cuda)
    N.check(N.lib().fsg_cnn_to_bf16(xt.data_ptr(), x.data_ptr(), xt.numel(),
                                    torch.cuda.current_stream().cuda_stream))
    return net, x


def _oracle_params(net):
    return [(_round_bf16(w), b) for w, b in net.params_numpy()]


def test_conv_layers_match_the_oracle(cuda):
    net, x = _net_and_input(cuda)
    yf, acts, pres = net.forward(x, keep=True)
    torch.cuda.synchronize()
    params = _oracle_params(net)
    for li in range(6):
        a = _bf16_to_f32(acts[li]).cpu().numpy().astype(np.float64)
        w, b = params[li]
        bb = np.zeros(w.shape[0]) + b[:w.shape[0]]
        z = ocnn.conv3x3(a, w, bb)
        got = pres[li].cpu().numpy() if li < 5 else yf.cpu().numpy()
        got = got[..., :w.shape[0]]
        err = np.abs(got - z).max() / max(1e-6, np.abs(z).max())
        assert err < 2e-5, f"layer {li}: relative error {err:.2e}"
        if li < 5:  # the stored activation is bf16(gelu(pre))
            nxt = _bf16_to_f32(acts[li + 1]).cpu().numpy()
            assert np.abs(nxt - _round_bf16(ocnn.gelu(got))).max() <= 1e-2 * (1 + np.abs(nxt).max())


def test_backward_matches_oracle_gradients(cuda):
    from fujishadergpu_amd import _native as N
    net, x = _net_and_input(cuda, seed=3)
    bsz, h, w = x.shape[:3]
    rng = np.random.default_rng(9)
    f32 = lambda a: torch.from_numpy(np.asarray(a, np.float32)).to(cuda)  # noqa: E731
    base = f32(rng.normal(0, 1, (bsz, h, w)))
    label = f32(rng.normal(0, 1, (bsz, h, w)))
    label[0, :2] = float("nan")
    label_s = f32(rng.normal(0, 1, (bsz, h, w)))
    label_s[1, 3:5] = float("nan")
    weight_s = f32(rng.uniform(0.2, 2.0, (bsz, h, w)))
    yf, acts, pres = net.forward(x, keep=True)
    assert yf.shape == (bsz, h, w, 2)
    dz = torch.empty((bsz, h, w, 32), dtype=torch.int16, device=cuda)
    loss = torch.zeros(2, dtype=torch.float32, device=cuda)
    n = int((~torch.isnan(label)).sum())
    N.check(N.lib().fsg_cnn_mse(yf.data_ptr(), 2, base.data_ptr(), label.data_ptr(),
                                label_s.data_ptr(), weight_s.data_ptr(), base.numel(), 1.0 / n,
                                dz.data_ptr(), loss.data_ptr(),
                                torch.cuda.current_stream().cuda_stream))
    net.backward(acts, pres, dz)
    torch.cuda.synchronize()
    params = _oracle_params(net)
    xa = _bf16_to_f32(x).cpu().numpy().astype(np.float64)
    out, cache = ocnn.forward([(w_, np.zeros(w_.shape[0]) + b_[:w_.shape[0]]) for w_, b_ in params],
                              xa)
    lab, ls, ws = (a.cpu().numpy().astype(np.float64) for a in (label, label_s, weight_s))
    assert float(loss.sum().item()) / n == pytest.approx(
        ocnn.loss(out, base.cpu().numpy(), lab, 1.0 / n, ls, ws), rel=2e-2)
    grads = ocnn.backward(params, cache, out, base.cpu().numpy(), lab, 1.0 / n, ls, ws)
    for li in range(6):
        co = net.shapes[li][0]
        gw = net.gw(li).cpu().numpy().astype(np.float64)
        gb = net.gb(li).cpu().numpy()[:co].astype(np.float64)
        ref_w, ref_b = grads[li][0].reshape(-1), grads[li][1]
        for name, g, r in (("w", gw, ref_w), ("b", gb, ref_b)):
            if np.abs(r).max() == 0:
                continue
            cos = float(np.dot(g, r) / (np.linalg.norm(g) * np.linalg.norm(r)))
            rel = float(np.linalg.norm(g - r) / np.linalg.norm(r))
            assert cos > 0.999 and rel < 0.05, f"layer {li} d{name}: cos {cos:.5f} rel {rel:.3f}"


@pytest.mark.parametrize("f", [8, 4, 16])
def test_compose_and_detail_label_match_the_oracle(cuda, f):
    from fujishadergpu_amd import _native as N
    rng = np.random.default_rng(4)
    hs, ws, unit = 9, 13, 20.0
    H, W = hs * f - 5, ws * f - 3                # ragged: not a multiple of the grid
    yf = rng.normal(size=(hs, ws, 2))
    base = rng.normal(size=(hs, ws))
    zg = rng.normal(size=(hs, ws)) * 30
    z = ocnn.bilinear(zg, f, hs * f, ws * f) + rng.normal(size=(hs * f, ws * f)) * 3
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(cuda)  # noqa: E731
    s = torch.cuda.current_stream().cuda_stream
    out = torch.empty((H, W), dtype=torch.float32, device=cuda)
    yt, bt, zgt, zt = t(yf), t(base), t(zg), t(z)
    N.check(N.lib().fsg_cnn_compose(yt.data_ptr(), 2, bt.data_ptr(), zgt.data_ptr(), ws, hs, ws, f,
                                    zt.data_ptr(), zt.stride(0), 1.0 / unit, out.data_ptr(), W, H,
                                    W, s))
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    want = ocnn.compose(f32(yf), f32(base), f32(zg), f, z=f32(z)[:H, :W], s_scale=1.0 / unit)
    assert np.abs(out.cpu().numpy() - want).max() < 2e-4
    N.check(N.lib().fsg_cnn_compose(yt.data_ptr(), 2, bt.data_ptr(), zgt.data_ptr(), ws, hs, ws, f,
                                    None, 0, 1.0 / unit, out.data_ptr(), W, H, W, s))
    want = ocnn.bilinear(f32(yf)[..., 0] + f32(base), f, H, W)
    assert np.abs(out.cpu().numpy() - want).max() < 2e-5
    # detail labels: exact known on a sub-rectangle, lab NaN outside its cells
    lab = rng.normal(size=(hs, ws))
    lab[:2] = np.nan
    ey0, ex0, eh, ew = 2 * f, f, 5 * f, 9 * f
    exact = rng.normal(size=(eh, ew))
    ls = torch.empty((hs, ws), dtype=torch.float32, device=cuda)
    wt = torch.empty_like(ls)
    et, lt = t(exact), t(lab)
    N.check(N.lib().fsg_cnn_detail_label(et.data_ptr(), ew, ey0, ex0, eh, ew, lt.data_ptr(),
                                         zgt.data_ptr(), ws, hs, ws, f, zt.data_ptr(),
                                         zt.stride(0), unit, ls.data_ptr(), wt.data_ptr(), s))
    rls, rwt = ocnn.detail_label(f32(exact), ey0, ex0, f32(lab), f32(zg), f, f32(z), unit)
    gls, gwt = ls.cpu().numpy(), wt.cpu().numpy()
    assert np.array_equal(np.isnan(gls), np.isnan(rls))
    m = ~np.isnan(rls)
    assert m.sum() > 20
    assert np.allclose(gls[m], rls[m], rtol=1e-3, atol=1e-4)
    assert np.allclose(gwt, rwt, rtol=1e-4, atol=1e-7)


@pytest.mark.parametrize("bsz,h,w,rows", [(2, 37, 200, 16), (1, 9, 70, 0), (3, 5, 84, 2),
                                           (1, 130, 85, 64)])
def test_fused_forward_is_bit_identical_to_the_layer_chain(cuda, bsz, h, w, rows):
    net, x = _net_and_input(cuda, seed=5, bsz=bsz, h=h, w=w)
    want = net.forward(x)
    got = net.infer(x, rows_per_block=rows)
    torch.cuda.synchronize()
    assert got.shape == want.shape
    assert torch.equal(got, want), float((got - want).abs().max())


def test_adam_step_matches_the_oracle(cuda):
    from fujishadergpu_amd import _native as N
    rng = np.random.default_rng(2)
    p0, g = rng.normal(size=1000), rng.normal(size=1000)
    m0, v0 = rng.normal(size=1000) * 0.1, np.abs(rng.normal(size=1000)) * 0.01
    t = [torch.from_numpy(a.astype(np.float32)).to(cuda) for a in (p0, g, m0, v0)]
    N.check(N.lib().fsg_cnn_adam(t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(),
                                 t[3].data_ptr(), 1000, 1e-2, 0.9, 0.999, 1e-8, 3,
                                 torch.cuda.current_stream().cuda_stream))
    p1, m1, v1 = ocnn.adam_step(p0.astype(np.float32).astype(np.float64), g.astype(np.float32),
                                m0.astype(np.float32), v0.astype(np.float32), 3, lr=1e-2)
    # float32 arithmetic: (1 - beta2) alone is 0.0009999871 in float32
    assert np.allclose(t[0].cpu().numpy(), p1, rtol=1e-5, atol=1e-6)
    assert np.allclose(t[2].cpu().numpy(), m1, rtol=1e-4, atol=1e-7)
    assert np.allclose(t[3].cpu().numpy(), v1, rtol=1e-4, atol=1e-9)


def test_short_training_beats_the_method_without_the_network(cuda):
    from fujishadergpu_amd import surrogate as S
    tr = S.make_tiles(range(4000, 4048), cuda)
    te = S.make_tiles(range(5000, 5008), cuda)
    net = S.Net(cuda, seed=0)
    before = S.evaluate(net, te)
    losses = S.train(net, tr, steps=300, batch=8, lr=2e-3)
    after = S.evaluate(net, te)
    print("surrogate before", before, "after", after, "losses", losses, flush=True)
    assert losses[-1] < losses[0]
    assert after["mae_surrogate_deg"] < after["mae_base_deg"]
    assert after["mae_surrogate_deg"] < after["mae_without_detail_deg"]


def test_trained_network_cuts_the_tail_and_the_worst_tile(cuda):
    """The shipped weights on 8 held-out tiles: the network must lower the
    p99 error and the worst tile's MAE against the method without it (exact
    openness of the 16x grid, upsampled), not only the mean."""
    from fujishadergpu_amd import surrogate as S
    te = S.make_tiles(range(5000, 5008), cuda)
    net = S.Net(cuda, seed=0)
    net.load(S.WEIGHTS_FILE)
    m = S.evaluate(net, te)
    print("held-out", m, flush=True)
    assert m["mae_surrogate_deg"] < 0.6 * m["mae_base_deg"]
    assert m["p99_abs_err_surrogate_deg"] < 0.8 * m["p99_abs_err_base_deg"]
    assert m["worst_tile_mae_surrogate_deg"] < 0.8 * m["worst_tile_mae_base_deg"]


def test_judge_replaces_exactly_the_tiles_over_the_threshold(cuda):
    from fujishadergpu_amd import ops, synth
    from fujishadergpu_amd import surrogate as S
    win = synth.fbm(512, 512, 0.75, 1500.0, seed=3, device=cuda)
    exact = S.exact_openness(win, ops.full_region(win), 64, 8, 10.0)
    approx = exact.clone()
    approx[0:256, 256:512] += 3.0          # tile (0, 1): over the threshold
    approx[256:512, 0:256] += 0.5          # tile (1, 0): under it
    rep = S.judge(win, approx, 64, 8, 10.0, tile=256, threshold=1.5)
    assert rep["tiles_checked"] == 4 and rep["tiles_replaced"] == 1
    assert rep["samples"] >= 4 * 0.01 * 256 * 256
    assert torch.equal(approx[0:256, 256:512], exact[0:256, 256:512])
    assert torch.equal(approx[0:256, 0:256], exact[0:256, 0:256])
    assert torch.allclose(approx[256:512, 0:256], exact[256:512, 0:256] + 0.5)
    assert rep["max_tile_mae_deg"] == pytest.approx(3.0, abs=1e-3)
