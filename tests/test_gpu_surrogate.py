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
        got = pres[li].cpu().numpy() if li < 5 else yf.cpu().numpy()[..., None]
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
    base = torch.from_numpy(rng.normal(0, 1, (bsz, h, w)).astype(np.float32)).to(cuda)
    label = torch.from_numpy(rng.normal(0, 1, (bsz, h, w)).astype(np.float32)).to(cuda)
    label[0, :2] = float("nan")
    yf, acts, pres = net.forward(x, keep=True)
    dz = torch.empty((bsz, h, w, 32), dtype=torch.int16, device=cuda)
    loss = torch.zeros(1, dtype=torch.float32, device=cuda)
    n = int((~torch.isnan(label)).sum())
    N.check(N.lib().fsg_cnn_mse(yf.data_ptr(), base.data_ptr(), label.data_ptr(), yf.numel(),
                                1.0 / n, dz.data_ptr(), loss.data_ptr(),
                                torch.cuda.current_stream().cuda_stream))
    net.backward(acts, pres, dz)
    torch.cuda.synchronize()
    params = _oracle_params(net)
    xa = _bf16_to_f32(x).cpu().numpy().astype(np.float64)
    out, cache = ocnn.forward([(w_, np.zeros(w_.shape[0]) + b_[:w_.shape[0]]) for w_, b_ in params],
                              xa)
    lab = label.cpu().numpy().astype(np.float64)
    assert float(loss.item()) / n == pytest.approx(
        ocnn.loss(out, base.cpu().numpy(), lab, 1.0 / n), rel=2e-2)
    grads = ocnn.backward(params, cache, out, base.cpu().numpy(), lab, 1.0 / n)
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
