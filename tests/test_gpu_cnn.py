"""The warm-start U-Net on the GPU (through the C ABI) against the NumPy
oracle (oracle/cnn_ref.py, itself pinned to torch autograd and finite
differences), its training, and the LEM warm start it drives.

Tolerances.  Single layers: the device and the oracle round the same
BF16 operands and differ only in accumulation (FP32 MFMA vs FP64), so
outputs agree except for rare one-unit BF16 differences and gradients to
1e-5.  The whole network: the two round to BF16 at the same points, but a
one-unit difference in any stored activation changes every later
pre-activation slightly, which moves other values across BF16 rounding
boundaries; across 14 layers this divergence reaches a few 1e-3 in the
output and a few 1e-2 in the deepest layers' gradients (measured: up to
5e-3 and 6e-2, against 1e-2 and 2e-1 between the device and an unrounded
FP64 network).  So the network test asks for output and loss within 1e-2,
the gradient vector within 5e-2 and each layer's gradient within 1e-1
(relative L2), a cosine above 0.995 per layer, and that the device is
closer to the BF16-emulating oracle than to the FP64 one.
"""
import numpy as np
import pytest

from oracle import cnn_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def learned():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    from lsdtt3_amd import learned as Lm
    return Lm


def rel(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("base,levels,H,W", [(32, 4, 32, 32), (64, 3, 32, 48), (32, 2, 16, 16)])
def test_cnn_forward_backward_match_oracle(learned, base, levels, H, W):
    print(f"[gpu cnn parity] base={base} levels={levels} {H}x{W}", flush=True)
    rng = np.random.default_rng(base + levels)
    net = learned.UNet(4, base, levels, seed=3, tile=H, halo=0)
    p = net.get_params()
    # nonzero biases so that the bias paths are exercised
    lay = R.Layout(4, base, levels)
    for pair in lay.enc + lay.dec:
        for L in pair:
            p[L["boff"]:L["boff"] + L["cout"]] = rng.normal(0, 0.05, L["cout"])
    p[lay.head_b] = 0.1
    net.set_params(p)
    x = rng.normal(size=(2, H, W, 4)).astype(np.float32)
    t = rng.normal(size=(2, H, W)).astype(np.float32)
    mask = (rng.random((2, H, W)) < 0.8).astype(np.uint8)
    y = net.forward(x)
    net.zero_grad()
    loss = net.backward(t, mask)
    g = net.get_grads()
    ref = R.UNetRef(4, base, levels, emulate=True)
    lo, yo, go = ref.loss_and_grad(x, t, mask, p)
    _, yf, gf = R.UNetRef(4, base, levels, emulate=False).loss_and_grad(x, t, mask, p)
    assert rel(y, yo) <= 1e-2, rel(y, yo)
    assert abs(loss - lo) <= 1e-2 * lo
    assert rel(g, go) <= 5e-2, rel(g, go)
    assert rel(g, go) < rel(g, gf)
    for pair in lay.enc + lay.dec:
        for L in pair:
            w = slice(L["woff"], L["woff"] + L["cout"] * 9 * L["cin_pad"])
            b = slice(L["boff"], L["boff"] + L["cout"])
            assert rel(g[w], go[w]) <= 1e-1, (L, rel(g[w], go[w]))
            assert rel(g[b], go[b]) <= 1e-1, (L, rel(g[b], go[b]))
            cos = float(np.dot(g[w], go[w]) / (np.linalg.norm(g[w]) * np.linalg.norm(go[w])))
            assert cos >= 0.995, (L, cos)
    hw = slice(lay.head_w, lay.head_w + base)
    assert rel(g[hw], go[hw]) <= 1e-2
    assert abs(g[lay.head_b] - go[lay.head_b]) <= 1e-2 * abs(go[lay.head_b]) + 1e-6
    # padded input channels of the first layer keep a zero gradient
    L0 = lay.enc[0][0]
    W0 = g[L0["woff"]:L0["woff"] + L0["cout"] * 9 * L0["cin_pad"]].reshape(L0["cout"], 9, L0["cin_pad"])
    assert np.all(W0[..., 4:] == 0.0)


def test_cnn_gradient_accumulates_and_zero_grad_clears(learned):
    rng = np.random.default_rng(11)
    net = learned.UNet(4, 32, 2, seed=1, tile=16, halo=0)
    x = rng.normal(size=(1, 16, 16, 4)).astype(np.float32)
    t = rng.normal(size=(1, 16, 16)).astype(np.float32)
    net.forward(x)
    net.zero_grad()
    net.backward(t)
    g1 = net.get_grads()
    net.backward(t)
    g2 = net.get_grads()
    assert rel(g2, 2 * g1) <= 1e-5
    net.zero_grad()
    assert not np.any(net.get_grads())


def test_cnn_adam_step_matches_formula(learned):
    rng = np.random.default_rng(12)
    net = learned.UNet(4, 32, 2, seed=2, tile=16, halo=0)
    x = rng.normal(size=(2, 16, 16, 4)).astype(np.float32)
    t = rng.normal(size=(2, 16, 16)).astype(np.float32)
    p0 = net.get_params()
    net.forward(x)
    net.zero_grad()
    net.backward(t)
    g = net.get_grads().astype(np.float64)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    net.adam_step(lr, b1, b2, eps)
    p1 = net.get_params()
    m = (1 - b1) * g
    v = (1 - b2) * g * g
    want = p0 - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
    assert np.allclose(p1, want, rtol=1e-5, atol=1e-7)


def test_cnn_training_fits_a_fixed_batch(learned):
    """Adam on one batch: the target is a smooth function of the inputs
    (a local average of channel 2), which the network can represent"""
    print("[gpu cnn training]", flush=True)
    rng = np.random.default_rng(13)
    net = learned.UNet(4, 32, 3, seed=4, tile=32, halo=0)
    x = rng.normal(size=(4, 32, 32, 4)).astype(np.float32)
    c = x[..., 2]
    t = (c + np.roll(c, 1, 1) + np.roll(c, -1, 1) + np.roll(c, 1, 2) + np.roll(c, -1, 2)) / 5.0
    losses = []
    for _ in range(150):
        net.forward(x)
        net.zero_grad()
        losses.append(net.backward(t))
        net.adam_step(1e-3)
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])


def test_cnn_shape_and_argument_errors(learned):
    import lsdtt3_amd._native as N
    with pytest.raises(N.GpuError):
        learned.UNet(4, 48, 4)          # base not a multiple of 32
    with pytest.raises(N.GpuError):
        learned.UNet(4, 64, 4)          # 64 * 8 channels > 256
    net = learned.UNet(4, 32, 4, tile=32, halo=0)
    with pytest.raises(N.GpuError):
        net.forward(np.zeros((1, 12, 16, 4), np.float32))  # 12 not divisible by 8
    with pytest.raises(N.GpuError):
        net.adam_step(lr=-1.0)


# ---------------------------------------------------------------------------
# the LEM warm start
# ---------------------------------------------------------------------------
def _lem(gpu_pkg, z, **kw):
    base = dict(timestep_years=500.0, erodibility=3e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=5e-4,
                hillslope_diffusivity=0.005, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
                boundary_conditions="nnon", newton_rel_tol=1e-8)
    base.update(kw)
    return gpu_pkg.LandscapeModel(z, 30.0, **base)


def test_warm_start_explicit_keeps_the_solution(learned):
    """the explicit-Euler start lowers the starting residual and the step
    still converges to the same tolerance, to the same surface"""
    import lsdtt3_amd
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(128, 160, dx=30.0, seed=21, max_slope=0.6)
    a, b = _lem(lsdtt3_amd, z), _lem(lsdtt3_amd, z)
    learned.set_warm_start(b, learned.WARM_EXPLICIT)
    for _ in range(3):
        sa, sb = a.step(), b.step()
        assert sb["warm_start_mode"] == 1
        assert sb["hillslope_converged"] == 1 and sb["hillslope_residual"] <= sb["hillslope_tolerance"]
        assert sb["warm_start_accepted"] == 1
        assert sb["warm_start_residual"] < sb["hillslope_initial_residual"]
        assert sb["newton_iterations"] <= sa["newton_iterations"]
    tol = 4 * 500.0 * sa["hillslope_tolerance"]
    assert np.abs(a.elevation - b.elevation).max() <= max(tol, 1e-8)


def test_warm_start_previous_step_keeps_the_solution(learned):
    """the previous step's hillslope change as the start: not used on the
    first step (nothing to repeat) nor after the surface is set, accepted
    later with a lower starting residual, and the steps still converge to the
    same tolerance, to the same surface (each within 4 dt tol of the root,
    accumulated over the steps)"""
    import lsdtt3_amd
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(128, 160, dx=30.0, seed=22, max_slope=0.6)
    a, b = _lem(lsdtt3_amd, z), _lem(lsdtt3_amd, z)
    learned.set_warm_start(b, learned.WARM_PREVIOUS)
    bound = 0.0
    for k in range(6):
        sa, sb = a.step(), b.step()
        assert sb["hillslope_converged"] == 1 and sb["hillslope_residual"] <= sb["hillslope_tolerance"]
        if k == 0:
            assert sb["warm_start_mode"] == 0      # no previous step yet
        else:
            assert sb["warm_start_mode"] == 3 and sb["warm_start_accepted"] == 1
            assert sb["warm_start_residual"] < sb["hillslope_initial_residual"]
        bound += 4 * 500.0 * max(sa["hillslope_tolerance"], sb["hillslope_tolerance"])
    assert np.abs(a.elevation - b.elevation).max() <= max(bound, 1e-8)
    b.elevation = b.elevation      # a new surface: the next step starts plainly again
    assert b.step()["warm_start_mode"] == 0


def test_warm_start_learned_trained_on_lem_pairs(learned):
    """train the network on pairs captured from the LEM's own steps and
    show what it gains: its start lowers ||F||inf by a large factor over
    the plain start z^n, and the solves it starts converge to the same
    tolerance and surface as plain Newton"""
    print("[gpu cnn warm start] generate + train", flush=True)
    import lsdtt3_amd
    from lsdtt3_amd.synthetic import lem_surface
    ds = learned.Dataset()
    gens = [_lem(lsdtt3_amd, lem_surface(192, 192, dx=30.0, seed=s, max_slope=0.6)) for s in (31, 32)]
    learned.generate(gens, ds, 12)
    assert len(ds) == 24
    net = learned.UNet(4, 32, 4, seed=5, tile=64, halo=16)
    before, zero = net.evaluate(ds, 2, 8)
    losses = net.train(ds, 300, batch=8, lr=1e-3, seed=1)
    after, zero2 = net.evaluate(ds, 2, 8)
    print(f"  loss zero-pred {zero:.3e} untrained {before:.3e} trained {after:.3e} (last batch {losses[-1]:.3e})",
          flush=True)
    assert after < 0.2 * zero
    z = lem_surface(192, 192, dx=30.0, seed=33, max_slope=0.6)
    a, b = _lem(lsdtt3_amd, z), _lem(lsdtt3_amd, z)
    learned.set_warm_start(b, learned.WARM_LEARNED, net)
    ratios = []
    for _ in range(3):
        sa, sb = a.step(), b.step()
        assert sb["warm_start_mode"] == 2 and sb["warm_start_accepted"] == 1
        assert sb["hillslope_converged"] == 1 and sb["hillslope_residual"] <= sb["hillslope_tolerance"]
        ratios.append(sb["warm_start_residual"] / sb["hillslope_initial_residual"])
        assert sb["newton_iterations"] <= sa["newton_iterations"]
    print(f"  start residual / plain start residual: {ratios}", flush=True)
    assert max(ratios) < 0.5
    tol = 4 * 500.0 * sa["hillslope_tolerance"]
    assert np.abs(a.elevation - b.elevation).max() <= max(tol, 1e-8)


@pytest.mark.parametrize("cin,cout,relu,shape", [(4, 32, 1, (2, 16, 24)), (32, 64, 1, (1, 20, 20)),
                                                 (96, 32, 1, (2, 8, 16)), (64, 32, 0, (3, 7, 16)),
                                                 (256, 256, 1, (2, 4, 4))])
def test_cnn_single_layer_kernels_match_oracle(learned, cin, cout, relu, shape):
    """one layer at a time, where no BF16 rounding can diverge between the
    device and the oracle except on the layer's own output: forward
    outputs equal (a one-unit BF16 difference allowed on at most 0.1% of
    the elements, where FP32 and FP64 accumulation straddle a rounding
    boundary); weight, bias and input gradients to 1e-5 relative (FP32
    accumulation of the same BF16 products)"""
    print(f"[gpu cnn layer] cin={cin} cout={cout} relu={relu} {shape}", flush=True)
    rng = np.random.default_rng(cin * 7 + cout)
    n, H, W = shape
    x = rng.normal(size=(n, H, W, cin)).astype(np.float32)
    w = rng.normal(0, np.sqrt(2.0 / (9 * cin)), size=(cout, 9, cin)).astype(np.float32)
    b = rng.normal(0, 0.1, size=cout).astype(np.float32)
    gy = rng.normal(size=(n, H, W, cout)).astype(np.float32)
    y, gx, gw, gb = learned.conv3x3_layer(x, w, b, relu=bool(relu), gy=gy, input_grad=(cin % 32 == 0))
    xb, wb = R.bf16(x), R.bf16(w)
    pre = R.conv3x3(xb, wb) + b.astype(np.float64)
    yo = R.bf16((np.maximum(pre, 0.0) if relu else pre).astype(np.float32))
    differ = y != yo
    assert differ.mean() <= 1e-3, differ.mean()
    if differ.any():
        ulp = np.abs(yo[differ]) * 2.0 ** -7
        assert np.all(np.abs(y[differ] - yo[differ]) <= ulp * 1.01)
    # backward from the device's own forward output (same mask)
    gm = R.bf16(np.where(y > 0, gy, 0.0).astype(np.float32)) if relu else R.bf16(gy)
    gwo = R.conv3x3_wgrad(gm, xb)
    gbo = gm.reshape(-1, cout).astype(np.float64).sum(axis=0)
    assert rel(gw, gwo) <= 1e-5, rel(gw, gwo)
    assert rel(gb, gbo) <= 1e-5, rel(gb, gbo)
    if cin % 32 == 0:
        gxo = R.conv3x3_xgrad(gm, wb, H, W)
        assert rel(gx, gxo) <= 1e-5, rel(gx, gxo)
