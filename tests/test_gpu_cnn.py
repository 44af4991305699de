"""The warm-start network's HIP kernels (bf16 MFMA GEMMs, fp32 accumulation)
against the FP64 oracle (oracle/cnn_oracle.py), and a short training run.

The GPU rounds weights, inputs, the stored activations of every layer and the
gradients it feeds its GEMMs to bf16.  Against the oracle evaluated with the
same rounding points (everything else FP64) each tensor must agree to a
relative L2 error of 1e-2 (fp32 accumulation order, and the rare bf16 value
that rounds to a neighbour); against the exact FP64 network, forward to 4e-2
and the last layers' gradients to 6e-2.  A transposed operand, a wrong im2col index or
a missing ReLU mask gives errors of order 1.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bf16_round(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


@pytest.fixture(scope="module")
def setup():
    import torch
    from oracle import cnn_oracle as co
    from slam_amd.warmstart import WarmStartNet
    net = WarmStartNet(max_batch=8)
    net.init_params(seed=5)
    flat = net.get_params()
    # the GEMMs read bf16 weights (biases stay fp32)
    p = co.unpack(flat.astype(np.float64))
    for k in p:
        if k.endswith(".W"):
            p[k] = bf16_round(p[k]).astype(np.float64)
    rng = np.random.default_rng(6)
    x = bf16_round(rng.uniform(0, 1, size=(4, 60, 80, 4)))
    xt = torch.from_numpy(x).to(torch.bfloat16).cuda().contiguous()
    target = rng.normal(size=(4, 6)).astype(np.float32)
    return dict(net=net, co=co, p=p, x=x, xt=xt, target=target)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def cos(a, b):
    return float(np.dot(a.ravel(), b.ravel()) / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))


def test_forward_matches_oracle(setup):
    s = setup
    out = s["net"].forward(s["xt"]).cpu().numpy()
    ref, _ = s["co"].forward(s["p"], s["x"])
    assert out.shape == (4, 6)
    assert rel(out, ref) < 4e-2 and cos(out, ref) > 0.998, (rel(out, ref), cos(out, ref))


def test_forward_matches_rounded_oracle(setup):
    s = setup
    out = s["net"].forward(s["xt"]).cpu().numpy()
    ref, _ = s["co"].forward(s["p"], s["x"], rnd=s["co"].bf16)
    assert rel(out, ref) < 1e-2, rel(out, ref)


def test_backward_matches_oracle(setup):
    import torch
    s = setup
    net, co = s["net"], s["co"]
    w = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0], np.float32)
    net.forward(s["xt"])
    loss = net.backward(4, torch.from_numpy(s["target"]).cuda(), w).cpu().numpy()
    G = co.unpack(net.get_grads().astype(np.float64))
    ref_loss, g = co.loss_and_grads(s["p"], s["x"], s["target"], w, rnd=co.bf16)
    assert abs(loss.mean() - ref_loss) < 1e-2 * ref_loss
    for name in g:
        a, b = G[name], g[name]
        assert rel(a, b) < 1e-2 and cos(a, b) > 0.9999, (name, rel(a, b), cos(a, b))
    # against the exact network: the layers next to the loss (the bf16
    # rounding of the 3840 hidden activations and of the weights is the whole
    # difference; it stays below 6 %, and below 1 % above with the same
    # rounding points)
    _, ge = co.loss_and_grads(s["p"], s["x"], s["target"], w)
    for name in ("fc2.W", "fc2.b", "fc1.W", "fc1.b", "conv5.W"):
        assert rel(G[name], ge[name]) < 6e-2, (name, rel(G[name], ge[name]))


def test_training_reduces_loss():
    """2048 rendered pairs, 10 epochs on the GPU with the library's kernels:
    the training loss falls by 40 %, the validation loss falls, and the
    predicted rotation beats the zero-motion guess by 20 % on held-out pairs (translation needs far more data;
    slam_amd/train_warmstart.py trains the shipped weights)."""
    from slam_amd import warmstart as ws
    net = ws.WarmStartNet(max_batch=128)
    net.init_params(seed=1)
    tr = ws.to_device(ws.render_pairs_parallel(3072, seed=11, workers=8, chunk=256))
    va = ws.to_device(ws.render_pairs_parallel(256, seed=12, workers=8, chunk=256))
    hist = ws.train(net, tr, va, epochs=12, batch=128, lr=2e-3)
    first, last = hist[0], hist[-1]
    assert last["train_loss"] < 0.7 * first["train_loss"], (first, last)
    assert last["val_loss"] < first["val_loss"], (first, last)
    assert last["val_rot_err_deg"] < 0.8 * last["zero_guess_rot_deg"], last


def test_forward_is_the_same_in_any_inference_batch():
    """A sample's prediction does not depend on the batch it runs in (up to
    64 samples): the frame loop batches sequences without changing any of
    them.  Batch 8 against 8 forwards of batch 1, and batch 3, bit for bit."""
    import torch
    from slam_amd.warmstart import WarmStartNet
    net = WarmStartNet(max_batch=64)
    net.init_params(seed=9)
    rng = np.random.default_rng(10)
    x = torch.from_numpy(rng.uniform(0, 1, size=(8, 60, 80, 4)).astype(np.float32)).to(torch.bfloat16).cuda()
    full = net.forward(x.contiguous()).cpu().numpy()
    for i in range(8):
        one = net.forward(x[i:i + 1].contiguous()).cpu().numpy()
        assert np.array_equal(one[0], full[i]), i
    three = net.forward(x[2:5].contiguous()).cpu().numpy()
    assert np.array_equal(three, full[2:5])
    net.close()


def test_fused_inference_matches_forward_and_oracle(setup):
    """cnn_infer (the convolutions in one workgroup per sample with the
    activations in LDS, implicit-GEMM; the FC layers in a second launch)
    computes the network of cnn_forward with the same rounding points: the
    two agree to far below the bf16 tolerance, and the fused result meets the
    forward's bound against the rounded oracle."""
    import torch
    net, co, p, x, xt = setup["net"], setup["co"], setup["p"], setup["x"], setup["xt"]
    fwd = net.forward(xt).cpu().numpy().astype(np.float64)
    inf = net.infer(xt).cpu().numpy().astype(np.float64)
    assert rel(inf, fwd) < 2e-3, rel(inf, fwd)
    ref, _ = co.forward(p, x, rnd=co.bf16)
    assert rel(inf, ref) < 1e-2, rel(inf, ref)
    exact, _ = co.forward(p, x)
    assert rel(inf, exact) < 4e-2 and cos(inf, exact) > 0.998


def test_fused_inference_is_per_sample():
    """Every sample of a launch runs in its own workgroup: batch 8 equals 8
    launches of one, bit for bit."""
    import torch
    from slam_amd.warmstart import WarmStartNet
    net = WarmStartNet(max_batch=8)
    net.init_params(seed=11)
    rng = np.random.default_rng(12)
    x = torch.from_numpy(rng.uniform(0, 1, size=(8, 60, 80, 4)).astype(np.float32)).to(torch.bfloat16).cuda()
    full = net.infer(x.contiguous()).cpu().numpy()
    for i in range(8):
        assert np.array_equal(net.infer(x[i:i + 1].contiguous()).cpu().numpy()[0], full[i]), i
    net.close()


def test_split_head_equals_one_launch_head():
    """The FC layers in 16 workgroups per sample (the default) give the bits
    of the FC layers in the convolution workgroup, at batch 1 and 8, and on a
    second call (the arrival tickets are back at 0)."""
    import torch
    from slam_amd.warmstart import WarmStartNet
    net = WarmStartNet(max_batch=8)
    net.init_params(seed=13)
    rng = np.random.default_rng(14)
    x = torch.from_numpy(rng.uniform(0, 1, size=(8, 60, 80, 4)).astype(np.float32)).to(torch.bfloat16).cuda()
    x = x.contiguous()
    net.set_infer_split(False)
    one = net.infer(x).cpu().numpy()
    one1 = net.infer(x[:1].contiguous()).cpu().numpy()
    net.set_infer_split(True)
    for _ in range(2):
        assert np.array_equal(net.infer(x).cpu().numpy(), one)
        assert np.array_equal(net.infer(x[:1].contiguous()).cpu().numpy(), one1)
    net.close()
