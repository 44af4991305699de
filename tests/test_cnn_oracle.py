"""The network oracle (oracle/cnn_oracle.py) on the CPU: its backward pass
equals central finite differences of its own forward (FP64), im2col/col2im
are adjoint, and the layout matches the library's flat parameter vector."""
from __future__ import annotations

import numpy as np
import pytest


@pytest.fixture(scope="module")
def co():
    from oracle import cnn_oracle
    return cnn_oracle


def test_layout_sizes(co):
    lay, total = co.layout()
    assert len(lay) == 16
    names = [n for n, _, _ in lay]
    assert names[0] == "conv0.W" and names[-1] == "fc2.b"
    assert lay[0][2] == (64, 32)            # 9 * 4 = 36 rows padded to 64
    assert total % 64 == 0


def test_im2col_col2im_adjoint(co):
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 15, 20, 8))
    cols = co.im2col(x, 8, 10, co.kp(8))
    y = rng.normal(size=cols.shape)
    y[:, 9 * 8:] = 0.0
    lhs = np.sum(cols * y)
    rhs = np.sum(x * co.col2im(y, x.shape, 8, 10))
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_conv_matches_direct_sum(co):
    """A 3x3 stride-2 pad-1 convolution written as an explicit sum."""
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 8, 10, 3))
    W = rng.normal(size=(co.kp(3), 5))
    cols = co.im2col(x, 4, 5, co.kp(3))
    y = cols @ W
    for ho in range(4):
        for wo in range(5):
            ref = np.zeros(5)
            for kh in range(3):
                for kw in range(3):
                    hi, wi = 2 * ho - 1 + kh, 2 * wo - 1 + kw
                    if 0 <= hi < 8 and 0 <= wi < 10:
                        ref += x[0, hi, wi] @ W[(kh * 3 + kw) * 3:(kh * 3 + kw + 1) * 3]
            np.testing.assert_allclose(y[ho * 5 + wo], ref, rtol=1e-12, atol=1e-12)


def test_backward_equals_finite_differences(co):
    rng = np.random.default_rng(2)
    p = co.init_params(3)
    x = rng.uniform(0, 1, size=(2, 60, 80, 4))
    t = rng.normal(size=(2, 6))
    w = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
    loss, g = co.loss_and_grads(p, x, t, w)
    flat = co.pack(p)
    gflat = co.pack(g)
    lay, _ = co.layout()
    eps = 1e-6
    checked = bad = 0
    for name, off, shape in lay:
        n = int(np.prod(shape))
        valid = [i for i in range(n)]
        if name.endswith(".W") and name.startswith("conv"):
            cin = {0: 4, 1: 32, 2: 64, 3: 96, 4: 128, 5: 128}[int(name[4])]
            valid = [i for i in range(n) if i // shape[1] < 9 * cin]
        if name.startswith("fc2"):
            valid = [i for i in range(n) if i % (shape[1] if len(shape) == 2 else 8) < 6]
        for i in rng.choice(valid, size=min(4, len(valid)), replace=False):
            fp = flat.copy()
            fm = flat.copy()
            fp[off + i] += eps
            fm[off + i] -= eps
            lp, _ = co.loss_and_grads(co.unpack(fp), x, t, w)
            lm, _ = co.loss_and_grads(co.unpack(fm), x, t, w)
            fd = (lp - lm) / (2 * eps)
            an = gflat[off + i]
            checked += 1
            if abs(fd - an) > 1e-5 * max(1.0, abs(fd), abs(an)) + 1e-7:
                bad += 1
    # a ReLU kink inside the +-eps interval is the only admissible mismatch
    assert checked >= 60 and bad <= 1, (checked, bad)
