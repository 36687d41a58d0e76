"""CPU pins of the FP64 network oracle (oracle/mlp_oracle.py): a hand-
computed forward pass, finite differences of the loss, BF16 rounding on
known values, and loss/gradient against torch autograd (float64, CPU,
in a subprocess so that torch's HIP runtime never shares a process with
the clexmc library)."""
import json
import os
import subprocess
import sys

import numpy as np

import mlp_oracle as M


def _params(W1=None, b1=None, W2=None, b2=0.0):
    W1 = np.zeros((64, 48)) if W1 is None else W1
    b1 = np.zeros(64) if b1 is None else b1
    W2 = np.zeros(64) if W2 is None else W2
    return np.concatenate([W1.ravel(), b1, W2, [b2]])


def test_forward_hand_computed():
    W1 = np.zeros((64, 48))
    W1[0, 0] = 2.0
    W1[0, 1] = 0.5
    W1[0, 47] = 100.0          # beyond n_env: must be ignored
    b1 = np.zeros(64)
    b1[0] = -1.0
    W2 = np.zeros(64)
    W2[0] = 3.0
    p = _params(W1, b1, W2, 0.5)
    X = np.zeros((2, 48))
    X[0, :3] = [1, -1, 1]
    X[1, :3] = [-1, -1, 1]
    X[:, 47] = 1
    out = M.logits(p, X, n_env=42)
    # row 0: h0 = 2 - 0.5 - 1 = 0.5 -> 3*0.5 + 0.5 = 2.0; flipped h0 = -2 + 0.5 - 1 < 0 -> 0.5
    # row 1: h0 = -2 - 0.5 - 1 < 0 -> 0.5; flipped h0 = 2 + 0.5 - 1 = 1.5 -> 5.0
    assert np.allclose(out, [[2.0, 0.5], [0.5, 5.0]])


def test_loss_closed_form():
    p = _params(b2=0.7)                    # constant logit 0.7
    X = np.zeros((3, 48))
    y = np.array([0.0, 1.0, 0.25])
    loss, g = M.loss_and_grad(p, X, y, 42)
    s = 1 / (1 + np.exp(-0.7))
    ref = -np.mean(y * np.log(s) + (1 - y) * np.log(1 - s))
    assert abs(loss - ref) < 1e-14
    assert abs(g[-1] - np.mean(s - y)) < 1e-14
    assert np.all(g[:-1] == 0)


def test_gradient_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, M.N_PARAMS)
    X = np.where(rng.random((50, 48)) < 0.5, -1.0, 1.0)
    X[:, 42:] = 0
    y = rng.random(50)
    idx = [0, 5, 41, 100, 47 * 64 + 3, 3072, 3100, 3136, 3150, 3200]
    for t in (None, rng.uniform(0.5, 5.0, 50)):
        _, g = M.loss_and_grad(p, X, y, 42, t)
        fd = M.finite_difference(p, X, y, 42, idx, t=t)
        assert np.abs(fd - g[idx]).max() <= 1e-8


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -1.5, 1.0 + 2.0 ** -7 + 2.0 ** -8, 3.0e-3],
                 dtype=np.float32)
    r = M.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0            # tie -> even
    assert r[2] == np.float32(1.0 + 2.0 ** -7)
    assert r[3] == -1.5
    assert r[4] == np.float32(1.0 + 2.0 ** -6)    # tie -> even (odd mantissa rounds up)
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8


_TORCH_SCRIPT = r"""
import json, sys
import numpy as np
import torch
d = json.loads(sys.stdin.read())
p = torch.tensor(d["p"], dtype=torch.float64, requires_grad=True)
X = torch.tensor(d["X"], dtype=torch.float64)
y = torch.tensor(d["y"], dtype=torch.float64)
t = torch.tensor(d["t"], dtype=torch.float64)
n_env = d["n_env"]
mask = torch.zeros(64, 48, dtype=torch.float64); mask[:, :n_env] = 1
W1 = p[:3072].reshape(64, 48) * mask
b1 = p[3072:3136]; W2 = p[3136:3200]; b2 = p[3200]
z = (torch.relu(X @ W1.T + b1) @ W2 + b2) * t
loss = torch.nn.functional.binary_cross_entropy_with_logits(z, y)
loss.backward()
xb = torch.tensor(d["xb"], dtype=torch.float32).to(torch.bfloat16).to(torch.float32)
print(json.dumps({"loss": loss.item(), "grad": p.grad.tolist(), "xb": xb.tolist()}))
"""


def test_loss_gradient_and_bf16_vs_torch():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.4, M.N_PARAMS)
    X = np.where(rng.random((64, 48)) < 0.5, -1.0, 1.0)
    X[:, 42:] = 0
    y = rng.random(64)
    t = rng.uniform(0.5, 5.0, 64)
    xb = rng.normal(0, 1, 1000).astype(np.float32)
    inp = json.dumps({"p": p.tolist(), "X": X.tolist(), "y": y.tolist(), "t": t.tolist(), "n_env": 42,
                      "xb": xb.tolist()})
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    r = subprocess.run([sys.executable, "-c", _TORCH_SCRIPT], input=inp, capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    loss, g = M.loss_and_grad(p, X, y, 42, t)
    assert abs(loss - out["loss"]) <= 1e-12
    assert np.abs(g - np.array(out["grad"])).max() <= 1e-12
    assert np.array_equal(M.bf16_round(xb), np.array(out["xb"], dtype=np.float32))
