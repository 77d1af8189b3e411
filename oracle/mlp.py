"""CPU restatement of the learned step proposer's network -- TEST
INFRASTRUCTURE ONLY (tests/ and nothing else use it).

The 20-64-64-1 tanh MLP of instaheat_amd/csrc/ihg_mlp.hip in float64:
forward pass, mean-squared loss and its gradient by back-propagation, with
the same flat parameter layout
    W1[64][20] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3.
The network has no counterpart in the reference (whose controller is
success(), src/dopri853.c:505-540).  The gradient here is pinned by central
finite differences in tests/test_oracle.py.
"""
import numpy as np

NI, NH = 20, 64
OFF_W1 = 0
OFF_B1 = NH * NI
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NH, NI), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(NH, NH), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3 + b3


def loss(p, x, t):
    return float(np.mean((forward(p, x) - t) ** 2))


def loss_grad(p, x, t):
    """(loss, gradient) of mean (y - t)^2 over the rows of x."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ W3 + b3
    e = y - t
    dy = 2.0 * e / n
    d2 = np.outer(dy, W3) * (1.0 - h2 ** 2)
    d1 = (d2 @ W2) * (1.0 - h1 ** 2)
    g = np.empty(NPARAM)
    g[OFF_W1:OFF_B1] = (d1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    return float(np.mean(e ** 2)), g
