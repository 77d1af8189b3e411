"""CPU restatement of DOPRI853's dense output -- TEST INFRASTRUCTURE ONLY
(tests/ and nothing else use it).

Hairer's continuous extension of DOP853 (the coefficients sit unused in the
reference, src/dopri853_constants.c:23-25, 98-173): three extra stages and a
degree-7 interpolant in the nested form
    y(theta) = y + theta (F0 + (1-theta) (F1 + theta (F2 + (1-theta) (F3 +
               theta (F4 + (1-theta) (F5 + theta F6))))))
F0 = y_new - y, F1 = h k1 - F0, F2 = 2 F0 - h (k13 + k1), F3..F6 = h D K.
K holds the 16 stages (K[12] = f(t + h, y_new), K[13..15] the extra ones).
Pinned in tests/test_dense.py against scipy's DOP853 dense output.
"""
import numpy as np


def extra_stage_input(y, K, h, a_row):
    """y + h sum_j a_j K_j over the stages before the extra one"""
    out = y.copy()
    for j, a in enumerate(a_row):
        if a != 0.0:
            out = out + h * a * K[j]
    return out


def interpolate(y, y_new, K, h, theta, D):
    F = [y_new - y]
    F.append(h * K[0] - F[0])
    F.append(2.0 * F[0] - h * (K[12] + K[0]))
    for r in range(4):
        acc = np.zeros_like(y)
        for j in range(16):
            if D[r][j] != 0.0:
                acc = acc + D[r][j] * K[j]
        F.append(h * acc)
    out = np.zeros_like(y)
    for i, f in enumerate(reversed(F)):
        out = out + f
        out = out * (theta if i % 2 == 0 else 1.0 - theta)
    return y + out


def step(rhs, t, y, k1, h, A, B, C):
    """one DOP853 step with the 12-stage table (A rows 0..11, weights B):
    the stage derivatives K[0..11] and y_new"""
    K = [k1]
    for s in range(1, 12):
        ys = y.copy()
        for j in range(s):
            if A[s][j] != 0.0:
                ys = ys + h * A[s][j] * K[j]
        K.append(rhs(t + C[s] * h, ys))
    y_new = y.copy()
    for j in range(12):
        if B[j] != 0.0:
            y_new = y_new + h * B[j] * K[j]
    return K, y_new
