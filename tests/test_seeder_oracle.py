"""The learned seeder's CPU restatement (oracle/src/seeder.cpp, test
infrastructure): its analytic gradient against central finite differences of
its own loss, in double precision, on parameters of every layer."""
import ctypes

import numpy as np

from helpers import P, oracle, oracle_seeder, seeder_batch, seeder_params


def test_oracle_gradient_matches_finite_differences():
    params = seeder_params(3)
    x, t = seeder_batch(2, 4)
    loss, _, g = oracle_seeder(params, x, t)
    lib = oracle()
    lib.orc_seeder_loss_at.restype = ctypes.c_double
    lib.orc_seeder_loss_at.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double]
    n = len(params)
    rng = np.random.default_rng(5)
    # a sample of every block: conv weights and biases of each layer, the linear layer
    blocks = [(0, 576), (576, 608), (608, 9824), (9824, 9856), (9856, 19072), (19072, 19104), (19104, 28320),
              (28320, 28352), (28352, 31424), (31424, 31430)]
    assert blocks[-1][1] == n
    h = 1e-5
    worst = 0.0
    for a, b in blocks:
        for i in rng.choice(np.arange(a, b), size=min(6, b - a), replace=False):
            lp = lib.orc_seeder_loss_at(P(params), P(x), P(t), 2, int(i), h)
            lm = lib.orc_seeder_loss_at(P(params), P(x), P(t), 2, int(i), -h)
            fd = (lp - lm) / (2 * h)
            err = abs(fd - g[i]) / max(1e-3, abs(fd))
            worst = max(worst, err)
            assert err <= 1e-4, (i, fd, g[i])
    print(f"loss {loss:.4f}, worst relative error {worst:.2e}")


def test_oracle_loss_is_the_mean_squared_error():
    params = seeder_params(1)
    x, t = seeder_batch(3, 2)
    loss, y, _ = oracle_seeder(params, x, t)
    assert abs(loss - float(np.mean((y.astype(np.float64) - t) ** 2))) <= 1e-6 * max(1.0, loss)
