"""Independent float64 numpy statement of a mass-action network, used to
compute FP64 truth solutions with scipy.  Reaction r: a_r + b_r -> c_r + d_r
with rate w_r = k_r s y_a y_b (s = the system's rate scale); species
a_r, b_r lose w_r each, c_r, d_r gain w_r each (a repeated index counts
once per appearance)."""
import numpy as np


class MassAction64:
    def __init__(self, N, rx, k, scale):
        rx = np.asarray(rx, np.int64).reshape(-1, 4)
        self.N = N
        self.a, self.b, self.c, self.d = rx.T
        # float32 constants widened exactly as the float32 kernels see them
        # (k * scale is rounded to float32 there, as here).
        self.ks = (np.asarray(k, np.float32) * np.float32(scale)).astype(np.float64)
        R = len(self.ks)
        S = np.zeros((N, R))
        for col, sign in ((self.a, -1.0), (self.b, -1.0), (self.c, 1.0), (self.d, 1.0)):
            np.add.at(S, (col, np.arange(R)), sign)
        self.S = S

    def rhs(self, t, y):
        return self.S @ (self.ks * y[self.a] * y[self.b])

    def jac(self, t, y):
        R = len(self.ks)
        D = np.zeros((R, self.N))
        np.add.at(D, (np.arange(R), self.a), self.ks * y[self.b])
        np.add.at(D, (np.arange(R), self.b), self.ks * y[self.a])
        return self.S @ D
