"""Independent float64 numpy statement of the hhe6 network (H/He ionisation +
internal energy), used to pin the oracle's float and double versions and to
compute FP64 truth solutions with scipy.  Equations (time in years, y =
[x_H0, x_H+, x_He0, x_He+, x_He++, e], e in k_B * 1e4 K per H nucleus):

    x_e = x_H+ + x_He+ + 2 x_He++,  mu = sum(x) + x_e,  T_raw = 1e4 e / (1.5 mu)
    T = max(T_raw, 10),  phi = min(1, T_raw / 10)
    R_j = A_j T^b_j exp(-C_j / T) * n_H * yr,  G_j = Gamma_j * yr
    w0 = (R0 x_e + G0) x_H0   w1 = (R1 x_e + G1) x_He0   w2 = (R2 x_e + G2) x_He+
    w3 = R3 x_e x_H+          w4 = (R4 + R5) x_e x_He+   w5 = R6 x_e x_He++
    f = [w3-w0, w0-w3, w4-w1, w1+w5-w2-w4, w2-w5,
         heat - phi (cool_ci + cool_rec + cool_ff)]
    heat = G0 E0 x_H0 + G1 E1 x_He0 + G2 E2 x_He+
    cool_ci = R0 x_e x_H0 chi0 + R1 x_e x_He0 chi1 + R2 x_e x_He+ chi2
    cool_rec = (w3 + w4 + w5) * 0.75 * T / 1e4
    cool_ff = Lff n_H yr sqrt(T) x_e (x_H+ + x_He+ + 4 x_He++)
"""
import numpy as np

YR = 3.15576e7
E = np.array([4.2, 5.8, 11.6])
CHI = np.array([15.78, 28.55, 63.13])
LFF = 1.0285e-15
RECC = 0.75
TFLOOR = 10.0


def _as64(v):
    """float32 constants of the network, widened exactly as the GPU/oracle code sees them."""
    return np.float64(np.float32(v))


def rhs(t, y, nH, p):
    p = np.asarray(p, np.float64)
    lnA, b, C, G = p[0:7], p[7:14], p[14:21], p[21:24] * _as64(YR)
    x0, x1, x2, x3, x4, e = y
    xe = x1 + x3 + 2 * x4
    mu = x0 + x1 + x2 + x3 + x4 + xe
    Traw = 1e4 * e / (1.5 * mu)
    T = max(Traw, TFLOOR)
    phi = min(1.0, Traw * _as64(0.1))
    nHyr = nH * _as64(YR)
    R = np.exp(lnA + b * np.log(T) - C / T) * nHyr
    w0 = (R[0] * xe + G[0]) * x0
    w1 = (R[1] * xe + G[1]) * x2
    w2 = (R[2] * xe + G[2]) * x3
    w3 = R[3] * xe * x1
    w4 = (R[4] + R[5]) * xe * x3
    w5 = R[6] * xe * x4
    heat = G[0] * _as64(E[0]) * x0 + G[1] * _as64(E[1]) * x2 + G[2] * _as64(E[2]) * x3
    cci = (R[0] * xe * x0 * _as64(CHI[0]) + R[1] * xe * x2 * _as64(CHI[1]) + R[2] * xe * x3 * _as64(CHI[2]))
    crec = (w3 + w4 + w5) * RECC * T * _as64(1e-4)
    cff = _as64(LFF) * nHyr * np.sqrt(T) * xe * (x1 + x3 + 4 * x4)
    return np.array([w3 - w0, w0 - w3, w4 - w1, w1 + w5 - w2 - w4, w2 - w5, heat - (cci + crec + cff) * phi])


def jac_fd(y, nH, p, rel=1e-7):
    """Central finite-difference Jacobian, J[i, j] = d f_i / d y_j."""
    y = np.asarray(y, np.float64)
    J = np.zeros((6, 6))
    for j in range(6):
        d = rel * max(abs(y[j]), 1e-8)
        yp, ym = y.copy(), y.copy()
        yp[j] += d
        ym[j] -= d
        J[:, j] = (rhs(0, yp, nH, p) - rhs(0, ym, nH, p)) / (2 * d)
    return J
