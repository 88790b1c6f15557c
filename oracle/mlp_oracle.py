"""CPU restatement of the step-size proposal network (test infrastructure
only: used by tests/ as the checker of the device kernels, never by the
product).  Binary64 numpy forward pass and analytic gradient of the mean
squared error of the normalised outputs, for the parameter layout of
unicycle_amd/csrc/mlp_dev.h; pinned by central finite differences in
tests/test_mlp_oracle.py.
"""
import numpy as np

NI, NH = 8, 64
OFF_B1 = NH * NI
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NP = OFF_B3 + 1


def unpack(th):
    th = np.asarray(th, np.float64)
    return (th[:OFF_B1].reshape(NH, NI), th[OFF_B1:OFF_W2], th[OFF_W2:OFF_B2].reshape(NH, NH),
            th[OFF_B2:OFF_W3], th[OFF_W3:OFF_B3], th[OFF_B3])


def forward(th, Xn):
    W1, b1, W2, b2, w3, b3 = unpack(th)
    H1 = np.tanh(Xn @ W1.T + b1)
    H2 = np.tanh(H1 @ W2.T + b2)
    return H2 @ w3 + b3, H1, H2


def loss_grad(th, X, T, norm):
    """norm = (x_mean, x_scale, y_mean, y_scale)"""
    xm, xs, ym, ys = norm
    Xn = (np.asarray(X, np.float64) - xm) / xs
    Tn = (np.asarray(T, np.float64) - ym) / ys
    W1, b1, W2, b2, w3, b3 = unpack(th)
    y, H1, H2 = forward(th, Xn)
    n = Xn.shape[0]
    e = y - Tn
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    D2 = (dy[:, None] * w3[None, :]) * (1.0 - H2 ** 2)
    D1 = (D2 @ W2) * (1.0 - H1 ** 2)
    g = np.empty(NP)
    g[:OFF_B1] = (D1.T @ Xn).ravel()
    g[OFF_B1:OFF_W2] = D1.sum(axis=0)
    g[OFF_W2:OFF_B2] = (D2.T @ H1).ravel()
    g[OFF_B2:OFF_W3] = D2.sum(axis=0)
    g[OFF_W3:OFF_B3] = H2.T @ dy
    g[OFF_B3] = dy.sum()
    return loss, g


def predict(th, X, norm):
    xm, xs, ym, ys = norm
    y, _, _ = forward(th, (np.asarray(X, np.float64) - xm) / xs)
    return y * ys + ym
