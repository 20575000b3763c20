"""float64 central finite differences of the density fit loss, the reference the GPU adjoint is
checked against at the 1M fit configuration.

rho(v) = sum over the pairs (g, v) of o_g exp(-|diag(1/s_g) R_g^T (x_v - mu_g)|^2 / 2), with
x_v = (ijk + 1/2) vs, s = max(exp(log-scale), vs/8), R from the normalised quaternion (w, x, y, z),
o = sigmoid(logit); the loss is L = sum_v (rho(v) - target(v))^2 over the voxels of the pairs.
Each Gaussian's pair support is held at the unperturbed one (the derivative the adjoint defines),
so perturbing one parameter of one Gaussian changes only that Gaussian's terms:
  dL/dtheta = sum_{v in pairs(g)} [(r_v + d+_v)^2 - (r_v + d-_v)^2] / 2h
with r_v the unperturbed residual and d+-_v the change of g's term at v. Steps: 1e-4 of the
smallest scale for means, 1e-4 for log-scales and logits, 1e-4 of |q| for quaternion components,
combined by Richardson extrapolation of the steps h and h/2 (truncation error O(h^4)).
"""
import numpy as np

import oracle
from helpers import preprocessed

# (name, first column, end column, index of the adjoint's output) over theta's 11 columns
FD_GROUPS = (("means", 0, 3, 0), ("log_scales", 3, 6, 1), ("quats", 6, 10, 2), ("opacity_logits", 10, 11, 3))


def rotation(qn):
    W, X, Y, Z = qn[:, 0], qn[:, 1], qn[:, 2], qn[:, 3]
    return np.stack([np.stack([1 - 2 * (Y * Y + Z * Z), 2 * (X * Y - Z * W), 2 * (X * Z + Y * W)], -1),
                     np.stack([2 * (X * Y + Z * W), 1 - 2 * (X * X + Z * Z), 2 * (Y * Z - X * W)], -1),
                     np.stack([2 * (X * Z - Y * W), 2 * (Y * Z + X * W), 1 - 2 * (X * X + Y * Y)], -1)], 1)


def terms_f64(theta, gid, ijk, vs, normalize=True):
    """Density term of every pair from the raw parameters theta [G, 11] = (means 3, log-scales 3,
    quats 4, opacity logit 1); normalize=False takes the quaternions as already normalised."""
    mu, ls, q, ol = theta[:, 0:3], theta[:, 3:6], theta[:, 6:10], theta[:, 10]
    R = rotation(q / np.linalg.norm(q, axis=1, keepdims=True) if normalize else q)
    s = np.maximum(np.exp(ls), float(np.float32(0.125) * np.float32(vs)))
    M = R / s[:, None, :]                                    # u = M^T d
    o = 1.0 / (1.0 + np.exp(-ol))
    d = (ijk.astype(np.float64) + 0.5) * vs - mu[gid]
    u = (M[gid] * d[:, :, None]).sum(1)
    return o[gid] * np.exp(-0.5 * (u * u).sum(1))


def voxel_key(ijk):
    """Sortable int64 key of int voxel coordinates [M, 3] (each within +-2^20)."""
    a = np.asarray(ijk, np.int64) + (1 << 20)
    return (a[:, 0] << 42) | (a[:, 1] << 21) | a[:, 2]


def support(cloud, nthreads=0):
    """The oracle's pair list (gid, ijk) of a raw cloud and its float32 voxel size as a double."""
    vs = float(np.float32(cloud.voxel_size))
    m, op, q, s, col = preprocessed(cloud)
    gid, ijk, _ = oracle.pairs(m, q, s, col, op, vs, want_rgba=False, nthreads=nthreads)
    return gid, ijk, vs


def fd_l2_gradients(cloud, voxels, target=None, residual=None, nthreads=0, pairs=None):
    """[G, 11] central-difference gradients of the L2 loss; voxels [V, 3] lists the grid's voxels.
    The unperturbed residual r_v is `residual` [V] when given (e.g. the forward pass's own
    values - target, so the check covers exactly the adjoint's Jacobian product), else the float64
    density of the raw parameters minus `target` [V]."""
    gid, ijk, vs = pairs if pairs is not None else support(cloud, nthreads)
    G = cloud.count
    theta = np.concatenate([cloud.means, cloud.log_scales, cloud.quats, cloud.opacity_logits[:, None]],
                           1).astype(np.float64)
    vk = voxel_key(voxels)
    order = np.argsort(vk)
    pk = voxel_key(ijk)
    vidx = order[np.searchsorted(vk[order], pk)]
    assert np.array_equal(vk[vidx], pk), "a pair voxel is not among the grid's voxels"
    base = terms_f64(theta, gid, ijk, vs)
    if residual is None:
        rho = np.bincount(vidx, weights=base, minlength=vk.size)
        residual = rho - np.asarray(target, np.float64)
    r_pair = np.asarray(residual, np.float64)[vidx]
    fd = np.zeros((G, 11))
    sig = np.exp(theta[:, 3:6]).min(1)
    qn = np.linalg.norm(theta[:, 6:10], axis=1)
    def central(j, h):
        tp, tm = theta.copy(), theta.copy()
        tp[:, j] += h
        tm[:, j] -= h
        dp = terms_f64(tp, gid, ijk, vs) - base
        dm = terms_f64(tm, gid, ijk, vs) - base
        contrib = ((r_pair + dp) ** 2 - (r_pair + dm) ** 2) / (2 * h[gid])
        return np.bincount(gid, weights=contrib, minlength=G)

    for j in range(11):
        h = 1e-4 * (sig if j < 3 else (qn if 6 <= j < 10 else np.ones(G)))
        # Richardson extrapolation of the central differences at h and h/2 (error O(h^4))
        fd[:, j] = (4.0 * central(j, h / 2) - central(j, h)) / 3.0
    return fd
