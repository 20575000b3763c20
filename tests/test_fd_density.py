"""The float64 finite-difference reference of the fit gradients (fd_density.py) against the oracle:
its per-voxel density sums equal the oracle's float64 density, and its gradients of the L2 fit
loss equal the oracle's float64 adjoint on Gaussians of the 1M fit cloud (which pins both)."""
import numpy as np

import oracle
from fd_density import FD_GROUPS, fd_l2_gradients, support, terms_f64, voxel_key
from helpers import oracle_threads, preprocessed, small_uniform
from nanovdb_editor_amd import synthetic as syn


def _theta(c):
    return np.concatenate([c.means, c.log_scales, c.quats, c.opacity_logits[:, None]], 1).astype(np.float64)


def test_terms_sum_to_oracle_density():
    c = small_uniform(200, 0.01, 91)
    gid, ijk, vs = support(c)
    m, op, q, s, col = preprocessed(c)
    keys, dens = oracle.density_f64(gid, ijk, m, q, s, op, vs)
    vk = voxel_key(keys)
    order = np.argsort(vk)
    vidx = order[np.searchsorted(vk[order], voxel_key(ijk))]
    # the oracle evaluates in double from the float32-preprocessed parameters (the float32-normalised
    # quaternion as it is): written back as raw parameters (log of the scale, logit of the opacity)
    # the terms agree to double rounding
    n = c.count
    pre = np.concatenate([m.reshape(n, 3), np.log(s.reshape(n, 3).astype(np.float64)), q.reshape(n, 4),
                          np.log(op / (1.0 - op.astype(np.float64)))[:, None]], 1).astype(np.float64)
    t = terms_f64(pre, gid, ijk, vs, normalize=False)
    np.testing.assert_allclose(np.bincount(vidx, weights=t, minlength=vk.size), dens, rtol=1e-12, atol=0)
    # from the raw parameters the float32 preprocessing (expf, 1/sqrtf, sigmoid) moves the sums by
    # at most a few float32 ulps of the exponent
    t = terms_f64(_theta(c), gid, ijk, vs)
    np.testing.assert_allclose(np.bincount(vidx, weights=t, minlength=vk.size), dens, rtol=3e-5, atol=0)


def test_fd_matches_oracle_adjoint_on_fit_cloud():
    fit = syn.workload_fit_1m(31)
    idx = np.sort(np.random.default_rng(128).choice(fit.count, 96, replace=False))
    sub = fit.subset(idx)
    gid, ijk, vs = support(sub, oracle_threads())
    m, op, q, s, col = preprocessed(sub)
    keys, dens = oracle.density_f64(gid, ijk, m, q, s, op, vs)
    rng = np.random.default_rng(5)
    target = dens * rng.uniform(0.3, 1.2, dens.size)            # some voxels above, some below
    residual = dens - target
    fd = fd_l2_gradients(sub, keys, residual=residual, pairs=(gid, ijk, vs))
    vk = voxel_key(keys)
    order = np.argsort(vk)
    vidx = order[np.searchsorted(vk[order], voxel_key(ijk))]
    res = 2.0 * residual[vidx]
    adj = oracle.density_grad_f64(gid, ijk, sub.means, sub.log_scales, sub.quats, sub.opacity_logits, vs, res)
    for name, a, b, k in FD_GROUPS:
        ref = adj[k].reshape(sub.count, -1)
        got = fd[:, a:b]
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"FD vs oracle adjoint, {name}: {err:.2e} of max |g|")
        assert err < 1e-8, f"{name}: {err:.2e}"
