"""Seeded synthetic Gaussian clouds for the benchmark and test workloads.

Arrays are float32 in the reference ``raster_to_nanovdb_from_arrays`` order and layout
(pymodule/nanovdb_editor/raster.py:231-275): means[N,3], opacity logits[N],
quaternions[N,4] stored (w, x, y, z) and not normalised, log-scales[N,3], sh_0[N,3],
sh_n empty. The arrays are the raw on-disk form; the entry point preprocesses them.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class GaussianCloud:
    means: np.ndarray
    opacity_logits: np.ndarray
    quats: np.ndarray
    log_scales: np.ndarray
    sh0: np.ndarray
    voxel_size: float

    @property
    def count(self) -> int:
        return int(self.means.shape[0])

    def arrays(self):
        """The six arrays in reference order (sh_n empty)."""
        return [self.means, self.opacity_logits, self.quats, self.log_scales, self.sh0,
                np.zeros(0, np.float32)]

    def subset(self, idx) -> "GaussianCloud":
        return GaussianCloud(self.means[idx].copy(), self.opacity_logits[idx].copy(), self.quats[idx].copy(),
                             self.log_scales[idx].copy(), self.sh0[idx].copy(), self.voxel_size)


def _common(rng, n, log_scale_mean, log_scale_std):
    quats = rng.standard_normal((n, 4)).astype(np.float32)
    log_scales = (log_scale_mean + log_scale_std * rng.standard_normal((n, 3))).astype(np.float32)
    opacity_logits = rng.standard_normal(n).astype(np.float32)
    sh0 = rng.standard_normal((n, 3)).astype(np.float32)
    return quats, log_scales, opacity_logits, sh0


def uniform_cloud(n, voxel_size=0.01, log_scale_mean=-4.0, log_scale_std=0.5, seed=1) -> GaussianCloud:
    """Means uniform in [-1,1]^3; log-scales N(mean, std^2) per axis; random quaternions;
    opacity logits N(0,1); sh_0 N(0,1)."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    quats, log_scales, op, sh0 = _common(rng, n, log_scale_mean, log_scale_std)
    return GaussianCloud(means, op, quats, log_scales, sh0, float(voxel_size))


def clustered_cloud(n, clusters=64, sigma=0.1, voxel_size=0.005, log_scale_mean=-5.0, log_scale_std=0.7,
                    seed=2) -> GaussianCloud:
    """Cluster centres uniform in [-1,1]^3; members offset N(0, sigma^2) per axis."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-1.0, 1.0, (clusters, 3))
    member = rng.integers(0, clusters, n)
    means = (centres[member] + sigma * rng.standard_normal((n, 3))).astype(np.float32)
    quats, log_scales, op, sh0 = _common(rng, n, log_scale_mean, log_scale_std)
    return GaussianCloud(means, op, quats, log_scales, sh0, float(voxel_size))


# Named workloads (sizes as stated next to each)
def workload_uniform_100k(seed=1):
    """100k Gaussians, uniform means, log-scales N(-4, 0.5^2), voxel size 0.01."""
    return uniform_cloud(100_000, 0.01, -4.0, 0.5, seed)


def workload_clustered_2m(seed=2):
    """2M Gaussians, 64-cluster mixture (sigma 0.1), log-scales N(-5, 0.7^2), voxel size 0.005."""
    return clustered_cloud(2_000_000, 64, 0.1, 0.005, -5.0, 0.7, seed)


def workload_fit_1m(seed):
    """1M Gaussians as the uniform generator, voxel size 0.005 (fit and target use two seeds)."""
    return uniform_cloud(1_000_000, 0.005, -4.0, 0.5, seed)


def workload_clustered_16m(seed=5):
    """16M Gaussians, 256-cluster mixture, log-scales N(-5.5, 0.6^2), voxel size 0.002."""
    return clustered_cloud(16_000_000, 256, 0.1, 0.002, -5.5, 0.6, seed)


def workload_batch_256(seed=4, count=256, max_size=500_000):
    """256 clouds, sizes log-uniform in [1k, 500k] (capped at ``max_size``), coin flip uniform /
    clustered (64 clusters, sigma 0.1), voxel size uniform in [0.004, 0.02]. Scale law, frozen:
    log-scales N(ln(1.5 * voxel size), 0.4^2) per axis, so every cloud is resolved at about 1.5
    voxels per standard deviation whatever its voxel size."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(1e3), np.log(5e5), count))).astype(np.int64)
    sizes = np.minimum(sizes, max_size)
    kinds = rng.integers(0, 2, count)
    vss = rng.uniform(0.004, 0.02, count)
    clouds = []
    for i in range(count):
        s = int(rng.integers(0, 2**31))
        vs = float(vss[i])
        mean = float(np.log(1.5 * vs))
        if kinds[i] == 0:
            clouds.append(uniform_cloud(int(sizes[i]), vs, mean, 0.4, s))
        else:
            clouds.append(clustered_cloud(int(sizes[i]), 64, 0.1, vs, mean, 0.4, s))
    return clouds
