"""Shared helpers of the test suite: cloud construction and oracle runs."""
import numpy as np

import oracle
from nanovdb_editor_amd import synthetic as syn


def preprocessed(cloud):
    """Oracle preprocessing of a raw cloud: (means, opacities, quats, scales, colors) float32."""
    return oracle.preprocess(cloud.means, cloud.opacity_logits, cloud.quats, cloud.log_scales, cloud.sh0)


def small_uniform(n, vs=0.01, seed=3, log_scale_mean=-4.0, log_scale_std=0.5):
    return syn.uniform_cloud(n, vs, log_scale_mean, log_scale_std, seed)


def small_clustered(n, vs=0.005, seed=4):
    return syn.clustered_cloud(n, 8, 0.1, vs, -5.0, 0.7, seed)


def sort_pairs(ijk, gid):
    """Canonical order: voxel (i, j, k) then gaussian id."""
    order = np.lexsort((gid, ijk[:, 2], ijk[:, 1], ijk[:, 0]))
    return ijk[order], gid[order]


def oracle_threads():
    """Host threads for oracle runs: the process's CPU share, at most 16 (the share of one GPU)."""
    import os
    try:
        cpus = len(os.sched_getaffinity(0))
    except AttributeError:
        cpus = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        cpus = min(cpus, int(env))
    return max(1, min(cpus, 16))
