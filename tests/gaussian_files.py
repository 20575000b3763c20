"""Writers of Gaussian PLY / NPZ files for the ingest tests (3DGS property layout)."""
import numpy as np


def write_ply(path, means, opacity, quats, log_scales, sh0, sh_rest=None, big_endian=False, extra_props=0,
              fmt=None):
    """Binary PLY with one vertex element: x y z [extra_k] f_dc_0..2 [f_rest_*] opacity
    scale_0..2 rot_0..3 (the property order of common 3DGS exports, extras interleaved)."""
    n = means.shape[0]
    cols, names = [], []

    def add(name, col):
        names.append(name)
        cols.append(np.asarray(col, np.float32).reshape(n))

    for a, nm in enumerate("xyz"):
        add(nm, means[:, a])
    for e in range(extra_props):
        add(f"nx_{e}", np.full(n, 7.0 + e))
    for a in range(3):
        add(f"f_dc_{a}", sh0[:, a])
    if sh_rest is not None:
        for a in range(sh_rest.shape[1]):
            add(f"f_rest_{a}", sh_rest[:, a])
    add("opacity", opacity)
    for a in range(3):
        add(f"scale_{a}", log_scales[:, a])
    for a in range(4):
        add(f"rot_{a}", quats[:, a])
    data = np.stack(cols, 1).astype(">f4" if big_endian else "<f4")
    form = fmt or ("binary_big_endian" if big_endian else "binary_little_endian")
    header = ["ply", f"format {form} 1.0", f"element vertex {n}"]
    header += [f"property float {nm}" for nm in names]
    header += ["element face 0", "property list uchar int vertex_indices", "end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode())
        f.write(data.tobytes())
