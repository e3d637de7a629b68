"""Writes tests/golden/golden.json: frozen vectors of the CPU oracle for the
configured workloads, so that later changes to the oracle or to the HIP path
are caught against a fixed record (the reference itself has no fixtures).

    python tests/golden/make_golden.py

Contents: raw noise samples (as float32 bit patterns) per terrain, and SHA-256
digests of whole outputs (block ids, InstanceRaw records, vertex and index
buffers in batch order) for small instances of the configs, plus a frustum
mask digest.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402

TERRAINS = {
    "config1_heightmap": dict(kind=2, seed=42, octaves=1, frequency=1 / 32, lacunarity=2.0, gain=0.5,
                              height_scale=16.0, height_max=15, y_scale=0.0),
    "config2_density": dict(kind=3, seed=7, octaves=4, frequency=1 / 64, lacunarity=2.0, gain=0.5,
                            height_scale=0.0, height_max=0, y_scale=1 / 128),
    "config3_surface": dict(kind=2, seed=11, octaves=4, frequency=1 / 32, lacunarity=2.0, gain=0.5,
                            height_scale=32.0, height_max=31, y_scale=0.0),
    "config5_density": dict(kind=3, seed=5, octaves=4, frequency=1 / 64, lacunarity=2.0, gain=0.5,
                            height_scale=0.0, height_max=0, y_scale=1 / 128),
    "reference_params": dict(kind=3, seed=1, octaves=3, frequency=0.09, lacunarity=0.5, gain=2.5,
                             height_scale=0.0, height_max=0, y_scale=0.0),
    # the reference's own rule and axis order: grass iff fbm3(wx*f, wz*f, wy*f) < 0
    "reference_rule": dict(kind=4, seed=1, octaves=3, frequency=0.09, lacunarity=0.5, gain=2.5,
                           height_scale=0.0, height_max=0, y_scale=0.0),
}


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def grid_nb(nx, nz):
    nb = np.full((nx * nz, 4), -1, np.int32)
    for x in range(nx):
        for z in range(nz):
            i = x * nz + z
            if x > 0:
                nb[i, 0] = (x - 1) * nz + z
            if x < nx - 1:
                nb[i, 1] = (x + 1) * nz + z
            if z > 0:
                nb[i, 2] = x * nz + z - 1
            if z < nz - 1:
                nb[i, 3] = x * nz + z + 1
    return nb


def batch_digest(t, coords, edges, nb):
    ids = [O.chunk_fill(t, *c, e) for c, e in zip(coords, edges)]
    inst_all, vert_all, idx_all, counts = [], [], [], []
    for i, (c, e) in enumerate(zip(coords, edges)):
        adj = [ids[j] if j >= 0 else None for j in nb[i]]
        f, inst, vert, idx = O.mesh_buffers(ids[i], e, tuple(v * e for v in c), adj)
        inst_all.append(inst)
        vert_all.append(vert)
        idx_all.append(idx)
        counts.append(len(f))
    return {
        "coords": [list(map(int, c)) for c in coords],
        "edges": [int(e) for e in edges],
        "neighbours": nb.tolist(),
        "faces": counts,
        "ids_sha256": sha(*ids),
        "instances_sha256": sha(*inst_all),
        "vertices_sha256": sha(*vert_all),
        "indices_sha256": sha(*idx_all),
    }


def main():
    out = {"terrains": TERRAINS, "noise": {}, "batches": {}}
    rng = np.random.default_rng(2024)
    for name, kw in TERRAINS.items():
        t = O.terrain(**kw)
        p = O.perm(kw["seed"])
        pts = rng.integers(-3000, 3000, size=(64, 3))
        if kw["kind"] == 3:
            vals = [O.fbm3(p, float(x), float(y), float(z), kw["octaves"], kw["frequency"],
                           kw["lacunarity"], kw["gain"]) for x, y, z in pts]
        elif kw["kind"] == 4:  # noise sampled at (wx, wz, wy)
            vals = [O.fbm3(p, float(x), float(z), float(y), kw["octaves"], kw["frequency"],
                           kw["lacunarity"], kw["gain"]) for x, y, z in pts]
        else:
            vals = [O.fbm2(p, float(x), float(z), kw["octaves"], kw["frequency"], kw["lacunarity"],
                           kw["gain"]) for x, _, z in pts]
        out["noise"][name] = {"points": pts.tolist(),
                              "bits": np.array(vals, np.float32).view(np.uint32).tolist()}
        del t
    # config 1: the whole 4x4 grid of 16^3 chunks
    t1 = O.terrain(**TERRAINS["config1_heightmap"])
    c1 = [(x, 0, z) for x in range(4) for z in range(4)]
    out["batches"]["config1_grid4x4_e16"] = dict(terrain="config1_heightmap",
                                                 **batch_digest(t1, c1, [16] * 16, grid_nb(4, 4)))
    # config 2: a 3x3 sub-grid of 32^3 chunks at the grid origin
    t2 = O.terrain(**TERRAINS["config2_density"])
    c2 = [(x, 0, z) for x in range(3) for z in range(3)]
    out["batches"]["config2_grid3x3_e32"] = dict(terrain="config2_density",
                                                 **batch_digest(t2, c2, [32] * 9, grid_nb(3, 3)))
    # config 3 shape: mixed edges, no neighbours
    t3 = O.terrain(**TERRAINS["config3_surface"])
    c3 = [(0, 0, 0), (1, 0, 0), (-2, 0, 3)]
    out["batches"]["config3_surface_e32"] = dict(terrain="config3_surface",
                                                 **batch_digest(t3, c3, [32] * 3, np.full((3, 4), -1)))
    # config 5 terrain, 64^3 chunk with neighbours in a row
    t5 = O.terrain(**TERRAINS["config5_density"])
    c5 = [(0, 0, 0), (1, 0, 0)]
    nb5 = np.array([[-1, 1, -1, -1], [0, -1, -1, -1]], np.int32)
    out["batches"]["config5_pair_e64"] = dict(terrain="config5_density",
                                              **batch_digest(t5, c5, [64] * 2, nb5))
    # the reference's own rule, 3x3 grid of 32^3 with in-grid neighbours
    tr = O.terrain(**TERRAINS["reference_rule"])
    cr = [(x - 1, 0, z - 1) for x in range(3) for z in range(3)]
    out["batches"]["reference_rule_grid3x3_e32"] = dict(terrain="reference_rule",
                                                        **batch_digest(tr, cr, [32] * 9, grid_nb(3, 3)))
    # frustum: seeded camera, 64x64 chunk grid at y = 0 and y = -1
    sys.path.insert(0, os.path.join(ROOT, "wgpu-voxel-engine_amd"))
    import camera as T
    eye = np.array([1000.0, 40.0, 1000.0])
    m = T.colmajor(T.perspective(np.radians(70), 16 / 9, 0.1, 1024) @
                   T.look_at(eye, eye + np.array([0.6, -0.1, 0.8])))
    planes = O.frustum_planes(m)
    coords = np.array([(x, y, z) for y in (-1, 0) for x in range(64) for z in range(64)], np.int32)
    mask = O.frustum_contains(planes, coords, 32)
    out["frustum"] = {"matrix_bits": m.view(np.uint32).tolist(), "edge": 32,
                      "coords": "y in (-1, 0), x in 0..63, z in 0..63 (y, x, z order)",
                      "visible": int(mask.sum()), "mask_sha256": sha(mask)}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
