"""The autotuner's cache file (VX_TUNE_CACHE): vx_batch_autotune writes one
line per batch shape, "bucket max_edge uniform fused_ok has_nb mode grid
ratio fill_split emit_split"; the first batch a process creates reads the
file once.  Each step below runs in a fresh Python process so that the file
is read again; every run checks that the form the cache selects produces
the same face count.
"""
import os
import subprocess
import sys
import textwrap

import pytest

import voxel_amd as V

pytestmark = pytest.mark.gpu

PKG = os.path.dirname(os.path.abspath(V.__file__))

SCRIPT = textwrap.dedent("""
    import sys
    sys.path.insert(0, {pkg!r})
    import numpy as np
    import voxel_amd as V
    ctx = V.Context(0)
    t = V.Terrain(kind=V.TERRAIN_DENSITY, seed=7, octaves=4, frequency=1 / 64, lacunarity=2.0,
                  gain=0.5, y_scale=1 / 128)
    coords = np.array([(x, 0, z) for x in range(8) for z in range(8)], np.int32)
    b = V.Batch(ctx, coords, 32, ctx.terrain(t), V.grid_neighbours(8, 8), formats=V.FMT_VERTICES)
    if sys.argv[1] == "tune":
        b.autotune(2)
    b.run()
    faces = b.wait()
    print("RESULT", b.mode, faces, flush=True)
    b.close()
    ctx.close()
""").format(pkg=PKG)


def run(step, cache):
    print(f"[tune cache] {step}", flush=True)
    env = dict(os.environ, VX_TUNE_CACHE=cache)
    r = subprocess.run([sys.executable, "-c", SCRIPT, step], env=env, capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT")][-1].split()
    return int(line[1]), int(line[2])


def test_tune_cache_round_trip_and_old_lines(tmp_path):
    cache = str(tmp_path / "tune.txt")
    mode, faces = run("tune", cache)
    lines = [x.split() for x in open(cache).read().splitlines() if x.strip()]
    assert len(lines) == 1 and len(lines[0]) == 10, lines
    key = lines[0][:5]
    assert int(lines[0][5]) == mode
    assert int(lines[0][7]) in (0, 2, 3, 4)
    assert int(lines[0][8]) in (0, 1, 2, 4, 8) and int(lines[0][9]) in range(5)

    def write(*rows):
        with open(cache, "w") as f:
            f.write("".join(" ".join(map(str, r)) + "\n" for r in rows))

    other = ["99999"] + key[1:]
    # pipelined, one workgroup per chunk, one mesher in four; around it a
    # line that is no entry, an entry of the earlier seven-field form and one
    # with a role split outside 2..8 (both for another shape)
    write(["not", "an", "entry"], key + [V.BATCH_PIPELINED, 0, 4], other + [V.BATCH_FUSED, 0],
          other + [V.BATCH_PIPELINED, 0, 1])
    assert run("run", cache) == (V.BATCH_PIPELINED, faces)
    # the earlier seven-field form still reads (default role split)
    write(key + [V.BATCH_FUSED, 0])
    assert run("run", cache) == (V.BATCH_FUSED, faces)
    # the staged form at a tile shape (fill 1 and emit 4 workgroups)
    write(key + [V.BATCH_STAGED, 0, 0, 1, 4])
    assert run("run", cache) == (V.BATCH_STAGED, faces)
    # a tile shape no setter accepts is not loaded: the batch takes the rule
    # (staged, as with no entry), shown with a pipelined entry
    write(key + [V.BATCH_PIPELINED, 0, 3, 3, 0])
    assert run("run", cache) == (V.BATCH_STAGED, faces)
    # a role split outside 2..8 is not loaded: the batch runs staged
    write(key + [V.BATCH_PIPELINED, 0, 1])
    assert run("run", cache) == (V.BATCH_STAGED, faces)
    # an unknown form is not loaded either
    write(key + [7, 0, 0])
    assert run("run", cache) == (V.BATCH_STAGED, faces)
