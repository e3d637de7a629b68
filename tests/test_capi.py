"""The C-ABI library loads and exports every entry point include/voxel.h
declares; host-only behaviour that needs no GPU.  CPU only."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import voxel_amd as V


def test_header_lists_entry_points():
    names = V.header_functions()
    assert "vx_chunk_new" in names and "vx_chunk_create_mesh" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    L = C.CDLL(V.LIB_PATH)
    missing = [n for n in V.header_functions() if not hasattr(L, n)]
    assert missing == []


def test_binding_covers_header():
    assert sorted(V._SIGS) == V.header_functions()


def test_status_strings_and_abi():
    L = V.lib()
    assert L.vx_abi_version() == 1
    assert L.vx_status_string(0) == b"ok"
    assert L.vx_status_string(-5) == b"HBM budget exceeded"
    assert L.vx_status_string(-99) == b"unknown status"


def test_invalid_arguments_are_rejected_without_a_device():
    L = V.lib()
    assert L.vx_ctx_create(None, None) == V.ERR_INVALID
    assert L.vx_chunk_new(None, 0, V.Coord(0, 0, 0), 32, None) == V.ERR_INVALID
    assert L.vx_frustum_from_matrix(None, None) == V.ERR_INVALID


def null_args(argtypes):
    """Zero / NULL / default-structure arguments for a binding signature."""
    out = []
    for a in argtypes:
        if isinstance(a, type) and issubclass(a, C.Structure):
            out.append(a())
        elif a in (C.c_int, C.c_int32, C.c_uint32, C.c_uint64, C.c_int64, C.c_float, C.c_double):
            out.append(0)
        else:
            out.append(None)
    return out


def test_every_entry_point_rejects_null_handles():
    """No call panics across the ABI: each entry point given NULL handles and
    pointers (and zero sizes) returns VX_ERR_INVALID, with no device."""
    L = V.lib()
    skip = {"vx_abi_version", "vx_status_string", "vx_ctx_last_error"}
    got = {}
    for name, (_, argtypes) in sorted(V._SIGS.items()):
        if name in skip:
            continue
        got[name] = getattr(L, name)(*null_args(argtypes))
    bad = {n: st for n, st in got.items() if st != V.ERR_INVALID}
    assert bad == {} and len(got) >= 60
    assert L.vx_ctx_last_error(None) == b"no context"


def test_frustum_planes_on_host_match_oracle(oracle_mod):
    rng = np.random.default_rng(0)
    for _ in range(20):
        m = rng.normal(size=16).astype(np.float32)
        f = V.Frustum(m)
        assert (f.planes.view(np.uint32) == oracle_mod.frustum_planes(m).view(np.uint32)).all()
        coords = rng.integers(-4, 4, size=(50, 3))
        host = [f.contains(tuple(int(v) for v in c)) for c in coords]
        assert host == list(oracle_mod.frustum_contains(f.planes, coords).astype(bool))
        # far coordinates: corner positions are exact 64-bit products rounded
        # to f32, defined even where the 32-bit product would overflow
        far = np.concatenate([rng.integers(-2 ** 31, 2 ** 31, size=(20, 3)),
                              [[2 ** 31 - 1, -2 ** 31, 2 ** 26]]]).astype(np.int64)
        host = [f.contains(tuple(int(v) for v in c)) for c in far]
        assert host == list(oracle_mod.frustum_contains(f.planes, far.astype(np.int32)).astype(bool))


def test_no_device_reports_nodev_here():
    if V.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(V.VoxelError) as e:
        V.Context()
    assert e.value.status == V.ERR_NODEV


# ABI struct -> its ctypes mirror in the binding (layout must agree)
MIRRORS = {"vx_coord": "Coord", "vx_terrain": "TerrainC", "vx_ctx_desc": "CtxDesc",
           "vx_frustum": "FrustumC", "vx_request": "Request", "vx_completion": "Completion",
           "vx_server_desc": "ServerDesc", "vx_server_stats": "ServerStats",
           "vx_sim_desc": "SimDesc", "vx_sim_result": "SimResult",
           "vx_tune_result": "TuneResult", "vx_world_desc": "WorldDesc",
           "vx_world_stats": "WorldStats", "vx_batch_req": "BatchReq"}

C_PROGRAM = r"""
#include <stdio.h>
#include "voxel.h"
#define SZ(t) printf(#t " %zu %zu\n", sizeof(t), _Alignof(t));
int main(void) {
    if (vx_abi_version() != VX_ABI_VERSION) return 2;
    if (vx_chunk_new(NULL, 0, (vx_coord){0, 0, 0}, 32, NULL) != VX_ERR_INVALID) return 3;
    SZ(vx_coord) SZ(vx_terrain) SZ(vx_ctx_desc) SZ(vx_frustum) SZ(vx_request)
    SZ(vx_completion) SZ(vx_server_desc) SZ(vx_server_stats) SZ(vx_sim_desc)
    SZ(vx_sim_result) SZ(vx_tune_result) SZ(vx_world_desc) SZ(vx_world_stats)
    SZ(vx_batch_req) SZ(vx_draw_indexed)
    return 0;
}
"""


def c_layouts(tmp_path):
    """(size, alignment) of every ABI struct as a C compiler lays it out."""
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "abi.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "abi"
    libdir = os.path.dirname(V.LIB_PATH)
    inc = os.path.join(os.path.dirname(libdir), "include")
    subprocess.run([cc, "-std=c11", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", inc,
                    str(src), "-o", str(exe), "-L", libdir, "-lvoxel_hip",
                    "-Wl,-rpath," + libdir], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {ln.split()[0]: (int(ln.split()[1]), int(ln.split()[2])) for ln in out.splitlines()}


def header_arity():
    """Parameter count of every function the header declares."""
    text = re.sub(r"/\*.*?\*/", "", open(V.HEADER).read(), flags=re.S)
    out = {}
    for name, params in re.findall(r"(vx_[a-z0-9_]+)\s*\(([^()]*(?:\([^()]*\)[^()]*)*)\)\s*;", text):
        params = params.strip()
        out[name] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def rust_stub():
    text = open(os.path.join(os.path.dirname(os.path.dirname(V.LIB_PATH)), "INTEGRATION.md")).read()
    return re.search(r"```rust\n(.*?)```", text, re.S).group(1)


def rust_layout(ty, structs):
    """(size, alignment) of a Rust field type under #[repr(C)]."""
    ty = ty.strip()
    prim = {"i32": 4, "u32": 4, "f32": 4, "i64": 8, "u64": 8}
    if ty in prim:
        return prim[ty], prim[ty]
    m = re.fullmatch(r"\[(.+);\s*(\d+)\]", ty)
    if m:
        sz, al = rust_layout(m.group(1), structs)
        return sz * int(m.group(2)), al
    return structs[ty]


def test_integration_rust_stub_matches_the_header(tmp_path):
    """The Rust binding shown in INTEGRATION.md declares only functions the
    header has, with the same number of parameters, and every #[repr(C)]
    struct it spells out has the C layout (size and alignment)."""
    stub = rust_stub()
    arity = header_arity()
    fns = re.findall(r"pub fn (vx_[a-z0-9_]+)\((.*?)\)\s*->", stub, re.S)
    assert len(fns) >= 20
    for name, params in fns:
        n = 0 if not params.strip() else params.count(",") + 1 - params.strip().endswith(",")
        assert name in arity, name
        assert n == arity[name], (name, n, arity[name])
    c = c_layouts(tmp_path)
    structs = {}
    for name, body in re.findall(r"#\[repr\(C\)\][^\n]*?pub struct (vx_[a-z_]+) \{(.*?)\}", stub, re.S):
        if "_p: [u8; 0]" in body:
            continue  # opaque handle
        off, align = 0, 1
        for fname, fty in re.findall(r"pub (\w+):\s*([^,]+?)\s*(?:,|$)", body.strip()):
            sz, al = rust_layout(fty, structs)
            off = (off + al - 1) // al * al + sz
            align = max(align, al)
        structs[name] = ((off + align - 1) // align * align, align)
        if name in c:
            assert structs[name] == c[name], (name, structs[name], c[name])
    assert {"vx_coord", "vx_terrain", "vx_request", "vx_completion", "vx_world_desc",
            "vx_draw_indexed"} <= set(structs)


def test_header_is_plain_c_and_links(tmp_path):
    """A C11 program (what a Rust bindgen or cgo consumer sees) compiles
    against include/voxel.h with -Wall -Wextra -pedantic -Werror, links the
    library, calls into it, and sees the same struct sizes and alignments as
    the ctypes binding."""
    sizes = c_layouts(tmp_path)
    for cname, pyname in MIRRORS.items():
        st = getattr(V, pyname)
        assert sizes[cname] == (C.sizeof(st), C.alignment(st)), cname
    assert sizes["vx_draw_indexed"][0] == V.DRAW_DTYPE.itemsize == 20


def test_missing_library_fails_loudly():
    """With the HIP library absent, the binding raises on first use (no CPU
    fallback exists for any entry point)."""
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import voxel_amd as V\n"
            "V.LIB_PATH = '/nonexistent/libvoxel_hip.so'\n"
            "try:\n"
            "    V.Context(0)\n"
            "except RuntimeError as e:\n"
            "    print('raised:', e)\n"
            "    sys.exit(0)\n"
            "sys.exit(1)\n") % os.path.dirname(V.__file__)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "is missing" in r.stdout
