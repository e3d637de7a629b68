"""The C-ABI library loads and exports every entry point include/voxel.h
declares; host-only behaviour that needs no GPU.  CPU only."""
import ctypes as C

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


def test_no_device_reports_nodev_here():
    if V.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(V.VoxelError) as e:
        V.Context()
    assert e.value.status == V.ERR_NODEV
