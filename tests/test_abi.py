"""CPU tests of the C ABI: the gfx950 libraries load, export every function
declared in include/*.h, and fail loudly (no CPU fallback) without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "return", "sizeof")))


def _lib(name):
    path = os.path.join(ROOT, "corruscant_amd", "bin", name)
    if not os.path.exists(path):
        pytest.skip("library not built: " + name)
    return ctypes.CDLL(path)


def test_kdtree_exports_every_declared_function():
    lib = _lib("libkdtree.so")
    names = _declared("corruscant_amd.h")
    assert {"tree_construct", "pair_count", "destroy", "pair_count_jobs", "kd_last_error"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n


def test_coords_exports_every_declared_function():
    lib = _lib("libcoords.so")
    names = _declared("corruscant_coords.h")
    assert {"radec2cart64", "radecdist2cart64", "radec2sph64"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n


def test_kdtree_t_layout_matches_reference_offsets():
    """node_data/size/memsize/num_fields at the offsets of src/kdtree.h:33-41
    on x86-64, 64 bytes in total."""
    import corruscant_amd.clustering as C
    assert ctypes.sizeof(C.kdtree) == 64
    assert C.kdtree.node_data.offset == 0
    assert C.kdtree.size.offset == 8
    assert C.kdtree.memsize.offset == 12
    assert C.kdtree.num_fields.offset == 16
    assert ctypes.sizeof(C.node) == 40 and ctypes.sizeof(C.datum) == 32


def test_no_gpu_means_an_error_not_a_fallback():
    import torch  # noqa: F401  (only to ask whether a GPU is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from corruscant_amd import clustering
    with pytest.raises(RuntimeError, match="HIP device"):
        clustering.tree(np.random.default_rng(0).random((100, 3)))


def test_error_paths_before_any_device_work():
    lib = _lib("libkdtree.so")
    lib.kd_last_error.restype = ctypes.c_char_p
    lib.kd_set_tile_size.restype = ctypes.c_int
    assert lib.kd_set_tile_size(0) == -1
    assert b"tile size" in lib.kd_last_error()
    assert lib.kd_set_tile_size(100) == -1
    lib.kd_set_stage_pool.restype = ctypes.c_int
    lib.kd_set_stage_pool.argtypes = [ctypes.c_longlong]
    assert lib.kd_set_stage_pool((1 << 31) + 1) == -1
    assert b"stage_pool" in lib.kd_last_error()
    assert lib.kd_set_stage_pool(0) == 0


def test_plan_entry_points_refuse_bad_arguments_before_any_device_work():
    lib = _lib("libkdtree.so")
    lib.kd_last_error.restype = ctypes.c_char_p
    lib.pair_count_plan.restype = ctypes.c_void_p
    lib.pair_count_plan.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p]
    assert lib.pair_count_plan(None, 0, None, 0, 0, None) is None
    assert b"no jobs" in lib.kd_last_error()
    lib.pair_count_plan_run.restype = ctypes.c_int
    lib.pair_count_plan_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    assert lib.pair_count_plan_run(None, None) == -1
    assert b"no plan" in lib.kd_last_error()
    lib.pair_count_plan_items.restype = ctypes.c_longlong
    lib.pair_count_plan_items.argtypes = [ctypes.c_void_p]
    assert lib.pair_count_plan_items(None) == -1
    lib.pair_count_plan_free.argtypes = [ctypes.c_void_p]
    lib.pair_count_plan_free(None)


def test_tree_field_validation_before_any_device_work():
    """clustering.tree rejects field ids that are not exactly 0..N-1, and a
    single field, before it touches the device (clustering.py:304-313)."""
    import numpy as np
    import pytest
    from corruscant_amd import clustering
    pts = np.random.default_rng(0).random((10, 3))
    for bad in ([0, 1, 3] * 3 + [0], [-1, 0] * 5, [1, 2] * 5):
        with pytest.raises(ValueError, match="0 to the number of unique fields"):
            clustering.tree(pts, np.array(bad, dtype=np.int32))
    with pytest.raises(ValueError, match="At least two unique field IDs"):
        clustering.tree(pts, np.zeros(10, dtype=np.int32))
    with pytest.raises(ValueError, match="length equal to the number of points"):
        clustering.tree(pts, np.zeros(9, dtype=np.int32))


def test_tree_file_checks_before_any_device_work(tmp_path):
    """tree.load refuses files that are not tree files, or of another
    version, before it builds anything."""
    from corruscant_amd import clustering
    p = str(tmp_path / "x.npz")
    np.savez(p, points=np.zeros((4, 3)))
    with pytest.raises(ValueError, match="not a tree file"):
        clustering.tree.load(p)
    np.savez(p, format=np.array([99, 1, 4, 4]), points=np.zeros((4, 3)), nodes=np.zeros((8, 40), np.uint8),
             perm=np.zeros(8, np.int32))
    with pytest.raises(ValueError, match="version"):
        clustering.tree.load(p)
    np.savez(p, format=np.array([clustering.TREE_FILE_VERSION, 1, 5, 4]), points=np.zeros((4, 3)),
             nodes=np.zeros((8, 40), np.uint8), perm=np.zeros(8, np.int32))
    with pytest.raises(ValueError, match="points must be"):
        clustering.tree.load(p)
