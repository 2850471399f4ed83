"""The C-ABI library loads on the CPU host and exports every declared symbol."""
import ctypes

from voxel_amd import _lib


def test_library_exports_every_declared_entry_point():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _lib.header_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_declared_entry_point_has_a_python_signature():
    import voxel_amd.dag  # noqa: F401
    import voxel_amd.hash  # noqa: F401
    import voxel_amd.primitives  # noqa: F401
    import voxel_amd.svo  # noqa: F401
    names = set(_lib.header_functions())
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))


def test_version_string():
    assert b"gfx950" in _lib.lib().vx_version()


def test_workspace_queries_need_no_gpu():
    import voxel_amd.primitives as P
    assert P.required_workspace_size(0) == 0
    assert P.required_workspace_size(1000) == 0  # one tile: no workspace (scan.cuh semantics)
    assert P.required_workspace_size(1 << 20) > 0
    assert P.required_sort_workspace_size(1, 4) == 0  # n <= 1 needs none
    assert P.required_sort_workspace_size(1 << 20, 8, 42) > (1 << 20) * 12
