"""CPU checks of the C ABI: the library loads, exports every symbol that
include/fasttree_c.h declares, and rejects bad arguments with the reference's
error kinds before touching a GPU."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "fasttree_c.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ft_[a-z0-9_]+)\s*\(", src)))


def test_header_and_python_export_lists_agree(ft):
    assert declared_symbols() == sorted(ft.EXPORTS)


def test_library_exports_every_declared_symbol(ft):
    L = ft.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing


def test_version_string(ft):
    assert "gfx950" in ft.version()


def _fake_tree(ft, n, ct=3):
    t = ft._Tree()
    t.num_leaves = n
    t.num_internal = max(n - 1, 0)
    t.coord_type = ct
    t.config = ft.Config()
    return t


def test_k_above_max_k_is_invalid_argument(ft):
    # self_knn_query<_MAX_K_=32> throws std::invalid_argument for k > 32
    # (src/hlbvh.hpp:1940-1944); the check precedes any device work.
    t = _fake_tree(ft, 10)
    o = ft.QueryOpts()
    rc = ft.lib().ft_self_knn_query(None, C.byref(t), 33, None, None, 1, C.byref(o))
    assert rc == ft.FT_EINVAL
    assert "_MAX_K_" in ft.lib().ft_last_error().decode()
    o.max_k = 64
    rc = ft.lib().ft_self_knn_query_subset(None, C.byref(t), None, 0, 65, None, None, 1,
                                           C.byref(o))
    assert rc == ft.FT_OK  # num_ids == 0 returns before the check, as the reference


def test_empty_inputs_are_noops(ft):
    t = _fake_tree(ft, 0)
    o = ft.QueryOpts()
    assert ft.lib().ft_self_knn_query(None, C.byref(t), 4, None, None, 1, C.byref(o)) == 0
    assert ft.lib().ft_knn_query(None, C.byref(t), None, None, None, 4, 10, None, None,
                                 C.byref(o)) == 0
    assert ft.lib().ft_range_query(None, C.byref(t), None, None, None, None, None, 5, None,
                                   None, 3) == 0
    assert ft.lib().ft_build_bvh(None, C.byref(t), None, None, None, None, None, None, 0, None,
                                 None) == 0


def test_oversize_tree_is_runtime_error(ft):
    t = ft._Tree()
    cfg = ft.Config()
    rc = ft.lib().ft_tree_alloc(C.byref(t), (1 << 31), 3, C.byref(cfg), None)
    assert rc == ft.FT_ERANGE
    assert "INT_MAX / 2" in ft.lib().ft_last_error().decode()


def test_bad_config_rejected(ft):
    t = ft._Tree()
    for cfg in (ft.Config(bits_per_dim=22), ft.Config(bits_per_dim=0), ft.Config(sfc=7),
                ft.Config(periodic=True, box_length=0.0)):
        ct = 1 if cfg.periodic else 3
        assert ft.lib().ft_tree_alloc(C.byref(t), 10, ct, C.byref(cfg), None) == ft.FT_EINVAL


def test_python_wrapper_maps_errors(ft):
    with pytest.raises(ValueError):
        ft._check(ft.FT_EINVAL)
    with pytest.raises(RuntimeError):
        ft._check(ft.FT_ELDS)
