"""The C ABI library loads and exports every entry point include/niq.h
declares (CPU: no compute calls), and the host-side drop-in logic works
without a GPU."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "niq.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(niq_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_hot_path():
    syms = declared_symbols()
    for s in ("niq_model_create", "niq_classify_boxes", "niq_tree_build", "niq_cast_rays",
              "niq_eval_points", "niq_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    import ctypes
    from niq import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libniq.so is not built (run neural-implicit-queries_amd/build.py)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= set(declared_symbols())


def test_lowering_and_validation_on_host():
    from niq import mlp, models, kd_tree, implicit_mlp_utils
    p = models.config_model("c1_relu_3x32")
    layers = mlp.lower(p)
    assert mlp.dims_of(layers) == [3, 32, 32, 32, 1]
    with pytest.raises(ValueError):
        mlp.lower({"0000.dense.A": np.zeros((3, 2), np.float32), "0001.squeeze_last._": np.zeros(0)})
    with pytest.raises(RuntimeError):
        implicit_mlp_utils.implicit_from_params(p, "bogus")
    f, _ = implicit_mlp_utils.implicit_from_params(p, "affine_fixed")
    with pytest.raises(ValueError):
        kd_tree.construct_uniform_unknown_levelset_tree(f, p, -np.ones(3), np.ones(3))
    with pytest.raises(ValueError):
        kd_tree.construct_uniform_unknown_levelset_tree(f, p, -np.ones(3), np.ones(3), split_depth=3,
                                                        batch_process_size=3000)


def test_models_are_seeded_and_median_shifted():
    from niq import models
    from oracle import mlp as omlp
    a = models.random_mlp([16, 16], seed=4)
    b = models.random_mlp([16, 16], seed=4)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    v = omlp.eval_points(a, models.median_grid(), np.float64)
    assert abs(np.median(v)) < 1e-6


def test_kernel_mc_table_matches_oracle_table():
    """The C library's marching-cubes table equals the oracle's (both are
    Bourke's public table in independent encodings)."""
    import ctypes
    from niq import _lib
    from oracle import mc as omc
    lib = _lib.load_library()
    buf = (ctypes.c_byte * (256 * 16))()
    assert lib.niq_mc_table(buf) == 0
    got = np.frombuffer(bytes(buf), dtype=np.int8).reshape(256, 16)[:, :15].reshape(256, 5, 3)
    np.testing.assert_array_equal(got, omc.tri_table())
