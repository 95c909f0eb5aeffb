"""The C-ABI library loads and exports every entry point its headers declare
(CPU only: no call here reaches the GPU)."""
import ctypes
import glob
import math
import os
import re

import pytest

from kinc_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(kinc_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "kinc_sim_create" in names and "kinc_sim_execute" in names
    assert len(names) >= 10


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name, native):
    lib = ctypes.CDLL(_native.LIB_PATH)
    assert hasattr(lib, name), name


def test_binding_covers_headers():
    from kinc_amd import formats, rmt, serving
    bound = ({n for n, _, _ in _native.SIGNATURES} | set(formats.SIGNATURES)
             | set(serving.SIGNATURES) | set(rmt.SIGNATURES))
    missing = set(declared_functions()) - bound
    assert not missing, missing


def test_default_params_match_reference(native):
    # src/core/similarity.h:55-71
    p = native.default_params()
    assert p.clus_method == 0 and p.corr_method == 0
    assert math.isinf(p.min_expr) and p.min_expr < 0
    assert p.min_samples == 30 and p.min_clusters == 1 and p.max_clusters == 5
    assert p.criterion == native.CRIT["ICL"]
    assert p.pre_out == 0 and p.post_out == 0
    assert abs(p.min_corr - 0.5) < 1e-7 and p.max_corr == 1.0 and p.ksize == 4096


def test_pair_index_abi(native):
    assert native.pair_count(20000) == 199_990_000
    assert native.pair_index(0) == (1, 0)
    assert native.pair_index(3) == (3, 0)
    assert native.pair_index(199_989_999) == (19999, 19998)


@pytest.mark.parametrize("bad,code", [
    (dict(minclus=3, maxclus=2), _native.KINC_E_INVALID),
    (dict(minsamp=0), _native.KINC_E_INVALID),
    (dict(maxclus=65), _native.KINC_E_INVALID),
])
def test_create_rejects_bad_params_before_device(native, bad, code):
    import numpy as np
    p = native.make_params(**bad)
    with pytest.raises(native.KincError) as e:
        native.Engine(np.zeros((4, 8), np.float32), p)
    assert e.value.code == code


def test_create_rejects_more_samples_than_the_device_takes(native):
    import numpy as np
    with pytest.raises(native.KincError) as e:
        native.Engine(np.zeros((2, 32769), np.float32), native.make_params())
    assert e.value.code == native.KINC_E_UNSUPPORTED


@pytest.mark.parametrize("field,value", [("pearson_fp64", 2), ("pearson_fp64", -1),
                                         ("batch_capacity", -1), ("batch_capacity", (1 << 24) + 1)])
def test_create_rejects_bad_extension_fields(native, field, value):
    # the fields the reference does not have are range-checked too
    import numpy as np
    p = native.make_params()
    setattr(p, field, value)
    with pytest.raises(native.KincError) as e:
        native.Engine(np.zeros((4, 8), np.float32), p)
    assert e.value.code == _native.KINC_E_INVALID


def test_create_rejects_nonzero_reserved(native):
    import numpy as np
    p = native.make_params()
    p.reserved[1] = 7
    with pytest.raises(native.KincError) as e:
        native.Engine(np.zeros((4, 8), np.float32), p)
    assert e.value.code == _native.KINC_E_INVALID


def test_rmt_default_params_match_reference(native):
    # src/core/rmt.h:31-40
    from kinc_amd import rmt
    p = rmt.make_params()
    assert abs(p.threshold_start - 0.99) < 1e-7 and abs(p.threshold_step - 0.001) < 1e-9
    assert p.threshold_stop == 0.5 and abs(p.chi_square1 - 99.607) < 1e-4
    assert p.chi_square2 == 200 and p.min_eigenvalues == 50
    assert (p.min_pace, p.max_pace, p.bins) == (10, 40, 60)


def test_group_autotune_rejects_bad_groups_before_device(native):
    # no handles, too many, a null handle: refused without touching a device
    fn = native.lib().kinc_sim_autotune_group
    null2 = (native._P * 2)(None, None)
    assert fn(None, 1, 0, 1, None) == native.KINC_E_INVALID
    assert fn(ctypes.cast(null2, native._P), 0, 0, 1, None) == native.KINC_E_INVALID
    assert fn(ctypes.cast(null2, native._P), 17, 0, 1, None) == native.KINC_E_INVALID
    assert fn(ctypes.cast(null2, native._P), 2, 0, 1, None) == native.KINC_E_INVALID
