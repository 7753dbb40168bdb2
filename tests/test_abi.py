"""The C ABI library loads and exports exactly what include/cosmo_hip.h declares."""
import os
import re

import cosmo_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "cosmo_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(cosmo_[a-z0-9_]+)\s*\(", txt))


def test_header_matches_binding_table():
    assert declared() == set(cosmo_amd.SIGNATURES)


def test_library_exports_every_symbol():
    L = cosmo_amd.lib()
    for name in declared():
        assert hasattr(L, name), name


def test_error_path_reports_message(tmp_path):
    import pytest
    with pytest.raises(cosmo_amd.CosmoError) as e:
        cosmo_amd.System.from_xml(str(tmp_path / "missing.xml"))
    assert e.value.code == cosmo_amd.EIO
    assert "missing.xml" in str(e.value)


def test_malformed_upper_rows_rejected_before_any_device_work():
    """cosmo_op_create_upper checks its row offsets on the host before allocating or
    uploading anything (they bound every read of cols / vals): a first offset other than 0
    or a decreasing offset is COSMO_EINVAL, here without a GPU."""
    import numpy as np
    import pytest
    cols = np.array([0, 1, 1, 2], dtype=np.int64)
    vals = np.ones(4)
    for rp in ([1, 2, 3, 4], [0, 3, 2, 4], [0, 2, 5, 4]):
        with pytest.raises(cosmo_amd.CosmoError) as e:
            cosmo_amd.Operator.from_upper(3, np.array(rp, dtype=np.int64), cols, vals)
        assert e.value.code == cosmo_amd.EINVAL
