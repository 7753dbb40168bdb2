import json
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
GOLDEN_CASES = ["sd2p2n_s11", "sd3p1n_m1_s7", "fp2p2n_s5", "fp3p2n_m1_s3", "sd4p4n_s0"]
FULL_CASES = ["sd2p2n_s11", "sd3p1n_m1_s7", "fp2p2n_s5"]
# truncated bases (hw window, hw cut with parity -, an hw window above its minimum, the
# hwmax attribute on odd A,
# occupation rejections in sd and fp); fixtures from tests/golden/make_golden.py trunc
TRUNC_CASES = ["psd2p2n_hw2", "psd2p2n_neg_hw1", "psd2p2n_hw22", "psd3p2n_m1_hwmax2", "sd3p3n_rej",
               "fp3p2n_m1_rej"]


def trunc_case(name):
    """(xml path, meta, arrays) of a truncated-basis fixture."""
    meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
    arrays = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    return os.path.join(GOLDEN, name + ".xml"), meta, arrays


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


class Golden:
    def __init__(self, name, tmpdir):
        self.name = name
        self.meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
        self.arrays = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
        for ext in (".xml", ".int"):
            shutil.copy(os.path.join(GOLDEN, name + ext), os.path.join(tmpdir, name + ext))
        self.xml = os.path.join(tmpdir, name + ".xml")
        self.dir = tmpdir


@pytest.fixture(scope="session")
def golden(tmp_path_factory):
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = Golden(name, str(tmp_path_factory.mktemp(name)))
        return cache[name]

    return get


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import cosmo_amd
    cosmo_amd.lib()
    return cosmo_amd
