import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)


_ensure_oracle()


@pytest.fixture(scope="session")
def oracle():
    from oracle import bind
    return bind


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_vectors.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def bv_opencl_n45(oracle):
    kn = oracle.knobs(45, 10, semantics=oracle.SEM_OPENCL)
    return kn, oracle.precompute_bv(kn), oracle.mod9_rows(35, 10)


@pytest.fixture(scope="session")
def bv_cuda_n45(oracle):
    kn = oracle.knobs(45, 10, semantics=oracle.SEM_CXX)
    return kn, oracle.precompute_bv(kn), oracle.mod9_rows(35, 10)


def read_cases(name):
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return [[int(x) for x in line.split()] for line in f if line.strip()]
