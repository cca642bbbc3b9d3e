import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP)")


def _ensure_built():
    # Built in-tree before the tests; only a missing library is built here (CPU-side, no GPU needed).
    from oracle import build as ob
    if not os.path.exists(ob.ORACLE_LIB):
        ob.build_oracle()
    if not os.path.exists(ob.REF_LIB):
        ob.build_ref()
    from glslppm_amd import build as pb
    if not os.path.exists(pb.LIB):
        pb.build()


_ensure_built()


@pytest.fixture(scope="session")
def reference_root():
    root = os.environ.get("PPM_REFERENCE_ROOT", "/root/reference")
    if not os.path.exists(os.path.join(root, "cornell_metal.obj")):
        pytest.skip("reference checkout not present on this machine")
    return root
