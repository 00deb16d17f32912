# SPDX-License-Identifier: Apache-2.0
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libnph_ref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def ref_api():
    """The reference's own step, compiled from its sources (oracle/_ref)."""
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    from nepheline_amd import physics as P
    return P.api(REF_LIB, "ref_shim_allocator")


@pytest.fixture(scope="session")
def gpu_api():
    """This package's HIP world step. Fails loudly when the library is missing."""
    from nepheline_amd import physics as P
    return P.product_api()


@pytest.fixture(scope="session")
def hull_bank():
    from nepheline_amd import scenes as S
    return S.load_hull_bank(os.path.join(GOLDEN, "hulls_12v.bin"))
