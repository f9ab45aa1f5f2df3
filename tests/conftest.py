import os
import sys
import warnings

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# The reference library is built with -Ofast, whose startup code turns on FTZ/DAZ in the
# loading process; numpy then warns about the smallest subnormal. Harmless here.
warnings.filterwarnings("ignore", message="The value of the smallest subnormal")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref/libwema_ref.so (compiled reference)")


@pytest.fixture(scope="session")
def has_gpu():
    import torch
    return torch.cuda.is_available()


def pytest_collection_modifyitems(config, items):
    ref_lib = os.path.join(ROOT, "oracle", "_ref", "libwema_ref.so")
    skip_ref = pytest.mark.skip(reason="oracle/_ref/libwema_ref.so not built")
    for it in items:
        if "ref" in it.keywords and not os.path.exists(ref_lib):
            it.add_marker(skip_ref)
