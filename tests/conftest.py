import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # The oracle is plain C; build it if this checkout has not yet.
    if not (ROOT / "oracle" / "liblcoracle.so").exists():
        os.system(f"make -s -C {ROOT / 'oracle'}")


def _has_gpu() -> bool:
    try:
        from linkcell_amd import _lib

        return bool(_lib.load().lc_gpu_available())
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Device handle for GPU tests: fails loudly if the HIP library or the
    device is missing (no CPU fallback is acceptable in a GPU test)."""
    import torch

    from linkcell_amd import _lib

    lib = _lib.load()
    assert lib.lc_gpu_available() == 1, "no HIP device visible to liblinkcell_amd"
    assert torch.cuda.is_available()
    return torch.device("cuda:0")
