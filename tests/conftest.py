"""Shared pytest configuration.

Tests marked ``gpu`` need an MI355X (gfx950) and the built HIP library; the
rest run on the CPU: the oracle against the reference's golden vectors, the
host-side logic, and the C-ABI library's exports (no compute calls).
"""

import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def kats():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def fa():
    """The HIP library's Python mirror; GPU tests only (fails loudly without
    the built library or a gfx950 device)."""
    import fullmag_amd
    from fullmag_amd import _abi

    _abi.lib()
    assert fullmag_amd.is_available(), "no gfx950 device visible"
    return fullmag_amd
