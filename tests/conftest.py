"""Shared test setup: markers, package/oracle loading, fixture helpers."""
from __future__ import annotations

import base64
import importlib
import re
import sys
import zlib
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


def _ensure_built():
    build = importlib.import_module("test-_amd.build")
    if not build.ORACLE_LIB.exists():
        build.build_oracle()
    if not build.LIB.exists():
        build.build_library()


_ensure_built()


@pytest.fixture(scope="session")
def femrt():
    return importlib.import_module("test-_amd")


@pytest.fixture(scope="session")
def oracle():
    return importlib.import_module("oracle.oracle")


@pytest.fixture(scope="session")
def gpu_ctx(femrt):
    ctx = femrt.Context(0)
    yield ctx
    ctx.close()


def decode_fld(path):
    """FLD payload -> float64 array (base64 + zlib little-endian doubles, or ASCII)."""
    text = Path(path).read_text()
    m = re.search(r"<ELEMENTS[^>]*>(.*?)</ELEMENTS>", text, re.S)
    payload = m.group(1).strip()
    try:
        return np.array([float(v) for v in payload.split()])
    except ValueError:
        raw = zlib.decompress(base64.b64decode(payload))
        return np.frombuffer(raw, "<f8").copy()


def hex_curved_face_geom(n=5):
    """Exact modal coefficients of the Hex_CurvedFace map
    x=(r+1)/2, y=(s+1)/2, z=(t+1)/2 - 0.3(1-r^2)(1-s^2)(1-t)/2
    in the Modified_A basis: (r+1)/2 = phi_1, 1 = phi_0 + phi_1,
    (1-r^2) = 4 phi_2, (1-t)/2 = phi_0.  Mode (i,j,k) at (k*n+j)*n+i."""
    g = np.zeros((3, n, n, n), np.float32)  # [comp][k][j][i]
    for a in (0, 1):
        for b in (0, 1):
            g[0, b, a, 1] = 1.0   # x: i=1, j=a, k=b
            g[1, b, 1, a] = 1.0   # y: j=1, i=a, k=b
            g[2, 1, b, a] = 1.0   # z: k=1, i=a, j=b
    g[2, 0, 2, 2] = -4.8
    return g.reshape(3, n ** 3)


# amplitude the float32 coefficient -4.8 actually encodes (4.8 is not exact in
# binary): z bubble = -(A/16)(1-r^2)(1-s^2)(1-t)/2 with A = float32(4.8)
HCF_AMP = float(np.float32(4.8)) / 16.0


def hex_curved_face_map(r, s, t, amp=HCF_AMP):
    return np.array([(r + 1) / 2, (s + 1) / 2,
                     (t + 1) / 2 - amp * (1 - r * r) * (1 - s * s) * (1 - t) / 2])
