"""The reference's test_vectors.cu flow compiled against the C++ drop-in header
(include/cudilithium_amd/api.hpp) with -DDILITHIUM_MODE: the tvecs SHA-256 must
be the reference's pin (tests/CMakeLists.txt:21-23)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests import kat

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx", "bin")


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_cxx_dropin_reproduces_reference_tvecs(mode, tmp_path):
    exe = os.path.join(BIN, f"kat_dropin_{mode}")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built; run __graft_entry__.build()")
    n = kat.NVECTORS
    m = kat.messages()
    out = str(tmp_path / f"tvecs{mode}")
    m.tofile(out + ".msg")
    r = subprocess.run([exe, out, str(n)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    p = oracle.params(mode)
    raw = np.fromfile(out, dtype=np.uint8)
    o = 0
    pk = raw[o:o + n * p.pk_bytes].reshape(n, -1); o += n * p.pk_bytes
    sk = raw[o:o + n * p.sk_bytes].reshape(n, -1); o += n * p.sk_bytes
    sig = raw[o:o + n * p.sig_bytes].reshape(n, -1); o += n * p.sig_bytes
    ret = raw[o:].view(np.int32)
    assert (ret == 0).all()
    assert kat.tvecs_sha256(m, pk, sk, sig, ret) == kat.PINS[mode]
