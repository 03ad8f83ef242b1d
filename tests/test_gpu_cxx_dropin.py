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


@pytest.mark.parametrize("mode", [2, 3, 5])
def test_cxx_dropin_randomized_signing_with_caller_source(mode, tmp_path):
    """api.hpp built with -DDILITHIUM_RANDOMIZED_SIGNING: rhoprime for item i
    is bytes [64 i, 64 i + 64) of the caller's random source
    (dil_set_randombytes), and each signature equals the oracle's randomized
    signature with that rhoprime; the exhausted source's error is returned."""
    exe = os.path.join(BIN, f"rnd_dropin_{mode}")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built; run __graft_entry__.build()")
    n, mlen, thr = 300, 57, 64
    pk, sk = oracle.keypair_counter_batch(mode, n, rand_index=600 + mode)
    rng = np.random.default_rng(mode + 40)
    m = rng.integers(0, 256, (n, mlen), dtype=np.uint8)
    rnd = rng.integers(0, 256, (n, 64), dtype=np.uint8)
    sk.tofile(str(tmp_path / "sk"))
    m.tofile(str(tmp_path / "msg"))
    rnd.tofile(str(tmp_path / "rnd"))
    r = subprocess.run([exe, str(tmp_path), str(n), str(mlen), str(thr)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    sig = np.fromfile(str(tmp_path / "sig"), dtype=np.uint8).reshape(n, -1)
    exp, _ = oracle.sign_rnd_batch(mode, m, sk, rnd)
    assert np.array_equal(sig, exp)
    assert (oracle.verify_batch(mode, sig, m, pk) == 0).all()
    det, _ = oracle.sign_batch(mode, m, sk)
    assert not np.array_equal(sig, det)
