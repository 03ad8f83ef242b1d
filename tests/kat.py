"""Known-answer-test file format of the reference (tests/test_vectors.cu:177-209)
and its SHA-256 pins (tests/CMakeLists.txt:21-23).

Message i is SHAKE128(le64(2i))[0:32] and keypair i uses the counter seed
SHAKE128(le64(2i+1))[0:32] (tests/test_vectors.cu:30-38,58-60;
src/keypair.cu:50-60), i.e. rand_index = 0.
"""
import hashlib

import numpy as np

NVECTORS = 10000
MLEN = 32
PINS = {
    2: "a0835b0f41084881409e42866012996d2bf72fb8694535f190b2a2bfaf271660",
    3: "98c70f95be840482fd656b264a559fa770c0e693629d019504b42c8229bab123",
    5: "d702f8ae15b5d9f0234120a3a073e617f7aebc9d52df0beae7bb47e14a17c1d3",
}


def messages(n=NVECTORS):
    return np.stack([np.frombuffer(hashlib.shake_128(int(2 * i).to_bytes(8, "little")).digest(MLEN),
                                   np.uint8) for i in range(n)])


def tvecs_text(m, pk, sk, sig, ret):
    out = []
    for i in range(len(ret)):
        if ret[i]:
            out.append("Signature verification failed! %u\n" % i)
    h = lambda b: hashlib.shake_256(b.tobytes()).digest(32).hex()
    for i in range(len(ret)):
        out.append("count = %u\n" % i)
        out.append("m = %s\n" % m[i].tobytes().hex())
        out.append("pk = %s\n" % h(pk[i]))
        out.append("sk = %s\n" % h(sk[i]))
        out.append("sig = %s\n" % h(sig[i]))
    return "".join(out)


def tvecs_sha256(m, pk, sk, sig, ret):
    return hashlib.sha256(tvecs_text(m, pk, sk, sig, ret).encode()).hexdigest()
