// The reference's tests/test_vectors.cu flow written against the drop-in
// header: counter-seeded keys, SHAKE128(le64(2i)) messages, sign with
// exec_threshold 2048, verify; writes the tvecs file to the path in argv[1].
// Built and run by tests/test_gpu_cxx_dropin.py.
#include <cstdio>
#include <cstring>
#include <vector>

#include "cudilithium_amd/api.hpp"

#define MLEN 32

int main(int argc, char **argv) {
    const size_t n = argc > 2 ? (size_t)atoi(argv[2]) : 10000;
    const char *path = argc > 1 ? argv[1] : "tvecs";
    std::vector<uint8_t> pk(n * CRYPTO_PUBLICKEYBYTES), sk(n * CRYPTO_SECRETKEYBYTES), m(n * MLEN),
        sig(n * CRYPTO_BYTES);
    std::vector<int> ret(n);
    // messages: the library's own SHAKE is not exported, so they come from a file
    FILE *fm = fopen((std::string(path) + ".msg").c_str(), "rb");
    if (!fm || fread(m.data(), 1, m.size(), fm) != m.size()) return 2;
    fclose(fm);

    uint8_t *d_kp;
    size_t kp_pitch;
    size_t kp_row = ALIGN_TO_256_BYTES(CRYPTO_PUBLICKEYBYTES) + ALIGN_TO_256_BYTES(CRYPTO_SECRETKEYBYTES) +
                    DILITHIUM_K * DILITHIUM_N * sizeof(int32_t);
    if (hipMallocPitch((void **)&d_kp, &kp_pitch, kp_row, n)) return 3;
    if (crypto_sign_keypair(pk.data(), sk.data(), d_kp, kp_pitch, n)) return 4;
    hipDeviceSynchronize();
    hipFree(d_kp);

    uint8_t *d_sp, *d_tp;
    size_t sp_pitch, tp_pitch, thr = 2048;
    size_t sp_row = ALIGN_TO_256_BYTES(CRYPTO_BYTES) + ALIGN_TO_256_BYTES(CRYPTO_SECRETKEYBYTES) +
                    ALIGN_TO_256_BYTES(SEEDBYTES) + ALIGN_TO_256_BYTES(SEEDBYTES + MLEN) +
                    ALIGN_TO_256_BYTES(SEEDBYTES + CRHBYTES + DILITHIUM_K * POLYW1_PACKEDBYTES) +
                    ALIGN_TO_256_BYTES(CRHBYTES) +
                    (DILITHIUM_K * DILITHIUM_L + 3 * DILITHIUM_L + 4 * DILITHIUM_K + 1) * DILITHIUM_N * 4;
    size_t tp_row = ALIGN_TO_256_BYTES(CRYPTO_BYTES) + ALIGN_TO_256_BYTES(CRHBYTES + DILITHIUM_K * POLYW1_PACKEDBYTES) +
                    (2 * DILITHIUM_L + 3 * DILITHIUM_K) * DILITHIUM_N * 4;
    if (hipMallocPitch((void **)&d_sp, &sp_pitch, sp_row, n)) return 5;
    if (hipMallocPitch((void **)&d_tp, &tp_pitch, tp_row, thr)) return 6;
    task_lut lut{};
    lut.sign_lut.resize(n);
    hipHostMalloc((void **)&lut.h_exec_lut, sizeof(uint32_t) * thr);
    hipHostMalloc((void **)&lut.h_done_lut, thr);
    hipHostMalloc((void **)&lut.h_copy_lut, sizeof(copy_lut_element) * thr);
    hipMalloc((void **)&lut.d_exec_lut, sizeof(uint32_t) * thr);
    hipMalloc((void **)&lut.d_done_lut, thr);
    hipMalloc((void **)&lut.d_copy_lut, sizeof(copy_lut_element) * thr);
    size_t siglen = 0;
    if (crypto_sign_signature(sig.data(), CRYPTO_BYTES, &siglen, m.data(), MLEN, MLEN, sk.data(), d_sp, sp_pitch,
                              d_tp, tp_pitch, lut, thr, n))
        return 7;
    hipDeviceSynchronize();

    uint8_t *d_vp;
    size_t vp_pitch;
    size_t vp_row = ALIGN_TO_256_BYTES(CRYPTO_BYTES) + ALIGN_TO_256_BYTES(CRYPTO_PUBLICKEYBYTES) +
                    ALIGN_TO_256_BYTES(SEEDBYTES + MLEN) +
                    ALIGN_TO_256_BYTES(CRHBYTES + DILITHIUM_K * POLYW1_PACKEDBYTES) + ALIGN_TO_256_BYTES(SEEDBYTES) +
                    (DILITHIUM_K * DILITHIUM_L + DILITHIUM_L + 3 * DILITHIUM_K + 1) * DILITHIUM_N * 4 + sizeof(int);
    if (hipMallocPitch((void **)&d_vp, &vp_pitch, vp_row, n)) return 8;
    crypto_sign_verify(ret.data(), sig.data(), CRYPTO_BYTES, siglen, m.data(), MLEN, MLEN, pk.data(), d_vp, vp_pitch,
                       n);
    hipDeviceSynchronize();
    FILE *fo = fopen(path, "wb");
    fwrite(pk.data(), 1, pk.size(), fo);
    fwrite(sk.data(), 1, sk.size(), fo);
    fwrite(sig.data(), 1, sig.size(), fo);
    fwrite(ret.data(), sizeof(int), n, fo);
    fclose(fo);
    printf("siglen %zu\n", siglen);
    return 0;
}
