// Randomized signing through the C++ drop-in header, as a caller of the
// reference builds it with -DDILITHIUM_RANDOMIZED_SIGNING (api.cu:143-148):
// the random source is a caller function (dil_set_randombytes) that serves
// the bytes of a file, so the test can recompute every signature.
//   rnd_dropin_<mode> <dir> <n> <mlen> <exec_threshold>
// reads <dir>/sk, <dir>/msg, <dir>/rnd and writes <dir>/sig.
// Built and run by tests/test_gpu_cxx_dropin.py.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cudilithium_amd/api.hpp"

static std::vector<uint8_t> g_rnd;
static size_t g_pos = 0;

static int file_randombytes(uint8_t *out, size_t outlen) {
    if (g_pos + outlen > g_rnd.size()) return -100;
    for (size_t i = 0; i < outlen; ++i) out[i] = g_rnd[g_pos + i];
    g_pos += outlen;
    return 0;
}

static bool read_file(const std::string &path, std::vector<uint8_t> &v) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    const bool ok = fread(v.data(), 1, v.size(), f) == v.size();
    fclose(f);
    return ok;
}

int main(int argc, char **argv) {
    if (argc < 5) return 1;
    const std::string dir = argv[1];
    const size_t n = (size_t)atol(argv[2]), mlen = (size_t)atol(argv[3]), thr = (size_t)atol(argv[4]);
    std::vector<uint8_t> sk(n * CRYPTO_SECRETKEYBYTES), m(n * mlen + 1), sig(n * CRYPTO_BYTES);
    g_rnd.resize(n * CRHBYTES);
    m.resize(n * mlen);
    if (!read_file(dir + "/sk", sk) || !read_file(dir + "/msg", m) || !read_file(dir + "/rnd", g_rnd)) return 2;
    dil_set_randombytes(file_randombytes);

    uint8_t *d_sp, *d_tp;
    size_t sp_pitch, tp_pitch;
    size_t sp_row = ALIGN_TO_256_BYTES(CRYPTO_BYTES) + ALIGN_TO_256_BYTES(CRYPTO_SECRETKEYBYTES) +
                    ALIGN_TO_256_BYTES(SEEDBYTES) + ALIGN_TO_256_BYTES(SEEDBYTES + mlen) +
                    ALIGN_TO_256_BYTES(SEEDBYTES + CRHBYTES + DILITHIUM_K * POLYW1_PACKEDBYTES) +
                    ALIGN_TO_256_BYTES(CRHBYTES) +
                    (DILITHIUM_K * DILITHIUM_L + 3 * DILITHIUM_L + 4 * DILITHIUM_K + 1) * DILITHIUM_N * 4;
    size_t tp_row = ALIGN_TO_256_BYTES(CRYPTO_BYTES) + ALIGN_TO_256_BYTES(CRHBYTES + DILITHIUM_K * POLYW1_PACKEDBYTES) +
                    (2 * DILITHIUM_L + 3 * DILITHIUM_K) * DILITHIUM_N * 4;
    if (hipMallocPitch((void **)&d_sp, &sp_pitch, sp_row, n)) return 3;
    if (hipMallocPitch((void **)&d_tp, &tp_pitch, tp_row, thr)) return 4;
    task_lut lut{};
    lut.sign_lut.resize(n);
    hipHostMalloc((void **)&lut.h_exec_lut, sizeof(uint32_t) * thr);
    hipHostMalloc((void **)&lut.h_done_lut, thr);
    hipHostMalloc((void **)&lut.h_copy_lut, sizeof(copy_lut_element) * thr);
    hipMalloc((void **)&lut.d_exec_lut, sizeof(uint32_t) * thr);
    hipMalloc((void **)&lut.d_done_lut, thr);
    hipMalloc((void **)&lut.d_copy_lut, sizeof(copy_lut_element) * thr);
    size_t siglen = 0;
    const int r = crypto_sign_signature(sig.data(), CRYPTO_BYTES, &siglen, m.data(), mlen, mlen, sk.data(), d_sp,
                                        sp_pitch, d_tp, tp_pitch, lut, thr, n);
    if (r) {
        fprintf(stderr, "crypto_sign_signature: %d\n", r);
        return 5;
    }
    hipDeviceSynchronize();
    if (g_pos != n * CRHBYTES) return 6;  // one draw of 64 bytes per item
    // a second call with the source exhausted reports its error and signs nothing
    if (crypto_sign_signature(sig.data(), CRYPTO_BYTES, &siglen, m.data(), mlen, mlen, sk.data(), d_sp, sp_pitch,
                              d_tp, tp_pitch, lut, thr, n) != -100)
        return 7;
    FILE *fo = fopen((dir + "/sig").c_str(), "wb");
    if (!fo) return 8;
    fwrite(sig.data(), 1, sig.size(), fo);
    fclose(fo);
    printf("siglen %zu\n", siglen);
    return 0;
}
