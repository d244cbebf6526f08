/*
 * Drives the host-side paths of this repository's library that run without a GPU -- the C17
 * sources under orkan_amd/csrc, built with AddressSanitizer and UndefinedBehaviorSanitizer and
 * linked with the (uninstrumented) kernel objects: kraus_to_superop for 1- to 3-qubit Kraus sets,
 * state_len for every layout and size, and orkan_state_load on every kind of malformed state file
 * (all rejected before any device call, including headers whose byte counts would wrap 64 bits).
 * Exit status 0 when every result was as expected.
 */
#include "orkan.h"
#include "orkan_amd.h"

#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int fails = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            fprintf(stderr, "san_host: check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                      \
        }                                                                 \
    } while (0)

static void header(unsigned char *h, const char *magic, uint32_t ver, uint32_t type, uint32_t qubits, uint32_t lt,
                   uint64_t length, uint32_t eb, uint32_t bom) {
    memset(h, 0, 64);
    memcpy(h, magic, 8);
    memcpy(h + 8, &ver, 4);
    memcpy(h + 12, &type, 4);
    memcpy(h + 16, &qubits, 4);
    memcpy(h + 20, &lt, 4);
    memcpy(h + 24, &length, 8);
    memcpy(h + 32, &eb, 4);
    memcpy(h + 36, &bom, 4);
}

static int load_blob(const char *dir, const char *name, const unsigned char *h, size_t payload) {
    char path[512];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    FILE *f = fopen(path, "wb");
    if (!f) return 999;
    fwrite(h, 1, 64, f);
    static const unsigned char zeros[256];
    while (payload) {
        const size_t k = payload < sizeof zeros ? payload : sizeof zeros;
        fwrite(zeros, 1, k, f);
        payload -= k;
    }
    fclose(f);
    state_t st = {.type = PURE, .data = NULL, .qubits = 0};
    const int rc = orkan_state_load(&st, path);
    CHECK(st.data == NULL && st.qubits == 0);
    return rc;
}

int main(int argc, char **argv) {
    const char *dir = argc > 1 ? argv[1] : ".";
    /* kraus_to_superop: S[(i,j),(k,l)] = sum_a K_a[i,k] conj(K_a[j,l]) */
    for (int nq = 1; nq <= 3; ++nq) {
        const int N = 1 << nq, M = N * N, terms = 3;
        cplx_t *K = malloc(sizeof(cplx_t) * terms * M);
        for (int i = 0; i < terms * M; ++i) K[i] = sin(1.0 + i) + I * cos(2.0 * i);
        kraus_t kr = {.n_qubits = (qubit_t)nq, .n_terms = (uint64_t)terms, .data = K};
        superop_t so = kraus_to_superop(&kr);
        CHECK(so.n_qubits == nq && so.data != NULL);
        double worst = 0.0;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < N; ++k)
                    for (int l = 0; l < N; ++l) {
                        cplx_t want = 0;
                        for (int a = 0; a < terms; ++a) want += K[a * M + i * N + k] * conj(K[a * M + j * N + l]);
                        const double d = cabs(so.data[(i * N + j) * M + k * N + l] - want);
                        worst = d > worst ? d : worst;
                    }
        CHECK(worst <= 1e-12);
        free(so.data);
        free(K);
    }
    /* state_len */
    for (qubit_t n = 0; n <= 20; ++n) {
        const idx_t D = (idx_t)1 << n, nt = (D + TILE_DIM - 1) / TILE_DIM;
        state_t a = {.type = PURE, .data = NULL, .qubits = n}, b = {.type = MIXED_PACKED, .data = NULL, .qubits = n},
                c = {.type = MIXED_TILED, .data = NULL, .qubits = n};
        CHECK(state_len(&a) == D);
        CHECK(state_len(&b) == D * (D + 1) / 2);
        CHECK(state_len(&c) == nt * (nt + 1) / 2 * TILE_SIZE);
    }
    /* malformed state files: QS_ERR_FORMAT, nothing allocated */
    unsigned char h[64];
    const uint32_t BOM = 0x01020304u;
    header(h, "NOTORKAN", 1, 0, 3, LOG_TILE_DIM, 8, 16, BOM);
    CHECK(load_blob(dir, "magic.bin", h, 128) == QS_ERR_FORMAT);
    header(h, "ORKAN-ST", 2, 0, 3, LOG_TILE_DIM, 8, 16, BOM);
    CHECK(load_blob(dir, "version.bin", h, 128) == QS_ERR_FORMAT);
    header(h, "ORKAN-ST", 1, 0, 3, LOG_TILE_DIM, 9, 16, BOM);
    CHECK(load_blob(dir, "length.bin", h, 144) == QS_ERR_FORMAT);
    header(h, "ORKAN-ST", 1, 0, 3, LOG_TILE_DIM, 8, 16, BOM);
    CHECK(load_blob(dir, "short.bin", h, 100) == QS_ERR_FORMAT);
    header(h, "ORKAN-ST", 1, 0, 60, LOG_TILE_DIM, (uint64_t)1 << 60, 16, BOM);
    CHECK(load_blob(dir, "wrap_pure.bin", h, 0) == QS_ERR_FORMAT);
    header(h, "ORKAN-ST", 1, 1, 62, LOG_TILE_DIM, ((uint64_t)1 << 61) * (((uint64_t)1 << 62) + 1), 16, BOM);
    CHECK(load_blob(dir, "wrap_packed.bin", h, 0) == QS_ERR_FORMAT);
    header(h, "ORKAN-ST", 1, 2, 63, LOG_TILE_DIM, 0, 16, BOM);
    CHECK(load_blob(dir, "wrap_tiled.bin", h, 0) == QS_ERR_FORMAT);
    header(h, "ORKAN-ST", 1, 0x100, 61, LOG_TILE_DIM, (uint64_t)1 << 61, 8, BOM);
    CHECK(load_blob(dir, "wrap_fp32.bin", h, 0) == QS_ERR_FORMAT);
    header(h, "ORKAN-ST", 1, 7, 3, LOG_TILE_DIM, 8, 16, BOM);
    CHECK(load_blob(dir, "type.bin", h, 128) == QS_ERR_FORMAT);
    state_t none = {.type = PURE, .data = NULL, .qubits = 0};
    char missing[512];
    snprintf(missing, sizeof missing, "%s/does-not-exist.bin", dir);
    CHECK(orkan_state_load(&none, missing) == QS_ERR_FILE);
    CHECK(orkan_state_load(NULL, missing) == QS_ERR_NULL);
    printf("san_host %s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}
