// Mutation fuzz of the loaders that read untrusted files: the PNG and JPEG
// decoders and the glTF / GLB loader.  Built by tests/test_fuzz.py with g++
// under AddressSanitizer and UBSan (host code only); any memory error or
// undefined behaviour aborts the run.
//
//   fuzz_loaders image ITERS SEED_FILE...          (.png or .jpg seeds)
//   fuzz_loaders gltf ITERS SCRATCH_PATH SEED_FILE...
//
// Each iteration applies 1-8 random edits to a seed (bit flips, random or
// JSON-syntax bytes, truncation, insertion, 0xFF runs) and loads the result;
// loading must either succeed with a consistent size or fail with an error.
#include "png.h"
#include "swraster.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace sr {
// the library's error reporter lives with the HIP runtime code; the loaders
// only need its return value here
int fail(int code, const char *, ...) { return code; }
}  // namespace sr

static std::vector<uint8_t> read_all(const char *path) {
    std::vector<uint8_t> v;
    FILE *f = fopen(path, "rb");
    if (!f) return v;
    int c;
    while ((c = fgetc(f)) != EOF) v.push_back((uint8_t)c);
    fclose(f);
    return v;
}

static void mutate(std::vector<uint8_t> &d, std::mt19937_64 &rng, bool json) {
    static const char SYNTAX[] = "{}[],:\"0123456789-.eE";
    int n = 1 + (int)(rng() % 8);
    for (int m = 0; m < n; m++) {
        size_t pos = d.empty() ? 0 : rng() % d.size();
        switch (rng() % 5) {
            case 0: if (!d.empty()) d[pos] ^= (uint8_t)(1u << (rng() % 8)); break;
            case 1: if (!d.empty()) d[pos] = json ? (uint8_t)SYNTAX[rng() % (sizeof(SYNTAX) - 1)] : (uint8_t)rng(); break;
            case 2: d.resize(pos); break;
            case 3: d.insert(d.begin() + (long)pos, (uint8_t)rng()); break;
            default:
                for (size_t k = 0; k < 1 + rng() % 4 && pos + k < d.size(); k++) d[pos + k] = 0xFF;
        }
    }
}

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const bool gltf = strcmp(argv[1], "gltf") == 0;
    const int iters = atoi(argv[2]);
    const int first = gltf ? 4 : 3;
    std::mt19937_64 rng(20261203);
    long loaded = 0, refused = 0;
    for (int a = first; a < argc; a++) {
        const std::vector<uint8_t> seed = read_all(argv[a]);
        if (seed.empty()) return 3;
        const bool png = std::string(argv[a]).find(".png") != std::string::npos;
        for (int it = 0; it < iters; it++) {
            std::vector<uint8_t> d = seed;
            mutate(d, rng, gltf && std::string(argv[a]).find(".glb") == std::string::npos);
            if (gltf) {
                FILE *f = fopen(argv[3], "wb");
                fwrite(d.data(), 1, d.size(), f);
                fclose(f);
                sr_gltf *g = nullptr;
                if (sr_gltf_load(argv[3], &g) == SR_OK) {
                    loaded++;
                    sr_gltf_destroy(g);
                } else {
                    refused++;
                }
            } else {
                uint32_t w = 0, h = 0;
                std::vector<uint32_t> px;
                std::string err;
                bool ok = png ? sr::png_decode_rgba8(d.data(), d.size(), w, h, px, err)
                              : sr::jpeg_decode_rgba8(d.data(), d.size(), w, h, px, err);
                if (ok && px.size() != (size_t)w * h) {
                    printf("%s iteration %d: %zu texels for %ux%u\n", argv[a], it, px.size(), w, h);
                    return 1;
                }
                (ok ? loaded : refused)++;
            }
        }
    }
    printf("loaded %ld refused %ld\n", loaded, refused);
    return 0;
}
