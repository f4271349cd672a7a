// cache.cpp -- the reference's load-time cache files:
//   <scene>.gi   VGI0 (try_load_gi_cache / save_gi_cache, src/gi.rs:30-117)
//   cubemap.ggx  GGX0 (try_load_prefiltered_specular_cubemap_cache /
//                save_prefiltered_specular_cubemap_cache, src/texture.rs:426-552)
// Both are a 20-byte little-endian header (magic, version, three u32) and a
// brotli stream of the little-endian payload.  Brotli comes from the
// system's libbrotlienc / libbrotlidec, opened at run time; without them
// these entry points fail with SR_ERR_UNSUPPORTED.

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "runtime.h"
#include "swraster.h"

namespace sr {
namespace {

// The public brotli C API (RFC 7932 reference library, encode.h / decode.h).
typedef int (*DecompressFn)(size_t, const uint8_t *, size_t *, uint8_t *);
typedef int (*CompressFn)(int, int, int, size_t, const uint8_t *, size_t *, uint8_t *);
typedef size_t (*BoundFn)(size_t);
struct Brotli {
    DecompressFn decompress = nullptr;
    CompressFn compress = nullptr;
    BoundFn bound = nullptr;
};

const Brotli &brotli() {
    static Brotli b;
    static std::once_flag once;
    std::call_once(once, [] {
        if (void *d = dlopen("libbrotlidec.so.1", RTLD_NOW | RTLD_LOCAL))
            b.decompress = (DecompressFn)dlsym(d, "BrotliDecoderDecompress");
        if (void *e = dlopen("libbrotlienc.so.1", RTLD_NOW | RTLD_LOCAL)) {
            b.compress = (CompressFn)dlsym(e, "BrotliEncoderCompress");
            b.bound = (BoundFn)dlsym(e, "BrotliEncoderMaxCompressedSize");
        }
    });
    return b;
}

const int BROTLI_QUALITY = 6, BROTLI_LGWIN = 22, BROTLI_MODE_GENERIC = 0, BROTLI_DECODER_SUCCESS = 1;

void put_u32(std::vector<uint8_t> &b, uint32_t v) {
    for (int k = 0; k < 4; k++) b.push_back((uint8_t)(v >> (8 * k)));
}
uint32_t get_u32(const uint8_t *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

int write_cache(const char *path, const char magic[4], uint32_t version, const uint32_t dims[3], const void *payload,
                size_t bytes) {
    const Brotli &B = brotli();
    if (!B.compress || !B.bound) return fail(SR_ERR_UNSUPPORTED, "libbrotlienc.so.1 is not available");
    std::vector<uint8_t> out;
    out.insert(out.end(), magic, magic + 4);
    put_u32(out, version);
    for (int k = 0; k < 3; k++) put_u32(out, dims[k]);
    size_t cap = B.bound(bytes);
    if (cap == 0) cap = bytes + (bytes >> 2) + 1024;
    size_t hdr = out.size();
    out.resize(hdr + cap);
    size_t n = cap;
    if (!B.compress(BROTLI_QUALITY, BROTLI_LGWIN, BROTLI_MODE_GENERIC, bytes, (const uint8_t *)payload, &n,
                    out.data() + hdr))
        return fail(SR_ERR_INVALID_ARGUMENT, "brotli compression failed");
    out.resize(hdr + n);
    FILE *f = fopen(path, "wb");
    if (!f) return fail(SR_ERR_NOT_FOUND, "cannot write %s", path);
    size_t w = fwrite(out.data(), 1, out.size(), f);
    int c = fclose(f);
    if (w != out.size() || c != 0) return fail(SR_ERR_NOT_FOUND, "short write to %s", path);
    return SR_OK;
}

// Reads the header and payload.  SR_ERR_NOT_FOUND: no file, or a file the
// reference would not use (short, other magic, version or dimensions, or a
// payload of another size); SR_ERR_INVALID_ARGUMENT: a corrupt brotli stream
// (the reference reports the read error and recomputes).
int read_cache(const char *path, const char magic[4], uint32_t version, const uint32_t dims[3], void *payload,
               size_t bytes) {
    const Brotli &B = brotli();
    if (!B.decompress) return fail(SR_ERR_UNSUPPORTED, "libbrotlidec.so.1 is not available");
    FILE *f = fopen(path, "rb");
    if (!f) return fail(SR_ERR_NOT_FOUND, "no cache file %s", path);
    std::vector<uint8_t> in;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) in.insert(in.end(), buf, buf + n);
    fclose(f);
    if (in.size() < 20 || memcmp(in.data(), magic, 4) != 0 || get_u32(&in[4]) != version)
        return fail(SR_ERR_NOT_FOUND, "%s: not a version %u %.4s cache", path, version, magic);
    for (int k = 0; k < 3; k++)
        if (get_u32(&in[8 + 4 * k]) != dims[k]) return fail(SR_ERR_NOT_FOUND, "%s: other dimensions", path);
    // one spare byte tells a longer payload from an exact one
    std::vector<uint8_t> out(bytes + 1);
    size_t got = out.size();
    int r = B.decompress(in.size() - 20, in.data() + 20, &got, out.data());
    if (r != BROTLI_DECODER_SUCCESS) {
        // a stream that decodes past the expected size reports that it needs more output
        if (r == 3) return fail(SR_ERR_NOT_FOUND, "%s: payload larger than expected", path);
        return fail(SR_ERR_INVALID_ARGUMENT, "%s: corrupt brotli stream", path);
    }
    if (got != bytes) return fail(SR_ERR_NOT_FOUND, "%s: payload of %zu bytes, expected %zu", path, got, bytes);
    memcpy(payload, out.data(), bytes);
    return SR_OK;
}

const char GI_MAGIC[4] = {'V', 'G', 'I', '0'};
const uint32_t GI_VERSION = 8;
const char GGX_MAGIC[4] = {'G', 'G', 'X', '0'};
const uint32_t GGX_VERSION = 1;

uint32_t ilog2(uint32_t v) { return 31u - (uint32_t)__builtin_clz(v); }

}  // namespace
}  // namespace sr

using namespace sr;

extern "C" {

SR_API int sr_gi_cache_save(const char *path, const sr_voxel_grid *grid) {
    if (!path || !grid || !grid->sh4) return fail(SR_ERR_INVALID_ARGUMENT, "the grid has no GI coefficients");
    uint32_t dims[3] = {(uint32_t)grid->dims[0], (uint32_t)grid->dims[1], (uint32_t)grid->dims[2]};
    size_t n = (size_t)dims[0] * dims[1] * dims[2];
    if (n == 0) return fail(SR_ERR_INVALID_ARGUMENT, "empty grid");
    return write_cache(path, GI_MAGIC, GI_VERSION, dims, grid->sh4, n * 16 * sizeof(float));
}

SR_API int sr_gi_cache_load(const char *path, const int32_t dims[3], float *sh4) {
    if (!path || !dims || !sh4) return fail(SR_ERR_INVALID_ARGUMENT, "null argument");
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return fail(SR_ERR_INVALID_ARGUMENT, "bad grid dimensions");
    uint32_t d[3] = {(uint32_t)dims[0], (uint32_t)dims[1], (uint32_t)dims[2]};
    size_t n = (size_t)d[0] * d[1] * d[2];
    return read_cache(path, GI_MAGIC, GI_VERSION, d, sh4, n * 16 * sizeof(float));
}

SR_API int sr_specular_cache_save(const char *path, const sr_environment *env) {
    if (!path || !env || !env->specular || env->specular_size == 0)
        return fail(SR_ERR_INVALID_ARGUMENT, "the environment has no specular cubemap");
    uint32_t dims[3] = {env->specular_size, env->specular_size, env->specular_mips};
    size_t n = (size_t)env->specular_size * env->specular_size * 6 * env->specular_mips;
    return write_cache(path, GGX_MAGIC, GGX_VERSION, dims, env->specular, n * 4);
}

SR_API int sr_specular_cache_load(const char *path, uint32_t face_size, uint32_t *specular) {
    if (!path || !specular || face_size == 0) return fail(SR_ERR_INVALID_ARGUMENT, "bad specular cache arguments");
    uint32_t mips = 1 + ilog2(face_size);
    uint32_t d[3] = {face_size, face_size, mips};
    size_t n = (size_t)face_size * face_size * 6 * mips;
    return read_cache(path, GGX_MAGIC, GGX_VERSION, d, specular, n * 4);
}

}  // extern "C"
