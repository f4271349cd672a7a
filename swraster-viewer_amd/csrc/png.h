// png.h -- zlib inflate and PNG decoding for the glTF loader.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace sr {

// Inflate a zlib stream (RFC 1950 wrapper around RFC 1951 deflate); appends
// to out.  Returns false with a message on malformed input.
bool zlib_inflate(const uint8_t *data, size_t size, std::vector<uint8_t> &out, std::string &err);

// Decode a non-interlaced PNG into packed r<<24 | g<<16 | b<<8 | a texels.
bool png_decode_rgba8(const uint8_t *data, size_t size, uint32_t &width, uint32_t &height, std::vector<uint32_t> &out,
                      std::string &err);

// Decode a JPEG (baseline, extended sequential or progressive Huffman,
// 8-bit, one or three components) into packed r<<24 | g<<16 | b<<8 | 0xFF.
bool jpeg_decode_rgba8(const uint8_t *data, size_t size, uint32_t &width, uint32_t &height, std::vector<uint32_t> &out,
                       std::string &err);

}  // namespace sr
