// Image files of a rendered pass: the window image (RGBA8, re.fs) as PNG and the linear radiance
// as PFM.  The reference shows its image in a GLUT window only (glslppm.cpp:506-526); these are
// the file forms of the same buffers.  Pixel rows arrive bottom-up (GL's gl_FragCoord order):
// PNG stores rows top-down, so they are flipped; PFM stores rows bottom-up, so they are not.
//
// The PNG is written without a compression library: its zlib stream holds "stored" deflate blocks
// (RFC 1951 block type 0), with the Adler-32 of the raw bytes (RFC 1950) and a CRC-32 per chunk
// (PNG specification, ISO/IEC 15948).

#include "../../include/ppm.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

uint32_t crc_table[256];
bool crc_ready = false;

void crc_init() {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
    crc_ready = true;
}

uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    for (size_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c;
}

void put32(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put32(out, (uint32_t)data.size());
    std::vector<uint8_t> body(type, type + 4);
    body.insert(body.end(), data.begin(), data.end());
    out.insert(out.end(), body.begin(), body.end());
    put32(out, crc32(body.data(), body.size()) ^ 0xFFFFFFFFu);
}

int write_file(const char* path, const void* data, size_t n) {
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return PPM_E_IO;
    const size_t w = std::fwrite(data, 1, n, fp);
    const int rc = std::fclose(fp);
    return (w == n && rc == 0) ? PPM_OK : PPM_E_IO;
}

}  // namespace

extern "C" {

int ppm_write_png(const char* path, const uint8_t* rgba, int32_t width, int32_t height) {
    if (!path || !rgba || width <= 0 || height <= 0 || (int64_t)width * height > (1ll << 28)) return PPM_E_INVAL;
    if (!crc_ready) crc_init();
    // raw scanlines, top row first, each with filter byte 0
    const size_t row = (size_t)width * 4;
    std::vector<uint8_t> raw;
    raw.reserve((row + 1) * height);
    for (int j = height - 1; j >= 0; j--) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba + (size_t)j * row, rgba + (size_t)(j + 1) * row);
    }
    // zlib stream of stored blocks
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(len & 0xFF));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xFF));
        z.push_back((uint8_t)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + (long)pos, raw.begin() + (long)(pos + len));
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    put32(z, (b << 16) | a);

    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)width);
    put32(ihdr, (uint32_t)height);
    ihdr.push_back(8);  // bit depth
    ihdr.push_back(6);  // colour type RGBA
    ihdr.push_back(0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    return write_file(path, out.data(), out.size());
}

int ppm_write_pfm(const char* path, const float* rgba, int32_t width, int32_t height) {
    if (!path || !rgba || width <= 0 || height <= 0 || (int64_t)width * height > (1ll << 28)) return PPM_E_INVAL;
    // colour PFM: "PF", size, negative scale for little-endian samples, rows bottom-up
    const std::string head = "PF\n" + std::to_string(width) + " " + std::to_string(height) + "\n-1.0\n";
    std::vector<uint8_t> out(head.begin(), head.end());
    const size_t n = (size_t)width * height;
    const size_t off = out.size();
    out.resize(off + n * 12);
    for (size_t i = 0; i < n; i++) std::memcpy(&out[off + i * 12], rgba + i * 4, 12);
    return write_file(path, out.data(), out.size());
}

}  // extern "C"
