// rc_export.hip — frame export: the display image (RGBA8) flipped to top-down
// rows on the device, read back once, and written as a PNG
// (FrameExporter.exportFrame, src/main/java/com/retrocam/export/FrameExporter.java:37-65,
// which reads the GL texture and lets stb write it with rows flipped).
//
// The PNG is 8-bit RGBA, filter type 0 on every row, in a zlib stream of
// stored (uncompressed) deflate blocks: no compression library is needed, and
// any PNG reader decodes it. CRC-32 (ISO 3309 polynomial 0xEDB88320, as the PNG
// specification requires) and Adler-32 (RFC 1950) are computed here.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rc_internal.h"
#include "retrocam_hip.h"

namespace rc {
namespace {

__global__ void flip_rows_kernel(const uchar4 *src, uchar4 *dst, int w, int h) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x < w) dst[(size_t)y * w + x] = src[(size_t)(h - 1 - y) * w + x];
}

uint32_t crc_table_entry(uint32_t n) {
    uint32_t c = n;
    for (int k = 0; k < 8; ++k) c = (c & 1u) ? (0xEDB88320u ^ (c >> 1)) : (c >> 1);
    return c;
}

struct Crc32 {
    uint32_t table[256];
    Crc32() {
        for (uint32_t n = 0; n < 256; ++n) table[n] = crc_table_entry(n);
    }
    uint32_t update(uint32_t crc, const uint8_t *p, size_t n) const {
        uint32_t c = crc ^ 0xFFFFFFFFu;
        for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
        return c ^ 0xFFFFFFFFu;
    }
};

void put_be32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void put_chunk(std::vector<uint8_t> &out, const Crc32 &crc, const char type[4], const std::vector<uint8_t> &data) {
    put_be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put_be32(out, crc.update(0u, out.data() + start, out.size() - start));
}

}  // namespace

// PNG bytes of a w x h RGBA8 image whose rows are given top-down.
std::vector<uint8_t> encode_png(const uint8_t *rgba, int w, int h) {
    static const Crc32 crc;
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)w);
    put_be32(ihdr, (uint32_t)h);
    ihdr.push_back(8);  // bit depth
    ihdr.push_back(6);  // colour type: RGBA
    ihdr.push_back(0);  // deflate
    ihdr.push_back(0);  // adaptive filtering (every row uses filter 0)
    ihdr.push_back(0);  // no interlace
    put_chunk(out, crc, "IHDR", ihdr);
    // raw scanlines: filter byte 0 + the row
    const size_t rowBytes = (size_t)w * 4;
    std::vector<uint8_t> raw;
    raw.reserve((rowBytes + 1) * (size_t)h);
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba + (size_t)y * rowBytes, rgba + (size_t)(y + 1) * rowBytes);
    }
    // zlib stream: CMF 0x78 (deflate, 32 KiB window), FLG 0x01 (check bits), stored blocks, Adler-32
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t n = raw.size() - pos < 65535 ? raw.size() - pos : 65535;
        const bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(n & 0xFF));
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)(~n & 0xFF));
        z.push_back((uint8_t)((~n >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + (long)pos, raw.begin() + (long)(pos + n));
        pos += n;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    put_chunk(out, crc, "IDAT", z);
    put_chunk(out, crc, "IEND", {});
    return out;
}

}  // namespace rc

using namespace rc;

static int write_file(rc_ctx *ctx, const char *path, const std::vector<uint8_t> &bytes) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return ctx ? set_error(ctx, RC_ERR_INVALID_ARG, std::string("cannot open ") + path) : RC_ERR_INVALID_ARG;
    const size_t n = std::fwrite(bytes.data(), 1, bytes.size(), f);
    const int closed = std::fclose(f);
    if (n != bytes.size() || closed != 0)
        return ctx ? set_error(ctx, RC_ERR_INVALID_ARG, std::string("write failed: ") + path) : RC_ERR_INVALID_ARG;
    return RC_OK;
}

extern "C" {

int rc_encode_png(const uint8_t *rgba, int32_t width, int32_t height, int32_t flipRows, const char *path) {
    if (!rgba || !path || width <= 0 || height <= 0 || width > 16384 || height > 16384) return RC_ERR_INVALID_ARG;
    const uint8_t *src = rgba;
    std::vector<uint8_t> flipped;
    if (flipRows) {
        const size_t row = (size_t)width * 4;
        flipped.resize(row * (size_t)height);
        for (int y = 0; y < height; ++y)
            std::memcpy(flipped.data() + (size_t)y * row, rgba + (size_t)(height - 1 - y) * row, row);
        src = flipped.data();
    }
    return write_file(nullptr, path, encode_png(src, width, height));
}

int rc_display_size(const rc_ctx *ctx, int32_t *width, int32_t *height) {
    if (!ctx || !width || !height) return RC_ERR_INVALID_ARG;
    if (!ctx->img.postOut) return RC_ERR_STATE;
    *width = ctx->img.postW;
    *height = ctx->img.postH;
    return RC_OK;
}

int rc_export_frame(rc_ctx *ctx, const char *path, uint8_t *rgbaTopDown) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc = capture_block(ctx, "rc_export_frame")) return rc;
    const Images &im = ctx->img;
    if (!im.postOut || !im.display) return set_error(ctx, RC_ERR_STATE, "no displayed frame: run rc_post first");
    if (!path && !rgbaTopDown) return set_error(ctx, RC_ERR_INVALID_ARG, "path and rgbaTopDown are both NULL");
    const int w = im.postW, h = im.postH;
    const size_t bytes = (size_t)w * (size_t)h * 4;
    (void)hipSetDevice(ctx->device);
    uchar4 *dFlip = nullptr;
    uint8_t *hBuf = nullptr;
    hipError_t e = hipMallocAsync((void **)&dFlip, bytes, ctx->stream);
    if (e == hipSuccess) e = hipHostMalloc((void **)&hBuf, bytes, hipHostMallocDefault);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(flip_rows_kernel, dim3((w + 255) / 256, h), dim3(256), 0, ctx->stream, im.display, dFlip, w,
                           h);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hBuf, dFlip, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (dFlip) (void)hipFreeAsync(dFlip, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    int rc = RC_OK;
    if (e != hipSuccess) {
        rc = hip_fail(ctx, e, "rc_export_frame");
    } else {
        if (rgbaTopDown) std::memcpy(rgbaTopDown, hBuf, bytes);
        if (path) rc = write_file(ctx, path, encode_png(hBuf, w, h));
    }
    if (hBuf) (void)hipHostFree(hBuf);
    return rc;
}

}  // extern "C"
