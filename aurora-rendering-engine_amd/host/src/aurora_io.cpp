// aurora_io.cpp — on-disk formats of the are:: host API: Portable FloatMap
// dumps of FP32 radiance and AOV images, and the accumulation checkpoint of
// a progressive render (see are/aurora.hpp for the layouts).
//
// Reference anchors: the reference's only image file format is the P6 PPM of
// are::Texture (src/texture.cpp:9-50 read, :362-395 write); these formats
// keep the float precision the path tracer's sums need, with the same error
// contract (std::runtime_error naming the file).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "are/aurora.hpp"

namespace are {

namespace {

struct file {
    FILE* f = nullptr;
    file(const std::string& path, const char* mode) : f(fopen(path.c_str(), mode)) {}
    ~file() {
        if (f) fclose(f);
    }
};

void put_u32(std::vector<unsigned char>& b, uint32_t v) {
    for (int k = 0; k < 4; ++k) b.push_back((unsigned char)(v >> (8 * k)));
}
void put_f32(std::vector<unsigned char>& b, float x) {
    uint32_t v;
    std::memcpy(&v, &x, 4);
    put_u32(b, v);
}
uint32_t get_u32(const unsigned char* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
float get_f32(const unsigned char* p) {
    const uint32_t v = get_u32(p);
    float x;
    std::memcpy(&x, &v, 4);
    return x;
}

// CRC-32 of the zlib / IEEE 802.3 polynomial (reflected 0xEDB88320), by table
uint32_t crc32(const unsigned char* p, size_t n) {
    static uint32_t table[256];
    static const bool init = [] {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        return true;
    }();
    (void)init;
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

constexpr char kMagic[8] = {'A', 'U', 'R', 'A', 'C', 'C', '0', '1'};
constexpr uint32_t kHeader = 128;

}  // namespace

// ------------------------------------------------------------ PFM -------
void write_pfm(const std::string& path, int width, int height, const std::vector<float>& rgb) {
    if (width <= 0 || height <= 0 || rgb.size() != (size_t)width * height * 3)
        throw std::invalid_argument("image size does not match width*height*3");
    file fp(path, "wb");
    if (!fp.f) throw std::runtime_error("Failed to open file for writing: " + path);
    bool ok = fprintf(fp.f, "PF\n%d %d\n-1.0\n", width, height) > 0;
    // little-endian floats (negative scale), scanlines from the bottom row up
    std::vector<unsigned char> row;
    row.reserve((size_t)width * 12);
    for (int y = height - 1; y >= 0 && ok; --y) {
        row.clear();
        for (size_t i = 0; i < (size_t)width * 3; ++i) put_f32(row, rgb[(size_t)y * width * 3 + i]);
        ok = fwrite(row.data(), 1, row.size(), fp.f) == row.size();
    }
    if (!ok) throw std::runtime_error("Failed to write file: " + path);
}

std::vector<float> read_pfm(const std::string& path, int& width, int& height) {
    file fp(path, "rb");
    if (!fp.f) throw std::runtime_error("Failed to open image file: " + path);
    char fmt[3] = {0};
    if (fscanf(fp.f, "%2s", fmt) != 1 || std::string(fmt) != "PF")
        throw std::runtime_error("Unsupported PFM format in file: " + path);
    int w = 0, h = 0;
    double scale = 0.0;
    if (fscanf(fp.f, "%d %d %lf", &w, &h, &scale) != 3 || w <= 0 || h <= 0 || scale == 0.0 || !std::isfinite(scale))
        throw std::runtime_error("Invalid PFM header in file: " + path);
    fgetc(fp.f);
    const bool little = scale < 0.0;
    std::vector<unsigned char> bytes((size_t)w * h * 12);
    if (fread(bytes.data(), 1, bytes.size(), fp.f) != bytes.size())
        throw std::runtime_error("Unexpected end of file while reading pixel data in file: " + path);
    std::vector<float> rgb((size_t)w * h * 3);
    for (int y = 0; y < h; ++y)
        for (size_t i = 0; i < (size_t)w * 3; ++i) {
            unsigned char* p = &bytes[((size_t)(h - 1 - y) * w * 3 + i) * 4];
            if (!little) {
                std::swap(p[0], p[3]);
                std::swap(p[1], p[2]);
            }
            rgb[(size_t)y * w * 3 + i] = get_f32(p);
        }
    width = w;
    height = h;
    return rgb;
}

// ---------------------------------------------------- checkpoint -------
void accumulation::save(const std::string& path) const {
    const size_t n = (size_t)camera.image_width * camera.image_height * 3;
    const bool aov = !albedo_sum.empty();
    if (n == 0 || rgb_sum.size() != n) throw std::invalid_argument("accumulation: rgb_sum does not match the camera");
    if (aov && (albedo_sum.size() != n || normal_sum.size() != n))
        throw std::invalid_argument("accumulation: AOV sums do not match the camera");
    std::vector<unsigned char> payload;
    payload.reserve(n * 4 * (aov ? 3 : 1));
    for (float x : rgb_sum) put_f32(payload, x);
    if (aov) {
        for (float x : albedo_sum) put_f32(payload, x);
        for (float x : normal_sum) put_f32(payload, x);
    }
    std::vector<unsigned char> h(kMagic, kMagic + 8);
    put_u32(h, kHeader);
    put_u32(h, camera.image_width);
    put_u32(h, camera.image_height);
    put_u32(h, aov ? 1u : 0u);
    put_f32(h, camera.vfov);
    for (float v : camera.lookfrom) put_f32(h, v);
    for (float v : camera.lookat) put_f32(h, v);
    for (float v : camera.vup) put_f32(h, v);
    put_f32(h, camera.defocus_angle);
    put_f32(h, camera.focus_dist);
    put_u32(h, spp);
    put_u32(h, render.sample_offset);
    put_u32(h, render.max_depth);
    put_u32(h, render.seed);
    for (float v : render.background) put_f32(h, v);
    put_f32(h, render.tmin);
    while (h.size() < kHeader - 8) put_u32(h, 0u);
    put_u32(h, crc32(payload.data(), payload.size()));
    put_u32(h, 0u);
    file fp(path, "wb");
    if (!fp.f) throw std::runtime_error("Failed to open file for writing: " + path);
    if (fwrite(h.data(), 1, h.size(), fp.f) != h.size() ||
        fwrite(payload.data(), 1, payload.size(), fp.f) != payload.size())
        throw std::runtime_error("Failed to write file: " + path);
}

accumulation accumulation::load(const std::string& path) {
    file fp(path, "rb");
    if (!fp.f) throw std::runtime_error("Failed to open checkpoint file: " + path);
    unsigned char h[kHeader];
    if (fread(h, 1, kHeader, fp.f) != kHeader || std::memcmp(h, kMagic, 8) != 0 || get_u32(h + 8) != kHeader)
        throw std::runtime_error("Not an accumulation checkpoint: " + path);
    accumulation a;
    const unsigned char* p = h + 12;
    a.camera.image_width = get_u32(p);
    a.camera.image_height = get_u32(p + 4);
    const bool aov = (get_u32(p + 8) & 1u) != 0;
    p += 12;
    a.camera.vfov = get_f32(p);
    p += 4;
    for (float& v : a.camera.lookfrom) v = get_f32(p), p += 4;
    for (float& v : a.camera.lookat) v = get_f32(p), p += 4;
    for (float& v : a.camera.vup) v = get_f32(p), p += 4;
    a.camera.defocus_angle = get_f32(p);
    a.camera.focus_dist = get_f32(p + 4);
    p += 8;
    a.spp = get_u32(p);
    a.render.sample_offset = get_u32(p + 4);
    a.render.max_depth = get_u32(p + 8);
    a.render.seed = get_u32(p + 12);
    p += 16;
    for (float& v : a.render.background) v = get_f32(p), p += 4;
    a.render.tmin = get_f32(p);
    const uint32_t want = get_u32(h + kHeader - 8);
    const uint64_t n = (uint64_t)a.camera.image_width * a.camera.image_height * 3;
    if (n == 0 || n > (1ull << 32)) throw std::runtime_error("Invalid checkpoint header in file: " + path);
    std::vector<unsigned char> payload(n * 4 * (aov ? 3 : 1));
    if (fread(payload.data(), 1, payload.size(), fp.f) != payload.size())
        throw std::runtime_error("Unexpected end of file while reading sums in file: " + path);
    if (fgetc(fp.f) != EOF) throw std::runtime_error("Trailing data in checkpoint file: " + path);
    if (crc32(payload.data(), payload.size()) != want) throw std::runtime_error("Checksum mismatch in file: " + path);
    auto take = [&](std::vector<float>& v, size_t off) {
        v.resize(n);
        for (size_t i = 0; i < n; ++i) v[i] = get_f32(&payload[(off + i) * 4]);
    };
    take(a.rgb_sum, 0);
    if (aov) {
        take(a.albedo_sum, n);
        take(a.normal_sum, 2 * n);
    }
    return a;
}

}  // namespace are
