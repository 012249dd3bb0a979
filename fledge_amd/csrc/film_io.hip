// Film output: Film::saveBuffer's ".exr" branch (src/film.cpp:13-49), which
// writes the colour buffer as a 3-channel float image through OpenImageIO
// starting from the film's last row (negative y stride).  OpenImageIO is not
// part of this build; the file is written directly in the OpenEXR layout
// (single-part scanline file, uncompressed, one scanline per block, FLOAT
// channels B, G, R in the order the format requires).  Host code only.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace flg {

namespace {

struct Bytes {
  std::vector<unsigned char> b;
  void raw(const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    b.insert(b.end(), c, c + n);
  }
  void i32(int32_t v) { raw(&v, 4); }  // the format is little-endian, as is the host
  void u64(uint64_t v) { raw(&v, 8); }
  void f32(float v) { raw(&v, 4); }
  void u8(unsigned char v) { b.push_back(v); }
  void str(const char* s) { raw(s, std::strlen(s) + 1); }
  void attr(const char* name, const char* type, int32_t size) {
    str(name);
    str(type);
    i32(size);
  }
};

}  // namespace

}  // namespace flg

using namespace flg;

extern "C" {

int flg_film_save_exr(const char* path, int32_t width, int32_t height, const float* h_rgb) {
  return host_guard(nullptr, "flg_film_save_exr", [&]() -> int {
    if (!path || !h_rgb || width <= 0 || height <= 0) return FLG_ERR_INVALID;
    if ((int64_t)width * 12 > ((int64_t)1 << 30)) return FLG_ERR_RANGE;
    Bytes f;
    const unsigned char magic[4] = {0x76, 0x2f, 0x31, 0x01};
    f.raw(magic, 4);
    f.i32(2);  // version 2, single-part scanline file, short names
    // header
    const char* chans[3] = {"B", "G", "R"};
    f.attr("channels", "chlist", 3 * (2 + 16) + 1);
    for (const char* c : chans) {
      f.str(c);
      f.i32(2);  // FLOAT
      f.u8(0);   // pLinear
      f.u8(0);
      f.u8(0);
      f.u8(0);
      f.i32(1);  // xSampling
      f.i32(1);  // ySampling
    }
    f.u8(0);
    f.attr("compression", "compression", 1);
    f.u8(0);  // NO_COMPRESSION
    f.attr("dataWindow", "box2i", 16);
    f.i32(0);
    f.i32(0);
    f.i32(width - 1);
    f.i32(height - 1);
    f.attr("displayWindow", "box2i", 16);
    f.i32(0);
    f.i32(0);
    f.i32(width - 1);
    f.i32(height - 1);
    f.attr("lineOrder", "lineOrder", 1);
    f.u8(0);  // INCREASING_Y
    f.attr("pixelAspectRatio", "float", 4);
    f.f32(1.0f);
    f.attr("screenWindowCenter", "v2f", 8);
    f.f32(0.0f);
    f.f32(0.0f);
    f.attr("screenWindowWidth", "float", 4);
    f.f32(1.0f);
    f.u8(0);  // end of header
    // offset table, then one block per scanline: y, byte count, B row, G row, R row
    const uint64_t table = f.b.size();
    const uint64_t block = 8 + (uint64_t)width * 12;
    for (int32_t y = 0; y < height; ++y) f.u64(table + 8 * (uint64_t)height + (uint64_t)y * block);
    std::vector<float> row((size_t)width);
    for (int32_t y = 0; y < height; ++y) {
      f.i32(y);
      f.i32(width * 12);
      const float* src = h_rgb + (size_t)(height - 1 - y) * width * 3;  // the film's last row first
      for (int c = 2; c >= 0; --c) {  // B, G, R
        for (int32_t x = 0; x < width; ++x) row[x] = src[3 * x + c];
        f.raw(row.data(), (size_t)width * 4);
      }
    }
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return FLG_ERR_INVALID;
    const size_t n = std::fwrite(f.b.data(), 1, f.b.size(), fp);
    const int closed = std::fclose(fp);
    return (n == f.b.size() && closed == 0) ? FLG_OK : FLG_ERR_INVALID;
  });
}

}  // extern "C"
