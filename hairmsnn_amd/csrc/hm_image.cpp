// Image output: uncompressed scanline OpenEXR (RGBA, FP32), written with the
// vertical flip of the reference's saveEXR (model.cpp:366-383): the render
// buffers hold the bottom row first, EXR files the top row first.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hmsnn.h"
#include "hm_api_util.h"

namespace {

void put_u32(std::vector<uint8_t>& b, uint32_t v) {
  for (int k = 0; k < 4; ++k) b.push_back((uint8_t)(v >> (8 * k)));
}
void put_i32(std::vector<uint8_t>& b, int32_t v) { put_u32(b, (uint32_t)v); }
void put_u64(std::vector<uint8_t>& b, uint64_t v) {
  for (int k = 0; k < 8; ++k) b.push_back((uint8_t)(v >> (8 * k)));
}
void put_f32(std::vector<uint8_t>& b, float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  put_u32(b, u);
}
void put_str(std::vector<uint8_t>& b, const char* s) {
  while (*s) b.push_back((uint8_t)*s++);
  b.push_back(0);
}
void attr(std::vector<uint8_t>& b, const char* name, const char* type, const std::vector<uint8_t>& val) {
  put_str(b, name);
  put_str(b, type);
  put_u32(b, (uint32_t)val.size());
  b.insert(b.end(), val.begin(), val.end());
}

}  // namespace

extern "C" int hmsnn_save_exr(const char* path, const float* rgba, int width, int height) {
  if (!path || !rgba || width <= 0 || height <= 0) return hm::fail(HMSNN_EINVAL, "hmsnn_save_exr: bad arguments");
  std::vector<uint8_t> b;
  put_u32(b, 20000630u);  // magic
  put_u32(b, 2u);         // version 2, single-part scanline
  // channels: A, B, G, R (alphabetical), FLOAT (2), no linear, sampling 1 x 1
  std::vector<uint8_t> ch;
  for (const char* c : {"A", "B", "G", "R"}) {
    put_str(ch, c);
    put_i32(ch, 2);
    ch.push_back(0);
    ch.push_back(0);
    ch.push_back(0);
    ch.push_back(0);
    put_i32(ch, 1);
    put_i32(ch, 1);
  }
  ch.push_back(0);
  attr(b, "channels", "chlist", ch);
  attr(b, "compression", "compression", {0});  // NO_COMPRESSION
  std::vector<uint8_t> box;
  put_i32(box, 0);
  put_i32(box, 0);
  put_i32(box, width - 1);
  put_i32(box, height - 1);
  attr(b, "dataWindow", "box2i", box);
  attr(b, "displayWindow", "box2i", box);
  attr(b, "lineOrder", "lineOrder", {0});  // INCREASING_Y
  std::vector<uint8_t> par;
  put_f32(par, 1.f);
  attr(b, "pixelAspectRatio", "float", par);
  std::vector<uint8_t> swc;
  put_f32(swc, 0.f);
  put_f32(swc, 0.f);
  attr(b, "screenWindowCenter", "v2f", swc);
  std::vector<uint8_t> sww;
  put_f32(sww, 1.f);
  attr(b, "screenWindowWidth", "float", sww);
  b.push_back(0);  // end of header
  const size_t line_bytes = (size_t)width * 4 * 4;
  const size_t chunk = 8 + line_bytes;  // y + size + data
  uint64_t offset = b.size() + 8 * (size_t)height;
  for (int y = 0; y < height; ++y) {
    put_u64(b, offset);
    offset += chunk;
  }
  for (int y = 0; y < height; ++y) {
    const int src = height - 1 - y;  // flip: file row 0 is the image top
    put_i32(b, y);
    put_u32(b, (uint32_t)line_bytes);
    for (int c : {3, 2, 1, 0})  // A, B, G, R planes
      for (int x = 0; x < width; ++x) put_f32(b, rgba[4 * ((size_t)src * width + x) + c]);
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return hm::fail(HMSNN_EIO, std::string("cannot write ") + path);
  const size_t w = std::fwrite(b.data(), 1, b.size(), f);
  std::fclose(f);
  return w == b.size() ? HMSNN_OK : hm::fail(HMSNN_EIO, std::string("short write to ") + path);
}
