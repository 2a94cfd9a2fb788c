// Image output: uncompressed scanline OpenEXR (RGBA, FP32), written with the
// vertical flip of the reference's saveEXR (model.cpp:366-383): the render
// buffers hold the bottom row first, EXR files the top row first.
#include <zlib.h>

#include <algorithm>
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

// ---------------------------------------------------------------------------
// Image input: scanline OpenEXR (single part), channels R, G, B[, A] or Y in
// HALF, FLOAT or UINT, compression NONE, RLE, ZIPS or ZIP (zlib), as the
// reference's environment maps are read with tinyexr's LoadEXR
// (model.cpp:158-180): RGBA floats, file row order (top row first), alpha 1
// when absent, a single luminance channel copied to R, G and B. Tiled,
// multi-part, deep and PIZ/PXR24/B44/DWA files are refused with a message.

namespace {

struct ExrReader {
  const uint8_t* p;
  size_t n, pos = 0;
  bool ok = true;
  // k more bytes at pos; written so that neither side can wrap (a chunk
  // offset from the file may be anywhere in 0 .. 2^64 - 1)
  bool need(size_t k) {
    if (k > n || pos > n - k) ok = false;
    return ok;
  }
  uint32_t u32() {
    if (!need(4)) return 0;
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) v |= (uint32_t)p[pos + k] << (8 * k);
    pos += 4;
    return v;
  }
  int32_t i32() { return (int32_t)u32(); }
  uint64_t u64() {
    const uint64_t lo = u32();
    return lo | ((uint64_t)u32() << 32);
  }
  uint8_t u8() { return need(1) ? p[pos++] : 0; }
  std::string str() {
    std::string s;
    while (need(1) && p[pos] != 0) s.push_back((char)p[pos++]);
    if (need(1)) ++pos;
    return s;
  }
};

float half_to_float(uint16_t h) {
  const uint32_t sign = (uint32_t)(h >> 15) << 31;
  const int exp = (h >> 10) & 0x1f;
  uint32_t mant = h & 0x3ffu;
  uint32_t bits;
  if (exp == 0) {
    if (mant == 0) {
      bits = sign;
    } else {  // subnormal: renormalise
      int e = -1;
      do {
        ++e;
        mant <<= 1;
      } while (!(mant & 0x400u));
      bits = sign | ((uint32_t)(127 - 15 - e) << 23) | ((mant & 0x3ffu) << 13);
    }
  } else if (exp == 31) {
    bits = sign | 0x7f800000u | (mant << 13);
  } else {
    bits = sign | ((uint32_t)(exp - 15 + 127) << 23) | (mant << 13);
  }
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

// OpenEXR's byte predictor and even/odd split, undone in place into out.
void exr_unpredict(std::vector<uint8_t>& t, uint8_t* out) {
  for (size_t i = 1; i < t.size(); ++i) t[i] = (uint8_t)((int)t[i - 1] + (int)t[i] - 128);
  const size_t n = t.size(), half = (n + 1) / 2;
  for (size_t i = 0, a = 0, b = half; i < n; ++i) out[i] = (i & 1) ? t[b++] : t[a++];
}

bool exr_rle(const uint8_t* in, size_t len, std::vector<uint8_t>& out, size_t expect) {
  out.clear();
  size_t i = 0;
  while (i < len) {
    const int c = (int8_t)in[i++];
    if (c < 0) {
      const size_t k = (size_t)(-c);
      if (i + k > len || out.size() + k > expect) return false;
      out.insert(out.end(), in + i, in + i + k);
      i += k;
    } else {
      if (i >= len || out.size() + (size_t)c + 1 > expect) return false;
      out.insert(out.end(), (size_t)c + 1, in[i++]);
    }
  }
  return out.size() == expect;
}

}  // namespace

namespace hm {

int read_exr(const std::string& path, int& width, int& height, std::vector<float>& rgba, std::string& err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    err = "cannot open " + path;
    return HMSNN_EIO;
  }
  std::vector<uint8_t> buf;
  {
    uint8_t tmp[1 << 16];
    size_t k;
    while ((k = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + k);
    std::fclose(f);
  }
  ExrReader r{buf.data(), buf.size()};
  if (r.u32() != 20000630u) {
    err = path + ": not an OpenEXR file";
    return HMSNN_EIO;
  }
  const uint32_t ver = r.u32();
  if ((ver & 0xff) != 2 || (ver & 0x200u) || (ver & 0x800u) || (ver & 0x1000u)) {
    err = path + ": only single-part scanline OpenEXR files are supported (tiled, deep or multi-part given)";
    return HMSNN_EUNSUPPORTED;
  }
  struct Ch {
    std::string name;
    int type;
  };
  std::vector<Ch> chans;
  int comp = -1;
  int32_t x0 = 0, y0 = 0, x1 = -1, y1 = -1;
  for (;;) {
    const std::string name = r.str();
    if (!r.ok) break;
    if (name.empty()) break;
    const std::string type = r.str();
    const uint32_t size = r.u32();
    const size_t end = r.pos + size;
    if (!r.need(size)) break;
    if (name == "channels" && type == "chlist") {
      while (r.pos < end && r.p[r.pos] != 0) {
        Ch c;
        c.name = r.str();
        c.type = r.i32();
        r.u8();
        r.u8();
        r.u8();
        r.u8();
        const int xs = r.i32(), ys = r.i32();
        if (xs != 1 || ys != 1) {
          err = path + ": subsampled channels are not supported";
          return HMSNN_EUNSUPPORTED;
        }
        chans.push_back(c);
      }
    } else if (name == "compression") {
      if (size < 1) break;
      comp = r.p[r.pos];
    } else if (name == "dataWindow" && type == "box2i") {
      if (size < 16) break;
      x0 = r.i32();
      y0 = r.i32();
      x1 = r.i32();
      y1 = r.i32();
    }
    r.pos = end;
  }
  if (!r.ok || chans.empty() || comp < 0 || x1 < x0 || y1 < y0) {
    err = path + ": malformed OpenEXR header";
    return HMSNN_EIO;
  }
  const int lines = comp == 0 || comp == 1 || comp == 2 ? 1 : comp == 3 ? 16 : -1;
  if (lines < 0) {
    err = path + ": OpenEXR compression " + std::to_string(comp) +
          " is not supported (NONE, RLE, ZIPS and ZIP are)";
    return HMSNN_EUNSUPPORTED;
  }
  const int64_t W64 = (int64_t)x1 - x0 + 1, H64 = (int64_t)y1 - y0 + 1;
  if (W64 > 65536 || H64 > 65536 || W64 * H64 > (int64_t)16384 * 16384) {
    err = path + ": image too large";
    return HMSNN_EINVAL;
  }
  const int W = (int)W64, H = (int)H64;
  std::vector<size_t> coff(chans.size());
  size_t line_bytes = 0;
  int ci[4] = {-1, -1, -1, -1}, cy = -1;  // R, G, B, A; Y
  for (size_t c = 0; c < chans.size(); ++c) {
    const int t = chans[c].type;
    if (t < 0 || t > 2) {
      err = path + ": unknown channel type";
      return HMSNN_EIO;
    }
    coff[c] = line_bytes;
    line_bytes += (size_t)W * (t == 1 ? 2 : 4);
    const std::string& nm = chans[c].name;
    if (nm == "R") ci[0] = (int)c;
    if (nm == "G") ci[1] = (int)c;
    if (nm == "B") ci[2] = (int)c;
    if (nm == "A") ci[3] = (int)c;
    if (nm == "Y") cy = (int)c;
  }
  if (ci[0] < 0 && cy < 0) {
    err = path + ": no R or Y channel";
    return HMSNN_EUNSUPPORTED;
  }
  rgba.assign((size_t)W * H * 4, 0.f);
  for (size_t i = 0; i < (size_t)W * H; ++i) rgba[4 * i + 3] = 1.f;
  const int nchunks = (H + lines - 1) / lines;
  std::vector<uint64_t> offs(nchunks);
  for (int k = 0; k < nchunks; ++k) offs[k] = r.u64();
  if (!r.ok) {
    err = path + ": truncated OpenEXR offset table";
    return HMSNN_EIO;
  }
  std::vector<uint8_t> raw, tmp;
  auto value = [&](const uint8_t* line, int c, int x) -> float {
    const uint8_t* q = line + coff[c];
    const int t = chans[c].type;
    if (t == 1) return half_to_float((uint16_t)(q[2 * x] | (q[2 * x + 1] << 8)));
    uint32_t u = (uint32_t)q[4 * x] | ((uint32_t)q[4 * x + 1] << 8) | ((uint32_t)q[4 * x + 2] << 16) |
                 ((uint32_t)q[4 * x + 3] << 24);
    if (t == 0) return (float)u;
    float fv;
    std::memcpy(&fv, &u, 4);
    return fv;
  };
  for (int k = 0; k < nchunks; ++k) {
    r.pos = (size_t)offs[k];
    const int32_t ys = r.i32();
    const uint32_t size = r.u32();
    if (!r.ok || !r.need(size) || ys < y0 || ys > y1) {
      err = path + ": truncated or malformed OpenEXR chunk";
      return HMSNN_EIO;
    }
    const int nl = std::min(lines, y1 - ys + 1);
    const size_t expect = line_bytes * (size_t)nl;
    const uint8_t* src = r.p + r.pos;
    raw.resize(expect);
    if (comp == 0 || size == expect) {
      if (size != expect) {
        err = path + ": OpenEXR chunk size mismatch";
        return HMSNN_EIO;
      }
      std::memcpy(raw.data(), src, expect);
    } else if (comp == 1) {
      if (!exr_rle(src, size, tmp, expect)) {
        err = path + ": corrupt RLE chunk";
        return HMSNN_EIO;
      }
      exr_unpredict(tmp, raw.data());
    } else {
      tmp.resize(expect);
      uLongf dl = (uLongf)expect;
      if (uncompress(tmp.data(), &dl, src, (uLong)size) != Z_OK || dl != expect) {
        err = path + ": corrupt ZIP chunk";
        return HMSNN_EIO;
      }
      exr_unpredict(tmp, raw.data());
    }
    for (int l = 0; l < nl; ++l) {
      const uint8_t* line = raw.data() + (size_t)l * line_bytes;
      float* dst = rgba.data() + (size_t)(ys - y0 + l) * W * 4;
      for (int x = 0; x < W; ++x) {
        if (ci[0] >= 0) {
          for (int c = 0; c < 3; ++c) dst[4 * x + c] = ci[c] >= 0 ? value(line, ci[c], x) : 0.f;
        } else {
          const float v = value(line, cy, x);
          dst[4 * x] = dst[4 * x + 1] = dst[4 * x + 2] = v;
        }
        if (ci[3] >= 0) dst[4 * x + 3] = value(line, ci[3], x);
      }
    }
  }
  width = W;
  height = H;
  return HMSNN_OK;
}

}  // namespace hm

extern "C" int hmsnn_exr_size(const char* path, int* width, int* height) {
  if (!path || !width || !height) return hm::fail(HMSNN_EINVAL, "hmsnn_exr_size: bad arguments");
  std::vector<float> img;
  std::string err;
  const int rc = hm::read_exr(path, *width, *height, img, err);
  return rc ? hm::fail(rc, err) : HMSNN_OK;
}

extern "C" int hmsnn_load_exr(const char* path, float* rgba, int width, int height) {
  if (!path || !rgba || width <= 0 || height <= 0) return hm::fail(HMSNN_EINVAL, "hmsnn_load_exr: bad arguments");
  std::vector<float> img;
  std::string err;
  int w = 0, h = 0;
  const int rc = hm::read_exr(path, w, h, img, err);
  if (rc) return hm::fail(rc, err);
  if (w != width || h != height) return hm::fail(HMSNN_EINVAL, "hmsnn_load_exr: size mismatch");
  std::memcpy(rgba, img.data(), sizeof(float) * img.size());
  return HMSNN_OK;
}
