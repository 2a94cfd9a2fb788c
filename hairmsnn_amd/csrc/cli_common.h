// Shared helpers of the headless command-line renderers.
#pragma once

#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "hmsnn.h"

namespace cli {

inline int die(const char* what) {
  std::fprintf(stderr, "%s: %s\n", what, hmsnn_last_error());
  return -1;  // the reference's main returns -1 on a failed scene load
}

inline std::string exr_name(std::string png) {
  // the reference saves EXR next to image_output with ".png" -> ".exr"
  const size_t pos = png.find(".png");
  if (pos != std::string::npos) png.replace(pos, 4, ".exr");
  else if (png.size() < 4 || png.substr(png.size() - 4) != ".exr") png += ".exr";
  return png;
}

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline void write_stats(const std::string& path, const std::string& body) {
  if (path.empty()) return;
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) {
    std::fprintf(stderr, "cannot write stats to %s\n", path.c_str());
    return;
  }
  std::fputs(body.c_str(), f);
  std::fclose(f);
}

}  // namespace cli
