// Radiance-cache network weight files: the file behind the scene's
// tcnn.init_weights (scene.cpp:330-331), which the reference hands to
// TINY_MLP::loadWeights (neural_network.cu:23-32) as tiny-cuda-nn's JSON
// serialisation of its trainer. The network here is not tiny-cuda-nn's, so
// the file is this project's own JSON document:
//   {"format": "hairmsnn_amd.mlp", "version": 1,
//    "architecture": {"inputs": 12, "features": 64, "hidden": [64, 64, 64, 64, 64],
//                     "outputs": 3, "activation": "ReLU", "output_activation": "Exponential",
//                     "bias": false},
//    "params": [20672 numbers: 5 x (64 x 64) hidden then 3 x 64 output, row-major [out][in]]}
// Values are written with 9 significant digits, which round-trips FP32.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hmsnn.h"
#include "hm_internal.h"
#include "hm_json.h"

namespace hm {

constexpr int kMlpParams = 5 * 64 * 64 + 3 * 64;
constexpr const char* kMlpFormat = "hairmsnn_amd.mlp";

int mlp_weights_write(const std::string& path, const float* w, std::string& err) {
  std::ostringstream os;
  os << "{\"format\": \"" << kMlpFormat << "\", \"version\": 1,\n"
     << " \"architecture\": {\"inputs\": 12, \"features\": 64, \"hidden\": [64, 64, 64, 64, 64], \"outputs\": 3, "
        "\"activation\": \"ReLU\", \"output_activation\": \"Exponential\", \"bias\": false},\n"
     << " \"params\": [";
  char buf[32];
  for (int k = 0; k < kMlpParams; ++k) {
    std::snprintf(buf, sizeof(buf), "%.9g", (double)w[k]);
    os << (k ? (k % 8 ? ", " : ",\n  ") : "") << buf;
  }
  os << "]}\n";
  std::ofstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot write the network weight file " + path;
    return kEIO;
  }
  const std::string s = os.str();
  f.write(s.data(), (std::streamsize)s.size());
  if (!f) {
    err = "error writing the network weight file " + path;
    return kEIO;
  }
  return kOk;
}

int mlp_weights_read(const std::string& path, float* w, std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open the network weight file " + path;
    return kEIO;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  Json j;
  try {
    j = Json::parse(ss.str());
  } catch (const JsonError& e) {
    err = "network weight file " + path + ": " + e.what();
    return kEInval;
  }
  if (!j.is_object() || !j.has("format") || !j.at("format").is_string() || j.at("format").str() != kMlpFormat) {
    err = "network weight file " + path +
          " is not a hairmsnn_amd network file (the radiance cache here is not tiny-cuda-nn's, so a "
          "tiny-cuda-nn serialisation cannot be loaded)";
    return kEInval;
  }
  if (!j.has("version") || !j.at("version").is_number() || j.at("version").i() != 1) {
    err = "network weight file " + path + ": unsupported version";
    return kEInval;
  }
  if (!j.has("params") || !j.at("params").is_array() || (int)j.at("params").size() != kMlpParams) {
    err = "network weight file " + path + ": expects " + std::to_string(kMlpParams) + " params";
    return kEInval;
  }
  const Json& p = j.at("params");
  for (int k = 0; k < kMlpParams; ++k) {
    const Json& v = p.at((size_t)k);
    if (!v.is_number()) {
      err = "network weight file " + path + ": param " + std::to_string(k) + " is not a number";
      return kEInval;
    }
    w[k] = (float)v.num();
  }
  return kOk;
}

}  // namespace hm
