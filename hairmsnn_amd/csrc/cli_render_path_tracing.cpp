// render_path_tracing <scene.json> [spp]
// Headless counterpart of the reference's render_path_tracing executable
// (render_path_tracing.cu:817-841): loads the scene, renders `spp` samples per
// pixel (integrator.spp by default), writes the mean image as EXR next to
// integrator.image_output and a stats JSON to integrator.stats_output.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cli_common.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <scene.json> [spp]\n", argv[0]);
    return -1;
  }
  hmsnn_ctx* ctx = nullptr;
  std::printf("Loading scene %s\n", argv[1]);
  if (hmsnn_create(argv[1], &ctx)) return cli::die("Error loading scene");
  hmsnn_info info;
  hmsnn_get_info(ctx, &info);
  const int spp = argc >= 3 ? std::atoi(argv[2]) : info.spp;
  std::vector<float> img((size_t)info.width * info.height * 4);
  const double t0 = cli::now_s();
  if (hmsnn_render_path_tracing(ctx, spp, img.data())) return cli::die("render failed");
  const double dt = cli::now_s() - t0;
  const double paths = (double)info.width * info.height * spp;
  std::printf("%d x %d, %d spp, %d segments: %.3f s, %.2f M paths/s\n", info.width, info.height, spp,
              info.num_segments, dt, paths / dt / 1e6);
  char out[4096], stats[4096];
  hmsnn_get_string(ctx, 0, out, sizeof(out));
  hmsnn_get_string(ctx, 1, stats, sizeof(stats));
  const std::string exr = cli::exr_name(out);
  if (hmsnn_save_exr(exr.c_str(), img.data(), info.width, info.height)) return cli::die("save failed");
  std::printf("(EXR) Accum buffer saved to %s\n", exr.c_str());
  char body[1024];
  std::snprintf(body, sizeof(body),
                "{\"renderer\": \"path_tracing\", \"width\": %d, \"height\": %d, \"spp\": %d, \"seconds\": %.6f, "
                "\"paths_per_second\": %.1f, \"segments\": %d}\n",
                info.width, info.height, spp, dt, paths / dt, info.num_segments);
  cli::write_stats(stats, body);
  hmsnn_destroy(ctx);
  return 0;
}
