// render_nrc <scene.json> [spp]
// Headless counterpart of the reference's render_nrc executable
// (render_nrc.cu:1080-1104): neural radiance caching, spp frames accumulated,
// the mean image saved as EXR next to image_output and a stats JSON.
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
  if (hmsnn_nrc_begin(ctx)) return cli::die("NRC set-up failed");
  const double t1 = cli::now_s();
  hmsnn_nrc_stats st{};
  if (hmsnn_nrc_frames(ctx, spp, &st)) return cli::die("render failed");
  if (hmsnn_nrc_get_image(ctx, img.data())) return cli::die("readback failed");
  const double t2 = cli::now_s();
  std::printf("%d frames: %.3f s (G-buffer %.1f ms, infer %.1f ms, render %.1f ms, train %.1f ms)\n", st.frames,
              t2 - t1, st.gbuffer_ms, st.infer_ms, st.render_ms, st.train_ms);
  char out[4096], stats[4096];
  hmsnn_get_string(ctx, 0, out, sizeof(out));
  hmsnn_get_string(ctx, 1, stats, sizeof(stats));
  const std::string exr = cli::exr_name(out);
  if (hmsnn_save_exr(exr.c_str(), img.data(), info.width, info.height)) return cli::die("save failed");
  std::printf("(EXR) Accum buffer saved to %s\n", exr.c_str());
  char body[1024];
  std::snprintf(body, sizeof(body),
                "{\"renderer\": \"nrc\", \"width\": %d, \"height\": %d, \"frames\": %d, \"setup_seconds\": %.6f, "
                "\"render_seconds\": %.6f, \"gbuffer_ms\": %.3f, \"infer_ms\": %.3f, \"render_ms\": %.3f, "
                "\"train_ms\": %.3f, \"train_steps\": %d, \"cache_queries_last_frame\": %d}\n",
                info.width, info.height, st.frames, t1 - t0, t2 - t1, st.gbuffer_ms, st.infer_ms, st.render_ms,
                st.train_ms, st.train_steps, st.cache_queries);
  cli::write_stats(stats, body);
  hmsnn_destroy(ctx);
  return 0;
}
