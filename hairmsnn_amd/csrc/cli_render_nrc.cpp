// render_nrc <scene.json>
// The reference's neural-radiance-cache comparator (render_nrc.cu:1080-1104)
// is not built yet; the entry point exists so scripts find it and fail with
// a clear message.
#include <cstdio>

int main(int argc, char** argv) {
  std::fprintf(stderr, "%s: the NRC renderer is not implemented in this build; use render_hair_msnn or "
               "render_path_tracing\n", argc > 0 ? argv[0] : "render_nrc");
  return 2;
}
