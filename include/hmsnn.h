/*
 * hmsnn.h - C ABI of the MI355X-native HairMSNN hair path tracer.
 *
 * The reference exposes its hot path as three executables driven by one
 * scene-config JSON (render_path_tracing.cu:817-841, render_hair_msnn.cu:
 * 1146-1175, render_nrc.cu:1080-1104); inside, the boundary is an
 * owlLaunch2D of a ray-generation program per pass and the tiny-cuda-nn
 * calls (neural_network.cuh:33-50). This header is the library behind
 * drop-in versions of those executables: every entry point is plain C,
 * takes plain pointers and sizes, and returns an hmsnn_status.
 *
 * Conventions
 *  - Status: 0 on success, otherwise an HMSNN_E* code; hmsnn_last_error()
 *    returns a thread-local message for the last failure on this thread.
 *  - Ownership: the caller owns every buffer it passes. A context owns its
 *    geometry, BVH, environment tables, path state and network.
 *  - Threading: one context per host thread; calls on a context are
 *    serialised. Functions named *_dev take device pointers and run on the
 *    given HIP stream (NULL = the library's stream) without synchronising;
 *    everything else is synchronous.
 *  - Angles are radians unless a name says _deg.
 */
#ifndef HMSNN_H
#define HMSNN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hmsnn_status {
  HMSNN_OK = 0,
  HMSNN_EINVAL = 1,      /* bad argument or malformed scene */
  HMSNN_EIO = 2,         /* file could not be read or written */
  HMSNN_EDEVICE = 3,     /* HIP runtime or kernel failure, or no GPU */
  HMSNN_ENOMEM = 4,      /* device allocation failed */
  HMSNN_EUNSUPPORTED = 5 /* input the library does not implement */
} hmsnn_status;

typedef struct hmsnn_ctx hmsnn_ctx;

/* ---- library ------------------------------------------------------------ */

/* Message of the last failure on the calling thread ("" if none). */
const char* hmsnn_last_error(void);
/* ABI version: major * 10000 + minor * 100 + patch. */
int hmsnn_version(void);
/* Number of visible GPUs; writes the device name of device 0. */
int hmsnn_device_count(int* count, char* name, int name_len);

/* Device memory helpers for harnesses that have no allocator of their own. */
int hmsnn_dev_alloc(size_t bytes, void** ptr);
int hmsnn_dev_free(void* ptr);
int hmsnn_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
int hmsnn_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
int hmsnn_dev_synchronize(void);

/* ---- hair BSDF (disney_hair.cuh:25-386) --------------------------------- */

/* Global hair material, as parsed by parseScene (scene.cpp:193-227):
 * alpha in radians (the JSON holds degrees), gains R, TT, TRT, TRRT. */
typedef struct hmsnn_hair_params {
  float beta_m, beta_n, alpha;
  float gain[4];
} hmsnn_hair_params;

/* One BSDF query in the hair frame: x = tangent, wo.y = 0 in renderer use.
 * h = dot(frame y, surface normal) in [-1, 1]. */
typedef struct hmsnn_bsdf_query {
  float wo[3];
  float wi[3];
  float h;
  float sigma_a[3];
} hmsnn_bsdf_query;

/* f * |cos theta_i| (n x RGB) and the lobe-mixture pdf (n) at wi
 * (replaces disney_hair, disney_hair.cuh:185-266). */
int hmsnn_hair_bsdf_eval_dev(const hmsnn_hair_params* params, const hmsnn_bsdf_query* q_dev, int n,
                             float* f_rgb_dev, float* pdf_dev, void* stream);
/* Importance-samples wi from four uniforms per query (u4_dev: n x 4) and
 * returns wi (n x 3), f * |cos| (n x 3), pdf (n) and the chosen lobe (n)
 * (replaces sample_disney_hair, disney_hair.cuh:276-386; q.wi is ignored). */
int hmsnn_hair_bsdf_sample_dev(const hmsnn_hair_params* params, const hmsnn_bsdf_query* q_dev,
                               const float* u4_dev, int n, float* wi_dev, float* f_rgb_dev, float* pdf_dev,
                               int* lobe_dev, void* stream);

/* ---- scene context ------------------------------------------------------- */

/* Parses a scene JSON (the reference's schema, scene.cpp:119-339, plus the
 * procedural keys hair.procedural, surface.sphere,
 * lights.environment.constant / .procedural and the hairmsnn block),
 * generates the geometry and environment, builds the BVH and the
 * environment tables on the GPU. Fails like parseScene fails. */
int hmsnn_create(const char* scene_json_path, hmsnn_ctx** out);
int hmsnn_create_from_json(const char* json_text, const char* base_dir, hmsnn_ctx** out);
int hmsnn_destroy(hmsnn_ctx* ctx);

typedef struct hmsnn_info {
  int width, height, spp, path_v1, path_v2;
  int num_segments, num_strands;
  int bvh_nodes, bvh_depth; /* 4-wide tree that traversal reads: node count, inner-node levels */
  int env_width, env_height;
  int num_dir_lights, has_env, has_sphere, mis, env_pdf;
  float scene_scale;
  float sigma_a[3];
  float beta_m, beta_n, alpha;
  float gain[4];
  float env_scale, env_rot_phi;
  float sphere[4];        /* centre xyz, radius */
  float sphere_albedo[3];
  float sphere_alpha;     /* after clamp to [0.01, 1] and squaring */
  float camera[12];       /* pos, dir_00, dir_du, dir_dv (render_path_tracing.cu:767-799) */
  float dir_from[8][3];
  float dir_emit[8][3];
  int num_triangles;      /* OBJ head mesh (surface.geometry), 0 without one */
  int num_curves;         /* round Catmull-Rom hair segments (.hair, primitive "curve"), 0 otherwise */
} hmsnn_info;
int hmsnn_get_info(const hmsnn_ctx* ctx, hmsnn_info* info);
/* Strings of the scene: which = 0 integrator.image_output, 1
 * integrator.stats_output (both resolved against the scene file's
 * directory when relative), 2 the scene file path. */
int hmsnn_get_string(const hmsnn_ctx* ctx, int which, char* buf, int len);
/* Parses a scene JSON on the host only (no GPU needed) and reports what a
 * context would be created with; fields that need generated content
 * (BVH, environment size) are left 0 except num_segments/num_strands of
 * procedural hair. Same failure rules as hmsnn_create (parseScene,
 * scene.cpp:119-339). */
int hmsnn_parse_scene(const char* json_text, hmsnn_info* info);
/* Generates a scene's procedural content on the host only (no GPU): hair
 * segments (8 floats each, as hmsnn_export_segments) and the RGBA32F
 * environment map. Capacities are in floats; *num_segments / *env_w /
 * *env_h report the sizes (call with NULL buffers to query). */
int hmsnn_generate_host(const char* json_text, float* seg_host, size_t seg_cap, int* num_segments,
                        float* env_rgba_host, size_t env_cap, int* env_w, int* env_h);

/* Copies generated content to the host: segments as 8 floats each
 * (p0 xyz, radius, p1 xyz, strand id); the RGBA32F environment map
 * (env_width x env_height x 4) and its tables (conditional pdf/cdf of
 * (env_width + 1) x env_height, marginal pdf/cdf of env_height + 1). Any
 * pointer may be NULL to skip that array. */
int hmsnn_export_segments(const hmsnn_ctx* ctx, float* seg_host);
/* Round Catmull-Rom hair curves (.hair strands with hair.primitive "curve",
 * the default; the reference's owlCurvesSetControlPoints /
 * owlCurvesSetSegmentIndices input, render_path_tracing.cu:236-244), 20
 * floats per segment: four control points (x, y, z, radius), strand id, end
 * flags (1: first segment of its strand, 2: last), 0, 0. With curves,
 * hmsnn_export_segments gives each segment's chord. */
int hmsnn_export_curves(const hmsnn_ctx* ctx, float* crv_host);
/* The same curves from the scene file on the host only (no GPU); capacity
 * in floats, *num_curves reports the count (NULL buffer to query). */
int hmsnn_generate_curves_host(const char* json_text, float* crv_host, size_t crv_cap, int* num_curves);
/* The OBJ head mesh as loaded (loadOBJ, model.cpp:233-358): per triangle 3
 * vertices (9 floats), 3 vertex normals (9 floats) and the diffuse colour (3
 * floats); NULL skips an array. */
int hmsnn_export_mesh(const hmsnn_ctx* ctx, float* vertices_host, float* normals_host, float* diffuse_host);
/* Host-only OBJ reader of the same mesh layout (no GPU needed): call with
 * NULL arrays for the count, then with arrays of `capacity` triangles. */
int hmsnn_obj_load(const char* path, int* num_triangles, float* vertices_host, float* normals_host,
                   float* diffuse_host, int capacity);
int hmsnn_export_env(const hmsnn_ctx* ctx, float* rgba_host, float* c_pdf, float* c_cdf, float* m_pdf,
                     float* m_cdf);
/* Per-strand absorption (num_strands x 3); NULL restores the global value. */
int hmsnn_set_strand_sigma_a(hmsnn_ctx* ctx, const float* sigma_host, int num_strands);

/* ---- path tracing (render_path_tracing: render_path_tracing.cu:404-423,
 *      cuda/path_tracing.cu:19-67) --------------------------------------- */

/* Adds offset to the pass index that seeds each pixel's random stream
 * (path_tracing.cu:29 seeds with accumId + 10007), for every later pass of
 * this context (path tracing, HairMSNN frames and the adjoint's replay).
 * Default 0 reproduces the reference's streams. No reference counterpart:
 * it lets a reference image and a test render use independent samples. */
int hmsnn_set_seed_offset(hmsnn_ctx* ctx, int offset);
/* Concurrency of path tracing (1..4, default 4): with the per-pass schedule
 * the passes kept in flight at once, with the streaming schedule the stream
 * groups the pixels are dealt to; each has its own path buffers and stream.
 * Accumulation stays in pass order, so the image does not depend on it. 1
 * runs everything on one stream (per-kernel timings of hmsnn_pt_time_passes
 * are then those of kernels running alone). No reference counterpart: the
 * reference renders one launch per frame. */
int hmsnn_set_pass_concurrency(hmsnn_ctx* ctx, int passes_in_flight);
/* Schedule of hmsnn_pt_passes. HMSNN_PT_STREAM (default): (pass, pixel)
 * jobs are handed out in pass order to a compacted wavefront of two paths per
 * pixel; a finished path's radiance waits in a 16-pass ring until its pass is
 * complete, so each wavefront launch stays full until the call's last jobs
 * drain. hmsnn_set_pass_concurrency sets the number of stream groups the
 * pixels are dealt to. HMSNN_PT_PASSES: one wavefront per pass, passes in
 * flight per hmsnn_set_pass_concurrency. Both accumulate each pixel's passes
 * in pass order: the images are equal. The setting also selects how
 * hmsnn_msnn_begin traces the warm-up's training records: HMSNN_PT_STREAM
 * on a compacted wavefront refilled with new records as paths finish,
 * HMSNN_PT_PASSES one wavefront per pass; the records, network and images
 * are equal.
 * No reference counterpart (the reference launches one megakernel per pass). */
typedef enum hmsnn_pt_schedule { HMSNN_PT_STREAM = 0, HMSNN_PT_PASSES = 1 } hmsnn_pt_schedule;
int hmsnn_set_pt_schedule(hmsnn_ctx* ctx, int schedule);

/* Clears the accumulation (accumId = 0), like cameraChanged does. */
int hmsnn_pt_reset(hmsnn_ctx* ctx);
/* Traces `passes` more samples per pixel into the accumulation buffers.
 * Asynchronous on the context's stream; hmsnn_ctx_synchronize waits. */
int hmsnn_pt_passes(hmsnn_ctx* ctx, int passes);
int hmsnn_ctx_synchronize(hmsnn_ctx* ctx);
/* Samples accumulated so far (the reference's accumId). */
int hmsnn_pt_accum_count(const hmsnn_ctx* ctx, int* count);
/* Copies width x height x 4 floats: which = 0 mean, 1 running sum. Row 0 is
 * the bottom row of the image (the reference's framebuffer order). */
int hmsnn_get_image(hmsnn_ctx* ctx, int which, float* rgba_host);
/* Reset, trace spp samples per pixel, copy the mean image. */
int hmsnn_render_path_tracing(hmsnn_ctx* ctx, int spp, float* rgba_host);
/* Per-pixel radiance of the last pass before accumulation (w x h x 4). */
int hmsnn_get_last_pass(hmsnn_ctx* ctx, float* rgba_host);

/* Traversal statistics of `passes` passes (not accumulated into the image):
 * out[0] closest-hit rays, out[1] nodes visited by them, out[2] segments
 * tested by them, out[3] any-hit rays, out[4] nodes, out[5] segments. */
int hmsnn_pt_trace_stats(hmsnn_ctx* ctx, int passes, double out[6]);
/* Times `passes` accumulated passes with HIP events on the context's
 * stream: ms_total, and per-kernel-class averages (ms per launch):
 * out[0] total ms, out[1] extend, out[2] shade, out[3] shadow,
 * out[4] launches of extend. */
int hmsnn_pt_time_passes(hmsnn_ctx* ctx, int passes, double out[5]);

/* Traces an explicit ray list against the scene (rays_dev: n x 8 floats,
 * origin xyz, pad, direction xyz, pad). Closest hit writes prim (segment id,
 * -1 miss, -2 head sphere), t and u; any hit writes prim = 1 if occluded. */
int hmsnn_trace_rays_dev(hmsnn_ctx* ctx, const float* rays_dev, int n, int any_hit, int* prim_dev,
                         float* t_dev, float* u_dev);

/* ---- HairMSNN (render_hair_msnn: render_hair_msnn.cu:98-771,
 *      cuda/hair_msnn.cu:17-357) ------------------------------------------ */

typedef struct hmsnn_msnn_stats {
  double gbuffer_ms, train_ms, infer_ms, render_ms, total_ms; /* summed over the call's frames */
  double loss;
  int frames, train_steps;
  double setup_ms, warmup_ms; /* hmsnn_msnn_begin: scene samples + buffers, warm-up training */
} hmsnn_msnn_stats;

/* Sets BETA exactly as the CLI argument (kernel beta = BETA - 1,
 * render_hair_msnn.cu:1170) and runs the fixed-length warm-up training. */
int hmsnn_msnn_begin(hmsnn_ctx* ctx, int BETA);
/* Renders `frames` frames (shuffle, G_BUFFER, one training step, inference,
 * RENDER composite), accumulating. */
int hmsnn_msnn_frames(hmsnn_ctx* ctx, int frames, hmsnn_msnn_stats* stats);
/* which = 0 final mean, 1 short-path mean, 2 network mean (w x h x 4). */
int hmsnn_msnn_get_image(hmsnn_ctx* ctx, int which, float* rgba_host);
/* Writes the radiance cache of the HairMSNN state to a weight file (the
 * format hmsnn_mlp_save writes), to be named by tcnn.init_weights
 * (scene.cpp:330-331 -> render_hair_msnn.cu:109-110). */
int hmsnn_msnn_save_weights(hmsnn_ctx* ctx, const char* path);
/* sizes[0] training records per frame (nnTrainInput rows,
 * render_hair_msnn.cu:112-120), [1] every_nth (render_hair_msnn.cu:106),
 * [2] frames rendered (accumId), [3] training steps taken. */
int hmsnn_msnn_get_sizes(hmsnn_ctx* ctx, int sizes[4]);
/* The last frame's buffers, for inspection and tests (the reference's
 * nnTrainInput / nnTrainOutput / nnFrameInput / GBuffer / nnFrameOutput,
 * render_hair_msnn.cu:112-145); `bytes` must equal the buffer's size. */
typedef enum hmsnn_msnn_buffer {
  HMSNN_MSNN_TRAIN_INPUT = 0,   /* float[n_train][12] training records */
  HMSNN_MSNN_TRAIN_TARGET = 1,  /* float[n_train][3] full - short, scrubbed (hair_msnn.cu:243-276) */
  HMSNN_MSNN_TRAIN_PIXEL = 2,   /* int[n_train] pixel of each training record */
  HMSNN_MSNN_GB_INPUT = 3,      /* float[w*h][12] network input at each pixel's primary hit */
  HMSNN_MSNN_GB_HIT = 4,        /* uint8[w*h] bit 0 hit, bit 1 surface */
  HMSNN_MSNN_GB_COLOR = 5,      /* float4[w*h] short-path colour (training pixels: full path) */
  HMSNN_MSNN_NN_OUTPUT = 6      /* float[w*h][3] network output */
} hmsnn_msnn_buffer;
int hmsnn_msnn_get_buffer(hmsnn_ctx* ctx, int which, void* host, size_t bytes);
/* Reset, warm up, render spp frames, copy the final mean. */
int hmsnn_render_hair_msnn(hmsnn_ctx* ctx, int BETA, int spp, float* rgba_host, hmsnn_msnn_stats* stats);

/* ---- neural radiance caching (render_nrc: render_nrc.cu:116-701,
 *      cuda/nrc.cu:18-398) -------------------------------------------------- */

typedef struct hmsnn_nrc_stats {
  double gbuffer_ms, infer_ms, render_ms, train_ms; /* summed over the call's frames (GPU events) */
  double total_ms;                                   /* host wall clock of the call */
  int frames, train_steps;
  int cache_queries;      /* last frame: pixels whose path ended in a cache query */
  int train_records_used; /* last frame: training records holding a path vertex */
} hmsnn_nrc_stats;

/* Creates the cache network and frame buffers (RenderWindowNRC::initialize,
 * render_nrc.cu:116-180): training records from the scene's "nrc" block
 * (default 65,536), training pixels = records / 40, one in every_nth pixels
 * (integer divisions as the reference). */
int hmsnn_nrc_begin(hmsnn_ctx* ctx);
/* Spread threshold multiplier c (setBounceThresholdControl,
 * render_nrc.cu:42-46) and unbiased suffixes for every training pixel
 * (setAllUnbiasedPaths, render_nrc.cu:13-17). */
int hmsnn_nrc_set_params(hmsnn_ctx* ctx, float c, int all_unbiased);
/* Renders `frames` frames (RenderWindowNRC::render, render_nrc.cu:640-701):
 * shuffle, G_BUFFER, inference, RENDER (composite + training data), one
 * training step; accumulating. */
int hmsnn_nrc_frames(hmsnn_ctx* ctx, int frames, hmsnn_nrc_stats* stats);
/* Mean image (w x h x 4). */
int hmsnn_nrc_get_image(hmsnn_ctx* ctx, float* rgba_host);
/* sizes[0] training pixels, [1] every_nth, [2] cache queries per frame
 * (nnFrameSize), [3] training records. */
int hmsnn_nrc_get_sizes(hmsnn_ctx* ctx, int sizes[4]);
/* The last frame's buffers, for inspection and tests; `bytes` must equal the
 * buffer's size. */
typedef enum hmsnn_nrc_buffer {
  HMSNN_NRC_QUERY = 0,          /* float[n_query][12] cache-query records */
  HMSNN_NRC_CACHE = 1,          /* float[n_query][3] network outputs */
  HMSNN_NRC_GB_RADIANCE = 2,    /* float4[w*h] radiance before the query */
  HMSNN_NRC_GB_BETA = 3,        /* float4[w*h] throughput at the query (0: none) */
  HMSNN_NRC_GB_HIT = 4,         /* uint8[w*h] camera ray hit */
  HMSNN_NRC_TRAIN_INPUT = 5,    /* float[n_rec][12] */
  HMSNN_NRC_TRAIN_TARGET = 6,   /* float[n_rec][3] */
  HMSNN_NRC_TRAIN_IDX = 7,      /* int[n_tpx] shuffled training indices */
  HMSNN_NRC_TRAIN_BOUNCES = 8,  /* int[n_tpx] recorded vertices per training pixel */
  HMSNN_NRC_TRAIN_HIT = 9       /* uint8[n_tpx] suffix ended in a cache query */
} hmsnn_nrc_buffer;
int hmsnn_nrc_get_buffer(hmsnn_ctx* ctx, int which, void* host, size_t bytes);
/* The cache network's weights (layout of hmsnn_mlp_get_weights). */
int hmsnn_nrc_get_mlp_weights(hmsnn_ctx* ctx, float* w_host, int n);
int hmsnn_nrc_set_mlp_weights(hmsnn_ctx* ctx, const float* w_host, int n);
/* Writes the cache network to a weight file (the format hmsnn_mlp_save
 * writes), for the scene's tcnn.init_weights (scene.cpp:330-331). */
int hmsnn_nrc_save_weights(hmsnn_ctx* ctx, const char* path);
/* Begin, render spp frames, copy the mean image. */
int hmsnn_render_nrc(hmsnn_ctx* ctx, int spp, float* rgba_host, hmsnn_nrc_stats* stats);

/* ---- radiance-cache MLP (replaces TINY_MLP, neural_network.cuh:33-50) -- */

typedef struct hmsnn_mlp hmsnn_mlp;
typedef enum hmsnn_precision { HMSNN_FP16 = 0, HMSNN_BF16 = 1, HMSNN_FP8 = 2, HMSNN_FP32 = 3 } hmsnn_precision;
typedef enum hmsnn_loss { HMSNN_LOSS_RELATIVE_L2_LUMINANCE = 0, HMSNN_LOSS_RELATIVE_L2 = 1, HMSNN_LOSS_L2 = 2 } hmsnn_loss;

/* 12-float input records -> 64 encoded features -> 5 x 64 ReLU -> 3, exp. */
int hmsnn_mlp_create(int precision, uint64_t seed, hmsnn_mlp** out);
int hmsnn_mlp_destroy(hmsnn_mlp* mlp);
/* Weights as FP32 host arrays: 5 x (64 x 64) hidden then 3 x 64 output,
 * each row-major [out][in]. n = 20672. */
int hmsnn_mlp_get_weights(const hmsnn_mlp* mlp, float* w_host, int n);
int hmsnn_mlp_set_weights(hmsnn_mlp* mlp, const float* w_host, int n);
/* Weight files, the file behind the scene's tcnn.init_weights, which the
 * reference loads with TINY_MLP::loadWeights (neural_network.cu:23-32,
 * render_hair_msnn.cu:109-110, render_nrc.cu:149-150). The network is not
 * tiny-cuda-nn's, so the file is this library's JSON document
 * ({"format": "hairmsnn_amd.mlp", "version": 1, "architecture": {...},
 * "params": [20672 numbers in the layout above]}); a tiny-cuda-nn
 * serialisation is refused with HMSNN_EINVAL. hmsnn_msnn_begin and
 * hmsnn_nrc_begin load the scene's tcnn.init_weights (relative to the scene
 * file) into their network before the warm-up. */
int hmsnn_mlp_save(const hmsnn_mlp* mlp, const char* path);
int hmsnn_mlp_load(hmsnn_mlp* mlp, const char* path);
/* Encodes 12-float records into 64 features (n x 64 FP32), for tests. */
int hmsnn_mlp_encode_dev(const float* in_dev, int n, float* feat_dev, void* stream);
/* Inference: in n x 12 FP32, out n x 3 FP32 (TINY_MLP::inference). */
int hmsnn_mlp_forward_dev(hmsnn_mlp* mlp, const float* in_dev, int n, float* out_dev, void* stream);
/* FP8 networks only (no reference counterpart: tiny-cuda-nn has no FP8 path):
 * the 12 power-of-two exponents the last FP8 inference used, [0..5] per layer
 * weights, [6..11] per layer input; a layer computes
 * 2^(ew + ex) * sum(e4m3(W / 2^ew) * e4m3(x / 2^ex)). Synchronises the
 * device. HMSNN_EINVAL before the first FP8 inference after a weight change. */
int hmsnn_mlp_get_fp8_scales(const hmsnn_mlp* mlp, int* exponents12);
/* Loss and gradient w.r.t. the weights (FP32, 20672) for one batch; does not
 * update the weights. */
int hmsnn_mlp_backward_dev(hmsnn_mlp* mlp, const float* in_dev, const float* target_dev, int n, int loss,
                           float* grad_dev, float* loss_out, void* stream);
/* One Adam step (trainer->training_step): forward, backward, update. */
int hmsnn_mlp_train_step_dev(hmsnn_mlp* mlp, const float* in_dev, const float* target_dev, int n, int loss,
                             float lr, float* loss_out, void* stream);
/* hmsnn_mlp_train_step_dev without allocations (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under
 * graph capture, tight serving or training loops). The same kernels as
 * hmsnn_mlp_train_step_dev, with their scratch (per-block weight-gradient
 * partial sums and losses) in a static 12.6 MB device array per GPU instead
 * of a buffer allocated on demand: n is limited to 1..16384 records
 * (HMSNN_EINVAL beyond), nothing needs clearing (every block writes all of
 * its partials; no memset nodes), and the batch loss is written to loss_dev (device float,
 * optional) instead of being copied to the host. Every call on a GPU shares
 * that GPU's scratch, so calls that overlap on one GPU race. */
int hmsnn_mlp_train_step_dev_noalloc(hmsnn_mlp* mlp, const float* in_dev, const float* target_dev, int n, int loss,
                                     float lr, float* loss_dev, void* stream);

/* ---- image output ------------------------------------------------------ */

/* Writes width x height x 4 floats (render-buffer order, bottom row first) as
 * an uncompressed RGBA FP32 OpenEXR file, flipped like saveEXR
 * (model.cpp:366-383). */
int hmsnn_save_exr(const char* path, const float* rgba, int width, int height);

/* Reads a scanline OpenEXR image as tinyexr's LoadEXR does for the
 * reference's environment maps (model.cpp:158-180): RGBA floats in file row
 * order, alpha 1 when absent, a lone Y channel copied to RGB. Channels HALF,
 * FLOAT or UINT; compression NONE, RLE, ZIPS or ZIP. hmsnn_exr_size gives the
 * size, hmsnn_load_exr fills width x height x 4 floats. */
int hmsnn_exr_size(const char* path, int* width, int* height);
int hmsnn_load_exr(const char* path, float* rgba_host, int width, int height);

/* ---- differentiable appearance (path-replay adjoint) --------------------- */

/* Sets the global hair roughness and scale tilt (the values parseScene reads,
 * scene.cpp:201-208; alpha in radians). */
int hmsnn_set_hair_params(hmsnn_ctx* ctx, float beta_m, float beta_n, float alpha);
/* Gradient of a loss of the spp-sample mean image, given dL/d(image) as
 * width x height x 4 floats (w ignored): per-strand d/d sigma_a
 * (num_strands x 3, FP64, NULL to skip), d/d (beta_m, beta_n, alpha), and
 * optionally each pixel's share of the latter (width x height x 3 floats,
 * NULL to skip). Paths are replayed from the passes 0..spp-1 of the path
 * tracer's streams (the image hmsnn_render_path_tracing(ctx, spp) makes);
 * sampled directions, sampling pdfs and Russian roulette are held fixed,
 * while f and the MIS weights are differentiated. No reference counterpart:
 * the reference renders only. */
int hmsnn_adjoint_render(hmsnn_ctx* ctx, int spp, const float* dl_dimage_host, double* grad_sigma_host,
                         double grad_global[3], float* grad_pixel_host);

#ifdef __cplusplus
}
#endif

#endif /* HMSNN_H */
