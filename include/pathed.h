/* pathed.h -- C ABI of the MI355X path tracer (libpathed.so).
 *
 * Plain pointers and sizes only; no torch or HIP types cross this boundary.  Every call
 * returns PT_OK (0) or a negative pt_status; pt_last_error() describes the last failure of
 * the calling thread.  Host pointers are host memory unless a name says otherwise.
 *
 * The renderer replaces the reference's Renderer plugin (include/renderers/renderer.h:14-35)
 * and the device-side scene upload (src/cuda_globals.cu:10-139); the scene calls replace
 * SceneData/SceneAdapter (include/scene_data.h:18-61, src/scene_data.cu:9-118); the C++
 * RenderSession/SceneModel and the `path_tracer` Python module are built on top of this ABI.
 *
 * Threading: one host thread per renderer at a time.  Calls on different renderers may run
 * concurrently.
 */
#ifndef PATHED_H
#define PATHED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_ABI_VERSION 1

typedef enum {
    PT_OK = 0,
    PT_ERR_INVALID = -1,    /* bad argument (null pointer, out-of-range id, bad size) */
    PT_ERR_DEVICE = -2,     /* HIP runtime error or no device */
    PT_ERR_MEMORY = -3,     /* allocation failed */
    PT_ERR_STATE = -4,      /* call not valid in the current state (no scene, pass running) */
    PT_ERR_NUMERIC = -5,    /* degenerate input (e.g. camera looks along its up vector) */
    PT_ERR_IO = -6          /* a file could not be read, parsed or written */
} pt_status;

/* Material kinds: values of the reference's MaterialType (include/materials/types.h:5-11). */
typedef enum {
    PT_LAMBERTIAN = 0,
    PT_MIRROR = 1,
    PT_GLASS = 2,
    PT_MICROFACET = 3,
    PT_MATERIAL_NONE = 4
} pt_material_type;

/* Renderer kinds: the reference's RendererType (include/scene_model.h:15-19) plus the
 * guided wavefront.  CUDA and Optix both select the plain wavefront integrator. */
typedef enum {
    PT_RENDERER_CUDA = 0,
    PT_RENDERER_OPTIX = 1,
    PT_RENDERER_NORMALS = 2,
    PT_RENDERER_GUIDED = 3
} pt_renderer_type;

typedef struct pt_scene pt_scene;
typedef struct pt_renderer pt_renderer;

/* ------------------------------------------------------------------------------------ */
/* Errors and version                                                                   */
/* ------------------------------------------------------------------------------------ */
const char *pt_last_error(void);
int pt_abi_version(void);
/* Number of visible HIP devices (0 on a host without a GPU). */
int pt_device_count(int *count);

/* ------------------------------------------------------------------------------------ */
/* Scene description (host memory).  Replaces SceneData + SceneAdapter::createSceneData   */
/* (include/scene_data.h:18-61, src/scene_data.cu:9-118).                                 */
/* ------------------------------------------------------------------------------------ */
int pt_scene_create(pt_scene **out);
int pt_scene_destroy(pt_scene *scene);

/* Appends a material (params.h:15-107 LambertianParams/MirrorParams/GlassParams/
 * MicrofacetParams); albedo/emit are used by Lambertian, ior by Glass, alpha by Microfacet.
 * *out_id receives the material id (dense, in insertion order). */
int pt_scene_add_material(pt_scene *scene, int type, const float albedo[3], const float emit[3],
                          float ior, float alpha, int *out_id);

/* Appends n triangles: positions[9n] (p0 p1 p2), normals[9n] per-vertex shading normals or
 * NULL for flat normals from the winding (obj_parser.cpp:185-194), material_ids[n]. */
int pt_scene_add_triangles(pt_scene *scene, const float *positions, const float *normals,
                           const int *material_ids, int64_t n);

/* Appends n spheres: centers[3n], radii[n], material_ids[n] (primitives/sphere.h:10-63). */
int pt_scene_add_spheres(pt_scene *scene, const float *centers, const float *radii,
                         const int *material_ids, int64_t n);

/* Equirectangular y-up environment, RGBA float texels row-major (rgba[4*w*h]), as loaded
 * by createEnvironmentLight (environment_light.cu:14-56).  sin_theta_weight = 0 keeps the
 * reference's sampling weight r+g+b; 1 multiplies it by sin(theta). */
int pt_scene_set_environment(pt_scene *scene, const float *rgba, int width, int height,
                             int sin_theta_weight);

/* Pinhole camera (core/camera.h:16-61, transform.cu:7-33).  vfov in radians. */
int pt_scene_set_camera(pt_scene *scene, const float origin[3], const float target[3],
                        const float up[3], float vfov, int flip_handedness);

/* Procedural workloads (seeded):
 *   0  Cornell box: 5 walls (10 triangles), 2-triangle ceiling light (17,12,4), diffuse and
 *      glass (ior 1.5) spheres of radius 0.3; the reference's default camera (scene.cu:176-184)
 *   1  displaced icosphere (subdivision level `param` (default 8 -> 1,310,720 triangles),
 *      3-octave value noise of amplitude 0.05), per-face Beckmann alpha U[0.05,0.5] with 30%
 *      Lambert faces, synthetic 1024x512 HDR environment (log-normal sigma 1 + sun lobe)
 *   2  closed room of ~`param` triangles (cubes and ellipsoids, Beckmann alpha U[0.02,0.3])
 *      lit through a 0.3-wide slot by an area light (1e3) and the environment, glass sphere
 *      r = 0.5
 *   3  synthetic environment only (width = param, height = param/2) for sampling tests
 *   4  displaced geodesic sphere of about `param` triangles (default 1e6; 20 n^2 with
 *      n = round(sqrt(param / 20))), displacement, materials, environment and camera as 1
 */
int pt_scene_preset(pt_scene *scene, int preset, uint64_t seed, int64_t param);

typedef struct {
    int64_t n_triangles;
    int64_t n_spheres;
    int64_t n_materials;
    int64_t n_emitters;
    int env_width, env_height;
    int has_normals;
    float camera_origin[3], camera_target[3], camera_up[3];
    float camera_vfov;
    int camera_flip;
    float bounds_lo[3], bounds_hi[3];
} pt_scene_info;
int pt_scene_get_info(const pt_scene *scene, pt_scene_info *info);

/* Copies the scene arrays out (sizes from pt_scene_get_info; any pointer may be NULL):
 * positions[9T], normals[9T] (filled with the flat normals if none were given),
 * tri_mat[T], spheres[4S] (cx cy cz r), sph_mat[S], mat_type[M], mat_albedo[3M],
 * mat_emit[3M], mat_ior[M], mat_alpha[M], env_rgba[4*W*H]. */
int pt_scene_export(const pt_scene *scene, float *positions, float *normals, int *tri_mat,
                    float *spheres, int *sph_mat, int *mat_type, float *mat_albedo,
                    float *mat_emit, float *mat_ior, float *mat_alpha, float *env_rgba);

/* ------------------------------------------------------------------------------------ */
/* Renderer.  Replaces Renderer/PathTracer/OptixTracer (renderer.h:14-35, path_tracer.cu, */
/* optix_tracer.cu) and CUDAGlobals (cuda_globals.cu:10-139).                            */
/* ------------------------------------------------------------------------------------ */
/* Renderer::init (renderer.cpp:10-18) + PathTracer::init (path_tracer.cu:266-283). */
int pt_renderer_create(int width, int height, pt_renderer **out);
int pt_renderer_destroy(pt_renderer *r);

/* CUDAGlobals::initMaterials/mallocWorld/copySceneData (cuda_globals.cu:51-139): uploads the
 * scene, builds the BVH and the environment alias tables on the GPU, resets accumulation. */
int pt_renderer_set_scene(pt_renderer *r, const pt_scene *scene);

/* Scene edits (render_session.cu:82-116 callback).  Each resets accumulation. */
int pt_renderer_set_camera(pt_renderer *r, const float origin[3], const float target[3],
                           const float up[3], float vfov, int flip_handedness);
int pt_renderer_set_material(pt_renderer *r, int material_id, int type, const float albedo[3],
                             const float emit[3], float ior, float alpha);
int pt_renderer_set_max_depth(pt_renderer *r, int max_depth);
int pt_renderer_set_nee(pt_renderer *r, int enabled);
int pt_renderer_set_type(pt_renderer *r, int renderer_type);
int pt_renderer_set_seed(pt_renderer *r, uint64_t seed);

/* Tuning knobs (queue capacity in paths, default 2^26; block counts); values <= 0 restore
 * defaults. */
typedef enum {
    PT_OPT_POOL_PATHS = 0,
    PT_OPT_TRACE_BLOCKS_PER_CU = 1,
    PT_OPT_SHADE_BLOCKS_PER_CU = 2,
    PT_OPT_PROFILE_KERNELS = 3,  /* 1: time each kernel launch with HIP events */
    PT_OPT_CONSISTENT_MIS = 4,   /* 1: the light-sample MIS weight uses the BSDF sampling density
                                    (the reference uses Microfacet::pdf, microfacet.h:124-134,
                                    whose widened alpha differs from sample()'s) */
    PT_OPT_GUIDE_TRAIN_STEPS = 5,/* guided: Adam steps after each chunk (default 16) */
    PT_OPT_GUIDE_BATCH = 6,      /* guided: training records per step (default 2^18) */
    PT_OPT_GUIDE_LR_MICRO = 7,   /* guided: Adam learning rate in units of 1e-6 (default 10000) */
    PT_OPT_GUIDE_RESET = 8,      /* guided: re-initialise the network with seed = value;
                                    get: number of training steps taken */
    PT_OPT_TRACE_LDS_STACK = 9,  /* traversal stack entries per lane kept in LDS (16 or 32);
                                    deeper entries spill to global memory */
    PT_OPT_BATCH_GRAPH = 10,     /* batches only: 1 (default) = capture the whole batch in one
                                    hipGraph and launch it once; 0 = launch kernels directly */
    PT_OPT_GUIDE_MAX_FRACTION_MILLI = 11, /* guided: cap, in thousandths, on the probability of
                                    sampling the learned lobes (the rest samples the BSDF);
                                    default 800; 1000 leaves the network's own clamp (0.98) */
    PT_OPT_TRACE_REFILL_LANES = 12, /* a trace wave takes new rays from the queue once this many
                                    of its 64 lanes are idle (1..64; 64: whole batches of 64
                                    neighbouring rays); 0 (default) = per launch: camera rays
                                    48, later closest-hit queues and shadow rays 40 */
    PT_OPT_GUIDE_BOUNCES = 13,   /* guided: only vertices of the first N bounces draw from the
                                    learned lobes (their network inference and training records);
                                    later vertices sample the BSDF; 0 (default) = every bounce */
    PT_OPT_BVH_BUILDER = 14,     /* binary tree under the 4-wide BVH, used from the next
                                    pt_renderer_set_scene: 1 = LBVH (Karras radix tree),
                                    2 = PLOC (locally-ordered clustering); 0 = the default */
    PT_OPT_BATCH_WAVEFRONT = 15, /* batches only: 1 (default) = one wavefront over every render
                                    (paths of different renders share the queues; chunks hold
                                    sample rows of several renders, a chunk of one render runs
                                    that render's own kernels); 0 = each render's own chunks,
                                    one render after the other */
    PT_OPT_TRACE_STATS = 16      /* 1: instrumented trace kernels count rays, BVH node visits
                                    and primitive tests (pt_render_stats.trace_*); slower, for
                                    measurement only.  Images are unchanged. */
} pt_option;
int pt_renderer_set_option(pt_renderer *r, int option, int64_t value);
int pt_renderer_get_option(pt_renderer *r, int option, int64_t *value);

/* Renderer::renderAsync (path_tracer.cu:285-336): enqueues one progressive pass of `spp`
 * samples per pixel on the renderer's stream and returns at once. */
int pt_renderer_render_async(pt_renderer *r, int spp);
/* Renderer::pollRender (path_tracer.cu:338-356): *finished = 1 once the pass completed;
 * *milliseconds = its GPU time (begin/end events). */
int pt_renderer_poll(pt_renderer *r, int *finished, float *milliseconds);
int pt_renderer_synchronize(pt_renderer *r);
/* The HIP stream (a hipStream_t) that the renderer's passes run on, so that a caller can order
 * its own work (events, copies) after a pass: the events of RenderRecord (render_record.h:7-11)
 * are recorded on it.  Owned by the renderer; valid until pt_renderer_destroy. */
int pt_renderer_get_stream(pt_renderer *r, void **stream);
/* Renderer::reset: accumulation restarts at the next pass. */
int pt_renderer_reset(pt_renderer *r);
/* Renderer::getSpp */
int pt_renderer_get_spp(pt_renderer *r, int64_t *spp);

/* Renderer::getRadianceBuffer (renderer.cpp:20-42): mean radiance, 3*W*H floats,
 * row-major, row 0 = bottom. */
int pt_renderer_get_radiance(pt_renderer *r, float *out);
/* FP64 per-pixel radiance sums over all accumulated samples (3*W*H doubles). */
int pt_renderer_get_radiance_sum(pt_renderer *r, double *out);
/* The uchar4 framebuffer of framebuffer.cu:35-40 (sRGB, 4*W*H bytes). */
int pt_renderer_get_rgba8(pt_renderer *r, uint8_t *out);

/* hitTest (hit_test.cu:49-82): material id under pixel (x, y from the bottom). */
int pt_renderer_hit_test(pt_renderer *r, int x, int y, int *material_id, int *is_hit);

typedef struct {
    int64_t rays_total;        /* extension + shadow rays since the last reset of stats */
    int64_t extension_rays;
    int64_t shadow_rays;
    int64_t paths;
    int64_t passes;
    double gpu_ms;             /* sum of pass times */
    int bvh_depth;
    int64_t bvh_nodes;
    int stack_overflows;       /* traversal stack overflow events (must stay 0) */
    int64_t guide_steps;       /* guided: training steps taken since the network was created */
    double guide_loss;         /* guided: mean loss of the last training step */
    int64_t bounce_rays[16];   /* extension rays traced at bounce 0..15 (0 = camera rays) */
    int64_t bounce_shadow[16]; /* shadow rays traced at bounce 0..15 */
    /* With PT_OPT_TRACE_STATS on: rays traced, BVH node visits and primitive tests of the
     * closest-hit ([0]) and shadow ([1]) kernels (0 otherwise). */
    int64_t trace_rays[2];
    int64_t trace_nodes[2];
    int64_t trace_prims[2];
} pt_render_stats;
int pt_renderer_get_stats(pt_renderer *r, pt_render_stats *stats);
int pt_renderer_reset_stats(pt_renderer *r);

/* Per-kernel timing of the last pass when PT_OPT_PROFILE_KERNELS is on: for kernel k,
 * launches, total ms and algorithmic HBM bytes.  Names: "raygen", "extend", "shade",
 * "shadow", "accumulate", "tonemap". */
typedef struct {
    char name[16];
    int64_t launches;
    int64_t units;            /* rays / segments / paths / pixels processed */
    double ms;
    double bytes;             /* algorithmic bytes (per-unit constants x units) */
} pt_kernel_stat;
int pt_renderer_get_kernel_stats(pt_renderer *r, pt_kernel_stat *out, int capacity, int *count);

/* Radiance of every path of the last pass in path order (sample-major: path = s*W*H + pixel),
 * 3 floats each; only when the pass fit in one chunk (paths <= pool).  For parity tests. */
int pt_renderer_get_path_radiance(pt_renderer *r, float *out, int64_t capacity, int64_t *count);

/* ------------------------------------------------------------------------------------ */
/* Unit-level entry points (parity harness).  Host pointers in and out.                  */
/* ------------------------------------------------------------------------------------ */
/* Closest hit (any_hit = 0) or occlusion (any_hit = 1) for n rays (8 floats each:
 * o.xyz tmin d.xyz tmax) through the renderer's BVH (watertight triangle test), traced by
 * the renderer's own extend and shadow kernels.
 * any_hit = 0: over [tmin, tmax]; out[4n]: t, u, v, primitive (int bits: >= 0 triangle index,
 * < 0 ~sphere index, INT32_MIN miss) -- the indexing of World::hit's PrimitiveIndex
 * (world.cu:5-41); t = -1 on a miss.
 * any_hit = 1: the shadow-ray query of directSampleLights (path_tracer.cu:104-110), over
 * [1e-4, tmax] whatever tmin says; out[4n]: occluded rays 0, 0, 0, primitive 0; unoccluded
 * rays -1, 0, 0, INT32_MIN (which occluder stopped the ray is not reported). */
int pt_renderer_intersect(pt_renderer *r, const float *rays, int64_t n, int any_hit, float *out);

/* Path-pool layout of a pass (host only, no device needed): for a width x height image at `spp`
 * samples per pixel under PT_OPT_POOL_PATHS = pool_option (<= 0: the default), *capacity = the
 * paths the renderer allocates per SoA buffer, *chunks = the pass's chunk count and *max_slot =
 * the highest queue slot any chunk addresses (camera, linear and appended queues).
 * max_slot < capacity for every input.  No reference counterpart (the reference allocates one
 * record per pixel, path_tracer.cu:266-283). */
int pt_pool_layout(int width, int height, int spp, int64_t pool_option, int64_t *capacity, int64_t *chunks,
                   int64_t *max_slot);

/* Single ray/triangle tests on the GPU (triangle.cu:7-48): tris[9n], rays[8n] -> out[4n]
 * (hit flag, t, u, v).  mode 0 = Moeller-Trumbore in the reference's operation order,
 * mode 1 = watertight.  Built without fused multiply-adds for bit-level comparison. */
int pt_unit_triangle_test(const float *tris, const float *rays, int64_t n, int mode, float *out);

/* BSDF f/pdf (bsdf.h:20-61) and sample (bsdf.h:87-115) on the GPU.  mat[6n] = type, albedo
 * rgb, ior, alpha; wo/wi[3n]; xi[2n].  eval: f[3n], pdf[n]; sample: out[8n] = wi[3], pdf,
 * f[3], is_delta. */
int pt_unit_bsdf_eval(const float *mat, const float *wo, const float *wi, int64_t n, float *f,
                      float *pdf);
int pt_unit_bsdf_sample(const float *mat, const float *wo, const float *xi, int64_t n,
                        float *out);

/* Environment tables built on the GPU in FP64 (alias per row + marginal) for an RGBA map,
 * copied back: row_prob[w*h] (double), row_alias[w*h], marg_prob[h], marg_alias[h],
 * texel_pdf[w*h] (pmf * w*h / (2 pi^2), float). */
int pt_unit_env_tables(const float *rgba, int width, int height, int sin_theta_weight,
                       double *row_prob, int *row_alias, double *marg_prob, int *marg_alias,
                       float *texel_pdf);
/* The tables of pt_unit_env_tables without any device allocation, for callers that cannot
 * allocate: code under HIP graph capture, serving loops, harnesses that forbid hipMalloc.  No
 * counterpart in the reference (environment_light.cu:14-56 builds its CMF tables on the host).
 * All pointers are device memory of the current HIP device: rgba[4*w*h] (RGBA float); outputs
 * rows[2*w*h] (per texel: row alias probability as float, alias index as int bits),
 * marginal[2*h] (the same over rows), texel_pdf[w*h] (pmf * w*h / (2 pi^2)).  The FP64
 * working arrays live in one fixed array per device (about 42 MB, held from the library's
 * first use on a device until the process ends), so w*h <= PT_ENV_NOALLOC_MAX_TEXELS and
 * h <= PT_ENV_NOALLOC_MAX_ROWS; larger maps fail with PT_ERR_INVALID.  Enqueued on `stream`
 * (a hipStream_t; NULL = the null stream) and returns without waiting.  Calls that overlap on
 * one device share that array and race: serialise them (one stream, or events). */
#define PT_ENV_NOALLOC_MAX_TEXELS (1 << 21)
#define PT_ENV_NOALLOC_MAX_ROWS (1 << 12)
int pt_env_tables_noalloc(const float *rgba, int width, int height, int sin_theta_weight,
                          float *rows, float *marginal, float *texel_pdf, void *stream);
/* Environment sampling and pdf on the GPU: xi[2n] -> dir[3n], pdf[n], emit[3n];
 * pdf_dirs[3m] -> pdf_out[m]. */
int pt_unit_env_sample(const float *rgba, int width, int height, int sin_theta_weight,
                       const float *xi, int64_t n, float *dir, float *pdf, float *emit,
                       const float *pdf_dirs, int64_t m, float *pdf_out);

/* The framebuffer bytes of framebuffer.cu:35-40 / color.h:7-28 for n radiance triples
 * (radiance[3n]) through the renderer's tonemap device function: rgba[4n] (alpha 255). */
int pt_unit_tonemap(const float *radiance, int64_t n, uint8_t *rgba);

/* PCG32 path streams on the GPU: first k uniforms of (seed, pixel, sample) for n queries. */
int pt_unit_path_uniforms(uint64_t seed, const uint32_t *pixel, const uint32_t *sample, int64_t n,
                          int k, float *out);

/* Guiding network on the GPU (BF16 MFMA, FP32 accumulate).  params: PT_GUIDE_PARAMS floats,
 * 5 layers of (W[64 out][64 in], b[64]); the last layer's rows 48..63 are padding.
 * forward: features[8n] raw inputs in [0, 1] -> out[48n] raw outputs.
 * gradients: records[16n] = features[8], omega[3], bsdf pdf, -, -, -, weight -> grad of the
 * mean loss (PT_GUIDE_PARAMS floats) and the mean loss.
 * adam: one Adam step (0.9, 0.999, 1e-8) in place on params/m/v.
 * mfma_tile: c[32x32] = a[32x16] x b[16x32] (row-major) through one v_mfma_f32_32x32x16_bf16,
 * the lane layout the network kernels rely on. */
#define PT_GUIDE_PARAMS 20800
int pt_unit_guide_forward(const float *params, const float *features, int64_t n, float *out);
/* The inference kernel's mixture records for n feature rows (16 uint32 words per row, the
 * record the guided shade reads: per lobe k, word 2k = octahedral unit mean as two unorm16
 * (u low, v high), word 2k+1 = kappa (FP16, low half) and the lobe weight (FP16, high half;
 * for lobe 7 the guided-sampling probability instead, lobe 7's weight being 1 minus the
 * others).  No counterpart in the reference (learned guiding); for the parity tests. */
int pt_unit_guide_records(const float *params, const float *features, int64_t n, uint32_t *records);
int pt_unit_guide_gradients(const float *params, const float *records, int64_t n, float *grad, float *loss);
int pt_unit_guide_adam(float *params, float *m, float *v, const float *grad, int step, float lr);
int pt_unit_mfma_tile(const float *a, const float *b, float *c);
/* One v_mfma_f32_16x16x32_bf16 with the lane layout of the guide inference kernel:
 * c[16x16] = bf16(a[16x32]) * bf16(b[32x16]), row-major FP32. */
int pt_unit_mfma_tile16(const float *a, const float *b, float *c);
/* von Mises-Fisher lobes as the guided vertices use them: for n queries (mean mu[3n], unit;
 * concentration kappa[n]; uniforms xi[2n]) -> sampled directions dir[3n] and the density
 * of the lobe at them, pdf[n]. */
int pt_unit_vmf_sample(const float *mu, const float *kappa, const float *xi, int64_t n, float *dir, float *pdf);

/* ---------------------------------------------------------------------------------------
 * Scene and image files (host only; no device needed).
 * ------------------------------------------------------------------------------------- */

/* Wavefront OBJ (+ MTL) appended to the scene, as src/parsers/obj_parser.cpp:20-269 and
 * src/scene_data.cu:20-52 read it: v, vn, f (quads of plain indices split (0,1,2) (0,2,3);
 * triangles of v/t/n or v//n corners; triangles of plain indices are read as well), negative
 * indices relative to the end, mtllib / usemtl.  The file's materials (newmtl, Kd, Ke;
 * src/parsers/mtl_parser.cpp:15-109) become Lambertians appended in name order; faces of a
 * file without materials take default_material.  first_material (may be NULL): id of the
 * first appended material, or -1. */
int pt_scene_load_obj(pt_scene *scene, const char *path, int default_material, int *first_material);
/* Binary little-endian PLY (float x y z vertices, uchar-count int-index triangles), every
 * face with `material` (src/parsers/ply_parser.cpp:9-112). */
int pt_scene_load_ply(pt_scene *scene, const char *path, int material);
/* Environment map from an OpenEXR file (src/lights/environment_light.cu:14-56): single-part
 * scanline files with NONE, ZIPS or ZIP compression and HALF / FLOAT / UINT R, G, B [, A]
 * channels (PIZ and other codecs return PT_ERR_IO). */
int pt_scene_load_exr_environment(pt_scene *scene, const char *path, int sin_theta_weight);
/* The reference's built-in scenes 0..3 (src/scene.cu:14-176: Cornell box OBJ, glossy Cornell
 * box with a mirror box and a glass ball, plane + 1-pixel environment, Veach MIS plates) read
 * from `scenes_dir` (the reference's scenes/ folder layout); the scene must be empty. */
int pt_scene_reference(pt_scene *scene, int index, const char *scenes_dir);
/* OpenEXR images: size; RGBA floats (4*w*h, scanline order, alpha 1 when absent);
 * writing an RGB radiance buffer (3*w*h, row 0 = bottom, as pt_renderer_get_radiance) as
 * uncompressed B, G, R scanlines, HALF when `half` else FLOAT (src/io/image.cpp:10-78). */
int pt_exr_info(const char *path, int *width, int *height);
int pt_exr_read_rgba(const char *path, float *rgba, int64_t capacity);
int pt_exr_write_rgb(const char *path, const float *radiance, int width, int height, int half);

/* ---------------------------------------------------------------------------------------
 * Autotuning and batches of renders.  No counterpart in the reference (which renders one
 * session at a time, render_session.h:59-69, with a fixed megakernel launch shape,
 * path_tracer.cu:295-311); these serve callers that render many views or scenes at once.
 * ------------------------------------------------------------------------------------- */

/* Tuned launch configuration: path-pool capacity, trace/shade blocks per CU, LDS stack
 * entries, and the rate (primary + extension + shadow rays per microsecond = Mrays/s) of
 * the best probe. */
typedef struct {
    int64_t pool_paths;
    int trace_blocks_per_cu;
    int shade_blocks_per_cu;
    int lds_stack;
    double mrays_per_s;
    int probes;
} pt_tune_result;

/* Coordinate descent over the four knobs, each candidate timed on a pass of `spp` samples
 * (the probe samples are discarded: accumulation restarts at the next pass); the best
 * configuration stays set on the renderer. */
int pt_renderer_autotune(pt_renderer *r, int spp, pt_tune_result *out);

/* A batch: independent renders (own scene, size, spp, depth; plain NEE + MIS) that share one
 * path pool and one stream.  pt_batch_run queues them shortest-first (pixels x spp x log2
 * triangles) and, with PT_OPT_BATCH_GRAPH on, captures the whole batch (per render:
 * accumulator clear, per chunk raygen, depth x (extend, shade, shadow), accumulate, and a
 * completion stamp from the device wall clock) in one hipGraph that is launched once and
 * re-launched while the batch and its options are unchanged. */
typedef struct pt_batch pt_batch;
int pt_batch_create(pt_batch **out);
int pt_batch_destroy(pt_batch *b);
/* The scene is copied to the device; it may be destroyed afterwards.  index: position of the
 * render in the batch. */
int pt_batch_add(pt_batch *b, const pt_scene *scene, int width, int height, int spp, int max_depth, int *index);
/* PT_OPT_POOL_PATHS, PT_OPT_TRACE_BLOCKS_PER_CU, PT_OPT_SHADE_BLOCKS_PER_CU,
 * PT_OPT_TRACE_LDS_STACK, PT_OPT_BATCH_GRAPH */
int pt_batch_set_option(pt_batch *b, int option, int64_t value);
int pt_batch_get_option(pt_batch *b, int option, int64_t *value);
/* Tunes the batch's options on probe runs of the same batch with spp / probe_divisor. */
int pt_batch_autotune(pt_batch *b, int probe_divisor, pt_tune_result *out);
/* Renders every render of the batch from scratch; blocks until done. */
int pt_batch_run(pt_batch *b);
/* Mean radiance of render `index` after a run: 3 floats per pixel, row 0 = bottom. */
int pt_batch_get_radiance(pt_batch *b, int index, float *out);
typedef struct {
    int width, height, spp;
    int64_t triangles;
    int64_t rays;        /* primary + extension + shadow rays of the last run */
    double latency_ms;   /* device time from the start of the last run to this render's end */
} pt_batch_render_stats;
int pt_batch_get_render_stats(pt_batch *b, int index, pt_batch_render_stats *out);
typedef struct {
    int renders;
    int chunks;
    int64_t graph_nodes;   /* nodes of the captured graph (0: direct launches) */
    int64_t rays;
    double wall_ms;        /* host wall time of the last run's launch to completion */
    int stack_overflows;
} pt_batch_stats;
int pt_batch_get_stats(pt_batch *b, pt_batch_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* PATHED_H */
