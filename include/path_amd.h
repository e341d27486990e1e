/*
 * path_amd: C ABI of the MI355X path tracer.
 *
 * Plain pointers and sizes only. Every function returns PT_OK (0) or a
 * non-zero status; pt_last_error() returns the message of the last failure on
 * the calling thread. The reference aborts on these conditions
 * (include/lib/assert.h:14-50, include/lib/cuda/utils.h:18-26); returning a
 * status instead is a deliberate deviation so callers and tests can observe
 * them.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   scene::Scene + ScenefileLoader   include/scene/scene.h:29-130,
 *                                    src/scene/scenefile_compat/scenefile_loader.cpp:23-319
 *   render::Settings + config_io     include/render/settings.h:27-70,
 *                                    src/render/config_io.cpp:14-34
 *   render::Renderer::render         include/render/renderer.h:43-70
 *   RendererFromFiles                include/render/renderer_from_files.h:22-43
 */
#ifndef PATH_AMD_H
#define PATH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pt_status {
  PT_OK = 0,
  PT_ERR_INVALID = 1,     /* bad argument / unsupported combination */
  PT_ERR_IO = 2,          /* file not found or unparsable */
  PT_ERR_HIP = 3,         /* HIP runtime error */
  PT_ERR_NO_DEVICE = 4,   /* no GPU visible */
  PT_ERR_UNSUPPORTED = 5, /* setting accepted by the reference but not built */
};

const char *pt_last_error(void);

/* ------------------------------------------------------------------------ */
/* Materials: POD mirror of bsdf::UnionBSDF + emission after the MTL        */
/* conversion (include/scene/tinyobj_material_conversion.h:13-101).         */
/* ------------------------------------------------------------------------ */
enum pt_bsdf_type {
  PT_BSDF_DIFFUSE = 0,
  PT_BSDF_GLOSSY = 1,
  PT_BSDF_DIFFUSE_GLOSSY = 2,
  PT_BSDF_DIFFUSE_MIRROR = 3,
  PT_BSDF_MIRROR = 4,
  PT_BSDF_DIELECTRIC_REFRACTIVE = 5,
};

typedef struct pt_material {
  uint32_t type;
  float diffuse[3];
  float specular[3];
  float shininess;
  float ior;
  float diffuse_weight; /* Combined lobe weights (DiffuseGlossy/DiffuseMirror) */
  float glossy_weight;
  float emission[3];
} pt_material;

/* ------------------------------------------------------------------------ */
/* Scene (scene::Scene)                                                      */
/* ------------------------------------------------------------------------ */
typedef struct pt_scene pt_scene;

/* Load a CS123 XML scene with its OBJ/MTL meshes (ScenefileLoader::load_scene,
 * src/scene/scenefile_compat/scenefile_loader.cpp:23-62). film_to_world is the
 * camera transform (row-major 3x4) for the given aspect ratio. */
int pt_scene_load(const char *path, float width_height_ratio, int quiet,
                  pt_scene **out, float film_to_world[12]);
/* Empty scene; meshes are added with pt_scene_add_mesh. */
int pt_scene_create(pt_scene **out);
/* Add one mesh + one instance of it. verts: n_tris*9 object-space vertices;
 * normals: n_tris*9 vertex normals or NULL for face normals; tri_mats: index
 * into mats; transform: row-major 3x4 object_to_world or NULL for identity.
 * Emissive clusters are derived as in scenefile_loader.cpp:181-276. */
int pt_scene_add_mesh(pt_scene *scene, const float *verts, const float *normals,
                      const uint32_t *tri_mats, uint32_t n_tris,
                      const pt_material *mats, uint32_t n_mats,
                      const float *transform);
/* Add a mesh from an OBJ file (tinyobjloader-compatible triangulation). */
int pt_scene_add_obj(pt_scene *scene, const char *path, const float *transform);
void pt_scene_destroy(pt_scene *scene);
/* Flattened world-space sizes (flat_triangle.h:79-146). */
int pt_scene_counts(const pt_scene *scene, uint32_t *n_tris,
                    uint32_t *n_materials, uint32_t *n_light_tris);
/* Flattened world-space arrays; any pointer may be NULL.
 * tris: n_tris*9 (v0, e1 = v1 - v0, e2 = v2 - v0); normals: n_tris*9;
 * light_tris: n_light_tris*9; light_cdf: n_light_tris (inclusive, last = 1). */
int pt_scene_flat(const pt_scene *scene, float *tris, float *normals,
                  uint32_t *tri_mat, pt_material *mats, float *light_tris,
                  float *light_cdf);
/* get_camera_transform (src/scene/camera.cpp:4-27). */
int pt_camera_from_look(const float look[3], const float up[3],
                        const float pos[3], float height_angle_deg,
                        float width_height_ratio, float film_to_world[12]);

/* ------------------------------------------------------------------------ */
/* Settings: POD mirror of render::Settings (include/render/settings.h).     */
/* ------------------------------------------------------------------------ */
typedef struct pt_work_division_settings {
  uint32_t block_size;
  uint32_t target_x_block_size;
  uint32_t force_target_samples;
  uint32_t forced_target_samples_per_thread;
  uint32_t base_num_threads;
  float samples_per_thread_scaling_power;
  uint32_t max_samples_per_thread;
} pt_work_division_settings;

enum pt_kernel_approach { PT_MEGA_KERNEL = 0, PT_STREAMING = 1 };
enum pt_accel_type {
  PT_ACCEL_LOOP_ALL = 0,
  PT_ACCEL_NAIVE_PARTITION_BVH = 1,
  PT_ACCEL_SBVH = 2,
  PT_ACCEL_DIRECTION_GRID = 3,
};
enum pt_light_sampler { PT_LIGHT_NO_LIGHT_SAMPLING = 0, PT_LIGHT_RANDOM_TRIANGLE = 1 };
enum pt_dir_sampler { PT_DIR_UNIFORM = 0, PT_DIR_BSDF = 1 };
enum pt_term_prob {
  PT_TERM_CONSTANT = 0,
  PT_TERM_N_ITERS = 1,
  PT_TERM_MULTIPLIER_FUNC = 2,
  PT_TERM_NORMALIZE = 3,
};
enum pt_rng { PT_RNG_UNIFORM = 0, PT_RNG_SOBEL = 1 };

typedef struct pt_settings {
  uint32_t kernel_approach; /* pt_kernel_approach */
  /* MegaKernelSettings::ComputationSettings (mega_kernel_settings.h:9-24) */
  uint32_t max_blocks_per_launch_gpu;
  uint32_t max_blocks_per_launch_cpu;
  pt_work_division_settings render_work_division;
  pt_work_division_settings reduce_work_division;
  /* StreamingSettings::ComputationSettings (streaming_settings.h:9-16) */
  pt_work_division_settings init_samples_division;
  uint32_t max_num_samples_per_launch;
  uint32_t min_num_init_blocks;
  uint32_t to_bulk_max_size; /* ToBulkSettings::max_size (to_bulk.h:10-14) */
  /* accel (IndividuallyIntersectableSettings, sbvh/settings.h) */
  uint32_t accel; /* pt_accel_type */
  float overlap_threshold;
  uint32_t use_spatial_splits;
  float traversal_per_intersect_cost;
  uint32_t target_objects;
  uint32_t print_stats;
  /* light sampler */
  uint32_t light_sampler;
  uint32_t binary_search_threshold;
  /* dir sampler */
  uint32_t dir_sampler;
  /* term prob */
  uint32_t term_prob;
  float constant_prob;
  uint32_t n_iters;
  float multiplier_func_exp;
  float multiplier_func_min_prob;
  float normalize_min_prob;
  /* rng (RngFromSequenceGenSettings::max_sample_size) */
  uint32_t rng;
  uint32_t max_sample_size;
  /* RenderingEquationSettings */
  uint32_t back_cull_emission;
} pt_settings;

/* Reference defaults (the initializers of the settings structs). */
void pt_settings_default(pt_settings *s);
/* Load the reference YAML schema (configs/benchmark.yaml; enum names as in
 * the reference). Unknown enum names are an error. */
int pt_settings_load_yaml(const char *path, pt_settings *s);
/* print_config: write the YAML form into buf (NUL-terminated, truncated to
 * cap); *needed receives the full length. */
int pt_settings_to_yaml(const pt_settings *s, char *buf, size_t cap,
                        size_t *needed);

/* ------------------------------------------------------------------------ */
/* Renderer (render::Renderer)                                               */
/* ------------------------------------------------------------------------ */
typedef struct pt_renderer pt_renderer;

enum pt_sample_spec_type {
  PT_SAMPLE_SQUARE_IMAGE = 0,       /* SquareImage{x_dim, y_dim, film_to_world} */
  PT_SAMPLE_INITIAL_RAYS = 1,       /* Span<const Ray> */
  PT_SAMPLE_INITIAL_IDX_AND_DIR = 2 /* Span<const InitialIdxAndDirSpec> */
};

typedef struct pt_sample_spec {
  uint32_t type;
  uint32_t x_dim, y_dim;
  float film_to_world[12];
  const float *rays; /* INITIAL_RAYS / INITIAL_IDX_AND_DIR: n_rays*6 (o, d) */
  const uint32_t *initial_idx; /* INITIAL_IDX_AND_DIR: flattened triangle idx */
  uint32_t n_rays;
} pt_sample_spec;

enum pt_output_type {
  PT_OUTPUT_BGRA = 0,     /* uint32 per pixel, memory order B,G,R,0 */
  PT_OUTPUT_FLOAT_RGB = 1, /* float[3] per pixel */
  PT_OUTPUT_PER_STEP = 2,  /* n_steps float[3] images, one per path vertex */
  PT_OUTPUT_DEVICE_FLOAT_RGB = 3, /* float[3] per pixel in device memory of the
                                     renderer's GPU (no reference counterpart:
                                     the image stays in HBM for GPU consumers) */
};

typedef struct pt_output {
  uint32_t type;
  void *data;         /* BGRA / FLOAT_RGB: host buffer of x_dim*y_dim pixels;
                         DEVICE_FLOAT_RGB: device buffer of as many pixels */
  float **per_step;   /* PER_STEP: n_steps host buffers of float[3] pixels */
  uint32_t n_steps;
} pt_output;

int pt_renderer_create(int device, pt_renderer **out);
void pt_renderer_destroy(pt_renderer *r);

/* Renderer::render (include/render/renderer.h:51-54). *seconds receives the
 * render time excluding the acceleration structure and sampler build, as in
 * general_render_impl.h:174. */
int pt_render(pt_renderer *r, const pt_sample_spec *spec, const pt_output *out,
              const pt_scene *scene, uint32_t samples_per,
              const pt_settings *settings, int show_progress, int show_times,
              double *seconds);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): a MegaKernel SquareImage render into `device_rgb`
 * (x_dim * y_dim * 3 floats of device memory on the renderer's GPU) that
 * allocates nothing and uploads nothing. The scene and, for the Sobel rng,
 * the Sobol table for (max_sample_size, samples_per) must already be
 * resident: an earlier pt_render of the same scene with the same samples_per
 * and settings leaves them there; otherwise PT_ERR_INVALID. Scratch (the ray
 * counter and the per-sample-block partial sums) is a 16 MiB device array per
 * GPU: renders whose pixels x sample blocks exceed 1,398,080 (only when a
 * pixel's samples span several sample blocks) are rejected with
 * PT_ERR_INVALID. All calls on one GPU share that array; calls that overlap
 * on one GPU race. Same image bits as pt_render. */
int pt_render_noalloc(pt_renderer *r, const pt_sample_spec *spec, float *device_rgb,
                      const pt_scene *scene, uint32_t samples_per,
                      const pt_settings *settings, double *seconds);

/* ------------------------------------------------------------------------ */
/* Lower-level entry points of the same path (tests, benchmarks, tools).     */
/* ------------------------------------------------------------------------ */
/* Closest hit of n host rays (o, d: 6 floats each) against the scene's GPU
 * BVH (BVH::intersect_objects, include/intersect/accel/detail/bvh/bvh_impl.h:12-89).
 * idx = flattened triangle index or 0xffffffff for a miss. */
int pt_intersect(pt_renderer *r, const pt_scene *scene, const float *rays,
                 uint32_t n, float *t, uint32_t *idx, uint32_t *back);
/* Sobol table vals[sample * dim_bound + dim] computed on the GPU
 * (replaces SobelSequenceGen::gen, src/rng/sobel/sobel.cu:19-93). */
int pt_sobol_table(pt_renderer *r, uint32_t dim_bound, uint32_t count,
                   float *out);
/* The Joe-Kuo direction numbers used: dim_bound*32 words. */
int pt_sobol_direction_numbers(uint32_t dim_bound, uint32_t *out);
/* Work division (src/kernel/work_division.cpp:12-69) as 11 words:
 * block_size, base_samples_per_thread, n_threads_per_unit_extra,
 * sample_block_size, x_block_size, y_block_size, num_sample_blocks,
 * num_x_blocks, num_y_blocks, x_dim, y_dim. */
int pt_work_division(const pt_work_division_settings *s, uint32_t samples_per,
                     uint32_t x_dim, uint32_t y_dim, uint32_t out[11]);
/* Stable radix sort of n (key, value) pairs by key bits [0, end_bit) on the
 * GPU (host arrays in and out). No reference counterpart (the reference
 * builds its trees on the host): it is the sort of the device tree builder
 * (pt_set_bvh_builder), which orders triangles by (scene, Morton code) keys,
 * exposed so that it can be checked on its own. */
int pt_sort_pairs(pt_renderer *r, const uint64_t *keys, const uint32_t *vals, uint32_t n,
                  uint32_t end_bit, uint64_t *keys_out, uint32_t *vals_out);
/* The renderer's per-pixel sample reduction on given values: each
 * work-division thread sums values[sample + location * samples_per] over its
 * samples in order, then a location's threads are combined by the same lane
 * reduction the mega-kernel uses; out[location] for num_locations x 1
 * "pixels". The division must have one sample block. Replaces the reduction
 * the reference's reduce test drives (RuntimeConstantsReducer,
 * include/kernel/runtime_constants_reducer.h:24-133, test
 * tests/src/reduce.cu:49-158). */
int pt_reduce_samples(pt_renderer *r, const pt_work_division_settings *s,
                      const float *values, uint32_t num_locations, uint32_t samples_per,
                      float *out);
/* Host SAH BVH of the flattened scene in the reference node layout
 * (node.h:22-126): per node lo[3], hi[3] (float) and is_leaf, a, b (uint32);
 * a leaf covers perm[a..b), perm[n_refs] holding triangle indices (n_refs =
 * n_tris, or more when spatial splits reference a triangle from several
 * leaves). This is the binary tree the GPU's 4-wide tree is collapsed from,
 * built with the accel settings of the scene's last render (or
 * pt_scene_set_accel). Pass NULL arrays to query n_nodes and n_refs only. */
int pt_scene_bvh(const pt_scene *scene, uint32_t *n_nodes, uint32_t *n_refs, float *lo,
                 float *hi, uint32_t *is_leaf, uint32_t *a, uint32_t *b, uint32_t *perm,
                 uint32_t *max_depth);
/* Tree build settings for the scene from the accel fields of settings
 * (traversal_per_intersect_cost, use_spatial_splits, overlap_threshold;
 * include/intersect/accel/sbvh/settings.h:9-19). Renders apply their own
 * settings the same way; a change rebuilds the tree on next use. A scene
 * must not be rendered from two threads at once with different accel
 * settings (the tree is rebuilt in place). */
int pt_scene_set_accel(pt_scene *scene, const pt_settings *settings);
/* One node of the GPU tree from n (1-4) child boxes (lo, hi: n*3 floats) and
 * child refs (>= 0 inner node, < 0 ~first triangle of a leaf): the 64-byte
 * record the traversal reads (node), with the boxes it decodes for the four
 * slots (decoded_lo / decoded_hi: 12 floats each, optional). Replaces the
 * reference's packed node value (include/intersect/accel/detail/bvh/node.h:
 * 22-126) in this build's 4-wide layout; PT_ERR_INVALID if a decoded box
 * would not contain its child's box. */
int pt_pack_node(const float *lo, const float *hi, uint32_t n, const int32_t *child,
                 void *node, float *decoded_lo, float *decoded_hi);
/* Bounds of the part of triangle tri (3 vertices) inside the box [lo, hi]
 * (empty, out_lo > out_hi, when they do not meet): the clipping a spatial
 * split applies to a triangle reference (ClippedTriangle::new_bounds,
 * include/intersect/accel/detail/clipped_triangle_impl.h:13-124). */
int pt_clip_triangle_bounds(const float tri[9], const float lo[3], const float hi[3],
                            float out_lo[3], float out_hi[3]);

/* ------------------------------------------------------------------------ */
/* Batched rendering for training-data generation. No reference counterpart */
/* renders several scenes per launch: the reference calls Renderer::render  */
/* once per scene (src/generate_data/full_scene/generate_data.cpp:280-294). */
/* ------------------------------------------------------------------------ */
/* Seeded random scene: 2-40 boxes/spheres, triangle budget log-uniform in
 * [20, 20000], diffuse albedos U[0.05, 0.95], 1-3 quad emitters of radiance
 * U[5, 30], square resolution from {32, 64, 96, 128}. */
int pt_scene_generate(uint64_t seed, pt_scene **out, float film_to_world[12],
                      uint32_t *resolution);

typedef struct pt_batch_item {
  const pt_scene *scene;
  uint32_t x_dim, y_dim;
  float film_to_world[12];
  float *out_rgb;     /* x_dim*y_dim*3 floats (host) */
  float *out_gbuffer; /* NULL or x_dim*y_dim*12 floats (host): first hit of
                         sample 0 per pixel - position, shading normal, albedo,
                         emission; zeros where the camera ray misses */
} pt_batch_item;

/* Render every item with the wavefront integrator in one launch sequence
 * (all scenes share the path pool). Each image is bit-identical to
 * pt_render of that scene alone with kernel_approach = Streaming. *seconds
 * covers scene packing, upload, render and download (BVH builds run on a
 * host thread pool and are included). */
int pt_render_batch(pt_renderer *r, const pt_batch_item *items, uint32_t n,
                    uint32_t samples_per, const pt_settings *settings,
                    double *seconds);

/* Ray sample specs of many scenes in one launch sequence: the batched form
 * of pt_render with an InitialRays or InitialIdxAndDir sample spec and a
 * FloatRGB (n_steps = 0) or OutputPerStep (n_steps images) output, which the
 * reference's training-data generator calls once per scene
 * (src/generate_data/full_scene/generate_data.cpp:280-294). Each item's values
 * are bit-identical to pt_render of that item alone with the MegaKernel
 * integrator (same work division, sample sequence and reduction order). All
 * items are InitialRays (initial_idx NULL) or all are InitialIdxAndDir; items
 * with n_rays = 0 are skipped. */
typedef struct pt_ray_batch_item {
  const pt_scene *scene;
  const float *rays;           /* n_rays * 6 floats (origin, direction), host */
  const uint32_t *initial_idx; /* NULL, or n_rays flattened triangle indices: the
                                  first intersection is that triangle at
                                  distance 1 (initial_ray_sample.h:58-75) */
  uint32_t n_rays;
  float *out;                  /* host, max(n_steps, 1) * n_rays * 3 floats,
                                  step-major (the per-step images in order) */
} pt_ray_batch_item;
int pt_render_rays_batch(pt_renderer *r, const pt_ray_batch_item *items, uint32_t n,
                         uint32_t samples_per, uint32_t n_steps,
                         const pt_settings *settings, double *seconds);

/* Launch parameters of the wavefront integrator. Results never depend on
 * them. The path pool (max_num_samples_per_launch paths) runs as
 * independent pipelines (pt_set_pipelines) of an equal share of the pool
 * each, on their own streams driven by their own host threads, so one's
 * shading overlaps another's traversal.
 * trace_blocks_per_cu: traversal grid per pipeline = CUs x this x 256
 * threads (default 3 for a render, 6 for a batch; 0 restores the
 * defaults); iters_per_sync: iterations enqueued between host
 * checks; use_graphs: replay those iterations as one captured hipGraph. */
int pt_set_launch_params(pt_renderer *r, uint32_t trace_blocks_per_cu,
                         uint32_t iters_per_sync, int use_graphs);
/* Tree builder of pt_render_batch (no reference counterpart: the reference
 * builds its SBVH on the host, src/intersect/accel/sbvh/detail/generator.cu).
 * PT_BVH_HOST_SAH (default): each scene's SAH tree is built on host threads
 * and cached with the scene. PT_BVH_GPU_LBVH / PT_BVH_GPU_PLOC: all trees of
 * a batch are built on the GPU in one pass over Morton-sorted triangles (the
 * binary radix tree, or locally-ordered clustering by surface area, whose
 * trees trace faster; a batch with a scene above 2^18 triangles uses the
 * radix tree), collapsed to the same 4-wide quantised nodes; a batch with a
 * tree deeper than the traversal stack allows falls back to the host
 * builder. Images do not depend on the tree. */
enum { PT_BVH_HOST_SAH = 0, PT_BVH_GPU_LBVH = 1, PT_BVH_GPU_PLOC = 2 };
/* Independent wavefront pipelines a Streaming render splits its path pool
 * into (1..4, default 3; each at least 256 paths), each on its own stream and
 * host thread. Results never depend on it. */
int pt_set_pipelines(pt_renderer *r, uint32_t n);
/* MegaKernel launch shape (no reference counterpart; the reference runs one
 * lane per work-division thread, kernel_launch_impl_gpu.cuh:12-33). enable
 * (default 1): where a pixel's threads fit one wave and there is no per-step
 * output, four lanes share each work-division thread's samples and the
 * thread's sum is formed in sample order afterwards. Results never depend on
 * it. */
int pt_set_megakernel_split(pt_renderer *r, int enable);
/* Device memory on the renderer's GPU for PT_OUTPUT_DEVICE_FLOAT_RGB (no
 * reference counterpart): allocate, free, and copy back to the host. */
int pt_device_alloc(pt_renderer *r, size_t bytes, void **out);
int pt_device_free(pt_renderer *r, void *ptr);
int pt_device_read(pt_renderer *r, const void *src, void *host_dst, size_t bytes);
int pt_device_write(pt_renderer *r, void *dst, const void *host_src, size_t bytes);
/* d_image = 2 (image - target) for n floats, all in device memory of the
 * renderer's GPU: the image gradient of an L2 loss, ready for
 * pt_render_adjoint, whose d_image may be device memory too (no reference
 * counterpart; the gradient step stays in HBM). */
int pt_l2_loss_gradient(pt_renderer *r, const float *image, const float *target, uint64_t n,
                        float *d_image);
int pt_set_bvh_builder(pt_renderer *r, int builder);
/* Wall time (ms, host clock around the device work) and levels of the 4-wide
 * trees of the last GPU tree build. */
int pt_last_bvh_build(const pt_renderer *r, double *ms, uint32_t *levels);
/* Host wall time (ms) of the phases of the last batch whose scenes were not
 * already resident: ms[0] scene preparation (flattening, shading records and,
 * with host trees, the SAH builds; 16 host threads), ms[1] packing into the
 * pinned staging buffer, ms[2] upload, ms[3] GPU tree builds (0 with host
 * trees). */
int pt_last_batch_phases(const pt_renderer *r, double ms[4]);
/* Time candidate (path pool, traversal grid, iterations per sync, pipelines)
 * on the given items, keep the fastest on the renderer; out[5] = path pool,
 * blocks per CU, iterations per sync, graphs used, pipelines. */
int pt_autotune_batch(pt_renderer *r, const pt_batch_item *items, uint32_t n,
                      uint32_t samples_per, pt_settings *settings,
                      uint32_t out[5]);

/* ------------------------------------------------------------------------ */
/* Differentiable rendering. No reference counterpart (the reference has no  */
/* gradients); the forward pass is pt_render's mega-kernel estimator.        */
/* ------------------------------------------------------------------------ */
/* Replace material `idx` (geometry, BVH and the emitter list are kept; a
 * material may not switch between emissive and non-emissive). Renderers
 * re-upload only the material table and the emitter CDF. */
int pt_scene_set_material(pt_scene *scene, uint32_t idx, const pt_material *m);

/* Path-replay adjoint of a SquareImage render with the same scene, camera,
 * spp and settings: given d_image = dLoss/dImage (host, y_dim*x_dim*3 floats,
 * FloatRGB layout), writes dLoss/d(albedo color) and dLoss/d(emission) for
 * every scene material (host, n_materials*3 floats each). The albedo color is
 * `diffuse` for Diffuse materials and `specular` for Glossy, Mirror and
 * DielectricRefractive ones. Paths are recomputed from their Sobol
 * coordinates; nothing is stored per vertex. Requirements: term_prob NIters or
 * Constant, no two-lobe BSDFs, at most 40 materials. Emission gradients hold
 * the emitter list and its selection distribution fixed (exact when one
 * material emits; for a non-emissive material only the hits its surface gets
 * from BSDF-sampled rays count). Channels whose albedo is 0 report 0. */
int pt_render_adjoint(pt_renderer *r, const pt_scene *scene,
                      const float film_to_world[12], uint32_t x_dim,
                      uint32_t y_dim, uint32_t samples_per,
                      const pt_settings *settings, const float *d_image,
                      float *grad_albedo, float *grad_emission, double *seconds);

/* ------------------------------------------------------------------------ */
/* Radiance MLP (no reference counterpart: the reference trains a different  */
/* network in PyTorch, neural_render/src/model.py:276-351). 9 features       */
/* (position, shading normal, albedo) -> 12-frequency encoding (225, padded  */
/* to 256) -> 6 ReLU layers of width 256 -> 3 outputs (reflected radiance).  */
/* Hand-written BF16 MFMA kernels with FP32 accumulation; FP32 master        */
/* weights and Adam. Weights: 7 row-major [out][in] matrices, in order, with */
/* (in, out) = (256, 256) x 6 and (256, 16); no biases.                      */
/* ------------------------------------------------------------------------ */
typedef struct pt_mlp pt_mlp;
int pt_mlp_create(int device, uint64_t seed, pt_mlp **out);
void pt_mlp_destroy(pt_mlp *m);
int pt_mlp_param_count(uint64_t *count);
int pt_mlp_get_weights(pt_mlp *m, float *weights);
/* Also resets the Adam state. */
int pt_mlp_set_weights(pt_mlp *m, const float *weights);
/* features: n x 9 (host), out: n x 3 (host). */
int pt_mlp_forward(pt_mlp *m, const float *features, uint32_t n, float *out);
/* Inference kernel choice: fused != 0 (default) runs the whole forward in one
 * kernel per 128-row tile with the activations in LDS; 0 runs the encoding and
 * one GEMM per layer through HBM (the training forward). Both give the same
 * bits. */
int pt_mlp_set_forward_kernel(pt_mlp *m, int fused);
/* Gradient of the relative L2 loss mean((p - t)^2 / (p^2 + 0.01)) (p held
 * constant in the denominator) over one batch; grad has pt_mlp_param_count
 * floats. */
int pt_mlp_gradient(pt_mlp *m, const float *features, const float *targets,
                    uint32_t n, float *grad, float *loss);
/* One Adam step (beta 0.9 / 0.99, eps 1e-8) on one host batch. */
int pt_mlp_train_step(pt_mlp *m, const float *features, const float *targets,
                      uint32_t n, float lr, float *loss);
/* Keep a training set resident on the device (rows x 9 features, rows x 3
 * targets), then run `steps` Adam steps on batches drawn from it on the
 * device; losses (may be NULL) receives one loss per step. */
int pt_mlp_set_dataset(pt_mlp *m, const float *features, const float *targets,
                       uint64_t rows);
int pt_mlp_train(pt_mlp *m, uint32_t steps, uint32_t batch, float lr,
                 uint64_t seed, float *losses, double *seconds);

/* Inference timing on `rows` device-resident rows: average duration of one
 * hidden 256x256 layer GEMM and of a whole forward pass (HIP events). */
int pt_mlp_time_forward(pt_mlp *m, uint32_t rows, uint32_t iters,
                        float *gemm_ms, float *forward_ms);

/* Radiance-cache render of a SquareImage (FloatRGB, host out_rgb x*y*3):
 * each sample traces the camera ray and one bounce with the light sample of
 * the first vertex; the radiance reflected at the second vertex is the MLP's
 * prediction for that point (clamped at 0). Same samples as pt_render. */
int pt_render_mlp(pt_renderer *r, pt_mlp *m, const pt_scene *scene,
                  const float film_to_world[12], uint32_t x_dim, uint32_t y_dim,
                  uint32_t samples_per, const pt_settings *settings,
                  float *out_rgb, double *seconds);

/* Radiance-cache render that queries the MLP once per pixel, at the first
 * camera hit through the pixel centre. samples_per > 0: the image is the
 * tracer's render terminated after the first vertex (emission seen plus its
 * light sample, samples_per samples, the same samples as pt_render) plus the
 * MLP's prediction there, for a network trained on the indirect part of the
 * reflected radiance. samples_per = 0: emission seen plus the prediction,
 * for a network trained on all reflected radiance (no tracing beyond the
 * camera ray). Host out_rgb of x_dim*y_dim*3 floats. */
int pt_render_mlp_first_vertex(pt_renderer *r, pt_mlp *m, const pt_scene *scene,
                               const float film_to_world[12], uint32_t x_dim, uint32_t y_dim,
                               uint32_t samples_per, const pt_settings *settings,
                               float *out_rgb, double *seconds);

/* Per-pixel means of the last Streaming SquareImage render re-summed in FP64
 * (sequentially over each pixel's per-sample radiance, in sample order):
 * host out_rgb of x_dim*y_dim*3 doubles. For ground-truth images. */
int pt_last_render_fp64(pt_renderer *r, double *out_rgb);

/* Statistics of the last pt_render. kernel_ms spans the launches with HIP
 * events on the renderer's stream. The per-kernel fields are filled only when
 * profiling is enabled (pt_set_profiling), which brackets each traversal or
 * mega-kernel launch with its own events. */
typedef struct pt_render_stats {
  double kernel_ms;
  double trace_ms;         /* Streaming: traversal launches (profiled); each
                              traces one iteration's extension and shadow
                              rays together */
  uint64_t rays;           /* camera + extension + shadow rays traced */
  uint64_t samples;
  uint64_t trace_launches; /* traversal launches enqueued (all pipelines) */
  uint64_t busy_trace_launches; /* those that traced at least one ray (the
                                   launches trace_ms covers) */
  uint64_t ext_rays, shadow_rays;
  double mega_ms;          /* mega-kernel launches */
  uint64_t mega_launches;
  uint32_t profiled;
  uint32_t pipelines;      /* wavefront pipelines that ran at once (Streaming);
                              their launches overlap, so a launch's duration
                              covers its share of the rays under contention */
  uint32_t replays;        /* pt_render_adjoint: replays of each path (1 when
                              path lengths are bounded, else 2) */
} pt_render_stats;
int pt_set_profiling(pt_renderer *r, int enable);
int pt_last_render_stats(const pt_renderer *r, pt_render_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* PATH_AMD_H */
