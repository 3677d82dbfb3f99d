/* muonx — MI355X (gfx950) wavefront path tracer behind muon's render surface.
 *
 * C ABI: plain C types, opaque handles, int status codes (0 = ok, < 0 =
 * error; muonx_last_error() returns a thread-local message). No exception
 * crosses this boundary. The caller owns host arrays; the library owns the
 * device memory of a scene until muonx_scene_destroy().
 *
 * The reference (voithos/muon) has no C ABI; its seams are C++ interfaces.
 * Each entry point below names the reference interface it replaces.
 */
#ifndef MUONX_H_
#define MUONX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MUONX_OK 0
#define MUONX_ERR_INVALID (-1)   /* bad argument */
#define MUONX_ERR_HIP (-2)       /* HIP runtime error (no device, launch failure) */
#define MUONX_ERR_NOMEM (-3)     /* device or host allocation failed */
#define MUONX_ERR_IO (-4)        /* file could not be read */
#define MUONX_ERR_STATE (-5)     /* call not valid in the object's state */
#define MUONX_ERR_CAPACITY (-6)  /* request exceeds an admission limit */
#define MUONX_ERR_CANCELLED (-7) /* the job was cancelled before it started rendering */

/* Integrator selection (muon/parser.cc:330-355). */
#define MUONX_INTEGRATOR_NORMALS 0
#define MUONX_INTEGRATOR_ALBEDO 1
#define MUONX_INTEGRATOR_DEPTH 2
#define MUONX_INTEGRATOR_RAYTRACER 3
#define MUONX_INTEGRATOR_ANALYTICDIRECT 4
#define MUONX_INTEGRATOR_PATHTRACER 5

/* muon/nee.h, muon/importance_sampling.h */
#define MUONX_NEE_OFF 0
#define MUONX_NEE_ON 1
#define MUONX_NEE_MIS 2
#define MUONX_IS_HEMISPHERE 0
#define MUONX_IS_COSINE 1
#define MUONX_IS_BRDF 2

/* Material BRDF tags. "lambertian" in a scene file yields PHONG
 * (muon/parser.cc:165-178 falls through). */
#define MUONX_BRDF_LAMBERTIAN 0
#define MUONX_BRDF_PHONG 1
#define MUONX_BRDF_GGX 2

#define MUONX_LIGHT_DIRECTIONAL 0
#define MUONX_LIGHT_POINT 1
#define MUONX_LIGHT_QUAD 2

#define MUONX_PRIM_TRI 0
#define MUONX_PRIM_SPHERE 1

typedef struct muonx_ctx muonx_ctx;
typedef struct muonx_scene muonx_scene;
typedef struct muonx_host_scene muonx_host_scene;
typedef struct muonx_server muonx_server;

/* Renderer options; replaces muon::Options (muon/options.h:11-23). */
typedef struct muonx_options {
  int32_t acceleration;       /* 0 = linear, 1 = bvh (muon/acceleration_type.h:11-17) */
  int32_t partition_strategy; /* 0 = uniform, 1 = midpoint, 2 = sah */
  int32_t parallelism;        /* host threads for CPU-side work */
  int32_t show_stats;
  int64_t default_seed;       /* >= 0 replaces the std::random_device seed (muon/parser.cc:181) */
  int32_t device_bvh;         /* 1 (with acceleration = bvh): skip the host SAH build; muonx_scene_create
                                 then builds a BVH on the device (fast admission, fast traversal mode
                                 only). 0 (default): the reference's SAH tree is built on the host */
} muonx_options;

/* Scene descriptor handed to the device: host arrays in the reference's data
 * model (muon/scene.h:22-95, muon/objects.h, muon/materials.h:125-148,
 * muon/lighting.h), plus the binary SAH BVH built on the host
 * (muon/acceleration.cc:204-458) in pre-order. Primitive ids are parse
 * (AddPrimitive) order. All arrays are read during muonx_scene_create only. */
typedef struct muonx_scene_desc {
  uint32_t num_prims;
  const uint8_t *prim_type;        /* MUONX_PRIM_* */
  const float *prim_geom;          /* 12 per prim: tri v0,v1,v2 (object space); sphere cx,cy,cz,r */
  const float *prim_vnormal;       /* 9 per prim: vertex normals N0,N1,N2, or NULL */
  const uint8_t *prim_use_vnormal; /* 1 = interpolate vertex normals (muon/objects.cc:163-166), or NULL */
  const int32_t *prim_transform;   /* index into transforms */
  const int32_t *prim_material;    /* index into materials */
  const int32_t *prim_light;       /* quad-light index for the emitter tris, else -1 */

  uint32_t num_transforms;
  const float *transforms;         /* 48 per transform: M, M^-1, M^-T, column-major */

  uint32_t num_materials;
  const float *materials;          /* 16 per material: ambient3 diffuse3 specular3 emission3
                                      shininess roughness brdf(MUONX_BRDF_*) 0 */
  uint32_t num_lights;
  const float *lights;             /* 16 per light: kind color3 a3 b3 c3 0 — directional: a=direction;
                                      point: a=position b=attenuation; quad: a=corner b=edge0 c=edge1 */
  uint32_t num_nodes;              /* binary BVH, pre-order, root = 0 */
  const float *node_bounds;        /* 6 per node: min xyz, max xyz */
  const int32_t *node_left;        /* child index, -1 for a leaf */
  const int32_t *node_right;
  const int32_t *node_axis;        /* split axis of an internal node */
  const int32_t *node_start;       /* leaf: first index into leaf_prims */
  const int32_t *node_count;       /* leaf: primitive count; 0 for internal nodes */
  const int32_t *leaf_prims;       /* num_prims parse-order ids in leaf order */
  int32_t acceleration;            /* 1: BVH; 0: the linear structure (muon/acceleration.cc:14-45),
                                      one leaf in parse order, whose queries count tests as
                                      Linear does */
} muonx_scene_desc;
/* A descriptor with num_nodes = 0 (the node arrays and leaf_prims unused)
 * hands over primitives without a tree: muonx_scene_create then builds a
 * binary BVH on the device (a linear BVH over Morton codes of the primitive
 * centroids) and uses it as it uses a host-built one. Such a scene is not
 * traversed in the reference's node order, so the exact traversal mode and
 * the integrators that need it (normals, albedo, depth, raytracer,
 * analyticdirect) are refused with MUONX_ERR_STATE; the path tracer and fast
 * queries run as on a host-built tree. */

/* Per-render settings: the scene's render properties (muon/scene.h:31-51)
 * and the pinhole camera precomputed on the host (muon/camera.cc:15-35). */
typedef struct muonx_render_params {
  int32_t width, height, spp;
  int32_t min_depth, max_depth;
  int32_t integrator;           /* MUONX_INTEGRATOR_* */
  int32_t nee;                  /* MUONX_NEE_* */
  int32_t importance_sampling;  /* MUONX_IS_* */
  int32_t russian_roulette;
  int32_t light_samples;
  int32_t light_stratify;
  uint32_t seed;                /* path-sampling seed (muon/parser.cc:347-349) */
  float gamma;
  float eye[3], cam_u[3], cam_v[3], cam_w[3];
  float tan_fov_x, tan_fov_y;
  int32_t has_camera;
  /* 0: sample 0 of a pixel at its centre, later samples jittered by the
     counter-based RNG (statistically the reference's Sampler,
     muon/sampling.cc:86-108). */
  int32_t sample_mode;
} muonx_render_params;

/* Counters of muon's TraceStats (muon/stats.h:11-44) plus shadow rays, which
 * the reference does not count. */
typedef struct muonx_trace_stats {
  uint64_t primary_rays;
  uint64_t secondary_rays;
  uint64_t shadow_rays;
  uint64_t render_ns;   /* device time of the render, ns */
  uint64_t culled_rays; /* of secondary_rays, MIS rays resolved without a BVH traversal: their
                           direction misses every emitting primitive, so their closest hit
                           cannot be a light (path tracer; 0 for the other integrators) */
  /* TraceStats' test counters (muon/stats.h:11-45, counted at
   * muon/acceleration.cc:92-114,162-179). The deterministic integrators'
   * exact traversal visits the reference's nodes, so theirs equal the
   * reference's. The path tracer walks the fast mode's 4-wide tree: with
   * tuning.count_work its counters are that traversal's work (child boxes
   * tested and passed, primitive tests and hits), not the reference's
   * binary-tree counts; without it they are 0. */
  uint64_t object_tests;
  uint64_t object_hits;
  uint64_t bounds_tests;
  uint64_t bounds_hits;
} muonx_trace_stats;

typedef struct muonx_scene_info {
  uint64_t device_bytes;  /* HBM footprint of the resident scene */
  uint32_t num_prims;
  uint32_t num_nodes;     /* device BVH nodes */
  uint32_t bvh_depth;     /* binary SAH tree depth */
  uint32_t num_lights;
  uint32_t num_quad_lights;
  uint32_t device_built;  /* 1: the BVH was built on the device (descriptor without a tree) */
  float build_ms;         /* that build: kernels, sort, copies and the host's pre-order pass */
} muonx_scene_info;

/* ------------------------------------------------------------ misc */
const char *muonx_last_error(void);
const char *muonx_version(void);

/* ------------------------------------------------------------ host ingest (CPU)
 * Replaces muon::Parser::Parse (muon/parser.cc:230-880) together with
 * acceleration::BVH::Init (muon/acceleration.cc:204-229): parses a .muon
 * scene and builds the SAH BVH on the host. Needs no GPU. `load` reads OBJ
 * and PLY meshes, relative to the scene file's directory (parse_file) or to
 * the working directory (parse_text); a file that cannot be read counts as
 * a warning and is skipped, as the reference logs and skips it. */
int muonx_host_scene_parse_file(const char *path, const muonx_options *opts, muonx_host_scene **out);
int muonx_host_scene_parse_text(const char *text, size_t len, const muonx_options *opts,
                                muonx_host_scene **out);
void muonx_host_scene_destroy(muonx_host_scene *hs);
/* Views into the host scene; valid until muonx_host_scene_destroy(). */
int muonx_host_scene_desc(const muonx_host_scene *hs, muonx_scene_desc *out);
int muonx_host_scene_params(const muonx_host_scene *hs, muonx_render_params *out);
/* Count of malformed / unknown lines (muon logs them and continues). */
int muonx_host_scene_warnings(const muonx_host_scene *hs);
/* The scene's `output` file name (muon/scene.h:36, set by the `output`
 * command, muon/parser.cc:309-318; default "raytrace.png"), which the render
 * driver writes unless the caller overrides it (muon/renderer.cc:26-27).
 * *out is valid until muonx_host_scene_destroy(). */
int muonx_host_scene_output(const muonx_host_scene *hs, const char **out);

/* ------------------------------------------------------------ device context / scene */
/* One context per device; calls on a context are serialised internally. */
int muonx_ctx_create(int device, muonx_ctx **out);
void muonx_ctx_destroy(muonx_ctx *ctx);

/* Launch tuning of a context (no counterpart in the reference, whose only
 * knob is --parallelism). Traversal stack: 0 = whole stack in LDS (capacity
 * from the BVH depth); 8 or 16 = LDS ring of that many entries per lane with
 * older entries spilled to a per-thread HBM region (more waves per CU).
 * while_while: lanes walk internal nodes until each holds a leaf, then test
 * leaves together. blocks_per_cu: persistent grid size. path_budget_mb:
 * HBM for the path tracer's per-path state and queues. bvh_width: node
 * width of the fast traversal mode (the exact mode always walks the
 * reference's binary tree). */
typedef struct muonx_tuning {
  int32_t stack_entries;
  int32_t while_while;
  int32_t blocks_per_cu;
  int32_t path_budget_mb;
  int32_t bvh_width;  /* fast mode: 2 = binary nodes, 4 = 4-wide nodes collapsed from the same tree */
  int32_t ray_fetch;  /* fast-mode queries: 0 = one ray per lane per grid step; 1..64 = a lane whose
                         ray is done takes a new one once that many lanes of its wavefront are idle */
  int32_t node_bits;  /* 4-wide nodes: 32 = float child boxes (128 B), 8 = 8-bit quantized boxes (64 B) */
  int32_t count_work; /* path tracer: 1 = count its traversals' box and primitive tests and hits into
                         muonx_trace_stats (about 3 % slower), 0 = leave those counters at 0 (default) */
} muonx_tuning;
int muonx_ctx_set_tuning(muonx_ctx *ctx, const muonx_tuning *t);
int muonx_ctx_get_tuning(const muonx_ctx *ctx, muonx_tuning *t);

/* Autotuner (no counterpart in the reference): times candidate traversal
 * configurations (bvh_width, stack_entries, while_while, blocks_per_cu,
 * ray_fetch) on n_rays incoherent closest-hit rays over the scene's bounds
 * (0: 2^22) and, when `job` is not NULL, the engine's node_bits, ray_fetch,
 * blocks_per_cu and path-queue capacity (path_budget_mb) on that render (spp
 * reduced to stay under 2^26 samples); installs the fastest in the scene's
 * context (with a job: the job's fastest). report may be NULL. */
typedef struct muonx_autotune_report {
  muonx_tuning tuning;     /* installed */
  uint32_t candidates;     /* traversal configurations timed */
  uint32_t sample_rays;
  float best_ms, worst_ms; /* per launch of the sample rays */
  float previous_ms;       /* the tuning in place before the call */
  float render_ms;         /* the job at the chosen capacity, 0 without a job */
} muonx_autotune_report;
int muonx_autotune(muonx_scene *scene, const muonx_render_params *job, uint32_t n_rays,
                   muonx_autotune_report *report);

/* Uploads a scene; replaces the Structure::Init()'d acceleration structure
 * (muon/acceleration.h:26-53) with a device-resident BVH (the descriptor's
 * host-built tree, or one built on the device when it has none). Immutable
 * after creation and shareable by concurrent renders. */
int muonx_scene_create(muonx_ctx *ctx, const muonx_scene_desc *desc, muonx_scene **out);
/* Waits for server jobs that still render the scene, then frees it. */
void muonx_scene_destroy(muonx_scene *scene);
int muonx_scene_get_info(const muonx_scene *scene, muonx_scene_info *out);

/* ------------------------------------------------------------ queries (device pointers)
 * Batched acceleration::Structure::Intersect (muon/acceleration.cc:65-142).
 * ray_o / ray_d: 3n floats each (xyz per ray); tmax: n floats or NULL (+inf).
 * prim_id: parse-order id or -1 on a miss; t: world distance. */
int muonx_intersect_closest(muonx_scene *scene, const float *ray_o, const float *ray_d, const float *tmax,
                            uint32_t n, int32_t *prim_id, float *t, void *stream);
/* Batched acceleration::Structure::HasIntersection (muon/acceleration.cc:144-202). */
int muonx_intersect_any(muonx_scene *scene, const float *ray_o, const float *ray_d, const float *tmax,
                        uint32_t n, uint8_t *occluded, void *stream);

/* Traversal modes.
 * FAST: reciprocal slab test widened by 2^-20 relative, so every box the
 *   reference's slab test accepts is accepted; identical results except for
 *   distance ties and hits within rounding of a primitive's edge.
 * EXACT: the reference's division slab test (muon/bounds.cc:88-125) with the
 *   exact entry distance kept for the far child, so the traversal visits the
 *   reference's nodes and primitives in its order: results are the
 *   reference's, ties included. */
#define MUONX_TRAVERSAL_FAST 0
#define MUONX_TRAVERSAL_EXACT 1

typedef struct muonx_query {
  const float *ray_o;          /* device, 3n floats */
  const float *ray_d;          /* device, 3n floats */
  const float *tmax;           /* device, n floats, or NULL (+inf) */
  uint32_t n;
  int32_t any_hit;             /* 0 closest (Intersect), 1 any (HasIntersection) */
  int32_t mode;                /* MUONX_TRAVERSAL_* */
  int32_t *prim_id;            /* closest: parse-order id or -1 */
  float *t;                    /* closest: world distance, may be NULL */
  uint8_t *occluded;           /* any-hit: 1 if a primitive lies at 0 < d < tmax */
  unsigned long long *counters; /* device u64[2] or NULL: += internal nodes fetched,
                                   += primitive tests (cf. the TraceStats increments,
                                   muon/acceleration.cc:92,108) */
} muonx_query;

/* General batched query: closest or any hit, either traversal mode. */
int muonx_intersect(muonx_scene *scene, const muonx_query *q, void *stream);

/* muonx_intersect that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Results
 * are those of muonx_intersect. The traversal stacks' spill region and the
 * ray-claim counter live in a 32 MiB static device array per GPU (held until
 * the process ends) instead of context buffers; the grid is capped so that
 * the spill fits, and a scene whose traversal stack needs more than the
 * array holds for one block is rejected with MUONX_ERR_CAPACITY. Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads, other contexts) race and must be serialised by the
 * caller. The caller passes the rays and outputs as for muonx_intersect. */
int muonx_intersect_noalloc(muonx_scene *scene, const muonx_query *q, void *stream);

/* ------------------------------------------------------------ render
 * Replaces the tile loop of muon::Renderer::Render (muon/renderer.cc:40-76):
 * Camera::CastRay + Integrator::Trace per sample + Film::SetPixel.
 * accum_rgb: device pointer to H*W*3 floats, row-major, receives the
 * unnormalised sum over samples (zeroed by the call). stats may be NULL. */
int muonx_render(muonx_scene *scene, const muonx_render_params *params, float *accum_rgb,
                 muonx_trace_stats *stats, void *stream);

/* muonx_render into a host buffer: the library stages the film in its own
 * device memory and copies it back (for callers without a HIP toolchain,
 * such as the muonx CLI). accum_rgb: host pointer, H*W*3 floats. */
int muonx_render_host(muonx_scene *scene, const muonx_render_params *params, float *accum_rgb,
                      muonx_trace_stats *stats);

/* ------------------------------------------------------------ batched and served rendering
 * No counterpart in the reference, which renders one scene per process
 * (muon/muon.cc:48-49). Path-tracer jobs of any resident scenes share one
 * persistent wavefront engine: a fixed pool of path slots that every
 * iteration advances by one bounce and refills with camera samples of
 * admitted jobs, so jobs join and leave between bounces. Each job's film
 * is accumulated in 64-bit fixed point (2^-24 units) and its random
 * numbers are keyed by (seed, key, pixel, sample): the image of a job does
 * not depend on what else is in flight. */
#define MUONX_PENDING 1 /* muonx_server_wait: the job has not finished yet */

typedef struct muonx_engine_config {
  uint32_t path_slots;          /* 0: as many as tuning.path_budget_mb holds */
  uint32_t max_jobs;            /* in-flight job slots, at most 65536; 0: 1024 */
  int32_t shadow_per_vertex;    /* largest NEE shadow-ray count per vertex a job
                                   may need; 0: 4 */
  uint32_t iterations_per_poll; /* engine iterations between completion checks; 0: 2 */
  uint64_t film_budget_bytes;   /* HBM for in-flight films, 24 B per pixel; 0: 1 GiB */
  int32_t use_graph;            /* replay each iteration as a captured hipGraph */
  int32_t slice_order;          /* which running render gets the next slice of samples:
                                   MUONX_SLICES_OLDEST_FIRST (0),
                                   MUONX_SLICES_LEAST_WORK_FIRST (1: lower median latency
                                   under load, longer tail) or
                                   MUONX_SLICES_RESPONSE_RATIO (2: the highest ratio of
                                   work handed out since the render's submission to its
                                   own work left) */
} muonx_engine_config;
#define MUONX_SLICES_OLDEST_FIRST 0
#define MUONX_SLICES_LEAST_WORK_FIRST 1
#define MUONX_SLICES_RESPONSE_RATIO 2

typedef struct muonx_job {
  muonx_scene *scene;            /* resident scene (its context = the engine's) */
  muonx_render_params params;    /* integrator must be MUONX_INTEGRATOR_PATHTRACER */
  uint32_t key;                  /* random stream: equal (params, key) give equal images */
  float *accum_rgb;              /* device, H*W*3 float sums over samples, as muonx_render */
} muonx_job;

typedef struct muonx_job_result {
  int32_t status;                /* MUONX_OK or an error code */
  uint64_t submit_ns, admit_ns, done_ns; /* steady clock: arrival, admission, film written */
} muonx_job_result;

/* Renders n jobs through one engine and returns when all films are
 * written. results (n entries) and stats may be NULL. */
int muonx_render_batch(muonx_ctx *ctx, const muonx_job *jobs, uint32_t n, const muonx_engine_config *cfg,
                       muonx_job_result *results, muonx_trace_stats *stats);

/* Request loop: a thread that owns an engine, admits submitted jobs in
 * arrival order as film memory (film_budget_bytes) and job slots free up,
 * and keeps the engine iterating while work is in flight. */
int muonx_server_create(muonx_ctx *ctx, const muonx_engine_config *cfg, muonx_server **out);
/* Validates and queues a job; returns a ticket. Jobs that can never fit
 * (film above the budget, too many shadow rays per vertex) are rejected
 * with MUONX_ERR_CAPACITY. */
int muonx_server_submit(muonx_server *srv, const muonx_job *job, uint64_t *ticket);
/* Waits up to timeout_ms (< 0: forever, 0: poll) for a ticket. Returns
 * MUONX_OK and fills `result` (the ticket is then consumed), MUONX_PENDING
 * on timeout, MUONX_ERR_CANCELLED (result filled, ticket consumed) for a job
 * that was cancelled before it started rendering, or the job's error. */
int muonx_server_wait(muonx_server *srv, uint64_t ticket, int64_t timeout_ms, muonx_job_result *result);
/* Cancels a submitted job that has not started rendering: it is dropped
 * (its film is not written) and its wait returns MUONX_ERR_CANCELLED. A job
 * already admitted to the GPU finishes normally. Returns MUONX_OK when the
 * request is recorded, MUONX_ERR_INVALID for an unknown or collected ticket. */
int muonx_server_cancel(muonx_server *srv, uint64_t ticket);

typedef struct muonx_server_stats {
  uint64_t submitted, completed, pending, in_flight;
  uint64_t iterations;
  uint64_t primary_rays, secondary_rays, shadow_rays;
  uint64_t film_bytes_in_use, film_budget_bytes;
  uint32_t path_slots;
  int32_t shadow_per_vertex;
  uint64_t cancelled;           /* jobs dropped by muonx_server_cancel */
} muonx_server_stats;
int muonx_server_get_stats(muonx_server *srv, muonx_server_stats *out);
/* Finishes every submitted job, then stops the thread and frees the engine. */
void muonx_server_destroy(muonx_server *srv);

/* Film output, muon::Film::WriteOutput (muon/film.cc:21-43): /spp,
 * pow(v, 1/gamma), clamp [0,1], *255 truncated. Host arrays. */
int muonx_film_to_rgb8(const float *accum_rgb, int32_t width, int32_t height, int32_t spp, float gamma,
                       uint8_t *rgb8);
/* Writes an 8-bit RGB PNG (the reference saves through CImg). */
int muonx_write_png(const char *path, const uint8_t *rgb8, int32_t width, int32_t height);

#ifdef __cplusplus
}
#endif

#endif /* MUONX_H_ */
