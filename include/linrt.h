/*
 * linrt.h - C ABI of liblinrt_hip.so, the MI355X (gfx950) render hot path of
 * LinRT (Lincacc/Ray-Tracing).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (LRT_OK == 0, negative on error) and never lets a C++ exception
 * cross the boundary.  lrt_last_error() returns a thread-local message for the
 * most recent failure on the calling thread.
 *
 * Threading: a lrt_scene is externally synchronised (one host thread at a time
 * per scene).  All GPU work of a scene is ordered on the scene's own HIP stream
 * and every call blocks until its results are complete, unless stated.
 * Device pointers (device_ptrs != 0): the library's streams are non-blocking,
 * so such a call first makes its stream wait for the work already queued on
 * the device's null stream (torch's default stream), which may still be
 * writing the inputs or reading earlier outputs; work on other streams is
 * the caller's to complete before the call.
 *
 * Reference interfaces that these entry points stand in for are cited per
 * declaration as path:line of the reference repository.
 */
#ifndef LINRT_H
#define LINRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    LRT_OK = 0,
    LRT_ERR_ARG = -1,         /* invalid argument (null pointer, bad index, size) */
    LRT_ERR_HIP = -2,         /* a HIP runtime call failed */
    LRT_ERR_ALLOC = -3,       /* host or device allocation failed */
    LRT_ERR_UNSUPPORTED = -4, /* valid request this build does not implement */
    LRT_ERR_INTERNAL = -5
};

/* ---- BSDF flags (bsdf.h:18-28) ----------------------------------------- */
enum {
    LRT_FLAG_NONE = 0,
    LRT_FLAG_DIFFUSE = 1u << 0,
    LRT_FLAG_SPECULAR = 1u << 1,
    LRT_FLAG_REFLECTIVE = 1u << 2,
    LRT_FLAG_TRANSPARENT = 1u << 3,
    LRT_FLAG_GLOSSY = 1u << 4,
    LRT_FLAG_METALLIC = 1u << 5,
    LRT_FLAG_DIELECTRIC = 1u << 6,
    LRT_FLAG_REFRACTIVE = 1u << 7
};

/* measure (bsdf.h:9-16): only these two are used on the path */
enum { LRT_MEASURE_SOLID_ANGLE = 1, LRT_MEASURE_DISCRETE = 5 };
/* transport mode (bsdf.h:31-34) */
enum { LRT_MODE_RADIANCE = 0, LRT_MODE_IMPORTANCE = 1 };

/* ---- materials: one record per reference BSDF class ------------------- */
enum {
    LRT_BSDF_NONE = 0,            /* surface without BSDF: paths stop here */
    LRT_BSDF_DIFFUSE = 1,         /* diffuse_bsdf.h */
    LRT_BSDF_CONDUCTOR = 2,       /* rough_conductor_bsdf.h (alpha <= 0: smooth mirror) */
    LRT_BSDF_DIELECTRIC = 3,      /* dielectric_bsdf.h */
    LRT_BSDF_ROUGH_DIELECTRIC = 4,/* rough_dielectric_bsdf.h */
    LRT_BSDF_PLASTIC = 5,         /* plastic_bsdf.h */
    LRT_BSDF_ROUGH_PLASTIC = 6,   /* rough_plastic_bsdf.h */
    LRT_BSDF_COATING = 7,         /* coating_bsdf.h, over materials[nested] */
    LRT_BSDF_ROUGH_COATING = 8    /* rough_coating_bsdf.h, over materials[nested] */
};

/* Host-side precomputed data a material needs is carried inline:
 *   fdr_int    plastic: fresnel_diffuse_reflectance(1/eta) (fresnel.h:128-133),
 *              fill with lrt_fresnel_diffuse_reflectance().
 *   table_ext/table_int  rough plastic, rough coating: index of a 101-float
 *              rough-transmittance slice (100 theta samples + 1 diffuse value,
 *              rough_transmittance.h:129-185) in lrt_scene_desc.tables, or -1.
 * Constructor semantics follow the reference: eta = int_ior/ext_ior, GGX alpha
 * clamped at 1e-4 inside the distribution (microfacet.h:12-13). */
typedef struct lrt_material {
    int32_t type;
    int32_t nested;                 /* coatings: index of the inner material, else -1 */
    float int_ior, ext_ior;
    float alpha;
    int32_t nonlinear;              /* plastic / rough plastic */
    float diffuse[3];               /* diffuse albedo, plastic diffuse reflectance */
    float specular_reflectance[3];
    float specular_transmittance[3];
    float cond_eta[3], cond_k[3];   /* conductor complex IOR */
    float sigma_a[3];
    float thickness;                /* coatings */
    float fdr_int;                  /* plastic */
    int32_t table_ext, table_int;   /* rough plastic: both; rough coating: table_ext */
} lrt_material;

/* ---- scene description (host arrays the caller owns; they are copied) -- */
enum { LRT_SHAPE_MESH = 0, LRT_SHAPE_RECTANGLE = 1, LRT_SHAPE_SPHERE = 2 };

typedef struct lrt_object_desc {
    int32_t shape;
    int32_t material;               /* -1: no BSDF (rectangles never carry one, rectangle.h:50) */
    int32_t emitter;                /* index into lights (area light on this shape) or -1 */
    uint32_t n_vertices, n_triangles;
    const float* positions;         /* 3*n_vertices */
    const uint32_t* indices;        /* 3*n_triangles */
    const float* corner_normals;    /* 9*n_triangles, one normal per corner (triangle_mesh.h:13) */
    const uint16_t* triangle_material; /* optional per-triangle material, NULL: use .material */
    float corner[3], edge_u[3], edge_v[3]; /* rectangle (rectangle.h:9-16) */
    float center[3];
    float radius;                   /* sphere (sphere.h:10-13) */
} lrt_object_desc;

enum { LRT_LIGHT_AREA = 0, LRT_LIGHT_SKY = 1 };

typedef struct lrt_light_desc {
    int32_t kind;
    int32_t object;                 /* area light: emitting object */
    float radiance[3];              /* area light */
    float sun_dir[3];               /* sky: sun direction (normalised by the library) */
    float turbidity, ground_albedo; /* sky: model parameters (informative) */
    float scale;                    /* sky: exposure (env_emitter.h:13) */
    double config[3][9];            /* sky: cooked Hosek-Wilkie RGB coefficients */
    double channel_radiance[3];     /* sky: cooked per-channel radiance */
} lrt_light_desc;

typedef struct lrt_camera_desc {
    float origin[3], look_at[3], up[3];
    float fov_deg, aspect;          /* vertical field of view (camera.h:20-44) */
} lrt_camera_desc;

typedef struct lrt_scene_desc {
    uint32_t n_objects;
    const lrt_object_desc* objects;
    uint32_t n_materials;
    const lrt_material* materials;
    uint32_t n_lights;
    const lrt_light_desc* lights;   /* emitter list order is the reference's emitter order */
    uint32_t n_tables;
    const float* tables;            /* n_tables * 101 floats */
    lrt_camera_desc camera;
} lrt_scene_desc;

typedef struct lrt_scene lrt_scene; /* opaque, library-owned */

/* ---- errors / device ---------------------------------------------------- */
const char* lrt_last_error(void);
int lrt_version(void);
/* number of visible HIP devices (0 without a GPU; never fails) */
int lrt_device_count(void);

/* ---- accelerator: SAH kd-tree (sah_kdtree.h:35-57, 175-347) ------------ *
 * Builds the reference's SAH kd-tree over n boxes (6 floats each:
 * min xyz, max xyz) and returns it in the reference's own node numbering,
 * before flattening: per node leaf flag, axis, split, left, right, first,
 * count, plus the ordered primitive references.  Call with null outputs to
 * get the sizes (*n_nodes, *n_refs, *depth_limit), then again with arrays. */
int lrt_kdtree_build_export(const float* boxes, uint32_t n,
                            uint32_t* n_nodes, uint32_t* n_refs, int32_t* depth_limit,
                            uint8_t* leaf, int32_t* axis, float* split,
                            int32_t* left, int32_t* right, int32_t* first, int32_t* count,
                            uint32_t* refs);

/* ---- scene upload (scene.h:19-34, triangle_mesh.h:14-28) --------------- */
int lrt_scene_create(int hip_device, const lrt_scene_desc* desc, lrt_scene** out);
int lrt_scene_release(lrt_scene* scene);
/* statistics of the uploaded acceleration structures */
typedef struct lrt_scene_info {
    uint64_t n_triangles, n_nodes, n_refs, device_bytes;
    int32_t max_tree_depth;
    double build_seconds, upload_seconds;
    float bounding_radius;   /* scene::bounding_sphere_radius (scene.h:41-45) */
    int32_t has_delta_bsdf;  /* scene::has_delta_bsdf (scene.h:48-57) */
} lrt_scene_info;
int lrt_scene_info_get(const lrt_scene* scene, lrt_scene_info* info);

/* ---- ray queries (scene.h:36-38 -> hittable_list.h:19-34 ->
 *      triangle_mesh.h:30-99 / sah_kdtree.h:61-121) ------------------------
 * rays: SoA, 8 arrays of n floats back to back: ox oy oz dx dy dz tmin tmax.
 * hit_obj_prim: 2*n int32 (object, primitive; -1,-1 on a miss);
 * hit_t: n floats; hit_uv: 2*n floats (barycentric u, v; may be NULL).
 * device_ptrs != 0: every pointer is device memory on the scene's device and
 * the call is asynchronous on the scene's stream (use lrt_scene_sync).      */
int lrt_trace_closest(lrt_scene* scene, const float* rays, uint32_t n,
                      int32_t* hit_obj_prim, float* hit_t, float* hit_uv, int device_ptrs);
/* lrt_trace_closest on device pointers that allocates nothing: no
 * counterpart in the reference; it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same rays, outputs and results as lrt_trace_closest with
 * device_ptrs = 1 (rays, hit_obj_prim, hit_t and hit_uv are device memory on
 * the scene's device; asynchronous on the scene's stream, lrt_scene_sync
 * waits).  Its only scratch, the traversal's ray-fetch counter, is a
 * 256-byte static device array per device (resolved by lrt_scene_create);
 * every call on a device shares it, so calls that overlap on one device race.
 * n is limited to 2^31 rays per call (LRT_ERR_ARG beyond). */
int lrt_trace_closest_noalloc(lrt_scene* scene, const float* rays, uint32_t n,
                              int32_t* hit_obj_prim, float* hit_t, float* hit_uv);
/* occlusion query: occluded[i] = 1 if any surface other than those of
 * emitter ignore_emitter[i] (-1: none) lies in [tmin, tmax]
 * (path_integrator.h:113-116). ignore_emitter may be NULL. */
int lrt_trace_occluded(lrt_scene* scene, const float* rays, uint32_t n,
                       const int32_t* ignore_emitter, uint8_t* occluded, int device_ptrs);
/* algorithmic-work counters of the closest-hit traversal: per ray, interior
 * and leaf nodes visited, leaf references read, triangle tests run */
int lrt_trace_count(lrt_scene* scene, const float* rays, uint32_t n,
                    uint32_t* nodes, uint32_t* refs, uint32_t* tris, int device_ptrs);
int lrt_scene_sync(lrt_scene* scene);
/* mean duration (ms) of the closest-hit kernel launches since the last reset,
 * timed with HIP events on the scene's stream */
int lrt_scene_kernel_timing(lrt_scene* scene, int kernel, int reset,
                            double* total_ms, uint64_t* launches, uint64_t* items);
/* kernel kinds: CLOSEST = the path tracer's traversal launches (k_extend, k_trace_queues) and the batch
 * closest-hit queries; OCCLUDED = the batch shadow queries; SHADE = a bounce's binning and shading;
 * GATHER = the photon gathers; TAIL = the path tracer's tail launches (each finishes the paths
 * still live at a bounce at or past the roulette depth once few are left: traversal and shading) */
enum { LRT_KERNEL_CLOSEST = 0, LRT_KERNEL_OCCLUDED = 1, LRT_KERNEL_SHADE = 2, LRT_KERNEL_GATHER = 3,
       LRT_KERNEL_TAIL = 4, LRT_KERNEL_COUNT_ = 5 };

/* ---- BSDF batch evaluation (bsdf.h:78-88) -----------------------------
 * For n tuples i: material index mat[i] into the scene's (or the given)
 * material table, local wi/wo (3 floats each), front_face; outputs f*cos
 * (3 floats) and pdf for the given measure and transport mode.
 * component = -1 and type_mask = 0xffffffff reproduce the path integrator. */
int lrt_bsdf_eval_pdf(int hip_device, const lrt_material* materials, uint32_t n_materials,
                      const float* tables, uint32_t n_tables,
                      const int32_t* mat, const float* wi, const float* wo,
                      const uint8_t* front_face, uint32_t n,
                      int measure, int mode, int component, uint32_t type_mask,
                      float* f_cos, float* pdf);
/* d (f*cos) / d theta of BSDF tuples, as the differentiable renderer's adjoint
 * uses it (no counterpart in the reference): which 0 = GGX alpha, 1 = the
 * interior IOR; exact (dual numbers) for the rough conductor and the rough
 * dielectric, central differences (step 1e-3) for other types.  Solid-angle
 * measure, all components, df: 3 floats per tuple. */
int lrt_bsdf_param_grad(int hip_device, const lrt_material* materials, uint32_t n_materials,
                        const float* tables, uint32_t n_tables, const int32_t* mat,
                        const float* wi, const float* wo, const uint8_t* front_face, uint32_t n,
                        int mode, int which, float* df);
/* bsdf::sample for n tuples with sample u (2 floats) and lobe sample
 * (ctx.lobe_sample); outputs weight (3), wo (3), pdf, eta, sampled_type,
 * sampled_component. */
int lrt_bsdf_sample(int hip_device, const lrt_material* materials, uint32_t n_materials,
                    const float* tables, uint32_t n_tables,
                    const int32_t* mat, const float* wi, const uint8_t* front_face,
                    const float* u, const float* lobe_u, uint32_t n,
                    int mode, int component, uint32_t type_mask,
                    float* weight, float* wo, float* pdf, float* eta,
                    uint32_t* sampled_type, int32_t* sampled_component);
/* bsdf::flags / component_count for materials[index] (bsdf.h:58-62) */
int lrt_bsdf_flags(const lrt_material* materials, uint32_t n_materials, int32_t index,
                   uint32_t* flags, int32_t* component_count);

/* ---- host precompute the reference runs in material constructors ------- */
/* fresnel_diffuse_reflectance(eta): adaptive Gauss-Lobatto, FP32
 * (fresnel.h:128-133, gauss_lobatto.h:23-121) */
int lrt_fresnel_diffuse_reflectance(float eta, float* out);
/* rough_transmittance(eta, alpha) slice from a ggx.dat image held in memory
 * (rough_transmittance.h:17-24, 67-185): out[0..99] theta samples,
 * out[100] diffuse transmittance */
int lrt_rough_transmittance_slice(const void* table_file, size_t table_bytes,
                                  float eta, float alpha, float* out101);

/* ---- sky model state (ArHosekSkyModel.cc:147-289, 602-635) --------------
 * Cooks the Hosek-Wilkie RGB sky state that arhosek_rgb_skymodelstate_alloc_init
 * computes: config27 = 9 configuration values per channel (R, G, B), radiance3
 * = the radiance scale per channel, from the model's RGB datasets
 * (coefficients: 3 x 1080 doubles, radiances: 3 x 120 doubles, the layout of
 * ray-tracing_amd/data/hosek_rgb_dataset.json).  turbidity in [1, 10], albedo
 * in [0, 1]; an elevation below 0 gives NaN, as in the model.  The result fills
 * lrt_light_desc.config / channel_radiance of a sky light. */
int lrt_sky_cook(const double* coefficients, const double* radiances, double turbidity, double albedo,
                 double elevation, double* config27, double* radiance3);
/* the solar elevation env_emitter cooks the model at (env_emitter.h:13-30):
 * asin of the normalised (float) sun direction's y, clamped to [-1, 1] */
int lrt_sky_elevation(const float* sun_dir, double* elevation);

/* ---- sky emitter evaluation (env_emitter.h:71-88) ---------------------
 * rgb radiance of n world directions for light lights[0] of a scene
 * description (device evaluation) */
int lrt_sky_eval(int hip_device, const lrt_light_desc* sky, const float* dirs, uint32_t n, float* rgb);

/* ---- path tracer (path_integrator.h:18-170 over src/main.cpp:103-147) --- */
typedef struct lrt_render_params {
    uint32_t width, height;
    uint32_t spp;          /* samples per pixel rendered by this call */
    uint32_t spp_offset;   /* index of the first sample (progressive / resume) */
    uint32_t max_depth;    /* path_integrator.h:173 (reference: 50) */
    uint32_t rr_depth;     /* path_integrator.h:19  (reference: 5) */
    uint64_t seed;         /* Philox-4x32-10 key */
    uint32_t row_begin, row_end; /* rows [begin, end) of the image; 0,0 = all */
    uint32_t flags;        /* LRT_RENDER_COUNT: also count traversal work (stats);
                              LRT_RENDER_REFERENCE_SAMPLER: draw the samples from the
                              reference's own Sobol sampler (sampler.h, sobol.h) instead of
                              Philox (seed unused), so each sample follows the reference's path */
} lrt_render_params;
enum { LRT_RENDER_COUNT = 1, LRT_RENDER_REFERENCE_SAMPLER = 2,
       /* lrt_diff_render / lrt_diff_adjoint only: the per-sample kernels (one thread
          traces a whole sample) instead of the wavefront; the same estimator, an
          independent implementation kept for checking */
       LRT_DIFF_PER_SAMPLE = 4,
       /* lrt_render: every bounce as a wavefront launch (the last bounces of a pass are
          otherwise finished by one lane per path once few paths are left; the same
          arithmetic, the same results) */
       LRT_RENDER_WAVEFRONT_ONLY = 8 };

typedef struct lrt_render_stats {
    uint64_t samples, extension_rays, shadow_rays, segments;
    double seconds;        /* wall time of the render loop */
    double trace_ms, occluded_ms, shade_ms; /* summed kernel time, HIP events */
    uint64_t trace_launches, occluded_launches;
    /* with LRT_RENDER_COUNT: extension-ray traversal work (nodes visited,
       leaf references read, triangle tests), from a separate counting kernel */
    uint64_t nodes_visited, refs_read, tris_tested;
    /* with LRT_RENDER_COUNT: the same for the shadow rays, counted as the
       reference's closest-hit shadow query over [0.001, tmax]
       (path_integrator.h:113-116); lrt_render only */
    uint64_t shadow_nodes_visited, shadow_refs_read, shadow_tris_tested;
    /* with LRT_RENDER_COUNT, lrt_render only: the work the timed traversal
       kernels do themselves, counted by the same rules: extension rays
       traversed over the window clipped to each mesh's box, shadow rays clipped
       too and ended by their first blocking hit (closest-hit form toward a
       non-convex emitter) */
    uint64_t own_nodes_visited, own_refs_read, own_tris_tested;
    uint64_t own_shadow_nodes_visited, own_shadow_refs_read, own_shadow_tris_tested;
} lrt_render_stats;

/* Renders and ADDS the per-pixel sum of sample radiance into sum_rgb
 * (3*W*H doubles, row y=0 at the bottom like src/main.cpp:122) and, if
 * sumsq_rgb is non-null, the per-pixel sum of squares (for standard errors).
 * device_ptrs: both buffers are device memory. */
int lrt_render(lrt_scene* scene, const lrt_render_params* params,
               double* sum_rgb, double* sumsq_rgb, int device_ptrs,
               lrt_render_stats* stats);

/* Renders like lrt_render (the pixel sums are discarded) and copies the rays
 * the frame traces after its camera rays, in the layout lrt_trace_closest
 * takes (8 arrays of *n floats back to back): the extension rays of bounces
 * >= 1 (window [0.001, inf), path_integrator.h:140) into ext_rays and the
 * next-event shadow rays (window [0.001, tmax], path_integrator.h:104-116)
 * into shadow_rays, with the emitter each shadow ray ignores in
 * shadow_ignore.  Rays beyond a capacity are dropped; *n_ext / *n_shadow
 * return how many were stored.  Host pointers.  A test and benchmark source
 * of the frame's own incoherent rays; the reference has no counterpart. */
int lrt_render_capture_rays(lrt_scene* scene, const lrt_render_params* params,
                            float* ext_rays, uint32_t ext_cap, float* shadow_rays,
                            int32_t* shadow_ignore, uint32_t shadow_cap,
                            uint32_t* n_ext, uint32_t* n_shadow);

/* ---- photon mapping (photon_integrator.h, photon_gather.h, photon_tracer.h,
 *      photon_map.h) ---------------------------------------------------------
 * A scene holds two photon maps, as photon_integrator does: kind 0 the global
 * (surface) map and kind 1 the caustic map.  lrt_photon_build traces
 * particles 0, 1, 2, ... in order until `requested` photons are returned
 * (photon_gather.h:15-101 with one particle per work unit), keeps the first
 * `requested`, and builds the reference's median kd tree over them
 * (photon_map.h:83-114): by default on the device (see `tree`).  A gather that fails (photon_gather.h:99-101) leaves
 * the map empty, like photon_integrator::preprocess (:25-36). */
typedef struct lrt_photon_params {
    int32_t kind;          /* 0 global (surface) map, 1 caustic map */
    uint32_t requested;    /* photons to store (photon_integrator.h:62-63: 250000) */
    uint32_t max_depth;    /* photon path depth (photon_integrator.h:27 passes max_depth - 1 = 127) */
    uint32_t rr_depth;     /* photon_integrator.h:59: 5 */
    uint64_t seed;
    float radius;          /* query radius (photon_integrator.h:17-18: 0.05 / 0.0125 x bounding radius) */
    uint32_t k;            /* photons per lookup, <= 128 (photon_integrator.h:64: 120) */
    uint32_t tree;         /* LRT_PHOTON_TREE_DEVICE: the median tree built on the device from the
                              photons in HBM (split = the median photon's coordinate, exact search);
                              LRT_PHOTON_TREE_REFERENCE: the reference's std::nth_element tree and
                              split read, built on the host (positions copied out and back) */
} lrt_photon_params;
enum { LRT_PHOTON_TREE_DEVICE = 0, LRT_PHOTON_TREE_REFERENCE = 1 };

typedef struct lrt_photon_stats {
    uint64_t photons, shot, particles_traced, returned;
    int32_t full;          /* photons == requested (photon_gather.h:59) */
    double trace_ms, sort_ms, seconds;
} lrt_photon_stats;

int lrt_photon_build(lrt_scene* scene, const lrt_photon_params* params, lrt_photon_stats* stats);
/* the map's photons in particle order: 12 floats (position, direction,
 * normal, power) + (depth, particle) per photon; shot = particles traced */
int lrt_photon_export(lrt_scene* scene, int32_t kind, float* photons12, uint32_t* depth_particle, uint64_t capacity,
                      uint64_t* n, uint64_t* shot);
/* builds map `kind` from given photons (scale = 1 / shot; n = 0 empties it) */
int lrt_photon_map_load(lrt_scene* scene, int32_t kind, const float* photons12, const uint32_t* depth_particle,
                        uint64_t n, float scale, float radius, uint32_t k);
/* the kd tree of map `kind` (no counterpart in the reference, whose tree is
 * private): nodes, 4 words each in pre-order (interior {split bits, axis, left,
 * right}, leaf {0, 3, first, count}), and order[i] = the particle-order index
 * (lrt_photon_export's) of the photon at tree position i.  Null buffers: only
 * *n_nodes and *n_photons are returned. */
int lrt_photon_map_tree(lrt_scene* scene, int32_t kind, uint32_t* nodes4, uint64_t node_capacity,
                        uint32_t* order, uint64_t order_capacity, uint64_t* n_nodes, uint64_t* n_photons);
/* photon_map::estimate_irradiance of map `kind` at n (point, normal) queries (6 floats) */
int lrt_photon_estimate(lrt_scene* scene, int32_t kind, const float* queries6, uint32_t n, int max_depth, float* out3);
/* photon_integrator::radiance (photon_integrator.h:39-123) for every (pixel,
 * sample) with the scene's maps (an empty map is skipped); max_depth is the
 * integrator's depth limit (photon_integrator.h:57: 128), rr_depth unused.
 * stats->segments = photon gathers; with LRT_RENDER_COUNT, nodes_visited /
 * refs_read / tris_tested count the gathers' kd nodes, photon positions tested
 * and photons summed (the k-nearest sets).  Gather kernel time: timer LRT_KERNEL_GATHER. */
int lrt_render_photon(lrt_scene* scene, const lrt_render_params* params, double* sum_rgb, double* sumsq_rgb,
                      int device_ptrs, lrt_render_stats* stats);

/* ---- seeded scene generators for the benchmark configurations --------- */
/* Generated geometry lives in a library-owned lrt_generated; its desc()
 * pointer stays valid until lrt_generated_free. */
typedef struct lrt_generated lrt_generated;
enum {
    LRT_GEN_BOX_ICOSPHERE = 0,   /* closed box room + icosphere L5 + ceiling light */
    LRT_GEN_DISPLACED_SPHERE = 1, /* displaced sphere, 7 material regions, ground, sky */
    LRT_GEN_PHOTON_BOX = 2,       /* box room + displaced sphere (7 regions, radius 0.5) + 0.5x0.5 light */
    LRT_GEN_DISPLACED_SPHERE_VARIANT = 3, /* kind 1 with seeded material colours, roughness and glass index */
    LRT_GEN_INVERSE_SPHERE = 4    /* displaced sphere in one diffuse material, rough copper ground, rough
                                     glass ball, sky: the differentiable renderer's scene (no counterpart) */
};
typedef struct lrt_gen_params {
    int32_t kind;
    uint64_t seed;
    uint32_t segments, rings;    /* displaced sphere tessellation (1024 x 512 -> 1,048,576 tris) */
    uint32_t subdivisions;       /* icosphere level */
    float aspect;                /* camera aspect ratio */
} lrt_gen_params;
int lrt_generate_scene(const lrt_gen_params* params, const float* tables, uint32_t n_tables,
                       const lrt_light_desc* sky, lrt_generated** out);
/* the scene of src/main.cpp:55-87 (the README renders) around the mesh of an
 * OBJ file (load_obj, obj_loader.h:44-141): the mesh with materials[0] (the
 * others are materials it nests), the ground sphere (material n_materials),
 * the rectangle light and the camera of main.cpp; tables: the
 * rough-transmittance slices the materials need.  LRT_ERR_ARG when load_obj
 * fails. */
int lrt_generate_obj_scene(const char* obj_path, const lrt_material* materials, uint32_t n_materials,
                           const float* tables, uint32_t n_tables, float aspect, lrt_generated** out);
const lrt_scene_desc* lrt_generated_desc(const lrt_generated* g);
int lrt_generated_free(lrt_generated* g);

/* ---- OBJ meshes (obj_loader.h:10-141) ------------------------------------
 * lrt_load_obj reads "v", "vn" and "f" records (indices 1-based or negative,
 * corners p, p/t, p/t/n, p//n), fans polygons around their first corner,
 * drops zero-area triangles, gives each corner the file's normal or the face
 * normal, and centres and scales the mesh to a unit longest bounding-box
 * side, as the reference's load_obj.  Where that returns nullptr (file not
 * readable, index out of range, a face of fewer than 3 corners, no triangle)
 * or would throw (an unparsable index) the call returns LRT_ERR_ARG.
 * lrt_mesh_view's arrays stay valid until lrt_mesh_free. */
typedef struct lrt_mesh lrt_mesh;
int lrt_load_obj(const char* path, lrt_mesh** out);
int lrt_mesh_view(const lrt_mesh* mesh, uint32_t* n_vertices, uint32_t* n_triangles, const float** positions,
                  const uint32_t** indices, const float** corner_normals);
int lrt_mesh_free(lrt_mesh* mesh);

/* ---- image files (image_write.cpp:10-100) ---------------------------------
 * rgb: width x height float RGB, row 0 first.  8-bit formats map each channel
 * as the reference's to_byte (NaN -> 0, pow(v, 1/2.2), clamp to [0, 0.999],
 * x 256, truncated); PNG rows are flipped when asked (src/main.cpp:151 saves
 * the bottom-up frame flipped) and hold the same pixels as the reference's
 * (stored deflate blocks instead of stb's compression); PPM (its plain P3)
 * and HDR (Radiance RGBE) keep the rows in memory order, as the reference's
 * writer does; JPEG (lossy) is not written.  Lossless additions: PFM and
 * raw little-endian float32 with a JSON sidecar (path + ".json"), rows
 * bottom first (flipped first when asked); OpenEXR (uncompressed scanlines,
 * 32-bit float B, G, R channels, rows top first, so a bottom-up frame is
 * stored upright unless flipped). */
enum { LRT_IMAGE_PNG = 0, LRT_IMAGE_JPEG = 1, LRT_IMAGE_PPM = 2, LRT_IMAGE_HDR = 3, LRT_IMAGE_PFM = 4,
       LRT_IMAGE_RAW = 5, LRT_IMAGE_EXR = 6 };
int lrt_image_write(const char* path, const float* rgb, uint32_t width, uint32_t height, int format,
                    int flip_vertically);
int lrt_image_to_bytes(const float* rgb, uint32_t width, uint32_t height, int flip_vertically, uint8_t* out);

/* ---- FP64 accumulators of a progressive render (no counterpart in the
 *      reference, which keeps its FP32 frame in memory, src/main.cpp:105) --
 * The per-pixel sums (and sums of squares) lrt_render adds into, written as
 * little-endian float64 (sum, then sumsq when non-null; rows bottom first)
 * with a JSON sidecar path + ".json" holding the shape and what a resumed
 * render needs: continue with spp_offset = next_sample and the same seed and
 * depth limits.  lrt_accum_read with sum = NULL reads the sidecar only;
 * sumsq is zero-filled when the file holds no sums of squares. */
typedef struct lrt_accum_info {
    uint32_t width, height;
    uint64_t spp_done;     /* samples per pixel summed so far */
    uint64_t next_sample;  /* index of the next sample to render */
    uint64_t seed;
    uint32_t max_depth, rr_depth;
    int32_t has_sumsq;     /* read: the file holds sums of squares */
} lrt_accum_info;
int lrt_accum_write(const char* path, const double* sum, const double* sumsq, const lrt_accum_info* info);
int lrt_accum_read(const char* path, double* sum, double* sumsq, lrt_accum_info* info);

/* ---- Philox-4x32-10 (the counter-based sampler of every kernel) -------- */
int lrt_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- first-hit features for the denoiser (no counterpart in the reference:
 *      its renderer has no denoiser) --------------------------------------
 * Per pixel, averaged over the frame's jittered camera rays (the same rays
 * lrt_render traces first): albedo (3: diffuse colour, normal-incidence
 * Fresnel for conductors, 1 for dielectrics, emitters and misses), shading
 * normal (3, 0 on a miss) and hit distance (1, 0 on a miss); 7 floats per
 * pixel, rows bottom-up like the radiance.  device_ptrs: aov is device memory. */
int lrt_render_aov(lrt_scene* scene, const lrt_render_params* params, float* aov, int device_ptrs);
/* lrt_render and lrt_render_aov in one pass (no counterpart in the reference):
 * the features come from the render's own camera rays, found by its first
 * bounce, instead of a second trace of them; for Philox renders both outputs
 * equal the two separate calls' bit for bit (with LRT_RENDER_REFERENCE_SAMPLER
 * the features follow that sampler's camera rays) */
int lrt_render_features(lrt_scene* scene, const lrt_render_params* params, double* sum_rgb, double* sumsq_rgb,
                        float* aov, int device_ptrs, lrt_render_stats* stats);

/* ---- differentiable rendering (no counterpart in the reference) ---------
 * The path tracer of lrt_render as one kernel thread per pixel (identical
 * radiance per sample), with parameters theta: an albedo texture on one
 * diffuse material (nearest texel; a point's direction from the origin
 * projected along z onto the unit disk, points behind z = 0 mirrored, and the
 * disk mapped onto the texture square by the inverse concentric map), the
 * GGX alpha of up to two materials and the interior IOR of one.  theta holds
 * tex_height x tex_width RGB texels (row-major), then alpha[0], alpha[1] and
 * the IOR (3 slots always; unused ones are ignored).  lrt_diff_adjoint
 * returns d(sum(d_image * image)) / d theta by path replay (sampling, MIS
 * and roulette detached); images are width x height x 3 floats, bottom-up. */
typedef struct lrt_diff_params {
    int32_t tex_material;          /* diffuse material textured by theta, -1: none */
    uint32_t tex_width, tex_height;
    int32_t alpha_material[2];     /* conductor / rough dielectric whose alpha is a parameter, -1: none */
    int32_t ior_material;          /* dielectric / rough dielectric whose interior IOR is a parameter, -1: none */
} lrt_diff_params;
int lrt_diff_param_count(const lrt_diff_params* dp, uint32_t* n);
int lrt_diff_render(lrt_scene* scene, const lrt_diff_params* dp, const float* theta, const lrt_render_params* params,
                    float* image, int device_ptrs);
int lrt_diff_adjoint(lrt_scene* scene, const lrt_diff_params* dp, const float* theta, const lrt_render_params* params,
                     const float* d_image, float* grad, int device_ptrs);
/* the estimator lrt_diff_adjoint differentiates: BSDF values and texture at
 * theta, directions, pdfs and MIS weights sampled with theta_sampling's
 * scalars; its derivative in theta at theta == theta_sampling is the
 * adjoint's gradient for the same samples (host arrays; for gradient checks) */
int lrt_diff_render_held(lrt_scene* scene, const lrt_diff_params* dp, const float* theta, const float* theta_sampling,
                         const lrt_render_params* params, float* image);

/* ---- learned denoiser (no counterpart in the reference) -------------------
 * A 3-level U-Net (32 / 64 / 128 channels, 3x3 convolutions) over radiance +
 * first-hit features, run and trained with the library's BF16 matrix-core
 * kernels (FP32 accumulation, FP32 master weights, Adam).  Images are
 * B x H x W pixels (H, W multiples of 4), radiance 3 floats and aux 7 floats
 * (lrt_render_aov) per pixel; the output is the denoised radiance. */
typedef struct lrt_denoiser lrt_denoiser;
int lrt_denoiser_create(int hip_device, uint64_t seed, lrt_denoiser** out);
int lrt_denoiser_release(lrt_denoiser* d);
int lrt_denoiser_param_count(const lrt_denoiser* d, uint64_t* n);
/* copies the FP32 parameters out (get) and/or in (set; resets Adam) */
int lrt_denoiser_params(lrt_denoiser* d, float* get, const float* set);
int lrt_denoiser_forward(lrt_denoiser* d, const float* radiance, const float* aux, uint32_t B, uint32_t H, uint32_t W,
                         float* out, int device_ptrs);
/* L1 loss in log(1 + radiance) against target and its parameter gradient
 * (kept for lrt_denoiser_adam_step; copied to grads when non-null) */
int lrt_denoiser_loss_grad(lrt_denoiser* d, const float* radiance, const float* aux, const float* target, uint32_t B,
                           uint32_t H, uint32_t W, float* loss, float* grads, int device_ptrs);
int lrt_denoiser_adam_step(lrt_denoiser* d, float lr);
/* per-layer timing of lrt_denoiser_forward: returns the milliseconds each of
 * the 10 convolution layers took (HIP events on the denoiser's stream,
 * summed over the forwards since the previous call; layer_ms may be NULL)
 * and the number of those forwards, clears the sums, and turns timing on or
 * off for the following forwards */
int lrt_denoiser_timing(lrt_denoiser* d, int enable, double* layer_ms, uint64_t* calls);
/* one 3x3 convolution layer on host FP32 arrays (rounded to BF16 on the
 * device): op 0 out = act(conv(x, w) + b) [B,H,W,cout]; op 1 the data
 * gradient out = conv^T(g, w) [B,H,W,cin]; op 2 the weight and bias gradient
 * out = (dW [9][cin][cout], db [cout]) of x and g.  w: [9][cin][cout]. */
int lrt_conv3x3(int hip_device, int op, const float* x, const float* w, const float* b, const float* g, uint32_t B,
                uint32_t H, uint32_t W, uint32_t cin, uint32_t cout, int relu, float* out);

#ifdef __cplusplus
}
#endif
#endif /* LINRT_H */
