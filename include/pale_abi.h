/*
 * pale_abi.h -- C ABI of the MI355X beta-surfel adjoint renderer.
 *
 * This is the drop-in boundary for the reference's pybind11 module `pale`
 * (Pale/pale_bindings.cpp:1613-1650, class `Renderer`).  Every entry point
 * below replaces one method of that class; the comment on each names the
 * reference method it stands in for.  The Python shim
 * differentiablepointrendering_amd/pale.py binds these with ctypes and
 * restores the reference's Python call shapes (dicts of numpy arrays).
 *
 * Conventions
 *  - Only plain C types cross the boundary.  All arrays are float32 / int
 *    C-contiguous host buffers owned by the caller; the library never keeps a
 *    caller pointer after a call returns.  All device memory belongs to the
 *    handle.
 *  - Every function returns an int status: PALE_OK (0) or a negative code.
 *    pale_last_error() returns a thread-local message for the last failure.
 *    The Python shim maps PALE_ERR_RANGE to IndexError and every other code to
 *    RuntimeError, as pybind11 maps std::out_of_range / std::runtime_error in
 *    the reference.
 *  - Point arrays are in BVH order (the order the reference's
 *    get_point_parameters returns, pale_bindings.cpp:734-850).
 *  - Images are row-major (H, W, 4) float32 with row 0 at the top of the
 *    image (ray v = H - (py + 0.5 + jy), KernelHelpers.h:776).
 */
#ifndef PALE_ABI_H
#define PALE_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PALE_NAME_LEN 16

enum {
    PALE_OK = 0,
    PALE_ERR_INVALID_ARG = -1, /* bad pointer, shape, key or dtype      */
    PALE_ERR_SIZE = -2,        /* size / resolution mismatch             */
    PALE_ERR_RANGE = -3,       /* index out of range (IndexError)        */
    PALE_ERR_DEVICE = -4,      /* HIP runtime or allocation failure      */
    PALE_ERR_INTERNAL = -5     /* anything else                          */
};

/* One triangle-mesh instance (the reference's MeshComponent + Material +
 * Transform, SceneBuild.cpp:72-177).  A mesh with power > 0 is an area
 * emitter (GPUMaterial::isEmissive, GPUDataStructures.h:93-95); its emission
 * colour is base_color (MaterialBakerMitsuba.ixx:42-66). */
typedef struct pale_mesh_desc {
    const float* positions;      /* [vertex_count*3] object-space vertices */
    uint32_t vertex_count;
    const uint32_t* indices;     /* [triangle_count*3]                     */
    uint32_t triangle_count;
    float object_to_world[16];   /* row-major 4x4                          */
    float world_to_object[16];   /* row-major 4x4 (inverse of the above)   */
    float base_color[3];
    float power;
    char name[PALE_NAME_LEN];
} pale_mesh_desc;

/* One pinhole camera (CameraGPU, GPUDataStructures.h:126-153). */
typedef struct pale_camera_desc {
    char name[PALE_NAME_LEN];
    uint32_t width;
    uint32_t height;
    float fovy_deg;
    float view[16];              /* world -> camera, row-major            */
    float inv_view[16];          /* camera -> world, row-major            */
    float position[3];
    float forward[3];            /* unit world-space viewing direction    */
    int32_t use_for_adjoint;     /* CameraGPU::useForAdjointPass          */
} pale_camera_desc;

/* The surfel cloud (PointGeometry of the reference's PLY asset,
 * SceneBuild.cpp:181-227).  `shape` may be NULL (zeros). */
typedef struct pale_points_desc {
    uint32_t count;
    const float* position;   /* [count*3] */
    const float* tangent_u;  /* [count*3] */
    const float* tangent_v;  /* [count*3] */
    const float* scale;      /* [count*2] */
    const float* albedo;     /* [count*3] */
    const float* opacity;    /* [count]   */
    const float* beta;       /* [count]   */
    const float* shape;      /* [count] or NULL */
} pale_points_desc;

/* Whole scene: meshes become instances first (in order), the surfel cloud
 * is the last instance with an identity transform, as the reference's
 * "Gaussian" entity is (pale_bindings.cpp:90-101, SceneBuild.cpp:26-46). */
typedef struct pale_scene_desc {
    const pale_mesh_desc* meshes;
    uint32_t mesh_count;
    const pale_camera_desc* cameras;
    uint32_t camera_count;
    pale_points_desc points;
} pale_scene_desc;

/* Renderer settings: the reference's settings-dict keys
 * (pale_bindings.cpp:123-139) with the PathTracerSettings defaults
 * (GPUDataStructures.h:400-422), plus keys the reference hard-codes. */
typedef struct pale_settings {
    uint32_t photons;          /* "photons"          photonsPerLaunch (1e6)   */
    uint32_t bounces;          /* "bounces"          maxBounces (6)           */
    uint32_t forward_passes;   /* "forward_passes"   numForwardPasses (6)     */
    uint32_t gather_passes;    /* "gather_passes"    numGatherPasses (6)      */
    uint32_t adjoint_bounces;  /* "adjoint_bounces"  maxAdjointBounces (6)    */
    uint32_t adjoint_passes;   /* "adjoint_passes"   adjointSamplesPerPixel(6)*/
    uint64_t seed;             /* "seed"             Random::seed (42)        */
    float gather_radius;       /* "gather_radius"    PathTracer.cpp:187 (0.01)*/
    uint32_t rr_start;         /* "rr_start"         russianRouletteStart (6) */
    float q_null;              /* AdjointSampleSettings (0.5)                */
    float q_reflect;           /* (0.5)                                      */
    float q_transmit;          /* (0.0)                                      */
    int32_t spp_decorrelate;   /* "spp_decorrelate": 0 = every gather pass
                                  reuses the camera seed, as the reference
                                  does (PrimalKernels.cpp:516-520); 1 = pass
                                  k uses seed + k (independent samples)     */
    int32_t debug_images;      /* "debug_images"                             */
    uint32_t gather_entry_budget; /* "gather_entry_budget": camera-gather
                                  entries per pixel held at once (0 = 24, at
                                  least 65536 in all); deeper images take
                                  more rounds, with the same result         */
    uint32_t photon_map_capacity; /* "photon_map_capacity": stored photons at
                                  most (0 = photons * forward_passes *
                                  bounces, clamped to 10 GiB of 64-byte
                                  records, as the reference allocates); past
                                  it deposits replace stored photons by the
                                  reference's reservoir sampling            */
    int32_t full_gradients;    /* "full_gradients": 0 = opacity only, as the
                                  reference (the other gradient arrays are
                                  zeros); 1 = also position, tangents,
                                  scale, albedo and beta, through alpha_geom
                                  and the BRDF of each adjoint event        */
    int32_t batch_passes;      /* "batch_passes" (default 1): 1 = render_forward walks
                                  the gather passes of all its cameras, and
                                  render_backward traces the adjoint samples
                                  of all its cameras, as batches (the
                                  launch shape of pale_set_batch_tuning);
                                  0 = one pass / sample after another.
                                  Same images either way; gradients equal
                                  up to the order of float atomics          */
    int32_t adjoint_decorrelate; /* "adjoint_decorrelate" (default 0): 0 =
                                  every adjoint sample seeds from
                                  adjoint_seed (42), as the
                                  reference (AdjointKernels.cpp:27,94: its
                                  random.number is never updated), so the
                                  samples of a pixel trace one path; 1 =
                                  sample k seeds from 42 + k (independent
                                  samples)                                  */
    uint32_t adjoint_seed;     /* "adjoint_seed" (default 42): the adjoint's
                                  base seed, Random::number
                                  (GPUDataStructures.h:395-398); with
                                  adjoint_decorrelate sample k seeds from
                                  adjoint_seed + k                          */
    uint32_t bvh_build;        /* "bvh_build" (default 0): where the surfel
                                  BVH is built (SceneBuild.cpp:642-772 on
                                  the host in the reference); 0 = on the
                                  GPU from 65,536 surfels, else on the
                                  host; 1 = on the GPU from 5 surfels; 2 =
                                  always on the host.  Nodes and
                                  permutation are the same either way; a
                                  scene with non-finite or -0 box
                                  coordinates is always built on the host */
} pale_settings;

/* Work counters of the last render calls (no reference counterpart; the
 * reference only logs wall-clock, ScopedTimer.h:15-56). */
typedef struct pale_stats {
    uint64_t photon_queries;    /* closest-hit queries of the photon pass    */
    uint64_t gather_queries;    /* closest-hit queries of the camera gather  */
    uint64_t adjoint_queries;   /* closest-hit queries of the adjoint pass   */
    uint64_t photons_deposited; /* attempted photon deposits                 */
    uint64_t gradient_events;   /* completed adjoint events                  */
    uint64_t grid_cells;        /* cells of the photon grid                  */
    double forward_ms;          /* device time of the last render_forward    */
    double backward_ms;         /* device time of the last render_backward   */
    /* Device time per kernel class (sum over launches) and launch counts. */
    double photon_trace_ms, gather_ms, adjoint_trace_ms, adjoint_transport_ms, grid_ms;
    uint32_t photon_trace_launches, gather_launches, adjoint_trace_launches, adjoint_transport_launches;
    /* Camera gather stages: closest-hit layer walk, and the entry sort plus
     * compositing; gather_ms above is its irradiance gathers.  camera_rounds
     * = entry-buffer rounds of the last render_forward (one per pass when
     * the buffer suffices). */
    double camera_trace_ms, camera_sort_ms;
    uint32_t camera_trace_launches, camera_rounds;
    /* Counting mode only (pale_set_count_work): work units of the
     * reference algorithm per kernel class, summed over the last
     * render_forward and any render_backward after it.  work[c][u] with
     * class c = PALE_WORK_PHOTON / _GATHER / _ADJOINT and unit u = popped
     * nodes, AABB slab tests, primitive tests, irradiance gathers, gather
     * cells visited, non-empty cells, photons scanned. */
    uint64_t work[3][7];
    /* The last surfel-BVH build (pale_create, rebuild_bvh and the point
     * edits that rebuild): host wall time of the whole scene build and
     * upload, and 1 when the surfel BLAS was built on the GPU (setting
     * bvh_build), else 0. */
    double bvh_build_ms;
    uint32_t bvh_on_device, bvh_pad;
} pale_stats;

enum { PALE_WORK_PHOTON = 0, PALE_WORK_GATHER = 1, PALE_WORK_ADJOINT = 2 };
enum {
    PALE_WORK_POPS = 0, PALE_WORK_AABB = 1, PALE_WORK_PRIMS = 2, PALE_WORK_GATHERS = 3,
    PALE_WORK_CELLS = 4, PALE_WORK_CELLS_NONEMPTY = 5, PALE_WORK_PHOTONS = 6
};

typedef struct pale_handle pale_handle;

/* Thread-local text of the last failure in this thread ("" if none). */
const char* pale_last_error(void);

/* Fill `out` with the reference defaults (GPUDataStructures.h:400-422,
 * PathTracer.cpp:187). */
void pale_default_settings(pale_settings* out);

/* Renderer(assetRootDir, sceneXml, pointCloudFile, settings)
 * (pale_bindings.cpp:61-170).  Scene/PLY parsing happens in the Python
 * shim; this call receives the parsed scene. */
int pale_create(const pale_scene_desc* scene, const pale_settings* settings,
                pale_handle** out);

/* ~PythonRenderer (pale_bindings.cpp:172-177). */
int pale_destroy(pale_handle* h);

/* get_camera_names (pale_bindings.cpp:1414-1420) -- count + per-camera info. */
int pale_camera_count(const pale_handle* h, uint32_t* count);
int pale_camera_info(const pale_handle* h, uint32_t index,
                     char name_out[PALE_NAME_LEN], uint32_t* width,
                     uint32_t* height, int32_t* use_for_adjoint);

/* render_forward(camera_name="") (pale_bindings.cpp:180-277).
 * camera_name NULL or "" renders every camera, else the cameras with that
 * name.  For the k-th rendered camera (camera order) the LDR image goes to
 * ldr_out[k] and the HDR ("_raw") image to raw_out[k], H*W*4 floats each;
 * either array of pointers may be NULL.  *rendered = number of cameras. */
int pale_render_forward(pale_handle* h, const char* camera_name,
                        float* const* ldr_out, float* const* raw_out,
                        uint32_t* rendered);

/* render_backward(targetRgb32f) (pale_bindings.cpp:279-731).
 * dloss_rgb[k] is the (H,W,3) dL/dI image of camera camera_names[k]; names
 * that are not adjoint cameras are skipped, as the reference skips them.
 * grads[0..7] receive position (N*3), tangent_u (N*3), tangent_v (N*3),
 * scale (N*2), albedo (N*3), opacity (N), beta (N), shape (N); any may be
 * NULL.  adjoint_source_out[k] (may be NULL) receives the (H,W,4) adjoint
 * source image of camera k. */
int pale_render_backward(pale_handle* h, uint32_t n,
                         const char* const* camera_names,
                         const float* const* dloss_rgb, float* const* grads,
                         float* const* adjoint_source_out);

/* get_point_parameters (pale_bindings.cpp:734-850). */
int pale_point_count(const pale_handle* h, uint32_t* count);
int pale_get_points(const pale_handle* h, float* position, float* tangent_u,
                    float* tangent_v, float* scale, float* albedo,
                    float* opacity, float* beta, float* shape);

/* apply_point_optimization(parameters) (pale_bindings.cpp:853-1056):
 * same N required; writes parameters in BVH order and uploads them without
 * rebuilding the BVH. */
int pale_apply_points(pale_handle* h, const pale_points_desc* points);

/* add_points({"new": {...}}) (pale_bindings.cpp:1221-1412). */
int pale_add_points(pale_handle* h, const pale_points_desc* points);

/* remove_points({"indices": ...}) (pale_bindings.cpp:1081-1218). */
int pale_remove_points(pale_handle* h, const int64_t* indices, size_t n);

/* rebuild_bvh() (pale_bindings.cpp:1058-1079). */
int pale_rebuild_bvh(pale_handle* h);

/* set_point_properties(translation3, rotation_quat4 (x,y,z,w), scale3,
 * albedo3, opacity, beta, index=-1) (pale_bindings.cpp:1486-1587). */
int pale_set_point_properties(pale_handle* h, const float translation[3],
                              const float rotation_xyzw[4],
                              const float scale[3], const float albedo[3],
                              float opacity, float beta, int32_t index);

/* set_point_opacity / set_point_beta (pale_bindings.cpp:1422-1483). */
int pale_set_point_opacity(pale_handle* h, float opacity, int32_t index);
int pale_set_point_beta(pale_handle* h, float beta, int32_t index);

/* ---- device-memory variants: no reference counterpart (the reference's
 * bindings copy through host numpy arrays, pale_bindings.cpp:180-277,
 * 279-731).  For callers whose tensors live on the GPU (a
 * torch.autograd.Function).  Pointers are device addresses on the handle's
 * device; each call returns after its results are written, and reads its
 * inputs after the caller's own work on them has finished (the caller
 * synchronises its stream first). ---- */

/* images["debug"][camera]["opacity"] of render_backward when the setting
 * debug_images is on (pale_bindings.cpp:627-726; written by
 * launchAdjointTransportKernel, AdjointKernels.cpp:434-505): per pixel the
 * sum of the opacity-gradient contributions of its adjoint events in RGB,
 * alpha 1 where any event landed (atomicAddFloat4ToImage,
 * KernelHelpers.h:1489-1502), from the last pale_render_backward.
 * rgba_out: caller-allocated H*W*4 floats.  The reference's other seven
 * debug images (position_x/y/z, rotation, scale, albedo, beta) are never
 * written by its kernels and stay zero. */
int pale_get_debug_image(const pale_handle* h, const char* camera_name,
                         float* rgba_out);

/* render_forward with device outputs: H*W*4 floats per rendered camera. */
int pale_render_forward_device(pale_handle* h, const char* camera_name,
                               float* const* ldr_dev, float* const* raw_dev,
                               uint32_t* rendered);

/* render_backward with device dL/dI images (H,W,3) and device gradient
 * arrays (same order and sizes as pale_render_backward; any may be NULL). */
int pale_render_backward_device(pale_handle* h, uint32_t n,
                                const char* const* camera_names,
                                const float* const* dloss_rgb_dev,
                                float* const* grads_dev);

/* No counterpart in the reference: pale_render_view_backward_device's
 * opacity gradient for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It allocates nothing: every intermediate of the adjoint lives
 * in a static device array of the library (about 21 MB, one per GPU, held
 * until the process ends), so views are limited to 65536 pixels (256 x 256);
 * larger views return PALE_ERR_SIZE.  dloss_rgb_dev: (H,W,3) floats,
 * opacity_grad_dev: N floats (BVH order), both device memory.  Other
 * gradients are not computed.  All calls on one GPU share the array: calls
 * that overlap on one GPU (from several threads or handles) race. */
int pale_render_view_backward_noalloc(pale_handle* h,
                                      const pale_camera_desc* camera,
                                      uint32_t spp,
                                      const float* dloss_rgb_dev,
                                      float* opacity_grad_dev);

/* Opacity of all n surfels (BVH order) from device memory, without a BVH
 * rebuild (apply_point_optimization semantics for one parameter). */
int pale_set_opacity_device(pale_handle* h, const float* opacity_dev,
                            uint32_t n);

/* apply_point_optimization(parameters) (pale_bindings.cpp:853-1056) with the
 * seven parameter arrays in device memory (BVH order; position, tangent_u,
 * tangent_v: n*3 floats, scale: n*2, albedo: n*3, opacity, beta: n): same N
 * required, values written as given, no BVH rebuild -- the pale_apply_points
 * result without the host round trip.  Host copies of the parameters
 * (pale_get_points and the point-editing calls) are refreshed from the
 * device when next needed. */
int pale_apply_points_device(pale_handle* h, uint32_t n, const float* position,
                             const float* tangent_u, const float* tangent_v,
                             const float* scale, const float* albedo,
                             const float* opacity, const float* beta);

/* ---- served views: no reference counterpart (the reference renders only
 * the cameras of its scene, pale_bindings.cpp:180-277).  For a server that
 * keeps a scene resident and renders views of it on request: the photon map
 * is built once and every view gathers from it.  A view is any camera
 * description (its name and use_for_adjoint are ignored); its spp gather
 * passes draw the seeds seed + 0 .. seed + spp - 1, so the image depends only
 * on (scene, settings, camera, spp, seed) and never on other requests.
 * Calls on one handle are serialised by the caller; handles on their own
 * host threads run concurrently. ---- */

/* Device bytes held by all renderers of the process (buffers grow on demand
 * and are kept until pale_destroy): the measure of HBM admission. */
int pale_device_bytes(uint64_t* out);

/* Device bytes that serving a width x height view of spp passes (adjoint
 * != 0: with its adjoint) would add to this renderer's buffers; 0 when they
 * already fit (with batch_passes, the passes of a batch share the buffers:
 * the batch tuning's slots per batch bound them). */
int pale_view_bytes(pale_handle* h, uint32_t width, uint32_t height,
                    uint32_t spp, int32_t adjoint, uint64_t* out);

/* Device bytes that pale_render_views of n views (then pale_render_view_
 * backward of each view with adjoint[i] != 0) would add; needs batch_passes
 * = 1 (render_views always batches; single views then do too). */
int pale_views_bytes(pale_handle* h, uint32_t n, const uint32_t* widths,
                     const uint32_t* heights, const uint32_t* spp,
                     const int32_t* adjoint, uint64_t* out);

/* Device bytes of this renderer's image-sized buffers (camera gather, view
 * images, adjoint): what pale_release_view_buffers frees. */
int pale_view_buffer_bytes(pale_handle* h, uint64_t* out);

/* Frees the buffers that grow with the image size (camera gather, view
 * images, adjoint); scene, BVH and photon map stay resident. */
int pale_release_view_buffers(pale_handle* h);

/* Photon stage of render_forward only (emission, bounces, grid). */
int pale_build_photon_map(pale_handle* h);

/* One view over the current photon map: H*W*4 floats into each non-NULL
 * output (tone-mapped, linear).  Fails if no photon map was built. */
int pale_render_view(pale_handle* h, const pale_camera_desc* camera,
                     uint32_t spp, uint32_t seed, float* ldr_out,
                     float* raw_out);

/* pale_render_view into device memory. */
int pale_render_view_device(pale_handle* h, const pale_camera_desc* camera,
                            uint32_t spp, uint32_t seed, float* ldr_dev,
                            float* raw_dev);

/* Adjoint of one view: spp adjoint samples of dL/dI (H,W,3) through the
 * current photon map; gradients as pale_render_backward. */
int pale_render_view_backward(pale_handle* h, const pale_camera_desc* camera,
                              uint32_t spp, const float* dloss_rgb,
                              float* const* grads);

/* pale_render_view_backward with device dL/dI (H,W,3) and device gradient
 * arrays (as pale_render_backward_device). */
int pale_render_view_backward_device(pale_handle* h,
                                     const pale_camera_desc* camera,
                                     uint32_t spp, const float* dloss_rgb_dev,
                                     float* const* grads_dev);

/* ---- batched views: no reference counterpart (the reference renders one
 * camera after another, SyclBridge.cpp:144-157).  Many views of uneven
 * resolution and spp over the current photon map in one persistent
 * layer-walk launch per batch and round: pixel slots of all views and passes
 * are cut into batches of slots_per_batch; the rounds of a batch run as a
 * captured HIP graph when graph_rounds > 0.  Every image equals
 * pale_render_view(cameras[i], spp[i], seeds[i]) bit for bit, whatever the
 * tuning. ---- */

typedef struct pale_batch_tuning {
    uint32_t slots_per_batch;  /* pixel slots (pixels x passes) per batch; 0 = forward:
                                  16 passes of the largest view within [4Mi, 32Mi],
                                  adjoint: 4Mi                                         */
    uint32_t entries_per_slot; /* camera-gather entries per slot and round; 0 = 24        */
    uint32_t graph_rounds;     /* rounds in the captured graph; 0 = eager launches        */
    float grid_fraction;       /* layer-walk grid / full occupancy grid, (0,1]; 0 = 1     */
    uint32_t walk_lds_entries; /* layer-walk stack entries per lane held in LDS (the rest
                                  in a global spill area); 0 = 16, at most 64          */
} pale_batch_tuning;

int pale_set_batch_tuning(pale_handle* h, const pale_batch_tuning* tuning);
int pale_get_batch_tuning(const pale_handle* h, pale_batch_tuning* tuning);

/* n views; each output (H_i*W_i*4 floats) may be NULL.  batches_out (may be
 * NULL) receives the number of batches the views took. */
int pale_render_views(pale_handle* h, uint32_t n, const pale_camera_desc* cameras,
                      const uint32_t* spp, const uint32_t* seeds,
                      float* const* ldr_out, float* const* raw_out,
                      uint32_t* batches_out);

/* Adjoint of n views over the current photon map, gradients summed over the
 * views (arrays as pale_render_backward).  Each view contributes what
 * pale_render_view_backward(cameras[i], spp[i], dloss_rgb[i]) gives (the
 * same rays and events; float atomics add in another order). */
int pale_render_views_backward(pale_handle* h, uint32_t n,
                               const pale_camera_desc* cameras,
                               const uint32_t* spp,
                               const float* const* dloss_rgb,
                               float* const* grads, uint32_t* batches_out);

/* ---- diagnostics: no reference counterpart; used by the parity tests ---- */

/* Closest-hit query (the reference's intersectScene in FirstHit mode,
 * IntersectionKernels.h:954-1097) for n world rays.  Outputs per ray:
 * instance (-1 = miss), primitive (BVH-order surfel or triangle index),
 * world t, world hit position (3 floats) and alpha_geom. */
int pale_trace_rays(pale_handle* h, uint32_t n, const float* origins,
                    const float* directions, int32_t* instance_out,
                    int32_t* primitive_out, float* t_out, float* position_out,
                    float* alpha_out);

/* Irradiance estimate of the photon map of the last render_forward at n
 * points with n normals (3 floats each): the reference's device function
 * gatherDiffuseIrradianceAtPoint (KernelHelpers.h:990-1037), which no
 * binding exposes; a diagnostic for tests of the gather on its own. */
int pale_gather_irradiance(pale_handle* h, uint32_t n, const float* positions,
                           const float* normals, float* irradiance_out);

/* Surfel BLAS topology in the reference's BVHNode form
 * (GPUDataStructures.h:53-63): per node aabb min (3), aabb max (3),
 * leftFirst, triCount.  nodes_out may be NULL to query *node_count only.
 * permutation_out (may be NULL) receives, per BVH slot, the index of the
 * surfel in the order the cloud had before this build. */
int pale_get_surfel_bvh(const pale_handle* h, uint32_t* node_count,
                        float* bounds_out, uint32_t* left_first_out,
                        uint32_t* count_out, uint32_t* permutation_out);

/* Photon map of the last render_forward: *count stored photons (deposit
 * order); positions/powers/directions (3 floats each) may be NULL.
 * *attempted = the deposit counter T of the reference. */
int pale_get_photons(const pale_handle* h, uint32_t* count,
                     uint64_t* attempted, float* position, float* power,
                     float* direction);

/* Work counters and device times of the last render calls. */
int pale_get_stats(const pale_handle* h, pale_stats* out);

/* Counting mode on (1) / off (0): render calls run kernel variants that also
 * count the work units of pale_stats (slower; results are identical). */
int pale_set_count_work(pale_handle* h, int32_t on);

#ifdef __cplusplus
}
#endif

#endif /* PALE_ABI_H */
