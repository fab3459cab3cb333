/*
 * vgi.h — C ABI of the MI355X-native voxel global-illumination path.
 *
 * The path: triangles -> conservative voxelization into an integer colour/normal
 * accumulator -> radiance injection + 2x2x2 mip chain -> per-pixel diffuse and
 * specular cone tracing -> indirect-light buffer.
 *
 * Every entry point returns an int status (VGI_OK or a negative VGI_E_* code) and
 * never throws. vgi_last_error() returns the message of the calling thread's last
 * failure. All GPU work is asynchronous on the given stream (a hipStream_t passed as
 * void*; NULL selects the context's own stream): the voxelize, mips, trace and screen
 * entries enqueue their kernels and return without waiting for the device. The calls
 * that wait are named as such: vgi_sync, vgi_scene_get_stats (for the scene's last
 * voxelize), the parity readbacks, the glTF upload, the autotuner and the server's
 * drain/destroy; a call that must replace scratch grown too small first drains its own
 * stream (rare: scratch only grows).
 *
 * Reference interfaces each entry replaces (tobyc11/Final184, path:line):
 *   - CVoxelizeRenderer::RenderList / Render   Foreground/Renderer/VoxelizeRenderer.cpp:19-30,81-127
 *     (the voxelize draw per primitive, RG16UI imageStore in Pipelang/Internal/main.lua:242-275)
 *   - the per-frame clear of the voxel image    Foreground/Renderer/MegaPipeline.cpp:196
 *   - the lighting_indirect fullscreen pass     Foreground/Renderer/MegaPipeline.cpp:252-268
 *     (Shader/Lighting/indirect.frag:104-243)
 *   - the voxel image ownership                 Foreground/Renderer/MegaPipeline.cpp:470-490
 *   - G-buffer reconstruction in that pass      Shader/Lighting/indirect.frag:44-58,189-196
 *   - its temporal blend                        Shader/Lighting/indirect.frag:224-242
 *   - the indirect_blurX / blurY passes         Foreground/Renderer/MegaPipeline.cpp:270-284
 *     (Shader/Lighting/bilateralBlur.inc:15-72)
 *
 * Input/output buffers passed to vgi_voxelize / vgi_trace are DEVICE pointers
 * (the reference hands GPU vertex/index buffers and images to its passes); the
 * vgi_malloc/vgi_memcpy helpers exist for callers without their own allocator.
 */
#ifndef VGI_H
#define VGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VGI_ABI_VERSION 3u

#define VGI_MAX_LEVELS 11u   /* 1024^3 -> 1^3 */
#define VGI_MAX_LIGHTS 128u  /* the reference's LightLists holds 100 (MegaPipeline.cpp:101-105) */
#define VGI_MAX_DIFFUSE_CONES 32u
/* Snapped vertex coordinates must satisfy |v| <= VGI_GUARD_FIXED in 1/256-voxel units. */
#define VGI_GUARD_FIXED 262144
#define VGI_FIXED_BITS 8
/* Upper bound on samples per cone (a loop guard; real schedules stay far below). */
#define VGI_MAX_MARCH_STEPS 4096

typedef enum vgi_status {
  VGI_OK = 0,
  VGI_E_INVALID = -1,   /* bad argument */
  VGI_E_NOMEM = -2,     /* device or host allocation failed */
  VGI_E_HIP = -3,       /* HIP runtime error */
  VGI_E_BUDGET = -4,    /* refused by the memory budget */
  VGI_E_NOT_READY = -5, /* scene has no voxels / mips yet */
  VGI_E_RANGE = -6,     /* size beyond a documented maximum */
  VGI_E_NODEVICE = -7,  /* no usable gfx950 device */
  VGI_E_STATE = -8      /* object used in the wrong state */
} vgi_status;

typedef enum vgi_precision {
  VGI_PREC_FP32 = 0,  /* radiance RGBA fp32, 16 B/texel */
  VGI_PREC_FP16 = 1,  /* radiance RGB fp16 + opacity fp16, 8 B/texel */
  VGI_PREC_BF16 = 2,  /* radiance RGB bf16 + opacity fp16, 8 B/texel */
  VGI_PREC_FP16X2 = 3 /* the VGI_PREC_FP16 texels (same values), each stored beside its +x
                         neighbour as channel pairs (R_x R_x+1, G_x G_x+1, B.., A..: 16 B/texel)
                         so that the cone march interpolates along x with packed dot
                         products (FP16 products, FP32 accumulation) */
} vgi_precision;

typedef enum vgi_index_type {
  VGI_INDEX_NONE = 0, /* triangle i = vertices 3i, 3i+1, 3i+2 */
  VGI_INDEX_U16 = 1,
  VGI_INDEX_U32 = 2
} vgi_index_type;

typedef enum vgi_light_type { VGI_LIGHT_DIRECTIONAL = 0, VGI_LIGHT_POINT = 1 } vgi_light_type;

/* OR'd into vgi_light.type: the light casts shadows. Its contribution to an occupied voxel
 * is cut when a walk through the voxel occupancy toward the light meets an occupied voxel
 * first (the voxel-space counterpart of the reference's shadow-map test of each lit hit,
 * Shader/Lighting/indirect.frag:160-165). The walk, in voxel units of level 0:
 *  - start: the voxel centre moved one voxel along the dominant axis of the voxel's summed
 *    normal (ties to the lowest axis; no move when the sum is zero);
 *  - direction l toward the light: -vec (directional), or the light's position in voxel
 *    units minus the start (point); dominant axis a of l (ties to the lowest axis);
 *  - samples at start + m * l / |l_a|, m = 0, 1, ..., in 16.16 fixed point (the step rounded
 *    to nearest), each naming the voxel that contains it; the voxel itself is passed over;
 *  - lit when a sample leaves the grid, or (point) when m reaches ceil(|l_a|), the light's
 *    distance along a; unlit at the first occupied sample.
 * Every sample advances one voxel along a, so a walk takes at most n + 1 samples. */
#define VGI_LIGHT_SHADOW 0x100u

typedef struct vgi_context vgi_context;
typedef struct vgi_scene vgi_scene;

/* Voxel grid: the cube [origin, origin + extent]^3 split into n^3 voxels (n a power
 * of two, 8 <= n <= 1024), with `levels` mip levels (level k has n>>k voxels per axis,
 * 1 <= levels, (n >> (levels-1)) >= 1). Replaces the voxelizer camera's VoxelView /
 * VoxelProj pair (App/MainBehaviour.cpp:54-64, Foreground/SceneGraph/Camera.cpp:104-111). */
typedef struct vgi_grid {
  float origin[3];
  float extent;
  uint32_t n;
  uint32_t levels;
} vgi_grid;

/* The reference's voxelizer camera as a grid: the world-space image of the clip box
 * [-1, 1]^2 x [0, 1] under inverse(voxel_proj * voxel_view) (GLSL column-major mat4s as
 * the reference uploads VoxelProjection / VoxelView, MegaPipeline.cpp:241-246; the
 * camera of App/MainBehaviour.cpp:54-64 with Camera.cpp:104-111's orthographic matrix),
 * held in the cube with the box's centre and its longest side as edge, n voxels per axis.
 * Needs an orthographic projection and a view rotation that maps world axes to world axes
 * (VGI_E_INVALID otherwise). Replaces the VoxelView / VoxelProj pair as the voxelize
 * pass's grid input; the reference's 32 x 19 x 32 box at 128^3 becomes a 32-unit cube
 * of 0.25-unit voxels (the reference's are 0.25 x 0.148 x 0.25). Host only. */
int vgi_grid_from_voxel_camera(const float voxel_view[16], const float voxel_proj[16], uint32_t n,
                               uint32_t levels, vgi_grid* out);

/* One drawable primitive: CTriangleMesh (vertex + index buffer, element count;
 * Foreground/Shape/TriangleMesh.h:29-89) + CBasicMaterial albedo
 * (Foreground/Material/BasicMaterial.h) + the model matrix RenderList passes
 * (Foreground/Renderer/VoxelizeRenderer.cpp:114-120). All pointers are device pointers. */
typedef struct vgi_mesh {
  const float* positions;     /* xyz float32 per vertex */
  uint32_t position_stride;   /* bytes between vertices; 0 means 12 */
  uint32_t num_vertices;
  const void* indices;        /* NULL for VGI_INDEX_NONE */
  uint32_t index_type;        /* vgi_index_type */
  uint32_t num_triangles;
  const uint8_t* albedo_rgb8; /* 3 bytes per triangle, or NULL to use base_color */
  uint8_t base_color[4];      /* RGB8 (+ unused) used when albedo_rgb8 is NULL */
  float model[12];            /* row-major 3x4: world = M * (x, y, z, 1) */
  /* Texture (BasicMaterial with UseTextures, Pipelang/Internal/main.lua:188-205), used when
   * texcoords and texture_rgba8 are both set; albedo_rgb8 and base_color are then ignored.
   * A voxel covered by a triangle takes the albedo texture(uv) * base_color_factor, sampled
   * at the voxel centre projected along the triangle's dominant axis: barycentric weights
   * from the exact snapped coordinates (a centre outside the triangle, which conservative
   * coverage admits, has its negative weights clamped to zero), uv interpolated in single
   * precision, wrap addressing, bilinear filtering between texel centres at level 0. A
   * voxel whose sampled alpha is below 0.05 receives nothing from the triangle (the
   * reference's discard): no colour, normal or count, so alone it stays unoccupied.
   * RGB is then rounded to 8 bits for the integer sums. */
  const float* texcoords;       /* uv float32 x 2 per vertex (TEXCOORD_0) */
  uint32_t texcoord_stride;     /* bytes between vertices' uv; 0 means 8 */
  uint32_t texture_width;       /* 1 .. 16384 texels */
  uint32_t texture_height;      /* 1 .. 16384 texels */
  uint32_t texture_reserved;    /* 0 */
  const uint8_t* texture_rgba8; /* RGBA8 texels, row-major, row 0 at v = 0; 4-byte aligned */
  float base_color_factor[4];   /* RGBA multiplier of the sample */
} vgi_mesh;

typedef struct vgi_light {
  uint32_t type;      /* vgi_light_type, optionally | VGI_LIGHT_SHADOW */
  float vec[3];       /* directional: unit direction the light travels; point: position */
  float color[3];     /* RGB radiant intensity (colour x intensity) */
  float min_distance; /* point: distance clamp for 1/d^2 falloff (world units) */
} vgi_light;

typedef struct vgi_lights {
  uint32_t count;
  vgi_light light[VGI_MAX_LIGHTS];
} vgi_lights;

/* Cone set and march schedule. Diffuse cone 0 points along the normal, cones
 * 1..num_diffuse-1 sit on a ring at diffuse_ring_deg from the normal, evenly in
 * azimuth; weights are proportional to cos(polar angle) and sum to one. */
typedef struct vgi_cone_params {
  uint32_t num_diffuse;        /* 0..VGI_MAX_DIFFUSE_CONES */
  uint32_t specular;           /* 0: no specular cone */
  float diffuse_aperture_deg;  /* full cone angle */
  float diffuse_ring_deg;      /* polar angle of the ring cones */
  float specular_aperture_deg; /* full angle, used when the G-buffer has no roughness */
  float step_factor;           /* step = step_factor * cone diameter */
  float max_distance;          /* world units */
  float alpha_cutoff;          /* stop when accumulated opacity reaches this */
  float normal_offset;         /* cone origin offset along the normal, in voxels */
  float start_offset;          /* first sample distance along the cone, in voxels */
  float eye[3];                /* camera position for the specular reflection */
} vgi_cone_params;

/* G-buffer: world position and normal float32x3 per pixel, row-major width x height;
 * optional per-pixel roughness (tan of the specular half-angle). Device pointers. */
typedef struct vgi_gbuffer {
  const float* position;
  const float* normal;
  const float* roughness; /* may be NULL */
  uint32_t width;
  uint32_t height;
} vgi_gbuffer;

/* Launch-shape knobs (the autotuner's search space). */
/* Cone-march pixel order. TILES: workgroups take screen tiles (coherent G-buffers from a
 * camera). MORTON: the pixels of each G-buffer are first sorted by the Morton code of
 * their position in the grid and workgroups take runs of 256 sorted pixels, so that the
 * cones of a wavefront start close together even when the G-buffer is not spatially
 * coherent (e.g. points sampled on triangles). Results are identical either way. */
enum { VGI_ORDER_TILES = 0, VGI_ORDER_MORTON = 1 };
/* Cone-march texel addressing: AUTO uses 32-bit buffer offsets when every mip level of a
 * launch is under 4 GiB and 64-bit addresses otherwise (grids of 1024^3); WIDE uses 64-bit
 * addresses always. The results agree within FP32 rounding (the fixed-aperture march
 * schedule is computed per lane instead of once per aperture); WIDE exists to exercise and
 * compare the large-grid form on small scenes. */
enum { VGI_ADDRESS_AUTO = 0, VGI_ADDRESS_WIDE = 1 };

typedef struct vgi_tuning {
  uint32_t bin_size;    /* voxelize triangle bin edge: 8, 16 or 32 voxels */
  uint32_t chunk_tris;  /* triangles per voxelize work item: 1..64 */
  uint32_t tile_w;      /* cone-march screen tile: 8x8, 16x16 or 32x8 (TILES order) */
  uint32_t tile_h;
  uint32_t pixel_order; /* VGI_ORDER_* */
  uint32_t address_mode; /* VGI_ADDRESS_* */
} vgi_tuning;

typedef struct vgi_scene_stats {
  uint64_t triangles_in;       /* triangles submitted */
  uint64_t triangles_rejected; /* non-finite, outside the guard band or bad index */
  uint64_t triangles_degenerate; /* zero area after snapping */
  uint64_t bin_pairs;          /* (triangle, bin) pairs after the bin-level overlap test */
  uint64_t work_items;         /* voxelize work units (bin, up to 16 chunks of chunk_tris triangles) */
  uint64_t device_bytes;       /* device memory held by the scene */
  uint32_t n;
  uint32_t levels;
  uint32_t precision;
  uint32_t has_voxels;
  uint32_t has_mips;
  uint32_t bin_size;
  uint32_t voxelize_fallback; /* 1: the last voxelize took the per-triangle path (its pair
                                 list did not fit the size planned for the call) */
  uint32_t pad;
} vgi_scene_stats;

/* Per-phase GPU durations of the last call on a context with profiling enabled,
 * measured with HIP events on the stream the kernels ran on. */
typedef struct vgi_timings {
  float voxelize_setup_ms; /* snap + triangle records + bin counts */
  float voxelize_sort_ms;  /* scan + emit + work list */
  float voxelize_raster_ms;/* the brick voxelize kernel */
  float voxelize_total_ms;
  float mips_ms;           /* inject + all mip levels */
  float trace_ms;          /* the cone-march kernel */
  uint64_t trace_trilinear_samples; /* trilinear gathers executed (8 texels each) */
} vgi_timings;

/* ---- library / context ---------------------------------------------------- */
const char* vgi_version(void);
uint32_t vgi_abi_version(void);
/* Message of the calling thread's last failed call ("" if none). */
const char* vgi_last_error(void);

/* Create a context on a HIP device (replaces CForegroundBootstrapper::Init's device
 * globals, Foreground/ForegroundBootstrapper.cpp:8-17). */
int vgi_create(vgi_context** out, int device);
int vgi_destroy(vgi_context* ctx);
int vgi_set_tuning(vgi_context* ctx, const vgi_tuning* tuning);
int vgi_get_tuning(const vgi_context* ctx, vgi_tuning* tuning);
int vgi_set_profiling(vgi_context* ctx, int enabled);
int vgi_get_timings(const vgi_context* ctx, vgi_timings* out);
int vgi_sync(vgi_context* ctx, void* stream);

/* ---- scenes (resident voxel volume + mip chain) ---------------------------- */
/* Replaces CMegaPipeline::CreateVoxelizePass (MegaPipeline.cpp:470-490): the scene owns
 * the accumulator and the mip levels; memory is sized on the first voxelize. */
int vgi_scene_create(vgi_context* ctx, vgi_scene** out);
int vgi_scene_release(vgi_scene* scene);
/* Reads the statistics of the scene's last voxelize call. When that call's bin scan has
 * not yet run on the device, it first waits for the device's queued work
 * (hipDeviceSynchronize), so it may wait for other streams too. */
int vgi_scene_get_stats(const vgi_scene* scene, vgi_scene_stats* out);
/* Device bytes a scene of this grid and precision holds while resident: accumulator,
 * occupancy mask, bin table and mip levels (admission control). The voxelize scratch
 * (about 112 bytes per triangle, 4 per bin pair and 44 per bin) belongs to the stream
 * the call runs on and is reported by vgi_workspace_bytes. num_triangles is unused. */
uint64_t vgi_scene_bytes_estimate(const vgi_grid* grid, uint32_t precision, uint64_t num_triangles);

/* Clear + voxelize all meshes into the scene's accumulator (one frame's voxelize pass).
 * Replaces MegaPipeline.cpp:196 (clear) + :218-223 (VoxelizeRenderer.RenderList), i.e.
 * VoxelizeRenderer.cpp:19-30,81-127 and the VoxelGS/VoxelPS pipeline
 * (Pipelang/Internal/main.lua:83-144,242-275). Conservative, order-independent.
 * Asynchronous like the reference's recorded voxelize pass (committed with the frame,
 * MegaPipeline.cpp:332): the (triangle, bin) pair list and the work map are sized on the
 * host before the call (4 pairs per triangle + 64 Ki, or 5/4 of the most an earlier
 * call on the same stream needed); a call whose pairs do not fit is detected on the
 * device and voxelized by a per-triangle path instead, with the same result (reported by
 * vgi_scene_stats.voxelize_fallback), and later calls on the stream are sized from it. */
int vgi_voxelize(vgi_context* ctx, vgi_scene* scene, const vgi_mesh* meshes, uint32_t num_meshes,
                 const vgi_grid* grid, void* stream);

/* Voxelize without allocating. No counterpart in the reference: for callers that cannot
 * allocate device memory while they run (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same kernels and the same result as
 * vgi_voxelize for one scene, but every per-call buffer (descriptor table, triangle
 * records, bin counters and scan, pair list, work map) is carved from one static device
 * array of VGI_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from vgi_create
 * until the process ends, and the descriptors reach it through a kernel argument instead
 * of a host copy. No hipMalloc, hipFree, pinned allocation, host copy or memset is
 * issued. Limits: at most VGI_NOALLOC_MAX_MESHES meshes; the call's scratch (about 100
 * bytes per triangle, 16 per (triangle, bin) pair planned as 4 per triangle + 64 Ki, 48
 * per bin) must fit the array (a scene of 1,048,576 triangles on a 512^3 grid with 8^3
 * bins does), else VGI_E_RANGE; the scene's resident volume must already exist for this
 * grid size (one earlier vgi_voxelize with the same n and bin size), else VGI_E_STATE.
 * Every call on one GPU uses that GPU's one array: calls that overlap on a GPU (on two
 * streams, or from two threads, without ordering between them) race and corrupt each
 * other; order them on one stream. */
#define VGI_NOALLOC_SCRATCH_BYTES (144u << 20)
#define VGI_NOALLOC_MAX_MESHES 16u
int vgi_voxelize_noalloc(vgi_context* ctx, vgi_scene* scene, const vgi_mesh* meshes, uint32_t num_meshes,
                         const vgi_grid* grid, void* stream);

/* Inject direct light into level 0 and build the 2x2x2 box mip chain in the given
 * precision. No reference code (planned in docs/index.md:36). */
int vgi_build_mips(vgi_context* ctx, vgi_scene* scene, const vgi_lights* lights,
                   uint32_t precision, void* stream);

/* Cone-trace every G-buffer pixel. out_diffuse: float4 per pixel (irradiance RGB,
 * ambient visibility); out_specular: float4 per pixel (radiance RGB, opacity) or NULL.
 * Replaces the lighting_indirect pass (MegaPipeline.cpp:252-268,
 * Shader/Lighting/indirect.frag:104-222). Device pointers. */
int vgi_trace(vgi_context* ctx, const vgi_scene* scene, const vgi_gbuffer* gbuffer,
              const vgi_cone_params* cones, float* out_diffuse, float* out_specular,
              void* stream);

/* ---- batches: many scenes through one launch per stage ---------------------- */
/* Scenes of uneven triangle count and grid size are voxelized together: one setup
 * launch, one scan over all their bins and one persistent voxelize launch over all
 * their work items (a scene may appear once per batch). No reference counterpart: the
 * reference voxelizes one scene with one draw per primitive (VoxelizeRenderer.cpp:19-30). */
typedef struct vgi_voxelize_item {
  vgi_scene* scene;
  const vgi_mesh* meshes;
  uint32_t num_meshes;
  vgi_grid grid;
} vgi_voxelize_item;

typedef struct vgi_mips_item {
  vgi_scene* scene;
  const vgi_lights* lights; /* may be NULL: no direct light */
} vgi_mips_item;

typedef struct vgi_trace_item {
  const vgi_scene* scene;
  vgi_gbuffer gbuffer;
  vgi_cone_params cones;
  float* out_diffuse;
  float* out_specular; /* may be NULL */
} vgi_trace_item;

int vgi_voxelize_batch(vgi_context* ctx, const vgi_voxelize_item* items, uint32_t count, void* stream);
int vgi_build_mips_batch(vgi_context* ctx, const vgi_mips_item* items, uint32_t count,
                         uint32_t precision, void* stream);
/* Scenes whose mips have different precisions are traced in one launch per precision. */
int vgi_trace_batch(vgi_context* ctx, const vgi_trace_item* items, uint32_t count, void* stream);

/* ---- screen-space passes around the cone march ------------------------------ */
/* The reference's indirect-lighting frame is: reconstruct world position and normal from
 * the depth and normal targets, march the cones, blend with last frame's result
 * (temporal), then a cross-bilateral blur along y and along x (MegaPipeline.cpp:252-284).
 * These entry points are those steps on device buffers, one thread per pixel, in FP32
 * with a fixed operation order (no fused multiply-add contraction), so that results are
 * reproducible bit for bit. Images are row-major width x height, 4 channels per texel,
 * stored as fp32 or fp16 (the reference's render targets are RGBA16F). */
enum { VGI_IMAGE_RGBA32F = 0, VGI_IMAGE_RGBA16F = 1 };

/* The reference's G-buffer inputs: the D32 depth target and the RGBA16 UNORM view-space
 * normal target (rgb = n * 0.5 + 0.5), with the inverse projection and inverse
 * model-view matrices (GLSL mat4, column-major: m[column * 4 + row]). */
typedef struct vgi_gbuffer_ref {
  const float* depth;      /* width x height */
  const uint16_t* normals; /* width x height x 4 */
  float inv_proj[16];
  float inv_model_view[16];
  uint32_t width;
  uint32_t height;
} vgi_gbuffer_ref;

/* World position and normal (float32 x 3 per pixel, the vgi_gbuffer layout) and view
 * depth (-z in view space, float per pixel; may be NULL) of every pixel. Pixel (x, y) is
 * taken at uv = ((x + 0.5) / width, (y + 0.5) / height): view position =
 * InvProj * (uv * 2 - 1, depth, 1) divided by its w; world position = InvModelView *
 * (view position, 1); normal = mat3(InvModelView) * normalize(unorm * 2 - 1), or 0 where
 * the stored normal has zero length (the trace skips such pixels). Replaces
 * Shader/Lighting/indirect.frag:44-58 (getDepth, getCSpos, getNormal) and :189-196. */
int vgi_gbuffer_decode(vgi_context* ctx, const vgi_gbuffer_ref* in, float* out_position,
                       float* out_normal, float* out_view_depth, void* stream);

/* Temporal accumulation: each pixel's world position is reprojected with last frame's
 * projection and model-view; where it lands inside the screen, last frame's result
 * (bilinear, wrapping at the edges as the reference's repeat sampler does) is blended in
 * with weight 0.95 * smoothstep(0, 1, 1 - |history depth - view depth|) and the colour is
 * clamped to [0, 16]. out = (colour, view depth), the next frame's history. Replaces
 * Shader/Lighting/indirect.frag:224-242. Device pointers; out must not alias history. */
typedef struct vgi_temporal_params {
  const float* indirect;   /* float4 per pixel, rgb used (e.g. vgi_trace's out_diffuse) */
  const float* position;   /* world position float32 x 3 per pixel */
  const float* view_depth; /* float per pixel */
  const void* history;     /* last frame's out, history_format */
  void* out;               /* out_format */
  uint32_t history_format;
  uint32_t out_format;
  float prev_projection[16];  /* column-major */
  float prev_model_view[16];  /* column-major */
  uint32_t width;
  uint32_t height;
} vgi_temporal_params;
int vgi_temporal_resolve(vgi_context* ctx, const vgi_temporal_params* params, void* stream);

/* Cross-bilateral blur along one screen axis: the centre texel plus taps at +-2, 4, 6, 8
 * pixels (distance weights for r = 1..4) and +-11, 15 pixels (r = 5, 7), each weighted by
 * exp2(-r^2 / 32 - ((d_centre - d_tap) * 512)^2) on the raw depth, edges clamped;
 * out = (weighted mean rgb, 0). axis 0 blurs along x (the reference's blurY.frag), axis 1
 * along y (blurX.frag); the reference runs axis 1 then axis 0. exp2 is evaluated as
 * 2^k * p(f) with k = rint(x), f = x - k and p the degree-7 Taylor polynomial of 2^f,
 * 0 for x <= -125; the depth square and the weighted sums are fused multiply-adds.
 * Replaces Shader/Lighting/bilateralBlur.inc:15-72. out must not alias color. */
typedef struct vgi_blur_params {
  const void* color;  /* color_format, rgb used */
  const float* depth; /* the D32 depth target */
  void* out;          /* out_format */
  uint32_t color_format;
  uint32_t out_format;
  uint32_t width;
  uint32_t height;
  uint32_t axis;
} vgi_blur_params;
int vgi_bilateral_blur(vgi_context* ctx, const vgi_blur_params* params, void* stream);

/* ---- autotuner -------------------------------------------------------------- */
/* Times one batch (the same scenes in vox, mips and trace, count each) under every bin
 * size {8,16,32} and every pixel schedule (screen tiles 8x8, 16x16, 32x8, or Morton
 * order), keeps the fastest pair in the context's tuning (chunk_tris unchanged) and
 * reports all timings. The sweep is separable: the bin size changes only voxelize + mips
 * and the schedule only the trace (the trace reads mip levels, which are bit-identical
 * for every bin size), so a frame time is build_ms[bin] + trace_ms[schedule]. Each
 * candidate runs once untimed, then `reps` timed runs of which the fastest counts.
 * Synchronous; leaves the scenes built and traced. No reference counterpart (the
 * reference has fixed 128^3 draws). */
typedef struct vgi_autotune_report {
  vgi_tuning best;
  uint32_t bin_sizes[3];
  uint32_t tiles[4][3];  /* tile_w, tile_h, pixel_order of each schedule */
  float build_ms[3];     /* voxelize + mips of the batch, per bin size */
  float trace_ms[4];     /* trace of the batch, per schedule */
  float frame_ms[3][4];  /* [bin][schedule] = build_ms[bin] + trace_ms[schedule] */
} vgi_autotune_report;

int vgi_autotune_batch(vgi_context* ctx, const vgi_voxelize_item* vox, const vgi_mips_item* mips,
                       const vgi_trace_item* trace, uint32_t count, uint32_t precision, uint32_t reps,
                       vgi_autotune_report* report, void* stream);

/* ---- serving: continuous batching over resident scenes ---------------------- */
/* A server owns `slots` HIP streams, each driven by its own host thread, and a scheduler
 * thread. Submitted requests wait in arrival order; whenever a slot is idle the scheduler
 * fills it with every waiting request it can admit now (up to the batch limits): one
 * vgi_voxelize_batch + vgi_build_mips_batch for the scenes that must be (re)built and one
 * vgi_trace_batch for all their frames. Slots run concurrently, so requests that arrive
 * while others are on the GPU overlap with them. Built scenes stay resident, keyed by
 * scene_key, until evicted least-recently-used to admit another within memory_budget
 * (bytes per vgi_scene_bytes_estimate); an evicted scene of the same grid shape is
 * recycled without freeing. A scene is rebuilt when its request sets revoxelize, when its
 * grid changed, or after eviction; lights take effect when a scene is built. Requests
 * whose scene is being built or rebuilt by an in-flight batch wait for it. A result does
 * not depend on what else is in its batch or in flight. No reference counterpart (the
 * reference renders one scene per frame, Game.cpp:248-345). */
typedef struct vgi_server vgi_server;

typedef struct vgi_server_config {
  uint64_t memory_budget;      /* bytes of resident scenes */
  uint32_t precision;          /* mip storage of every scene (VGI_PREC_*) */
  uint32_t slots;              /* concurrent batches, 1..16 */
  uint32_t max_batch_requests; /* requests per batch, >= 1 */
  uint64_t max_batch_pixels;   /* stop filling a batch past this many pixels (0: no limit) */
} vgi_server_config;

typedef struct vgi_frame {
  vgi_gbuffer gbuffer;
  vgi_cone_params cones;
  float* out_diffuse;
  float* out_specular; /* may be NULL */
} vgi_frame;

typedef struct vgi_request {
  uint64_t scene_key;
  const vgi_mesh* meshes; /* device buffers stay valid until the request completes */
  uint32_t num_meshes;
  uint32_t revoxelize;    /* rebuild even if resident (dynamic geometry) */
  vgi_grid grid;
  vgi_lights lights;
  const vgi_frame* frames;
  uint32_t num_frames;
  uint32_t pad;
  double arrival_s;       /* vgi_now() time the request arrived (0: at submit) */
  uint64_t tag;           /* returned in its completion */
} vgi_request;

typedef struct vgi_completion {
  uint64_t tag;
  int32_t status;         /* VGI_OK, or the error that failed the request */
  uint32_t voxelized;     /* 1 if this request built its scene */
  double arrival_s, start_s, done_s; /* vgi_now() times */
} vgi_completion;

typedef struct vgi_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t builds, evictions, batches;
  uint64_t frames, pixels;
  uint64_t resident_scenes, resident_bytes, peak_resident_bytes;
  uint32_t max_concurrent_batches;
  uint32_t max_batch_requests_seen;
} vgi_server_stats;

double vgi_now(void); /* monotonic seconds */
int vgi_server_create(vgi_context* ctx, const vgi_server_config* cfg, vgi_server** out);
/* Copies the request (mesh and frame tables); returns at once. */
int vgi_server_submit(vgi_server* server, const vgi_request* req);
/* Moves up to max completions into out, waiting up to wait_us for the first one. */
int vgi_server_poll(vgi_server* server, vgi_completion* out, uint32_t max, uint32_t* count, uint32_t wait_us);
/* Blocks until every submitted request has completed (completions stay to be polled). */
int vgi_server_drain(vgi_server* server);
int vgi_server_get_stats(vgi_server* server, vgi_server_stats* out);
/* Drains, stops the threads and releases the resident scenes and streams. */
int vgi_server_destroy(vgi_server* server);

/* ---- glTF 2.0 ingestion (host side) ------------------------------------------ */
/* Reads a .gltf (external or base64 data-URI buffers) or .glb file: the default scene's
 * node tree (matrix or TRS, composed parent * local in double, rounded to float), every
 * triangle primitive's POSITION (float VEC3, any byteStride) and indices (u8/u16/u32 or
 * none), with strips and fans assembled into lists as Vulkan assembles them (strip
 * triangle i = (i, i+1+i%2, i+2-i%2), fan = (i+1, i+2, 0)); points and lines are
 * skipped. A mesh placed by several nodes is stored once. Replaces the mesh and node part
 * of CglTFSceneImporter::ImportFile (Foreground/SceneGraph/glTFSceneImporter.cpp:210-326
 * GetMesh, :333-372 Visit, :387-425 ImportFile).
 * Albedo: VGI_GLTF_ALBEDO_REFERENCE reproduces the reference importer, whose material
 * lookup of "baseColor" (:180) never matches glTF's "baseColorFactor", so every material
 * is white, and whose voxelizer draws only primitives that have a material
 * (VoxelizeRenderer.cpp:40,103); VGI_GLTF_ALBEDO_FACTOR uses baseColorFactor (glTF's
 * default material, white, for primitives without one). Colours are rounded to RGB8. */
typedef struct vgi_gltf vgi_gltf;
enum { VGI_GLTF_ALBEDO_REFERENCE = 0, VGI_GLTF_ALBEDO_FACTOR = 1 };

typedef struct vgi_gltf_info {
  uint32_t num_primitives;     /* placed primitives (node x mesh primitive) */
  uint32_t num_geometries;     /* distinct mesh primitives */
  uint32_t skipped_primitives; /* points / lines */
  uint32_t num_images;         /* the file's images array */
  uint64_t num_triangles;      /* over placed primitives */
} vgi_gltf_info;

typedef struct vgi_gltf_primitive {
  const float* positions;   /* host memory owned by the handle: xyz per vertex, object space */
  const uint32_t* indices;  /* host memory owned by the handle: 3 per triangle */
  uint32_t num_vertices;
  uint32_t num_triangles;
  float model[12];          /* row-major 3x4 world transform of the node */
  uint8_t base_color[4];    /* RGBA8 */
  int32_t node;
  int32_t mesh;
  int32_t primitive;
  int32_t material;         /* -1: none */
  const float* texcoords;   /* host memory owned by the handle: uv per vertex (TEXCOORD_0), or NULL */
  int32_t image;            /* image of the material's baseColorTexture, -1: none; the reference
                               mode indexes images with the texture's index, as the reference's
                               importer does (glTFSceneImporter.cpp:110-131,201), the factor
                               mode with the texture's source */
  uint32_t reserved;
  float base_color_factor[4]; /* multiplier of that texture: white (reference mode) or baseColorFactor */
} vgi_gltf_primitive;

/* An image as the file stores it (PNG, JPEG, ...): decoding is the caller's, as the
 * reference decodes with stb_image inside tinygltf. */
typedef struct vgi_gltf_image {
  const uint8_t* data;   /* encoded bytes, owned by the handle */
  size_t bytes;
  const char* mime_type; /* "" when the file does not say */
  const char* uri;       /* relative file name, "" when embedded */
} vgi_gltf_image;

int vgi_gltf_load_file(const char* path, uint32_t albedo_mode, vgi_gltf** out);
/* base_dir: directory for relative buffer URIs (NULL: only embedded buffers). */
int vgi_gltf_load_memory(const void* data, size_t bytes, const char* base_dir, uint32_t albedo_mode,
                         vgi_gltf** out);
int vgi_gltf_get_info(const vgi_gltf* g, vgi_gltf_info* out);
int vgi_gltf_get_primitive(const vgi_gltf* g, uint32_t index, vgi_gltf_primitive* out);
/* The image's encoded bytes (an external file is read on the first request). With
 * vgi_gltf_set_image_rgba8 this replaces GetImage (glTFSceneImporter.cpp:110-131), whose
 * decoded RGBA8 image the importer uploads as the material's albedo image. */
int vgi_gltf_get_image(vgi_gltf* g, uint32_t index, vgi_gltf_image* out);
/* The image's decoded texels, RGBA8 row-major width x height (copied; 1..16384 per side).
 * A primitive whose material's baseColorTexture is this image and whose geometry has
 * TEXCOORD_0 is then uploaded as a textured mesh (texture x base_color_factor, with the
 * alpha discard); primitives whose image has no texels keep their base colour. */
int vgi_gltf_set_image_rgba8(vgi_gltf* g, uint32_t index, const uint8_t* rgba8, uint32_t width, uint32_t height);
/* Copies every geometry (and texcoords and decoded images) to device memory of ctx's
 * device (kept by the handle until vgi_gltf_release; a second call with the same ctx
 * reuses it, uploading images set since) and writes one vgi_mesh per placed primitive
 * (u32 indices, the node's model matrix, base colour or texture) into meshes[0..capacity);
 * *count receives the number of primitives. Synchronous. */
int vgi_gltf_upload(vgi_context* ctx, vgi_gltf* g, vgi_mesh* meshes, uint32_t capacity, uint32_t* count);
int vgi_gltf_release(vgi_gltf* g);

/* ---- parity readback (synchronous) ----------------------------------------- */
/* Accumulator as n^3 voxels x 8 uint32 (x fastest): R sum, G sum, B sum, count,
 * sum(nx+1024), sum(ny+1024), sum(nz+1024), 0 (normals quantized to 1/1024). */
int vgi_scene_read_voxels(vgi_context* ctx, const vgi_scene* scene, uint32_t* host_out);
/* Mip level as (n>>level)^3 texels x 4 float (x fastest), decoded from storage. */
int vgi_scene_read_level(vgi_context* ctx, const vgi_scene* scene, uint32_t level, float* host_out);

/* ---- device memory / stream helpers ---------------------------------------- */
int vgi_malloc(vgi_context* ctx, void** out, size_t bytes);
int vgi_free(vgi_context* ctx, void* ptr);
int vgi_memcpy_h2d(vgi_context* ctx, void* dst, const void* src, size_t bytes);
int vgi_memcpy_d2h(vgi_context* ctx, void* dst, const void* src, size_t bytes);
int vgi_memset(vgi_context* ctx, void* dst, int value, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16 B per lane, grid-stride), async on
 * the stream; bytes must be a multiple of 16. Also the known-size reference kernel for
 * calibrating HBM traffic counters. */
int vgi_memcpy_d2d(vgi_context* ctx, void* dst, const void* src, size_t bytes, void* stream);
int vgi_stream_create(vgi_context* ctx, void** out);
/* Synchronizes and destroys the stream, and frees the scratch the context held for it. */
int vgi_stream_destroy(vgi_context* ctx, void* stream);
/* Makes the work enqueued on `waiting` from now on wait for the work enqueued on
 * `signalling` so far (an event recorded on one, waited on by the other; the host does not
 * block). NULL selects the context's stream. No counterpart in the reference, which runs
 * one Vulkan queue: it lets a caller overlap stages of different frames on two streams,
 * e.g. voxelize frame k+1 while frame k is traced (a scene's voxelize must not overlap its
 * own mip build, which reads the accumulator, and its mip build must not overlap its
 * trace, which reads the levels). */
int vgi_stream_wait(vgi_context* ctx, void* waiting, void* signalling);
/* VGI_OK when every call enqueued on the stream has completed, VGI_E_NOT_READY while work
 * is pending (hipStreamQuery; does not block). */
int vgi_stream_query(vgi_context* ctx, void* stream);
/* Device bytes of voxelize scratch and descriptor staging the context holds, summed over
 * the streams it has run on (each stream keeps its own, grown to its largest batch). */
int vgi_workspace_bytes(vgi_context* ctx, uint64_t* out);
/* ---- calibration kernels (ceilings for the counter-based rooflines) ---------- */
/* No reference counterpart. One launch of `blocks` workgroups of 256 lanes on the stream:
 * VGI_PROBE_FMA: each lane runs `iters` rounds of 8 independent FP32 fused multiply-adds
 *   and stores one float (buf holds blocks * 256 floats): the VALU issue ceiling;
 * VGI_PROBE_ATOMIC_RETURN / _NORETURN: each lane adds 1 `iters` times to the first word of
 *   its own 64-byte line (buf holds blocks * 256 * 64 bytes), with returning atomics whose
 *   results are summed into the line's second word, or non-returning ones: the atomic
 *   rate without contention;
 * VGI_PROBE_INT: as VGI_PROBE_FMA with 8 independent 32-bit integer (exclusive-or)
 *   chains and one store per lane: the integer VALU issue rate. */
enum { VGI_PROBE_FMA = 0, VGI_PROBE_ATOMIC_RETURN = 1, VGI_PROBE_ATOMIC_NORETURN = 2, VGI_PROBE_INT = 3 };
int vgi_probe(vgi_context* ctx, uint32_t kind, void* buf, size_t buf_bytes, uint32_t blocks, uint32_t iters,
              void* stream);

/* Device-side elapsed time between two points on a stream (HIP events). */
int vgi_timer_start(vgi_context* ctx, void* stream);
int vgi_timer_stop(vgi_context* ctx, void* stream, float* out_ms);

#ifdef __cplusplus
}
#endif

#endif /* VGI_H */
