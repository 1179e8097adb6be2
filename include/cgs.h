/*
 * cgs.h — C ABI of the MI355X (gfx950) ray–bicubic-patch library (libcgs.so).
 *
 * This is the drop-in boundary for cgspline's ray–patch path.  Every entry
 * point lists the reference interface it replaces (file:line in
 * MachSilva/cgspline).  Plain pointers and sizes only: "device" pointers are
 * HIP device allocations (hipMalloc, or any allocator's device memory), host
 * pointers are ordinary memory.  Streams are hipStream_t passed as void*.
 *
 * Errors: every call returns a cgs_status; nothing aborts.  The C++ facade
 * (include/cgspline/) maps the codes back to the reference's behaviour
 * (throws, int returns, warnings).
 */
#ifndef CGS_H
#define CGS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGS_ABI_VERSION 2

typedef enum cgs_status
{
    CGS_OK = 0,
    CGS_ERR_INVALID = 1,        /* bad argument (null pointer, size 0 where not allowed, ...) */
    CGS_ERR_HIP = 2,            /* a HIP runtime call failed; see cgs_last_error() */
    CGS_ERR_HASH_FAILED = 3,    /* perfect hash: all (A,B) attempts failed (reference: runtime_error) */
    CGS_ERR_HASH_DUPLICATE = 4, /* perfect hash: final function not injective (reference: build() == 1) */
    CGS_ERR_BUCKET_OVERFLOW = 5,/* perfect hash: a bucket holds more than 4 keys (reference: __trap) */
    CGS_ERR_BUSY = 6,           /* render already in flight on this context (reference: warning, no-op) */
    CGS_ERR_EMPTY_SPLIT = 7,    /* BVH build over an empty range (reference: logic_error("invalid")) */
    CGS_ERR_STACK_OVERFLOW = 8, /* a Bézier-clipping region stack exceeded CGS_CLIP_STACK entries */
    CGS_ERR_NO_DEVICE = 9,      /* no HIP device */
    CGS_ERR_OUT_OF_MEMORY = 10
} cgs_status;

/* Region-stack capacity of the GPU Bézier clipper (the reference's device
 * variant traps at 32 entries, Spline_BezierClipping.cu:736-737). */
#define CGS_CLIP_STACK 32

/* Hit patch id reported for a ray whose clipping stack overflowed. */
#define CGS_PATCH_OVERFLOW 0xFFFFFFFEu
#define CGS_EMPTY 0xFFFFFFFFu

/* ---------------------------------------------------------------------- */
/* Plain-old-data records                                                  */
/* ---------------------------------------------------------------------- */

/* 32-byte ray: origin, max range, direction (rt/Intersection.h:16-24 Ray). */
typedef struct cgs_ray
{
    float ox, oy, oz, tmax;
    float dx, dy, dz, pad;
} cgs_ray;

/* 16-byte closest hit: t, patch (u,v), patch id (CGS_EMPTY on a miss)
 * (rt/Intersection.h:28-34 Intersection: t, coordinates, index). */
typedef struct cgs_hit
{
    float t, u, v;
    uint32_t patch;
} cgs_hit;

/* 64-byte BVH node, the reference's rt::BVH::Node layout (rt/BVH.h:23-39):
 * child boxes, then the left/right node offsets; a leaf has left = CGS_EMPTY
 * and right = first element slot in the index array. */
typedef struct cgs_bvh_node
{
    float lmin[3], lmax[3];
    float rmin[3], rmax[3];
    uint32_t left, right;
    uint32_t pad[2];
} cgs_bvh_node;

/* BVH element (rt/BVH.h:41-46 ElementData). */
typedef struct cgs_bvh_element
{
    float bmin[3], bmax[3];
    float centroid[3];
    uint32_t index;
} cgs_bvh_element;

/* Perfect hash function parameters (rt/PerfectHashFunction.h:30-70):
 * h(k) = (A·k + B + D[k & mask]) mod table_size in 32-bit arithmetic. */
typedef struct cgs_phf
{
    uint32_t A, B, mask, table_size;
    uint32_t attempts;      /* (A,B) pairs tried */
    uint32_t displacement_size;
} cgs_phf;

/* Scene description (host memory), the content of rt::Scene
 * (rt/Scene.h:97-159) as MainWindow::convertScene fills it
 * (MainWindow.cpp:363-492).  Vertices of all surfaces are concatenated;
 * patch indices are relative to the owning surface's first vertex;
 * point(i,j) of a patch is its index [4j+i] (Spline.h:137-141). */
typedef struct cgs_scene_desc
{
    const float* vertices;        /* nv × 4 (x, y, z, w) */
    const uint32_t* indices;      /* npatch × 16 */
    const uint32_t* surf_vbase;   /* nsurf + 1, prefix over vertices */
    const uint32_t* surf_pbase;   /* nsurf + 1, prefix over patches */
    uint32_t nsurf;
    const uint32_t* obj_surface;  /* nobj: surface of each object; CGS_EMPTY: an object without
                                   * geometry on this path (the reference's Mesh and Sphere, whose
                                   * bounds() are empty, Primitive.cu:15-23,47-56): it takes part in
                                   * the top-level BVH with an empty box and is never hit */
    const float* obj_local2world; /* nobj × 16, column-major */
    const float* obj_world2local; /* nobj × 16, column-major */
    const float* obj_material;    /* nobj × 8: diffuse rgb, specular rgb, metalness, roughness */
    uint32_t nobj;
    const float* lights;          /* nlight × 12: color rgb, strength, position xyz, range, direction xyz, angle */
    uint32_t nlight;
    float background[3];
    float ambient[3];
    uint32_t hash_seed;           /* PerfectHashFunction seed (reference default 42) */
} cgs_scene_desc;

/* Pinhole camera (rt/Scene.h:29-45 Camera; RayTracer.cu:59-71,141-163). */
typedef struct cgs_camera
{
    float to_world[16];   /* camera→world, column-major (Transform::matrix()) */
    float position[3];
    float fovy;           /* radians */
    int32_t flip_y;
} cgs_camera;

/* RayTracer::Options (rt/RayTracer.h:26-34), non-Monte-Carlo path. */
typedef struct cgs_render_options
{
    float eps;             /* secondary-ray offset, 1e-4 */
    int32_t recursion_depth;
    uint32_t frame_stride; /* pixels from one frame row to the next (rt::Frame::stride,
                            * Frame.h:29-94, a row padded to an alignment); 0: the width */
} cgs_render_options;

/* Leaf method of a ray query (cgs_trace_closest_ex / cgs_trace_any_ex).
 * CGS_METHOD_CLIPPING: Bézier clipping, the reference's leaf call
 * (Primitive.cu:234,260); CGS_METHOD_SUBDIVISION: subdivision, its
 * commented-out alternative (Primitive.cu:236,263; Spline_Subdivision.cu:168-288). */
#define CGS_METHOD_CLIPPING 0u
#define CGS_METHOD_SUBDIVISION 1u

/* Options of a ray query.  A zeroed record, or a null pointer, is the
 * default: clipping at the reference's tolerance 2^-12 (Spline.h:493), no
 * cap, no polish; its results are the reference algorithm's bit for bit.
 *   clip_tol: clipping tolerance (0: 2^-12);
 *   subdiv_depth: subdivision halvings per parameter side, 1..16 (0: the
 *     reference's threshold 2^-4, which is depth 5: cells of 1/32); depth D
 *     is the threshold 2^(1-D);
 *   max_clip_iters: stop each patch test after this many clipping-loop
 *     passes, keeping what it found (0: no cap; an opt-in trade of accuracy
 *     for a bounded cost per patch, no counterpart in the reference);
 *   polish: refine the reported closest hit by FP64 Newton steps on
 *     S(u,v) = o_l + t M d in the hit object's frame, after the query (no
 *     counterpart in the reference; hits, patches and objects are the
 *     unpolished query's, t/u/v come within ~1e-7 of the true root where
 *     the clipper's 2^-12 stopping rule leaves up to ~1e-3 in (u,v)); a hit
 *     whose steps do not converge inside the patch keeps its clipped values. */
typedef struct cgs_trace_options
{
    uint32_t method;
    float clip_tol;
    uint32_t subdiv_depth;
    uint32_t max_clip_iters;
    uint32_t polish;
    uint32_t reserved[3];
} cgs_trace_options;

/* Accuracy limits of cgs_batch_autotune_accuracy, against the reference's
 * clipping (polished when allow_polish is set, so the reference is then the
 * FP64-refined root): a ray is outside the limits when its hit or miss,
 * object or occlusion differ, when its patch differs with t beyond max_rel_t
 * relative (a seam hit at the same distance agrees), or when on the same
 * patch t differs beyond max_rel_t relative or u or v beyond max_uv.  A
 * candidate is admissible when at most the fraction max_disagree of the
 * rays is outside (0: every ray within the limits). */
typedef struct cgs_accuracy_limits
{
    float max_rel_t;
    float max_uv;
    float max_disagree;
    uint32_t allow_polish;
} cgs_accuracy_limits;

/* One candidate of cgs_batch_autotune_accuracy: its options, the batched
 * launch's time (ms), the largest relative t and (u,v) differences over rays
 * hit on the same patch as the reference, and the fraction of rays outside
 * the limits. */
typedef struct cgs_autotune_entry
{
    cgs_trace_options options;
    float ms;
    float max_rel_t;
    float max_uv;
    float disagree;
    uint32_t admissible;
} cgs_autotune_entry;

/* Per-launch ray counters written by cgs_render (device memory, 4 × u64):
 * primary, shadow, secondary, overflowed clip stacks. */
typedef struct cgs_ray_counts
{
    uint64_t primary, shadow, secondary, overflow;
} cgs_ray_counts;

typedef struct cgs_ctx cgs_ctx;
typedef struct cgs_scene cgs_scene;
typedef struct cgs_batch cgs_batch;
typedef struct cgs_server cgs_server;

/* One job of a batched query: a ray range against one resident scene
 * (device pointers).  any = 0: closest hits into hits (and objects, nullable);
 * any = 1: occlusion flags into occluded (1 / 0, 2 on a clip-stack overflow). */
typedef struct cgs_batch_job
{
    cgs_scene* scene;
    const cgs_ray* rays;
    uint32_t n;
    int32_t any;
    cgs_hit* hits;
    int32_t* objects;
    uint32_t* occluded;
} cgs_batch_job;

/* Launch shape of batched queries: rays taken per wave fetch (a multiple of
 * 64) and resident workgroups per CU (0 = as many as fit).  Neither changes
 * any result. */
typedef struct cgs_batch_config
{
    uint32_t chunk;
    uint32_t blocks_per_cu;
} cgs_batch_config;

/* ---------------------------------------------------------------------- */
/* Context                                                                 */
/* ---------------------------------------------------------------------- */

int cgs_abi_version(void);
const char* cgs_status_string(int status);
/* Thread-local description of the last error on this thread. */
const char* cgs_last_error(void);

/* Opens device `device` (RayTracer::Options::device, rt/RayTracer.h:32). */
cgs_status cgs_open(int device, cgs_ctx** out);
cgs_status cgs_close(cgs_ctx* ctx);

/* ---------------------------------------------------------------------- */
/* Spline intersection entry points (Spline.h:488-508)                     */
/* ---------------------------------------------------------------------- */

/* spline::doBezierClipping (Spline.h:488-493, Spline_BezierClipping.cu:46-114)
 * batched: ray r against patch r.  rays: n cgs_ray (device), directions are
 * normalized as Ray3f does; patches: n × 16 float4 control points (device);
 * hits: n cgs_hit (device), t starts at ray.tmax, patch = 1 when found else
 * CGS_EMPTY (CGS_PATCH_OVERFLOW when the region stack overflowed).  Host
 * (exact-region) semantics of the reference's CPU path. */
cgs_status cgs_bezier_clip(cgs_ctx* ctx, const cgs_ray* rays, const float* patches, uint32_t n,
                           float tol, cgs_hit* hits, void* stream);

/* spline::doSubdivision (Spline.h:495-500, Spline_Subdivision.cu:168-288),
 * batched like cgs_bezier_clip; `threshold` is the parametric size below
 * which a region is reported (reference default 2^-4). */
cgs_status cgs_subdivision(cgs_ctx* ctx, const cgs_ray* rays, const float* patches, uint32_t n,
                           float threshold, cgs_hit* hits, void* stream);

/* Patch evaluation, batched: patch r (16 float4 control points, device) at
 * (u, v) = uv[2r], uv[2r+1] (device) into out[4r..4r+3] (device).  The kinds
 * are the reference's header functions, FP32, same operation order:
 *   CGS_EVAL_INTERPOLATE       spline::interpolate, de Casteljau (Spline.h:197-214)
 *   CGS_EVAL_INTERPOLATE_MAT   spline::mat::interpolate (SplineMat.h:125-140)
 *   CGS_EVAL_DERIVATIVE_U/V    spline::derivativeU/V (Spline.h:216-254): the
 *                              hodograph without its factor 3
 *   CGS_EVAL_DERIVATIVE_U/V_MAT  spline::mat::derivativeU/V (SplineMat.h:142-176)
 *   CGS_EVAL_NORMAL            spline::normal with its degenerate-edge step of
 *                              1e-4 (Spline.cu:10-27); w = 0
 *   CGS_EVAL_NORMAL_MAT        spline::mat::normal (SplineMat.h:178-187); w = 0 */
#define CGS_EVAL_INTERPOLATE 0u
#define CGS_EVAL_INTERPOLATE_MAT 1u
#define CGS_EVAL_DERIVATIVE_U 2u
#define CGS_EVAL_DERIVATIVE_V 3u
#define CGS_EVAL_DERIVATIVE_U_MAT 4u
#define CGS_EVAL_DERIVATIVE_V_MAT 5u
#define CGS_EVAL_NORMAL 6u
#define CGS_EVAL_NORMAL_MAT 7u
cgs_status cgs_patch_eval(cgs_ctx* ctx, uint32_t kind, const float* patches, const float* uv, uint32_t n, float* out,
                          void* stream);

/* spline::subpatch (Spline.h:384-396), batched: patch r (device) cut to
 * [umin, umax] × [vmin, vmax] = ranges[4r..4r+3] (device), subpatchV then
 * subpatchU, into out (n × 16 float4, device). */
cgs_status cgs_patch_subpatch(cgs_ctx* ctx, const float* patches, const float* ranges, uint32_t n, float* out,
                              void* stream);

/* cg::PatchBVH (PatchBVH.h:13-28, PatchBVH.cpp:7-121) over one surface of
 * bicubic Bézier patches (host points nv × 4, indices np × 16, point(i,j) =
 * [4j+i]): one patch per leaf, boxed by its control points
 * (PatchBVH.cpp:66-75), resident on ctx's device.  The tree is the
 * reference's rt::BVH (median split) with its perfect hash; the reference's
 * BVHBase comes from its absent Ds library, and only the hits, not the tree,
 * are its results. */
typedef struct cgs_patchbvh cgs_patchbvh;
cgs_status cgs_patchbvh_create(cgs_ctx* ctx, const float* vertices, uint32_t nv, const uint32_t* indices, uint32_t np,
                               cgs_patchbvh** out);
/* cg::PatchBVH over a patch table (PatchBVH(const Surface*) of a `.se`
 * surface): the tree over the given element boxes (cgs_patch_table_elements)
 * and the leaf test on the 16 control points of each patch
 * (cgs_patch_table_patches, np × 16 indices into the nv points). */
cgs_status cgs_patchbvh_create_elements(cgs_ctx* ctx, const float* vertices, uint32_t nv, const uint32_t* indices,
                                        uint32_t np, const cgs_bvh_element* elements, cgs_patchbvh** out);
cgs_status cgs_patchbvh_destroy(cgs_patchbvh* bvh);
/* BVHBase::intersect(ray, hit) with PatchBVH::intersectLeaf (PatchBVH.cpp:104-121):
 * the closest patch hit of each ray (a Ray3f: its direction normalized once),
 * hit.distance starting at ray.tmax.  hits (device): t, (u, v), patch index
 * (CGS_EMPTY: a miss; CGS_PATCH_OVERFLOW: a region-stack overflow). */
cgs_status cgs_patchbvh_intersect(cgs_ctx* ctx, cgs_patchbvh* bvh, const cgs_ray* rays, uint32_t n, cgs_hit* hits,
                                  void* stream);
/* BVHBase::intersect(ray) with PatchBVH::intersectLeaf (PatchBVH.cpp:84-102):
 * 1 when a patch whose box the traversal reaches within ray.tmax has a hit in
 * front of the origin (at any distance: the leaf clips from distance inf),
 * else 0; 2 on a region-stack overflow.  occluded: n u32 (device). */
cgs_status cgs_patchbvh_intersect_any(cgs_ctx* ctx, cgs_patchbvh* bvh, const cgs_ray* rays, uint32_t n,
                                      uint32_t* occluded, void* stream);

/* ---------------------------------------------------------------------- */
/* BVH and perfect hash                                                    */
/* ---------------------------------------------------------------------- */

/* rt::BVH::build (rt/BVH.h:59, rt/BVH.cpp:9-108) on the host.  The library
 * allocates *nodes (node_count), *keys (node_count) and *indices
 * (index_count) with malloc; release them with cgs_free_host. */
cgs_status cgs_bvh_build(const cgs_bvh_element* elements, uint32_t n, uint32_t elements_per_node,
                         cgs_bvh_node** nodes, uint32_t** keys, uint32_t* node_count,
                         uint32_t** indices, uint32_t* index_count);
/* BezierSurface::buildBVH element data (rt/Primitive.cu:199-220): bounds and
 * centroid of the projected control points of each patch.  vertices: nv × 4
 * host floats; indices: npatch × 16.  elements: npatch entries (caller). */
cgs_status cgs_patch_elements(const float* vertices, const uint32_t* indices, uint32_t npatch,
                              cgs_bvh_element* elements);
void cgs_free_host(void* p);
/* Mesh::buildBVH element data (rt/Primitive.cu:90-108): per triangle the box
 * of its vertices and the centroid (1/3)((p0+p1)+p2).  vertices: nvertices
 * × 3 host floats; indices: ntriangles × 3.  Build the mesh BVH from them with
 * cgs_bvh_build(..., elements_per_node = 16), as the reference does. */
cgs_status cgs_mesh_elements(const float* vertices, uint32_t nvertices, const uint32_t* indices, uint32_t ntriangles,
                             cgs_bvh_element* elements);

/* Surface::load_be (Surface.cpp:149-213): parses the text of a `.be` model
 * (the patch count; 16 one-based control-point indices per patch, u
 * fastest; the point count; x, y, z per point; one separator character after
 * each number).  The library allocates *vertices (npoints × 4 floats, w = 1)
 * and *indices (npatches × 16, zero-based) with malloc; release them with
 * cgs_free_host.  Malformed text fails with CGS_ERR_INVALID and the
 * reference's messages; an index equal to the point count is rejected (the
 * reference lets it through). */
cgs_status cgs_load_be(const char* text, size_t length, float** vertices, uint32_t* npoints,
                       uint32_t** indices, uint32_t* npatches);

/* A surface's patch table (PatchTable, Surface.h:45-95) as Surface::load_se
 * fills it (Surface.cpp:214-348, PatchTable::set :36-148): points (x, y, z,
 * w), node indices and basis matrices of every element, grouped by type,
 * node count and matrix presence in order of first appearance; patches[i]
 * is the i-th patch in that order (PatchTable::forEachPatch / find). */
#define CGS_PATCH_BEZIER 0u
#define CGS_PATCH_BSPLINE 1u
#define CGS_PATCH_GREGORY 2u
typedef struct cgs_patch_ref
{
    uint32_t offset;         /* first node index in `indices` */
    int32_t matrix_offset;   /* first matrix entry, or -1: no matrix */
    uint32_t size;           /* node count */
    uint32_t type;           /* CGS_PATCH_* */
} cgs_patch_ref;
typedef struct cgs_patch_table
{
    float* points;           /* npoints × 4 */
    uint32_t npoints;
    uint32_t* indices;
    uint32_t nindices;
    float* matrices;         /* per patch with one: (16 or 20) × size, column-major */
    uint32_t nmatrices;
    cgs_patch_ref* patches;
    uint32_t npatches;
} cgs_patch_table;
/* Surface::load_se (Surface.cpp:214-348): parses a `.se` text (nodes with
 * ids and (x, y, z, w); elements of type 1 = B-spline, 2 = Gregory, degree
 * 3, with node indices, a discarded value per node, and a basis matrix of 16
 * (boundary B-spline) or 20 (Gregory) rows, transposed on reading; '#'
 * comments).  Fails with CGS_ERR_INVALID and the reference's messages; an
 * index equal to the node count is rejected (the reference lets it through),
 * and a truncated text fails where the reference would read undefined
 * values.  Release with cgs_patch_table_free. */
cgs_status cgs_load_se(const char* text, size_t length, cgs_patch_table** out);
void cgs_patch_table_free(cgs_patch_table* table);
/* The element boxes cg::PatchBVH gives a table's patches (PatchBVH.cpp:40-80):
 * a patch with a basis matrix is boxed by its 16 or 20 basis-transformed
 * points, projected by w; one without by its first 16 points.  Centroid:
 * the box centre.  A patch whose leaf test would read past the index table
 * (the reference reads undefined memory there) gets the empty box: no ray
 * reaches it. */
cgs_status cgs_patch_table_elements(const cgs_patch_table* table, cgs_bvh_element* elements);
/* The 16 control-point indices PatchBVH::intersectLeaf clips for each patch
 * (PatchBVH.cpp:84-121): the 16 from its offset, whatever its node count or
 * type (a patch of fewer nodes reads on into the next one's); 16 zeros for a
 * patch whose 16 would run past the index table (its box is empty). */
cgs_status cgs_patch_table_patches(const cgs_patch_table* table, uint32_t* indices16);

/* A scene read from the reference's scene description language (`.dl`;
 * reader/DescriptionLanguage.cpp, reader/SceneReader.cpp), flattened the way
 * MainWindow::convertScene fills rt::Scene (MainWindow.cpp:363-492): one
 * object per surface-carrying scene object, depth first, with its world
 * transform and material, one light per light-carrying object.  Surfaces
 * are read with cgs_load_be; objects naming the same Surface asset share it.
 * Triangle meshes, textures and `.se` surfaces are refused (CGS_ERR_INVALID
 * with the reason).  Euler angles are applied yaw (y), pitch (x), roll (z). */
typedef struct cgs_dl_scene
{
    cgs_scene_desc scene;            /* ready for cgs_scene_create; arrays owned by this record */
    float view_to_world[16];         /* the file's view: camera→world, column-major */
    float view_position[3];
    const char* const* object_names; /* scene.nobj names, in object order */
    const char* const* light_names;  /* scene.nlight names */
    void* storage;                   /* owned by the library */
} cgs_dl_scene;

/* Reads the `.dl` file at `path`; model paths in it are relative to its
 * directory.  Release the result with cgs_dl_free.  Errors carry
 * "file:line:column: message". */
cgs_status cgs_dl_read(const char* path, cgs_dl_scene** out);
/* The same from text; `name` labels error messages, `working_dir` (NULL:
 * the current directory) resolves model paths. */
cgs_status cgs_dl_parse(const char* text, size_t length, const char* name, const char* working_dir,
                        cgs_dl_scene** out);
void cgs_dl_free(cgs_dl_scene* scene);

/* PerfectHashFunction::build (rt/PerfectHashFunction.h:53, .cu:229-503) on
 * the GPU, deterministic.  keys: n device u32; displacement: device u32
 * buffer of cgs_phf_displacement_size(n) entries (written); out: parameters.
 * Returns CGS_ERR_HASH_FAILED / _DUPLICATE / _BUCKET_OVERFLOW like the
 * reference's non-zero return. */
uint32_t cgs_phf_displacement_size(uint32_t n);
cgs_status cgs_phf_build(cgs_ctx* ctx, const uint32_t* keys, uint32_t n, uint32_t seed,
                         uint32_t* displacement, cgs_phf* out, void* stream);
/* No counterpart in the reference: cgs_phf_build for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * cgs_phf_build; it allocates nothing.  Its working memory is one static
 * device array per GPU (17.3 MiB, held until the process ends), so n is at
 * most 524,288 keys (CGS_ERR_INVALID above that).  Every call on a GPU
 * shares that array: calls that overlap on one GPU race, and a caller must
 * order them (one stream, or events).  Synchronizes `stream`. */
cgs_status cgs_phf_build_static(cgs_ctx* ctx, const uint32_t* keys, uint32_t n, uint32_t seed,
                                uint32_t* displacement, cgs_phf* out, void* stream);
/* BVH::buildHashTable's table fill (rt/BVH.cpp:24-38): table[h(keys[i])] = i,
 * CGS_EMPTY elsewhere.  table: out->table_size device u32. */
cgs_status cgs_phf_fill_table(cgs_ctx* ctx, const cgs_phf* h, const uint32_t* displacement,
                              const uint32_t* keys, uint32_t n, uint32_t* table, void* stream);
/* PerfectHashFunction::hash (rt/PerfectHashFunction.h:58-63) for n keys. */
cgs_status cgs_phf_hash(cgs_ctx* ctx, const cgs_phf* h, const uint32_t* displacement,
                        const uint32_t* keys, uint32_t n, uint32_t* out, void* stream);

/* The wide family (no counterpart in the reference, which widens its
 * PerfectHashFunction.h:58-63 family): the same function, sizes, (A,B) draws
 * and build, evaluated in exact integer arithmetic on keys of up to 64 bits,
 * h(k) = (A·(k mod T) + B + D[k & mask]) mod T.  The reference's 32-bit
 * A·k wraps, which breaks the function's near-bijection on dense heap keys:
 * every reference attempt fails at 2^21 keys (one BVH over 2^20 patches),
 * where the wide family succeeds.  Buckets are the reference's (k & mask, at
 * most 4 keys), so key sets whose low bits spread like heap keys hash;
 * others may overflow a bucket (CGS_ERR_BUCKET_OVERFLOW).  keys: n device
 * u64; n below 2^30. */
cgs_status cgs_phf_build_wide(cgs_ctx* ctx, const uint64_t* keys, uint32_t n, uint32_t seed,
                              uint32_t* displacement, cgs_phf* out, void* stream);
cgs_status cgs_phf_fill_table_wide(cgs_ctx* ctx, const cgs_phf* h, const uint32_t* displacement,
                                   const uint64_t* keys, uint32_t n, uint32_t* table, void* stream);
cgs_status cgs_phf_hash_wide(cgs_ctx* ctx, const cgs_phf* h, const uint32_t* displacement,
                             const uint64_t* keys, uint32_t n, uint32_t* out, void* stream);

/* ---------------------------------------------------------------------- */
/* Scenes (rt::Scene + MainWindow::convertScene)                            */
/* ---------------------------------------------------------------------- */

/* Builds every surface BVH (host), the top-level BVH over object world
 * bounds (Scene::buildBVH, rt/Scene.cu:39-53), uploads the scene and
 * builds every perfect hash and table on the GPU (BVH::buildHashTable,
 * rt/BVH.cpp:24-38).  The desc's host arrays are copied. */
cgs_status cgs_scene_create(cgs_ctx* ctx, const cgs_scene_desc* desc, cgs_scene** out);
cgs_status cgs_scene_destroy(cgs_scene* scene);

/* Sizes and copies (host memory) of one surface's BVH/hash (surface >= 0) or
 * of the top-level BVH (surface = -1), for inspection and tests. */
cgs_status cgs_scene_bvh_info(cgs_scene* scene, int32_t surface, uint32_t* node_count,
                              uint32_t* index_count, cgs_phf* hash);
cgs_status cgs_scene_bvh_copy(cgs_scene* scene, int32_t surface, cgs_bvh_node* nodes, uint32_t* keys,
                              uint32_t* indices, uint32_t* displacement, uint32_t* table);

/* 1 when the scene's ray queries find the objects a ray reaches through a
 * cull tree (an SAH BVH over the top-level leaf boxes) instead of walking the
 * reference's top-level BVH; 0 otherwise, -1 for a null scene.  No
 * counterpart in the reference, and no effect on any result: a ray whose
 * answer could depend on the visit order (an exact tie between objects, a
 * hit before its object's box entry, a clip-stack overflow, a direction
 * component of zero) is traced by the reference's traversal.  Built when the
 * scene has two or more objects, every control point has weight 1 and every
 * object transform is affine with a consistent inverse; CGS_CULL=0 in the
 * environment at scene creation turns it off. */
int cgs_scene_culled(cgs_scene* scene);
/* 1 when that BVH (surface >= 0, or -1: the top level, as cgs_scene_culled)
 * has a cull tree, 0 when not, -1 for a bad argument.  A surface of at least
 * 4096 patches whose control points all have weight 1 gets one (unless
 * CGS_CULL=0): closest and any hits by clipping find its patches through
 * the tree, and fall back to the reference's traversal of the surface where
 * the visit order could decide (ties, a hit before its patch's box entry, an
 * overflow, a ray with a zero component).  The results are the same. */
int cgs_scene_bvh_culled(cgs_scene* scene, int32_t surface);
/* 1 when that BVH (surface >= 0, or -1: the top level) is hashed with the
 * wide family, 0 with the reference's, -1 for a bad argument.  A scene
 * hashes a BVH with the wide family when the reference family's build ends
 * with a duplicate (from 2^21 heap keys: a surface of 2^20 patches); the
 * reference would go on with that broken table. */
int cgs_scene_bvh_wide(cgs_scene* scene, int32_t surface);

/* Scenes are read-only once created: any context on the scene's device may
 * trace or render it, and several contexts may do so at the same time.
 * The trace, clip, render and batched calls of one context share its region
 * stacks and work counters: each launch waits on the GPU for the context's
 * previous one, whatever streams they were issued on.  Contexts overlap
 * freely with each other. */

/* Closest hit of n world rays (RayTracer::Context::intersect(hit, ray),
 * RayTracer.cu:296-350: top-level hashIntersect → per-object local ray →
 * BezierSurface hashIntersect → doBezierClipping).  rays, hits, objects
 * (nullable, int32 object id or -1) are device pointers. */
cgs_status cgs_trace_closest(cgs_ctx* ctx, cgs_scene* scene, const cgs_ray* rays, uint32_t n,
                             cgs_hit* hits, int32_t* objects, void* stream);
/* Occlusion of n world rays up to ray.tmax (Context::intersect(ray),
 * RayTracer.cu:352-383).  occluded: n device u32 (1 / 0). */
cgs_status cgs_trace_any(cgs_ctx* ctx, cgs_scene* scene, const cgs_ray* rays, uint32_t n,
                         uint32_t* occluded, void* stream);

/* cgs_trace_closest / cgs_trace_any with a leaf method and its parameters
 * (options may be null: the default).  CGS_ERR_INVALID for a depth above 16
 * or a negative or non-finite tolerance. */
cgs_status cgs_trace_closest_ex(cgs_ctx* ctx, cgs_scene* scene, const cgs_ray* rays, uint32_t n, cgs_hit* hits,
                                int32_t* objects, const cgs_trace_options* options, void* stream);
cgs_status cgs_trace_any_ex(cgs_ctx* ctx, cgs_scene* scene, const cgs_ray* rays, uint32_t n, uint32_t* occluded,
                            const cgs_trace_options* options, void* stream);

/* RayTracer::render (rt/RayTracer.h:42-43, RayTracer.cu:107-180 and the
 * render kernel :194-544, non-Monte-Carlo): camera rays through integer
 * pixel centres, closest hit, PBR shading with one shadow ray per light and
 * the reflection bounce loop.  frame: `height` rows of packed pixels
 * (device), options->frame_stride (or width) pixels apart, pixel (x, y) at
 * frame[stride*y + x] like rt::Frame::at; the padding is not written.
 * color: optional width*height*3 floats (device, unclamped radiance, dense);
 * counts: optional device cgs_ray_counts (accumulated).
 * Asynchronous on `stream`; CGS_ERR_BUSY while a previous render of this
 * context is still running (RayTracer.cu:109-115).
 * Scheduling: each render records how long each 8×8 pixel tile took, and the
 * next render of a frame of the same size on this scene takes the tiles
 * longest first by those times; a scene's first render at a size takes them
 * row by row.  The frame, the colours and the counts are the same in any
 * order. */
cgs_status cgs_render(cgs_ctx* ctx, cgs_scene* scene, const cgs_camera* camera,
                      const cgs_render_options* options, uint32_t width, uint32_t height,
                      uint32_t* frame, float* color, cgs_ray_counts* counts, void* stream);
/* RayTracer::running() (rt/RayTracer.h:45-48): 1 while the last render runs. */
int cgs_render_running(cgs_ctx* ctx);

/* The sampled (Monte Carlo) render: RayTracer::render built with
 * USE_MONTECARLO_SAMPLING (rt/StaticConfig.h:5; RayTracer.cu:212-246 and the
 * sampled branches of closestHit, RayTracer.cu:398-407,473-522), with
 * RayTracer::Options::samples (rt/RayTracer.h:28, default 8).  Each pixel
 * traces `samples` camera rays through (i + s/n, j + vdc(s)) and averages
 * them; a hit has no ambient term and its reflection mirrors about a GGX
 * microfacet normal drawn from the pixel's random stream (the cell
 * (i mod 256, j mod 256) of a 256×256 grid, as the reference's).  The draws
 * are Philox4x32-10 (Salmon et al., SC'11) under `seed` (the reference
 * seeds 42): the same for the same seed, scene, camera and options, but not
 * cuRAND's stream, so frames equal the reference's in distribution only.
 * Arguments and errors as cgs_render; samples outside 1..CGS_MAX_SAMPLES is
 * CGS_ERR_INVALID. */
#define CGS_MAX_SAMPLES 65536u
cgs_status cgs_render_sampled(cgs_ctx* ctx, cgs_scene* scene, const cgs_camera* camera,
                              const cgs_render_options* options, uint32_t samples, uint32_t seed, uint32_t width,
                              uint32_t height, uint32_t* frame, float* color, cgs_ray_counts* counts, void* stream);

/* Shaded ray queries: each of n world rays is traced and shaded as the render
 * kernel traces and shades a camera ray (the trace loop of RayTracer.cu:
 * 194-294 with Context::trace / miss / closestHit, :385-544, non-Monte-Carlo;
 * the reference starts it only at camera rays): closest hit up to ray.tmax,
 * background on a miss, else ambient plus one shadow ray per light with the
 * PBR terms, then the reflection bounce (to infinity) while
 * options->recursion_depth remains and the attenuation exceeds 2^-8.
 * Directions are used as given: unit directions give the render's shading,
 * as its camera rays are unit.  options: eps and recursion_depth
 * (frame_stride unused).  Device pointers: color (3 floats per ray,
 * unclamped radiance), pixels (nullable: the frame's packed sRGB), hits and
 * objects (nullable: the first segment's closest hit as cgs_trace_closest
 * reports it; patch CGS_PATCH_OVERFLOW when any segment overflowed its clip
 * stack); counts (nullable, accumulated): rays of each kind as cgs_render. */
cgs_status cgs_shade(cgs_ctx* ctx, cgs_scene* scene, const cgs_ray* rays, uint32_t n, const cgs_render_options* options,
                     float* color, uint32_t* pixels, cgs_hit* hits, int32_t* objects, cgs_ray_counts* counts,
                     void* stream);

/* Work-event totals of the renders on `ctx`'s device since the last reset,
 * for the work model (no counterpart in the reference): cull-tree node and
 * element boxes tested, BVH inner nodes visited, perfect-hash probes, leaf
 * elements visited, patch tests, clipping passes that clipped / split /
 * pruned, converged regions, candidates evaluated, object-local rays (12
 * values, in that order, into out[0..n)).  Only a library built with
 * CGS_WORK_COUNTS=1 counts (CGS_ERR_INVALID from any other build). */
cgs_status cgs_work_counts(cgs_ctx* ctx, uint64_t* out, uint32_t n, int reset);

/* ---------------------------------------------------------------------- */
/* Batched queries over many resident scenes (no counterpart in the        */
/* reference: each job is Context::intersect over a ray range,             */
/* RayTracer.cu:296-383, and all jobs share one persistent launch)         */
/* ---------------------------------------------------------------------- */

/* One launch over every job.  Asynchronous on `stream`; batched launches of
 * one context run one at a time (a call waits for the previous one's kernel
 * before reusing the context's job table).  Refused (CGS_ERR_INVALID) when
 * the jobs' rays, plus one chunk per resident wave, reach 2^32. */
cgs_status cgs_trace_batch(cgs_ctx* ctx, const cgs_batch_job* jobs, uint32_t njobs, void* stream);
/* A fixed batch with its own job table, counter and region stacks, captured
 * once in a HIP graph; cgs_batch_launch replays it (results as
 * cgs_trace_batch).  Different batches may run concurrently on different
 * streams; one batch must not be relaunched before its previous launch ends. */
cgs_status cgs_batch_create(cgs_ctx* ctx, const cgs_batch_job* jobs, uint32_t njobs, cgs_batch** out);
cgs_status cgs_batch_launch(cgs_batch* batch, void* stream);
cgs_status cgs_batch_destroy(cgs_batch* batch);
/* Leaf method of the context's later batched calls and batches
 * (cgs_trace_options' meaning; null: the reference's clipping).  Closest-hit
 * jobs are polished when the options ask for it. */
cgs_status cgs_batch_set_trace_options(cgs_ctx* ctx, const cgs_trace_options* options);
cgs_status cgs_batch_get_trace_options(cgs_ctx* ctx, cgs_trace_options* options);
/* Accuracy-gated tuning of the leaf method (no counterpart in the reference):
 * runs the jobs with the reference's clipping (the reference results,
 * polished when limits->allow_polish), then with each candidate: clipping
 * with a pass cap of 8, 12, 16, 24 or 32, subdivision at depth 4, 5, 6, 8,
 * 10 or 12, and the reference's clipping, each also polished when allowed;
 * times each (the faster of two launches at the context's launch shape),
 * compares its results with the reference's, and keeps the fastest candidate
 * within the limits as the context's batched leaf method (the reference
 * candidate always is).  report (up to report_cap entries, nullable) and
 * report_count receive every candidate.  The jobs' outputs are overwritten
 * with valid results of some candidate. */
cgs_status cgs_batch_autotune_accuracy(cgs_ctx* ctx, const cgs_batch_job* jobs, uint32_t njobs,
                                       const cgs_accuracy_limits* limits, cgs_trace_options* best,
                                       cgs_autotune_entry* report, uint32_t report_cap, uint32_t* report_count);
/* Launch shape of the context's later batched calls and batches. */
cgs_status cgs_batch_set_config(cgs_ctx* ctx, const cgs_batch_config* config);
cgs_status cgs_batch_get_config(cgs_ctx* ctx, cgs_batch_config* config);
/* Times every candidate shape (chunk 64/128/256, all to one workgroup per CU
 * below the maximum) on the given jobs, keeps the fastest in the context and
 * returns it.  The jobs' outputs are overwritten with valid results. */
cgs_status cgs_batch_autotune(cgs_ctx* ctx, const cgs_batch_job* jobs, uint32_t njobs, cgs_batch_config* best);

/* ---------------------------------------------------------------------- */
/* Serving: continuous batching of ray-query requests against resident     */
/* scenes (no counterpart in the reference)                                */
/* ---------------------------------------------------------------------- */

typedef struct cgs_server_options
{
    size_t hbm_budget;      /* bytes of device ray + result buffers in flight (0: 1 GiB) */
    uint32_t slots;         /* launches in flight, each on its own stream (0: 3) */
    uint32_t max_batch_rays;/* rays per launch (0: no limit beyond the budget) */
} cgs_server_options;

/* A server over resident scenes: a worker thread admits pending requests
 * into batched launches whenever a slot is free and the byte budget allows,
 * so requests that arrive while others run join the next launch. */
cgs_status cgs_server_create(cgs_ctx* ctx, cgs_scene* const* scenes, uint32_t nscenes,
                             const cgs_server_options* options, cgs_server** out);
/* Queues a request: n host rays against scene `scene`; any = 0: closest hits
 * into host `hits` (and `objects`, nullable); any = 1: occlusion flags into
 * host `occluded`.  The buffers must stay valid until the request is done.
 * Fails with CGS_ERR_INVALID if the request alone exceeds the byte budget or
 * one launch's 32-bit work counter (n plus one chunk per resident wave). */
cgs_status cgs_server_submit(cgs_server* server, uint32_t scene, const cgs_ray* rays, uint32_t n, int32_t any,
                             cgs_hit* hits, int32_t* objects, uint32_t* occluded, uint64_t* id);
/* Queues a shaded request: n host rays traced and shaded as cgs_shade does,
 * with options' eps and recursion_depth, into host color (3 floats per ray)
 * and, when not null, pixels, hits and objects.  Shaded requests join the
 * same launches as the others (a second kernel of the launch).  Same
 * failures as cgs_server_submit. */
cgs_status cgs_server_submit_shaded(cgs_server* server, uint32_t scene, const cgs_ray* rays, uint32_t n,
                                    const cgs_render_options* options, float* color, uint32_t* pixels, cgs_hit* hits,
                                    int32_t* objects, uint64_t* id);
/* 1 when request `id` is done (outputs written), 0 while pending, -1 when
 * its launch failed or the id is unknown or already collected; latency_ns
 * (nullable): submit to done.  A result of 1 or -1 collects the request: the
 * server forgets it, so a long-running server keeps no record per request. */
int cgs_server_poll(cgs_server* server, uint64_t id, uint64_t* latency_ns);
/* Waits for request `id` and collects it like cgs_server_poll;
 * CGS_ERR_INVALID for an unknown or already collected id. */
cgs_status cgs_server_wait(cgs_server* server, uint64_t id, uint64_t* latency_ns);
/* Finishes every queued request, then stops the worker. */
cgs_status cgs_server_destroy(cgs_server* server);

/* ---------------------------------------------------------------------- */
/* Device memory helpers (for callers without their own allocator)         */
/* ---------------------------------------------------------------------- */

cgs_status cgs_malloc(cgs_ctx* ctx, size_t bytes, void** dptr);
cgs_status cgs_free(cgs_ctx* ctx, void* dptr);
cgs_status cgs_memcpy_h2d(cgs_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
cgs_status cgs_memcpy_d2h(cgs_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
cgs_status cgs_memset(cgs_ctx* ctx, void* dst, int value, size_t bytes, void* stream);
cgs_status cgs_stream_sync(cgs_ctx* ctx, void* stream);
/* A non-blocking stream on the context's device (hipStream_t as void*). */
cgs_status cgs_stream_create(cgs_ctx* ctx, void** stream);
cgs_status cgs_stream_destroy(cgs_ctx* ctx, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CGS_H */
