/*
 * femrt.h -- C ABI of the MI355X curved-hex FEM ray tracer.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  No C++
 * or torch types cross this boundary.  Every call returns FEMRT_OK (0) or a
 * negative FEMRT_E* code; femrt_last_error() returns the message of the last
 * failure on a context.  A context is externally synchronised: one host
 * thread per context.  Device work is queued on the context's HIP stream;
 * calls that hand results back to host memory synchronise that stream.
 *
 * Reference interfaces replaced (paths relative to FEM-AABB/ of the
 * reference tree):
 *   femrt_mesh_upload_hex   <- Scene::add_hex + the host pointer patching and
 *                              cudaMemcpy of HexElementData / coefficient pools
 *                              (raytracer.h:118-166, main.cu:336-511)
 *   femrt_mesh_build_bvh    <- Scene::build_bvh -> BVH::build
 *                              (raytracer.h:338-344, BVH.cu:180-247)
 *   femrt_eval_field        <- EvaluateModalField / EvaluateModalGradientWorld
 *                              (fem/HexElement.cu:274-330)
 *   femrt_invert_points     <- WorldToReferenceNewton (fem/HexElement.cu:227-272)
 *   femrt_trace             <- render_kernel -> ray_color -> bvh_intersect_device
 *                              hex branch (raytracer.h:313-331,
 *                              raytracer.cu:102-254, 410-476)
 *   femrt_gen_structured_hex, femrt_load_xml_fld
 *                           <- load_xml_fld (xml_fld_loader.h:14-16) and the
 *                              geometry fit it calls (BuildHexGeometry,
 *                              xml_fld_geometry_builder.cpp:68-381)
 */
#ifndef FEMRT_H
#define FEMRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEMRT_API __attribute__((visibility("default")))

enum {
  FEMRT_OK = 0,
  FEMRT_EINVAL = -1,   /* bad argument (null pointer, size, mode count)   */
  FEMRT_EHIP = -2,     /* a HIP runtime call failed                        */
  FEMRT_ENOMEM = -3,   /* device or host allocation failed                 */
  FEMRT_EUNSUP = -4,   /* mode counts / options this build has no kernel for */
  FEMRT_EIO = -5,      /* file missing or malformed                        */
  FEMRT_ESTATE = -6    /* call out of order (e.g. trace before BVH build)  */
};

typedef struct femrt_ctx femrt_ctx;
typedef struct femrt_mesh femrt_mesh;
typedef struct femrt_frame femrt_frame;

/* ---------------------------------------------------------------- context */
FEMRT_API int femrt_version(void);
FEMRT_API int femrt_ctx_create(int device, femrt_ctx** out);
FEMRT_API int femrt_ctx_destroy(femrt_ctx* ctx);
FEMRT_API const char* femrt_last_error(const femrt_ctx* ctx);
FEMRT_API int femrt_sync(femrt_ctx* ctx);
/* Kernel timing: when enabled, every kernel launch is bracketed by HIP events
 * on the context stream and its duration accumulated per kernel name. */
FEMRT_API int femrt_timing_enable(femrt_ctx* ctx, int on);
FEMRT_API int femrt_timing_reset(femrt_ctx* ctx);
/* Copies up to cap entries: names (NUL-separated, 48 bytes each), total ms,
 * launch counts.  Returns the number of distinct kernels (>=0). */
FEMRT_API int femrt_timing_read(femrt_ctx* ctx, int cap, char* names48, double* total_ms,
                                long long* launches);

/* ------------------------------------------------------------------ mesh */
/* Uniform-order hex mesh in SoA host arrays (the reference stores per-element
 * mode counts; every mesh of the benchmark configurations is uniform, and a
 * mixed-order mesh is uploaded as several meshes).
 *   vertices  n_elem x 8 x 3   corner vertices, XML/VTK order (HexElement.cuh:10)
 *   geom      n_elem x 3 x ng^3 modal geometry coefficients, x block, y block,
 *             z block; mode (i,j,k) at (k*ng+j)*ng+i (StdHexExp.cpp:516-532)
 *   field     n_elem x nf^3     modal field coefficients, same ordering
 *   adjacency n_elem x 6 x 2    (neighbour element, neighbour local face) per
 *             local face 0..5 = t-,s-,r+,s+,r-,t+ (high_order.h:172-184), -1 on
 *             the boundary; may be NULL (element walk then re-queries the BVH)
 * The mesh is copied to device memory; the host arrays may be freed after the
 * call returns. */
FEMRT_API int femrt_mesh_upload_hex(femrt_ctx* ctx, int n_elem, int ng, int nf,
                                    const float* vertices, const float* geom, const float* field,
                                    const int* adjacency, femrt_mesh** out);
/* Curved quadrilateral surface elements (the reference's QuadElementData,
 * fem/QuadElement.cuh:9-18, Scene::add_quad, raytracer.h:213-257, traversal
 * branch raytracer.cu:297-320).  vertices: n x 4 x 3 floats, vertex q = i + 2 j
 * at (r, s) = (2 i - 1, 2 j - 1); geom: n x 3 x ng^2 Modified_A coefficients
 * (x, y, z), mode (i, j) at j ng + i, or NULL for the bilinear map of the
 * vertices; field: n x nf^2, or NULL (zero field).  The mesh traces like a
 * hex mesh (same BVH, face Newton with sub-patch seeds, FP64 polish) with
 * each quad as face 0 of a map constant in the third coordinate: hit face 0,
 * rst = (r, s, -1), u = the quad's field.  No isosurfaces (iso_enable = 0). */
FEMRT_API int femrt_mesh_upload_quads(femrt_ctx* ctx, int n_quad, int ng, int nf, const float* vertices,
                                      const float* geom, const float* field, femrt_mesh** out);
/* Prisms (the reference's PrismElementData, fem/PrismElement.cuh:9-18,
 * Scene::add_prism, raytracer.h:167-212, traversal branch raytracer.cu:255-295,
 * where the isosurface search of prisms is switched off).  vertices: n x 6 x 3
 * floats; geom: n x 3 blocks of Modified_A(i, r) Modified_A(j, s)
 * Modified_B(i, k, t) coefficients, i slowest, then j, then k < mz - i
 * (EvaluatePrismModal, fem/PrismElement.cu:16-37), with geom_modes = (mx, my,
 * mz), or NULL for the reference's linear map of the vertices
 * (fem/PrismElement.cu:194-226); field alike with field_modes, or NULL (zero
 * field).  (r, s, t) are the reference's collapsed coordinates, so a prism is
 * a hex map whose t = +1 face is the edge v4 v5: the mesh is traced as that
 * hex (exact embedding, hex modes max(mx, my, mz), coefficients rounded to
 * float once).  Hits report the hex face (0: t = -1, 1: s = -1, 2: r = +1,
 * 3: s = +1, 4: r = -1), rst in the prism's own coordinates and u.  No
 * isosurfaces (iso_enable = 0). */
FEMRT_API int femrt_mesh_upload_prisms(femrt_ctx* ctx, int n_prism, const int* geom_modes,
                                       const int* field_modes, const float* vertices, const float* geom,
                                       const float* field, femrt_mesh** out);
/* The embedding femrt_mesh_upload_prisms uploads, in double: ng, nf (hex
 * modes per direction: max of the mode counts, 2 and 1 for NULL geometry
 * and field), and, where the output pointers are not NULL, hex_vertices
 * (n x 8 x 3, XML corner order), hex_geom (n x 3 x ng^3) and hex_field
 * (n x nf^3) in the hex layout of femrt_mesh_upload_hex. */
FEMRT_API int femrt_host_prism_to_hex(int n_prism, const int* geom_modes, const int* field_modes,
                                      const float* vertices, const float* geom, const float* field, int* ng,
                                      int* nf, float* hex_vertices, double* hex_geom, double* hex_field);
/* Curved triangles (the reference's TriElementData, fem/TriElement.cuh:9-17,
 * Scene::add_curved_tri, raytracer.h:259-304).  vertices: n x 3 x 3 floats;
 * geom: n x 3 blocks of Modified_A(i, r) Modified_B(i, j, s) coefficients,
 * i slowest, j < my - i (EvaluateTriModal, fem/TriElement.cu:46-63), with
 * geom_modes = (mx, my), or NULL for the flat triangle v0 v1 v2; field alike
 * with field_modes, or NULL (zero).  (r, s) are collapsed coordinates, the
 * collapsed edge at s = +1 (for the flat triangle: barycentric weights of
 * v1, v2 = (1+r)(1-s)/4, (1+s)/2), so a triangle is a quad map with one
 * collapsed edge: uploaded as that quad (femrt_mesh_upload_quads; exact
 * embedding, coefficients rounded to float once).  Hits report face 0,
 * rst = (r, s, -1) and u.  No isosurfaces. */
FEMRT_API int femrt_mesh_upload_tris(femrt_ctx* ctx, int n_tri, const int* geom_modes, const int* field_modes,
                                     const float* vertices, const float* geom, const float* field,
                                     femrt_mesh** out);
/* The embedding femrt_mesh_upload_tris uploads, in double: ng, nf (quad
 * modes: max of the mode counts, 2 and 1 for NULL geometry and field) and,
 * where not NULL, quad_vertices (n x 4 x 3), quad_geom (n x 3 x ng^2) and
 * quad_field (n x nf^2) in the layout of femrt_mesh_upload_quads. */
FEMRT_API int femrt_host_tri_to_quad(int n_tri, const int* geom_modes, const int* field_modes,
                                     const float* vertices, const float* geom, const float* field, int* ng, int* nf,
                                     float* quad_vertices, double* quad_geom, double* quad_field);
FEMRT_API int femrt_mesh_destroy(femrt_mesh* mesh);
/* Conservative boxes from the exact Bernstein hull of the modal geometry,
 * split 3 ways per reference direction (replaces the sampled +15% box of
 * ComputeHexAabb, high_order.h:51-113).  out: n_elem x 61 x 6 floats: [0]
 * element box, [1..6] face boxes (Nektar face order), [7 + 9 f + 3 b + a] box
 * of sub-patch (a, b) of face f on the 3 x 3 split of its (u, v) square; each
 * lo.xyz, hi.xyz. */
FEMRT_API int femrt_mesh_boxes(const femrt_mesh* mesh, float* out);
/* builder FEMRT_BVH_HOST_SAH (0): binned SAH on the host, 16 bins, leaves of
 * <= 4 elements (BVH.cu:35-177, replaces BVH::build called by
 * Scene::build_bvh, raytracer.h:338-344).
 * builder FEMRT_BVH_GPU_LBVH (1): linear BVH built on the device on the
 * context's stream (Morton codes of the element-box centroids, radix sort,
 * binary radix tree, boxes bottom-up, subtrees of <= 4 elements as leaves);
 * faster to build, slower to trace than SAH.  The reference has no GPU
 * builder.  Both give the traversal the same answers (the nearest hit does
 * not depend on the tree, except which of two equally near hits is kept).  Returns the node count via *n_nodes (may be NULL;
 * LBVH node arrays keep 2 n - 1 nodes, some unreferenced). */
#define FEMRT_BVH_HOST_SAH 0
#define FEMRT_BVH_GPU_LBVH 1
FEMRT_API int femrt_mesh_build_bvh(femrt_mesh* mesh, int builder, int* n_nodes);
/* Host copy of the BVH: nodes x 8 floats (lo.xyz, a as int bits, hi.xyz, b
 * as int bits).  Internal node: a, b = child node indices.  Leaf: b < 0,
 * a = first slot in the element index list, -b = element count.  prims:
 * n_elem element indices in leaf order. */
FEMRT_API int femrt_mesh_get_bvh(const femrt_mesh* mesh, float* nodes8, int* prims);

/* ---------------------------------------------------------- point kernels */
/* Field value and world gradient at (elem, rst) points: out is 4 x n SoA
 * (u, du/dx, du/dy, du/dz).  rst is 3 x n SoA.  precision 0 = FP32, 1 = FP64
 * arithmetic (results stored as float / double respectively: out points to
 * float or double). Host buffers; synchronous. */
FEMRT_API int femrt_eval_field(femrt_mesh* mesh, int n, const int* elem, const float* rst,
                               int precision, void* out);
/* femrt_eval_field without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Inputs and
 * outputs are staged in a fixed device array of the library (3 MiB per GPU,
 * held from the first call on that GPU until the process ends), so n is at
 * most FEMRT_NOALLOC_MAX_POINTS (larger n: FEMRT_EINVAL).  Same arguments and
 * results as femrt_eval_field.  Calls on one GPU share the array: calls that
 * overlap on one GPU (contexts of the same device used from several host
 * threads) race. */
#define FEMRT_NOALLOC_MAX_POINTS 65536
FEMRT_API int femrt_eval_field_noalloc(femrt_mesh* mesh, int n, const int* elem, const float* rst,
                                       int precision, void* out);
/* Point inversion x -> rst: the preimage inside the element.  Newton from
 * rst = 0 as WorldToReferenceNewton, restarted from the 8 octant centres when
 * it lands outside the element (Newton from 0 finds spurious preimages of the
 * polynomial map near faces), then an FP64 polish.  x is 3 x n SoA; rst_out
 * 3 x n SoA doubles; iters_out (may be NULL) the FP32 iteration count, -1
 * when no preimage inside the element was found. */
FEMRT_API int femrt_invert_points(femrt_mesh* mesh, int n, const int* elem, const float* x,
                                  double* rst_out, int* iters_out);

/* ------------------------------------------------------------------ trace */
enum {
  FEMRT_GUESS_PLAIN = 0,   /* reference initial guesses: face seeds {-0.8,0,0.8}^2,
                              point inversion from rst = 0 */
  FEMRT_GUESS_WARM = 1,    /* engineered: the predicted exit, sub-patch face seeds,
                              and scan samples started from the cubic Hermite
                              interpolant of rst(t) between the segment's ends
                              (FEMRT_TRACE_COOP_SCAN), else extrapolated from the
                              two previous samples */
  FEMRT_GUESS_LEARNED = 2  /* MLP initial guesses (femrt_mlp_*) for the point
                              inversions with no earlier sample to start from
                              (all of them with FEMRT_TRACE_LEARNED_EVERY_SAMPLE),
                              warm starts elsewhere, plain fallback (never with
                              FEMRT_TRACE_COOP_SCAN) */
};

#define FEMRT_MAX_ISO 8

typedef struct {
  int n_iso;                     /* isovalues searched (0..FEMRT_MAX_ISO)         */
  int iso_enable;                /* 0 disables the iso stage (0.0 is a valid iso) */
  double iso[FEMRT_MAX_ISO];      /* exact isovalues (the FP32 scan uses their
                                    rounding, the FP64 endgame the values)  */
  int samples;                   /* uniform sub-intervals per element segment (16) */
  int guess;                     /* FEMRT_GUESS_*                                 */
  float t_min;                   /* 0.001 in the reference (raytracer.cu:221)     */
  float t_max;
  int max_march;                 /* element-walk step cap per ray (at most 65535)  */
  int flags;                     /* FEMRT_TRACE_* bits                            */
} femrt_trace_params;

#define FEMRT_TRACE_RGBA 1             /* shade into the frame's RGBA                 */
#define FEMRT_TRACE_BINNED_WALK 2      /* walk in rounds with the rays binned by
                                          element, re-entries in their own kernel,
                                          instead of one persistent kernel that
                                          walks each ray to its end                 */
#define FEMRT_TRACE_LEARNED_EVERY_SAMPLE 4 /* GUESS_LEARNED: network guesses for every
                                         interior scan sample; by default only
                                         the first of each segment takes one (no
                                         earlier sample to extrapolate from) and
                                         the rest extrapolated warm starts */
#define FEMRT_TRACE_NO_SAMPLE_SKIP 8   /* evaluate every scan sample, even where the
                                          element's bounds certify its signs from
                                          an earlier sample (same results)          */
#define FEMRT_TRACE_COOP_SCAN 16       /* with plain or warm guesses and at most 16
                                          samples per segment: spread the samples of
                                          a wave's segments over its lanes, each
                                          started from the cubic Hermite
                                          interpolant of rst(t) between the
                                          segment's ends; by default each segment
                                          is scanned in its ray's lane, sample after
                                          sample (warm starts extrapolated from the
                                          two previous samples, certified skipping) */
#define FEMRT_TRACE_ISO_BISECT 32      /* FP64 iso endgame by plain bisection of the
                                          scan's bracket (default: Newton from the
                                          bracket's secant point, bisection
                                          safeguard); stops within 1e-10 (1 + |t|)
                                          with |u - iso| <= 1e-8 */
#define FEMRT_TRACE_ISO_LEARNED_BRACKET 64 /* FP64 bisection endgame started from the
                                          bracket network's predicted [t_lo, t_hi]
                                          (femrt_mesh_set_bracket_mlp) where its ends
                                          narrow the scan's bracket; the crossing
                                          is still the scan's first.  Records the
                                          segments like FEMRT_TRACE_SEG_INFO.  Not
                                          with FEMRT_TRACE_BINNED_WALK.  No
                                          counterpart in the reference (its endgame:
                                          fem/RayIsosurface.cu:121-171) */
#define FEMRT_TRACE_NESTED_FIRST_HIT 256 /* WARM / LEARNED: run each lane's
                                          first-hit face searches to completion
                                          before it traverses on (as the walk's
                                          re-entries do); by default the
                                          traversal hands face-Newton jobs to a
                                          loop that runs one Newton iteration on
                                          all the wave's lanes at a time.  Same
                                          hits either way                         */
#define FEMRT_TRACE_NO_HERMITE_START 512 /* WARM / LEARNED: start the first
                                          interior sample of a segment from
                                          the entry point (warm) or the
                                          network (learned) even where the
                                          predicted exit left the tangents
                                          drst/dt at both ends; by default
                                          the cubic Hermite interpolant of
                                          rst(t) between entry and exit is
                                          the start there, and the network
                                          guesses only segments without
                                          tangents.  Same results (the FP64
                                          endgame), other FP32 iterates     */
#define FEMRT_TRACE_NO_SKIP_JUMP 1024 /* walk every element along the ray; by
                                          default, after two segments in a
                                          row whose field bounds exclude
                                          every open isovalue, leading into a
                                          neighbour whose bounds do too, a
                                          BVH query that passes such
                                          elements (entry roots only) finds
                                          the next element that can hold a
                                          crossing, or ends the walk.  The
                                          elements passed hold no crossing   */
#define FEMRT_TRACE_SEG_INFO 128       /* record each bracket's ends (rst and the
                                          FP32 g at ta and tb), the bracket
                                          network's inputs, for
                                          femrt_frame_bracket_pairs */

FEMRT_API void femrt_trace_params_default(femrt_trace_params* p);

/* A frame owns device buffers for n_rays rays and their results.
 * Per-ray results (SoA, device):
 *   hit_elem int32 (-1 miss), hit_face int32, hit_t f64, hit_rst 3 x f64,
 *   hit_u f64 (field at the hit), and per isovalue k: iso_elem int32,
 *   iso_t f64, iso_rst 3 x f64, iso_u f64; rgba uint32. */
FEMRT_API int femrt_frame_create(femrt_ctx* ctx, int n_rays, femrt_frame** out);
FEMRT_API int femrt_frame_destroy(femrt_frame* f);
/* rays: 6 x n SoA floats (ox, oy, oz, dx, dy, dz), host memory */
FEMRT_API int femrt_frame_set_rays(femrt_frame* f, const float* rays);
/* pinhole camera (raytracer.h:31-48), pixel-centre rays, y = 0 bottom row,
 * generated on the device.  cam: lookfrom.xyz, lookat.xyz, vup.xyz, vfov_deg */
FEMRT_API int femrt_frame_camera_rays(femrt_frame* f, const float* cam10, int width, int height);
FEMRT_API int femrt_frame_get_rays(femrt_frame* f, float* rays);
/* Queue the trace of every ray of the frame on the context stream (async). */
FEMRT_API int femrt_trace(femrt_mesh* mesh, femrt_frame* f, const femrt_trace_params* p);
/* Copy results to host (synchronises).  Any pointer may be NULL.  iso arrays
 * are n_iso x n_rays.  A ray without a hit reports elem -1, face -1 and NaN
 * t, rst and u; a (ray, isovalue) pair without a crossing reports elem -1 and
 * NaN t, rst and u. */
FEMRT_API int femrt_frame_get_hits(femrt_frame* f, int* hit_elem, int* hit_face, double* hit_t,
                                   double* hit_rst, double* hit_u);
FEMRT_API int femrt_frame_get_iso(femrt_frame* f, int n_iso, int* iso_elem, double* iso_t,
                                  double* iso_rst, double* iso_u);
FEMRT_API int femrt_frame_get_rgba(femrt_frame* f, uint32_t* rgba);
/* Work counters of the last trace (validation runs): [0] rays, [1] hit rays,
 * [2] face-Newton iterations FP32 of the element walk (exits and BVH
 * re-entries), [3] point-inversion iterations FP32,
 * [4] field evaluations FP32, [5] FP64 geometry iterations, [6] FP64 field
 * evaluations, [7] slab tests, [8] MLP inferences, [9] elements marched,
 * [10] segment exits found by prediction, [11] by face search, [12] by
 * marching, [13] BVH re-entries of the walk, [14] redone segments,
 * [15] isovalue brackets found, [16..18] timeline of the persistent walk on
 * the constant 100 MHz clock: start, the moment its hit queue first ran dry
 * (after that lanes only finish the rays they hold), end; [19..26] rays
 * by segments walked, log2 bins 1, 2-3, 4-7, ..., 128+; [27] segments whose
 * scan the element's field bounds ruled out; [28] scan samples whose signs
 * the field and geometry bounds certified without evaluating them; [29] walks
 * ended as stalled (8 stalls skipped already; a stall is a run of more than
 * 16 exits in a row found by the marching fallback, as for a ray inside a
 * face two elements share); [30] face-Newton iterations FP32 of the
 * first-hit pass, [31] its slab tests ([7] counts the slab tests of the
 * other passes); [32] stalls skipped past (the entry moved 1e-3 of the
 * element along the ray); [33] runs of elements that cannot hold a crossing
 * passed by a BVH query (FEMRT_TRACE_NO_SKIP_JUMP; also counted in [13]). */
#define FEMRT_NCOUNTERS 34
FEMRT_API int femrt_frame_counters(femrt_frame* f, int enable, long long* out);
/* Ray ids of the walks of the last counted trace that stalled (skipped past
 * a stall or ended stalled, counters [32] and [29]): copies up to
 * min(cap, FEMRT_MAX_STALLED_IDS) of them, in no set order, and returns how
 * many walks stalled (>= 0). */
#define FEMRT_MAX_STALLED_IDS 64
FEMRT_API int femrt_frame_stalled_rays(femrt_frame* f, int cap, int* ids);

/* ------------------------------------------------ shading and bounces */
/* Materials of scatter (raytracer.cu:61-97): type 0 Lambertian (cosine-
 * weighted), 1 metal (roughness = fuzz), 2 dielectric (ior); color is the
 * attenuation, emission ends a path with color * emission.  A geometry hit on
 * local face 0 takes material 1, other faces material 0 (raytracer.cu:248). */
typedef struct {
  float color[3];
  float emission[3];
  float roughness;
  float ior;
  int type;
} femrt_material;

/* Jittered primary rays of a pinhole camera (render_kernel,
 * raytracer.cu:431-443): the frame must hold width*height*spp rays; path
 * p = (y*width + x)*spp + s goes through ((x + u)/width, (y + v)/height)
 * with u, v drawn from a counter-based generator keyed by (seed, p).
 * cam12: origin, lower_left_corner, horizontal, vertical (raytracer.h:22-49). */
FEMRT_API int femrt_frame_camera_rays_jitter(femrt_frame* f, const float* cam12, int width,
                                             int height, int spp, unsigned long long seed);
/* Path state: throughput 1 and radiance 0 for every ray of the frame. */
FEMRT_API int femrt_frame_path_begin(femrt_frame* f);
/* Uploads the frame's material table (at most 64) ahead of scatter; scatter
 * uploads it itself when the table it is given differs from the frame's. */
FEMRT_API int femrt_frame_set_materials(femrt_frame* f, const femrt_material* mats, int n_mats);
/* One bounce after a geometry trace of the frame (femrt_trace, iso stage
 * off): misses add throughput * sky (raytracer.cu:398-402) and end; hits
 * scatter off the surface (normal from the hit face's tangents, outward and
 * front face from the element's orientation) into the frame's rays for the
 * next trace.  depth numbers the bounce (0 = primary hit); last = 1 ends the
 * paths still alive with black (the depth cap of ray_color). */
FEMRT_API int femrt_frame_scatter(femrt_mesh* mesh, femrt_frame* f, const femrt_material* mats,
                                  int n_mats, unsigned long long seed, int depth, int last);
/* Mean radiance over each pixel's samples, gamma sqrt, clamp, 8-bit RGBA
 * (raytracer.cu:460-475) into the frame's RGBA buffer; copied to rgba
 * (width*height) when rgba is not NULL, else read later with
 * femrt_frame_get_rgba. */
FEMRT_API int femrt_frame_resolve(femrt_frame* f, int width, int height, int spp, uint32_t* rgba);
/* camera_rays_jitter + path_begin + max_depth x (trace + scatter) + resolve. */
FEMRT_API int femrt_render_path(femrt_mesh* mesh, femrt_frame* f, const float* cam12, int width,
                                int height, int spp, int max_depth, const femrt_material* mats,
                                int n_mats, unsigned long long seed, uint32_t* rgba);

/* ---------------------------------------------- batches and launch shapes */
/* No counterpart in the reference (render_kernel renders one mesh per
 * launch, raytracer.cu:410-476): a batch records traces and path-traced
 * renders of many meshes of uneven size, each into its own frame, into one
 * HIP graph that femrt_batch_launch replays with one launch on the context
 * stream.  Jobs on one frame run in the order they were added; jobs on
 * different frames may run side by side (the graph has up to four parallel
 * branches, each frame on one), so a job must not read a frame another job
 * writes.  A mesh may serve several jobs.  Not capturable: frames
 * with work counters on, the binned walk (FEMRT_TRACE_BINNED_WALK).  Replays
 * rewrite each job's frame from scratch; read results after femrt_sync. */
typedef struct femrt_batch femrt_batch;
FEMRT_API int femrt_batch_create(femrt_ctx* ctx, femrt_batch** out);
FEMRT_API int femrt_batch_destroy(femrt_batch* b);
FEMRT_API int femrt_batch_add_trace(femrt_batch* b, femrt_mesh* mesh, femrt_frame* f,
                                    const femrt_trace_params* p);
/* camera_rays_jitter + path_begin + max_depth x (trace + scatter) + resolve,
 * as femrt_render_path; the RGBA stays in the frame (femrt_frame_get_rgba). */
FEMRT_API int femrt_batch_add_paths(femrt_batch* b, femrt_mesh* mesh, femrt_frame* f, const float* cam12,
                                    int width, int height, int spp, int max_depth,
                                    const femrt_material* mats, int n_mats, unsigned long long seed);
FEMRT_API int femrt_batch_capture(femrt_batch* b);
FEMRT_API int femrt_batch_launch(femrt_batch* b);

/* Launch shapes of a mesh's trace kernels, FEMRT_LAUNCH_CFG_N ints in this
 * order:
 *   [0] first-hit block size (64/128/256),
 *   [1] persistent-walk workgroups per CU (1..64),
 *   [2] polish workgroups per CU (1..64),
 *   [3] tile edge in pixels of the first hit's ray order for camera frames
 *       (4/8/16: square tiles, rays of a tile on neighbouring lanes, and the
 *       hit queue the walk consumes in that order),
 *   [4] walk hit-queue segments per XCD (1/2/4: each XCD's band of the queue
 *       split into that many sub-bands with a counter each).
 * Default {256, 4, 8, 8, 1}.  The shape decides only which lane traces which
 * ray and when: every result is the same under every shape.  get/set take
 * the first n_cfg (1..FEMRT_LAUNCH_CFG_N) values; set leaves the others as
 * they are and rejects a value outside the lists above (nothing changed).
 * femrt_autotune times candidate values on the given trace (reps timed
 * traces per candidate, one value at a time over the kernels it affects),
 * keeps the fastest on the mesh and returns the first n_cfg values in cfg
 * (cfg may be NULL). */
#define FEMRT_LAUNCH_CFG_N 5
FEMRT_API int femrt_mesh_get_launch_cfg(const femrt_mesh* mesh, int* cfg, int n_cfg);
FEMRT_API int femrt_mesh_set_launch_cfg(femrt_mesh* mesh, const int* cfg, int n_cfg);
FEMRT_API int femrt_autotune(femrt_mesh* mesh, femrt_frame* f, const femrt_trace_params* p, int reps,
                             int* cfg, int n_cfg);

/* ------------------------------------------------- learned Newton guess */
/* No counterpart in the reference (its Newton solves start from fixed seeds,
 * HexElement.cu:227-272, high_order.h:206): a small network maps a point in
 * an element to a reference-coordinate guess for point inversion.
 * FEMRT_MLP_GUESS: 27 inputs (the point and the 8 corner vertices, each
 * normalised to the element's box as 2 (x - lo) / (hi - lo) - 1) -> 64 tanh
 * -> 64 tanh -> 3.  BF16 operands, FP32 accumulation, on the matrix cores.
 * Flat FP32 parameters (6147): W1 [28][64] (row 27 = first-layer bias),
 * W2 [64][64] (in, out), b2 [64], W3 [64][3], b3 [3]. */
#define FEMRT_MLP_GUESS 0
/* FEMRT_MLP_BRACKET: the isosurface-bracket network (no counterpart in the
 * reference, whose bracket is the first sign change of 16 uniform samples,
 * fem/RayIsosurface.cu:67-188).  32 inputs -> 64 tanh (bias b1) -> 64 tanh
 * -> 2: the 5% and 95% quantiles of where in the scan's bracket [ta, tb] the
 * crossing lies (s in [0, 1]), trained by the pinball loss.  Inputs: rst
 * where the ray enters and leaves the bracket (at ta and tb), the secant
 * weight g(ta) / (g(ta) - g(tb)) of the FP32 samples there, and the 25 modal
 * coefficients of u - iso with (i, j, k) in {0, 1, 2}^3 other than
 * (2, 2, 2) and (2, 2, 1), scaled to unit maximum.  Flat FP32 parameters
 * (6402): W1 [32][64], b1 [64], W2 [64][64], b2 [64], W3 [64][2], b3 [2]. */
#define FEMRT_MLP_BRACKET 1
typedef struct femrt_mlp femrt_mlp;
typedef struct {
  int n_samples;      /* training pairs generated from the mesh             */
  int batch;          /* multiple of 64                                      */
  int epochs;
  float lr;           /* Adam step, cosine-decayed to lr_final               */
  float lr_final;
  uint64_t seed;
} femrt_mlp_train_params;

FEMRT_API int femrt_mlp_num_params(int kind);
/* Glorot-uniform weights from a keyed generator, zero biases. */
FEMRT_API int femrt_mlp_create(femrt_ctx* ctx, int kind, unsigned long long seed, femrt_mlp** out);
FEMRT_API int femrt_mlp_destroy(femrt_mlp* m);
FEMRT_API int femrt_mlp_get_params(femrt_mlp* m, float* params);
/* Replaces the parameters and resets the optimiser state. */
FEMRT_API int femrt_mlp_set_params(femrt_mlp* m, const float* params);
/* Inference with the tracer's own device routine: feats n x IN, out n x OUT
 * (host arrays, synchronous; IN, OUT = 27, 3 for the guess network and
 * 32, 2 for the bracket network). */
FEMRT_API int femrt_mlp_forward(femrt_mlp* m, int n, const float* feats, float* out);
/* Loss and its gradient with respect to the flat parameters, from the
 * training kernel's forward and backward pass (no update).  Guess network:
 * sum_i sum_c (y_ic - target_ic)^2 / n; bracket network: sum_i sum_c
 * rho_tau_c(target_ic - y_ic) / n, rho_tau(r) = max(tau r, (tau - 1) r),
 * tau = (0.05, 0.95).  feats n x IN, targets n x OUT, grad
 * femrt_mlp_num_params(kind) floats. */
FEMRT_API int femrt_mlp_loss_grad(femrt_mlp* m, int n, const float* feats, const float* targets,
                                  float* loss, float* grad);
/* Training pairs from the tracer's own point inversions: a seeded element and
 * reference point rst0, x = X_e(rst0), target = the FP32 Newton preimage of x
 * inside the element.  feats n x 27, targets n x 3 (host). */
FEMRT_API int femrt_mlp_make_samples(femrt_mesh* mesh, int n, unsigned long long seed, float* feats,
                                     float* targets);
FEMRT_API void femrt_mlp_train_params_default(femrt_mlp_train_params* p);
/* Generates p->n_samples pairs on the device and runs p->epochs of Adam over
 * them (MSE loss); epoch_loss (may be NULL) receives the mean batch loss of
 * each epoch.  Synchronous. */
FEMRT_API int femrt_mlp_train(femrt_mlp* m, femrt_mesh* mesh, const femrt_mlp_train_params* p,
                              float* epoch_loss);
/* Adam over caller-given training pairs (host arrays feats n x IN, targets
 * n x OUT), p->epochs epochs in batches of p->batch (p->n_samples unused);
 * either kind of network.  Synchronous. */
FEMRT_API int femrt_mlp_train_pairs(femrt_mlp* m, int n, const float* feats, const float* targets,
                                    const femrt_mlp_train_params* p, float* epoch_loss);
/* Network used by FEMRT_GUESS_LEARNED traces of this mesh (NULL detaches);
 * it must stay alive while such traces run. */
FEMRT_API int femrt_mesh_set_guess_mlp(femrt_mesh* mesh, const femrt_mlp* m);
/* Bracket network used by FEMRT_TRACE_ISO_LEARNED_BRACKET traces of this mesh
 * (NULL detaches); it must stay alive while such traces run. */
FEMRT_API int femrt_mesh_set_bracket_mlp(femrt_mesh* mesh, const femrt_mlp* m);
/* World-space gradient of the field at every isosurface crossing of the
 * last trace of frame f, from the crossing's FP64 reference coordinates, in
 * FP32: the reference shades an iso hit with the normal +-grad u / |grad u|
 * (raytracer.cu:207-216; the gradient: fem/HexElement.cu:274-330).
 * engine FEMRT_GRAD_MFMA: the crossings binned by element (sorted by element
 * id), each element's field coefficients staged in LDS, the contraction over
 * the r modes on the FP32 matrix cores for 16 crossings at a time, the rest
 * on the vector ALUs; FEMRT_GRAD_VALU: one crossing per lane on the vector
 * ALUs.  out: n_iso x 3 x n floats (host array; NaN where there is no
 * crossing).  Synchronous. */
#define FEMRT_GRAD_MFMA 0
#define FEMRT_GRAD_VALU 1
FEMRT_API int femrt_frame_iso_gradients(femrt_mesh* mesh, femrt_frame* f, int n_iso, int engine, float* out);
/* Training pairs of the bracket network from the last trace of frame f,
 * which must have recorded segments (FEMRT_TRACE_SEG_INFO or
 * FEMRT_TRACE_ISO_LEARNED_BRACKET) with isovalues iso[0 .. n_iso): for each
 * crossing found, the network inputs (feats, 32 floats) and twice the
 * position s of the FP64 root in the scan's bracket (targets, 2 floats).  At
 * most max_pairs pairs (host arrays), their count in *n_pairs. */
FEMRT_API int femrt_frame_bracket_pairs(femrt_mesh* mesh, femrt_frame* f, int n_iso, const double* iso,
                                        int max_pairs, float* feats, float* targets, int* n_pairs);

/* ------------------------------------------------ host-only utilities */
/* The same box and BVH builders the mesh upload uses, on host arrays (no
 * device needed).  boxes: n_elem x 61 x 6 floats (layout of femrt_mesh_boxes).  nodes8 must hold
 * 2 * n_elem nodes; *n_nodes receives the count. */
FEMRT_API int femrt_host_boxes(int n_elem, int ng, const float* geom, float* boxes);
FEMRT_API int femrt_host_build_bvh(int n_elem, const float* boxes, int* n_nodes, float* nodes8,
                                   int* prims);
/* Per element, bounds (lo, hi) of the field over the reference cube, as the
 * walk uses them to skip scans that can cross no open isovalue: the extreme
 * tensor-Bernstein coefficients over the 8 half-cubes, widened by 1e-5 of
 * sum |c|.  out: 2 n_elem. */
FEMRT_API int femrt_host_field_range(int n_elem, int nf, const float* field, float* out);
/* Per element, a bound L of |u(a) - u(b)| / max_c |a_c - b_c| over the
 * reference cube, as the walk uses it to accept a sample's signs after one
 * Newton step: over the 8 half-cubes, the largest sum over r, s, t of the
 * largest |Bernstein coefficient| of the partial derivative, plus 5%.
 * out: n_elem.  (No counterpart in the
 * reference, whose scan converges every sample: fem/RayIsosurface.cu:57.) */
FEMRT_API int femrt_host_field_lipschitz(int n_elem, int nf, const float* field, float* out);
/* Per element, K >= |J^-1|_inf (max row sum, J = dx/drst) everywhere in the
 * reference cube, or 0 when none is found: interval bounds of J on the 8
 * half-cubes (Bernstein coefficients of the geometry's partial derivatives)
 * and a Neumann-series bound of the interval matrix's inverse, plus 5%.  With
 * the field bound it certifies the signs of scan samples the walk then skips.
 * out: n_elem.  (No counterpart in the reference.) */
FEMRT_API int femrt_host_geom_jinv_bound(int n_elem, int ng, const float* geom, float* out);

/* ------------------------------------------------------- synthetic meshes */
typedef struct {
  int n;          /* elements per axis on the unit cube                     */
  int q;          /* geometry order (ng = q + 1)                            */
  int p;          /* field order (nf = p + 1)                               */
  int pert;       /* 0 none; 1 A sin(2 pi a) sin(2 pi b) of the node's two
                     in-face world coordinates; 2 seeded bubble per face
                     A h xi_f (1-u^2)(1-v^2), xi_f ~ N(0,1) clipped to +-2,
                     (u, v) the node's face coordinates; 3 the same bubble
                     with xi_f uniform on (-1, 1) (peak amp h)               */
  double amp;
  uint64_t seed;
} femrt_gen_params;

/* Structured n^3 hex mesh of [0,1]^3: straight edges, interior face nodes
 * (evenly spaced, as PolyEvenlySpaced CURVED faces) displaced along the face
 * normal, geometry fitted on GLL nodes exactly as BuildHexGeometry does;
 * field c_ijk = N(0,1)/(1+i+j+k)^2 from a counter-based generator keyed by
 * (seed, element, mode).  Output arrays sized as for femrt_mesh_upload_hex. */
FEMRT_API int femrt_gen_structured_hex(const femrt_gen_params* gp, float* vertices, float* geom,
                                       float* field, int* adjacency);

/* Seeded multi-frequency field on a given geometry (n_elem x 3 x ng^3 modal
 * coefficients): u(x) = sum of n_waves (1..64) plane waves sin(k.x + phi)
 * with 1, 2, 4, 8 periods over the unit cube in turn (times a seeded factor
 * in [1, 1.5)), amplitude 1/frequency, seeded directions and phases,
 * normalised to unit RMS; sampled at each element's nf^3 GLL nodes and
 * fitted to nf modes per direction.  field: n_elem x nf^3.  The reference
 * reads fields from FLD files only (xml_fld_field_parser.cpp:165-289); this
 * generator provides smooth fields with isosurfaces that span many elements
 * for the same loader-shaped arrays. */
FEMRT_API int femrt_gen_field_waves(int n_elem, int ng, const float* geom, int nf, int n_waves,
                                    uint64_t seed, float* field);

/* Drop-in loader: Nektar++ XML mesh + FLD field (xml_fld_loader.h:14-16).
 * Two-phase: call with NULL arrays to get sizes, then with arrays.  fld_path
 * may be NULL or empty (field modes then come from the XML EXPANSIONS and the
 * field is zero).  Hex elements only; uniform field order.
 * flags bit0: replicate the reference's vertex-0 parse defect
 * (xml_fld_geometry_parser.cpp:54-77); bit1: replicate geomModes inflation
 * to max(field, geometry) modes (xml_fld_loader.cpp:252-254). */
FEMRT_API int femrt_load_xml_fld(const char* xml_path, const char* fld_path, int flags,
                                 int* n_elem, int* ng, int* nf, float* vertices, float* geom,
                                 float* field, int* adjacency);
/* Writer of the same formats (the reference has none; Nektar++ itself writes
 * FLD as Hex_CurvedFace.fld:12 shows): vertices and edges shared by exact
 * position, every face CURVED with ng^2 evenly spaced points of the modal
 * geometry, one FLD ELEMENTS block B64Z-LittleEndian 64-bit,
 * NUMMODESPERDIR="UNIORDER:nf,nf,nf". */
FEMRT_API int femrt_save_xml_fld(const char* xml_path, const char* fld_path, int n_elem, int ng,
                                 int nf, const float* vertices, const float* geom,
                                 const float* field);
/* Message of the last failed load/save call on this thread. */
FEMRT_API const char* femrt_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
