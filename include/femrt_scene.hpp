// femrt_scene.hpp -- C++ host layer with the reference's entry points for the
// hex path, above the C ABI of femrt.h.
//
// Mirrors FEM-AABB/raytracer.h (Camera :22-49, Scene :53-309 hex subset,
// save_image_ppm :334) and FEM-AABB/xml_fld_loader.h:14-16 (load_xml_fld),
// with the same names, argument meaning and error behaviour (bool returns,
// messages on stderr, no exceptions).  What differs by design: device memory
// belongs to C-ABI handles owned by the Scene (never raw pointers patched
// into host structs, main.cu:362-420), and rendering is a host call,
// render(), instead of the render_kernel launch (raytracer.h:313-331).
// Vector types (float3, uint3, uchar4) are HIP's host-side definitions.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "femrt.h"

#define FEMRT_CXX_API __attribute__((visibility("default")))

namespace zidingyi {
// Hex element record (FEM-AABB/fem/HexElement.cuh:8-21): corner vertices in
// XML/VTK order, modal field and geometry coefficients with their mode counts.
struct HexElementData {
  float3 vertices[8];
  const float* fieldCoefficients;
  uint3 fieldModes;
  const float* geomCoefficients[3];
  uint3 geomModes;
  float3 aabbMin;
  float3 aabbMax;
};

// Prism record (FEM-AABB/fem/PrismElement.cuh:9-18): vertices v0..v3 the
// t = -1 quad, v4 v5 the collapsed t = +1 edge; coefficients in the order
// i (r) slowest, j (s), k (t) < modes.z - i fastest.
struct PrismElementData {
  float3 vertices[6];
  const float* fieldCoefficients{nullptr};
  uint3 fieldModes{0, 0, 0};
  const float* geomCoefficients[3]{nullptr, nullptr, nullptr};
  uint3 geomModes{0, 0, 0};
};

// Curved triangle record (FEM-AABB/fem/TriElement.cuh:9-17): coefficients
// in the order i (r) slowest, j (s) < modes.y - i fastest.
struct TriElementData {
  float3 vertices[3];
  const float* fieldCoefficients{nullptr};
  uint2 fieldModes{0, 0};
  const float* geomCoefficients[3]{nullptr, nullptr, nullptr};
  uint2 geomModes{0, 0};
};

// Quad surface element record (FEM-AABB/fem/QuadElement.cuh:9-18): corner
// vertices (corner i at r = +-1 by bit 0, s = +-1 by bit 1), modal field and
// geometry coefficients (mode (i, j) at j * modes.x + i).
struct QuadElementData {
  float3 vertices[4];
  const float* fieldCoefficients{nullptr};
  uint2 fieldModes{0, 0};
  const float* geomCoefficients[3]{nullptr, nullptr, nullptr};
  uint2 geomModes{0, 0};
};
}  // namespace zidingyi

// Pinhole camera (raytracer.h:22-49); the lens is ignored (the reference
// renders with aperture 0).
struct FEMRT_CXX_API Camera {
  float3 origin;
  float3 lower_left_corner;
  float3 horizontal;
  float3 vertical;
  float3 u, v, w;
  float lens_radius;

  Camera() {}
  Camera(float3 lookfrom, float3 lookat, float3 vup, float vfov, float aspect_ratio, float aperture,
         float focus_dist);
};

// Per-pixel results that the reference's HitRecord does not carry: the hit
// element, face, distance, reference coordinates and field value, and the
// first crossing of each isovalue (element -1 when none).
struct HitAOV {
  int elem, face;
  double t, rst[3], u;
  int iso_elem[FEMRT_MAX_ISO];
  double iso_t[FEMRT_MAX_ISO];
};

struct FEMRT_CXX_API Scene {
  struct FieldSlice {
    std::string name;
    bool is3D{true};
    uint3 modes3{0, 0, 0};
    uint2 modes2{0, 0};
    size_t count{0};
    int offset{-1};
  };

  // pooled host copies, as the reference keeps them (raytracer.h:118-166)
  std::vector<zidingyi::HexElementData> hexes;
  std::vector<float> hex_coefficients;
  std::vector<int> hex_coeff_offsets;
  std::vector<float> hex_geom_coefficients;
  std::vector<int> hex_geom_offsets;
  std::vector<std::vector<FieldSlice>> hex_field_slices;
  std::vector<zidingyi::PrismElementData> prisms;
  std::vector<float> prism_coefficients;
  std::vector<int> prism_coeff_offsets;
  std::vector<float> prism_geom_coefficients;
  std::vector<int> prism_geom_offsets;
  std::vector<std::vector<FieldSlice>> prism_field_slices;
  std::vector<zidingyi::TriElementData> curved_tris;
  std::vector<float> tri_coefficients;
  std::vector<int> tri_coeff_offsets;
  std::vector<float> tri_geom_coefficients;
  std::vector<int> tri_geom_offsets;
  std::vector<std::vector<FieldSlice>> tri_field_slices;
  std::vector<zidingyi::QuadElementData> quads;
  std::vector<float> quad_coefficients;
  std::vector<int> quad_coeff_offsets;
  std::vector<float> quad_geom_coefficients;
  std::vector<int> quad_geom_offsets;
  std::vector<std::vector<FieldSlice>> quad_field_slices;

  // materials (raytracer.h Material subset as femrt_material); empty means
  // the reference's defaults, grey and blue Lambertian (main.cu:39-41)
  std::vector<femrt_material> materials;
  void add_material(const femrt_material& m) { materials.push_back(m); }

  Camera camera;
  int width{0}, height{0};
  // samples_per_pixel > 1 or max_depth > 1 selects path tracing in render()
  int samples_per_pixel{1};
  int max_depth{1};
  float iso_value{0.0f};
  bool use_iso{true};
  // further isovalues searched after iso_value (up to FEMRT_MAX_ISO in all)
  std::vector<double> extra_iso_values;

  Scene() = default;
  Scene(const Scene&) = delete;
  Scene& operator=(const Scene&) = delete;
  ~Scene();

  void add_hex(const zidingyi::HexElementData& h, const float* fieldCoeffs = nullptr,
               uint3 fieldModes = {0, 0, 0}, size_t field_count = 0,
               const float* geomX = nullptr, const float* geomY = nullptr,
               const float* geomZ = nullptr, uint3 geomModes = {0, 0, 0}, size_t geom_count = 0,
               const std::vector<FieldSlice>* slices = nullptr);

  // Prism (raytracer.h:167-212): pooled like add_hex.  Without geometry
  // coefficients the prism is the reference's linear map of its vertices.
  // A prism scene renders first hits only, as the reference's does (its
  // prism isosurface branch is off, raytracer.cu:266).
  void add_prism(const zidingyi::PrismElementData& p, const float* fieldCoeffs = nullptr,
                 uint3 fieldModes = {0, 0, 0}, size_t field_count = 0, const float* geomX = nullptr,
                 const float* geomY = nullptr, const float* geomZ = nullptr, uint3 geomModes = {0, 0, 0},
                 size_t geom_count = 0, const std::vector<FieldSlice>* slices = nullptr);

  // Curved quad surface element (raytracer.h:213-257): pooled like add_hex.
  // Without geometry coefficients the quad is bilinear in its corners.  A
  // scene holds one element type; a quad scene renders first hits only
  // (use_iso is ignored).
  void add_quad(const zidingyi::QuadElementData& q, const float* fieldCoeffs = nullptr,
                uint2 fieldModes = {0, 0}, size_t field_count = 0, const float* geomX = nullptr,
                const float* geomY = nullptr, const float* geomZ = nullptr, uint2 geomModes = {0, 0},
                size_t geom_count = 0, const std::vector<FieldSlice>* slices = nullptr);

  // Curved triangle (raytracer.h:259-304): pooled like add_hex.  Without
  // geometry coefficients the triangle is flat.  A triangle scene renders
  // first hits only.
  void add_curved_tri(const zidingyi::TriElementData& tri, const float* fieldCoeffs = nullptr,
                      uint2 fieldModes = {0, 0}, size_t field_count = 0, const float* geomX = nullptr,
                      const float* geomY = nullptr, const float* geomZ = nullptr, uint2 geomModes = {0, 0},
                      size_t geom_count = 0, const std::vector<FieldSlice>* slices = nullptr);

  // Uploads the hexes (uniform mode counts), or the prisms, quads or
  // triangles of a scene of that element type, to the GPU and builds the
  // BVH.  On failure prints to stderr and leaves
  // the scene unbuilt (render then returns false).
  void build_bvh();

  // device objects (owned)
  femrt_ctx* ctx{nullptr};
  femrt_mesh* mesh{nullptr};
  femrt_mlp* guess_mlp{nullptr};
  int device{0};
};

struct RenderParams {
  int guess{FEMRT_GUESS_WARM};   // FEMRT_GUESS_*; LEARNED needs Scene::guess_mlp
  int samples{16};               // iso samples per element segment
  int max_march{4096};
  unsigned long long seed{1};    // path tracing: random stream
};

// Drop-in of load_xml_fld (xml_fld_loader.h:14-16): appends the hexes of a
// Nektar++ XML mesh with its FLD field to the scene.
FEMRT_CXX_API bool load_xml_fld(const std::string& xml_path, const std::string& fld_path, Scene& scene);

// Traces one pixel-centre ray per pixel (y = 0 is the bottom row, as the
// reference's image buffer): geometry first hit, the scene's isovalues,
// shading into out (width * height).  aov (may be null) receives the
// per-pixel results.  With scene.samples_per_pixel > 1 or scene.max_depth > 1
// it path-traces the mesh surface instead (jittered samples, material
// scatter, sky; render_kernel, raytracer.cu:410-476) and aov must be null.
// Synchronous.
FEMRT_CXX_API bool render(Scene& scene, const Camera& cam, int width, int height, std::vector<uchar4>& out,
            std::vector<HitAOV>* aov = nullptr, const RenderParams* params = nullptr);

// The rays render() traces: pixel centres (x + 0.5, y + 0.5) of the camera
// (raytracer.cu:431-443 without the random jitter), unit directions; 6 x
// (width * height) SoA floats (ox, oy, oz, dx, dy, dz).
FEMRT_CXX_API void camera_rays(const Camera& cam, int width, int height, std::vector<float>& rays);

// P6 writer (raytracer.cu:480-494).
FEMRT_CXX_API void save_image_ppm(const std::vector<uchar4>& image, int width, int height,
                    const std::string& filename);
