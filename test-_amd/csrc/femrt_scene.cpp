// femrt_scene.cpp -- the C++ host layer of include/femrt_scene.hpp: Scene,
// Camera, load_xml_fld, render and save_image_ppm over the C ABI.
#include "../../include/femrt_scene.hpp"

#include <algorithm>
#include <array>
#include <cstdio>
#include <fstream>
#include <iostream>

#include "femrt_internal.hpp"

namespace {

float3 sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
float3 scale(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
float3 cross3(float3 a, float3 b)
{
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
float3 unit(float3 a)
{
  const float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return l > 0.f ? scale(a, 1.f / l) : a;
}

}  // namespace

Camera::Camera(float3 lookfrom, float3 lookat, float3 vup, float vfov, float aspect_ratio,
               float aperture, float focus_dist)
{
  const float theta = vfov * 3.14159265f / 180.0f;
  const float h = std::tan(theta / 2.0f);
  const float vh = 2.0f * h, vw = aspect_ratio * vh;
  w = unit(sub(lookfrom, lookat));
  u = unit(cross3(vup, w));
  v = cross3(w, u);
  origin = lookfrom;
  horizontal = scale(u, focus_dist * vw);
  vertical = scale(v, focus_dist * vh);
  lower_left_corner =
      sub(sub(sub(origin, scale(horizontal, 0.5f)), scale(vertical, 0.5f)), scale(w, focus_dist));
  lens_radius = aperture / 2.0f;
}

Scene::~Scene()
{
  if (mesh) femrt_mesh_destroy(mesh);
  if (ctx) femrt_ctx_destroy(ctx);
}

void Scene::add_hex(const zidingyi::HexElementData& h, const float* fieldCoeffs, uint3 fieldModes,
                    size_t field_count, const float* geomX, const float* geomY, const float* geomZ,
                    uint3 geomModes, size_t geom_count, const std::vector<FieldSlice>* slices)
{
  zidingyi::HexElementData e = h;
  if (fieldCoeffs && field_count > 0) {
    hex_coeff_offsets.push_back((int)hex_coefficients.size());
    hex_coefficients.insert(hex_coefficients.end(), fieldCoeffs, fieldCoeffs + field_count);
    if (fieldModes.x > 0) e.fieldModes = fieldModes;
  } else {
    hex_coeff_offsets.push_back(-1);
  }
  if (geomX && geomY && geomZ && geom_count > 0) {
    hex_geom_offsets.push_back((int)hex_geom_coefficients.size());
    hex_geom_coefficients.insert(hex_geom_coefficients.end(), geomX, geomX + geom_count);
    hex_geom_coefficients.insert(hex_geom_coefficients.end(), geomY, geomY + geom_count);
    hex_geom_coefficients.insert(hex_geom_coefficients.end(), geomZ, geomZ + geom_count);
    if (geomModes.x > 0) e.geomModes = geomModes;
  } else {
    hex_geom_offsets.push_back(-1);
  }
  e.fieldCoefficients = nullptr;
  e.geomCoefficients[0] = e.geomCoefficients[1] = e.geomCoefficients[2] = nullptr;
  hexes.push_back(e);
  if (slices)
    hex_field_slices.push_back(*slices);
  else
    hex_field_slices.emplace_back();
  if (mesh) {  // the device copy is stale now
    femrt_mesh_destroy(mesh);
    mesh = nullptr;
  }
}

void Scene::add_quad(const zidingyi::QuadElementData& q, const float* fieldCoeffs, uint2 fieldModes,
                     size_t field_count, const float* geomX, const float* geomY, const float* geomZ,
                     uint2 geomModes, size_t geom_count, const std::vector<FieldSlice>* slices)
{
  zidingyi::QuadElementData e = q;
  if (fieldCoeffs && field_count > 0) {
    quad_coeff_offsets.push_back((int)quad_coefficients.size());
    quad_coefficients.insert(quad_coefficients.end(), fieldCoeffs, fieldCoeffs + field_count);
    if (fieldModes.x > 0) e.fieldModes = fieldModes;
  } else {
    quad_coeff_offsets.push_back(-1);
  }
  if (geomX && geomY && geomZ && geom_count > 0) {
    quad_geom_offsets.push_back((int)quad_geom_coefficients.size());
    quad_geom_coefficients.insert(quad_geom_coefficients.end(), geomX, geomX + geom_count);
    quad_geom_coefficients.insert(quad_geom_coefficients.end(), geomY, geomY + geom_count);
    quad_geom_coefficients.insert(quad_geom_coefficients.end(), geomZ, geomZ + geom_count);
    if (geomModes.x > 0) e.geomModes = geomModes;
  } else {
    quad_geom_offsets.push_back(-1);
  }
  e.fieldCoefficients = nullptr;
  e.geomCoefficients[0] = e.geomCoefficients[1] = e.geomCoefficients[2] = nullptr;
  quads.push_back(e);
  quad_field_slices.push_back(slices ? *slices : std::vector<FieldSlice>{});
  if (mesh) {
    femrt_mesh_destroy(mesh);
    mesh = nullptr;
  }
}

void Scene::add_prism(const zidingyi::PrismElementData& p, const float* fieldCoeffs, uint3 fieldModes,
                      size_t field_count, const float* geomX, const float* geomY, const float* geomZ,
                      uint3 geomModes, size_t geom_count, const std::vector<FieldSlice>* slices)
{
  zidingyi::PrismElementData e = p;
  if (fieldCoeffs && field_count > 0) {
    prism_coeff_offsets.push_back((int)prism_coefficients.size());
    prism_coefficients.insert(prism_coefficients.end(), fieldCoeffs, fieldCoeffs + field_count);
    if (fieldModes.x > 0) e.fieldModes = fieldModes;
  } else {
    prism_coeff_offsets.push_back(-1);
  }
  if (geomX && geomY && geomZ && geom_count > 0) {
    prism_geom_offsets.push_back((int)prism_geom_coefficients.size());
    prism_geom_coefficients.insert(prism_geom_coefficients.end(), geomX, geomX + geom_count);
    prism_geom_coefficients.insert(prism_geom_coefficients.end(), geomY, geomY + geom_count);
    prism_geom_coefficients.insert(prism_geom_coefficients.end(), geomZ, geomZ + geom_count);
    if (geomModes.x > 0) e.geomModes = geomModes;
  } else {
    prism_geom_offsets.push_back(-1);
  }
  e.fieldCoefficients = nullptr;
  e.geomCoefficients[0] = e.geomCoefficients[1] = e.geomCoefficients[2] = nullptr;
  prisms.push_back(e);
  prism_field_slices.push_back(slices ? *slices : std::vector<FieldSlice>{});
  if (mesh) {
    femrt_mesh_destroy(mesh);
    mesh = nullptr;
  }
}

// Prism scene upload: uniform mode counts; prisms without geometry
// coefficients take the linear map of their vertices, which needs every
// prism of the scene to be so.
static bool build_prisms(Scene& S)
{
  const int n = (int)S.prisms.size();
  const uint3 fm = S.prisms[0].fieldModes, gm = S.prisms[0].geomModes;
  const bool curved = S.prism_geom_offsets[0] >= 0;
  bool has_field = false;
  for (int e = 0; e < n; ++e) has_field |= S.prism_coeff_offsets[e] >= 0;
  const int gmi[3] = {(int)gm.x, (int)gm.y, (int)gm.z}, fmi[3] = {(int)fm.x, (int)fm.y, (int)fm.z};
  int gc = 0, fc = 0;
  for (int i = 0; i < gmi[0]; ++i) gc += gmi[1] * std::max(gmi[2] - i, 0);
  for (int i = 0; i < fmi[0]; ++i) fc += fmi[1] * std::max(fmi[2] - i, 0);
  std::vector<float> verts((size_t)n * 18), geom(curved ? (size_t)n * 3 * gc : 0), field((size_t)n * fc, 0.f);
  for (int e = 0; e < n; ++e) {
    const auto& p = S.prisms[e];
    const bool gsame = p.geomModes.x == gm.x && p.geomModes.y == gm.y && p.geomModes.z == gm.z;
    const bool fsame = p.fieldModes.x == fm.x && p.fieldModes.y == fm.y && p.fieldModes.z == fm.z;
    if ((S.prism_geom_offsets[e] >= 0) != curved || (curved && !gsame) || (S.prism_coeff_offsets[e] >= 0 && !fsame)) {
      std::cerr << "Scene::build_bvh: prisms of different orders (upload them as separate scenes)" << std::endl;
      return false;
    }
    for (int c = 0; c < 6; ++c) {
      verts[(size_t)e * 18 + 3 * c] = p.vertices[c].x;
      verts[(size_t)e * 18 + 3 * c + 1] = p.vertices[c].y;
      verts[(size_t)e * 18 + 3 * c + 2] = p.vertices[c].z;
    }
    if (curved) std::copy_n(&S.prism_geom_coefficients[S.prism_geom_offsets[e]], 3 * gc, &geom[(size_t)e * 3 * gc]);
    if (S.prism_coeff_offsets[e] >= 0) std::copy_n(&S.prism_coefficients[S.prism_coeff_offsets[e]], fc, &field[(size_t)e * fc]);
  }
  if (!S.ctx && femrt_ctx_create(S.device, &S.ctx) != FEMRT_OK) {
    std::cerr << "Scene::build_bvh: no GPU context" << std::endl;
    S.ctx = nullptr;
    return false;
  }
  if (femrt_mesh_upload_prisms(S.ctx, n, curved ? gmi : nullptr, has_field ? fmi : nullptr, verts.data(),
                               curved ? geom.data() : nullptr, has_field ? field.data() : nullptr,
                               &S.mesh) != FEMRT_OK ||
      femrt_mesh_build_bvh(S.mesh, 0, nullptr) != FEMRT_OK) {
    std::cerr << "Scene::build_bvh: " << femrt_last_error(S.ctx) << std::endl;
    if (S.mesh) femrt_mesh_destroy(S.mesh);
    S.mesh = nullptr;
    return false;
  }
  return true;
}

void Scene::add_curved_tri(const zidingyi::TriElementData& tri, const float* fieldCoeffs, uint2 fieldModes,
                           size_t field_count, const float* geomX, const float* geomY, const float* geomZ,
                           uint2 geomModes, size_t geom_count, const std::vector<FieldSlice>* slices)
{
  zidingyi::TriElementData e = tri;
  if (fieldCoeffs && field_count > 0) {
    tri_coeff_offsets.push_back((int)tri_coefficients.size());
    tri_coefficients.insert(tri_coefficients.end(), fieldCoeffs, fieldCoeffs + field_count);
    if (fieldModes.x > 0) e.fieldModes = fieldModes;
  } else {
    tri_coeff_offsets.push_back(-1);
  }
  if (geomX && geomY && geomZ && geom_count > 0) {
    tri_geom_offsets.push_back((int)tri_geom_coefficients.size());
    tri_geom_coefficients.insert(tri_geom_coefficients.end(), geomX, geomX + geom_count);
    tri_geom_coefficients.insert(tri_geom_coefficients.end(), geomY, geomY + geom_count);
    tri_geom_coefficients.insert(tri_geom_coefficients.end(), geomZ, geomZ + geom_count);
    if (geomModes.x > 0) e.geomModes = geomModes;
  } else {
    tri_geom_offsets.push_back(-1);
  }
  e.fieldCoefficients = nullptr;
  e.geomCoefficients[0] = e.geomCoefficients[1] = e.geomCoefficients[2] = nullptr;
  curved_tris.push_back(e);
  tri_field_slices.push_back(slices ? *slices : std::vector<FieldSlice>{});
  if (mesh) {
    femrt_mesh_destroy(mesh);
    mesh = nullptr;
  }
}

// Triangle scene upload: uniform mode counts; flat triangles (no geometry
// coefficients) need every triangle of the scene to be flat.
static bool build_tris(Scene& S)
{
  const int n = (int)S.curved_tris.size();
  const uint2 fm = S.curved_tris[0].fieldModes, gm = S.curved_tris[0].geomModes;
  const bool curved = S.tri_geom_offsets[0] >= 0;
  bool has_field = false;
  for (int e = 0; e < n; ++e) has_field |= S.tri_coeff_offsets[e] >= 0;
  const int gmi[2] = {(int)gm.x, (int)gm.y}, fmi[2] = {(int)fm.x, (int)fm.y};
  int gc = 0, fc = 0;
  for (int i = 0; i < gmi[0]; ++i) gc += std::max(gmi[1] - i, 0);
  for (int i = 0; i < fmi[0]; ++i) fc += std::max(fmi[1] - i, 0);
  std::vector<float> verts((size_t)n * 9), geom(curved ? (size_t)n * 3 * gc : 0), field((size_t)n * fc, 0.f);
  for (int e = 0; e < n; ++e) {
    const auto& t = S.curved_tris[e];
    const bool gsame = t.geomModes.x == gm.x && t.geomModes.y == gm.y;
    const bool fsame = t.fieldModes.x == fm.x && t.fieldModes.y == fm.y;
    if ((S.tri_geom_offsets[e] >= 0) != curved || (curved && !gsame) || (S.tri_coeff_offsets[e] >= 0 && !fsame)) {
      std::cerr << "Scene::build_bvh: triangles of different orders (upload them as separate scenes)" << std::endl;
      return false;
    }
    for (int c = 0; c < 3; ++c) {
      verts[(size_t)e * 9 + 3 * c] = t.vertices[c].x;
      verts[(size_t)e * 9 + 3 * c + 1] = t.vertices[c].y;
      verts[(size_t)e * 9 + 3 * c + 2] = t.vertices[c].z;
    }
    if (curved) std::copy_n(&S.tri_geom_coefficients[S.tri_geom_offsets[e]], 3 * gc, &geom[(size_t)e * 3 * gc]);
    if (S.tri_coeff_offsets[e] >= 0) std::copy_n(&S.tri_coefficients[S.tri_coeff_offsets[e]], fc, &field[(size_t)e * fc]);
  }
  if (!S.ctx && femrt_ctx_create(S.device, &S.ctx) != FEMRT_OK) {
    std::cerr << "Scene::build_bvh: no GPU context" << std::endl;
    S.ctx = nullptr;
    return false;
  }
  if (femrt_mesh_upload_tris(S.ctx, n, curved ? gmi : nullptr, has_field ? fmi : nullptr, verts.data(),
                             curved ? geom.data() : nullptr, has_field ? field.data() : nullptr, &S.mesh) != FEMRT_OK ||
      femrt_mesh_build_bvh(S.mesh, 0, nullptr) != FEMRT_OK) {
    std::cerr << "Scene::build_bvh: " << femrt_last_error(S.ctx) << std::endl;
    if (S.mesh) femrt_mesh_destroy(S.mesh);
    S.mesh = nullptr;
    return false;
  }
  return true;
}

// Quad scene upload: uniform mode counts; a quad without geometry
// coefficients takes the bilinear map of its corners (order 1, two modes),
// which needs every quad of the scene to be so.
static bool build_quads(Scene& S)
{
  const int n = (int)S.quads.size();
  const uint2 fm = S.quads[0].fieldModes, gm = S.quads[0].geomModes;
  const bool curved = S.quad_geom_offsets[0] >= 0;
  if (fm.x != fm.y || (curved && gm.x != gm.y)) {
    std::cerr << "Scene::build_bvh: anisotropic mode counts are not supported" << std::endl;
    return false;
  }
  const int nf = fm.x > 0 ? (int)fm.x : 1, ng = curved ? (int)gm.x : 2;
  const size_t f2 = (size_t)nf * nf, g2 = (size_t)ng * ng;
  std::vector<float> verts((size_t)n * 12), geom(curved ? (size_t)n * 3 * g2 : 0), field((size_t)n * f2, 0.f);
  for (int e = 0; e < n; ++e) {
    const auto& q = S.quads[e];
    if ((S.quad_geom_offsets[e] >= 0) != curved || (curved && (q.geomModes.x != gm.x || q.geomModes.y != gm.y)) ||
        (S.quad_coeff_offsets[e] >= 0 && (q.fieldModes.x != fm.x || q.fieldModes.y != fm.y))) {
      std::cerr << "Scene::build_bvh: quads of different orders (upload them as separate scenes)" << std::endl;
      return false;
    }
    for (int c = 0; c < 4; ++c) {
      verts[(size_t)e * 12 + 3 * c] = q.vertices[c].x;
      verts[(size_t)e * 12 + 3 * c + 1] = q.vertices[c].y;
      verts[(size_t)e * 12 + 3 * c + 2] = q.vertices[c].z;
    }
    if (curved) std::copy_n(&S.quad_geom_coefficients[S.quad_geom_offsets[e]], 3 * g2, &geom[(size_t)e * 3 * g2]);
    if (S.quad_coeff_offsets[e] >= 0)
      std::copy_n(&S.quad_coefficients[S.quad_coeff_offsets[e]], f2, &field[(size_t)e * f2]);
  }
  if (!S.ctx && femrt_ctx_create(S.device, &S.ctx) != FEMRT_OK) {
    std::cerr << "Scene::build_bvh: no GPU context" << std::endl;
    S.ctx = nullptr;
    return false;
  }
  if (femrt_mesh_upload_quads(S.ctx, n, ng, nf, verts.data(), curved ? geom.data() : nullptr, field.data(),
                              &S.mesh) != FEMRT_OK ||
      femrt_mesh_build_bvh(S.mesh, 0, nullptr) != FEMRT_OK) {
    std::cerr << "Scene::build_bvh: " << femrt_last_error(S.ctx) << std::endl;
    if (S.mesh) femrt_mesh_destroy(S.mesh);
    S.mesh = nullptr;
    return false;
  }
  return true;
}

void Scene::build_bvh()
{
  if (mesh) {
    femrt_mesh_destroy(mesh);
    mesh = nullptr;
  }
  const int n = (int)hexes.size();
  if ((n > 0) + !quads.empty() + !prisms.empty() + !curved_tris.empty() > 1) {
    std::cerr << "Scene::build_bvh: one element type per scene (hexahedra, prisms, quads or triangles)"
              << std::endl;
    return;
  }
  if (!curved_tris.empty()) {
    build_tris(*this);
    return;
  }
  if (!quads.empty()) {
    build_quads(*this);
    return;
  }
  if (!prisms.empty()) {
    build_prisms(*this);
    return;
  }
  if (n == 0) {
    std::cerr << "Scene::build_bvh: no hexahedra" << std::endl;
    return;
  }
  const uint3 fm = hexes[0].fieldModes, gm = hexes[0].geomModes;
  if (fm.x != fm.y || fm.y != fm.z || gm.x != gm.y || gm.y != gm.z) {
    std::cerr << "Scene::build_bvh: anisotropic mode counts are not supported" << std::endl;
    return;
  }
  const int nf = (int)fm.x, ng = (int)gm.x;
  const size_t f3 = (size_t)nf * nf * nf, g3 = (size_t)ng * ng * ng;
  std::vector<float> verts((size_t)n * 24), geom((size_t)n * 3 * g3), field((size_t)n * f3, 0.f);
  for (int e = 0; e < n; ++e) {
    const auto& h = hexes[e];
    if (h.fieldModes.x != fm.x || h.geomModes.x != gm.x || h.fieldModes.y != fm.y ||
        h.fieldModes.z != fm.z || h.geomModes.y != gm.y || h.geomModes.z != gm.z) {
      std::cerr << "Scene::build_bvh: elements of different orders (upload them as separate scenes)"
                << std::endl;
      return;
    }
    if (hex_geom_offsets[e] < 0) {
      std::cerr << "Scene::build_bvh: hex " << e << " has no geometry coefficients" << std::endl;
      return;
    }
    for (int c = 0; c < 8; ++c) {
      verts[(size_t)e * 24 + 3 * c] = h.vertices[c].x;
      verts[(size_t)e * 24 + 3 * c + 1] = h.vertices[c].y;
      verts[(size_t)e * 24 + 3 * c + 2] = h.vertices[c].z;
    }
    std::copy_n(&hex_geom_coefficients[hex_geom_offsets[e]], 3 * g3, &geom[(size_t)e * 3 * g3]);
    if (hex_coeff_offsets[e] >= 0)
      std::copy_n(&hex_coefficients[hex_coeff_offsets[e]], f3, &field[(size_t)e * f3]);
  }
  // face adjacency from shared corner quadruples
  std::vector<int> adj((size_t)n * 12, -1);
  {
    static const int fc[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                 {3, 2, 6, 7}, {0, 3, 7, 4}, {4, 5, 6, 7}};
    struct Key {
      float v[12];
      bool operator<(const Key& o) const { return std::lexicographical_compare(v, v + 12, o.v, o.v + 12); }
    };
    std::vector<std::pair<Key, int>> keys;
    keys.reserve((size_t)n * 6);
    for (int e = 0; e < n; ++e)
      for (int f = 0; f < 6; ++f) {
        std::array<std::array<float, 3>, 4> p;
        for (int c = 0; c < 4; ++c)
          for (int k = 0; k < 3; ++k) p[c][k] = verts[(size_t)e * 24 + 3 * fc[f][c] + k];
        std::sort(p.begin(), p.end());
        Key K;
        for (int c = 0; c < 4; ++c)
          for (int k = 0; k < 3; ++k) K.v[3 * c + k] = p[c][k];
        keys.push_back({K, e * 6 + f});
      }
    std::sort(keys.begin(), keys.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t i = 0; i + 1 < keys.size(); ++i)
      if (!(keys[i].first < keys[i + 1].first) && !(keys[i + 1].first < keys[i].first)) {
        const int a = keys[i].second, b = keys[i + 1].second;
        adj[(size_t)a * 2] = b / 6;
        adj[(size_t)a * 2 + 1] = b % 6;
        adj[(size_t)b * 2] = a / 6;
        adj[(size_t)b * 2 + 1] = a % 6;
      }
  }
  if (!ctx && femrt_ctx_create(device, &ctx) != FEMRT_OK) {
    std::cerr << "Scene::build_bvh: no GPU context" << std::endl;
    ctx = nullptr;
    return;
  }
  if (femrt_mesh_upload_hex(ctx, n, ng, nf, verts.data(), geom.data(), field.data(), adj.data(), &mesh) !=
          FEMRT_OK ||
      femrt_mesh_build_bvh(mesh, 0, nullptr) != FEMRT_OK) {
    std::cerr << "Scene::build_bvh: " << femrt_last_error(ctx) << std::endl;
    if (mesh) femrt_mesh_destroy(mesh);
    mesh = nullptr;
    return;
  }
  if (guess_mlp) femrt_mesh_set_guess_mlp(mesh, guess_mlp);
}

bool load_xml_fld(const std::string& xml_path, const std::string& fld_path, Scene& scene)
{
  int n = 0, ng = 0, nf = 0;
  const char* fld = fld_path.empty() ? nullptr : fld_path.c_str();
  if (femrt_load_xml_fld(xml_path.c_str(), fld, 0, &n, &ng, &nf, nullptr, nullptr, nullptr, nullptr) !=
      FEMRT_OK) {
    std::cerr << "XML/FLD loader: " << femrt_io_last_error() << std::endl;
    return false;
  }
  const size_t f3 = (size_t)nf * nf * nf, g3 = (size_t)ng * ng * ng;
  std::vector<float> verts((size_t)n * 24), geom((size_t)n * 3 * g3), field((size_t)n * f3);
  std::vector<int> adj((size_t)n * 12);
  if (femrt_load_xml_fld(xml_path.c_str(), fld, 0, &n, &ng, &nf, verts.data(), geom.data(),
                         field.data(), adj.data()) != FEMRT_OK) {
    std::cerr << "XML/FLD loader: " << femrt_io_last_error() << std::endl;
    return false;
  }
  for (int e = 0; e < n; ++e) {
    zidingyi::HexElementData h{};
    float3 lo = make_float3(1e30f, 1e30f, 1e30f), hi = make_float3(-1e30f, -1e30f, -1e30f);
    for (int c = 0; c < 8; ++c) {
      h.vertices[c] = make_float3(verts[(size_t)e * 24 + 3 * c], verts[(size_t)e * 24 + 3 * c + 1],
                                  verts[(size_t)e * 24 + 3 * c + 2]);
      lo = make_float3(std::min(lo.x, h.vertices[c].x), std::min(lo.y, h.vertices[c].y),
                       std::min(lo.z, h.vertices[c].z));
      hi = make_float3(std::max(hi.x, h.vertices[c].x), std::max(hi.y, h.vertices[c].y),
                       std::max(hi.z, h.vertices[c].z));
    }
    h.aabbMin = lo;  // vertex box; the device builds exact curved boxes
    h.aabbMax = hi;
    const float* g = &geom[(size_t)e * 3 * g3];
    Scene::FieldSlice s;
    s.name = "u";
    s.modes3 = make_uint3(nf, nf, nf);
    s.count = f3;
    s.offset = (int)(e * f3);
    std::vector<Scene::FieldSlice> sl{s};
    scene.add_hex(h, &field[(size_t)e * f3], make_uint3(nf, nf, nf), f3, g, g + g3, g + 2 * g3,
                  make_uint3(ng, ng, ng), g3, &sl);
  }
  return true;
}

void camera_rays(const Camera& cam, int width, int height, std::vector<float>& rays)
{
  const size_t n = (size_t)width * height;
  rays.assign(6 * n, 0.f);
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      const size_t p = (size_t)y * width + x;
      const float s = (x + 0.5f) / width, t = (y + 0.5f) / height;
      float3 d = sub(make_float3(cam.lower_left_corner.x + s * cam.horizontal.x + t * cam.vertical.x,
                                 cam.lower_left_corner.y + s * cam.horizontal.y + t * cam.vertical.y,
                                 cam.lower_left_corner.z + s * cam.horizontal.z + t * cam.vertical.z),
                     cam.origin);
      d = unit(d);
      rays[p] = cam.origin.x;
      rays[n + p] = cam.origin.y;
      rays[2 * n + p] = cam.origin.z;
      rays[3 * n + p] = d.x;
      rays[4 * n + p] = d.y;
      rays[5 * n + p] = d.z;
    }
}

bool render(Scene& scene, const Camera& cam, int width, int height, std::vector<uchar4>& out,
            std::vector<HitAOV>* aov, const RenderParams* params)
{
  if (!scene.mesh) scene.build_bvh();
  if (!scene.mesh || width <= 0 || height <= 0) {
    std::cerr << "render: scene not built" << std::endl;
    return false;
  }
  const RenderParams P = params ? *params : RenderParams();
  if (scene.samples_per_pixel > 1 || scene.max_depth > 1) {
    if (aov) {
      std::cerr << "render: per-pixel AOVs are not produced by path tracing" << std::endl;
      return false;
    }
    std::vector<femrt_material> mats = scene.materials;
    if (mats.empty()) {
      femrt_material g{{0.5f, 0.5f, 0.5f}, {0.f, 0.f, 0.f}, 0.f, 0.f, 0};
      femrt_material b{{0.1f, 0.2f, 0.8f}, {0.f, 0.f, 0.f}, 0.f, 0.f, 0};
      mats = {g, b};
    }
    const int spp = std::max(1, scene.samples_per_pixel);
    femrt_frame* f = nullptr;
    if (femrt_frame_create(scene.ctx, width * height * spp, &f) != FEMRT_OK) {
      std::cerr << "render: " << femrt_last_error(scene.ctx) << std::endl;
      return false;
    }
    const float cam12[12] = {cam.origin.x, cam.origin.y, cam.origin.z,
                             cam.lower_left_corner.x, cam.lower_left_corner.y, cam.lower_left_corner.z,
                             cam.horizontal.x, cam.horizontal.y, cam.horizontal.z,
                             cam.vertical.x, cam.vertical.y, cam.vertical.z};
    out.resize((size_t)width * height);
    const bool ok = femrt_render_path(scene.mesh, f, cam12, width, height, spp, std::max(1, scene.max_depth),
                                      mats.data(), (int)mats.size(), P.seed,
                                      reinterpret_cast<uint32_t*>(out.data())) == FEMRT_OK;
    if (!ok) std::cerr << "render: " << femrt_last_error(scene.ctx) << std::endl;
    femrt_frame_destroy(f);
    return ok;
  }
  const size_t n = (size_t)width * height;
  std::vector<float> rays;
  camera_rays(cam, width, height, rays);
  femrt_frame* f = nullptr;
  if (femrt_frame_create(scene.ctx, (int)n, &f) != FEMRT_OK) {
    std::cerr << "render: " << femrt_last_error(scene.ctx) << std::endl;
    return false;
  }
  femrt_trace_params tp;
  femrt_trace_params_default(&tp);
  const bool iso = scene.use_iso && scene.hexes.size() > 0;
  tp.iso_enable = iso ? 1 : 0;
  tp.n_iso = iso ? 1 : 0;
  tp.iso[0] = scene.iso_value;
  for (double v : scene.extra_iso_values)
    if (tp.n_iso < FEMRT_MAX_ISO) tp.iso[tp.n_iso++] = v;
  tp.guess = P.guess;
  tp.samples = P.samples;
  tp.max_march = P.max_march;
  tp.flags = 1;
  out.resize(n);
  bool ok = femrt_frame_set_rays(f, rays.data()) == FEMRT_OK && femrt_trace(scene.mesh, f, &tp) == FEMRT_OK &&
            femrt_frame_get_rgba(f, reinterpret_cast<uint32_t*>(out.data())) == FEMRT_OK;
  if (ok && aov) {
    const int K = tp.n_iso;
    std::vector<int> he(n), hf(n), ie((size_t)std::max(K, 1) * n);
    std::vector<double> ht(n), hr(3 * n), hu(n), it((size_t)std::max(K, 1) * n);
    ok = femrt_frame_get_hits(f, he.data(), hf.data(), ht.data(), hr.data(), hu.data()) == FEMRT_OK &&
         (K == 0 || femrt_frame_get_iso(f, K, ie.data(), it.data(), nullptr, nullptr) == FEMRT_OK);
    if (ok) {
      aov->resize(n);
      for (size_t p = 0; p < n; ++p) {
        HitAOV& a = (*aov)[p];
        a.elem = he[p];
        a.face = hf[p];
        a.t = ht[p];
        for (int c = 0; c < 3; ++c) a.rst[c] = hr[c * n + p];
        a.u = hu[p];
        for (int k = 0; k < FEMRT_MAX_ISO; ++k) {
          a.iso_elem[k] = k < K ? ie[(size_t)k * n + p] : -1;
          a.iso_t[k] = k < K ? it[(size_t)k * n + p] : 0.0;
        }
      }
    }
  }
  if (!ok) std::cerr << "render: " << femrt_last_error(scene.ctx) << std::endl;
  femrt_frame_destroy(f);
  return ok;
}

void save_image_ppm(const std::vector<uchar4>& image, int width, int height, const std::string& filename)
{
  std::ofstream file(filename, std::ios::out | std::ios::binary);
  if (!file.is_open()) {
    std::cerr << "Error: Could not open file for writing: " << filename << std::endl;
    return;
  }
  file << "P6\n" << width << " " << height << "\n255\n";
  for (int i = 0; i < width * height && i < (int)image.size(); ++i) {
    const char px[3] = {(char)image[i].x, (char)image[i].y, (char)image[i].z};
    file.write(px, 3);
  }
}
