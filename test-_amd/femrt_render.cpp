// femrt_render -- command-line renderer over the C++ host layer, the
// counterpart of the reference's main program (main.cu:23-596): load a
// Nektar++ XML/FLD mesh, frame it with the automatic camera (main.cu:248-302),
// trace, write a PPM.
//
//   femrt_render mesh.xml field.fld out.ppm [width height [iso ...]]
//                [--guess plain|warm] [--aov file] [--rays file] [--load-only]
//                [--spp n] [--depth n] [--seed n]
//   femrt_render --quads scene.quads out.ppm [width height] [options]
//   femrt_render --prisms scene.prisms out.ppm [width height] [options]
//   femrt_render --tris scene.tris out.ppm [width height] [options]
//
// --quads reads a curved-quad surface scene (raw, host byte order: int32
// n, ng, nf; float32 corners n x 4 x 3, geometry n x 3 x ng^2, field
// n x nf^2) and adds it quad by quad with Scene::add_quad.  --prisms reads
// a prism scene (int32 n, gx, gy, gz, fx, fy, fz; float32 vertices
// n x 6 x 3, geometry n x 3 x count(g) when gx > 0, field n x count(f)
// when fx > 0, count(m) = sum over i < mx of my (mz - i)) and adds it with
// Scene::add_prism.  --tris reads a triangle scene (int32 n, gx, gy, fx, fy;
// float32 vertices n x 3 x 3, geometry n x 3 x count(g) when gx > 0, field
// n x count(f) when fx > 0, count(m) = sum over i < mx of my - i) and adds
// it with Scene::add_curved_tri.
//
// --spp / --depth > 1 path-trace the mesh surface (the reference's default
// grey and blue diffuse materials) instead of the primary-ray AOV render.
//
// --aov writes the per-pixel HitAOV records and --rays the 6 x W*H SoA ray
// floats (raw, host byte order); --load-only stops after loading and prints
// the element count, mode counts and coefficient sums.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../include/femrt_scene.hpp"

int main(int argc, char** argv)
{
  if (argc < 4) {
    std::cerr << "usage: femrt_render mesh.xml field.fld out.ppm [width height [iso ...]] "
                 "[--guess plain|warm] [--aov file]" << std::endl;
    return 2;
  }
  std::vector<std::string> pos;
  std::string aov_path, rays_path, quads_path, prisms_path, tris_path;
  bool load_only = false;
  int spp = 1, depth = 1;
  RenderParams rp;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--aov") && i + 1 < argc) {
      aov_path = argv[++i];
    } else if (!std::strcmp(argv[i], "--rays") && i + 1 < argc) {
      rays_path = argv[++i];
    } else if (!std::strcmp(argv[i], "--quads") && i + 1 < argc) {
      quads_path = argv[++i];
    } else if (!std::strcmp(argv[i], "--prisms") && i + 1 < argc) {
      prisms_path = argv[++i];
    } else if (!std::strcmp(argv[i], "--tris") && i + 1 < argc) {
      tris_path = argv[++i];
    } else if (!std::strcmp(argv[i], "--load-only")) {
      load_only = true;
    } else if (!std::strcmp(argv[i], "--spp") && i + 1 < argc) {
      spp = std::atoi(argv[++i]);
    } else if (!std::strcmp(argv[i], "--depth") && i + 1 < argc) {
      depth = std::atoi(argv[++i]);
    } else if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) {
      rp.seed = std::strtoull(argv[++i], nullptr, 10);
    } else if (!std::strcmp(argv[i], "--guess") && i + 1 < argc) {
      const std::string g = argv[++i];
      rp.guess = g == "plain" ? FEMRT_GUESS_PLAIN : FEMRT_GUESS_WARM;
    } else {
      pos.push_back(argv[i]);
    }
  }
  const size_t b = quads_path.empty() && prisms_path.empty() && tris_path.empty() ? 2 : 0;   // index of out.ppm
  if (pos.size() <= b) {
    std::cerr << "femrt_render: no output file" << std::endl;
    return 2;
  }
  const int width = pos.size() > b + 1 ? std::atoi(pos[b + 1].c_str()) : 640;
  const int height = pos.size() > b + 2 ? std::atoi(pos[b + 2].c_str()) : 360;
  Scene scene;
  scene.width = width;
  scene.height = height;
  scene.samples_per_pixel = spp;
  scene.max_depth = depth;
  std::vector<float> qdata;
  if (!quads_path.empty()) {
    FILE* f = std::fopen(quads_path.c_str(), "rb");
    int hdr[3] = {0, 0, 0};
    if (!f || std::fread(hdr, 4, 3, f) != 3 || hdr[0] <= 0 || hdr[1] < 1 || hdr[2] < 1) {
      std::cerr << "cannot read " << quads_path << std::endl;
      if (f) std::fclose(f);
      return 1;
    }
    const size_t n = hdr[0], g2 = (size_t)hdr[1] * hdr[1], f2 = (size_t)hdr[2] * hdr[2];
    qdata.resize(n * (12 + 3 * g2 + f2));
    const bool ok = std::fread(qdata.data(), 4, qdata.size(), f) == qdata.size();
    std::fclose(f);
    if (!ok) {
      std::cerr << "cannot read " << quads_path << std::endl;
      return 1;
    }
    const float* V = qdata.data();
    const float* G = V + n * 12;
    const float* F = G + n * 3 * g2;
    for (size_t e = 0; e < n; ++e) {
      zidingyi::QuadElementData q{};
      for (int c = 0; c < 4; ++c) q.vertices[c] = make_float3(V[e * 12 + 3 * c], V[e * 12 + 3 * c + 1], V[e * 12 + 3 * c + 2]);
      const float* g = G + e * 3 * g2;
      scene.add_quad(q, F + e * f2, make_uint2(hdr[2], hdr[2]), f2, g, g + g2, g + 2 * g2,
                     make_uint2(hdr[1], hdr[1]), g2);
    }
  } else if (!prisms_path.empty()) {
    FILE* f = std::fopen(prisms_path.c_str(), "rb");
    int hdr[7] = {0, 0, 0, 0, 0, 0, 0};
    if (!f || std::fread(hdr, 4, 7, f) != 7 || hdr[0] <= 0) {
      std::cerr << "cannot read " << prisms_path << std::endl;
      if (f) std::fclose(f);
      return 1;
    }
    auto count = [](const int* m) {
      size_t c = 0;
      for (int i = 0; i < m[0]; ++i) c += (size_t)m[1] * (size_t)std::max(m[2] - i, 0);
      return c;
    };
    const size_t n = hdr[0], gc = count(hdr + 1), fc = count(hdr + 4);
    qdata.resize(n * (18 + 3 * gc + fc));
    const bool ok = std::fread(qdata.data(), 4, qdata.size(), f) == qdata.size();
    std::fclose(f);
    if (!ok) {
      std::cerr << "cannot read " << prisms_path << std::endl;
      return 1;
    }
    const float* V = qdata.data();
    const float* G = V + n * 18;
    const float* F = G + n * 3 * gc;
    for (size_t e = 0; e < n; ++e) {
      zidingyi::PrismElementData p{};
      for (int c = 0; c < 6; ++c) p.vertices[c] = make_float3(V[e * 18 + 3 * c], V[e * 18 + 3 * c + 1], V[e * 18 + 3 * c + 2]);
      const float* g = G + e * 3 * gc;
      scene.add_prism(p, fc ? F + e * fc : nullptr, make_uint3(hdr[4], hdr[5], hdr[6]), fc, gc ? g : nullptr,
                      gc ? g + gc : nullptr, gc ? g + 2 * gc : nullptr, make_uint3(hdr[1], hdr[2], hdr[3]), gc);
    }
  } else if (!tris_path.empty()) {
    FILE* f = std::fopen(tris_path.c_str(), "rb");
    int hdr[5] = {0, 0, 0, 0, 0};
    if (!f || std::fread(hdr, 4, 5, f) != 5 || hdr[0] <= 0) {
      std::cerr << "cannot read " << tris_path << std::endl;
      if (f) std::fclose(f);
      return 1;
    }
    auto count = [](const int* m) {
      size_t c = 0;
      for (int i = 0; i < m[0]; ++i) c += (size_t)std::max(m[1] - i, 0);
      return c;
    };
    const size_t n = hdr[0], gc = count(hdr + 1), fc = count(hdr + 3);
    qdata.resize(n * (9 + 3 * gc + fc));
    const bool ok = std::fread(qdata.data(), 4, qdata.size(), f) == qdata.size();
    std::fclose(f);
    if (!ok) {
      std::cerr << "cannot read " << tris_path << std::endl;
      return 1;
    }
    const float* V = qdata.data();
    const float* G = V + n * 9;
    const float* F = G + n * 3 * gc;
    for (size_t e = 0; e < n; ++e) {
      zidingyi::TriElementData t{};
      for (int c = 0; c < 3; ++c) t.vertices[c] = make_float3(V[e * 9 + 3 * c], V[e * 9 + 3 * c + 1], V[e * 9 + 3 * c + 2]);
      const float* g = G + e * 3 * gc;
      scene.add_curved_tri(t, fc ? F + e * fc : nullptr, make_uint2(hdr[3], hdr[4]), fc, gc ? g : nullptr,
                           gc ? g + gc : nullptr, gc ? g + 2 * gc : nullptr, make_uint2(hdr[1], hdr[2]), gc);
    }
  } else if (!load_xml_fld(pos[0], pos[1], scene)) {
    return 1;
  }
  if (load_only) {
    double gs = 0.0, fs = 0.0;
    for (float v : scene.hex_geom_coefficients) gs += v;
    for (float v : scene.hex_coefficients) fs += v;
    std::printf("hexes %zu geom_modes %u field_modes %u geom_sum %.9g field_sum %.9g\n", scene.hexes.size(),
                scene.hexes.empty() ? 0u : scene.hexes[0].geomModes.x,
                scene.hexes.empty() ? 0u : scene.hexes[0].fieldModes.x, gs, fs);
    return 0;
  }
  scene.use_iso = pos.size() > b + 3;
  if (scene.use_iso) {
    scene.iso_value = (float)std::atof(pos[b + 3].c_str());
    for (size_t k = b + 4; k < pos.size(); ++k) scene.extra_iso_values.push_back(std::atof(pos[k].c_str()));
  }
  float3 lo = make_float3(1e30f, 1e30f, 1e30f), hi = make_float3(-1e30f, -1e30f, -1e30f);
  auto grow = [&](const float3& v) {
    lo = make_float3(std::min(lo.x, v.x), std::min(lo.y, v.y), std::min(lo.z, v.z));
    hi = make_float3(std::max(hi.x, v.x), std::max(hi.y, v.y), std::max(hi.z, v.z));
  };
  for (const auto& h : scene.hexes)
    for (const auto& v : h.vertices) grow(v);
  for (const auto& q : scene.quads)
    for (const auto& v : q.vertices) grow(v);
  for (const auto& p : scene.prisms)
    for (const auto& v : p.vertices) grow(v);
  for (const auto& t : scene.curved_tris)
    for (const auto& v : t.vertices) grow(v);
  const float3 c = make_float3(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z));
  const float md = std::max(hi.x - lo.x, std::max(hi.y - lo.y, hi.z - lo.z));
  const float3 from = make_float3(c.x + 2.5f * md, c.y - 2.5f * md, c.z - 2.0f * md);
  const Camera cam(from, c, make_float3(0, 1, 0), 20.0f, (float)width / height, 0.0f, 1.0f);
  std::vector<uchar4> img;
  std::vector<HitAOV> aov;
  const bool paths = spp > 1 || depth > 1;
  if (!render(scene, cam, width, height, img, aov_path.empty() || paths ? nullptr : &aov, &rp)) return 1;
  save_image_ppm(img, width, height, pos[b]);
  if (!rays_path.empty()) {
    std::vector<float> rays;
    camera_rays(cam, width, height, rays);
    FILE* f = std::fopen(rays_path.c_str(), "wb");
    if (!f || std::fwrite(rays.data(), 4, rays.size(), f) != rays.size()) {
      std::cerr << "cannot write " << rays_path << std::endl;
      return 1;
    }
    std::fclose(f);
  }
  if (!aov_path.empty()) {
    FILE* f = std::fopen(aov_path.c_str(), "wb");
    if (!f || std::fwrite(aov.data(), sizeof(HitAOV), aov.size(), f) != aov.size()) {
      std::cerr << "cannot write " << aov_path << std::endl;
      return 1;
    }
    std::fclose(f);
  }
  int hits = 0;
  for (const auto& a : aov) hits += a.elem >= 0;
  std::cout << "rendered " << width << "x" << height << " from ";
  if (!scene.quads.empty())
    std::cout << scene.quads.size() << " quads";
  else if (!scene.prisms.empty())
    std::cout << scene.prisms.size() << " prisms";
  else if (!scene.curved_tris.empty())
    std::cout << scene.curved_tris.size() << " triangles";
  else
    std::cout << scene.hexes.size() << " hexes";
  if (!aov_path.empty()) std::cout << ", " << hits << " hit pixels";
  std::cout << std::endl;
  return 0;
}
