// femrt_render -- command-line renderer over the C++ host layer, the
// counterpart of the reference's main program (main.cu:23-596): load a
// Nektar++ XML/FLD mesh, frame it with the automatic camera (main.cu:248-302),
// trace, write a PPM.
//
//   femrt_render mesh.xml field.fld out.ppm [width height [iso ...]]
//                [--guess plain|warm] [--aov file] [--rays file] [--load-only]
//                [--spp n] [--depth n] [--seed n]
//
// --spp / --depth > 1 path-trace the mesh surface (the reference's default
// grey and blue diffuse materials) instead of the primary-ray AOV render.
//
// --aov writes the per-pixel HitAOV records and --rays the 6 x W*H SoA ray
// floats (raw, host byte order); --load-only stops after loading and prints
// the element count, mode counts and coefficient sums.
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
  std::string aov_path, rays_path;
  bool load_only = false;
  int spp = 1, depth = 1;
  RenderParams rp;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--aov") && i + 1 < argc) {
      aov_path = argv[++i];
    } else if (!std::strcmp(argv[i], "--rays") && i + 1 < argc) {
      rays_path = argv[++i];
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
  const int width = pos.size() > 3 ? std::atoi(pos[3].c_str()) : 640;
  const int height = pos.size() > 4 ? std::atoi(pos[4].c_str()) : 360;
  Scene scene;
  scene.width = width;
  scene.height = height;
  scene.samples_per_pixel = spp;
  scene.max_depth = depth;
  if (!load_xml_fld(pos[0], pos[1], scene)) return 1;
  if (load_only) {
    double gs = 0.0, fs = 0.0;
    for (float v : scene.hex_geom_coefficients) gs += v;
    for (float v : scene.hex_coefficients) fs += v;
    std::printf("hexes %zu geom_modes %u field_modes %u geom_sum %.9g field_sum %.9g\n", scene.hexes.size(),
                scene.hexes.empty() ? 0u : scene.hexes[0].geomModes.x,
                scene.hexes.empty() ? 0u : scene.hexes[0].fieldModes.x, gs, fs);
    return 0;
  }
  scene.use_iso = pos.size() > 5;
  if (scene.use_iso) {
    scene.iso_value = (float)std::atof(pos[5].c_str());
    for (size_t k = 6; k < pos.size(); ++k) scene.extra_iso_values.push_back(std::atof(pos[k].c_str()));
  }
  float3 lo = make_float3(1e30f, 1e30f, 1e30f), hi = make_float3(-1e30f, -1e30f, -1e30f);
  for (const auto& h : scene.hexes)
    for (const auto& v : h.vertices) {
      lo = make_float3(std::min(lo.x, v.x), std::min(lo.y, v.y), std::min(lo.z, v.z));
      hi = make_float3(std::max(hi.x, v.x), std::max(hi.y, v.y), std::max(hi.z, v.z));
    }
  const float3 c = make_float3(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z));
  const float md = std::max(hi.x - lo.x, std::max(hi.y - lo.y, hi.z - lo.z));
  const float3 from = make_float3(c.x + 2.5f * md, c.y - 2.5f * md, c.z - 2.0f * md);
  const Camera cam(from, c, make_float3(0, 1, 0), 20.0f, (float)width / height, 0.0f, 1.0f);
  std::vector<uchar4> img;
  std::vector<HitAOV> aov;
  const bool paths = spp > 1 || depth > 1;
  if (!render(scene, cam, width, height, img, aov_path.empty() || paths ? nullptr : &aov, &rp)) return 1;
  save_image_ppm(img, width, height, pos[2]);
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
  std::cout << "rendered " << width << "x" << height << " from " << scene.hexes.size() << " hexes";
  if (!aov_path.empty()) std::cout << ", " << hits << " hit pixels";
  std::cout << std::endl;
  return 0;
}
