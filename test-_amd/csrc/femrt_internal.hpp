// femrt_internal.hpp -- private declarations shared by the library's sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/femrt.h"

namespace femrt {

// Per-mesh host-side products of the build (Bernstein boxes, BVH).
struct BvhHost {
  std::vector<float> nodes;  // n_nodes x 8 (lo.xyz, a, hi.xyz, b), a/b int bits
  std::vector<int> prims;
};

// Host algorithms (femrt_host.cpp)
// Boxes per element: [0] element, [1..6] faces (Nektar order), [7 + 9 f +
// 3 b + a] sub-patch (a, b) of face f on the 3 x 3 split of its (u, v)
// square; each box lo.xyz, hi.xyz.
#define FEMRT_NBOX 61
#define FEMRT_BOX_STRIDE (FEMRT_NBOX * 6)
void bernstein_boxes(int n_elem, int ng, const float* geom, float* out /* n x FEMRT_NBOX x 6 */);
void build_bvh_sah(int n_elem, const float* boxes /* n x FEMRT_NBOX x 6, element box first */,
                   BvhHost& out);
int gen_structured_hex(const femrt_gen_params* gp, float* vertices, float* geom, float* field,
                       int* adjacency);
void encode_adjacency(int n_elem, const float* vertices, const int* adjacency, int* out);
int gen_field_waves(int n_elem, int ng, const float* geom, int nf, int n_waves, uint64_t seed,
                    float* field);
void geom_jinv_bound(int n_elem, int ng, const float* geom, float* out /* n */);
void field_lipschitz(int n_elem, int nf, const float* field, float* out /* n */);
void field_range(int n_elem, int nf, const float* field, float* out /* n x (lo, hi) */);
std::vector<double> gll_nodes(int n);
bool modal_fit_1d_matrix(int n, const std::vector<double>& nodes, std::vector<double>& inv);
double modA(int p, double x);
std::vector<double> lagrange_matrix(const std::vector<double>& src, const std::vector<double>& dst);
void tensor_fit(int n, const std::vector<double>& inv, const double* nodal, double* out);
double modB(int i, int k, double x);
int prism_n_coeffs(const int* modes);
void prism_to_hex(int n_elem, const int* modes, int n, const float* coeffs, int comps, double* out);
void tri_to_quad(int n_elem, const int* modes, int n, const float* coeffs, int comps, double* out);
int save_xml_fld(const char* xml_path, const char* fld_path, int n, int ng, int nf,
                 const float* vertices, const float* geom, const float* field, std::string* err);
int load_xml_fld(const char* xml_path, const char* fld_path, int flags, int* n_elem, int* ng,
                 int* nf, float* vertices, float* geom, float* field, int* adjacency,
                 std::string* err);

}  // namespace femrt
