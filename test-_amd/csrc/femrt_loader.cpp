// femrt_loader.cpp -- Nektar++ XML/FLD loader (drop-in for load_xml_fld).
#include <string>

#include "femrt_internal.hpp"

namespace femrt {

int load_xml_fld(const char* xml_path, const char* fld_path, int flags, int* n_elem, int* ng,
                 int* nf, float* vertices, float* geom, float* field, int* adjacency,
                 std::string* err)
{
  (void)xml_path; (void)fld_path; (void)flags; (void)n_elem; (void)ng; (void)nf;
  (void)vertices; (void)geom; (void)field; (void)adjacency;
  if (err) *err = "load_xml_fld: not built yet";
  return FEMRT_EUNSUP;
}

}  // namespace femrt
