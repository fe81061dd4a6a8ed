// rhllc_make_geo: the RHLLC part of the reference's make_all_geo_data utility
// (utils/src/make_all_geo_data.cpp:12-41) without VTK:
//   rhllc_make_geo <base_address> <file.vtk>
// runs BuildDataFromVTK, then BinToGeo, writing neighbors.bin, normals.bin,
// areas.bin, volume.bin, centers.bin, geo_faces.bin and geo_cells.bin under
// base_address (a directory prefix, normally ending in '/').
#include <cstdio>
#include <filesystem>
#include <string>

#include "rhllc_dropin.h"

int main(int argc, char *argv[]) {
  if (argc < 3) {
    std::fprintf(stderr, "Need base_address and file_vtk\n");
    return e_completion_fail;
  }
  global_files_t files;
  files.base_address = argv[1];
  files.name_file_vtk = argv[2];
  files.Build();
  if (!std::filesystem::exists(files.base_address) || !std::filesystem::exists(files.name_file_vtk)) {
    std::fprintf(stderr, "bad input address\n");
    return e_completion_fail;
  }
  if (BuildDataFromVTK(files) != e_completion_success) {
    std::fprintf(stderr, "don't build start format\n");
    return e_completion_fail;
  }
  if (BinToGeo(files.base_address) != e_completion_success) {
    std::fprintf(stderr, "don't build geo format\n");
    return e_completion_fail;
  }
  return e_completion_success;
}
