// Drop-in replacement of the reference's RHLLC step interface, backed by the
// MI355X kernels of librhllc_hip.so.
//
// Same names, arguments and error behaviour as the reference:
//   rhllc::Hllc3dStab(tau, grid)        solvers/rhllc/include/rhllc_calc.h:32
//   rhllc::max_signal_speed             solvers/rhllc/include/rhllc_utils.h:17
//   rhllc::GetTimeStep(hllc_set, cells) solvers/rhllc/include/rhllc_utils.h:35
//   rhllc::HllcPhysToConv(cells)        solvers/rhllc/include/rhllc_utils.h:43
//   rhllc::Init(file, cells)            solvers/rhllc/include/rhllc_init.h:32
//   rhllc::RunRhllcModule()             solvers/rhllc/include/rhllc_main.h:21
//   rhllc::RunRhllcMpiModule()          solvers/rhllc/include/rhllc_main.h:22 (one rank)
//   files_sys::json::ReadStartSettings  lib/files_sys/include/reader_json.h:63-89
//   files_sys::bin::ReadGridGeo / ReadHllcInit / WriteSolution
//                                       lib/files_sys/include/reader_bin.h:113-135,232-252,
//                                       lib/files_sys/src/writer_bin.cpp:56-114
//   files_sys::bin::WriteSolutionMPI    lib/files_sys/src/writer_bin.cpp:120-170 (one rank)
//   BuildDataFromVTK(glb_files)         lib/geometry/include/make_internal_format.h
//   BinToGeo(base_address)              utils/template_functions/bin_files_to_geo.h:11-38
// Where the reference aborts (DIE_IF), these abort too, after a message on
// stderr; the functions that return e_type_completion return it.
//
// Device residency: by default every Hllc3dStab call uploads grid.cells,
// steps on the GPU and downloads cells and face fluxes, so the host grid is
// always current (a strict drop-in).  SetDeviceResident(true) keeps the state
// on the device between calls; SyncGridFromDevice brings it back (the module
// driver does this only before writing a solution).
#pragma once

#include <string>
#include <vector>

#include "rhllc_types.h"

extern hllc_value_t _hllc_cfg;      // lib/grids/src/solvers_struct.cpp:8
extern solve_mode_t _solve_mode;    // lib/grids/src/solvers_struct.cpp:7
extern global_files_t glb_files;    // lib/json/src/json_struct.cpp:10

namespace rhllc {

extern Type max_signal_speed;

void Hllc3dStab(const Type tau, grid_t &grid);
Type GetTimeStep(const hllc_value_t &hllc_set, const std::vector<elem_t> &cells);
void HllcPhysToConv(std::vector<elem_t> &cells);
int Init(std::string &file_init_value, std::vector<elem_t> &cells);
int RunRhllcModule();
int RunRhllcMpiModule();  // at one rank: the partition must put every cell on rank 0

// Extensions of the drop-in (no reference counterpart).
void SetDeviceResident(bool resident);
void SyncGridFromDevice(grid_t &grid);
void ReleaseDevice();

}  // namespace rhllc

namespace files_sys {
namespace json {
int Read(const std::string &file, global_files_t &st);
int Read(const std::string &file, solve_mode_t &st);
int Read(const std::string &file, hllc_value_t &st);
int ReadStartSettings(const std::string &file_set, global_files_t &glb, solve_mode_t *solve_mode = nullptr,
                      hllc_value_t *hllc_conf = nullptr);
}  // namespace json
namespace bin {
int ReadGridGeo(const std::string &name_file, std::vector<face_t> &faces);
int ReadGridGeo(const std::string &name_file, std::vector<elem_t> &cells);
int ReadHllcInit(const std::string &file_init_value, std::vector<elem_t> &cells);
int WriteSolution(const std::string &main_dir, const grid_t &grid);
int WriteSolutionMPI(const std::string &main_dir, const grid_t &grid);  // flux.bin, one rank
}  // namespace bin
}  // namespace files_sys

// Mesh preprocessing without VTK (legacy-format VTK files of tetrahedra), by
// libradrhd_mesh.so.  BuildDataFromVTK writes neighbors.bin, normals.bin,
// areas.bin, volume.bin and centers.bin under glb_files.base_address, with the
// GEOMETRY_TYPE Cone boundary codes; BinToGeo writes geo_faces.bin and
// geo_cells.bin from them.  Both return e_type_completion.
int BuildDataFromVTK(const global_files_t &glb_files);
int BinToGeo(const std::string &base_address);
