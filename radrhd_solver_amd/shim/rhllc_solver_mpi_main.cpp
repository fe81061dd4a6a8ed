// rhllc_solver_mpi: the RHLLC MPI module executable (src/RHD/rhllc_solver_mpi.cpp:6-31)
// at one rank on the MI355X step.  argv[1] = path to directories_cfg.json.  The exit
// status is the module's e_type_completion (the reference's main returns 0).
#include <cstdio>
#include <string>

#include "rhllc_dropin.h"

int main(int argc, char *argv[]) {
  std::string file_config = "directories_cfg.json";
  if (argc > 1) file_config = argv[1];
  if (files_sys::json::ReadStartSettings(file_config, glb_files, &_solve_mode, &_hllc_cfg)) return e_completion_fail;
  return rhllc::RunRhllcMpiModule();
}
