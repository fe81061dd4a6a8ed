// rhllc_solver: the RHLLC module executable on the MI355X step
// (src/RHD/rhllc_solver.cpp:6-27).  argv[1] = path to directories_cfg.json.
#include <cstdio>
#include <string>

#include "rhllc_dropin.h"

int main(int argc, char *argv[]) {
  std::string file_config = "directories_cfg.json";
  if (argc > 1) file_config = argv[1];
  if (files_sys::json::ReadStartSettings(file_config, glb_files, &_solve_mode, &_hllc_cfg)) return e_completion_fail;
  return rhllc::RunRhllcModule();
}
