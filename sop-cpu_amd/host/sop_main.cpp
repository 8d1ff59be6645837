// sop_main.cpp -- `sop_gpu.x input_file`: the reference's command-line driver
// (sop.cpp:17-53) with the hot path on the GPU.  Same input scripts, same
// output files, same closing timing report.
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <ctime>
#include <iostream>

#include "sop_state.h"

// `sop_gpu.x --check input_file`: parse the script and apply its load / set
// commands without touching the GPU, then print the derived state (one
// "name value" line each, %.17g) -- for checking the host layer on a CPU.
static int check_only(const char* path) {
  alloc_arrays();
  read_input(path);
  for (int i = 1; i <= ncmd; i++) {
    if (!std::strcmp(cmd[i], "load")) load(i);
    else if (!std::strcmp(cmd[i], "set")) set_params(i);
  }
  const struct { const char* n; double v; } d[] = {
      {"a1", a1}, {"a2", a2}, {"a3", a3}, {"a4", a4}, {"a5", a5}, {"h", h}, {"halfh", halfh}, {"zeta", zeta},
      {"T", T}, {"boxl", boxl}, {"ncell", ncell}, {"lcell", lcell}, {"nstep", nstep},
      {"istep_restart", istep_restart}, {"e_ang_ss_coeff", e_ang_ss_coeff}, {"f_ang_ss_coeff", f_ang_ss_coeff}};
  for (auto& e : d) std::printf("%s %.17g\n", e.n, e.v);
  const struct { const char* n; int v; } k[] = {
      {"nbead", nbead}, {"nbnd", nbnd}, {"nang", nang}, {"ncon_att", ncon_att}, {"ncon_rep", ncon_rep},
      {"nup", nup}, {"nnlup", nnlup}, {"sim_type", sim_type}, {"neighborlist", neighborlist},
      {"celllist", celllist}, {"restart", restart}, {"rgen_restart", rgen_restart}, {"run", run}, {"ncmd", ncmd}};
  for (auto& e : k) std::printf("%s %d\n", e.n, e.v);
  std::printf("first_gasdev %.17g\n", generator.gasdev());
  return 0;
}

int main(int argc, char* argv[]) {
  if (argc >= 3 && !std::strcmp(argv[1], "--check")) return check_only(argv[2]);
  if (argc < 2) {
    std::cerr << "Usage: " << argv[0] << " < input_file >" << std::endl;
    return -1;
  }
  const time_t tm0 = time(nullptr);
  std::cout << "CURRENT TIME IS: " << ctime(&tm0);
  if (getcwd(pathname, MAXPATHLEN) == nullptr) std::cerr << "PROBLEM GETTING PATH" << std::endl;
  else std::cout << "CURRENT WORKING DIRECTORY: " << pathname << std::endl;
  alloc_arrays();
  read_input(argv[1]);
  const clock_t ck0 = clock();
  ex_cmds();
  const time_t tm1 = time(nullptr);
  const clock_t ck1 = clock();
  std::cout << "+-------------------+" << std::endl
            << "| Simulation Stats: |" << std::endl
            << "+-------------------+" << std::endl;
  std::cout << "Wall Time              : " << difftime(tm1, tm0) << " sec" << std::endl;
  std::cout << "Total Computation Time : " << float(ck1 - ck0) / CLOCKS_PER_SEC << " sec" << std::endl;
  std::cout << "Computation Rate       : " << float(ck1 - ck0) / CLOCKS_PER_SEC / nstep << " sec / timestep"
            << std::endl;
  std::cout << "CURRENT TIME IS        : " << ctime(&tm1);
  return 0;
}
