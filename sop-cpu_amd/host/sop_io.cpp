// sop_io.cpp -- input script and output files of the host driver, in the
// reference's formats: the `;`-terminated command script (io.cpp:9-89), the
// parameter banner (io.cpp:91-154), "idx x y z" coordinate / velocity files
// and "x y z" trajectory frames (io.cpp:156-304).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "sop_state.h"

// Commands are whitespace (single space) separated words; a word containing
// ';' ends the current command, and every ';' is cut from the stored words.
// As in the reference, a final line without a newline is not read; blank lines
// are skipped (the reference dereferences a null token on them).
void read_input(const char* const ifile) {
  std::ifstream in(ifile);
  if (!in) {
    std::cerr << "CANNOT OPEN INPUT FILE " << ifile << std::endl;
    std::exit(-1);
  }
  ncmd = 0;
  int nopt = 0;
  bool newcmd = true;
  std::string line;
  while (std::getline(in, line)) {
    if (in.eof()) {
      std::cerr << "warning: last input line has no newline and is ignored (as by SOP-CPU): " << line << std::endl;
      break;
    }
    std::vector<char> buf(line.begin(), line.end());
    buf.push_back(0);
    for (char* t = std::strtok(buf.data(), " "); t; t = std::strtok(nullptr, " ")) {
      if (newcmd) {
        if (ncmd >= mcmd) { std::cerr << "too many commands (max " << mcmd << ")" << std::endl; std::exit(-1); }
        ncmd++;
        std::strncpy(cmd[ncmd], t, mwdsize - 1);
        opt_ptr[ncmd] = nopt + 1;
        newcmd = false;
      } else {
        if (nopt >= mopt_tot) { std::cerr << "too many options" << std::endl; std::exit(-1); }
        nopt++;
        std::strncpy(opt[nopt], t, mwdsize - 1);
      }
      if (std::strchr(t, ';')) newcmd = true;
    }
  }
  if (!newcmd && ncmd > 0) ncmd--;  // an unterminated trailing command is not counted
  opt_ptr[ncmd + 1] = nopt + 1;
  for (int c = 1; c <= ncmd; c++) {
    if (char* p = std::strchr(cmd[c], ';')) *p = 0;
    for (int o = opt_ptr[c]; o < opt_ptr[c + 1]; o++)
      if (char* p = std::strchr(opt[o], ';')) *p = 0;
  }
}

void print_sim_params() {
  char l[2048];
  std::cout << std::endl << "+------------------------+" << std::endl << "| Simulation Parameters: |" << std::endl
            << "+------------------------+" << std::endl;
  if (sim_type == 1) std::cout << "Simulation Type                   : Underdamped" << std::endl;
  else if (sim_type == 2) std::cout << "Simulation Type                   : Overdamped" << std::endl;
  else {
    std::cerr << "UNRECOGNIZED SIMULATION TYPE!" << std::endl;
    std::exit(-1);
  }
  std::snprintf(l, sizeof l, "Simulation Temperature            : %.3f", T); std::cout << l << std::endl;
  std::snprintf(l, sizeof l, "Start Time Step                   : %.0lf", istep_restart); std::cout << l << std::endl;
  std::snprintf(l, sizeof l, "Final Time Step                   : %.0lf", nstep); std::cout << l << std::endl;
  std::snprintf(l, sizeof l, "Output Frequency                  : %d", nup); std::cout << l << std::endl;
  std::snprintf(l, sizeof l, "Friction Coefficient              : %.0e", zeta); std::cout << l << std::endl;
  std::snprintf(l, sizeof l, "PBC Box Length                    : %.1f", boxl); std::cout << l << std::endl;
  if (neighborlist == 1) {
    std::cout << "Long-range Cutoff Type            : Neighbor List" << std::endl;
    std::snprintf(l, sizeof l, "Neighbor List Update Frequency    : %d", nnlup); std::cout << l << std::endl;
  } else if (celllist == 1) {
    std::cout << "Long-range Cutoff Type            : Cell List" << std::endl;
    std::snprintf(l, sizeof l, "Cell List Update Frequency        : %d", nnlup); std::cout << l << std::endl;
    std::snprintf(l, sizeof l, "Number of Cells Each Dimension    : %.0lf", ncell); std::cout << l << std::endl;
  } else {
    std::cout << "Long-range Cutoff Type            : None" << std::endl;
  }
  std::cout << std::endl;
}

// "idx x y z" at %f (io.cpp:156-178), or at %.17g under `set restart_format lossless;`
// so that a restart reads back the exact state.
static void write_idx_xyz(const char* fname, const coord* a) {
  FILE* f = std::fopen(fname, "w");
  if (!f) return;
  const char* fmt = sopg_restart_lossless ? "%d %.17g %.17g %.17g\n" : "%d %f %f %f\n";
  for (int i = 1; i <= nbead; i++) std::fprintf(f, fmt, i, a[i].x, a[i].y, a[i].z);
  std::fclose(f);
}

static void read_idx_xyz(const char* fname, coord* a) {
  std::ifstream in(fname);
  if (!in) {
    std::cerr << "CANNOT OPEN " << fname << std::endl;
    std::exit(-1);
  }
  char buf[1024];
  for (int i = 1; i <= nbead; i++) {
    buf[0] = 0;
    in.getline(buf, sizeof buf);
    char* t = std::strtok(buf, " ");
    double v[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 3; c++) {
      t = std::strtok(nullptr, " ");
      v[c] = t ? std::atof(t) : 0.0;
    }
    a[i].x = v[0]; a[i].y = v[1]; a[i].z = v[2];
  }
}

void save_coords(char* fname, char* fname2) {
  write_idx_xyz(fname, pos);
  write_idx_xyz(fname2, unc_pos);
}

void load_coords(char* fname, char* fname2) {
  read_idx_xyz(fname, pos);
  read_idx_xyz(fname2, unc_pos);
}

void load_vels(char* fname) { read_idx_xyz(fname, vel); }
void save_unccoords(char* fname) { write_idx_xyz(fname, unc_pos); }
void save_vels(char* fname) { write_idx_xyz(fname, vel); }

void record_traj(char* fname, char* fname2) {
  FILE* f = std::fopen(fname, "a");
  FILE* g = std::fopen(fname2, "a");
  for (int i = 1; i <= nbead; i++) {
    if (f) std::fprintf(f, "%f %f %f\n", pos[i].x, pos[i].y, pos[i].z);
    if (g) std::fprintf(g, "%f %f %f\n", unc_pos[i].x, unc_pos[i].y, unc_pos[i].z);
  }
  if (f) std::fclose(f);
  if (g) std::fclose(g);
}
