// sop_ctrl.cpp -- command dispatch and the simulation loops of the host driver
// (sop.cpp:55-198, 365-471), with the hot loop resident on the GPU.
//
// The loop structure, counters and outputs are the reference's: initial list
// build, restart coordinates loaded after it, a first energy + force
// evaluation, then per step a list rebuild every nnlup steps, the pair list,
// one iteration, and every nup steps an `update.out` line, trajectory frames,
// coordinate / velocity files and the generator state.  Between two output
// steps the state stays on the device: sopg_run advances the chunk of steps in
// one call, with the noise stream drawn from `generator` in consumption order
// (or the device generator with `set noise device;`).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "sop_shim.h"
#include "sop_state.h"

void ex_cmds() {   // sop.cpp:55-74
  for (int i = 1; i <= ncmd; i++) {
    if (!std::strcmp(cmd[i], "load")) load(i);
    else if (!std::strcmp(cmd[i], "set")) set_params(i);
    else if (!std::strcmp(cmd[i], "run")) simulation_ctrl();
  }
}

void simulation_ctrl() {   // sop.cpp:76-91
  switch (sim_type) {
    case 1: underdamped_ctrl(); break;
    case 2: overdamped_ctrl(); break;
    default:
      std::cerr << "UNRECOGNIZED SIM_TYPE!" << std::endl;
      std::exit(-1);
  }
}

namespace {

// One `update.out` line and the periodic files (sop.cpp:173-185).
void write_outputs(FILE* out, double istep) {
  if (out)
    std::fprintf(out, "%.0lf %f %f %f %f %f %f %f %d %f\n", istep, T, kinT, e_bnd, e_ang_ss, e_vdw_rr, rna_etot, Q,
                 contct_nat, rgsq);
  if (out) std::fflush(out);
  record_traj(binfname, uncbinfname);
  save_coords(cfname, unccfname);
  save_vels(vfname);
  generator.save_state();
}

// Shared body of underdamped_ctrl / overdamped_ctrl.  `inlup_io` is the list
// counter the loop uses: a local starting at 1 (underdamped, sop.cpp:100) or
// the global inlup (overdamped, sop.cpp:428).
void run_ctrl(int& first_time, int& inlup_io) {
  double istep = 1.0;
  int iup = 1;
  FILE* out = std::fopen(ufname, "a");
  if (!restart && first_time) {
    for (int i = 1; i <= nbead; i++) {
      vel[i].x = vel[i].y = vel[i].z = 0.0;
      force[i].x = force[i].y = force[i].z = 0.0;
    }
  }
  print_sim_params();
  if (neighborlist == 1) {
    update_neighbor_list();
    update_pair_list();
  } else if (celllist == 1) {
    update_cell_list();
    update_pair_list();
  }
  set_potential();
  set_forces();
  if (restart) {
    load_coords(cfname, unccfname);
    if (sim_type == 1) load_vels(vfname);
    istep = istep_restart + 1.0;
  }
  if (rgen_restart) generator.restart();
  if (first_time) {
    energy_eval();
    force_eval();
  }
  if (binsave) {
    if (first_time && !rgen_restart) record_traj(binfname, uncbinfname);
    sopg_ctx* ctx = sopg_shim_context();
    sopg_shim_put_state();
    sopg_shim_check(sopg_set_step(ctx, 0, istep, inlup_io, iup), "sopg_set_step");
    std::vector<sopg_record> rec(1);
    while (istep <= nstep) {
      // steps up to and including the next output step (or the last step)
      long chunk = (long)((nup - iup % nup) % nup) + 1;
      const long left = (long)(nstep - istep) + 1;
      if (chunk > left) chunk = left;
      long nrec = 0;
      sopg_shim_check(sopg_run(ctx, chunk, sopg_shim_noise(chunk), rec.data(), 1, &nrec), "sopg_run");
      // counters as the reference leaves them after `chunk` steps (the list
      // counter advances in every cutoff mode, sop.cpp:156-165)
      if (inlup_io == 0) inlup_io = nnlup;
      inlup_io = (int)((inlup_io - 1 + chunk) % nnlup) + 1;
      iup = (int)((iup - 1 + chunk) % nup) + 1;
      istep += (double)chunk;
      if (nrec > 0) {
        sopg_shim_get_state();
        kinT = rec[0].kinT;
        e_bnd = rec[0].e_bnd;
        e_ang_ss = rec[0].e_ang_ss;
        e_vdw_rr = rec[0].e_vdw_rr;
        rna_etot = rec[0].rna_etot;
        system_etot = rna_etot;
        Q = rec[0].Q;
        contct_nat = rec[0].contct_nat;
        rgsq = rec[0].rgsq;
        write_outputs(out, rec[0].istep);
      }
    }
    sopg_shim_get_state();
    if (out) std::fclose(out);
  } else if (out) {
    std::fclose(out);
  }
  if (first_time) first_time = 0;
}

}  // namespace

void underdamped_ctrl() {   // sop.cpp:93-198
  static int first_time = 1;
  int inlup_local = 1;
  run_ctrl(first_time, inlup_local);
}

void overdamped_ctrl() {    // sop.cpp:365-471
  static int first_time = 1;
  run_ctrl(first_time, inlup);
}
