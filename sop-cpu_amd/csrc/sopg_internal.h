// sopg_internal.h -- device-side data layout of a context (a ragged batch of
// chains) shared by the kernels and the C ABI implementation.
//
// Layout in HBM (SoA, FP64, all chains concatenated; global bead index g =
// chain offset + local bead index):
//   unc_pos x/y/z  two buffers (the step kernel reads one, writes the next)
//   pos x/y/z      two buffers (wrapped copy, output only)
//   vel, force x/y/z
//   incr x/y/z     two buffers, paired with the position buffers: incr[b] is the
//                  increment that moved the beads into positions b
//   per-bead incidence lists ([beg, end) ranges, in the reference's list order):
//     bonds, angles, native (attractive) contacts, and -- for explicit
//     non-native master lists -- non-native contacts
//   per-bead non-native neighbor slots: cap words per bead + count
//   master native-contact arrays (sigma powers, cutoffs, NL / PL flags)
#pragma once
#include <cstdint>

namespace sopg {

constexpr double kTol = 1.0e-7;  // energy.cpp:271

struct Tables {
  double fatt[3][3], frep[3][3], catt[3][3], crep[3][3];
  double s2[3][3], s6[3][3], s12[3][3];
  double nlrc2[3][3], plrc2[3][3];  // (3.2 sigma_rep)^2, (2.5 sigma_rep)^2
  double k_bnd, R0sq, e_bnd_coeff, e_ang_ss_coeff, f_ang_ss_coeff;
  double thr_rep;                   // max over types of sigma_rep2 * tol
  // Candidate lists (neighbor-list criterion, runs only): the non-native
  // pairs within 3.2 sigma + skin of the positions u0 of their build.  While
  // no bead of the chain has moved more than skin/2 from u0, every pair of
  // the neighbor list is among them, and a rebuild filters them instead of
  // scanning the grid.
  int cand_on;                      // skin > 0
  double candc2[3][3];              // (3.2 sigma_rep + skin)^2
  double cand_disp2;                // (skin/2)^2, less a margin far above rounding
  double cand_reach;                // bound on a valid candidate's distance: max 3.2 sigma + 2 skin + 1
};

struct ChainP {
  int off, n;          // first global bead, bead count
  int att_off, natt;   // native master entries [att_off, att_off+natt)
  int cutoff;          // SOPG_CUTOFF_*
  int rep_explicit;    // 1: non-native master list stored (CSR); 0: implicit
  int nnlup, nup;
  int phase, iup0;     // reference inlup / iup before the first step of a run
  int vel_on;          // T >= 1e-5 (sop.cpp:286, 316)
  int noise_mode;      // SOPG_NOISE_*
  double boxl, qbox, lcell, ncell, qncell;
  double a1, a2, a3, a4, a5, var, T;
  double thr_att;      // tol * max lj_nat_pdb_dist2 of the chain
  uint64_t seed, noise_base;  // Philox key, force evaluations done so far
  int sim_type;        // 1 underdamped, 2 overdamped
  double gcell;        // binned neighbor-list rebuild: grid edge >= largest 3.2 sigma_rep of the chain
  double gcand;        // the same for the candidate-list build: >= largest 3.2 sigma_rep + skin
  int resident;        // 1 while sopg_run advances the chain in the resident kernel (the
                       // per-step kernels then skip it)
};

// Per-chain grid of the binned rebuild, recomputed at every rebuild.  A chain
// uses it only when its extent on every axis is at most the period less the
// grid edge (ncell - 2 cells for the cell list): then every pair the criterion
// keeps lies in adjacent grid cells.
constexpr int kGridDim = 32;                       // cells per axis, at most
constexpr int kGridCells = kGridDim * kGridDim * kGridDim;
// What a list rebuild does for the chain (GridP::work, set by k_grid for every
// chain of the launch, 0 for a chain not due).
enum { W_BIN = 1,      // beads binned into the grid, k_rebuild_grid runs
       W_CAND = 2,     // ... and writes the candidate list instead of the neighbor list
       W_FILTER = 4,   // the neighbor list is filtered from the candidate list
       W_SCAN = 8 };   // the full scan of k_rebuild
struct GridP {
  double ox, oy, oz;   // origin (bead coordinate, or cell index for the cell list)
  double inv;          // 1 / grid cell edge (1 for the cell list: cells are the reference's)
  int nx, ny, nz;
  int work;            // W_* of the current rebuild
  int cand_valid;      // the candidate list holds the pairs within 3.2 sigma + skin of u0
  int periodic;        // the grid covers the box with wrapped coordinates (cells adjacent modulo n)
};

// Word of one native-contact incidence that the step kernel reads:
//   bits  0-24 partner bead (chain-local, 0-based)
//   bits 25-29 t = itype | jtype << 2 | side << 4 (side 0: this bead is i)
constexpr int kAttTShift = 25;

// Per-bead reduction fields (red[g * kRed + field]).
enum { R_EBND = 0, R_EANG, R_EATT, R_EREP, R_CONT, R_RG, R_V2, R_NLA, R_PLA, R_NLR, R_PLR, kRed = 12 };

// Bead tiles of the binned list-rebuild launches: with free beads in a
// rebuild's bead range, the launch takes a table of the tiles (kBeadTile
// consecutive beads, tile k = beads [k, k + 1) * kBeadTile) that hold a bead
// of an active chain.
constexpr int kBeadTile = 16;

// Records of sopg_run_noalloc are staged in a file-scope device array of
// this many doubles (4 MiB; one copy per GPU, held until the process ends).
constexpr size_t kScratchDoubles = (size_t)1 << 19;

struct Dev {
  // Added to every kernel's relative step argument: 0 for direct launches;
  // for a replayed graph segment the run-relative step at which it starts
  // (the segment's kernels are captured with steps relative to it).
  const int* relb;
  int nb;    // bead high-water mark: beads [0, nb) are live or free (chain -1)
  int cap;   // non-native slots per bead
  double *ux[2], *uy[2], *uz[2];
  double *px[2], *py[2], *pz[2];
  double *vx, *vy, *vz, *fx, *fy, *fz;
  double *ix[2], *iy[2], *iz[2];  // incr[b]: the increment that produced positions b
  int* chain;   // chain id of each bead, -1 for a free bead
  int8_t* type;
  // bonds, per incidence: partner << 1 | side (0: bead is i, 1: bead is j), r0
  int *bond_beg, *bond_end, *bond_pq; double* bond_r0;
  // angles (i, k) pairs, per incidence: partner << 1 | side (0: bead is i, 1: bead is k)
  int *ang_beg, *ang_end, *ang_pq;
  // natives: code = master << 1 | side; partner = the other bead (global).
  // Per incidence also what the step kernel reads, so that a bead's lanes load
  // it contiguously: the packed partner / type word (kAttTShift) and the
  // pair's native distance sigma, from which sigma^6, sigma^12 and the
  // pair-list cutoff (2.5 sigma)^2 follow with the host's own arithmetic.
  int *att_beg, *att_end, *att_code, *att_partner; uint8_t* att_flag;  // flag: bit0 NL, bit1+p PL(parity p)
  uint32_t* att_w; double* att_r;
  // master native entries (am_t = itype | jtype << 2; am_q0 = incidence of the first bead)
  int *am_i, *am_j, *am_q0; int8_t* am_t;
  double *am_r2, *am_nlrc2; float* am_rf;
  // explicit non-natives
  int *rep_beg, *rep_end, *rep_code, *rm_i, *rm_j; int8_t* rm_t;
  // non-native neighbor slots
  uint32_t* slot; int* slot_k; int* cnt;
  // candidate lists (cap words per bead, slot format without pair-list bits)
  // and the positions they were built at
  uint32_t* cand; int* ccnt;
  double *u0x, *u0y, *u0z;
  float4* fpos;  // single-precision copy of the positions at a candidate-list rebuild (w unused)
  double *cx, *cy, *cz;  // cell coordinates (cell list)
  double* red;           // per-bead reduction fields
  int* err;              // bit0: slot overflow
  // The (run-relative) step of the latest rebuild in which some chain needed
  // the binned kernels [0], the candidate filter [1], the full scan [2]: a
  // kernel of a rebuild in which no chain needs it returns at once.
  int* plan;
  ChainP* cp;
  // binned rebuild: per-chain grid (GridP), kGridCells cells per chain
  GridP* grid;
  int *cell_start, *cell_fill;             // per cell: first position and bead count
  int *bead_cell, *cell_beads;             // per bead: its cell, and beads sorted by cell
  int bitmap_words;                        // ceil(max chain size / 32)
};

}  // namespace sopg
