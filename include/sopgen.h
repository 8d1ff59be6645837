/*
 * sopgen.h -- synthetic SOP chains and the reference's input file formats.
 *
 * No counterpart in SOP-CPU, which ships no inputs: these produce the native
 * structures, unfolded starts and Maxwell velocities that the benchmarks and
 * tests run on, and write them in the formats the reference's loaders read
 * (params.cpp:128-259, io.cpp:180-242).  Exported from libsopg.so.
 */
#ifndef SOPGEN_H
#define SOPGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Self-avoiding walk: 3.8 A bonds, non-bonded pairs >= 4.0 A.  xyz: 3n doubles. */
int sopgen_saw(int n, uint64_t seed, double* xyz);
/* Unfolded start of a folding run: a SAW inside radius rmax (<= 0: unbounded)
 * that keeps every native pair of native_xyz (j-i>2, closer than 8 A there) at
 * least its native distance apart.  xyz: 3n doubles. */
int sopgen_unfolded(int n, uint64_t seed, const double* native_xyz, double rmax, double* xyz);
/* Native structure: SAW + Monte Carlo collapse into a sphere of 3.0*n^(1/3) A. */
int sopgen_native(int n, uint64_t seed, double* xyz);
/* Maxwell velocities at temperature kT (unit mass). */
int sopgen_maxwell(int n, double kT, uint64_t seed, double* vel);
/* Smallest non-bonded distance, largest |bond - 3.8|, largest radius. */
int sopgen_check(int n, const double* xyz, double* min_nonbonded, double* max_bond_dev, double* max_radius);
/* bonds / angles / vdw (every pair j-i>2, lexicographic) / init files at %.17g;
 * any path may be NULL.  Returns the number of native pairs (r < rcut_nat). */
int sopgen_write_topology(int n, const double* xyz, const char* bonds_path, const char* angles_path,
                          const char* vdw_path, const char* init_path, double rcut_nat);
/* Native contacts (pairs j-i>2 closer than rcut) in (i,j) order, 1-based, with
 * the distances the vdw file carries.  Returns the count (writes at most cap). */
long sopgen_natives(int n, const double* xyz, double rcut, int* out_i, int* out_j, double* out_r, long cap);
/* "idx x y z" lines (coordinates or velocities) at %.17g. */
int sopgen_write_xyz(int n, const double* xyz, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* SOPGEN_H */
