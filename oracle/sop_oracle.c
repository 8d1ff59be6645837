/*
 * sop_oracle.c -- CPU restatement of the SOP-CPU hot path (TEST INFRASTRUCTURE).
 *
 * Oracle only: imported by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker / CPU baseline; never by the product.
 * Each function names the reference lines it restates.  Expressions keep the
 * reference's evaluation order exactly (same operands, same association), since
 * the point is bit-identical FP64 results under gcc -O3 without -march.
 */
#define _POSIX_C_SOURCE 200809L
#include "sop_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- helpers */

/* global.cpp:223-227: round half away from zero via floor/ceil. */
double orc_rnd(double x) { return (x > 0) ? floor(x + 0.5) : ceil(x - 0.5); }

/* Power selection for the angular terms (energy.cpp:180, 382).  0: glibc pow,
 * exactly as the reference.  1: the correctly rounded d^6 / d^8 of
 * orc_pow_cr() -- the value the device kernels compute; glibc's pow differs
 * from it by one ulp on ~0.08% of inputs, so this switch lets tests separate
 * that single, known source of difference from everything else. */
int orc_pow_mode = 0;

/* Exact product a*b = p + e (fma is exact). */
static inline void two_prod(double a, double b, double* p, double* e) { *p = a * b; *e = fma(a, b, -*p); }
/* Square of the double-double (h,l), renormalised. */
static inline void dd_sq(double h, double l, double* oh, double* ol) {
  double p, e, c, s;
  two_prod(h, h, &p, &e);
  c = h * l;
  e += c + c;
  s = p + e;
  *ol = e - (s - p);
  *oh = s;
}
/* Correctly rounded d^k for k in {6, 8} via double-double squaring (relative
 * error of the unrounded value < 2^-100, so only exact near-midpoints could
 * round differently). */
double orc_pow_cr(double d, int k) {
  double h2, l2, h4, l4, h8, l8, p, e, s;
  dd_sq(d, 0.0, &h2, &l2);
  dd_sq(h2, l2, &h4, &l4);
  if (k == 8) { dd_sq(h4, l4, &h8, &l8); return h8; }
  two_prod(h4, h2, &p, &e);
  e += h4 * l2 + l4 * h2;
  s = p + e;
  return s;
}
static inline double angpow(double d, double k) { return orc_pow_mode ? orc_pow_cr(d, (int)k) : pow(d, k); }

/* min-image of one component: d -= boxl*rnd(d/boxl) (e.g. energy.cpp:287-289). */
static inline double mimg(double d, double boxl) { d -= boxl * orc_rnd(d / boxl); return d; }

/* ---------------------------------------------------------------- RNG */
/* Ran_Gen (random_generator.h:15-36, random_generator.cpp:8-17, 80-146):
 * Numerical Recipes ran2 (L'Ecuyer two-MLCG + 32-slot Bays-Durham shuffle)
 * and gasdev (polar Box-Muller with one cached deviate). */
enum { R_IM1 = 2147483563, R_IM2 = 2147483399, R_IA1 = 40014, R_IA2 = 40692,
       R_IQ1 = 53668, R_IQ2 = 52774, R_IR1 = 12211, R_IR2 = 3791, R_NTAB = 32,
       R_IMM1 = R_IM1 - 1, R_NDIV = 1 + R_IMM1 / R_NTAB };
static const double R_EPS = 3.0e-16;

void orc_rng_seed(orc_rng* g, int seed) {
  g->iset = 0; g->idum2 = 123456789; g->iy = 0; g->gset = 0.0;
  memset(g->iv, 0, sizeof g->iv);
  g->mseed = seed;
}

double orc_ran2(orc_rng* g) {
  const double rnmx = 1.0 - R_EPS;
  const double am = 1.0 / (double)(2147483563);
  int j, k;
  double t;
  if (g->mseed <= 0) {
    g->mseed = (g->mseed == 0 ? 1 : -g->mseed);
    g->idum2 = g->mseed;
    for (j = R_NTAB + 7; j >= 0; j--) {
      k = g->mseed / R_IQ1;
      g->mseed = R_IA1 * (g->mseed - k * R_IQ1) - k * R_IR1;
      if (g->mseed < 0) g->mseed += R_IM1;
      if (j < R_NTAB) g->iv[j] = g->mseed;
    }
    g->iy = g->iv[0];
  }
  k = g->mseed / R_IQ1;
  g->mseed = R_IA1 * (g->mseed - k * R_IQ1) - k * R_IR1;
  if (g->mseed < 0) g->mseed += R_IM1;
  k = g->idum2 / R_IQ2;
  g->idum2 = R_IA2 * (g->idum2 - k * R_IQ2) - k * R_IR2;
  if (g->idum2 < 0) g->idum2 += R_IM2;
  j = g->iy / R_NDIV;
  g->iy = g->iv[j] - g->idum2;
  g->iv[j] = g->mseed;
  if (g->iy < 1) g->iy += R_IMM1;
  if ((t = am * g->iy) > rnmx) return rnmx;
  return t;
}

double orc_gasdev(orc_rng* g) {
  double fac, rsq, v1, v2;
  if (g->iset == 0) {
    do {
      v1 = 2.0 * orc_ran2(g) - 1.0;
      v2 = 2.0 * orc_ran2(g) - 1.0;
      rsq = v1 * v1 + v2 * v2;
    } while (rsq >= 1.0 || rsq == 0.0);
    fac = sqrt(-2.0 * log(rsq) / rsq);
    g->gset = v1 * fac;
    g->iset = 1;
    return v2 * fac;
  }
  g->iset = 0;
  return g->gset;
}

/* ---------------------------------------------------------------- state */

static void* xcalloc(size_t n, size_t sz) { void* p = calloc(n ? n : 1, sz); return p; }

static void free_lists(orc_sys* s) {
  free(s->ai); free(s->aj); free(s->ait); free(s->ajt);
  free(s->ar); free(s->ar2); free(s->ar6); free(s->ar12);
  free(s->ri); free(s->rj); free(s->rit); free(s->rjt);
  free(s->nai); free(s->naj); free(s->nait); free(s->najt);
  free(s->nar); free(s->nar2); free(s->nar6); free(s->nar12);
  free(s->nri); free(s->nrj); free(s->nrit); free(s->nrjt);
  free(s->pai); free(s->paj); free(s->pait); free(s->pajt);
  free(s->par); free(s->par2); free(s->par6); free(s->par12);
  free(s->pri); free(s->prj); free(s->prit); free(s->prjt);
}

/* init_lj (params.cpp:496-543): master, neighbor and pair lists sized by the
 * header counts. */
static void alloc_lists(orc_sys* s, int na, int nr) {
  free_lists(s);
  s->ncon_att = na; s->ncon_rep = nr;
#define A_I(p, n) p = (int*)xcalloc((size_t)(n) + 1, sizeof(int))
#define A_D(p, n) p = (double*)xcalloc((size_t)(n) + 1, sizeof(double))
  A_I(s->ai, na); A_I(s->aj, na); A_I(s->ait, na); A_I(s->ajt, na);
  A_D(s->ar, na); A_D(s->ar2, na); A_D(s->ar6, na); A_D(s->ar12, na);
  A_I(s->ri, nr); A_I(s->rj, nr); A_I(s->rit, nr); A_I(s->rjt, nr);
  A_I(s->nai, na); A_I(s->naj, na); A_I(s->nait, na); A_I(s->najt, na);
  A_D(s->nar, na); A_D(s->nar2, na); A_D(s->nar6, na); A_D(s->nar12, na);
  A_I(s->nri, nr); A_I(s->nrj, nr); A_I(s->nrit, nr); A_I(s->nrjt, nr);
  A_I(s->pai, na); A_I(s->paj, na); A_I(s->pait, na); A_I(s->pajt, na);
  A_D(s->par, na); A_D(s->par2, na); A_D(s->par6, na); A_D(s->par12, na);
  A_I(s->pri, nr); A_I(s->prj, nr); A_I(s->prit, nr); A_I(s->prjt, nr);
#undef A_I
#undef A_D
}

static void alloc_beads(orc_sys* s, int n) {
  free(s->pos); free(s->unc_pos); free(s->vel); free(s->force); free(s->incr);
  s->nbead = n;
  s->pos = (orc_vec*)xcalloc((size_t)n + 1, sizeof(orc_vec));
  s->unc_pos = (orc_vec*)xcalloc((size_t)n + 1, sizeof(orc_vec));
  s->vel = (orc_vec*)xcalloc((size_t)n + 1, sizeof(orc_vec));
  s->force = (orc_vec*)xcalloc((size_t)n + 1, sizeof(orc_vec));
  s->incr = (orc_vec*)xcalloc((size_t)n + 1, sizeof(orc_vec));
}

static void update_coeffs_dynamics_underdamped(orc_sys* s) {
  /* params.cpp:21-27 */
  s->sim_type = 1;
  s->h = 2.5e-3;
  s->halfh = s->h / 2.0;
  s->a1 = s->h * (1.0 - s->zeta * s->halfh);
  s->a2 = s->h * s->halfh;
  s->a3 = (1.0 - s->h * s->zeta / 2.0 + (s->h * s->zeta) * (s->h * s->zeta) / 4.0) / s->h;
  s->a4 = s->halfh * (1.0 - s->h * s->zeta / 2.0);
}

orc_sys* orc_create(void) {
  orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
  int i, j;
  /* type tables: global.cpp:60-92 */
  static const double catt[3][3] = {{0, 0, 0}, {0, 0.7, 0.8}, {0, 0.8, 1.0}};
  static const double crep[3][3] = {{0, 0, 0}, {0, 1.0, 1.0}, {0, 1.0, 1.0}};
  static const double fatt[3][3] = {{0, 0, 0}, {0, -12.0 * 1.0, -12.0 * 0.8}, {0, -12.0 * 0.8, -12.0 * 0.7}};
  static const double frep[3][3] = {{0, 0, 0}, {0, -6.0 * 1.0, -6.0 * 1.0}, {0, -6.0 * 1.0, -6.0 * 1.0}};
  static const double srep[3][3] = {{0, 0, 0}, {0, 3.8, 5.4}, {0, 5.4, 7.0}};
  static const double rnat[3][3] = {{0, 0, 0}, {0, 8.0, 11.0}, {0, 11.0, 14.0}};
  memcpy(s->coeff_att, catt, sizeof catt);
  memcpy(s->coeff_rep, crep, sizeof crep);
  memcpy(s->force_coeff_att, fatt, sizeof fatt);
  memcpy(s->force_coeff_rep, frep, sizeof frep);
  memcpy(s->sigma_rep, srep, sizeof srep);
  memcpy(s->rcut_nat, rnat, sizeof rnat);
  /* alloc_arrays defaults: params.cpp:268-399 */
  s->k_bnd = 20.0; s->R0 = 2.0; s->R0sq = s->R0 * s->R0;
  s->e_bnd_coeff = s->k_bnd * s->R0sq / 2.0;
  s->sigma_ss = 3.5; s->sigma_ss6 = pow(s->sigma_ss, 6.0); s->epsilon_ss = 1.0;
  s->e_ang_ss_coeff = s->epsilon_ss * s->sigma_ss6;
  s->f_ang_ss_coeff = 6.0 * s->e_ang_ss_coeff;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      s->sigma_rep2[i][j] = s->sigma_rep[i][j] * s->sigma_rep[i][j];
      s->sigma_rep6[i][j] = s->sigma_rep2[i][j] * s->sigma_rep2[i][j] * s->sigma_rep2[i][j];
      s->sigma_rep12[i][j] = s->sigma_rep6[i][j] * s->sigma_rep6[i][j];
    }
  s->run = 1; orc_rng_seed(&s->rng, -100 - s->run);
  s->T = 0.6; s->neighborlist = 0; s->celllist = 0;
  s->boxl = 500.0; s->ncell = 55.0; s->lcell = s->boxl / s->ncell;
  s->zeta = 5.0e-2; s->nstep = 5e7; s->nup = 1000; s->nnlup = 50;
  s->h = 2.5e-3; s->halfh = s->h / 2.0;
  s->a1 = s->h * (1.0 - s->zeta * s->halfh);
  s->a2 = s->h * s->halfh;
  s->a3 = (1.0 - s->h * s->zeta / 2.0 + (s->h * s->zeta) * (s->h * s->zeta) / 4.0) / s->h;
  s->a4 = s->halfh * (1.0 - s->h * s->zeta / 2.0);
  s->a5 = s->h / s->zeta;
  s->sim_type = 1;
  alloc_lists(s, 0, 0);
  alloc_beads(s, 0);
  return s;
}

void orc_destroy(orc_sys* s) {
  if (!s) return;
  free_lists(s);
  free(s->pos); free(s->unc_pos); free(s->vel); free(s->force); free(s->incr);
  free(s->ibnd); free(s->jbnd); free(s->pdb_dist);
  free(s->iang); free(s->jang); free(s->kang); free(s->pdb_ang);
  free(s);
}

/* set_params (params.cpp:12-101).  File-name keys are accepted and ignored. */
int orc_set(orc_sys* s, const char* key, const char* val) {
  if (!strcmp(key, "dynamics")) {
    if (!strcmp(val, "underdamped")) update_coeffs_dynamics_underdamped(s);
    else if (!strcmp(val, "overdamped")) { s->sim_type = 2; s->h = 0.02; s->a5 = s->h / s->zeta; }
  } else if (!strcmp(key, "temp")) { s->T = atof(val);
  } else if (!strcmp(key, "nstep")) { s->nstep = atof(val);
  } else if (!strcmp(key, "istep_restart")) { s->istep_restart = atof(val);
  } else if (!strcmp(key, "nup")) { s->nup = atoi(val);
  } else if (!strcmp(key, "run")) { s->run = atoi(val); s->rng.mseed = -100 - s->run;
  } else if (!strcmp(key, "cutofftype")) {
    if (!strcmp(val, "neighborlist")) s->neighborlist = 1;
    else if (!strcmp(val, "celllist")) s->celllist = 1;
  } else if (!strcmp(key, "nnlup")) { s->nnlup = atoi(val);
  } else if (!strcmp(key, "boxl")) { s->boxl = atof(val);
  } else if (!strcmp(key, "ncell")) { s->ncell = atof(val); s->lcell = s->boxl / s->ncell;
  } else if (!strcmp(key, "restart")) { s->restart = !strcmp(val, "on");
  } else if (!strcmp(key, "rgen_restart")) { s->rgen_restart = !strcmp(val, "on");
  } else if (!strcmp(key, "t_step")) {
    s->h = atof(val);
    s->halfh = s->h / 2.0;
    s->a1 = s->h * (1.0 - s->zeta * s->halfh);
    s->a2 = s->h * s->halfh;
    s->a3 = (1.0 - s->h * s->zeta / 2.0 + (s->h * s->zeta) * (s->h * s->zeta) / 4.0) / s->h;
    s->a4 = s->halfh * (1.0 - s->h * s->zeta / 2.0);
    s->a5 = s->h / s->zeta;
  } else if (!strcmp(key, "zeta")) {
    /* params.cpp:92-98: resets h but NOT halfh/a2 (order-sensitive). */
    if (s->sim_type == 1) s->h = 2.5e-3; else if (s->sim_type == 2) s->h = 0.02;
    s->zeta = atof(val);
    s->a1 = s->h * (1.0 - s->zeta * s->halfh);
    s->a3 = (1.0 - s->h * s->zeta / 2.0 + (s->h * s->zeta) * (s->h * s->zeta) / 4.0) / s->h;
    s->a4 = s->halfh * (1.0 - s->h * s->zeta / 2.0);
    s->a5 = s->h / s->zeta;
  } else {
    return 1; /* unknown / file-name key: ignored like the reference */
  }
  return 0;
}

/* Line reader matching `in.getline(line,2048)` + strtok(line," "). */
static char* next_tok(char* line) { return strtok(line, " "); }

int orc_load(orc_sys* s, const char* kind, const char* path) {
  char line[2048];
  char* t;
  int i;
  FILE* f = fopen(path, "r");
  if (!f) return -1;
#define GETLINE() do { if (!fgets(line, sizeof line, f)) line[0] = 0; \
    line[strcspn(line, "\n")] = 0; } while (0)
  if (!strcmp(kind, "bonds")) {           /* params.cpp:128-147 */
    GETLINE(); t = next_tok(line); t = strtok(NULL, " ");
    s->nbnd = atoi(t);
    free(s->ibnd); free(s->jbnd); free(s->pdb_dist);
    s->ibnd = (int*)xcalloc(s->nbnd + 1, sizeof(int));
    s->jbnd = (int*)xcalloc(s->nbnd + 1, sizeof(int));
    s->pdb_dist = (double*)xcalloc(s->nbnd + 1, sizeof(double));
    for (i = 1; i <= s->nbnd; i++) {
      GETLINE();
      t = next_tok(line); s->ibnd[i] = atoi(t);
      t = strtok(NULL, " "); s->jbnd[i] = atoi(t);
      t = strtok(NULL, " "); s->pdb_dist[i] = atof(t);
    }
  } else if (!strcmp(kind, "angles")) {   /* params.cpp:148-169 */
    GETLINE(); t = next_tok(line); t = strtok(NULL, " ");
    s->nang = atoi(t);
    free(s->iang); free(s->jang); free(s->kang); free(s->pdb_ang);
    s->iang = (int*)xcalloc(s->nang + 1, sizeof(int));
    s->jang = (int*)xcalloc(s->nang + 1, sizeof(int));
    s->kang = (int*)xcalloc(s->nang + 1, sizeof(int));
    s->pdb_ang = (double*)xcalloc(s->nang + 1, sizeof(double));
    for (i = 1; i <= s->nang; i++) {
      GETLINE();
      t = next_tok(line); s->iang[i] = atoi(t);
      t = strtok(NULL, " "); s->jang[i] = atoi(t);
      t = strtok(NULL, " "); s->kang[i] = atoi(t);
      t = strtok(NULL, " "); s->pdb_ang[i] = atof(t);
    }
  } else if (!strcmp(kind, "vdw")) {      /* params.cpp:170-234 */
    int na, nr, ib, jb, it, jt, ia = 0, ir = 0;
    double r;
    GETLINE(); t = next_tok(line); t = strtok(NULL, " "); na = atoi(t);
    t = strtok(NULL, " "); t = strtok(NULL, " "); nr = atoi(t);
    alloc_lists(s, na, nr);
    for (i = 1; i <= na + nr; i++) {
      GETLINE();
      t = next_tok(line); ib = atoi(t);
      t = strtok(NULL, " "); jb = atoi(t);
      t = strtok(NULL, " "); r = atof(t);
      t = strtok(NULL, " "); it = atoi(t);
      t = strtok(NULL, " "); jt = atoi(t);
      if (r < s->rcut_nat[it][jt]) {
        ia++;
        s->ai[ia] = ib; s->aj[ia] = jb; s->ait[ia] = it; s->ajt[ia] = jt;
        s->ar[ia] = r;
        s->ar2[ia] = r * r;
        s->ar6[ia] = s->ar2[ia] * s->ar2[ia] * s->ar2[ia];
        s->ar12[ia] = s->ar6[ia] * s->ar6[ia];
        /* the loader seeds the pair list with the full master list (params.cpp:210-218) */
        s->nil_att++;
        s->pai[s->nil_att] = ib; s->paj[s->nil_att] = jb;
        s->pait[s->nil_att] = it; s->pajt[s->nil_att] = jt;
        s->par[s->nil_att] = r; s->par2[s->nil_att] = s->ar2[ia];
        s->par6[s->nil_att] = s->ar6[ia]; s->par12[s->nil_att] = s->ar12[ia];
      } else {
        ir++;
        s->ri[ir] = ib; s->rj[ir] = jb; s->rit[ir] = it; s->rjt[ir] = jt;
        s->nil_rep++;
        s->pri[s->nil_rep] = ib; s->prj[s->nil_rep] = jb;
        s->prit[s->nil_rep] = it; s->prjt[s->nil_rep] = jt;
      }
    }
  } else if (!strcmp(kind, "init")) {     /* params.cpp:235-259 */
    int n;
    GETLINE(); t = next_tok(line); t = strtok(NULL, " "); n = atoi(t);
    alloc_beads(s, n);
    for (i = 1; i <= n; i++) {
      GETLINE();
      t = next_tok(line);
      t = strtok(NULL, " "); s->pos[i].x = atof(t); s->unc_pos[i].x = s->pos[i].x;
      t = strtok(NULL, " "); s->pos[i].y = atof(t); s->unc_pos[i].y = s->pos[i].y;
      t = strtok(NULL, " "); s->pos[i].z = atof(t); s->unc_pos[i].z = s->pos[i].z;
    }
  } else {
    fclose(f);
    return 1;
  }
#undef GETLINE
  fclose(f);
  return 0;
}

/* ---------------------------------------------------------------- lists */

/* neighbor_list.cpp:6-117 */
void orc_update_neighbor_list(orc_sys* s) {
  double dx, dy, dz, d2, rcut, rcut2;
  int i, ib, jb, it, jt;
  const orc_vec* u = s->unc_pos;
  s->nnl_att = 0;
  s->nnl_rep = 0;
  for (i = 1; i <= s->ncon_att; i++) {
    ib = s->ai[i]; jb = s->aj[i]; it = s->ait[i]; jt = s->ajt[i];
    dx = mimg(u[jb].x - u[ib].x, s->boxl);
    dy = mimg(u[jb].y - u[ib].y, s->boxl);
    dz = mimg(u[jb].z - u[ib].z, s->boxl);
    d2 = dx * dx + dy * dy + dz * dz;
    rcut = 3.2 * s->ar[i];
    rcut2 = rcut * rcut;
    if (d2 < rcut2) {
      int k = ++s->nnl_att;
      s->nai[k] = ib; s->naj[k] = jb; s->nait[k] = it; s->najt[k] = jt;
      s->nar[k] = s->ar[i]; s->nar2[k] = s->ar2[i]; s->nar6[k] = s->ar6[i]; s->nar12[k] = s->ar12[i];
    }
  }
  for (i = 1; i <= s->ncon_rep; i++) {
    ib = s->ri[i]; jb = s->rj[i]; it = s->rit[i]; jt = s->rjt[i];
    dx = mimg(u[jb].x - u[ib].x, s->boxl);
    dy = mimg(u[jb].y - u[ib].y, s->boxl);
    dz = mimg(u[jb].z - u[ib].z, s->boxl);
    d2 = dx * dx + dy * dy + dz * dz;
    rcut = 3.2 * s->sigma_rep[it][jt];
    rcut2 = rcut * rcut;
    if (d2 < rcut2) {
      int k = ++s->nnl_rep;
      s->nri[k] = ib; s->nrj[k] = jb; s->nrit[k] = it; s->nrjt[k] = jt;
    }
  }
}

/* cell_list.cpp:6-98: cell-adjacency filter of the master lists. */
static int cell_adjacent(const orc_sys* s, int ib, int jb) {
  const orc_vec* u = s->unc_pos;
  double imx = floor(u[ib].x / s->lcell), imy = floor(u[ib].y / s->lcell), imz = floor(u[ib].z / s->lcell);
  double jmx = floor(u[jb].x / s->lcell), jmy = floor(u[jb].y / s->lcell), jmz = floor(u[jb].z / s->lcell);
  double cx = imx - jmx, cy = imy - jmy, cz = imz - jmz;
  cx -= s->ncell * orc_rnd(cx / s->ncell);
  cy -= s->ncell * orc_rnd(cy / s->ncell);
  cz -= s->ncell * orc_rnd(cz / s->ncell);
  return cx >= -1.0 && cx <= 1.0 && cy >= -1.0 && cy <= 1.0 && cz >= -1.0 && cz <= 1.0;
}

void orc_update_cell_list(orc_sys* s) {
  int i;
  s->nnl_att = 0;
  s->nnl_rep = 0;
  for (i = 1; i <= s->ncon_att; i++) {
    if (cell_adjacent(s, s->ai[i], s->aj[i])) {
      int k = ++s->nnl_att;
      s->nai[k] = s->ai[i]; s->naj[k] = s->aj[i]; s->nait[k] = s->ait[i]; s->najt[k] = s->ajt[i];
      s->nar[k] = s->ar[i]; s->nar2[k] = s->ar2[i]; s->nar6[k] = s->ar6[i]; s->nar12[k] = s->ar12[i];
    }
  }
  for (i = 1; i <= s->ncon_rep; i++) {
    if (cell_adjacent(s, s->ri[i], s->rj[i])) {
      int k = ++s->nnl_rep;
      s->nri[k] = s->ri[i]; s->nrj[k] = s->rj[i]; s->nrit[k] = s->rit[i]; s->nrjt[k] = s->rjt[i];
    }
  }
}

/* sop.cpp:473-553 */
void orc_update_pair_list(orc_sys* s) {
  double dx, dy, dz, d2, rcut, rcut2;
  int i, ib, jb, it, jt;
  const orc_vec* u = s->unc_pos;
  s->nil_att = 0;
  s->nil_rep = 0;
  for (i = 1; i <= s->nnl_att; i++) {
    ib = s->nai[i]; jb = s->naj[i]; it = s->nait[i]; jt = s->najt[i];
    dx = mimg(u[jb].x - u[ib].x, s->boxl);
    dy = mimg(u[jb].y - u[ib].y, s->boxl);
    dz = mimg(u[jb].z - u[ib].z, s->boxl);
    d2 = dx * dx + dy * dy + dz * dz;
    rcut = 2.5 * s->nar[i];
    rcut2 = rcut * rcut;
    if (d2 < rcut2) {
      int k = ++s->nil_att;
      s->pai[k] = ib; s->paj[k] = jb; s->pait[k] = it; s->pajt[k] = jt;
      s->par[k] = s->nar[i]; s->par2[k] = s->nar2[i]; s->par6[k] = s->nar6[i]; s->par12[k] = s->nar12[i];
    }
  }
  for (i = 1; i <= s->nnl_rep; i++) {
    ib = s->nri[i]; jb = s->nrj[i]; it = s->nrit[i]; jt = s->nrjt[i];
    dx = mimg(u[jb].x - u[ib].x, s->boxl);
    dy = mimg(u[jb].y - u[ib].y, s->boxl);
    dz = mimg(u[jb].z - u[ib].z, s->boxl);
    d2 = dx * dx + dy * dy + dz * dz;
    rcut = 2.5 * s->sigma_rep[it][jt];
    rcut2 = rcut * rcut;
    if (d2 < rcut2) {
      int k = ++s->nil_rep;
      s->pri[k] = ib; s->prj[k] = jb; s->prit[k] = it; s->prjt[k] = jt;
    }
  }
}

/* ---------------------------------------------------------------- forces */

void orc_clear_forces(orc_sys* s) {   /* energy.cpp:35-45 */
  int i;
  for (i = 1; i <= s->nbead; i++) { s->force[i].x = 0.0; s->force[i].y = 0.0; s->force[i].z = 0.0; }
}

static double next_gauss(orc_sys* s) {
  if (s->noise) return s->noise[s->noise_pos++];
  return orc_gasdev(&s->rng);
}

void orc_random_force(orc_sys* s) {   /* energy.cpp:449-465 */
  int i;
  double var = sqrt(2.0 * s->T * s->zeta / s->h);
  for (i = 1; i <= s->nbead; i++) {
    s->force[i].x += var * next_gauss(s);
    s->force[i].y += var * next_gauss(s);
    s->force[i].z += var * next_gauss(s);
  }
}

void orc_fene_forces(orc_sys* s) {    /* energy.cpp:402-447 */
  int i, ib, jb;
  double dx, dy, dz, d, dev, dev2, fx, fy, fz, temp;
  const orc_vec* u = s->unc_pos;
  for (i = 1; i <= s->nbnd; i++) {
    ib = s->ibnd[i]; jb = s->jbnd[i];
    dx = mimg(u[jb].x - u[ib].x, s->boxl);
    dy = mimg(u[jb].y - u[ib].y, s->boxl);
    dz = mimg(u[jb].z - u[ib].z, s->boxl);
    d = sqrt(dx * dx + dy * dy + dz * dz);
    dev = d - s->pdb_dist[i];
    dev2 = dev * dev;
    temp = -s->k_bnd * dev / d / (1.0 - dev2 / s->R0sq);
    fx = temp * dx; fy = temp * dy; fz = temp * dz;
    s->force[ib].x -= fx; s->force[ib].y -= fy; s->force[ib].z -= fz;
    s->force[jb].x += fx; s->force[jb].y += fy; s->force[jb].z += fz;
  }
}

void orc_angular_forces(orc_sys* s) { /* energy.cpp:354-400 */
  int i, ib, kb;
  double dx, dy, dz, d, d8, co1, fx, fy, fz;
  const orc_vec* u = s->unc_pos;
  for (i = 1; i <= s->nang; i++) {
    ib = s->iang[i]; kb = s->kang[i];
    dx = mimg(u[kb].x - u[ib].x, s->boxl);
    dy = mimg(u[kb].y - u[ib].y, s->boxl);
    dz = mimg(u[kb].z - u[ib].z, s->boxl);
    d = sqrt(dx * dx + dy * dy + dz * dz);
    d8 = angpow(d, 8.0);
    co1 = s->f_ang_ss_coeff / d8;
    fx = co1 * dx; fy = co1 * dy; fz = co1 * dz;
    s->force[ib].x -= fx; s->force[ib].y -= fy; s->force[ib].z -= fz;
    s->force[kb].x += fx; s->force[kb].y += fy; s->force[kb].z += fz;
  }
}

void orc_vdw_forces(orc_sys* s) {     /* energy.cpp:259-352 */
  const double tol = 1.0e-7;
  /* rep_tol is only assigned inside the att loop (energy.cpp:292); with an
   * empty att pair list the reference reads it uninitialised.  The oracle
   * defines that case as 0.0 (no rep pair is skipped). */
  double rep_tol = 0.0;
  int i, ib, jb, it, jt;
  double dx, dy, dz, d2, d6, d12, co1, fx, fy, fz;
  const orc_vec* u = s->unc_pos;
  for (i = 1; i <= s->nil_att; i++) {
    ib = s->pai[i]; jb = s->paj[i]; it = s->pait[i]; jt = s->pajt[i];
    dx = mimg(u[jb].x - u[ib].x, s->boxl);
    dy = mimg(u[jb].y - u[ib].y, s->boxl);
    dz = mimg(u[jb].z - u[ib].z, s->boxl);
    d2 = dx * dx + dy * dy + dz * dz;
    rep_tol = s->sigma_rep2[it][jt] * tol;
    /* the skip test indexes the MASTER array with the pair-list index (energy.cpp:293) */
    if (d2 < tol * s->ar2[i]) continue;
    d6 = d2 * d2 * d2;
    d12 = d6 * d6;
    co1 = s->force_coeff_att[it][jt] / d2 * ((s->par12[i] / d12) - (s->par6[i] / d6));
    fx = co1 * dx; fy = co1 * dy; fz = co1 * dz;
    s->force[ib].x += fx; s->force[ib].y += fy; s->force[ib].z += fz;
    s->force[jb].x -= fx; s->force[jb].y -= fy; s->force[jb].z -= fz;
  }
  for (i = 1; i <= s->nil_rep; i++) {
    ib = s->pri[i]; jb = s->prj[i]; it = s->prit[i]; jt = s->prjt[i];
    dx = mimg(u[jb].x - u[ib].x, s->boxl);
    dy = mimg(u[jb].y - u[ib].y, s->boxl);
    dz = mimg(u[jb].z - u[ib].z, s->boxl);
    d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < rep_tol) continue;
    d6 = d2 * d2 * d2;
    d12 = d6 * d6;
    co1 = s->force_coeff_rep[it][jt] / d2 *
          (2.0 * s->sigma_rep12[it][jt] / d12 + s->sigma_rep6[it][jt] / d6);
    fx = co1 * dx; fy = co1 * dy; fz = co1 * dz;
    s->force[ib].x += fx; s->force[ib].y += fy; s->force[ib].z += fz;
    s->force[jb].x -= fx; s->force[jb].y -= fy; s->force[jb].z -= fz;
  }
}

/* force_eval (energy.cpp:21-33) with the default set_forces table
 * (energy.cpp:78-113, global.cpp:200-203): random, FENE, angular, vdw. */
void orc_force_eval(orc_sys* s) {
  orc_clear_forces(s);
  orc_random_force(s);
  orc_fene_forces(s);
  orc_angular_forces(s);
  orc_vdw_forces(s);
}

/* ---------------------------------------------------------------- energies */

void orc_fene_energy(orc_sys* s) {    /* energy.cpp:115-152 */
  int i, ib, jb;
  double dx, dy, dz, d, dev;
  const orc_vec* u = s->unc_pos;
  s->e_bnd = 0.0;
  for (i = 1; i <= s->nbnd; i++) {
    ib = s->ibnd[i]; jb = s->jbnd[i];
    dx = mimg(u[jb].x - u[ib].x, s->boxl);
    dy = mimg(u[jb].y - u[ib].y, s->boxl);
    dz = mimg(u[jb].z - u[ib].z, s->boxl);
    d = sqrt(dx * dx + dy * dy + dz * dz);
    dev = d - s->pdb_dist[i];
    s->e_bnd += log1p(-dev * dev / s->R0sq);
  }
  s->e_bnd *= -s->e_bnd_coeff;
}

void orc_angular_energy(orc_sys* s) { /* energy.cpp:154-187 */
  int i, ib, kb;
  double dx, dy, dz, d, d6;
  const orc_vec* u = s->unc_pos;
  s->e_ang_ss = 0.0;
  for (i = 1; i <= s->nang; i++) {
    ib = s->iang[i]; kb = s->kang[i];
    dx = mimg(u[kb].x - u[ib].x, s->boxl);
    dy = mimg(u[kb].y - u[ib].y, s->boxl);
    dz = mimg(u[kb].z - u[ib].z, s->boxl);
    d = sqrt(dx * dx + dy * dy + dz * dz);
    d6 = angpow(d, 6.0);
    s->e_ang_ss += s->e_ang_ss_coeff / d6;
  }
}

void orc_vdw_energy(orc_sys* s) {     /* energy.cpp:189-257 */
  int i, ib, jb, it, jt;
  double dx, dy, dz, d2, d6, d12;
  const orc_vec* u = s->unc_pos;
  s->e_vdw_rr = 0.0; s->e_vdw_rr_att = 0.0; s->e_vdw_rr_rep = 0.0;
  for (i = 1; i <= s->nil_att; i++) {
    ib = s->pai[i]; jb = s->paj[i]; it = s->pait[i]; jt = s->pajt[i];
    dx = mimg(u[jb].x - u[ib].x, s->boxl);
    dy = mimg(u[jb].y - u[ib].y, s->boxl);
    dz = mimg(u[jb].z - u[ib].z, s->boxl);
    d2 = dx * dx + dy * dy + dz * dz;
    d6 = d2 * d2 * d2;
    d12 = d6 * d6;
    /* precedence as coded (energy.cpp:224): the coefficient scales only the r^-12 term */
    s->e_vdw_rr_att += s->coeff_att[it][jt] * (s->par12[i] / d12) - 2.0 * (s->par6[i] / d6);
  }
  for (i = 1; i <= s->nil_rep; i++) {
    ib = s->pri[i]; jb = s->prj[i]; it = s->prit[i]; jt = s->prjt[i];
    dx = mimg(u[jb].x - u[ib].x, s->boxl);
    dy = mimg(u[jb].y - u[ib].y, s->boxl);
    dz = mimg(u[jb].z - u[ib].z, s->boxl);
    d2 = dx * dx + dy * dy + dz * dz;
    d6 = d2 * d2 * d2;
    d12 = d6 * d6;
    s->e_vdw_rr_rep += s->coeff_rep[it][jt] * (s->sigma_rep12[it][jt] / d12 + s->sigma_rep6[it][jt] / d6);
  }
  s->e_vdw_rr = s->e_vdw_rr_att + s->e_vdw_rr_rep;
}

void orc_energy_eval(orc_sys* s) {    /* energy.cpp:6-19 with set_potential order */
  orc_fene_energy(s);
  orc_angular_energy(s);
  orc_vdw_energy(s);
  s->rna_etot = s->e_bnd + s->e_ang_ss + s->e_vdw_rr;
}

/* ---------------------------------------------------------------- dynamics */

void orc_underdamped_iteration(orc_sys* s) {  /* sop.cpp:282-329 */
  const double eps = 1.0e-5;
  int i;
  for (i = 1; i <= s->nbead; i++) {
    orc_vec* in = &s->incr[i];
    in->x = s->a1 * s->vel[i].x + s->a2 * s->force[i].x;
    in->y = s->a1 * s->vel[i].y + s->a2 * s->force[i].y;
    in->z = s->a1 * s->vel[i].z + s->a2 * s->force[i].z;
    s->pos[i].x += in->x; s->pos[i].y += in->y; s->pos[i].z += in->z;
    s->pos[i].x -= s->boxl * orc_rnd(s->pos[i].x / s->boxl);
    s->pos[i].y -= s->boxl * orc_rnd(s->pos[i].y / s->boxl);
    s->pos[i].z -= s->boxl * orc_rnd(s->pos[i].z / s->boxl);
    s->unc_pos[i].x += in->x; s->unc_pos[i].y += in->y; s->unc_pos[i].z += in->z;
  }
  orc_force_eval(s);
  if (s->T < eps) return;
  for (i = 1; i <= s->nbead; i++) {
    s->vel[i].x = s->a3 * s->incr[i].x + s->a4 * s->force[i].x;
    s->vel[i].y = s->a3 * s->incr[i].y + s->a4 * s->force[i].y;
    s->vel[i].z = s->a3 * s->incr[i].z + s->a4 * s->force[i].z;
  }
}

void orc_overdamped_iteration(orc_sys* s) {   /* sop.cpp:331-363 */
  int i;
  for (i = 1; i <= s->nbead; i++) {
    orc_vec* in = &s->incr[i];
    in->x = s->a5 * s->force[i].x;
    in->y = s->a5 * s->force[i].y;
    in->z = s->a5 * s->force[i].z;
    s->unc_pos[i].x += in->x; s->unc_pos[i].y += in->y; s->unc_pos[i].z += in->z;
    s->pos[i].x += in->x; s->pos[i].y += in->y; s->pos[i].z += in->z;
    s->pos[i].x -= s->boxl * orc_rnd(s->pos[i].x / s->boxl);
    s->pos[i].y -= s->boxl * orc_rnd(s->pos[i].y / s->boxl);
    s->pos[i].z -= s->boxl * orc_rnd(s->pos[i].z / s->boxl);
  }
  orc_force_eval(s);
}

void orc_calculate_observables(orc_sys* s) {  /* sop.cpp:200-280 */
  int i, j, ib, jb;
  double dx, dy, dz, d, sumvsq;
  float r_ij;
  const orc_vec* u = s->unc_pos;
  s->contct_nat = 0;
  for (i = 1; i <= s->ncon_att; i++) {
    ib = s->ai[i]; jb = s->aj[i];
    r_ij = (float)s->ar[i];
    dx = mimg(u[ib].x - u[jb].x, s->boxl);
    dy = mimg(u[ib].y - u[jb].y, s->boxl);
    dz = mimg(u[ib].z - u[jb].z, s->boxl);
    d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d / r_ij < 1.25) s->contct_nat++;
  }
  s->Q = (double)(s->contct_nat) / s->ncon_att;
  s->rgsq = 0.0;
  for (i = 1; i <= s->nbead - 1; i++) {
    for (j = i + 1; j <= s->nbead; j++) {
      dx = mimg(u[i].x - u[j].x, s->boxl);
      dy = mimg(u[i].y - u[j].y, s->boxl);
      dz = mimg(u[i].z - u[j].z, s->boxl);
      s->rgsq += (dx * dx + dy * dy + dz * dz);
    }
  }
  s->rgsq /= (double)(s->nbead * s->nbead);
  if (s->sim_type == 1) {
    sumvsq = 0.0;
    for (i = 1; i <= s->nbead; i++)
      sumvsq += s->vel[i].x * s->vel[i].x + s->vel[i].y * s->vel[i].y + s->vel[i].z * s->vel[i].z;
    s->kinT = sumvsq / (3.0 * (double)(s->nbead));
  } else if (s->sim_type == 2) {
    sumvsq = 0.0;
    for (i = 1; i <= s->nbead; i++)
      sumvsq += s->incr[i].x * s->incr[i].x + s->incr[i].y * s->incr[i].y + s->incr[i].z * s->incr[i].z;
    sumvsq *= s->zeta / (2.0 * s->h);
    s->kinT = sumvsq / (3.0 * (double)(s->nbead));
  }
}

/* ---------------------------------------------------------------- accessors */

int orc_get_int(orc_sys* s, const char* n) {
  if (!strcmp(n, "nbead")) return s->nbead;
  if (!strcmp(n, "nbnd")) return s->nbnd;
  if (!strcmp(n, "nang")) return s->nang;
  if (!strcmp(n, "ncon_att")) return s->ncon_att;
  if (!strcmp(n, "ncon_rep")) return s->ncon_rep;
  if (!strcmp(n, "nnl_att")) return s->nnl_att;
  if (!strcmp(n, "nnl_rep")) return s->nnl_rep;
  if (!strcmp(n, "nil_att")) return s->nil_att;
  if (!strcmp(n, "nil_rep")) return s->nil_rep;
  if (!strcmp(n, "contct_nat")) return s->contct_nat;
  if (!strcmp(n, "nup")) return s->nup;
  if (!strcmp(n, "nnlup")) return s->nnlup;
  if (!strcmp(n, "sim_type")) return s->sim_type;
  if (!strcmp(n, "neighborlist")) return s->neighborlist;
  if (!strcmp(n, "celllist")) return s->celllist;
  if (!strcmp(n, "restart")) return s->restart;
  return -999999;
}

double orc_get_double(orc_sys* s, const char* n) {
  if (!strcmp(n, "e_bnd")) return s->e_bnd;
  if (!strcmp(n, "e_ang_ss")) return s->e_ang_ss;
  if (!strcmp(n, "e_vdw_rr")) return s->e_vdw_rr;
  if (!strcmp(n, "e_vdw_rr_att")) return s->e_vdw_rr_att;
  if (!strcmp(n, "e_vdw_rr_rep")) return s->e_vdw_rr_rep;
  if (!strcmp(n, "rna_etot")) return s->rna_etot;
  if (!strcmp(n, "Q")) return s->Q;
  if (!strcmp(n, "rgsq")) return s->rgsq;
  if (!strcmp(n, "kinT")) return s->kinT;
  if (!strcmp(n, "T")) return s->T;
  if (!strcmp(n, "h")) return s->h;
  if (!strcmp(n, "halfh")) return s->halfh;
  if (!strcmp(n, "zeta")) return s->zeta;
  if (!strcmp(n, "a1")) return s->a1;
  if (!strcmp(n, "a2")) return s->a2;
  if (!strcmp(n, "a3")) return s->a3;
  if (!strcmp(n, "a4")) return s->a4;
  if (!strcmp(n, "a5")) return s->a5;
  if (!strcmp(n, "boxl")) return s->boxl;
  if (!strcmp(n, "ncell")) return s->ncell;
  if (!strcmp(n, "lcell")) return s->lcell;
  if (!strcmp(n, "nstep")) return s->nstep;
  if (!strcmp(n, "istep_restart")) return s->istep_restart;
  if (!strcmp(n, "e_ang_ss_coeff")) return s->e_ang_ss_coeff;
  if (!strcmp(n, "f_ang_ss_coeff")) return s->f_ang_ss_coeff;
  return -999999.0;
}

static orc_vec* vec_array(orc_sys* s, int which) {
  switch (which) {
    case 0: return s->pos;
    case 1: return s->unc_pos;
    case 2: return s->vel;
    case 3: return s->force;
    case 4: return s->incr;
    default: return NULL;
  }
}

void orc_get_coords(orc_sys* s, int which, double* out) {
  orc_vec* a = vec_array(s, which);
  int i;
  for (i = 1; i <= s->nbead; i++) { out[3 * i - 3] = a[i].x; out[3 * i - 2] = a[i].y; out[3 * i - 1] = a[i].z; }
}

void orc_set_coords(orc_sys* s, int which, const double* in) {
  orc_vec* a = vec_array(s, which);
  int i;
  for (i = 1; i <= s->nbead; i++) { a[i].x = in[3 * i - 3]; a[i].y = in[3 * i - 2]; a[i].z = in[3 * i - 1]; }
}

int orc_get_list(orc_sys* s, int which, int* oi, int* oj) {
  int n = 0, k;
  int *a = NULL, *b = NULL;
  switch (which) {
    case 0: n = s->nnl_att; a = s->nai; b = s->naj; break;
    case 1: n = s->nnl_rep; a = s->nri; b = s->nrj; break;
    case 2: n = s->nil_att; a = s->pai; b = s->paj; break;
    case 3: n = s->nil_rep; a = s->pri; b = s->prj; break;
    case 4: n = s->ncon_att; a = s->ai; b = s->aj; break;
    case 5: n = s->ncon_rep; a = s->ri; b = s->rj; break;
    default: return -1;
  }
  if (oi && oj) for (k = 1; k <= n; k++) { oi[k - 1] = a[k]; oj[k - 1] = b[k]; }
  return n;
}

void orc_set_noise(orc_sys* s, const double* noise) { s->noise = noise; s->noise_pos = 0; }
void orc_gasdev_n(orc_sys* s, double* out, long n) { long i; for (i = 0; i < n; i++) out[i] = orc_gasdev(&s->rng); }
double orc_sys_ran2(orc_sys* s) { return orc_ran2(&s->rng); }
void orc_set_seed(orc_sys* s, int seed) { s->rng.mseed = seed; }
