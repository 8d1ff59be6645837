// ran_gen.h -- the reference's random-number generator interface
// (random_generator.h:4-36): Numerical Recipes ran2 (L'Ecuyer two-MLCG with a
// 32-slot Bays-Durham shuffle) and gasdev (polar Box-Muller, one cached
// deviate), with the same state file (random_generator.cpp:24-70).  The class
// name and members match so that host code written against the reference's
// `generator` compiles unchanged; the stream is bit-identical (tests pin it
// against the reference's known-answer fixtures).
#ifndef SOPG_RAN_GEN_H
#define SOPG_RAN_GEN_H

class Ran_Gen {
 public:
  Ran_Gen();
  ~Ran_Gen();
  void save_state();
  void restart();
  void set_fname(char*);
  void set_seed(int);
  double gasdev();
  double ran2();

 private:
  char fname[2048];
  int mseed;
  int iset;
  double gset;
  int idum2;
  int iy;
  int iv[32];
};

#endif
