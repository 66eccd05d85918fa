// jms/rng.h -- random number generator of the library (interface of the
// reference's CRng, rng.h:22-65): a Mersenne twister (std::mt19937) behind
// rand() in [0, 0x7FFFFFFF], unirand() in the closed range, normrand() and
// poisrand().  Seeding with 0 seeds from the clock (rng.cpp:50-59).
#pragma once
#include <random>

#define JMS_RAND_MAXROLL 0x7FFFFFFF

class CRng {
 public:
  CRng();
  explicit CRng(int seed);
  ~CRng();
  // re-seed the generator (0 = from the clock)
  void seed(int seed = 0);
  // last seed used
  int get_seed(void);
  // integer pseudo random number in [0, JMS_RAND_MAXROLL]
  int rand();
  // uniform pseudo random number in the closed range [rmin, rmax]
  double unirand(double rmin = 0.0, double rmax = 1.0);
  // uniform pseudo random number in the half-open range [rmin, rmax)
  double unirand_oub(double rmin = 0.0, double rmax = 1.0);
  // normal pseudo random number with mean mu and standard deviation sdev
  double normrand(double mu = 0.0, double sdev = 1.0);
  // Poisson pseudo random number with mean mu
  int poisrand(double mu = 1.0);

 private:
  std::mt19937 m_eng;
  std::uniform_int_distribution<int> m_uni{0, JMS_RAND_MAXROLL};
  std::normal_distribution<double> m_norm{0.0, 1.0};
  int m_seed = 0;
};
