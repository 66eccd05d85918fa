// jms_rng.cpp -- CRng (interface of the reference's rng.cpp:30-89)
#include "jms/rng.h"

#include <ctime>

CRng::CRng() { seed(0); }

// explicit seed, used as given (0 included), as rng.cpp:37-41
CRng::CRng(int s) : m_seed(s) { m_eng.seed((std::mt19937::result_type)s); }

CRng::~CRng() {}

void CRng::seed(int s) {
  m_seed = (s == 0) ? (int)::time(nullptr) : s;
  // only the engine is re-seeded; a normal deviate cached by the polar method
  // survives, as in the reference (rng.cpp:50-59)
  m_eng.seed((std::mt19937::result_type)m_seed);
}

int CRng::get_seed(void) { return m_seed; }

int CRng::rand() { return m_uni(m_eng); }

double CRng::unirand(double rmin, double rmax) {
  return rmin + ((double)rand() / (double)((unsigned long)JMS_RAND_MAXROLL)) * (rmax - rmin);
}

double CRng::unirand_oub(double rmin, double rmax) {
  return rmin + ((double)rand() / (double)((unsigned long)JMS_RAND_MAXROLL + 1)) * (rmax - rmin);
}

double CRng::normrand(double mu, double sdev) { return mu + sdev * m_norm(m_eng); }

int CRng::poisrand(double mu) {
  std::poisson_distribution<int> pd(mu);
  return pd(m_eng);
}
