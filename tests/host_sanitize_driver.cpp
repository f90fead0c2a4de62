// host_sanitize_driver.cpp -- test infrastructure: drives the host set-up code
// of the engine (nlm_host.cpp: setLengths, resetGrids, setDispersion,
// setPoling, setPump and its chirp / delay, the Raman spectrum, the host
// transforms and the constructors' error paths; src/_NonlinearMedium.cpp:43-222
// and the solver constructors) without touching a GPU, so that a build with
// AddressSanitizer + UndefinedBehaviorSanitizer checks it for out-of-bounds
// accesses, leaks and undefined behaviour.  Built and run by
// tests/test_oracle_sanitizers.py; exits 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "nlm_host.hpp"

using nlm::cd;

static int failures = 0;

static void check_true(bool ok, const std::string& what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL %s\n", what.c_str());
    failures++;
  }
}

static std::vector<cd> seeded(int n, unsigned seed) {
  std::vector<cd> v(n);
  for (int k = 0; k < n; k++) {
    seed = seed * 1103515245u + 12345u;
    const double a = ((seed >> 8) & 0xffff) / 65536.0 - 0.5;
    seed = seed * 1103515245u + 12345u;
    const double b = ((seed >> 8) & 0xffff) / 65536.0 - 0.5;
    v[k] = cd(a, b);
  }
  return v;
}

struct Case {
  const char* name;
  std::vector<double> num;  // numeric constructor arguments in declaration order
  bool pumped;
};

static void exercise(const Case& c, unsigned N, int profile, int pulse, const std::vector<double>& poling,
                     unsigned nZ) {
  const std::string what = std::string(c.name) + " N=" + std::to_string(N);
  std::vector<double> profileZ;
  auto m = nlm::makeSolver(c.name, c.num, {}, pulse, profile, poling, N, 100, nZ, profileZ);
  check_true((unsigned)m->nFreqs() == N && (unsigned)m->nZSteps() == nZ, what + " sizes");
  check_true(m->omega().size() == N && m->tau().size() == N, what + " grids");
  double w = 0;
  for (double x : m->omega()) w += std::fabs(x);
  check_true(std::isfinite(w), what + " omega finite");
  if (c.pumped) {
    m->setPump(nlm::PulseType::Sinc, 0.3, 0.2, 0);
    m->setPump(nlm::PulseType::Sech, -0.1, 0.0, 0);
    m->setPump(seeded((int)N, 3u + N), 0.2, 0.1, 0);
    bool threw = false;
    try {
      m->setPump(seeded((int)N + 2, 5u), 0, 0, 0);  // wrong length
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    check_true(threw, what + " custom pump length check");
  }
  if (!poling.empty()) {
    std::vector<double> prof(nZ);
    for (unsigned i = 0; i < nZ; i++) prof[i] = std::cos(0.3 * i);
    m->setPolingProfile(prof);
  }
}

int main() {
  const double inf = INFINITY;
  const std::vector<Case> cases = {
      {"Chi3", {1.0, 2.0, -1.0, 0.01, inf, 10.0, 0.1}, true},
      {"Chi2PDC", {1.0, 2.0, 0.5, -0.4, 0.0, 0.3, 0.01, 0.02, 1.5, 0.7, 10.0, 0.2, 0.1}, true},
      {"Chi2SFG", {1.0, 2.0, 3.0, 0.5, -0.4, 0.2, 0.0, 0.3, -0.2, 0.01, 0.02, -0.01, 1.5, inf, 10.0, 0.0, 0.0},
       true},
      {"Chi2AFC", {1.0, 2.0, 3.0, 0.5, -0.4, 0.2, 0.0, 0.3, -0.2, 0.01, 0.02, -0.01, 1.5, -2.0, inf, 10.0, 0.0,
                   0.0},
       true},
      {"Chi2SHG", {1.0, 2.0, 3.0, -0.4, 0.2, 0.3, -0.2, 0.02, -0.01, 1.5, 0.7, 10.0}, false},
      {"Chi3GNLSE", {1.0, 0.5, 100.0, 0.18, 0.21, 0.0122, 0.032, 0.096, -1.0, 0.0, inf, 10.0}, false},
  };
  const std::vector<double> poling = {0.3, 0.45, 0.25};
  for (const Case& c : cases) {
    const bool poled = std::string(c.name) != "Chi3" && std::string(c.name) != "Chi2AFC" &&
                       std::string(c.name) != "Chi3GNLSE";
    for (int profile = 0; profile < 3; profile++) {
      exercise(c, 64, profile, 0, poled ? poling : std::vector<double>{}, 16);
      exercise(c, 30, profile, 2, std::vector<double>{}, 7);  // mixed-radix length
    }
  }
  // the constructors' error paths
  const std::vector<double> pdc = cases[1].num;
  auto expect_throw = [&](unsigned N, std::vector<double> num, const std::vector<double>& pol, const char* what) {
    bool threw = false;
    try {
      auto m = nlm::makeSolver("Chi2PDC", num, {}, 0, 0, pol, N, 100, 0, {});
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    check_true(threw, what);
  };
  expect_throw(63, pdc, {}, "odd nFreqs");
  std::vector<double> neg = pdc;
  neg[0] = -1.0;
  expect_throw(64, neg, {}, "negative length");
  expect_throw(64, pdc, {0.5, -0.1}, "invalid domain length");
  // host transforms: forward then inverse (scaled) returns the input
  for (int n : {2, 6, 30, 64, 210, 1000}) {
    std::vector<cd> x = seeded(n, 9u + n), y(n), z(n);
    nlm::host_fft(x.data(), y.data(), n, false);
    nlm::host_fft(y.data(), z.data(), n, true);
    double err = 0;
    for (int k = 0; k < n; k++) err = std::fmax(err, std::abs(z[k] - x[k]));
    check_true(err < 1e-13, "host transform round trip n=" + std::to_string(n));
  }
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("host sanitizer driver: all checks passed\n");
  return 0;
}
