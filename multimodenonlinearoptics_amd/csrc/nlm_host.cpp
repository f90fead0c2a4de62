// nlm_host.cpp -- host side of the `nonlinearmedium` drop-in (see nlm_host.hpp).
//
// Set-up arithmetic (grids, dispersion, poling, pump envelopes, the scalar
// factors of each DiffEq) is written to evaluate the same expressions in the
// same order as the reference, so the tables handed to the device match the
// reference's values to the last bit wherever the C++ library does.
#include "nlm_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace nlm {

static constexpr cd I1(0.0, 1.0);
static const double kInf = std::numeric_limits<double>::infinity();

void check(int status) {
  if (status == NLM_OK) return;
  const char* m = nlm_last_error();
  if (status == NLM_INVALID_ARGUMENT) throw std::invalid_argument(m);
  throw std::runtime_error(m);
}

// ---------------------------------------------------------------------------
// set-up transforms (not on the propagation path)
// ---------------------------------------------------------------------------

void host_fft(const cd* in, cd* out, int n, bool inverse) {
  const long double pi = 3.141592653589793238462643383279502884L;
  const double sgn = inverse ? 1.0 : -1.0;
  std::vector<cd> a(in, in + n);
  if (n > 0 && (n & (n - 1)) == 0) {
    // iterative radix-2, bit-reversed input order
    for (int i = 1, j = 0; i < n; i++) {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
      std::vector<cd> w(len / 2);
      for (int k = 0; k < len / 2; k++) {
        long double ang = 2.0L * pi * k / len;
        w[k] = cd((double)cosl(ang), sgn * (double)sinl(ang));
      }
      for (int i = 0; i < n; i += len)
        for (int k = 0; k < len / 2; k++) {
          cd u = a[i + k], v = a[i + k + len / 2] * w[k];
          a[i + k] = u + v;
          a[i + k + len / 2] = u - v;
        }
    }
  } else {
    std::vector<cd> b(n);
    for (int k = 0; k < n; k++) {
      cd acc = 0;
      for (int j = 0; j < n; j++) {
        long double ang = 2.0L * pi * (long double)(((long long)j * k) % n) / n;
        acc += a[j] * cd((double)cosl(ang), sgn * (double)sinl(ang));
      }
      b[k] = acc;
    }
    a.swap(b);
  }
  if (inverse) {
    const double s = 1. / n;
    for (auto& x : a) x *= s;
  }
  std::copy(a.begin(), a.end(), out);
}

// ---------------------------------------------------------------------------
// construction: _NonlinearMedium::_NonlinearMedium, src/_NonlinearMedium.cpp:8-40
// ---------------------------------------------------------------------------

Medium::Medium(const Args& a) : model_(a.model), nS_((int)a.nS), nP_((int)a.nP) {
  double rayleighLength = a.rayleighLength;
  if (a.profile == IntensityProfile::Constant) rayleighLength = kInf;
  setLengths(a, rayleighLength);
  resetGrids(a.tPrecision, a.tMax);
  setDispersion(a);
  if (a.canBePoled) {
    setPoling(a.poling);
    if (!a.polingProfile.empty()) setPolingProfile(a.polingProfile);
  }
  profile_ = (zR_ != kInf ? a.profile : IntensityProfile::Constant);
  if (nP_ > 0) {
    envelope_.assign(nP_, std::vector<cd>(N_, cd(0, 0)));
    if (!a.customPump.empty()) Medium::setPump(a.customPump, a.chirp, a.delay);
    else Medium::setPump(a.pulseType, a.chirp, a.delay);
  }
  pumpHostStale_.assign(nP_, false);
  pumpDevStale_.assign(nP_, false);
}

Medium::~Medium() {
  if (plan_) nlm_plan_destroy(plan_);
}

FullyNonlinearMedium::FullyNonlinearMedium(const Args& a) : Medium(a) {
  dzp_ = 0;
  nZP_ = 0;
}

// _NonlinearMedium::setLengths, src/_NonlinearMedium.cpp:43-102 (plus the
// nZSteps override extension)
void Medium::setLengths(const Args& a, double rayleighLength) {
  bool negativeLength = (a.relativeLength <= 0 || rayleighLength <= 0);
  for (double nl : a.nl) negativeLength |= (nl <= 0);
  if (negativeLength) throw std::invalid_argument("Non-positive length scale");

  bool allNonUnit = true;
  for (double b : a.beta1) allNonUnit &= (std::abs(b) != 1);
  for (double b : a.beta2) allNonUnit &= (std::abs(b) != 1);
  for (double b : a.beta3) allNonUnit &= (std::abs(b) != 1);
  for (double b : a.beta1s) allNonUnit &= (std::abs(b) != 1);
  for (double b : a.beta2s) allNonUnit &= (std::abs(b) != 1);
  for (double b : a.beta3s) allNonUnit &= (std::abs(b) != 1);
  for (double nl : a.nl) allNonUnit &= (nl != 1);
  allNonUnit &= (a.relativeLength != 1);
  allNonUnit &= (rayleighLength != 1);
  if (allNonUnit) throw std::invalid_argument("No unit length scale provided: please normalize variables");

  z_ = a.relativeLength;
  auto absComp = [](double x, double y) { return std::abs(x) < std::abs(y); };
  auto mx = [&](const std::vector<double>& v) { return v.empty() ? 0 : *std::max_element(v.begin(), v.end(), absComp); };
  const double minDispLength =
      1 / std::abs(std::max({mx(a.beta2), mx(a.beta2s), mx(a.beta1), mx(a.beta1s), mx(a.beta3), mx(a.beta3s)}, absComp));
  const double minNl = a.nl.empty() ? 0 : *std::min_element(a.nl.begin(), a.nl.end());
  if (a.nZOverride > 0)
    nZ_ = (int)a.nZOverride;
  else
    nZ_ = (int)static_cast<unsigned>(a.zPrecision * z_ / std::min({1., minDispLength, rayleighLength, minNl}));
  nZP_ = 2 * nZ_ - 1;
  dz_ = z_ / (unsigned)(nZ_ - 1);
  dzp_ = z_ / (unsigned)(nZP_ - 1);
  nlStep_.resize(a.nl.size());
  for (size_t p = 0; p < a.nl.size(); p++) nlStep_[p] = 1.0 * I1 / a.nl[p] * dz_;
  zR_ = rayleighLength;
}

static std::vector<double> fftshift(const std::vector<double>& in) {
  const size_t n = in.size(), h = n / 2;
  std::vector<double> out(n);
  for (size_t i = 0; i < h; i++) out[i] = in[n - h + i];
  for (size_t i = 0; i < h; i++) out[n - h + i] = in[i];
  return out;
}

// _NonlinearMedium::resetGrids, src/_NonlinearMedium.cpp:105-145
void Medium::resetGrids(unsigned nFreqs, double tMax) {
  if (nFreqs % 2 != 0 || nFreqs == 0) throw std::invalid_argument("Invalid number of Frequencies");
  if (nZ_ == 0) throw std::invalid_argument("Zero steps");
  if (tMax <= 0) throw std::invalid_argument("Negative time span");
  N_ = (int)nFreqs;
  tMax_ = tMax;
  const int Nt = N_;
  tau_.resize(Nt);
  omega_.resize(Nt);
  const double sT = 2 * tMax / Nt, sW = -M_PI / tMax_;
  for (int i = 0; i < Nt; i++) {
    const double lin = (double)(-Nt / 2 + i);
    tau_[i] = sT * lin;
    omega_[i] = sW * lin;
  }
  tau_ = fftshift(tau_);
  omega_ = fftshift(omega_);
  // host copies of the grids are allocated on first use (the device holds the
  // working grids; a batch of many media never touches these)
  pumpFreq_.assign(nP_, Grid());
  pumpTime_.assign(nP_, Grid());
  signalFreq_.assign(nS_, Grid());
  signalTime_.assign(nS_, Grid());
}

void Medium::allocPump(unsigned q) {
  if (pumpFreq_[q].rows != nZP_ || pumpFreq_[q].cols != N_) pumpFreq_[q].setZero(nZP_, N_);
  if (pumpTime_[q].rows != nZP_ || pumpTime_[q].cols != N_) pumpTime_[q].setZero(nZP_, N_);
}

void Medium::allocSignal() {
  for (int m = 0; m < nS_; m++) {
    if (signalFreq_[m].rows != nZ_ || signalFreq_[m].cols != N_) signalFreq_[m].setZero(nZ_, N_);
    if (signalTime_[m].rows != nZ_ || signalTime_[m].cols != N_) signalTime_[m].setZero(nZ_, N_);
  }
}

// _NonlinearMedium::setDispersion, src/_NonlinearMedium.cpp:148-177
void Medium::setDispersion(const Args& a) {
  beta2_ = a.beta2;
  beta1_ = a.beta1;
  diffBeta0_ = a.diffBeta0;
  dispPump_.assign(nP_, std::vector<double>(N_));
  dispStepPump_.assign(nP_, std::vector<cd>(N_));
  for (int m = 0; m < nP_; m++)
    for (int k = 0; k < N_; k++) {
      const double w = omega_[k];
      dispPump_[m][k] = w * (a.beta1[m] + w * (0.5 * a.beta2[m] + w * a.beta3[m] / 6));
      dispStepPump_[m][k] = std::exp((1.0 * I1 * dzp_) * dispPump_[m][k]);
    }
  dispSign_.assign(nS_, std::vector<double>(N_));
  dispStepSign_.assign(nS_, std::vector<cd>(N_));
  for (int m = 0; m < nS_; m++)
    for (int k = 0; k < N_; k++) {
      const double w = omega_[k];
      dispSign_[m][k] = w * (a.beta1s[m] + w * (0.5 * a.beta2s[m] + w * a.beta3s[m] / 6));
      dispStepSign_[m][k] = std::exp((1.0 * I1 * dz_) * dispSign_[m][k]);
    }
}

// _NonlinearMedium::setPoling, src/_NonlinearMedium.cpp:470-503
void Medium::setPoling(const std::vector<double>& poling) {
  if (poling.size() <= 1) {
    poling_.assign(nZ_, 1.0);
  } else {
    for (double p : poling)
      if (p <= 0) throw std::invalid_argument("Poling contains invalid domain length");
    std::vector<double> dom(poling.size());
    dom[0] = poling[0];
    for (size_t i = 1; i < poling.size(); i++) dom[i] = poling[i] + dom[i - 1];
    const double scale = (unsigned)nZ_ / dom.back();
    for (auto& d : dom) d *= scale;
    poling_.assign(nZ_, 0.0);
    unsigned prevInd = 0;
    int direction = 1;
    for (size_t i = 0; i < dom.size(); i++) {
      const double currInd = dom[i];
      const unsigned currIndRound = static_cast<unsigned>(currInd);
      if (currInd < prevInd) throw std::invalid_argument("Poling period too small for simulation resolution");
      for (unsigned j = prevInd; j < currIndRound; j++) poling_[j] = direction;
      if (currIndRound < (unsigned)nZ_) poling_[currIndRound] = direction * (2 * std::abs(std::fmod(currInd, 1)) - 1);
      direction *= -1;
      prevInd = currIndRound + 1;
    }
  }
  coefDirty_ = true;
}

void Medium::setPolingProfile(const std::vector<double>& profile) {
  if (poling_.empty()) throw std::invalid_argument("this medium has no poling");
  if ((int)profile.size() != nZ_)
    throw std::invalid_argument("polingProfile length does not match number of z steps");
  poling_ = profile;
  coefDirty_ = true;
}

// chirp and delay phase of setPump, src/_NonlinearMedium.cpp:198-203, 216-221
void Medium::applyChirpDelay(unsigned q, double chirp, double delay) {
  if (chirp == 0 && delay == 0) return;
  std::vector<cd> tmp(N_);
  host_fft(envelope_[q].data(), tmp.data(), N_, false);
  for (int k = 0; k < N_; k++) {
    const double w = omega_[k];
    tmp[k] *= std::exp(1.0 * I1 * (beta1_[q] * delay + 0.5 * beta2_[q] * chirp * w) * w);
  }
  host_fft(tmp.data(), envelope_[q].data(), N_, true);
}

// _NonlinearMedium::setPump(PulseType, ...), src/_NonlinearMedium.cpp:180-204
void Medium::setPump(PulseType pulseType, double chirpLength, double delayLength, unsigned pumpIndex) {
  if ((int)pumpIndex >= nP_) throw std::invalid_argument("Invalid pump index");
  auto& e = envelope_[pumpIndex];
  for (int k = 0; k < N_; k++) {
    const double t = tau_[k];
    switch (pulseType) {
      default:
      case PulseType::Gaussian: e[k] = std::exp(-0.5 * (t * t)); break;
      case PulseType::Sech: e[k] = 1 / std::cosh(t); break;
      case PulseType::Sinc: e[k] = std::sin(t) / t; break;
    }
  }
  if (pulseType == PulseType::Sinc) e[0] = 1;
  applyChirpDelay(pumpIndex, chirpLength, delayLength);
}

// _NonlinearMedium::setPump(customPump, ...), src/_NonlinearMedium.cpp:207-222
void Medium::setPump(const std::vector<cd>& customPump, double chirpLength, double delayLength, unsigned pumpIndex) {
  if ((int)customPump.size() != N_)
    throw std::invalid_argument("Custom pump array length does not match number of frequency/time bins");
  if ((int)pumpIndex >= nP_) throw std::invalid_argument("Invalid pump index");
  envelope_[pumpIndex] = customPump;
  applyChirpDelay(pumpIndex, chirpLength, delayLength);
}

// _NonlinearMedium::setPump(other, ...), src/_NonlinearMedium.cpp:506-534
// (the pump grid is interpolated on the host from the other medium's signal
// grid, then uploaded before the next propagation)
void Medium::setPump(const Medium& otherC, unsigned modeIndex, double delayLength, unsigned pumpIndex) {
  Medium& other = const_cast<Medium&>(otherC);
  if (other.N_ != N_ || other.tMax_ != tMax_)
    throw std::invalid_argument("Medium does not have same time and frequency axes as this one");
  if ((int)modeIndex >= nS_) throw std::invalid_argument("Mode index larger than number of modes in medium");
  if ((int)pumpIndex >= nP_) throw std::invalid_argument("Invalid pump index");
  if (other.nZ_ < nZ_) throw std::invalid_argument("Medium does not have sufficient resolution to be used with this one");
  const double bd = beta1_[pumpIndex] * delayLength;
  if (other.signalOnDevice_ && other.plan_) {
    // the source grid is on the device: build the pump grid there
    // (the last pump time row keeps the device copy of its current contents)
    if (pumpDevStale_[pumpIndex] || !pumpComputed_) uploadPump(pumpIndex);
    check(nlm_pump_import(plan(), (int)pumpIndex, other.plan_, (int)modeIndex, dispPump_[pumpIndex].data(),
                          other.dispSign_[modeIndex].data(), omega_.data(), dzp_, other.dz_, bd, z_, other.z_));
    pumpHostStale_[pumpIndex] = true;
    pumpDevStale_[pumpIndex] = false;
    pumpComputed_ = true;
    return;
  }
  syncPumpToHost(pumpIndex);
  const Grid& oF = other.signalFreq(modeIndex);
  Grid& pF = pumpFreq_[pumpIndex];
  Grid& pT = pumpTime_[pumpIndex];
  for (unsigned i = 0; i < (unsigned)nZP_ - 1; i++) {
    const double j_ = i * (static_cast<double>(other.nZ_ - 1) / (nZP_ - 1));
    const unsigned j = static_cast<unsigned>(j_);
    const double frac = j_ - j;
    for (int k = 0; k < N_; k++) {
      const cd ph = (1.0 * I1 * (i * dzp_)) * dispPump_[pumpIndex][k] -
                    (1.0 * I1 * (j_ + 0.5) * other.dz_) * other.dispSign_[modeIndex][k] + 1.0 * I1 * bd * omega_[k];
      pF.row(i)[k] = ((1 - frac) * oF.row(j)[k] + frac * oF.row(j + 1)[k]) * std::exp(ph);
    }
    host_fft(pF.row(i), pT.row(i), N_, true);
  }
  for (int k = 0; k < N_; k++) {
    const cd ph = (1.0 * I1 * z_) * dispPump_[pumpIndex][k] - (1.0 * I1 * other.z_) * other.dispSign_[modeIndex][k] +
                  1.0 * I1 * bd * omega_[k];
    pF.row(nZP_ - 1)[k] = oF.row(other.nZ_ - 1)[k] * std::exp(ph);
  }
  // the reference transforms row _nZSteps-1 here, leaving the last pump time row as it was
  host_fft(pF.row(nZ_ - 1), pT.row(nZ_ - 1), N_, true);
  pumpHostStale_[pumpIndex] = false;
  pumpDevStale_[pumpIndex] = true;
  pumpComputed_ = true;
}

std::vector<double> Medium::linSpaced(int n, double low, double high) {
  std::vector<double> v(n);
  if (n == 1) {
    v[0] = high;
    return v;
  }
  const double step = (high - low) / (double)(n - 1);
  const bool flip = std::abs(high) < std::abs(low);
  for (int i = 0; i < n; i++) {
    if (flip) v[i] = (i == 0) ? low : high - (double)(n - 1 - i) * step;
    else v[i] = (i == n - 1) ? high : low + (double)i * step;
  }
  return v;
}

// ---------------------------------------------------------------------------
// device plan and host/device grid coherence
// ---------------------------------------------------------------------------

static std::vector<double> as_doubles(const std::vector<cd>& v) {
  std::vector<double> d(2 * v.size());
  std::memcpy(d.data(), v.data(), sizeof(double) * d.size());
  return d;
}

nlm_plan* Medium::plan() {
  if (plan_) {
    if (coefDirty_) refreshCoef();
    return plan_;
  }
  std::vector<cd> coef, cconst, halfIn((size_t)nS_ * N_), halfOut((size_t)nS_ * N_), dstep((size_t)nS_ * N_);
  stepCoefficients(coef, cconst);
  for (int m = 0; m < nS_; m++)
    for (int k = 0; k < N_; k++) {
      dstep[(size_t)m * N_ + k] = dispStepSign_[m][k];
      halfIn[(size_t)m * N_ + k] = std::exp(((0.5 * I1) * dz_) * dispSign_[m][k]);
      halfOut[(size_t)m * N_ + k] = std::exp(((-0.5 * I1) * dz_) * dispSign_[m][k]);
    }
  std::vector<cd> raman, filter;
  nlm_plan_desc d;
  std::memset(&d, 0, sizeof d);
  d.model = model_;
  d.n_freqs = N_;
  d.n_zsteps = nZ_;
  d.disp_step = reinterpret_cast<const double*>(dstep.data());
  d.half_in = reinterpret_cast<const double*>(halfIn.data());
  d.half_out = reinterpret_cast<const double*>(halfOut.data());
  d.step_coef = reinterpret_cast<const double*>(coef.data());
  d.const_coef = cconst.empty() ? nullptr : reinterpret_cast<const double*>(cconst.data());
  if (model_ == NLM_CHI3GNLSE) {
    gnlseTables(raman, filter);
    d.raman = reinterpret_cast<const double*>(raman.data());
    d.gnlse_filter = reinterpret_cast<const double*>(filter.data());
  }
  nlm_plan* p = nullptr;
  check(nlm_plan_create(&d, &p));
  plan_ = p;
  if (stream_) check(nlm_plan_set_stream(plan_, stream_));
  coefDirty_ = false;
  for (int q = 0; q < nP_; q++)
    if (pumpComputed_) pumpDevStale_[q] = true;
  return plan_;
}

void Medium::setStream(void* s) {
  stream_ = s;
  if (plan_) check(nlm_plan_set_stream(plan_, s));
}

void Medium::refreshCoef() {
  std::vector<cd> coef, cconst;
  stepCoefficients(coef, cconst);
  check(nlm_plan_set_coef(plan_, reinterpret_cast<const double*>(coef.data())));
  coefDirty_ = false;
}

void Medium::uploadPump(unsigned q) {
  allocPump(q);
  check(nlm_pump_upload(plan(), (int)q, reinterpret_cast<const double*>(pumpTime_[q].v.data()),
                        reinterpret_cast<const double*>(pumpFreq_[q].v.data())));
  pumpDevStale_[q] = false;
}

void Medium::markPumpOnDevice() {
  pumpComputed_ = true;
  for (int q = 0; q < nP_; q++) {
    pumpHostStale_[q] = true;
    pumpDevStale_[q] = false;
  }
}

void Medium::syncPumpToHost(unsigned i) {
  if ((int)i < nP_) allocPump(i);
  if (i < pumpHostStale_.size() && pumpHostStale_[i] && plan_) {
    check(nlm_pump_download(plan_, (int)i, reinterpret_cast<double*>(pumpTime_[i].v.data()),
                            reinterpret_cast<double*>(pumpFreq_[i].v.data())));
    pumpHostStale_[i] = false;
  }
}

void Medium::syncSignalToHost() {
  allocSignal();
  if (signalHostStale_ && plan_) {
    for (int m = 0; m < nS_; m++)
      check(nlm_grid_download(plan_, m, reinterpret_cast<double*>(signalFreq_[m].v.data()),
                              reinterpret_cast<double*>(signalTime_[m].v.data())));
    signalHostStale_ = false;
  }
}

Grid& Medium::pumpFreq(unsigned i) {
  if ((int)i >= nP_) throw std::out_of_range("vector::_M_range_check");
  syncPumpToHost(i);
  return pumpFreq_[i];
}
Grid& Medium::pumpTime(unsigned i) {
  if ((int)i >= nP_) throw std::out_of_range("vector::_M_range_check");
  syncPumpToHost(i);
  return pumpTime_[i];
}
Grid& Medium::signalFreq(unsigned i) {
  if ((int)i >= nS_) throw std::out_of_range("vector::_M_range_check");
  syncSignalToHost();
  return signalFreq_[i];
}
Grid& Medium::signalTime(unsigned i) {
  if ((int)i >= nS_) throw std::out_of_range("vector::_M_range_check");
  syncSignalToHost();
  return signalTime_[i];
}
Grid& Medium::field(unsigned i) { return field_.at(i); }

static void ensure_pumps_on_device(Medium* m, std::vector<bool>& devStale, bool computed, int nP,
                                   void (Medium::*up)(unsigned)) {
  (void)computed;
  for (int q = 0; q < nP; q++)
    if (devStale[q]) (m->*up)(q);
}

// ---------------------------------------------------------------------------
// runs
// ---------------------------------------------------------------------------

// _NonlinearMedium::runPumpSimulation, src/_NonlinearMedium.cpp:225-250
void Medium::runPumpSimulation() {
  nlm_plan* p = plan();
  std::vector<cd> env((size_t)nP_ * N_);
  std::vector<double> Dp((size_t)nP_ * N_);
  for (int q = 0; q < nP_; q++) {
    std::copy(envelope_[q].begin(), envelope_[q].end(), env.begin() + (size_t)q * N_);
    std::copy(dispPump_[q].begin(), dispPump_[q].end(), Dp.begin() + (size_t)q * N_);
  }
  std::vector<double> strength;
  if (profile_ != IntensityProfile::Constant) {
    const std::vector<double> x = linSpaced(nZP_, -0.5 * z_, 0.5 * z_);
    strength.resize(nZP_);
    for (int i = 0; i < nZP_; i++) {
      const double r = x[i] / zR_;
      strength[i] = profile_ == IntensityProfile::GaussianApodization ? std::exp(-0.5 * (r * r)) : 1 / std::sqrt(1 + r * r);
    }
  }
  check(nlm_pump_linear(p, nP_, reinterpret_cast<const double*>(env.data()), Dp.data(), dzp_,
                        strength.empty() ? nullptr : strength.data()));
  markPumpOnDevice();
}

// _NonlinearMedium::runSignalSimulation, src/_NonlinearMedium.cpp:253-260
void Medium::runSignalSimulation(const std::vector<cd>& in, bool inTimeDomain, unsigned inputMode) {
  const size_t n = in.size();
  if (n % N_ != 0 || n / N_ == 0 || n / N_ > (size_t)nS_)
    throw std::invalid_argument("inputProf array size does not match number of frequency/time bins");
  if ((int)inputMode >= nS_) throw std::invalid_argument("inputModes does not match any mode in the system");
  nlm_plan* p = plan();
  ensure_pumps_on_device(this, pumpDevStale_, pumpComputed_, nP_, &Medium::uploadPump);
  if (nP_ > 0 && !pumpComputed_) {
    // the reference propagates against its zero-initialised pump grid
    for (int q = 0; q < nP_; q++) uploadPump(q);
    pumpComputed_ = true;
  }
  check(nlm_propagate_grid(p, reinterpret_cast<const double*>(in.data()), (int)n, inTimeDomain ? 1 : 0, (int)inputMode));
  check(nlm_plan_sync(p));
  signalHostStale_ = true;
  signalOnDevice_ = true;
}

// runSignalSimulation (src/_NonlinearMedium.cpp:253-260) of every input row at once.
std::pair<Grid, Grid> Medium::batchSignalGrids(const Grid& in, bool inTimeDomain, unsigned inputMode) {
  if (in.cols % N_ != 0 || in.cols / N_ == 0 || in.cols / N_ > nS_)
    throw std::invalid_argument("inputProf array size does not match number of frequency/time bins");
  if ((int)inputMode >= nS_) throw std::invalid_argument("inputModes does not match any mode in the system");
  nlm_plan* p = plan();
  ensure_pumps_on_device(this, pumpDevStale_, pumpComputed_, nP_, &Medium::uploadPump);
  if (nP_ > 0 && !pumpComputed_) {
    for (int q = 0; q < nP_; q++) uploadPump(q);
    pumpComputed_ = true;
  }
  std::pair<Grid, Grid> g;
  g.first.setZero(in.rows * nS_ * nZ_, N_);
  g.second.setZero(in.rows * nS_ * nZ_, N_);
  check(nlm_batch_grids(p, reinterpret_cast<const double*>(in.v.data()), in.rows, in.cols, inTimeDomain ? 1 : 0,
                        (int)inputMode, reinterpret_cast<double*>(g.first.v.data()),
                        reinterpret_cast<double*>(g.second.v.data())));
  // the plan's grid now holds row 0's propagation
  signalHostStale_ = true;
  signalOnDevice_ = in.rows > 0;
  return g;
}

// Mode selection of computeGreensFunction, src/_NonlinearMedium.cpp:266-299.
// The reference counts entries below nSignalModes but selects by flag; the two
// agree only for flag lists whose set entries are all < nSignalModes, which is
// what is accepted here (other lists are undefined behaviour there).
static std::vector<int32_t> select_modes(int nS, const std::vector<uint8_t>& use, const char* none) {
  std::vector<int32_t> list;
  size_t cnt = 0;
  if (!use.empty()) {
    for (auto v : use) cnt += (v < nS);
    if (cnt == 0) throw std::invalid_argument(none);
  } else {
    cnt = nS;
  }
  for (int m = 0; m < nS; m++)
    if (use.empty() || (m < (int)use.size() && use[m])) list.push_back(m);
  if (list.size() != cnt) throw std::invalid_argument("useInput/useOutput must be per-mode flags");
  return list;
}

// _NonlinearMedium::computeGreensFunction, src/_NonlinearMedium.cpp:263-375
std::pair<Grid, Grid> Medium::computeGreensFunction(bool inTimeDomain, bool runPump, unsigned nThreads, bool normalize,
                                                    const std::vector<uint8_t>& useInput,
                                                    const std::vector<uint8_t>& useOutput) {
  return greensImpl(inTimeDomain, runPump, nThreads, normalize, useInput, useOutput, true);
}

void Medium::greensOnDevice(bool inTimeDomain, bool runPump) {
  if (fullyNonlinear()) throw std::runtime_error("Object does not have this method.");
  greensImpl(inTimeDomain, runPump, 1, false, {}, {}, false);
}

std::pair<Grid, Grid> Medium::greensForward(bool inTimeDomain, bool runPump, int ckptEvery) {
  if (model_ != NLM_CHI2PDC) throw std::invalid_argument("the differentiable Green's function is implemented for Chi2PDC");
  if (runPump) runPumpSimulation();
  nlm_plan* p = plan();
  for (int q = 0; q < nP_; q++)
    if (pumpDevStale_[q]) uploadPump(q);
  if (nP_ > 0 && !pumpComputed_) {
    for (int q = 0; q < nP_; q++) uploadPump(q);
    pumpComputed_ = true;
  }
  std::pair<Grid, Grid> res;
  res.first.setZero(N_, N_);
  res.second.setZero(N_, N_);
  const int32_t m0 = 0;
  check(nlm_greens_forward(p, inTimeDomain ? 1 : 0, 1, &m0, 1, &m0, nullptr, ckptEvery,
                           reinterpret_cast<double*>(res.first.v.data()), reinterpret_cast<double*>(res.second.v.data())));
  ++forwardSerial_;
  return res;
}

// Chain rule from the step coefficients to the per-step poling values for
// Chi2PDC (solvers/Chi2PDC.cpp:35-46): c_prev = p[i-1] nl mis(i-1),
// c_intm = (p[i-1] + p[i]) / 2 nl mis(i-1/2), c_curr = p[i] nl mis(i).
std::pair<std::vector<cd>, std::vector<double>> Medium::greensAdjoint(const cd* gC, const cd* gS, int rows,
                                                                      int cols) {
  if (model_ != NLM_CHI2PDC) throw std::invalid_argument("the differentiable Green's function is implemented for Chi2PDC");
  if (rows != N_ || cols != N_) throw std::invalid_argument("gradient shapes do not match the Green's function");
  nlm_plan* p = plan();
  std::vector<cd> gcoef((size_t)nZ_ * 3), genv(N_);
  check(nlm_greens_adjoint(p, reinterpret_cast<const double*>(gC), reinterpret_cast<const double*>(gS),
                           reinterpret_cast<double*>(gcoef.data()), nullptr, reinterpret_cast<double*>(genv.data())));
  std::vector<double> gpol(nZ_, 0.0);
  for (unsigned i = 1; i < (unsigned)nZ_; i++) {
    const double x[3] = {(i - 1) * dz_, (i - .5) * dz_, i * dz_};
    const cd d0 = nlStep_[0] * mismatch(0, x[0]), d1 = nlStep_[0] * mismatch(0, x[1]), d2 = nlStep_[0] * mismatch(0, x[2]);
    const cd* g = gcoef.data() + (size_t)i * 3;
    const double a0 = (std::conj(g[0]) * d0).real(), a1 = (std::conj(g[1]) * d1).real(), a2 = (std::conj(g[2]) * d2).real();
    gpol[i - 1] += a0 + 0.5 * a1;
    gpol[i] += 0.5 * a1 + a2;
  }
  return {genv, gpol};
}

// Differentiable batchSignalSimulation of a fully nonlinear medium
// (src/_FullyNonlinearMedium.hpp:8-11) for Chi3GNLSE.
Grid Medium::batchForward(const Grid& in, bool inTimeDomain, int ckptEvery) {
  if (model_ != NLM_CHI3GNLSE) throw std::invalid_argument("the differentiable batch is implemented for Chi3GNLSE");
  if (in.cols != N_) throw std::invalid_argument("Signals not of correct length!");
  if (in.rows == 0) throw std::invalid_argument("empty batch");
  Grid out;
  out.setZero(in.rows, N_);
  check(nlm_batch_forward(plan(), reinterpret_cast<const double*>(in.v.data()), in.rows, inTimeDomain ? 1 : 0,
                          ckptEvery, reinterpret_cast<double*>(out.v.data())));
  batchRows_ = in.rows;
  ++forwardSerial_;
  return out;
}

Grid Medium::batchAdjoint(const Grid& gOut) {
  if (model_ != NLM_CHI3GNLSE) throw std::invalid_argument("the differentiable batch is implemented for Chi3GNLSE");
  if (batchRows_ == 0) throw std::runtime_error("run batchForward before batchAdjoint");
  if (gOut.rows != batchRows_ || gOut.cols != N_)
    throw std::invalid_argument("gradient shape does not match the last batchForward");
  Grid g;
  g.setZero(gOut.rows, N_);
  check(nlm_batch_adjoint(plan(), reinterpret_cast<const double*>(gOut.v.data()), reinterpret_cast<double*>(g.v.data())));
  return g;
}

void Medium::enableTiming(bool on) { check(nlm_plan_enable_timing(plan(), on ? 1 : 0)); }
void Medium::setStepTable(bool on) { check(nlm_plan_set_step_table(plan(), on ? 1 : 0)); }

std::vector<double> Medium::lastTiming() {
  std::vector<double> ms(3, 0.0);
  if (plan_) check(nlm_plan_last_timing(plan_, ms.data()));
  return ms;
}

void Medium::sync() {
  if (plan_) check(nlm_plan_sync(plan_));
}

std::pair<Grid, Grid> Medium::greensImpl(bool inTimeDomain, bool runPump, unsigned nThreads, bool normalize,
                                         const std::vector<uint8_t>& useInput, const std::vector<uint8_t>& useOutput,
                                         bool download) {
  if (useInput.size() > (size_t)nS_)
    throw std::invalid_argument("List of requested inputs indices longer than number of modes!");
  if (useOutput.size() > (size_t)nS_)
    throw std::invalid_argument("List of requested output indices longer than number of modes!");
  const std::vector<int32_t> ins = select_modes(nS_, useInput, "Requested no valid inputs!");
  const std::vector<int32_t> outs = select_modes(nS_, useOutput, "Requested no valid outputs!");
  const int nIn = (int)ins.size(), nOut = (int)outs.size();
  if (nThreads > (unsigned)(N_ * nIn)) throw std::invalid_argument("Too many threads requested!");
  if (nThreads == 0) throw std::length_error("vector::reserve");
  if (runPump) runPumpSimulation();
  nlm_plan* p = plan();
  ensure_pumps_on_device(this, pumpDevStale_, pumpComputed_, nP_, &Medium::uploadPump);
  if (nP_ > 0 && !pumpComputed_) {
    for (int q = 0; q < nP_; q++) uploadPump(q);
    pumpComputed_ = true;
  }
  const int nb = std::min(nIn, nOut);
  std::vector<cd> scale;
  if (normalize) {
    // the reference indexes _nlStep by block index; models with fewer nonlinear
    // lengths than modes read past the end there (undefined behaviour)
    if ((int)nlStep_.size() < std::max(nb, nOut))
      throw std::invalid_argument("normalize requires one nonlinear length per mode for this medium");
    scale.assign((size_t)nb * nOut, cd(1, 0));
    for (int im = 0; im < nb; im++)
      for (int om = 0; om < nOut; om++)
        if (im != om) scale[(size_t)im * nOut + om] = std::sqrt(nlStep_[im] / nlStep_[om]);
  }
  std::pair<Grid, Grid> res;
  lastGreensRows_ = nOut * N_;
  lastGreensCols_ = nIn * N_;
  if (download) {
    res.first.setZero(nOut * N_, nIn * N_);
    res.second.setZero(nOut * N_, nIn * N_);
  }
  check(nlm_greens(p, inTimeDomain ? 1 : 0, nIn, ins.data(), nOut, outs.data(),
                   normalize ? reinterpret_cast<const double*>(scale.data()) : nullptr,
                   download ? reinterpret_cast<double*>(res.first.v.data()) : nullptr,
                   download ? reinterpret_cast<double*>(res.second.v.data()) : nullptr));
  return res;
}

// _NonlinearMedium::batchSignalSimulation, src/_NonlinearMedium.cpp:378-444
Grid Medium::batchSignalSimulation(const Grid& in, bool inTimeDomain, bool runPump, unsigned nThreads,
                                   unsigned inputMode, const std::vector<uint8_t>& useOutput) {
  const int nInputs = in.rows, inCols = in.cols;
  if (inCols % N_ != 0 || inCols / N_ == 0 || inCols / N_ > nS_) throw std::invalid_argument("Signals not of correct length!");
  if (nThreads > (unsigned)nInputs) throw std::invalid_argument("Too many threads requested!");
  if ((int)inputMode >= nS_) throw std::invalid_argument("inputModes does not match any mode in the system");
  int nOut = 0;
  std::vector<int32_t> outs;
  if (!useOutput.empty()) {
    for (auto v : useOutput) nOut += (v != 0);
    if (nOut == 0) throw std::invalid_argument("Requested no outputs!");
  } else {
    nOut = nS_;
  }
  for (int m = 0; m < nS_; m++)
    if (useOutput.empty() || (m < (int)useOutput.size() && useOutput[m])) outs.push_back(m);
  if ((int)outs.size() != nOut) throw std::invalid_argument("useOutput must be per-mode flags");
  if (nThreads == 0) throw std::length_error("vector::reserve");
  if (runPump) runPumpSimulation();
  nlm_plan* p = plan();
  ensure_pumps_on_device(this, pumpDevStale_, pumpComputed_, nP_, &Medium::uploadPump);
  if (nP_ > 0 && !pumpComputed_) {
    for (int q = 0; q < nP_; q++) uploadPump(q);
    pumpComputed_ = true;
  }
  Grid out;
  out.setZero(nInputs, nOut * N_);
  check(nlm_batch(p, reinterpret_cast<const double*>(in.v.data()), nInputs, inCols, inTimeDomain ? 1 : 0,
                  (int)inputMode, nOut, outs.data(), reinterpret_cast<double*>(out.v.data())));
  return out;
}

// ---------------------------------------------------------------------------
// per-step scalar factors of the DiffEqs
// ---------------------------------------------------------------------------

cd Medium::mismatch(unsigned which, double x) const { return std::exp(1.0 * I1 * diffBeta0_[which] * x); }

// _FullyNonlinearMedium::relativeAmplitude / relativeIntensity,
// src/_FullyNonlinearMedium.hpp:40-63
double Medium::relativeAmplitude(double i) const {
  switch (profile_) {
    case IntensityProfile::GaussianBeam: return std::sqrt(1 / (1 + std::pow((i * dz_ - 0.5 * z_) / zR_, 2)));
    case IntensityProfile::Constant: return 1.;
    case IntensityProfile::GaussianApodization: return std::exp(-0.5 * std::pow((i * dz_ - 0.5 * z_) / zR_, 2));
  }
  return 0;
}
double Medium::relativeIntensity(double i) const {
  switch (profile_) {
    case IntensityProfile::GaussianBeam: return 1 / (1 + std::pow((i * dz_ - 0.5 * z_) / zR_, 2));
    case IntensityProfile::Constant: return 1.;
    case IntensityProfile::GaussianApodization: return std::exp(-std::pow((i * dz_ - 0.5 * z_) / zR_, 2));
  }
  return 0;
}

// Chi3GNLSE tables, solvers/Chi3GNLSE.cpp:20-36 (Raman) and :72 (filter)
void Medium::gnlseTables(std::vector<cd>& raman, std::vector<cd>& filter) const {
  raman = raman_;
  filter.resize(N_);
  for (int k = 0; k < N_; k++) filter[k] = nlStep_[0] + nlStep_[1] * omega_[k];
}

namespace {

// Positions of the three z points of RK step i: (i-1)dz, (i-1/2)dz, i dz.
struct StepPos {
  unsigned i;
  double x[3];
  double iv[3];
  StepPos(unsigned i_, double dz) : i(i_) {
    x[0] = (i - 1) * dz;
    x[1] = (i - .5) * dz;
    x[2] = i * dz;
    iv[0] = (double)(i - 1);
    iv[1] = i - .5;
    iv[2] = (double)i;
  }
};

template <class F>
void fill_steps(int nZ, int NC, std::vector<cd>& coef, F&& per_step) {
  coef.assign((size_t)nZ * 3 * std::max(NC, 1), cd(0, 0));
  for (unsigned i = 1; i < (unsigned)nZ; i++) per_step(i, coef.data() + (size_t)i * 3 * std::max(NC, 1));
}

}  // namespace

// ---------------------------------------------------------------------------
// the solvers (constructor mapping + DiffEq scalar factors)
// ---------------------------------------------------------------------------

#define POLS                                                     \
  const double prevPolDir = poling_[i - 1], currPolDir = poling_[i]; \
  const double intmPolDir = 0.5 * (prevPolDir + currPolDir);         \
  const double pol[3] = {prevPolDir, intmPolDir, currPolDir};

// Chi3, solvers/Chi3.cpp:18-24, 25-54, 57-69
Chi3::Chi3(const Args& a) : Medium(a) {}

// Chi3::runPumpSimulation (nonlinear pump), solvers/Chi3.cpp:25-54
void Chi3::runPumpSimulation() {
    nlm_plan* p = plan();
    std::vector<cd> relI(nZP_);
    const cd c0 = dzp_ / dz_ * nlStep_[0];
    const std::vector<double> x = linSpaced(nZP_, -0.5 * z_, 0.5 * z_);
    for (int i = 0; i < nZP_; i++) {
      const double r = x[i] / zR_;
      switch (profile_) {
        default:
        case IntensityProfile::Constant: relI[i] = c0; break;
        case IntensityProfile::GaussianBeam: relI[i] = c0 / (1 + r * r); break;
        case IntensityProfile::GaussianApodization: relI[i] = c0 * std::exp(-(r * r)); break;
      }
    }
    check(nlm_pump_chi3(p, reinterpret_cast<const double*>(envelope_[0].data()), dispPump_[0].data(), dzp_,
                        reinterpret_cast<const double*>(dispStepPump_[0].data()),
                        reinterpret_cast<const double*>(relI.data())));
    markPumpOnDevice();
}

void Chi3::stepCoefficients(std::vector<cd>& coef, std::vector<cd>& cconst) const {
    coef.assign((size_t)nZ_ * 3, cd(0, 0));
    cconst = {nlStep_[0]};
}

// Chi2PDC, solvers/Chi2PDC.cpp:27-47
Chi2PDC::Chi2PDC(const Args& a) : Medium(a) {}

void Chi2PDC::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 1, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) c[j] = pol[j] * nlStep_[0] * mismatch(0, s.x[j]);
    });
}

// Chi2SFG, solvers/Chi2SFG.cpp:28-56
Chi2SFG::Chi2SFG(const Args& a) : Medium(a) {}

void Chi2SFG::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 2, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) {
        const cd mis = mismatch(0, s.x[j]), inv = 1. / mis;
        c[2 * j] = pol[j] * nlStep_[0] * mis;
        c[2 * j + 1] = pol[j] * nlStep_[1] * inv;
      }
    });
}

// Chi2AFC, solvers/Chi2AFC.cpp:26-51 (no poling; linearly chirped mismatch)
Chi2AFC::Chi2AFC(const Args& a) : Medium(a) {}

void Chi2AFC::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 2, coef, [&](unsigned i, cd* c) {
      const double arg = 0.5 * (diffBeta0_[1] - diffBeta0_[0]) / z_;
      const cd mis[3] = {std::exp(0.5 * I1 * ((diffBeta0_[0] + (i - 1) * dz_ * arg) * (i - 1) * dz_)),
                         std::exp(0.5 * I1 * ((diffBeta0_[0] + (i - .5) * dz_ * arg) * (i - .5) * dz_)),
                         std::exp(0.5 * I1 * ((diffBeta0_[0] + i * dz_ * arg) * i * dz_))};
      for (int j = 0; j < 3; j++) {
        const cd inv = 1. / mis[j];
        c[2 * j] = nlStep_[0] * mis[j];
        c[2 * j + 1] = nlStep_[1] * inv;
      }
    });
}

// Chi2SFGPDC, solvers/Chi2SFGPDC.cpp:28-59
Chi2SFGPDC::Chi2SFGPDC(const Args& a) : Medium(a) {}

void Chi2SFGPDC::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 4, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) {
        const cd mis = mismatch(0, s.x[j]), inv = 1. / mis, miso = mismatch(1, s.x[j]);
        c[4 * j + 0] = pol[j] * nlStep_[0] * mis;
        c[4 * j + 1] = pol[j] * nlStep_[1];
        c[4 * j + 2] = inv;
        c[4 * j + 3] = miso;
      }
    });
}

// Chi2SFGII, solvers/Chi2SFGII.cpp:34-90; the reference's k4 uses prevPolDir
Chi2SFGII::Chi2SFGII(const Args& a) : Medium(a) {}

void Chi2SFGII::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 7, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      const double polk[3] = {prevPolDir, intmPolDir, prevPolDir};
      (void)pol;
      for (int j = 0; j < 3; j++) {
        const cd cz = mismatch(0, s.x[j]), cy = mismatch(1, s.x[j]), sz = mismatch(2, s.x[j]);
        const cd icz = 1. / cz, icy = 1. / cy, isz = 1. / sz;
        cd* o = c + 7 * j;
        o[0] = cd(polk[j], 0);
        o[1] = nlStep_[0] * cz;
        o[2] = nlStep_[1] * sz;
        o[3] = polk[j] * nlStep_[0] * cy;
        o[4] = nlStep_[0] * icy;
        o[5] = nlStep_[1] * isz;
        o[6] = polk[j] * nlStep_[0] * icz;
      }
    });
}

// Chi2PDCII, solvers/Chi2PDCII.cpp:29-55
Chi2PDCII::Chi2PDCII(const Args& a) : Medium(a) {}

void Chi2PDCII::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 2, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) {
        const cd mis = mismatch(0, s.x[j]);
        c[2 * j] = pol[j] * nlStep_[0] * mis;
        c[2 * j + 1] = pol[j] * nlStep_[1] * mis;
      }
    });
}

// Chi2SFGOPA, solvers/Chi2SFGOPA.cpp:31-77
Chi2SFGOPA::Chi2SFGOPA(const Args& a) : Medium(a) {}

void Chi2SFGOPA::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 4, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) {
        const cd ms = mismatch(0, s.x[j]), md = mismatch(1, s.x[j]), is = 1. / ms;
        c[4 * j + 0] = pol[j] * nlStep_[0] * ms;
        c[4 * j + 1] = pol[j] * nlStep_[2] * md;
        c[4 * j + 2] = pol[j] * nlStep_[1] * is;
        c[4 * j + 3] = pol[j] * nlStep_[3] * md;
      }
    });
}

// Chi2SFGXPM, solvers/Chi2SFGXPM.cpp:26-55
Chi2SFGXPM::Chi2SFGXPM(const Args& a) : Medium(a) {}

void Chi2SFGXPM::stepCoefficients(std::vector<cd>& coef, std::vector<cd>& cconst) const {
    fill_steps(nZ_, 2, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) {
        const cd mis = mismatch(0, s.x[j]), inv = 1. / mis;
        c[2 * j] = pol[j] * nlStep_[0] * mis;
        c[2 * j + 1] = pol[j] * nlStep_[1] * inv;
      }
    });
    cconst = {nlStep_[2], nlStep_[3]};
}

// Chi2SHG, solvers/Chi2SHG.cpp:25-60
Chi2SHG::Chi2SHG(const Args& a) : FullyNonlinearMedium(a) {}

void Chi2SHG::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 2, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) {
        const double r = pol[j] * relativeAmplitude(s.iv[j]);
        const cd mis = mismatch(0, s.x[j]), inv = 1. / mis;
        c[2 * j] = r * nlStep_[0] * mis;
        c[2 * j + 1] = r * nlStep_[1] * inv * .5;
      }
    });
}

// Chi2ASHG, solvers/Chi2ASHG.cpp:24-56 (no poling; linearly chirped mismatch)
Chi2ASHG::Chi2ASHG(const Args& a) : FullyNonlinearMedium(a) {}

void Chi2ASHG::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 2, coef, [&](unsigned i, cd* c) {
      StepPos s(i, dz_);
      const double arg = 0.5 * (diffBeta0_[1] - diffBeta0_[0]) / z_;
      const cd mis[3] = {std::exp(0.5 * I1 * ((diffBeta0_[0] + (i - 1) * dz_ * arg) * (i - 1) * dz_)),
                         std::exp(0.5 * I1 * ((diffBeta0_[0] + (i - .5) * dz_ * arg) * (i - .5) * dz_)),
                         std::exp(0.5 * I1 * ((diffBeta0_[0] + i * dz_ * arg) * i * dz_))};
      for (int j = 0; j < 3; j++) {
        const double r = relativeAmplitude(s.iv[j]);
        const cd inv = 1. / mis[j];
        c[2 * j] = r * nlStep_[0] * mis[j];
        c[2 * j + 1] = r * nlStep_[1] * inv * .5;
      }
    });
}

// Chi2SHGOPA, solvers/Chi2SHGOPA.cpp:32-84
Chi2SHGOPA::Chi2SHGOPA(const Args& a) : FullyNonlinearMedium(a) {}

void Chi2SHGOPA::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 6, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) {
        const double r = pol[j] * relativeAmplitude(s.iv[j]);
        const cd ms = mismatch(0, s.x[j]), is = 1. / ms, mo = mismatch(1, s.x[j]), io = 1. / mo;
        cd* o = c + 6 * j;
        o[0] = r * nlStep_[0] * ms;
        o[1] = r * nlStep_[1];
        o[2] = mo;
        o[3] = is * .5;
        o[4] = r * nlStep_[2] * io;
        o[5] = r * nlStep_[3] * io;
      }
    });
}

// Chi2DSFG, solvers/Chi2DSFG.cpp:25-65
Chi2DSFG::Chi2DSFG(const Args& a) : FullyNonlinearMedium(a) {}

void Chi2DSFG::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 3, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) {
        const double r = pol[j] * relativeAmplitude(s.iv[j]);
        const cd mis = mismatch(0, s.x[j]), inv = 1. / mis;
        c[3 * j + 0] = r * nlStep_[0] * mis;
        c[3 * j + 1] = r * nlStep_[1] * inv;
        c[3 * j + 2] = r * nlStep_[2] * mis;
      }
    });
}

// Chi2SHGXPM, solvers/Chi2SHGXPM.cpp:25-65
Chi2SHGXPM::Chi2SHGXPM(const Args& a) : FullyNonlinearMedium(a) {}

void Chi2SHGXPM::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 4, coef, [&](unsigned i, cd* c) {
      POLS StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) {
        const double amp = relativeAmplitude(s.iv[j]);
        const double r = pol[j] * amp;
        const cd mis = mismatch(0, s.x[j]), inv = 1. / mis;
        const cd chi3 = std::pow(amp, 2) * nlStep_[2];
        c[4 * j + 0] = r * nlStep_[0] * mis;
        c[4 * j + 1] = r * nlStep_[1] * inv * .5;
        c[4 * j + 2] = chi3;
        c[4 * j + 3] = 2. * chi3;
      }
    });
}

// Chi3GNLSE, solvers/Chi3GNLSE.cpp:20-36, 51-92
Chi3GNLSE::Chi3GNLSE(const Args& a, double fr, double fb, double tau1, double tau2, double tau3) : FullyNonlinearMedium(a) {
    const double coeff1 = fr * (1. - fb) * (tau1 / (tau2 * tau2) + 1. / tau1);
    const double coeff2 = fr * fb / (tau3 * tau3);
    std::vector<cd> rt(N_, cd(0, 0));
    for (int k = 0; k < N_ / 2; k++) {
      const double t = tau_[k];
      rt[k] = coeff1 * std::exp(-t / tau2) * std::sin(t / tau1) + coeff2 * std::exp(-t / tau3) * (2 * tau3 - t);
    }
    rt[0] += 1. - fr;
    raman_.resize(N_);
    host_fft(rt.data(), raman_.data(), N_, false);
}

void Chi3GNLSE::stepCoefficients(std::vector<cd>& coef, std::vector<cd>&) const {
    fill_steps(nZ_, 1, coef, [&](unsigned i, cd* c) {
      StepPos s(i, dz_);
      for (int j = 0; j < 3; j++) c[j] = cd(relativeIntensity(s.iv[j]), 0);
    });
}

// ---------------------------------------------------------------------------
// solver factory: the argument mapping of each solver constructor
// ---------------------------------------------------------------------------

std::unique_ptr<Medium> makeSolver(const std::string& name, const std::vector<double>& v,
                                   const std::vector<cd>& customPump, int pulseType, int intensityProfile,
                                   const std::vector<double>& poling, unsigned tPrecision, unsigned zPrecision,
                                   unsigned nZOverride, const std::vector<double>& polingProfile) {
  Medium::Args a{};
  a.customPump = customPump;
  a.pulseType = (PulseType)pulseType;
  a.profile = (IntensityProfile)intensityProfile;
  a.poling = poling;
  a.tPrecision = tPrecision;
  a.zPrecision = zPrecision;
  a.nZOverride = nZOverride;
  a.polingProfile = polingProfile;
  a.chirp = 0;
  a.delay = 0;
  auto need = [&](size_t n) {
    if (v.size() != n) throw std::invalid_argument("internal: wrong number of numeric arguments for " + name);
  };
  // numeric argument vectors hold the solver's double arguments in declaration order
  if (name == "Chi3") {  // solvers/Chi3.cpp:18-24
    need(7);             // relativeLength, nlLength, beta2, beta3, rayleighLength, tMax, chirp
    a.model = NLM_CHI3; a.nS = 1; a.nP = 1; a.canBePoled = false;
    a.relativeLength = v[0]; a.nl = {v[1]}; a.beta2 = {v[2]}; a.beta2s = {v[2]};
    a.beta1 = {0}; a.beta1s = {0}; a.beta3 = {v[3]}; a.beta3s = {v[3]}; a.diffBeta0 = {};
    a.rayleighLength = v[4]; a.tMax = v[5]; a.chirp = v[6]; a.delay = 0; a.poling = {};
    a.polingProfile = {};
    return std::unique_ptr<Medium>(new Chi3(a));
  }
  if (name == "Chi2PDC") {  // solvers/Chi2PDC.cpp:18-24
    need(13);  // relLen, nl, beta2, beta2s, beta1, beta1s, beta3, beta3s, diffBeta0, zR, tMax, chirp, delay
    a.model = NLM_CHI2PDC; a.nS = 1; a.nP = 1; a.canBePoled = true;
    a.relativeLength = v[0]; a.nl = {v[1]}; a.beta2 = {v[2]}; a.beta2s = {v[3]}; a.beta1 = {v[4]}; a.beta1s = {v[5]};
    a.beta3 = {v[6]}; a.beta3s = {v[7]}; a.diffBeta0 = {v[8]}; a.rayleighLength = v[9]; a.tMax = v[10];
    a.chirp = v[11]; a.delay = v[12];
    return std::unique_ptr<Medium>(new Chi2PDC(a));
  }
  if (name == "Chi2SFG" || name == "Chi2PDCII" || name == "Chi2AFC" || name == "Chi2SFGPDC") {
    // relLen, nl0, nl1, beta2, beta2s, beta2o, beta1, beta1s, beta1o, beta3, beta3s, beta3o, d0, [d1], zR, tMax, chirp, delay
    const bool two = (name == "Chi2AFC" || name == "Chi2SFGPDC");
    need(two ? 18 : 17);
    a.nS = 2; a.nP = 1;
    a.relativeLength = v[0];
    a.beta2 = {v[3]}; a.beta2s = {v[4], v[5]}; a.beta1 = {v[6]}; a.beta1s = {v[7], v[8]};
    a.beta3 = {v[9]}; a.beta3s = {v[10], v[11]};
    int o = 12;
    if (two) { a.diffBeta0 = {v[12], v[13]}; o = 14; }
    else a.diffBeta0 = {v[12]}, o = 13;
    a.rayleighLength = v[o]; a.tMax = v[o + 1]; a.chirp = v[o + 2]; a.delay = v[o + 3];
    if (name == "Chi2AFC") {
      a.model = NLM_CHI2AFC; a.canBePoled = false; a.nl = {0.5 * M_PI * v[1], 0.5 * M_PI * v[2]};
      a.poling = {}; a.polingProfile = {};
      return std::unique_ptr<Medium>(new Chi2AFC(a));
    }
    a.canBePoled = true; a.nl = {v[1], v[2]};
    if (name == "Chi2SFG") { a.model = NLM_CHI2SFG; return std::unique_ptr<Medium>(new Chi2SFG(a)); }
    if (name == "Chi2PDCII") { a.model = NLM_CHI2PDCII; return std::unique_ptr<Medium>(new Chi2PDCII(a)); }
    a.model = NLM_CHI2SFGPDC;
    return std::unique_ptr<Medium>(new Chi2SFGPDC(a));
  }
  if (name == "Chi2SFGII") {  // solvers/Chi2SFGII.cpp:20-31
    // relLen, nlZY, nlZZ, beta2, b2sz, b2sy, b2oz, b2oy, beta1, b1sz, b1sy, b1oz, b1oy, beta3, b3sz, b3sy, b3oz, b3oy,
    // d0z, d0y, d0, zR, tMax, chirp, delay
    need(25);
    a.model = NLM_CHI2SFGII; a.nS = 4; a.nP = 1; a.canBePoled = true;
    a.relativeLength = v[0]; a.nl = {v[1], v[2]};
    a.beta2 = {v[3]}; a.beta2s = {v[4], v[5], v[6], v[7]};
    a.beta1 = {v[8]}; a.beta1s = {v[9], v[10], v[11], v[12]};
    a.beta3 = {v[13]}; a.beta3s = {v[14], v[15], v[16], v[17]};
    a.diffBeta0 = {v[18], v[19], v[20]}; a.rayleighLength = v[21]; a.tMax = v[22]; a.chirp = v[23]; a.delay = v[24];
    return std::unique_ptr<Medium>(new Chi2SFGII(a));
  }
  if (name == "Chi2SFGOPA") {  // solvers/Chi2SFGOPA.cpp:26-34
    // relLen, nlSFGh, nlSFGf, nlDOPAh, nlDOPAf, b2F, b2H, b2h, b2f, b1F, b1H, b1h, b1f, b3F, b3H, b3h, b3f,
    // dSFG, dDOPA, zR, tMax, chirp, delay
    need(23);
    a.model = NLM_CHI2SFGOPA; a.nS = 2; a.nP = 2; a.canBePoled = true;
    a.relativeLength = v[0]; a.nl = {v[1], v[2], v[3], v[4]};
    a.beta2 = {v[5], v[6]}; a.beta2s = {v[7], v[8]}; a.beta1 = {v[9], v[10]}; a.beta1s = {v[11], v[12]};
    a.beta3 = {v[13], v[14]}; a.beta3s = {v[15], v[16]}; a.diffBeta0 = {v[17], v[18]};
    a.rayleighLength = v[19]; a.tMax = v[20]; a.chirp = v[21]; a.delay = v[22];
    return std::unique_ptr<Medium>(new Chi2SFGOPA(a));
  }
  if (name == "Chi2SFGXPM") {  // solvers/Chi2SFGXPM.cpp:16-23
    // relLen, nl, nlOrig, nlChi3, nlChi3Orig, beta2, beta2s, beta2o, beta1, beta1s, beta1o, beta3, beta3s, beta3o,
    // diffBeta0, zR, tMax, chirp, delay
    need(19);
    a.model = NLM_CHI2SFGXPM; a.nS = 2; a.nP = 1; a.canBePoled = true;
    a.relativeLength = v[0]; a.nl = {v[1], v[2], v[3] / 2, v[4] / 2};
    a.beta2 = {v[5]}; a.beta2s = {v[6], v[7]}; a.beta1 = {v[8]}; a.beta1s = {v[9], v[10]};
    a.beta3 = {v[11]}; a.beta3s = {v[12], v[13]}; a.diffBeta0 = {v[14]};
    a.rayleighLength = v[15]; a.tMax = v[16]; a.chirp = v[17]; a.delay = v[18];
    return std::unique_ptr<Medium>(new Chi2SFGXPM(a));
  }
  // fully nonlinear media: _FullyNonlinearMedium(nS, canBePoled, 0, relLen, nl, beta2s, beta1s, beta3s, diffBeta0, ...)
  a.nP = 0; a.beta2 = {}; a.beta1 = {}; a.beta3 = {}; a.customPump = {}; a.pulseType = PulseType::Gaussian;
  if (name == "Chi2SHG" || name == "Chi2SHGXPM" || name == "Chi2ASHG") {
    // Chi2SHG / Chi2ASHG: relLen, nlH, nlP, b2h, b2p, b1h, b1p, b3h, b3p, d0 [, d1], zR, tMax
    // Chi2SHGXPM:         relLen, nlH, nlP, nlChi3, b2h, b2p, b1h, b1p, b3h, b3p, d0, zR, tMax
    const bool xpm = name == "Chi2SHGXPM", ashg = name == "Chi2ASHG";
    need(ashg ? 13 : (xpm ? 13 : 12));
    a.nS = 2;
    a.relativeLength = v[0];
    int o = xpm ? 4 : 3;
    a.beta2s = {v[o + 1], v[o]}; a.beta1s = {v[o + 3], v[o + 2]}; a.beta3s = {v[o + 5], v[o + 4]};
    if (ashg) {
      a.model = NLM_CHI2ASHG; a.canBePoled = false;
      a.nl = {0.5 * M_PI * v[2], 0.5 * M_PI * v[1]};
      a.diffBeta0 = {v[9], v[10]}; a.rayleighLength = v[11]; a.tMax = v[12];
      a.poling = {}; a.polingProfile = {};
      return std::unique_ptr<Medium>(new Chi2ASHG(a));
    }
    a.canBePoled = true;
    if (xpm) {
      a.model = NLM_CHI2SHGXPM; a.nl = {v[2], v[1], v[3]};
      a.diffBeta0 = {v[10]}; a.rayleighLength = v[11]; a.tMax = v[12];
      return std::unique_ptr<Medium>(new Chi2SHGXPM(a));
    }
    a.model = NLM_CHI2SHG; a.nl = {v[2], v[1]};
    a.diffBeta0 = {v[9]}; a.rayleighLength = v[10]; a.tMax = v[11];
    return std::unique_ptr<Medium>(new Chi2SHG(a));
  }
  if (name == "Chi2SHGOPA") {  // solvers/Chi2SHGOPA.cpp:19-29
    // relLen, nlP, nlSH, nlPA1, nlPA2, b2p, b2sh, b2pa1, b2pa2, b1p.., b3p.., dshg, dopa, zR, tMax
    need(21);
    a.model = NLM_CHI2SHGOPA; a.nS = 4; a.canBePoled = true;
    a.relativeLength = v[0]; a.nl = {v[1], v[2], v[3], v[4]};
    a.beta2s = {v[5], v[6], v[7], v[8]}; a.beta1s = {v[9], v[10], v[11], v[12]}; a.beta3s = {v[13], v[14], v[15], v[16]};
    a.diffBeta0 = {v[17], v[18]}; a.rayleighLength = v[19]; a.tMax = v[20];
    return std::unique_ptr<Medium>(new Chi2SHGOPA(a));
  }
  if (name == "Chi2DSFG") {  // solvers/Chi2DSFG.cpp:16-22
    // relLen, nlP, nlS, nlD, b2p, b2s, b2d, b1p, b1s, b1d, b3p, b3s, b3d, d0, zR, tMax
    need(16);
    a.model = NLM_CHI2DSFG; a.nS = 3; a.canBePoled = true;
    a.relativeLength = v[0]; a.nl = {v[1], v[2], v[3]};
    a.beta2s = {v[4], v[5], v[6]}; a.beta1s = {v[7], v[8], v[9]}; a.beta3s = {v[10], v[11], v[12]};
    a.diffBeta0 = {v[13]}; a.rayleighLength = v[14]; a.tMax = v[15];
    return std::unique_ptr<Medium>(new Chi2DSFG(a));
  }
  if (name == "Chi3GNLSE") {  // solvers/Chi3GNLSE.cpp:13-36
    // relLen, nlLength, selfSteepLength, fr, fb, tau1, tau2, tau3, beta2, beta3, zR, tMax
    need(12);
    a.model = NLM_CHI3GNLSE; a.nS = 1; a.canBePoled = false;
    a.relativeLength = v[0]; a.nl = {v[1], v[2]};
    a.beta2s = {v[8]}; a.beta1s = {0}; a.beta3s = {v[9]}; a.diffBeta0 = {};
    a.rayleighLength = v[10]; a.tMax = v[11]; a.poling = {}; a.polingProfile = {};
    return std::unique_ptr<Medium>(new Chi3GNLSE(a, v[3], v[4], v[5], v[6], v[7]));
  }
  throw std::invalid_argument("unknown solver " + name);
}

// ---------------------------------------------------------------------------
// Cascade, src/Cascade.cpp:3-167
// ---------------------------------------------------------------------------

Cascade::Cascade(const std::vector<Medium*>& inputMedia, const std::vector<std::map<unsigned, unsigned>>& modeConnections,
                 bool sharePump) {
  if (inputMedia.empty()) throw std::invalid_argument("Cascade must contain at least one medium");
  N_ = inputMedia[0]->N_;
  tMax_ = inputMedia[0]->tMax_;
  nZ_ = 0;
  for (Medium* m : inputMedia) {
    if (m->N_ != N_ || m->tMax_ != tMax_)
      throw std::invalid_argument("Medium does not have same time and frequency axes as the first");
    media_.push_back(m);
    nZ_ += m->nZ_;
  }
  nZP_ = 0;
  if (modeConnections.size() != media_.size() - 1)
    throw std::invalid_argument("Must have one connection per pair of adjacent media");
  for (size_t i = 0; i < modeConnections.size(); i++) {
    if (modeConnections[i].empty()) throw std::invalid_argument("Connections missing between media!");
    for (const auto& sm : modeConnections[i])
      if ((int)sm.second >= media_[i]->nS_ || (int)sm.first >= media_[i + 1]->nS_)
        throw std::invalid_argument("Invalid connections, out of range");
    connections_.push_back(modeConnections[i]);
  }
  sharedPump_ = sharePump;
}

void Cascade::addMedium(Medium& medium, const std::map<unsigned, unsigned>& connection) {
  if (medium.N_ != N_ || medium.tMax_ != tMax_)
    throw std::invalid_argument("Medium does not have same time and frequency axes as the first");
  if (connection.empty()) throw std::invalid_argument("No connection!");
  for (const auto& sm : connection)
    if ((int)sm.second >= media_.back()->nS_ || (int)sm.first >= medium.nS_)
      throw std::invalid_argument("Invalid connections, out of range");
  media_.push_back(&medium);
  connections_.push_back(connection);
}

void Cascade::setPump(PulseType pulseType, double chirpLength, double delayLength, unsigned pumpIndex) {
  if (sharedPump_) media_[0]->setPump(pulseType, chirpLength, delayLength, pumpIndex);
  else
    for (Medium* m : media_) m->setPump(pulseType, chirpLength, delayLength, pumpIndex);
}

void Cascade::setPump(const std::vector<cd>& customPump, double chirpLength, double delayLength, unsigned pumpIndex) {
  if (sharedPump_) media_[0]->setPump(customPump, chirpLength, delayLength, pumpIndex);
  else
    for (Medium* m : media_) m->setPump(customPump, chirpLength, delayLength, pumpIndex);
}

void Cascade::runPumpSimulation() {
  if (!sharedPump_) {
    for (Medium* m : media_) m->runPumpSimulation();
  } else {
    media_[0]->runPumpSimulation();
    for (size_t i = 1; i < media_.size(); i++) {
      Grid& prev = media_[i - 1]->pumpTime(0);
      const cd* last = prev.row(prev.rows - 1);
      media_[i]->envelope_[0].assign(last, last + N_);
      media_[i]->runPumpSimulation();
    }
  }
}

void Cascade::runSignalSimulation(const std::vector<cd>& inputProf, bool inTimeDomain, unsigned inputMode) {
  media_[0]->runSignalSimulation(inputProf, inTimeDomain, inputMode);
  for (size_t i = 1; i < media_.size(); i++) {
    unsigned maxIn = 0;
    for (const auto& c : connections_[i - 1]) maxIn = std::max(maxIn, c.first);
    std::vector<cd> next((size_t)N_ * (maxIn + 1), cd(0, 0));
    for (const auto& c : connections_[i - 1]) {
      Grid& f = media_[i - 1]->signalFreq(c.second);
      std::copy(f.row(f.rows - 1), f.row(f.rows - 1) + N_, next.begin() + (size_t)c.first * N_);
    }
    media_[i]->runSignalSimulation(next, false);
  }
}


std::pair<Grid, Grid> Cascade::computeGreensFunction(bool inTimeDomain, bool runPump, unsigned nThreads, bool normalize,
                                                     const std::vector<uint8_t>& useInput,
                                                     const std::vector<uint8_t>& useOutput) {
  if (runPump) runPumpSimulation();
  if (media_.size() == 1)
    return media_[0]->computeGreensFunction(inTimeDomain, false, nThreads, normalize, useInput, useOutput);
  // every medium's Green's function stays on the device; the composition
  // (src/Cascade.cpp:128-129) runs there and only the result is copied out
  std::vector<nlm_plan*> plans;
  int rows = 0, cols = 0;
  for (size_t i = 0; i < media_.size(); i++) {
    const bool first = i == 0, last = i == media_.size() - 1;
    Medium& m = *media_[i];
    m.greensImpl(inTimeDomain, false, nThreads, normalize, first ? useInput : std::vector<uint8_t>{},
                 last ? useOutput : std::vector<uint8_t>{}, false);
    plans.push_back(m.plan());
    if (first) cols = m.lastGreensCols_;
    if (last) rows = m.lastGreensRows_;
  }
  std::pair<Grid, Grid> res;
  res.first.setZero(rows, cols);
  res.second.setZero(rows, cols);
  check(nlm_greens_cascade((int)plans.size(), plans.data(), reinterpret_cast<double*>(res.first.v.data()),
                           reinterpret_cast<double*>(res.second.v.data())));
  return res;
}

Grid Cascade::batchSignalSimulation(const Grid& inputProfs, bool inTimeDomain, bool runPump, unsigned nThreads,
                                    unsigned inputMode, const std::vector<uint8_t>& useOutput) {
  if (runPump) runPumpSimulation();
  Grid out = media_[0]->batchSignalSimulation(inputProfs, inTimeDomain, false, nThreads, inputMode, {});
  for (size_t i = 1; i < media_.size(); i++) {
    if (i != media_.size() - 1)
      out = media_[i]->batchSignalSimulation(out, inTimeDomain, false, nThreads, 0, {});
    else
      out = media_[i]->batchSignalSimulation(out, inTimeDomain, false, nThreads, 0, useOutput);
  }
  return out;
}

// ---------------------------------------------------------------------------
// MultiBatch
// ---------------------------------------------------------------------------

MultiBatch::MultiBatch(const std::vector<Medium*>& media, const std::vector<std::vector<cd>>& inputs,
                       bool inTimeDomain, bool autotune, bool stageRows)
    : media_(media) {
  if (media.size() != inputs.size()) throw std::invalid_argument("one input row per medium is required");
  const size_t n = media.size();
  std::vector<nlm_multi_item> items(n);
  std::vector<std::vector<double>> strengths(n);
  for (size_t i = 0; i < n; i++) {
    Medium* md = media[i];
    if (!md) throw std::invalid_argument("null medium");
    nlm_multi_item& it = items[i];
    std::memset(&it, 0, sizeof it);
    it.plan = md->plan();
    it.input = reinterpret_cast<const double*>(inputs[i].data());
    it.cols = (int32_t)inputs[i].size();
    it.in_time_domain = inTimeDomain ? 1 : 0;
    it.input_mode = 0;
    it.n_out = md->nS_;
    for (int q = 0; q < md->nS_ && q < 4; q++) it.outputs[q] = q;
    if (md->nP_ == 1 && md->model_ != NLM_CHI3) {  // single linear pump: generated in the batch
      it.envelope = reinterpret_cast<const double*>(md->envelope_[0].data());
      it.disp_pump = md->dispPump_[0].data();
      it.dzp = md->dzp_;
      if (md->profile_ != IntensityProfile::Constant) {
        const std::vector<double> x = md->linSpaced(md->nZP_, -0.5 * md->z_, 0.5 * md->z_);
        strengths[i].resize(md->nZP_);
        for (int r = 0; r < md->nZP_; r++) {
          const double u = x[r] / md->zR_;
          strengths[i][r] = md->profile_ == IntensityProfile::GaussianApodization ? std::exp(-0.5 * (u * u))
                                                                                  : 1 / std::sqrt(1 + u * u);
        }
        it.strength = strengths[i].data();
      }
    } else if (md->nP_ > 0) {
      // Chi3's nonlinear pump and Chi2SFGOPA's two pumps: the device pump grids
      // of the medium, computed here as batchSignalSimulation(runPump=True) does
      md->runPumpSimulation();
    }
    outLen_.push_back(md->nS_ * md->N_);
  }
  check(nlm_multi_create((int)n, items.data(), (autotune ? 1 : 0) | (stageRows ? 2 : 0), &m_));
}

MultiBatch::~MultiBatch() {
  if (m_) nlm_multi_destroy(m_);
}

void MultiBatch::launch() { check(nlm_multi_launch(m_)); }

double MultiBatch::sync() {
  double ms = 0;
  check(nlm_multi_sync(m_, &ms));
  return ms;
}

std::vector<cd> MultiBatch::output(int i) const {
  if (i < 0 || i >= (int)media_.size()) throw std::out_of_range("instance index out of range");
  std::vector<cd> out(outLen_[i]);
  check(nlm_multi_output(m_, i, reinterpret_cast<double*>(out.data())));
  return out;
}

std::vector<std::array<int, 4>> MultiBatch::classes() const {
  std::vector<std::array<int, 4>> r;
  const int nc = nlm_multi_classes(m_);
  for (int c = 0; c < nc; c++) {
    int32_t model, l, count, grid;
    check(nlm_multi_grid(m_, c, &model, &l, &count, &grid));
    r.push_back({model, l, count, grid});
  }
  return r;
}

}  // namespace nlm
