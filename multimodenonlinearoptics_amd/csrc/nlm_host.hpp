// nlm_host.hpp -- host object model of the `nonlinearmedium` drop-in.
//
// Mirrors the reference's class tree (src/_NonlinearMedium.hpp:20-142,
// src/_FullyNonlinearMedium.hpp:6-69, src/Cascade.hpp:6-44 and the solver
// classes in solvers/*.cpp).  Construction, grids, dispersion and poling are
// computed here on the host exactly as the reference does; every propagation
// (pump, signal, batch, Green's function) runs on the MI355X through the C-ABI
// of include/nlm.h.  There is no CPU propagation path: a run without a usable
// GPU raises an error.
#pragma once
#include <complex>
#include <cstdint>
#include <array>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "nlm.h"

namespace nlm {

using cd = std::complex<double>;

enum class PulseType : int { Gaussian = 0, Sech = 1, Sinc = 2 };
enum class IntensityProfile : int { GaussianBeam = 0, Constant = 1, GaussianApodization = 2 };

// Row-major 2-D complex array (the reference's Array2Dcd).
struct Grid {
  int rows = 0, cols = 0;
  std::vector<cd> v;
  void setZero(int r, int c) {
    rows = r;
    cols = c;
    v.assign((size_t)r * c, cd(0, 0));
  }
  cd* row(int i) { return v.data() + (size_t)i * cols; }
  const cd* row(int i) const { return v.data() + (size_t)i * cols; }
};

// Host transforms for set-up work only (pump chirp, Raman spectrum, pump import
// from another medium): forward unscaled, inverse scaled by 1/n.
void host_fft(const cd* in, cd* out, int n, bool inverse);

// Raise the reference's exception type for an engine status code.
void check(int status);

class Medium {
  friend class Cascade;
  friend class MultiBatch;

 public:
  virtual ~Medium();

  virtual void setPump(PulseType pulseType, double chirpLength = 0, double delayLength = 0, unsigned pumpIndex = 0);
  virtual void setPump(const std::vector<cd>& customPump, double chirpLength = 0, double delayLength = 0,
                       unsigned pumpIndex = 0);
  virtual void setPump(const Medium& other, unsigned modeIndex, double delayLength = 0, unsigned pumpIndex = 0);

  virtual void runPumpSimulation();
  virtual void runSignalSimulation(const std::vector<cd>& inputProf, bool inTimeDomain = true, unsigned inputMode = 0);
  virtual std::pair<Grid, Grid> computeGreensFunction(bool inTimeDomain = false, bool runPump = true,
                                                      unsigned nThreads = 1, bool normalize = false,
                                                      const std::vector<uint8_t>& useInput = {},
                                                      const std::vector<uint8_t>& useOutput = {});
  virtual Grid batchSignalSimulation(const Grid& inputProfs, bool inTimeDomain = false, bool runPump = true,
                                     unsigned nThreads = 1, unsigned inputMode = 0,
                                     const std::vector<uint8_t>& useOutput = {});

  // grid views (host copies refreshed from the device on access)
  Grid& pumpFreq(unsigned i = 0);
  Grid& pumpTime(unsigned i = 0);
  Grid& signalFreq(unsigned i = 0);
  Grid& signalTime(unsigned i = 0);
  Grid& field(unsigned i = 0);
  const std::vector<double>& tau() const { return tau_; }
  const std::vector<double>& omega() const { return omega_; }
  const std::vector<double>& poling() const { return poling_; }

  // extension: per-z-step poling value (continuous), length nZSteps
  void setPolingProfile(const std::vector<double>& profile);

  int nFreqs() const { return N_; }
  int nZSteps() const { return nZ_; }
  int nSignalModes() const { return nS_; }
  int nPumpModes() const { return nP_; }
  int model() const { return model_; }
  double dz() const { return dz_; }
  double length() const { return z_; }
  const std::vector<cd>& nlStep() const { return nlStep_; }

  // the device plan (created on first use); exposed for timing and streams
  nlm_plan* plan();
  void setStream(void* hipStream);
  // Green's function with C and S left in device memory (no host copy); the
  // default mode selection of computeGreensFunction.  For timing the engine.
  void greensOnDevice(bool inTimeDomain, bool runPump);
  // differentiable Green's function (single-mode Chi2PDC): forward with field
  // checkpoints every ckptEvery steps (0: ceil(sqrt(nZ-1))), then the adjoint;
  // returns (dL/d envelope [N] complex, dL/d poling [nZ] real) for dL/dC, dL/dS
  std::pair<Grid, Grid> greensForward(bool inTimeDomain, bool runPump, int ckptEvery);
  // gC, gS: rows x cols complex, row-major (dL/dC, dL/dS of the last greensForward)
  std::pair<std::vector<cd>, std::vector<double>> greensAdjoint(const cd* gC, const cd* gS, int rows, int cols);
  // differentiable batch (Chi3GNLSE): forward keeping checkpoints, then dL/d(inputs)
  Grid batchForward(const Grid& inputProfs, bool inTimeDomain, int ckptEvery);
  // z-resolved batch: runSignalSimulation of every row in one launch; grids of
  // shape (rows * nS * nZ, N), frequency and time domain
  std::pair<Grid, Grid> batchSignalGrids(const Grid& inputProfs, bool inTimeDomain, unsigned inputMode);
  Grid batchAdjoint(const Grid& gOut);
  void enableTiming(bool on);
  void setStepTable(bool on);
  std::vector<double> lastTiming();
  void sync();
  unsigned long long forwardSerial() const { return forwardSerial_; }

  const std::vector<cd>& raman() const { return raman_; }

  struct Args {
    int model;
    unsigned nS, nP;
    bool canBePoled;
    double relativeLength;
    std::vector<double> nl, beta2, beta2s, beta1, beta1s, beta3, beta3s, diffBeta0;
    double rayleighLength, tMax;
    unsigned tPrecision, zPrecision;
    IntensityProfile profile;
    double chirp, delay;
    std::vector<cd> customPump;
    PulseType pulseType;
    std::vector<double> poling;
    unsigned nZOverride = 0;
    std::vector<double> polingProfile;
  };

 protected:
  Medium() = default;
  explicit Medium(const Args& a);

  void setLengths(const Args& a, double rayleighLength);
  void resetGrids(unsigned nFreqs, double tMax);
  void setDispersion(const Args& a);
  void setPoling(const std::vector<double>& poling);
  void applyChirpDelay(unsigned q, double chirp, double delay);

  // per-step scalar factors of the model's DiffEq, [nZ][3][NC] and constants
  virtual void stepCoefficients(std::vector<cd>& coef, std::vector<cd>& cconst) const = 0;
  virtual bool fullyNonlinear() const { return false; }
  void gnlseTables(std::vector<cd>& raman, std::vector<cd>& filter) const;

  // scalar helpers used by stepCoefficients
  std::complex<double> mismatch(unsigned which, double x) const;
  double relativeAmplitude(double i) const;
  double relativeIntensity(double i) const;
  static std::vector<double> linSpaced(int n, double low, double high);

  std::pair<Grid, Grid> greensImpl(bool inTimeDomain, bool runPump, unsigned nThreads, bool normalize,
                                   const std::vector<uint8_t>& useInput, const std::vector<uint8_t>& useOutput,
                                   bool download);
  void markPumpOnDevice();
  void syncPumpToHost(unsigned i);
  void syncSignalToHost();
  void uploadPump(unsigned q);
  void refreshCoef();

  int model_ = 0;
  int nS_ = 0, nP_ = 0;
  double z_ = 0, dz_ = 0, dzp_ = 0, tMax_ = 0, zR_ = 0;
  int nZ_ = 0, nZP_ = 0, N_ = 0;
  IntensityProfile profile_ = IntensityProfile::Constant;
  std::vector<double> beta1_, beta2_, diffBeta0_;
  std::vector<cd> nlStep_;
  std::vector<std::vector<cd>> envelope_;
  std::vector<double> poling_;
  std::vector<double> tau_, omega_;
  std::vector<std::vector<double>> dispPump_, dispSign_;
  std::vector<std::vector<cd>> dispStepPump_, dispStepSign_;
  std::vector<Grid> pumpFreq_, pumpTime_, signalFreq_, signalTime_, field_;
  void allocPump(unsigned q);
  void allocSignal();
  std::vector<cd> raman_;

  nlm_plan* plan_ = nullptr;
  void* stream_ = nullptr;
  bool coefDirty_ = false;
  std::vector<bool> pumpHostStale_, pumpDevStale_;
  bool signalHostStale_ = false;
  bool signalOnDevice_ = false;  // the plan holds the signal grids of the last runSignalSimulation
  bool pumpComputed_ = false;
  int batchRows_ = 0;  // rows of the last batchForward
  // number of differentiable forwards run so far (greensForward, batchForward): an
  // autograd caller records it after its forward and checks it before the adjoint
  unsigned long long forwardSerial_ = 0;
  int lastGreensRows_ = 0, lastGreensCols_ = 0;  // shape of the last Green's function
};

class FullyNonlinearMedium : public Medium {
 public:
  Grid batchSignalSimulationFNLM(const Grid& inputProfs, bool inTimeDomain = false, unsigned nThreads = 1,
                                 unsigned inputMode = 0, const std::vector<uint8_t>& useOutput = {}) {
    return Medium::batchSignalSimulation(inputProfs, inTimeDomain, false, nThreads, inputMode, useOutput);
  }
  void setPump(PulseType, double, double, unsigned) override { noMethod(); }
  void setPump(const std::vector<cd>&, double, double, unsigned) override { noMethod(); }
  void setPump(const Medium&, unsigned, double, unsigned) override { noMethod(); }
  void runPumpSimulation() override { noMethod(); }
  std::pair<Grid, Grid> computeGreensFunction(bool, bool, unsigned, bool, const std::vector<uint8_t>&,
                                              const std::vector<uint8_t>&) override {
    noMethod();
    return {};
  }

 protected:
  explicit FullyNonlinearMedium(const Args& a);
  bool fullyNonlinear() const override { return true; }
  [[noreturn]] static void noMethod() { throw std::runtime_error("Object does not have this method."); }
};

// Factory for the fifteen solvers: builds the constructor mapping of the named
// class (kwargs in the reference's order, see the pybind11 module).
std::unique_ptr<Medium> makeSolver(const std::string& name, const std::vector<double>& num,
                                   const std::vector<cd>& customPump, int pulseType, int intensityProfile,
                                   const std::vector<double>& poling, unsigned tPrecision, unsigned zPrecision,
                                   unsigned nZOverride, const std::vector<double>& polingProfile);

// The fifteen solver classes (solvers/*.cpp).  Each maps its constructor
// arguments onto Args (see makeSolver) and supplies its DiffEq's per-step
// scalar factors.
#define NLM_DECLARE_SOLVER(NAME, BASE)                                                   \
  class NAME : public BASE {                                                            \
   public:                                                                              \
    explicit NAME(const Args& a);                                                       \
                                                                                        \
   protected:                                                                           \
    void stepCoefficients(std::vector<cd>& coef, std::vector<cd>& cconst) const override; \
  };
NLM_DECLARE_SOLVER(Chi2PDC, Medium)
NLM_DECLARE_SOLVER(Chi2SFG, Medium)
NLM_DECLARE_SOLVER(Chi2AFC, Medium)
NLM_DECLARE_SOLVER(Chi2SFGPDC, Medium)
NLM_DECLARE_SOLVER(Chi2SFGII, Medium)
NLM_DECLARE_SOLVER(Chi2PDCII, Medium)
NLM_DECLARE_SOLVER(Chi2SFGOPA, Medium)
NLM_DECLARE_SOLVER(Chi2SFGXPM, Medium)
NLM_DECLARE_SOLVER(Chi2SHG, FullyNonlinearMedium)
NLM_DECLARE_SOLVER(Chi2SHGOPA, FullyNonlinearMedium)
NLM_DECLARE_SOLVER(Chi2DSFG, FullyNonlinearMedium)
NLM_DECLARE_SOLVER(Chi2SHGXPM, FullyNonlinearMedium)
NLM_DECLARE_SOLVER(Chi2ASHG, FullyNonlinearMedium)
#undef NLM_DECLARE_SOLVER

class Chi3 : public Medium {
 public:
  explicit Chi3(const Args& a);
  void runPumpSimulation() override;

 protected:
  void stepCoefficients(std::vector<cd>& coef, std::vector<cd>& cconst) const override;
};

class Chi3GNLSE : public FullyNonlinearMedium {
 public:
  Chi3GNLSE(const Args& a, double fr, double fb, double tau1, double tau2, double tau3);

 protected:
  void stepCoefficients(std::vector<cd>& coef, std::vector<cd>& cconst) const override;
};

// Heterogeneous batch (extension): one input row for each of many media of
// different models, sizes and parameters, propagated together by one graph
// launch of per-(model, nFreqs) persistent kernels (nlm_multi_* in nlm.h).
// Each output equals media[i].batchSignalSimulation(inputs[i:i+1], inTimeDomain,
// runPump=True) of that medium (src/_NonlinearMedium.cpp:378-444).
class MultiBatch {
 public:
  // stageRows: nFreqs 4096 media with a generated pump stage their pump rows in a
  // device buffer at the start of each launch (see nlm_multi_create)
  MultiBatch(const std::vector<Medium*>& media, const std::vector<std::vector<cd>>& inputs, bool inTimeDomain,
             bool autotune, bool stageRows = false);
  ~MultiBatch();
  MultiBatch(const MultiBatch&) = delete;
  MultiBatch& operator=(const MultiBatch&) = delete;
  void launch();
  double sync();  // waits; device milliseconds of the last launch
  double run() {
    launch();
    return sync();
  }
  std::vector<cd> output(int i) const;
  int size() const { return (int)media_.size(); }
  std::vector<std::array<int, 4>> classes() const;  // (model, log2 N, instances, persistent workgroups)

 private:
  std::vector<Medium*> media_;
  std::vector<int> outLen_;
  nlm_multi* m_ = nullptr;
};

class Cascade : public Medium {
 public:
  Cascade(const std::vector<Medium*>& inputMedia, const std::vector<std::map<unsigned, unsigned>>& connections,
          bool sharePump);
  void addMedium(Medium& medium, const std::map<unsigned, unsigned>& connection);

  void setPump(PulseType pulseType, double chirpLength = 0, double delayLength = 0, unsigned pumpIndex = 0) override;
  void setPump(const std::vector<cd>& customPump, double chirpLength = 0, double delayLength = 0,
               unsigned pumpIndex = 0) override;
  void setPump(const Medium&, unsigned, double, unsigned) override {
    throw std::runtime_error("Object does not have this method.");
  }
  void runPumpSimulation() override;
  void runSignalSimulation(const std::vector<cd>& inputProf, bool inTimeDomain = true, unsigned inputMode = 0) override;
  std::pair<Grid, Grid> computeGreensFunction(bool inTimeDomain = false, bool runPump = true, unsigned nThreads = 1,
                                              bool normalize = false, const std::vector<uint8_t>& useInput = {},
                                              const std::vector<uint8_t>& useOutput = {}) override;
  Grid batchSignalSimulation(const Grid& inputProfs, bool inTimeDomain = false, bool runPump = true,
                             unsigned nThreads = 1, unsigned inputMode = 0,
                             const std::vector<uint8_t>& useOutput = {}) override;

  Medium& getMedium(unsigned i) { return *media_.at(i); }
  const std::vector<Medium*>& getMedia() const { return media_; }
  unsigned getNMedia() const { return (unsigned)media_.size(); }
  const std::vector<double>& cascadeTau() const { return media_.at(0)->tau(); }
  const std::vector<double>& cascadeOmega() const { return media_.at(0)->omega(); }

 protected:
  void stepCoefficients(std::vector<cd>&, std::vector<cd>&) const override {}

 private:
  std::vector<Medium*> media_;
  std::vector<std::map<unsigned, unsigned>> connections_;
  bool sharedPump_ = false;
};

}  // namespace nlm
