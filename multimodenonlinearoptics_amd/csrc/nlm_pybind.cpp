// nlm_pybind.cpp -- the `nonlinearmedium` Python module (drop-in for the
// reference's pybind11 binding, src/nlmModulePy.cpp:12-217 and the NLMMODULE
// block of every solver).  Class names, base classes, argument names, order
// and defaults follow the reference; the only additions are keyword-only
// extensions (nZSteps=, polingProfile=) and a setStream() method.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <limits>

#include "nlm_host.hpp"

namespace py = pybind11;
using namespace pybind11::literals;
using nlm::cd;
using nlm::Grid;
using nlm::Medium;

using carr = py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>;
using darr = py::array_t<double, py::array::c_style | py::array::forcecast>;

static std::vector<cd> to_vec(const carr& a) { return std::vector<cd>(a.data(), a.data() + a.size()); }
static std::vector<double> to_dvec(const darr& a) { return std::vector<double>(a.data(), a.data() + a.size()); }

static Grid to_grid(const carr& a) {
  Grid g;
  if (a.ndim() == 1) {
    g.rows = 1;
    g.cols = (int)a.shape(0);
  } else if (a.ndim() == 2) {
    g.rows = (int)a.shape(0);
    g.cols = (int)a.shape(1);
  } else {
    throw std::invalid_argument("expected a 1-D or 2-D array");
  }
  g.v.assign(a.data(), a.data() + a.size());
  return g;
}

// read-only view of a host grid that keeps `owner` alive
static py::array grid_view(Grid& g, py::handle owner) {
  py::array a(py::dtype::of<std::complex<double>>(), {(py::ssize_t)g.rows, (py::ssize_t)g.cols},
              {(py::ssize_t)(g.cols * sizeof(cd)), (py::ssize_t)sizeof(cd)}, g.v.data(), owner);
  a.attr("setflags")("write"_a = false);
  return a;
}

static py::array vec_view(const std::vector<double>& v, py::handle owner) {
  py::array a(py::dtype::of<double>(), {(py::ssize_t)v.size()}, {(py::ssize_t)sizeof(double)}, v.data(), owner);
  a.attr("setflags")("write"_a = false);
  return a;
}

static py::array cvec_view(const std::vector<cd>& v, py::handle owner) {
  py::array a(py::dtype::of<std::complex<double>>(), {(py::ssize_t)v.size()}, {(py::ssize_t)sizeof(cd)}, v.data(),
              owner);
  a.attr("setflags")("write"_a = false);
  return a;
}

// owning array from a grid (moved)
static py::array owned(Grid&& g) {
  auto* h = new Grid(std::move(g));
  py::capsule cap(h, [](void* p) { delete reinterpret_cast<Grid*>(p); });
  return py::array(py::dtype::of<std::complex<double>>(), {(py::ssize_t)h->rows, (py::ssize_t)h->cols},
                   {(py::ssize_t)(h->cols * sizeof(cd)), (py::ssize_t)sizeof(cd)}, h->v.data(), cap);
}

static std::vector<uint8_t> to_flags(const py::object& o) {
  std::vector<uint8_t> r;
  if (o.is_none()) return r;
  for (auto item : o) r.push_back(item.cast<uint8_t>());
  return r;
}

static const double inf = std::numeric_limits<double>::infinity();

template <class Cls>
static void bind_base_methods(Cls& c) {
  c.def("setPump",
        [](Medium& m, nlm::PulseType pt, double chirp, double delay, unsigned pumpIndex) {
          m.setPump(pt, chirp, delay, pumpIndex);
        },
        "Set the input shape of the pump\n"
        "pulseType Gaussian, Sech or Sinc profile; 0, 1, 2 respectively.\n"
        "chirp     Initial chirp of the pump, specified in dispersion lengths.\n"
        "delay     Initial time delay of the pump, specified in walk-off lengths.\n"
        "pumpIndex Index of the pump being set, if applicable.",
        "pulseType"_a, "chirp"_a = 0, "delay"_a = 0, "pumpIndex"_a = 0);
  c.def("setPump",
        [](Medium& m, const carr& custom, double chirp, double delay, unsigned pumpIndex) {
          m.setPump(to_vec(custom), chirp, delay, pumpIndex);
        },
        "Set the input shape of the pump\n"
        "customPump An arbitrary pump shape specified in the time domain, with self.tau as the axis.\n"
        "chirp     Initial chirp of the pump, specified in dispersion lengths.\n"
        "delay     Initial time delay of the pump, specified in walk-off lengths.\n"
        "pumpIndex Index of the pump being set, if applicable.",
        "customPump"_a, "chirp"_a = 0, "delay"_a = 0, "pumpIndex"_a = 0);
  c.def("setPump",
        [](Medium& m, Medium& other, unsigned modeIndex, double delay, unsigned pumpIndex) {
          m.setPump(other, modeIndex, delay, pumpIndex);
        },
        "Set the pump over the whole propagation length by copying from another simulation, accounting for the "
        "frame of reference\n"
        "Note: do not call runPumpSimulation after this function or the pump simulation will be overwritten.\n"
        "other     A NonlinearMedium instance with the same frequency axis, and resolution greater than or equal to "
        "the pump's.\n"
        "modeIndex The index of the mode in 'other' to use as pump in this simulation.\n"
        "delay     Initial time delay of the pump, specified in walk-off lengths.\n"
        "pumpIndex Index of the pump being set, if applicable.",
        "other"_a, "modeIndex"_a = 0, "delay"_a = 0, "pumpIndex"_a = 0);
  c.def("runPumpSimulation",
        [](Medium& m) {
          py::gil_scoped_release nogil;
          m.runPumpSimulation();
        },
        "Simulate propagation of the pump field");
  c.def("runSignalSimulation",
        [](Medium& m, const carr& in, bool inTimeDomain, unsigned inputMode) {
          std::vector<cd> v = to_vec(in);
          py::gil_scoped_release nogil;
          m.runSignalSimulation(v, inTimeDomain, inputMode);
        },
        "Simulate propagation of the signal field(s)\n"
        "inputProf    Profile of input pulse. May be time or frequency domain.\n"
        "             Note: Input is assumed to have self.omega or self.tau as the axis.\n"
        "inTimeDomain Specify if the input is in time or frequency domain.\n"
        "inputMode    Specify which signal mode of the nonlinear medium the input corresponds to.",
        "inputProf"_a, "inTimeDomain"_a = true, "inputMode"_a = 0);
  c.def("computeGreensFunction",
        [](Medium& m, bool inTimeDomain, bool runPump, unsigned nThreads, bool normalize, const py::object& useInput,
           const py::object& useOutput) {
          auto ui = to_flags(useInput), uo = to_flags(useOutput);
          std::pair<Grid, Grid> r;
          {
            py::gil_scoped_release nogil;
            r = m.computeGreensFunction(inTimeDomain, runPump, nThreads, normalize, ui, uo);
          }
          return py::make_tuple(owned(std::move(r.first)), owned(std::move(r.second)));
        },
        "Solve the Green's function matrices C and S where a(L) = C a(0) + S [a(0)]^t\n"
        "inTimeDomain Compute the Green's function in time or frequency domain.\n"
        "runPump      Whether to run the pump simulation before signal simulations.\n"
        "nThreads     Number of threads used to run simulations in parallel.\n"
        "normalize    Whether to adjust the final matrix so that amplitudes are converted 1-to-1 between different "
        "modes.\n"
        "useInput     Specify which inputs modes to include in the transformation. Default is all inputs.\n"
        "useOutput    Specify which output modes to include in the transformation. Default is all outputs.\n"
        "return: Green's function matrices C, S",
        "inTimeDomain"_a = false, "runPump"_a = true, "nThreads"_a = 1, "normalize"_a = false,
        "useInput"_a = py::list(), "useOutput"_a = py::list());
  c.def("batchSignalSimulation",
        [](Medium& m, const carr& in, bool inTimeDomain, bool runPump, unsigned nThreads, unsigned inputMode,
           const py::object& useOutput) {
          Grid g = to_grid(in);
          auto uo = to_flags(useOutput);
          Grid r;
          {
            py::gil_scoped_release nogil;
            r = m.batchSignalSimulation(g, inTimeDomain, runPump, nThreads, inputMode, uo);
          }
          return owned(std::move(r));
        },
        "Run multiple signal simulations.\n"
        "inputProfs   Profiles of input pulses. May be time or frequency domain.\n"
        "             Note: Input is assumed to have self.omega or self.tau as the axis.\n"
        "inTimeDomain Specify if the input is in time or frequency domain.\n"
        "runPump      Whether to run the pump simulation before signal simulations.\n"
        "nThreads     Number of threads used to run simulations in parallel.\n"
        "inputMode    Specify which signal mode of the nonlinear medium the input corresponds to.\n"
        "useOutput    Specify which signal mode outputs to return. Default is all outputs.\n"
        "return: Array of signal profiles at the output of the medium",
        "inputProfs"_a, "inTimeDomain"_a = false, "runPump"_a = true, "nThreads"_a = 1, "inputMode"_a = 0,
        "useOutput"_a = py::list());
}

PYBIND11_MODULE(nonlinearmedium, m) {
  m.doc() = "Module for numerical simulation of classical or quantum signals in pumped nonlinear media "
            "(MI355X engine).";

  py::class_<Medium> base(m, "_NonlinearMedium", "Base class for nonlinear medium solvers");
  py::class_<nlm::Cascade, Medium> cascade(m, "Cascade");
  py::class_<nlm::FullyNonlinearMedium, Medium> fnlm(m, "_FullyNonlinearMedium",
                                                     "Base class for fully nonlinear medium solvers");

  bind_base_methods(base);
  base.def_property_readonly("omega", [](py::object self) { return vec_view(self.cast<Medium&>().omega(), self); },
                             "Read-only frequency axis of the system.");
  base.def_property_readonly("tau", [](py::object self) { return vec_view(self.cast<Medium&>().tau(), self); },
                             "Read-only time axis of the system");
  base.def_property_readonly("signalFreq",
                             [](py::object self) { return grid_view(self.cast<Medium&>().signalFreq(0), self); },
                             "Read-only array of the signal frequency profile along the length of the medium.");
  base.def_property_readonly("signalTime",
                             [](py::object self) { return grid_view(self.cast<Medium&>().signalTime(0), self); },
                             "Read-only array of a signal time profile along the length of the medium.");
  base.def_property_readonly("pumpFreq",
                             [](py::object self) { return grid_view(self.cast<Medium&>().pumpFreq(0), self); },
                             "Read-only array of the pump frequency profile along the length of the medium.");
  base.def_property_readonly("pumpTime",
                             [](py::object self) { return grid_view(self.cast<Medium&>().pumpTime(0), self); },
                             "Read-only array of the pump time profile along the length of the medium.");
  base.def("signalFreqs", [](py::object self, unsigned i) { return grid_view(self.cast<Medium&>().signalFreq(i), self); },
           "Read-only array of a signal frequency profile along the length of the medium.", "i"_a = 0);
  base.def("signalTimes", [](py::object self, unsigned i) { return grid_view(self.cast<Medium&>().signalTime(i), self); },
           "Read-only array of a signal time profile along the length of the medium.", "i"_a = 0);
  base.def("pumpFreqs", [](py::object self, unsigned i) { return grid_view(self.cast<Medium&>().pumpFreq(i), self); },
           "Read-only array of a pump frequency profile along the length of the medium.", "i"_a = 0);
  base.def("pumpTimes", [](py::object self, unsigned i) { return grid_view(self.cast<Medium&>().pumpTime(i), self); },
           "Read-only array of a pump time profile along the length of the medium.", "i"_a = 0);
  base.def_property_readonly("poling", [](py::object self) { return vec_view(self.cast<Medium&>().poling(), self); },
                             "Read-only array of the domain poling along the length of a Chi(2) medium.");
  base.def("field",
           [](py::object self, unsigned i) {
             Grid& g = self.cast<Medium&>().field(i);
             return py::array(py::dtype::of<std::complex<double>>(), {(py::ssize_t)g.rows, (py::ssize_t)g.cols},
                              {(py::ssize_t)(g.cols * sizeof(cd)), (py::ssize_t)sizeof(cd)}, g.v.data(), self);
           },
           "Writeable array for some arbitrary field profile. Has the same shape as the pump arrays.", "i"_a = 0);
  // extensions
  base.def_property_readonly("nZSteps", &Medium::nZSteps, "Number of z steps of the signal grid.");
  base.def("setPolingProfile", [](Medium& md, const darr& p) { md.setPolingProfile(to_dvec(p)); },
           "Set the per-step poling value directly (length nZSteps; extension).", "profile"_a);
  base.def("setStream", [](Medium& md, std::uintptr_t s) { md.setStream(reinterpret_cast<void*>(s)); },
           "Run this medium's device work on the given hipStream_t (0: the medium's own stream).", "stream"_a);
  base.def("greensOnDevice",
           [](Medium& md, bool inTimeDomain, bool runPump) {
             py::gil_scoped_release nogil;
             md.greensOnDevice(inTimeDomain, runPump);
           },
           "Compute the Green's function and keep C and S in device memory (extension, for timing).",
           "inTimeDomain"_a = false, "runPump"_a = true);
  base.def("greensForward",
           [](Medium& md, bool inTimeDomain, bool runPump, int ckptEvery) {
             std::pair<Grid, Grid> r;
             {
               py::gil_scoped_release nogil;
               r = md.greensForward(inTimeDomain, runPump, ckptEvery);
             }
             return py::make_tuple(owned(std::move(r.first)), owned(std::move(r.second)));
           },
           "Green's function that keeps field checkpoints for greensAdjoint (extension; Chi2PDC).",
           "inTimeDomain"_a = false, "runPump"_a = true, "ckptEvery"_a = 0);
  base.def("greensAdjoint",
           [](Medium& md, const carr& gC, const carr& gS) {
             // the arrays are contiguous complex128 (converted if needed) and stay
             // alive for the call: hand their buffers over without a copy
             if (gC.ndim() != 2 || gS.ndim() != 2 || gC.shape(0) != gS.shape(0) || gC.shape(1) != gS.shape(1))
               throw std::invalid_argument("gradient shapes do not match the Green's function");
             const int rows = (int)gC.shape(0), cols = (int)gC.shape(1);
             std::pair<std::vector<cd>, std::vector<double>> r;
             {
               py::gil_scoped_release nogil;
               r = md.greensAdjoint(gC.data(), gS.data(), rows, cols);
             }
             py::array_t<std::complex<double>> ge(r.first.size());
             std::copy(r.first.begin(), r.first.end(), ge.mutable_data());
             py::array_t<double> gp(r.second.size());
             std::copy(r.second.begin(), r.second.end(), gp.mutable_data());
             return py::make_tuple(ge, gp);
           },
           "Gradients (dL/d envelope, dL/d poling profile) of the last greensForward for dL/dC, dL/dS "
           "(convention dL/dRe + i dL/dIm) (extension; Chi2PDC).", "gC"_a, "gS"_a);
  base.def("batchSignalGrids",
           [](Medium& md, const carr& in, bool inTimeDomain, unsigned inputMode) {
             Grid g = to_grid(in);
             std::pair<Grid, Grid> r;
             {
               py::gil_scoped_release nogil;
               r = md.batchSignalGrids(g, inTimeDomain, inputMode);
             }
             const py::ssize_t rows = g.rows, nS = (py::ssize_t)md.nSignalModes(), nZ = (py::ssize_t)md.nZSteps(),
                               N = (py::ssize_t)md.nFreqs();
             py::array f = owned(std::move(r.first)), t = owned(std::move(r.second));
             return py::make_tuple(f.reshape({rows, nS, nZ, N}), t.reshape({rows, nS, nZ, N}));
           },
           "runSignalSimulation of every input row in one launch (extension): signal grids "
           "[rows, nSignalModes, nZSteps, nFreqs] in the frequency and time domains.",
           "inputProfs"_a, "inTimeDomain"_a = true, "inputMode"_a = 0);
  base.def("batchForward",
           [](Medium& md, const carr& in, bool inTimeDomain, int ckptEvery) {
             Grid g = to_grid(in);
             Grid r;
             {
               py::gil_scoped_release nogil;
               r = md.batchForward(g, inTimeDomain, ckptEvery);
             }
             return owned(std::move(r));
           },
           "batchSignalSimulation that keeps field checkpoints on the device for batchAdjoint "
           "(extension; Chi3GNLSE).", "inputProfs"_a, "inTimeDomain"_a = false, "ckptEvery"_a = 0);
  base.def("batchAdjoint",
           [](Medium& md, const carr& gOut) {
             Grid g = to_grid(gOut);
             Grid r;
             {
               py::gil_scoped_release nogil;
               r = md.batchAdjoint(g);
             }
             return owned(std::move(r));
           },
           "dL/d(inputProfs) of the last batchForward for dL/d(outputs) (convention dL/dRe + i dL/dIm; "
           "extension; Chi3GNLSE).", "gOut"_a);
  base.def("planHandle", [](Medium& md) { return reinterpret_cast<std::uintptr_t>(md.plan()); },
           "The medium's nlm_plan* (include/nlm.h) as an integer, with its tables current, for callers of "
           "the C-ABI (extension).");
  base.def_property_readonly("forwardSerial", &Medium::forwardSerial,
                             "Number of greensForward / batchForward calls made on this medium (extension): "
                             "an adjoint belongs to the forward whose serial is current.");
  base.def("enableTiming", &Medium::enableTiming, "Record HIP-event kernel times of each run (extension).", "on"_a);
  base.def("setStepTable", &Medium::setStepTable,
           "One-mode undepleted media: propagate with precomputed RK4 step propagators (default) or evaluate "
           "the DiffEq per field (extension).", "on"_a);
  base.def("lastTiming", &Medium::lastTiming,
           "Kernel times of the last run in ms: [pump, propagation, assembly] (extension).");
  base.def("sync", [](Medium& md) {
    py::gil_scoped_release nogil;
    md.sync();
  }, "Wait for this medium's device work (extension).");

  py::enum_<nlm::PulseType>(m, "PulseType")
      .value("Gaussian", nlm::PulseType::Gaussian)
      .value("Sech", nlm::PulseType::Sech)
      .value("Sinc", nlm::PulseType::Sinc)
      .export_values();
  py::enum_<nlm::IntensityProfile>(m, "IntensityProfile")
      .value("GaussianBeam", nlm::IntensityProfile::GaussianBeam)
      .value("Constant", nlm::IntensityProfile::Constant)
      .value("GaussianApodization", nlm::IntensityProfile::GaussianApodization)
      .export_values();
  // the reference's scenario scripts pass pulseType as a plain integer
  // (tests/05 - Hosaka.py:48, tests/07 - Wasilewski.py:45).  IntensityProfile stays
  // strict: Chi2SFGII's swapped keyword slots (solvers/Chi2SFGII.cpp:97-106) must
  // reject the default zPrecision=100 landing in the profile slot.
  py::implicitly_convertible<int, nlm::PulseType>();

  // _FullyNonlinearMedium, src/nlmModulePy.cpp:140-164
  fnlm.def("batchSignalSimulation",
           [](nlm::FullyNonlinearMedium& md, const carr& in, bool inTimeDomain, unsigned nThreads, unsigned inputMode,
              const py::object& useOutput) {
             Grid g = to_grid(in);
             auto uo = to_flags(useOutput);
             Grid r;
             {
               py::gil_scoped_release nogil;
               r = md.batchSignalSimulationFNLM(g, inTimeDomain, nThreads, inputMode, uo);
             }
             return owned(std::move(r));
           },
           "Run multiple signal simulations.\n"
           "inputProfs   Profiles of input pulses. May be time or frequency domain.\n"
           "             Note: Input is assumed to have self.omega or self.tau as the axis.\n"
           "inTimeDomain Specify if the input is in time or frequency domain.\n"
           "nThreads     Number of threads used to run simulations in parallel.\n"
           "inputMode    Specify which signal mode of the nonlinear medium the input corresponds to.\n"
           "useOutput    Specify which signal mode outputs to return. Default is all outputs.\n"
           "return: Array of signal profiles at the output of the medium",
           "inputProfs"_a, "inTimeDomain"_a = false, "nThreads"_a = 1, "inputMode"_a = 0, "useOutput"_a = py::list());
  fnlm.def("setPump", [](Medium& md, nlm::PulseType pt, double c, double d, unsigned i) { md.setPump(pt, c, d, i); },
           "pulseType"_a, "chirp"_a = 0, "delay"_a = 0, "pumpIndex"_a = 0);
  fnlm.def("setPump", [](Medium& md, const carr& p, double c, double d, unsigned i) { md.setPump(to_vec(p), c, d, i); },
           "customPump"_a, "chirp"_a = 0, "delay"_a = 0, "pumpIndex"_a = 0);
  fnlm.def("setPump", [](Medium& md, Medium& o, unsigned s, double d, unsigned i) { md.setPump(o, s, d, i); },
           "other"_a, "signalIndex"_a = 0, "delayLength"_a = 0, "pumpIndex"_a = 0);
  fnlm.def("runPumpSimulation", [](Medium& md) { md.runPumpSimulation(); });
  fnlm.def("computeGreensFunction",
           [](Medium& md, bool a, bool b, unsigned c, bool d, const py::object& e, const py::object& f) {
             auto r = md.computeGreensFunction(a, b, c, d, to_flags(e), to_flags(f));
             return py::make_tuple(owned(std::move(r.first)), owned(std::move(r.second)));
           },
           "inTimeDomain"_a = false, "runPump"_a = true, "nThreads"_a = 1, "normalize"_a = false,
           "useInput"_a = py::list(), "useOutput"_a = py::list());

  // heterogeneous batch (extension)
  py::class_<nlm::MultiBatch>(m, "MultiBatch",
                              "One input row per medium, for many media of different models, sizes and parameters, "
                              "propagated together by one graph launch (extension).  output(i) equals "
                              "media[i].batchSignalSimulation(inputs[i][None], inTimeDomain)[0].")
      .def(py::init([](const py::list& media, const py::list& inputs, bool inTimeDomain, bool autotune,
                       bool stagePumpRows) {
             std::vector<Medium*> v;
             for (auto h : media) v.push_back(&h.cast<Medium&>());
             std::vector<std::vector<cd>> in;
             for (auto h : inputs) in.push_back(to_vec(h.cast<carr>()));
             return new nlm::MultiBatch(v, in, inTimeDomain, autotune, stagePumpRows);
           }),
           "media"_a, "inputs"_a, "inTimeDomain"_a = false, "autotune"_a = true, "stagePumpRows"_a = false,
           py::keep_alive<1, 2>())
      .def("launch", [](nlm::MultiBatch& b) {
        py::gil_scoped_release nogil;
        b.launch();
      })
      .def("sync", [](nlm::MultiBatch& b) {
        py::gil_scoped_release nogil;
        return b.sync();
      }, "Wait for the last launch; returns its device time in ms.")
      .def("run", [](nlm::MultiBatch& b) {
        py::gil_scoped_release nogil;
        return b.run();
      }, "launch() then sync(); returns the device time in ms.")
      .def("output", [](const nlm::MultiBatch& b, int i) {
        std::vector<cd> o = b.output(i);
        py::array_t<std::complex<double>> r(o.size());
        std::copy(o.begin(), o.end(), r.mutable_data());
        return r;
      }, "i"_a)
      .def("__len__", &nlm::MultiBatch::size)
      .def("classes", &nlm::MultiBatch::classes,
           "(model id, log2 nFreqs, instances, persistent workgroups) of each kernel class");

  // Cascade, src/nlmModulePy.cpp:170-211
  cascade.def(py::init([](const py::list& media, const std::vector<std::map<unsigned, unsigned>>& conns, bool share) {
                std::vector<Medium*> v;
                for (auto h : media) v.push_back(&h.cast<Medium&>());
                return new nlm::Cascade(v, conns, share);
              }),
              "inputMedia"_a, "modeConnections"_a, "sharePump"_a, py::keep_alive<1, 2>());
  cascade.def("setPump", [](nlm::Cascade& c, nlm::PulseType pt, double ch, double d, unsigned i) { c.setPump(pt, ch, d, i); },
              "pulseType"_a, "chirp"_a = 0, "delay"_a = 0, "pumpIndex"_a = 0);
  cascade.def("setPump", [](nlm::Cascade& c, const carr& p, double ch, double d, unsigned i) { c.setPump(to_vec(p), ch, d, i); },
              "customPump"_a, "chirp"_a = 0, "delay"_a = 0, "pumpIndex"_a = 0);
  cascade.def("setPump", [](nlm::Cascade& c, Medium& o, unsigned s, double d, unsigned i) { c.setPump(o, s, d, i); },
              "other"_a, "signalIndex"_a = 0, "delayLength"_a = 0, "pumpIndex"_a = 0);
  cascade.def("runPumpSimulation", [](nlm::Cascade& c) {
    py::gil_scoped_release nogil;
    c.runPumpSimulation();
  });
  cascade.def("runSignalSimulation",
              [](nlm::Cascade& c, const carr& in, bool t, unsigned im) {
                auto v = to_vec(in);
                py::gil_scoped_release nogil;
                c.runSignalSimulation(v, t, im);
              },
              "inputProf"_a, "inTimeDomain"_a = true, "inputMode"_a = 0);
  cascade.def("computeGreensFunction",
              [](nlm::Cascade& c, bool a, bool b, unsigned n, bool d, const py::object& e, const py::object& f) {
                auto ui = to_flags(e), uo = to_flags(f);
                std::pair<Grid, Grid> r;
                {
                  py::gil_scoped_release nogil;
                  r = c.computeGreensFunction(a, b, n, d, ui, uo);
                }
                return py::make_tuple(owned(std::move(r.first)), owned(std::move(r.second)));
              },
              "inTimeDomain"_a = false, "runPump"_a = true, "nThreads"_a = 1, "normalize"_a = false,
              "useInput"_a = py::list(), "useOutput"_a = py::list());
  cascade.def("batchSignalSimulation",
              [](nlm::Cascade& c, const carr& in, bool t, bool rp, unsigned n, unsigned im, const py::object& uo) {
                Grid g = to_grid(in);
                auto u = to_flags(uo);
                Grid r;
                {
                  py::gil_scoped_release nogil;
                  r = c.batchSignalSimulation(g, t, rp, n, im, u);
                }
                return owned(std::move(r));
              },
              "inputProfs"_a, "inTimeDomain"_a = false, "runPump"_a = true, "nThreads"_a = 1, "inputMode"_a = 0,
              "useOutput"_a = py::list());
  cascade.def("addMedium", [](nlm::Cascade& c, Medium& md, const std::map<unsigned, unsigned>& con) { c.addMedium(md, con); },
              "medium"_a, "connection"_a, py::keep_alive<1, 2>());
  cascade.def_property_readonly("omega", [](py::object self) { return vec_view(self.cast<nlm::Cascade&>().cascadeOmega(), self); });
  cascade.def_property_readonly("tau", [](py::object self) { return vec_view(self.cast<nlm::Cascade&>().cascadeTau(), self); });
  cascade.def("__getitem__", [](nlm::Cascade& c, unsigned i) -> Medium& { return c.getMedium(i); },
              py::return_value_policy::reference_internal);
  cascade.def_property_readonly("media", [](nlm::Cascade& c) {
    py::list l;
    for (Medium* md : c.getMedia()) l.append(py::cast(md, py::return_value_policy::reference));
    return l;
  });
  cascade.def_property_readonly("nMedia", &nlm::Cascade::getNMedia);

  // ---- solvers: class names, docstrings, argument names/defaults of each
  // solver's NLMMODULE block (solvers/*.cpp); extensions are keyword-only.
  using PT = nlm::PulseType;
  using IP = nlm::IntensityProfile;
  const carr noPump(std::vector<py::ssize_t>{0});
  const darr noPol(std::vector<py::ssize_t>{0});

  auto mk = [](const std::string& name, std::vector<double> num, const carr& custom, PT pt, IP ip, const darr& pol,
               unsigned tPrec, unsigned zPrec, unsigned nZ, const darr& polProf) {
    return nlm::makeSolver(name, num, to_vec(custom), (int)pt, (int)ip, to_dvec(pol), tPrec, zPrec, nZ,
                           to_dvec(polProf))
        .release();
  };

  // solvers/Chi3.cpp:75-80
  py::class_<nlm::Chi3, Medium>(m, "Chi3", "Single mode self phase modulation")
      .def(py::init([mk](double relativeLength, double nlLength, double beta2, const carr& customPump, PT pulseType,
                         double beta3, double rayleighLength, double tMax, unsigned tPrecision, unsigned zPrecision,
                         IP intensityProfile, double chirp, unsigned nZSteps) {
             return static_cast<nlm::Chi3*>(mk("Chi3", {relativeLength, nlLength, beta2, beta3, rayleighLength, tMax, chirp},
                                               customPump, pulseType, intensityProfile, darr(std::vector<py::ssize_t>{0}),
                                               tPrecision, zPrecision, nZSteps, darr(std::vector<py::ssize_t>{0})));
           }),
           "relativeLength"_a, "nlLength"_a, "beta2"_a, "customPump"_a = noPump, "pulseType"_a = PT::Gaussian,
           "beta3"_a = 0, "rayleighLength"_a = inf, "tMax"_a = 10, "tPrecision"_a = 512, "zPrecision"_a = 100,
           "intensityProfile"_a = IP::GaussianBeam, "chirp"_a = 0, py::kw_only(), "nZSteps"_a = 0)
      .def("runPumpSimulation",
           [](nlm::Chi3& c) {
             py::gil_scoped_release nogil;
             c.runPumpSimulation();
           },
           "Simulate propagation of the pump field");

  // solvers/Chi2PDC.cpp:51-60
  py::class_<nlm::Chi2PDC, Medium>(m, "Chi2PDC", "Degenerate optical parametric amplification")
      .def(py::init([mk](double relativeLength, double nlLength, double beta2, double beta2s, const carr& customPump,
                         PT pulseType, double beta1, double beta1s, double beta3, double beta3s, double diffBeta0,
                         double rayleighLength, double tMax, unsigned tPrecision, unsigned zPrecision,
                         IP intensityProfile, double chirp, double delay, const darr& poling, unsigned nZSteps,
                         const darr& polingProfile) {
             return static_cast<nlm::Chi2PDC*>(mk("Chi2PDC",
                                                  {relativeLength, nlLength, beta2, beta2s, beta1, beta1s, beta3, beta3s,
                                                   diffBeta0, rayleighLength, tMax, chirp, delay},
                                                  customPump, pulseType, intensityProfile, poling, tPrecision,
                                                  zPrecision, nZSteps, polingProfile));
           }),
           "relativeLength"_a, "nlLength"_a, "beta2"_a, "beta2s"_a, "customPump"_a = noPump,
           "pulseType"_a = PT::Gaussian, "beta1"_a = 0, "beta1s"_a = 0, "beta3"_a = 0, "beta3s"_a = 0,
           "diffBeta0"_a = 0, "rayleighLength"_a = inf, "tMax"_a = 10, "tPrecision"_a = 512, "zPrecision"_a = 100,
           "intensityProfile"_a = IP::GaussianBeam, "chirp"_a = 0, "delay"_a = 0, "poling"_a = noPol, py::kw_only(),
           "nZSteps"_a = 0, "polingProfile"_a = noPol);

  // Chi2SFG (solvers/Chi2SFG.cpp:60-70), Chi2PDCII (solvers/Chi2PDCII.cpp:194-204):
  // same argument list up to the names of the two nonlinear lengths.
  auto sfgLike = [&](auto* tag, const char* name, const char* doc, const char* nl0, const char* nl1) {
    using T = std::remove_pointer_t<decltype(tag)>;
    std::string nm = name;
    py::class_<T, Medium>(m, name, doc)
        .def(py::init([mk, nm](double relativeLength, double nlA, double nlB, double beta2, double beta2s,
                               double beta2o, const carr& customPump, PT pulseType, double beta1, double beta1s,
                               double beta1o, double beta3, double beta3s, double beta3o, double diffBeta0,
                               double rayleighLength, double tMax, unsigned tPrecision, unsigned zPrecision,
                               IP intensityProfile, double chirp, double delay, const darr& poling, unsigned nZSteps,
                               const darr& polingProfile) {
               return static_cast<T*>(mk(nm,
                                         {relativeLength, nlA, nlB, beta2, beta2s, beta2o, beta1, beta1s, beta1o, beta3,
                                          beta3s, beta3o, diffBeta0, rayleighLength, tMax, chirp, delay},
                                         customPump, pulseType, intensityProfile, poling, tPrecision, zPrecision,
                                         nZSteps, polingProfile));
             }),
             "relativeLength"_a, py::arg(nl0), py::arg(nl1), "beta2"_a, "beta2s"_a, "beta2o"_a,
             "customPump"_a = noPump, "pulseType"_a = PT::Gaussian, "beta1"_a = 0, "beta1s"_a = 0, "beta1o"_a = 0,
             "beta3"_a = 0, "beta3s"_a = 0, "beta3o"_a = 0, "diffBeta0"_a = 0, "rayleighLength"_a = inf,
             "tMax"_a = 10, "tPrecision"_a = 512, "zPrecision"_a = 100, "intensityProfile"_a = IP::GaussianBeam,
             "chirp"_a = 0, "delay"_a = 0, "poling"_a = noPol, py::kw_only(), "nZSteps"_a = 0,
             "polingProfile"_a = noPol);
  };
  sfgLike((nlm::Chi2SFG*)nullptr, "Chi2SFG", "Sum (or difference) frequency generation", "nlLength", "nlLengthOrig");

  // solvers/Chi2AFC.cpp:125-135
  py::class_<nlm::Chi2AFC, Medium>(m, "Chi2AFC",
                                   "Adiabatic sum (or difference) frequency generation, in a rotating frame with "
                                   "linearly varying poling frequency built into the solver")
      .def(py::init([mk](double relativeLength, double nlLength, double nlLengthOrig, double beta2, double beta2s,
                         double beta2o, const carr& customPump, PT pulseType, double beta1, double beta1s,
                         double beta1o, double beta3, double beta3s, double beta3o, double diffBeta0Start,
                         double diffBeta0End, double rayleighLength, double tMax, unsigned tPrecision,
                         unsigned zPrecision, IP intensityProfile, double chirp, double delay, unsigned nZSteps) {
             return static_cast<nlm::Chi2AFC*>(mk("Chi2AFC",
                                                  {relativeLength, nlLength, nlLengthOrig, beta2, beta2s, beta2o, beta1,
                                                   beta1s, beta1o, beta3, beta3s, beta3o, diffBeta0Start, diffBeta0End,
                                                   rayleighLength, tMax, chirp, delay},
                                                  customPump, pulseType, intensityProfile,
                                                  darr(std::vector<py::ssize_t>{0}), tPrecision, zPrecision, nZSteps,
                                                  darr(std::vector<py::ssize_t>{0})));
           }),
           "relativeLength"_a, "nlLength"_a, "nlLengthOrig"_a, "beta2"_a, "beta2s"_a, "beta2o"_a,
           "customPump"_a = noPump, "pulseType"_a = PT::Gaussian, "beta1"_a = 0, "beta1s"_a = 0, "beta1o"_a = 0,
           "beta3"_a = 0, "beta3s"_a = 0, "beta3o"_a = 0, "diffBeta0Start"_a = 0, "diffBeta0End"_a = 0,
           "rayleighLength"_a = inf, "tMax"_a = 10, "tPrecision"_a = 512, "zPrecision"_a = 100,
           "intensityProfile"_a = IP::GaussianBeam, "chirp"_a = 0, "delay"_a = 0, py::kw_only(), "nZSteps"_a = 0);

  // solvers/Chi2SFGPDC.cpp:267-277
  py::class_<nlm::Chi2SFGPDC, Medium>(m, "Chi2SFGPDC", "Simultaneous sum frequency generation and parametric amplification")
      .def(py::init([mk](double relativeLength, double nlLength, double nlLengthOrig, double beta2, double beta2s,
                         double beta2o, const carr& customPump, PT pulseType, double beta1, double beta1s,
                         double beta1o, double beta3, double beta3s, double beta3o, double diffBeta0,
                         double diffBeta0o, double rayleighLength, double tMax, unsigned tPrecision,
                         unsigned zPrecision, IP intensityProfile, double chirp, double delay, const darr& poling,
                         unsigned nZSteps, const darr& polingProfile) {
             return static_cast<nlm::Chi2SFGPDC*>(mk("Chi2SFGPDC",
                                                     {relativeLength, nlLength, nlLengthOrig, beta2, beta2s, beta2o,
                                                      beta1, beta1s, beta1o, beta3, beta3s, beta3o, diffBeta0,
                                                      diffBeta0o, rayleighLength, tMax, chirp, delay},
                                                     customPump, pulseType, intensityProfile, poling, tPrecision,
                                                     zPrecision, nZSteps, polingProfile));
           }),
           "relativeLength"_a, "nlLength"_a, "nlLengthOrig"_a, "beta2"_a, "beta2s"_a, "beta2o"_a,
           "customPump"_a = noPump, "pulseType"_a = PT::Gaussian, "beta1"_a = 0, "beta1s"_a = 0, "beta1o"_a = 0,
           "beta3"_a = 0, "beta3s"_a = 0, "beta3o"_a = 0, "diffBeta0"_a = 0, "diffBeta0o"_a = 0,
           "rayleighLength"_a = inf, "tMax"_a = 10, "tPrecision"_a = 512, "zPrecision"_a = 100,
           "intensityProfile"_a = IP::GaussianBeam, "chirp"_a = 0, "delay"_a = 0, "poling"_a = noPol, py::kw_only(),
           "nZSteps"_a = 0, "polingProfile"_a = noPol);

  // solvers/Chi2SFGII.cpp:94-106.  The reference's Python names "intensityProfile"
  // and "zPrecision" sit on the C++ slots (zPrecision: uint, intensityProfile:
  // enum) in that order; the same slot order is kept so calls behave alike.
  py::class_<nlm::Chi2SFGII, Medium>(m, "Chi2SFGII",
                                     "Type II or simultaneous 2-mode sum frequency generation with parametric amplification")
      .def(py::init([mk](double relativeLength, double nlLengthZY, double nlLengthZZ, double beta2, double beta2sz,
                         double beta2sy, double beta2oz, double beta2oy, const carr& customPump, PT pulseType,
                         double beta1, double beta1sz, double beta1sy, double beta1oz, double beta1oy, double beta3,
                         double beta3sz, double beta3sy, double beta3oz, double beta3oy, double diffBeta0z,
                         double diffBeta0y, double diffBeta0, double rayleighLength, double tMax, unsigned tPrecision,
                         unsigned zPrecisionSlot, IP intensityProfileSlot, double chirp, double delay,
                         const darr& poling, unsigned nZSteps, const darr& polingProfile) {
             return static_cast<nlm::Chi2SFGII*>(mk(
                 "Chi2SFGII",
                 {relativeLength, nlLengthZY, nlLengthZZ, beta2, beta2sz, beta2sy, beta2oz, beta2oy, beta1, beta1sz,
                  beta1sy, beta1oz, beta1oy, beta3, beta3sz, beta3sy, beta3oz, beta3oy, diffBeta0z, diffBeta0y, diffBeta0,
                  rayleighLength, tMax, chirp, delay},
                 customPump, pulseType, intensityProfileSlot, poling, tPrecision, zPrecisionSlot, nZSteps, polingProfile));
           }),
           "relativeLength"_a, "nlLengthZY"_a, "nlLengthZZ"_a, "beta2"_a, "beta2sz"_a, "beta2sy"_a, "beta2oz"_a,
           "beta2oy"_a, "customPump"_a = noPump, "pulseType"_a = PT::Gaussian, "beta1"_a = 0, "beta1sz"_a = 0,
           "beta1sy"_a = 0, "beta1oz"_a = 0, "beta1oy"_a = 0, "beta3"_a = 0, "beta3sz"_a = 0, "beta3sy"_a = 0,
           "beta3oz"_a = 0, "beta3oy"_a = 0, "diffBeta0z"_a = 0, "diffBeta0y"_a = 0, "diffBeta0"_a = 0,
           "rayleighLength"_a = inf, "tMax"_a = 10, "tPrecision"_a = 512, "intensityProfile"_a = IP::GaussianBeam,
           "zPrecision"_a = 100, "chirp"_a = 0, "delay"_a = 0, "poling"_a = noPol, py::kw_only(), "nZSteps"_a = 0,
           "polingProfile"_a = noPol);

  sfgLike((nlm::Chi2PDCII*)nullptr, "Chi2PDCII", "Type II or nondegenerate optical parametric amplification",
          "nlLengthI", "nlLengthII");

  // solvers/Chi2SFGOPA.cpp:187-197
  py::class_<nlm::Chi2SFGOPA, Medium>(m, "Chi2SFGOPA",
                                      "Simultaneous sum frequency generation and non-degenerate optical parametric "
                                      "amplification with two pumps")
      .def(py::init([mk](double relativeLength, double nlLengthSFGh, double nlLengthSFGf, double nlLengthDOPAh,
                         double nlLengthDOPAf, double beta2F, double beta2H, double beta2h, double beta2f,
                         const carr& customPump, PT pulseType, double beta1F, double beta1H, double beta1h,
                         double beta1f, double beta3F, double beta3H, double beta3h, double beta3f, double diffBeta0SFG,
                         double diffBeta0DOPA, double rayleighLength, double tMax, unsigned tPrecision,
                         unsigned zPrecision, IP intensityProfile, double chirp, double delay, const darr& poling,
                         unsigned nZSteps, const darr& polingProfile) {
             return static_cast<nlm::Chi2SFGOPA*>(mk(
                 "Chi2SFGOPA",
                 {relativeLength, nlLengthSFGh, nlLengthSFGf, nlLengthDOPAh, nlLengthDOPAf, beta2F, beta2H, beta2h,
                  beta2f, beta1F, beta1H, beta1h, beta1f, beta3F, beta3H, beta3h, beta3f, diffBeta0SFG, diffBeta0DOPA,
                  rayleighLength, tMax, chirp, delay},
                 customPump, pulseType, intensityProfile, poling, tPrecision, zPrecision, nZSteps, polingProfile));
           }),
           "relativeLength"_a, "nlLengthSFGh"_a, "nlLengthSFGf"_a, "nlLengthDOPAh"_a, "nlLengthDOPAf"_a, "beta2F"_a,
           "beta2H"_a, "beta2h"_a, "beta2f"_a, "customPump"_a = noPump, "pulseType"_a = PT::Gaussian,
           "beta1F"_a = 0, "beta1H"_a = 0, "beta1h"_a = 0, "beta1f"_a = 0, "beta3F"_a = 0, "beta3H"_a = 0,
           "beta3h"_a = 0, "beta3f"_a = 0, "diffBeta0SFG"_a = 0, "diffBeta0DOPA"_a = 0, "rayleighLength"_a = inf,
           "tMax"_a = 10, "tPrecision"_a = 512, "zPrecision"_a = 100, "intensityProfile"_a = IP::GaussianBeam,
           "chirp"_a = 0, "delay"_a = 0, "poling"_a = noPol, py::kw_only(), "nZSteps"_a = 0,
           "polingProfile"_a = noPol);

  // solvers/Chi2SFGXPM.cpp:259-269
  py::class_<nlm::Chi2SFGXPM, Medium>(m, "Chi2SFGXPM", "Sum (or difference) frequency generation with cross phase modulation")
      .def(py::init([mk](double relativeLength, double nlLength, double nlLengthOrig, double nlLengthChi3,
                         double nlLengthChi3Orig, double beta2, double beta2s, double beta2o, const carr& customPump,
                         PT pulseType, double beta1, double beta1s, double beta1o, double beta3, double beta3s,
                         double beta3o, double diffBeta0, double rayleighLength, double tMax, unsigned tPrecision,
                         unsigned zPrecision, IP intensityProfile, double chirp, double delay, const darr& poling,
                         unsigned nZSteps, const darr& polingProfile) {
             return static_cast<nlm::Chi2SFGXPM*>(mk("Chi2SFGXPM",
                                                     {relativeLength, nlLength, nlLengthOrig, nlLengthChi3,
                                                      nlLengthChi3Orig, beta2, beta2s, beta2o, beta1, beta1s, beta1o,
                                                      beta3, beta3s, beta3o, diffBeta0, rayleighLength, tMax, chirp,
                                                      delay},
                                                     customPump, pulseType, intensityProfile, poling, tPrecision,
                                                     zPrecision, nZSteps, polingProfile));
           }),
           "relativeLength"_a, "nlLength"_a, "nlLengthOrig"_a, "nlLengthChi3"_a, "nlLengthChi3Orig"_a, "beta2"_a,
           "beta2s"_a, "beta2o"_a, "customPump"_a = noPump, "pulseType"_a = PT::Gaussian, "beta1"_a = 0,
           "beta1s"_a = 0, "beta1o"_a = 0, "beta3"_a = 0, "beta3s"_a = 0, "beta3o"_a = 0, "diffBeta0"_a = 0,
           "rayleighLength"_a = inf, "tMax"_a = 10, "tPrecision"_a = 512, "zPrecision"_a = 100,
           "intensityProfile"_a = IP::GaussianBeam, "chirp"_a = 0, "delay"_a = 0, "poling"_a = noPol, py::kw_only(),
           "nZSteps"_a = 0, "polingProfile"_a = noPol);

  // ---- fully nonlinear solvers
  // solvers/Chi2SHG.cpp:205-212
  py::class_<nlm::Chi2SHG, nlm::FullyNonlinearMedium>(m, "Chi2SHG", "Fully nonlinear second harmonic generation")
      .def(py::init([mk](double relativeLength, double nlLengthH, double nlLengthP, double beta2h, double beta2p,
                         double beta1h, double beta1p, double beta3h, double beta3p, double diffBeta0,
                         double rayleighLength, double tMax, unsigned tPrecision, unsigned zPrecision,
                         IP intensityProfile, const darr& poling, unsigned nZSteps, const darr& polingProfile) {
             return static_cast<nlm::Chi2SHG*>(mk("Chi2SHG",
                                                  {relativeLength, nlLengthH, nlLengthP, beta2h, beta2p, beta1h, beta1p,
                                                   beta3h, beta3p, diffBeta0, rayleighLength, tMax},
                                                  carr(std::vector<py::ssize_t>{0}), PT::Gaussian, intensityProfile,
                                                  poling, tPrecision, zPrecision, nZSteps, polingProfile));
           }),
           "relativeLength"_a, "nlLengthH"_a, "nlLengthP"_a, "beta2h"_a, "beta2p"_a, "beta1h"_a = 0, "beta1p"_a = 0,
           "beta3h"_a = 0, "beta3p"_a = 0, "diffBeta0"_a = 0, "rayleighLength"_a = inf, "tMax"_a = 10,
           "tPrecision"_a = 512, "zPrecision"_a = 100, "intensityProfile"_a = IP::GaussianBeam, "poling"_a = noPol,
           py::kw_only(), "nZSteps"_a = 0, "polingProfile"_a = noPol);

  // solvers/Chi2SHGOPA.cpp:88-98
  py::class_<nlm::Chi2SHGOPA, nlm::FullyNonlinearMedium>(m, "Chi2SHGOPA",
                                                          "Fully nonlinear OPA driven by the second harmonic of the pump")
      .def(py::init([mk](double relativeLength, double nlLengthP, double nlLengthSH, double nlLengthPA1,
                         double nlLengthPA2, double beta2p, double beta2sh, double beta2pa1, double beta2pa2,
                         double beta1p, double beta1sh, double beta1pa1, double beta1pa2, double beta3p,
                         double beta3sh, double beta3pa1, double beta3pa2, double diffBeta0shg, double diffBeta0opa,
                         double rayleighLength, double tMax, unsigned tPrecision, unsigned zPrecision,
                         IP intensityProfile, const darr& poling, unsigned nZSteps, const darr& polingProfile) {
             return static_cast<nlm::Chi2SHGOPA*>(mk(
                 "Chi2SHGOPA",
                 {relativeLength, nlLengthP, nlLengthSH, nlLengthPA1, nlLengthPA2, beta2p, beta2sh, beta2pa1, beta2pa2,
                  beta1p, beta1sh, beta1pa1, beta1pa2, beta3p, beta3sh, beta3pa1, beta3pa2, diffBeta0shg, diffBeta0opa,
                  rayleighLength, tMax},
                 carr(std::vector<py::ssize_t>{0}), PT::Gaussian, intensityProfile, poling, tPrecision, zPrecision,
                 nZSteps, polingProfile));
           }),
           "relativeLength"_a, "nlLengthP"_a, "nlLengthSH"_a, "nlLengthPA1"_a, "nlLengthPA2"_a, "beta2p"_a,
           "beta2sh"_a, "beta2pa1"_a, "beta2pa2"_a, "beta1p"_a = 0, "beta1sh"_a = 0, "beta1pa1"_a = 0,
           "beta1pa2"_a = 0, "beta3p"_a = 0, "beta3sh"_a = 0, "beta3pa1"_a = 0, "beta3pa2"_a = 0,
           "diffBeta0shg"_a = 0, "diffBeta0opa"_a = 0, "rayleighLength"_a = inf, "tMax"_a = 10, "tPrecision"_a = 512,
           "zPrecision"_a = 100, "intensityProfile"_a = IP::GaussianBeam, "poling"_a = noPol, py::kw_only(),
           "nZSteps"_a = 0, "polingProfile"_a = noPol);

  // solvers/Chi2DSFG.cpp:169-177
  py::class_<nlm::Chi2DSFG, nlm::FullyNonlinearMedium>(
      m, "Chi2DSFG", "Sum (or difference) frequency generation with pump depletion (fully nonlinear)")
      .def(py::init([mk](double relativeLength, double nlLengthP, double nlLengthS, double nlLengthD, double beta2p,
                         double beta2s, double beta2d, double beta1p, double beta1s, double beta1d, double beta3p,
                         double beta3s, double beta3d, double diffBeta0, double rayleighLength, double tMax,
                         unsigned tPrecision, unsigned zPrecision, IP intensityProfile, const darr& poling,
                         unsigned nZSteps, const darr& polingProfile) {
             return static_cast<nlm::Chi2DSFG*>(mk("Chi2DSFG",
                                                   {relativeLength, nlLengthP, nlLengthS, nlLengthD, beta2p, beta2s,
                                                    beta2d, beta1p, beta1s, beta1d, beta3p, beta3s, beta3d, diffBeta0,
                                                    rayleighLength, tMax},
                                                   carr(std::vector<py::ssize_t>{0}), PT::Gaussian, intensityProfile,
                                                   poling, tPrecision, zPrecision, nZSteps, polingProfile));
           }),
           "relativeLength"_a, "nlLengthP"_a, "nlLengthS"_a, "nlLengthD"_a, "beta2p"_a, "beta2s"_a, "beta2d"_a,
           "beta1p"_a = 0, "beta1s"_a = 0, "beta1d"_a = 0, "beta3p"_a = 0, "beta3s"_a = 0, "beta3d"_a = 0,
           "diffBeta0"_a = 0, "rayleighLength"_a = inf, "tMax"_a = 10, "tPrecision"_a = 512, "zPrecision"_a = 100,
           "intensityProfile"_a = IP::GaussianBeam, "poling"_a = noPol, py::kw_only(), "nZSteps"_a = 0,
           "polingProfile"_a = noPol);

  // solvers/Chi2SHGXPM.cpp:246-254
  py::class_<nlm::Chi2SHGXPM, nlm::FullyNonlinearMedium>(
      m, "Chi2SHGXPM", "Fully nonlinear second harmonic generation with self and cross phase modulation")
      .def(py::init([mk](double relativeLength, double nlLengthH, double nlLengthP, double nlLengthChi3, double beta2h,
                         double beta2p, double beta1h, double beta1p, double beta3h, double beta3p, double diffBeta0,
                         double rayleighLength, double tMax, unsigned tPrecision, unsigned zPrecision,
                         IP intensityProfile, const darr& poling, unsigned nZSteps, const darr& polingProfile) {
             return static_cast<nlm::Chi2SHGXPM*>(mk("Chi2SHGXPM",
                                                     {relativeLength, nlLengthH, nlLengthP, nlLengthChi3, beta2h, beta2p,
                                                      beta1h, beta1p, beta3h, beta3p, diffBeta0, rayleighLength, tMax},
                                                     carr(std::vector<py::ssize_t>{0}), PT::Gaussian, intensityProfile,
                                                     poling, tPrecision, zPrecision, nZSteps, polingProfile));
           }),
           "relativeLength"_a, "nlLengthH"_a, "nlLengthP"_a, "nlLengthChi3"_a, "beta2h"_a, "beta2p"_a,
           "beta1h"_a = 0, "beta1p"_a = 0, "beta3h"_a = 0, "beta3p"_a = 0, "diffBeta0"_a = 0,
           "rayleighLength"_a = inf, "tMax"_a = 10, "tPrecision"_a = 512, "zPrecision"_a = 100,
           "intensityProfile"_a = IP::GaussianBeam, "poling"_a = noPol, py::kw_only(), "nZSteps"_a = 0,
           "polingProfile"_a = noPol);

  // solvers/Chi2ASHG.cpp:314-321
  py::class_<nlm::Chi2ASHG, nlm::FullyNonlinearMedium>(m, "Chi2ASHG", "Fully nonlinear adiabatic second harmonic generation")
      .def(py::init([mk](double relativeLength, double nlLengthH, double nlLengthP, double beta2h, double beta2p,
                         double beta1h, double beta1p, double beta3h, double beta3p, double diffBeta0Start,
                         double diffBeta0End, double rayleighLength, double tMax, unsigned tPrecision,
                         unsigned zPrecision, IP intensityProfile, unsigned nZSteps) {
             return static_cast<nlm::Chi2ASHG*>(mk("Chi2ASHG",
                                                   {relativeLength, nlLengthH, nlLengthP, beta2h, beta2p, beta1h, beta1p,
                                                    beta3h, beta3p, diffBeta0Start, diffBeta0End, rayleighLength, tMax},
                                                   carr(std::vector<py::ssize_t>{0}), PT::Gaussian, intensityProfile,
                                                   darr(std::vector<py::ssize_t>{0}), tPrecision, zPrecision, nZSteps,
                                                   darr(std::vector<py::ssize_t>{0})));
           }),
           "relativeLength"_a, "nlLengthH"_a, "nlLengthP"_a, "beta2h"_a, "beta2p"_a, "beta1h"_a = 0, "beta1p"_a = 0,
           "beta3h"_a = 0, "beta3p"_a = 0, "diffBeta0Start"_a = 0, "diffBeta0End"_a = 0, "rayleighLength"_a = inf,
           "tMax"_a = 10, "tPrecision"_a = 512, "zPrecision"_a = 100, "intensityProfile"_a = IP::GaussianBeam,
           py::kw_only(), "nZSteps"_a = 0);

  // solvers/Chi3GNLSE.cpp:308-318
  py::class_<nlm::Chi3GNLSE, nlm::FullyNonlinearMedium>(m, "Chi3GNLSE",
                                                         "Fully nonlinear general nonlinear Schrodinger equation")
      .def(py::init([mk](double relativeLength, double nlLength, double selfSteepLength, double fr, double fb,
                         double tau1, double tau2, double tau3, double beta2, double beta3, double rayleighLength,
                         double tMax, unsigned tPrecision, unsigned zPrecision, IP intensityProfile, unsigned nZSteps) {
             return static_cast<nlm::Chi3GNLSE*>(mk("Chi3GNLSE",
                                                    {relativeLength, nlLength, selfSteepLength, fr, fb, tau1, tau2,
                                                     tau3, beta2, beta3, rayleighLength, tMax},
                                                    carr(std::vector<py::ssize_t>{0}), PT::Gaussian, intensityProfile,
                                                    darr(std::vector<py::ssize_t>{0}), tPrecision, zPrecision,
                                                    nZSteps, darr(std::vector<py::ssize_t>{0})));
           }),
           "relativeLength"_a, "nlLength"_a, "selfSteepLength"_a, "fr"_a, "fb"_a, "tau1"_a, "tau2"_a, "tau3"_a,
           "beta2"_a, "beta3"_a = 0, "rayleighLength"_a = inf, "tMax"_a = 10, "tPrecision"_a = 512,
           "zPrecision"_a = 100, "intensityProfile"_a = IP::GaussianBeam, py::kw_only(), "nZSteps"_a = 0)
      .def_property_readonly("ramanResponse",
                             [](py::object self) { return cvec_view(self.cast<Medium&>().raman(), self); },
                             "Read-only array of the Raman response function in the frequency domain.");
}
