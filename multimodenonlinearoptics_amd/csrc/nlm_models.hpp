// nlm_models.hpp -- device nonlinear right-hand sides of the solver plug-ins.
//
// Each model supplies f<ST>(c, cc, x, P, k): the RK stage increment k for one
// time-domain point, given the signal state x (one value per signal mode), the
// pump values P at this z position, the per-position coefficients c (poling,
// nonlinear step, phase mismatch and beam profile already multiplied together
// on the host, in the reference's evaluation order) and the constants cc.
// ST is the RK stage (0..3).  The generic driver (nlm_device.hip) evaluates
//   k1 = f<0>(c_prev, x),  k2 = f<1>(c_intm, x + k1/2),
//   k3 = f<2>(c_intm, x + k2/2),  k4 = f<3>(c_curr, x + k3),
//   x += (k1 + 2 k2 + 2 k3 + k4) * (1/6),
// which is the structure every reference DiffEq shares.  Products are written
// in the same left-to-right order as the reference expressions.
#pragma once
#include "nlm_fft.hpp"
#include "../../include/nlm.h"

#define NLM_DEV __device__ __forceinline__

// Chi3::DiffEq, solvers/Chi3.cpp:57-69.  cc[0] = nlStep[0].
struct MChi3 {
  static constexpr int NS = 1, NP = 1, NC = 0, NCC = 1;
  template <int ST>
  NLM_DEV static void f(const cplx*, const cplx* cc, const cplx (&x)[NS], const cplx (&P)[1], cplx (&k)[NS]) {
    const double a2 = abs2(P[0]);
    k[0] = cc[0] * ((2 * a2) * x[0] + (P[0] * P[0]) * conj(x[0]));
  }
};

// Chi2PDC::DiffEq, solvers/Chi2PDC.cpp:27-47.  c[0] = polDir * nlStep * mismatch.
struct MChi2PDC {
  static constexpr int NS = 1, NP = 1, NC = 1, NCC = 0;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx*, const cplx (&x)[NS], const cplx (&P)[1], cplx (&k)[NS]) {
    k[0] = (c[0] * P[0]) * conj(x[0]);
  }
};

// Chi2SFG::DiffEq (solvers/Chi2SFG.cpp:28-56) and Chi2AFC::DiffEq
// (solvers/Chi2AFC.cpp:26-51; no poling, chirped mismatch folded into c).
// x = (S, O); c[0] = pol*nl0*mis, c[1] = pol*nl1/mis.
struct MChi2SFG {
  static constexpr int NS = 2, NP = 1, NC = 2, NCC = 0;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx*, const cplx (&x)[NS], const cplx (&P)[1], cplx (&k)[NS]) {
    k[0] = (c[0] * P[0]) * x[1];
    k[1] = (c[1] * conj(P[0])) * x[0];
  }
};

// Chi2SFGPDC::DiffEq, solvers/Chi2SFGPDC.cpp:28-59.
// c[0] = pol*nl0*mis, c[1] = pol*nl1, c[2] = 1/mis, c[3] = mis_o.
struct MChi2SFGPDC {
  static constexpr int NS = 2, NP = 1, NC = 4, NCC = 0;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx*, const cplx (&x)[NS], const cplx (&P)[1], cplx (&k)[NS]) {
    k[0] = (c[0] * P[0]) * x[1];
    k[1] = c[1] * ((c[2] * conj(P[0])) * x[0] + (c[3] * P[0]) * conj(x[1]));
  }
};

// Chi2SFGII::DiffEq, solvers/Chi2SFGII.cpp:34-90.  x = (Sz, Sy, Oz, Oy).
// c[0] = (pol, 0), c[1] = nl0*mis_cz, c[2] = nl1*mis_sz, c[3] = pol*nl0*mis_cy,
// c[4] = nl0/mis_cy, c[5] = nl1/mis_sz, c[6] = pol*nl0/mis_cz.  The reference's
// k4 uses the previous step's poling; the host table carries that.
struct MChi2SFGII {
  static constexpr int NS = 4, NP = 1, NC = 7, NCC = 0;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx*, const cplx (&x)[NS], const cplx (&P)[1], cplx (&k)[NS]) {
    const double pol = c[0].re;
    k[0] = (pol * P[0]) * (c[1] * x[3] + c[2] * x[2]);
    k[1] = (c[3] * P[0]) * x[2];
    k[2] = (pol * conj(P[0])) * (c[4] * x[1] + c[5] * x[0]);
    k[3] = (c[6] * conj(P[0])) * x[0];
  }
};

// Chi2PDCII::DiffEq, solvers/Chi2PDCII.cpp:29-55.  x = (S, O).
// c[0] = pol*nl0*mis, c[1] = pol*nl1*mis.
struct MChi2PDCII {
  static constexpr int NS = 2, NP = 1, NC = 2, NCC = 0;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx*, const cplx (&x)[NS], const cplx (&P)[1], cplx (&k)[NS]) {
    k[0] = (c[0] * P[0]) * conj(x[1]);
    k[1] = (c[1] * P[0]) * conj(x[0]);
  }
};

// Chi2SFGOPA::DiffEq, solvers/Chi2SFGOPA.cpp:31-77.  x = (H, F), two pumps.
// c[0] = pol*nl0*mis_sfg, c[1] = pol*nl2*mis_dopa, c[2] = pol*nl1/mis_sfg,
// c[3] = pol*nl3*mis_dopa.
struct MChi2SFGOPA {
  static constexpr int NS = 2, NP = 2, NC = 4, NCC = 0;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx*, const cplx (&x)[NS], const cplx (&P)[2], cplx (&k)[NS]) {
    k[0] = (c[0] * P[0]) * x[1] + (c[1] * P[1]) * conj(x[1]);
    k[1] = (c[2] * conj(P[0])) * x[0] + (c[3] * P[1]) * conj(x[0]);
  }
};

// Chi2SFGXPM::DiffEq, solvers/Chi2SFGXPM.cpp:26-55.  x = (S, O).
// c[0] = pol*nl0*mis, c[1] = pol*nl1/mis; cc[0] = nl2, cc[1] = nl3.
struct MChi2SFGXPM {
  static constexpr int NS = 2, NP = 1, NC = 2, NCC = 2;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx* cc, const cplx (&x)[NS], const cplx (&P)[1], cplx (&k)[NS]) {
    const double a2 = abs2(P[0]);
    k[0] = ((c[0] * P[0]) + cc[0] * a2) * x[1];
    k[1] = ((c[1] * conj(P[0])) + cc[1] * a2) * x[0];
  }
};

// Chi2SHG::DiffEq (solvers/Chi2SHG.cpp:25-60) and Chi2ASHG::DiffEq
// (solvers/Chi2ASHG.cpp:24-56).  x = (fundamental, second harmonic).
// c[0] = r*nl0*mis, c[1] = r*nl1/mis*0.5 (r = poling * beam amplitude).
struct MChi2SHG {
  static constexpr int NS = 2, NP = 0, NC = 2, NCC = 0;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx*, const cplx (&x)[NS], const cplx (&)[1], cplx (&k)[NS]) {
    k[0] = (c[0] * conj(x[0])) * x[1];
    k[1] = c[1] * (x[0] * x[0]);
  }
};

// Chi2SHGOPA::DiffEq, solvers/Chi2SHGOPA.cpp:32-84.  x = (P, SH, A1, A2).
// c[0] = r*nl0*mis_shg, c[1] = r*nl1, c[2] = mis_opa, c[3] = 0.5/mis_shg,
// c[4] = r*nl2/mis_opa, c[5] = r*nl3/mis_opa.  The first stage multiplies
// A2*A1, the later ones A1*A2, as the reference does.
struct MChi2SHGOPA {
  static constexpr int NS = 4, NP = 0, NC = 6, NCC = 0;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx*, const cplx (&x)[NS], const cplx (&)[1], cplx (&k)[NS]) {
    k[0] = (c[0] * conj(x[0])) * x[1];
    const cplx pa = ST == 0 ? (c[2] * x[3]) * x[2] : (c[2] * x[2]) * x[3];
    k[1] = c[1] * (pa + c[3] * (x[0] * x[0]));
    k[2] = (c[4] * x[1]) * conj(x[3]);
    k[3] = (c[5] * x[1]) * conj(x[2]);
  }
};

// Chi2DSFG::DiffEq, solvers/Chi2DSFG.cpp:25-65.  x = (P, S, D).
// c[0] = r*nl0*mis, c[1] = r*nl1/mis, c[2] = r*nl2*mis.
struct MChi2DSFG {
  static constexpr int NS = 3, NP = 0, NC = 3, NCC = 0;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx*, const cplx (&x)[NS], const cplx (&)[1], cplx (&k)[NS]) {
    k[0] = (c[0] * x[1]) * conj(x[2]);
    k[1] = (c[1] * x[0]) * x[2];
    k[2] = (c[2] * x[1]) * conj(x[0]);
  }
};

// Chi2SHGXPM::DiffEq, solvers/Chi2SHGXPM.cpp:25-65.  x = (fundamental, SH).
// c[0] = r*nl0*mis, c[1] = r*nl1/mis*0.5, c[2] = amp^2*nl2, c[3] = 2*c[2].
struct MChi2SHGXPM {
  static constexpr int NS = 2, NP = 0, NC = 4, NCC = 0;
  template <int ST>
  NLM_DEV static void f(const cplx* c, const cplx*, const cplx (&x)[NS], const cplx (&)[1], cplx (&k)[NS]) {
    const double a = abs2(x[0]), b = abs2(x[1]);
    k[0] = (c[0] * conj(x[0])) * x[1] + (c[2] * (a + 2 * b)) * x[0];
    k[1] = c[1] * (x[0] * x[0]) + (c[3] * (b + 2 * a)) * x[1];
  }
};

// Chi3GNLSE (solvers/Chi3GNLSE.cpp:51-92) needs transforms inside each RK
// stage; its stage evaluation lives in nlm_device.hip.  c[0] = (relI, 0).
struct MChi3GNLSE {
  static constexpr int NS = 1, NP = 0, NC = 1, NCC = 0;
};

template <int MODEL>
struct ModelOf;
template <> struct ModelOf<NLM_CHI3> { using type = MChi3; };
template <> struct ModelOf<NLM_CHI2PDC> { using type = MChi2PDC; };
template <> struct ModelOf<NLM_CHI2SFG> { using type = MChi2SFG; };
template <> struct ModelOf<NLM_CHI2AFC> { using type = MChi2SFG; };
template <> struct ModelOf<NLM_CHI2SFGPDC> { using type = MChi2SFGPDC; };
template <> struct ModelOf<NLM_CHI2SFGII> { using type = MChi2SFGII; };
template <> struct ModelOf<NLM_CHI2PDCII> { using type = MChi2PDCII; };
template <> struct ModelOf<NLM_CHI2SFGOPA> { using type = MChi2SFGOPA; };
template <> struct ModelOf<NLM_CHI2SFGXPM> { using type = MChi2SFGXPM; };
template <> struct ModelOf<NLM_CHI2SHG> { using type = MChi2SHG; };
template <> struct ModelOf<NLM_CHI2SHGOPA> { using type = MChi2SHGOPA; };
template <> struct ModelOf<NLM_CHI2DSFG> { using type = MChi2DSFG; };
template <> struct ModelOf<NLM_CHI2SHGXPM> { using type = MChi2SHGXPM; };
template <> struct ModelOf<NLM_CHI2ASHG> { using type = MChi2SHG; };
template <> struct ModelOf<NLM_CHI3GNLSE> { using type = MChi3GNLSE; };
