// Device physics for the CATO hot path on gfx950: field-operator arithmetic,
// primitive recovery, MUSCL/TVD2 and MLP3/MLP5 edge reconstruction, and the
// AUSM+/AUSMPW+/SLAU-family/Roe/HLLC edge fluxes.
//
// Every routine keeps the reference's operation order (citations inline) so
// that, compiled with -ffp-contract=off and IEEE division/sqrt, the FP64
// instantiation reproduces the reference arithmetic bit for bit.  The same
// templates instantiated with float give the reduced-precision flux form.
//
// Fortran semantics restated here:
//   MAX(a1,a2,..) keeps the running value unless a later argument compares
//   greater (MIN: less); SIGN(1,x) = copysign(1,x); x**2 = x*x; products and
//   sums associate left to right.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <type_traits>

#include "cato_types.hpp"

namespace cato {

template <typename R> struct Lim;
template <> struct Lim<double> {
  static constexpr double eps = 2.220446049250313080847e-16;  // epsilon(1.0_rk)
  static constexpr double tiny = 2.2250738585072013830902e-308; // tiny(1.0_rk)
};
template <> struct Lim<float> {
  // The reduced-precision form keeps the reference's absolute thresholds
  // (rounded to float); tiny() becomes the float minimum normal.
  static constexpr float eps = 2.220446049250313080847e-16f;
  static constexpr float tiny = 1.17549435082228750797e-38f;
};

// Arithmetic type of the FP32-flux precision form: IEEE binary32 except that
// a / b is a times the hardware reciprocal of b (v_rcp_f32, 1 ulp; the
// quotient within 1.5 ulp, and the IEEE result for zero, infinite and NaN
// operands) and the square root is v_sqrt_f32, instead of the correctly
// rounded sequences, whose scale and fix-up steps dominate the cost of the
// reconstruction and flux code; and a product feeding a sum may be fused into
// one FMA (the form's results are judged by their error against the FP64 path,
// not bit for bit).  Only explicit conversions exist, so an unintended double
// or int operand fails to compile.
struct f32x {
  float v;
  f32x() = default;
  __host__ __device__ constexpr explicit f32x(float x) : v(x) {}
  __host__ __device__ constexpr explicit f32x(double x) : v((float)x) {}
  __host__ __device__ constexpr explicit f32x(int x) : v((float)x) {}
  __host__ __device__ constexpr explicit operator double() const { return (double)v; }
  __host__ __device__ constexpr explicit operator float() const { return v; }
};
__device__ inline f32x operator+(f32x a, f32x b) {
#pragma clang fp contract(fast)
  return f32x(a.v + b.v);
}
__device__ inline f32x operator-(f32x a, f32x b) {
#pragma clang fp contract(fast)
  return f32x(a.v - b.v);
}
__device__ inline f32x operator*(f32x a, f32x b) {
#pragma clang fp contract(fast)
  return f32x(a.v * b.v);
}
__device__ inline f32x operator-(f32x a) { return f32x(-a.v); }
__device__ inline f32x operator/(f32x a, f32x b) { return f32x(a.v * __builtin_amdgcn_rcpf(b.v)); }
__device__ inline bool operator<(f32x a, f32x b) { return a.v < b.v; }
__device__ inline bool operator>(f32x a, f32x b) { return a.v > b.v; }
__device__ inline bool operator<=(f32x a, f32x b) { return a.v <= b.v; }
__device__ inline bool operator>=(f32x a, f32x b) { return a.v >= b.v; }
__device__ inline bool operator==(f32x a, f32x b) { return a.v == b.v; }
__device__ inline bool operator!=(f32x a, f32x b) { return a.v != b.v; }
template <> struct Lim<f32x> {
  static constexpr f32x eps = f32x(Lim<float>::eps);
  static constexpr f32x tiny = f32x(Lim<float>::tiny);
};
__device__ inline f32x absd(f32x x) { return f32x(__builtin_fabsf(x.v)); }
__device__ inline f32x sqrtr(f32x x) { return f32x(__builtin_amdgcn_sqrtf(x.v)); }
__device__ inline f32x copysignr(f32x a, f32x b) { return f32x(__builtin_copysignf(a.v, b.v)); }

template <typename R> __host__ __device__ inline R fmax2(R a, R b) { return (b > a) ? b : a; }
template <typename R> __host__ __device__ inline R fmin2(R a, R b) { return (b < a) ? b : a; }
// The precision form takes the hardware v_max_f32 / v_min_f32 (one instruction
// instead of a compare and a select): it differs from the compare-and-select
// only for NaN operands and the sign of a zero result, and that form is judged
// by its error against the FP64 path (a NaN state is still caught where the
// next stage recovers the primitives)
__device__ inline f32x fmax2(f32x a, f32x b) { return f32x(__builtin_fmaxf(a.v, b.v)); }
__device__ inline f32x fmin2(f32x a, f32x b) { return f32x(__builtin_fminf(a.v, b.v)); }
template <typename R> __host__ __device__ inline R sq(R x) { return x * x; }

__host__ __device__ inline double absd(double x) { return __builtin_fabs(x); }
__host__ __device__ inline float absd(float x) { return __builtin_fabsf(x); }
__host__ __device__ inline double sqrtr(double x) { return __builtin_sqrt(x); }
__host__ __device__ inline float sqrtr(float x) { return __builtin_sqrtf(x); }
__host__ __device__ inline double copysignr(double a, double b) { return __builtin_copysign(a, b); }
__host__ __device__ inline float copysignr(float a, float b) { return __builtin_copysignf(a, b); }

// field_2d_cpu_operators.f90:84-176: + and - flush |r| < epsilon to zero; field/field
// division returns 0 when |numerator| < tiny.
__device__ inline double op_add(double a, double b) { double r = a + b; return (absd(r) < Lim<double>::eps) ? 0.0 : r; }
__device__ inline double op_sub(double a, double b) { double r = a - b; return (absd(r) < Lim<double>::eps) ? 0.0 : r; }
__device__ inline double op_div(double a, double b) { return (absd(a) < Lim<double>::tiny) ? 0.0 : a / b; }

// Primitive state of one cell.
template <typename R> struct Prim { R rho, u, v, p; };

// fluid_t%calculate_derived_quantities for one cell (mod_fluid.f90:564-591 with
// equation_of_state.f90:286-292 and 250-251): returns rho,u,v,p (u,v Mach-filtered)
// and the sound speed.
__device__ inline Prim<double> cons_to_prim(double r, double ru, double rv, double rE, double gamma, double gm1,
                                            double *cs_out) {
  double u = op_div(ru, r);
  double v = op_div(rv, r);
  double t1 = r * gm1;
  double t2 = op_div(rE, r);
  double t5 = op_add(u * u, v * v);
  double t6 = t5 * 0.5;
  double t7 = op_sub(t2, t6);
  double p = t1 * t7;
  double cs = sqrtr(op_div(p, r) * gamma);
  // Mach filter (mod_fluid.f90:580-588): |u/cs| < 1e-13 => u = 0.  When |u| >= 2e-13*cs
  // the quotient is certainly >= 1e-13, so the division is skipped (identical result).
  const double MF = 1e-13;
  const double lim = 2.0 * MF * cs;
  if (!(absd(v) >= lim) && absd(op_div(v, cs)) < MF) v = 0.0;
  if (!(absd(u) >= lim) && absd(op_div(u, cs)) < MF) u = 0.0;
  if (cs_out) *cs_out = cs;
  return Prim<double>{r, u, v, p};
}

// 1/x in FP64 to about 1 ulp: hardware estimate plus two Newton steps (the
// precision form's epilogue; the FP64 form divides).
__device__ inline double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// Primitive recovery of the precision form: the same formulas in FP64 through
// one reciprocal of rho, without the field-operator flushes and the Mach
// filter (both act far below FP32 resolution); cs only when asked.
__device__ inline Prim<double> cons_to_prim_fast(double r, double ru, double rv, double rE, double gamma, double gm1,
                                                 int want_cs, double *cs_out) {
  const double ir = rcp64(r);
  const double u = ru * ir, v = rv * ir;
  const double p = (r * gm1) * (rE * ir - 0.5 * (u * u + v * v));
  *cs_out = want_cs ? __builtin_sqrt(gamma * p * ir) : 0.0;
  return Prim<double>{r, u, v, p};
}

// ---------------------------------------------------------------------------
// FP64 division and square root of the edge computations.
//
// IeeeOps: the correctly rounded operators.  hipcc lowers a / b to a range
// scaling of both operands (v_div_scale), the reciprocal with two Newton steps,
// the quotient with one correction (v_div_fmas) and a special-case fix-up
// (v_div_fixup); sqrt(x) to an input scaling below 2^-767, v_rsq, Goldschmidt
// steps, the inverse scaling and a select for 0 / inf.  While the scalings are
// identities and no special case arises, dropping those steps leaves the same
// operations on the same values, hence the same bits.  FastOps is that
// remainder, with the reciprocal of a shared denominator formed once.  The
// stage kernel uses it only on tiles whose operands keep every division in
// range (numerator 0 or |a| >= 2^-969, 2^-1021 <= |b| < 2^1021, normal
// quotient, exponents less than 768 apart; square roots of 2^-767 <= x < inf)
// and only in formulas where a numerator -0 over a positive denominator, which
// FastOps returns as +0, cannot change a result.
// ---------------------------------------------------------------------------
struct IeeeOps {
  template <typename T> struct Rcp { T b; };
  template <typename T> static __device__ Rcp<T> rcp(T b) { return Rcp<T>{b}; }
  template <typename T> static __device__ T div(T a, Rcp<T> r) { return a / r.b; }
  template <typename T> static __device__ T div(T a, T b) { return a / b; }
  template <typename T> static __device__ T sqrt_pos(T x) { return sqrtr(x); }
  template <typename T> static __device__ T sqrt0(T x) { return sqrtr(x); }
  template <typename T> static __device__ T min(T a, T b) { return fmin2(a, b); }
  template <typename T> static __device__ T max(T a, T b) { return fmax2(a, b); }
};
struct FastOps {
  template <typename T> struct Rcp { double b, y; };
  static __device__ Rcp<double> rcp(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    return Rcp<double>{b, y};
  }
  static __device__ double div(double a, Rcp<double> r) {
    const double q = a * r.y;
    const double e = __builtin_fma(-r.b, q, a);
    return __builtin_fma(e, r.y, q);
  }
  static __device__ double div(double a, double b) { return div(a, rcp(b)); }
  // v_min_f64 / v_max_f64 for the flux's min and max: on the finite operands of
  // an in-range tile the same value as the reference's compare-and-select, at
  // the sites that use them (no -0 operand meets a +0 one there, or the sign of
  // that zero is dropped: SLAU's g enters only through 1 - g)
  // sqrt of +0 or of an in-range x (sqrt_pos): SLAU's |velocity| from squares of
  // reconstructed velocities, which are 0 or at least 2^-218 in magnitude
  static __device__ double sqrt0(double x) { return x == 0.0 ? 0.0 : sqrt_pos(x); }
  static __device__ double min(double a, double b) { return __builtin_fmin(a, b); }
  static __device__ double max(double a, double b) { return __builtin_fmax(a, b); }
  static __device__ double sqrt_pos(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = r * 0.5;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    double d = __builtin_fma(-g, g, x);
    h = __builtin_fma(h, e, h);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
  }
};

// Operands for which the FastOps edge computations stay in range, by cell: rho
// and p in [2^-B, 2^B], u and v zero or 2^-150 <= |u| <= 2^B (gamma in (1, 4]).
//
// B = 60, TVD2 reconstruction with the SLAU family: every reconstructed value is
// 0 or within [2^-218, 2^62] (rho, p within [2^-61, 2^61]), differences are 0 or
// at least epsilon after their flush, and every quotient of those formulas lies
// within [2^-405, 2^248] with operand exponents at most 300 apart.
//
// B = 20, every other scheme: the reconstruction (TVD2 within the bound above)
// and, where it has a FastOps path, the flux (HLLC).  MLP: nonzero flushed
// differences lie in [2^-52, 2^21], so the slope ratios are 0 or within
// [2^-73, 2^121], r + 1e-30 stays at least 2^-100 from zero, and the quotients
// of tan(theta) and alpha lie within [2^-390, 2^250]; the reconstructed states
// leave [min, max] of their two cells by at most 5e-15 |dp| plus rounding, so
// rho and p stay within [2^-21, 2^21].  HLLC on such states: c in [2^-21, 2^22],
// |q| / c <= 2^42 < 2^52, so S_L - q_L and S_R - q_R keep at least half of c w
// and m_L, m_R are nonzero within [2^-43, 2^88]; |S_L| >= 2^-75 where it
// divides; S*, the star factors and the energy quotients lie within
// [2^-400, 2^170].  No numerator of either is -0.
template <int LOG2B> __device__ inline bool fast_ops_cell(const Prim<double> &q) {
  static_assert(LOG2B >= 1 && LOG2B <= 60, "fast_ops_cell bound");
  constexpr double HI = (double)(1ull << LOG2B), LO = 1.0 / HI, VLO = 0x1p-150;
  const double au = absd(q.u), av = absd(q.v);
  return q.rho >= LO && q.rho <= HI && q.p >= LO && q.p <= HI && au <= HI && av <= HI &&
         (au == 0.0 || au >= VLO) && (av == 0.0 || av >= VLO);
}

// cons_to_prim with the FastOps division and square root wherever a whole wave
// keeps them in range, bit for bit cons_to_prim: the four quotients by rho
// through one reciprocal while rho is in [2^-60, 2^60] and every numerator the
// field division does divide (|x| >= tiny) lies in [2^-600, 2^600] (quotients
// then normal, operand exponents at most 660 apart, and no -0 numerator: the
// tiny flush takes those); the sound speed's root while p is also in
// [2^-600, 2^600] (its argument then lies in [2^-661, 2^662]).  The Mach
// filter keeps its IEEE division (rarely taken).  Waves with any lane outside
// take cons_to_prim's IEEE sequences.
__device__ inline Prim<double> cons_to_prim_vote(double r, double ru, double rv, double rE, double gamma, double gm1,
                                                 double *cs_out) {
  constexpr double RLO = 0x1p-60, RHI = 0x1p60, XLO = 0x1p-600, XHI = 0x1p600;
  const double TINY = Lim<double>::tiny;
  auto num_ok = [&](double x) {
    const double ax = absd(x);
    return ax < TINY || (ax >= XLO && ax <= XHI);
  };
  const bool quot_fast = __all(r >= RLO && r <= RHI && num_ok(ru) && num_ok(rv) && num_ok(rE));
  double u, v, t2;
  FastOps::Rcp<double> dr{r, 0.0};
  if (quot_fast) {
    dr = FastOps::rcp(r);
    u = (absd(ru) < TINY) ? 0.0 : FastOps::div(ru, dr);
    v = (absd(rv) < TINY) ? 0.0 : FastOps::div(rv, dr);
    t2 = (absd(rE) < TINY) ? 0.0 : FastOps::div(rE, dr);
  } else {
    u = op_div(ru, r);
    v = op_div(rv, r);
    t2 = op_div(rE, r);
  }
  double t1 = r * gm1;
  double t5 = op_add(u * u, v * v);
  double t6 = t5 * 0.5;
  double t7 = op_sub(t2, t6);
  double p = t1 * t7;
  double cs;
  if (quot_fast && __all(p >= XLO && p <= XHI))
    cs = FastOps::sqrt_pos(FastOps::div(p, dr) * gamma);
  else
    cs = sqrtr(op_div(p, r) * gamma);
  const double MF = 1e-13;
  const double lim = 2.0 * MF * cs;
  if (!(absd(v) >= lim) && absd(op_div(v, cs)) < MF) v = 0.0;
  if (!(absd(u) >= lim) && absd(op_div(u, cs)) < MF) u = 0.0;
  *cs_out = cs;
  return Prim<double>{r, u, v, p};
}

// ---------------------------------------------------------------------------
// Reconstruction (edge states L/R at the interface between cells 0 and 1).
// Stencil values along the edge normal: qm3=q(-2), qm=q(-1), q0, q1, q2, q3=q(+3).
// ---------------------------------------------------------------------------
template <typename R, class D = IeeeOps> __device__ inline R tvd_phi(int lim, R r) {
  if (lim == LIM_MINMOD) return fmax2(R(0), fmin2(r, R(1)));  // muscl_tvd2.fypp:196
  if (lim == LIM_SUPERBEE) return fmax2(fmax2(R(0), fmin2(R(2) * r, R(1))), fmin2(r, R(2)));  // :198
  return D::div(r + absd(r), R(1) + absd(r));  // :200 (numerator +0 or positive)
}

// Precision form: phi(b / (a + 1e-16)) * a, the limited slope, written without
// the ratio: minmod and superbee as their min/max forms (no division), van Leer
// as (|a| b + a |b|) / (|a + 1e-16| + |b|) (one division instead of two).  With
// the eps flush a is 0 or |a| >= 2.2e-16, so a + 1e-16 keeps the sign of a and
// the forms agree with the ratio form up to 1e-16 / |a| relative.
template <typename R, int LIMC> __device__ inline R tvd_slope(R a, R b) {
  if constexpr (LIMC == LIM_VAN_LEER) {
    return (absd(a) * b + a * absd(b)) / (absd(a + R(1e-16)) + absd(b));
  } else {
    if (!(a * b > R(0))) return R(0);
    R m = (LIMC == LIM_MINMOD) ? fmin2(absd(a), absd(b))
                               : fmax2(fmin2(R(2) * absd(b), absd(a)), fmin2(absd(b), R(2) * absd(a)));
    return copysignr(m, a);
  }
}

// muscl_tvd2.fypp:177-220.  The flushes leave every difference +0 or at least
// epsilon in magnitude, so no numerator below is -0 and |dm + 1e-16| >= 1.2e-16.
template <typename R, int LIMC, class D = IeeeOps>
__device__ inline void recon_tvd2(R qm, R q0, R q1, R q2, R &L, R &Rv) {
  R dm = q0 - qm, dp = q1 - q0, d3 = q2 - q1;
  if (absd(dm) < Lim<R>::eps) dm = R(0);
  if (absd(dp) < Lim<R>::eps) dp = R(0);
  if (absd(d3) < Lim<R>::eps) d3 = R(0);
  if constexpr (!std::is_same<R, double>::value) {
    L = q0 + R(0.5) * tvd_slope<R, LIMC>(dm, dp);
    Rv = q1 - R(0.5) * tvd_slope<R, LIMC>(d3, dp);
  } else {
    R r_L = D::div(dp, dm + R(1e-16));
    R r_R = D::div(dp, d3 + R(1e-16));
    R phi = tvd_phi<R, D>(LIMC, r_L);
    L = q0 + R(0.5) * phi * dm;
    phi = tvd_phi<R, D>(LIMC, r_R);
    Rv = q1 - R(0.5) * phi * d3;
  }
}

// muscl_mlp.fypp:196-279.  t_* are the transverse neighbours for tan(theta):
// a0p = q(0,+1), a0m = q(0,-1), a1p = q(1,+1), a1m = q(1,-1) in (normal, transverse).
template <typename R, int ORDER, class D = IeeeOps>
__device__ inline void recon_mlp(R qm3, R qm, R q0, R q1, R q2, R q3, R a0p, R a0m, R a1p, R a1m, R &L, R &Rv) {
  const R SMALL = R(1e-30);
  R dm = q0 - qm, dp = q1 - q0, d3 = q2 - q1;
  if (absd(dm) < Lim<R>::eps) dm = R(0);
  if (absd(dp) < Lim<R>::eps) dp = R(0);
  if (absd(d3) < Lim<R>::eps) d3 = R(0);
  R r_R_p1 = D::div(dp, d3 + SMALL);
  R r_L = D::div(dp, dm + SMALL);
  R beta_L, beta_R;
  if constexpr (ORDER == 5) {
    R dm3 = qm - qm3, d5 = q3 - q2;
    if (absd(dm3) < Lim<R>::eps) dm3 = R(0);
    if (absd(d5) < Lim<R>::eps) d5 = R(0);
    const auto r_dp = D::rcp(dp + SMALL);
    R r_L_m1 = D::div(dm, dm3 + SMALL);
    R r_L_p1 = D::div(d3, r_dp);
    R r_R_p2 = D::div(d3, d5 + SMALL);
    R r_R = D::div(dm, r_dp);
    const auto r30 = D::rcp(R(30));
    beta_L = D::div(((-D::div(R(2), r_L_m1 + SMALL)) + R(11)) + R(24) * r_L - R(3) * r_L * r_L_p1, r30);
    beta_R = D::div(((-D::div(R(2), r_R_p2 + SMALL)) + R(11)) + R(24) * r_R_p1 - R(3) * r_R_p1 * r_R, r30);
  } else {
    (void)qm3;
    (void)q3;
    const auto r3 = D::rcp(R(3));
    beta_L = D::div(R(1) + R(2) * r_L, r3);
    beta_R = D::div(R(1) + R(2) * r_R_p1, r3);
  }
  R tt = D::div(absd(a0p - a0m), absd(q1 - qm) + SMALL);
  R tt1 = D::div(absd(a1p - a1m), absd(q2 - q0) + SMALL);
  // D::min / D::max: FastOps takes the hardware min/max.  On an in-range tile
  // no operand is NaN, and where a zero's sign could differ from the
  // compare-and-select (max(0, x) for x = +-0, and the last min) it is dropped
  // by 1 + (+-0) = 1 or by the outer compare-and-select max with 0, kept as is
  R aLt = D::div(R(2) * D::max(R(1), r_L) * (R(1) + D::max(R(0), D::div(tt1, r_R_p1 + SMALL))), R(1) + tt);
  R aRt = D::div(R(2) * D::max(R(1), r_R_p1) * (R(1) + D::max(R(0), D::div(tt, r_L + SMALL))), R(1) + tt1);
  R aL = D::max(R(1), D::min(R(2), aLt));
  R aR = D::max(R(1), D::min(R(2), aRt));
  L = q0 + R(0.5) * fmax2(R(0), D::min(D::min(aL * r_L, aL), beta_L)) * dm;
  Rv = q1 - R(0.5) * fmax2(R(0), D::min(D::min(aR * r_R_p1, aR), beta_R)) * d3;
}

// ---------------------------------------------------------------------------
// Edge fluxes of the cell's RIGHT edge (i-edge) and TOP edge (j-edge), whose
// normals quad_cell.f90:193-229 gives as exactly (1,0) and (0,1) on rectilinear
// grids (direction codes below).
// ---------------------------------------------------------------------------
template <typename R> struct EdgeState {
  R rL, uL, vL, pL, rR, uR, vR, pR;
  R nx, ny;  // unit normal of the edge (direction codes 2 to 4)
  int ax;    // index direction of the edge, 0 (i-edge) or 1 (j-edge) (direction code 4)
};

// Edge direction codes of the flux templates: 0 / 1, the i- / j-edge of a
// rectilinear grid (normal (1,0) / (0,1), folded at compile time); 2 / 3, the
// same edges with the normal (s.nx, s.ny) of a general quadrilateral grid; 4,
// either edge, its index direction s.ax and normal (s.nx, s.ny) given at run
// time, so that i- and j-edges in one wave run one code path (the normal
// (1,0) or (0,1) of a rectilinear grid times u or v is u or v exactly, or the
// same signed zero / NaN the folded form gives).  The direction-dependent
// choices of the reference (AUSMPW+'s normal velocity and get_f cells, the
// sensor stencils) follow the index direction.
template <typename R, int DIR> __device__ inline bool edge_ax0(const EdgeState<R> &s) {
  if constexpr (DIR == 4) return s.ax == 0;
  else return (DIR & 1) == 0;
}
template <typename R, int DIR> __device__ inline R edge_nx(const EdgeState<R> &s) {
  if constexpr (DIR >= 2) return s.nx;
  else return DIR == 0 ? R(1) : R(0);
}
template <typename R, int DIR> __device__ inline R edge_ny(const EdgeState<R> &s) {
  if constexpr (DIR >= 2) return s.ny;
  else return DIR == 0 ? R(0) : R(1);
}

// Neighbour cell pressures needed by the shock sensors, in (normal, transverse)
// coordinates relative to the edge's left cell 0: p(a,b).
template <typename R> struct PressNbr {
  R p00, p10, p20, pm10;   // (0,0) (1,0) (2,0) (-1,0)
  R p0p, p0m, p1p, p1m;    // (0,+1) (0,-1) (1,+1) (1,-1)
};

// Constants of the ideal-gas law the fluxes take, formed once per solve on the
// host in FP64 (gas_constants) instead of on every edge: gamma / (gamma - 1),
// 2 ((gamma - 1) / (gamma + 1)) (which equals (2 (gamma - 1)) / (gamma + 1): a
// product by 2 is exact) and (gamma + 1) / (2 gamma), each one IEEE operation
// sequence on gamma, so the FP64 values are the bits the per-edge formulas
// gave.  The precision form rounds them to binary32.
template <typename R> struct Gas {
  R gamma, gg, cr2, hf;
};
__host__ __device__ inline void gas_constants(double gamma, double out[3]) {
  out[0] = gamma / (gamma - 1.0);
  out[1] = 2.0 * ((gamma - 1.0) / (gamma + 1.0));
  out[2] = (gamma + 1.0) / (2.0 * gamma);
}
template <typename R> __device__ inline Gas<R> gas_of(double gamma, const double gc[3]) {
  return Gas<R>{R(gamma), R(gc[0]), R(gc[1]), R(gc[2])};
}

// SLAU family: slau_solver.fypp:355-490.  With D = FastOps: every numerator is
// +0 or positive except V_n_plus / V_n_minus, whose quotients M_plus / M_minus
// enter only Beta (even in a zero M) and g (whose sign of zero vanishes in 1 - g).
template <typename R, int VARIANT, int DIR, class D = IeeeOps>
__device__ inline void flux_slau(const Gas<R> &gas, const EdgeState<R> &s, const PressNbr<R> &pn, R out[4]) {
  const R n_x = edge_nx<R, DIR>(s), n_y = edge_ny<R, DIR>(s);
  const R THRESHOLD = R(5e-3), C_SD1 = R(0.01);
  const R eps = Lim<R>::eps;
  R rho_L = s.rL, u_L = s.uL, v_L = s.vL, p_L = s.pL;
  R rho_R = s.rR, u_R = s.uR, v_R = s.vR, p_R = s.pR;
  // precision form: one reciprocal per shared denominator (rho_L, rho_R, c_bar)
  // instead of a division per quotient; the FP64 form divides as the reference
  constexpr bool kRcp = !std::is_same<R, double>::value;
  R ir_L = R(0), ir_R = R(0), ic = R(0);
  if constexpr (kRcp) {
    ir_L = R(1) / rho_L;
    ir_R = R(1) / rho_R;
  }
  // FP64: IEEE quotients (D = IeeeOps) or the same quotients through one
  // reciprocal per shared denominator rho_L, rho_R, c_bar (D = FastOps)
  const auto dr_L = D::rcp(rho_L), dr_R = D::rcp(rho_R);
  auto over_rho = [&](R x, R ir, decltype(dr_L) dr) {
    if constexpr (kRcp) return x * ir;
    else return D::div(x, dr);
  };
  R V_n_plus = u_L * n_x + v_L * n_y;
  R V_n_minus = u_R * n_x + v_R * n_y;
  const R gamma = gas.gamma;
  R c_plus = D::sqrt_pos(over_rho(gamma * p_L, ir_L, dr_L));
  R c_minus = D::sqrt_pos(over_rho(gamma * p_R, ir_R, dr_R));
  R c_bar = R(0.5) * (c_plus + c_minus);
  if constexpr (kRcp) ic = R(1) / c_bar;
  const auto dc = D::rcp(c_bar);
  auto over_c = [&](R x) {
    if constexpr (kRcp) return x * ic;
    else return D::div(x, dc);
  };
  // x / (2 c_bar): with FastOps the quotient by c_bar halved, which is exact
  // (the quotients of an in-range tile stay normal), so no second reciprocal
  auto over_2c = [&](R x) {
    if constexpr (kRcp) return x * (R(0.5) * ic);
    else if constexpr (std::is_same<D, FastOps>::value) return D::div(x, dc) * R(0.5);
    else return D::div(x, R(2) * c_bar);
  };
  R V_n_bar = D::div(rho_L * absd(V_n_plus) + rho_R * absd(V_n_minus), rho_L + rho_R);
  R vel2 = sq(u_L) + sq(v_L) + sq(u_R) + sq(v_R);
  const R vmag = D::sqrt0(R(0.5) * vel2);
  R M_hat = D::min(R(1), over_c(vmag));
  R Chi = sq(R(1) - M_hat);
  R M_bar = over_c(V_n_bar);
  R p_bar = R(0.5) * (p_L + p_R);
  R rho_bar = R(0.5) * (rho_L + rho_R);
  R M_plus = over_c(V_n_plus);
  R M_minus = over_c(V_n_minus);
  R Beta_plus = (absd(M_plus) < R(1)) ? R(0.25) * (R(2) - M_plus) * sq(M_plus + R(1))
                                       : R(0.5) * (R(1) + copysignr(R(1), M_plus));
  R Beta_minus = (absd(M_minus) < R(1)) ? R(0.25) * (R(2) + M_minus) * sq(M_minus - R(1))
                                         : R(0.5) * (R(1) + copysignr(R(1), -M_minus));
  R delta_p = p_R - p_L;
  if (absd(delta_p) < eps || absd(delta_p) < THRESHOLD * p_bar) delta_p = R(0);
  R delta_rho = rho_R - rho_L;
  if (absd(delta_rho) < eps) delta_rho = R(0);
  R delta_Beta = Beta_plus - Beta_minus;
  if (absd(delta_Beta) < eps) delta_Beta = R(0);
  R p_L_minus_R = p_L - p_R;
  if (absd(p_L_minus_R) < eps) p_L_minus_R = R(0);
  constexpr bool SD = (VARIANT == FLUX_SD_SLAU || VARIANT == FLUX_SD_SLAU2);
  R theta = R(0);
  if constexpr (SD) {
    // slau_solver.fypp:410-448, cell pressures in the edge's (normal, transverse) frame
    R d[6];
    if (edge_ax0<R, DIR>(s)) {
      d[0] = absd(pn.p0p - pn.p00);   // p(i,j+1)   - p(i,j)
      d[1] = absd(pn.pm10 - pn.p00);  // p(i-1,j)   - p(i,j)
      d[2] = absd(pn.p0m - pn.p00);   // p(i,j-1)   - p(i,j)
      d[3] = absd(pn.p1m - pn.p10);   // p(i+1,j-1) - p(i+1,j)
      d[4] = absd(pn.p20 - pn.p10);   // p(i+2,j)   - p(i+1,j)
      d[5] = absd(pn.p1p - pn.p10);   // p(i+1,j+1) - p(i+1,j)
    } else {
      d[0] = absd(pn.p0m - pn.p00);   // p(i-1,j)   - p(i,j)
      d[1] = absd(pn.pm10 - pn.p00);  // p(i,j-1)   - p(i,j)
      d[2] = absd(pn.p0p - pn.p00);   // p(i+1,j)   - p(i,j)
      d[3] = absd(pn.p1p - pn.p10);   // p(i+1,j+1) - p(i,j+1)
      d[4] = absd(pn.p20 - pn.p10);   // p(i,j+2)   - p(i,j+1)
      d[5] = absd(pn.p1m - pn.p10);   // p(i-1,j+1) - p(i,j+1)
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (d[k] < eps || d[k] < THRESHOLD * p_bar) d[k] = R(0);
    }
    R mx6 = d[0];
#pragma unroll
    for (int k = 1; k < 6; ++k) mx6 = D::max(mx6, d[k]);
    R delta_p_max = D::max(absd(delta_p), mx6);
    const auto dp_bar = D::rcp(p_bar);
    theta = fmin2(R(1), sq(D::div((D::div(C_SD1 * absd(delta_p), dp_bar) + C_SD1),
                                  (D::div(delta_p_max, dp_bar) + C_SD1)))) *
            fmax2(R(0), R(1) - absd(M_bar));
  } else {
    (void)pn;
  }
  R p_tilde;
  if constexpr (VARIANT == FLUX_SLAU || VARIANT == FLUX_SD_SLAU)
    p_tilde = (p_bar + R(0.5) * delta_Beta * p_L_minus_R) + (R(1) - Chi) * (Beta_plus + Beta_minus - R(1)) * p_bar;
  else
    p_tilde = p_bar + (R(0.5) * delta_Beta * p_L_minus_R) +
              vmag * (Beta_plus + Beta_minus - R(1)) * rho_bar * c_bar;
  R g = -(D::max(D::min(M_plus, R(0)), R(-1)) * D::min(D::max(M_minus, R(0)), R(1)));
  R base = R(0.5) * ((rho_L * V_n_plus) + (rho_R * V_n_minus) - absd(V_n_bar) * delta_rho) * (R(1) - g);
  R mass_flux;
  if constexpr (SD) mass_flux = base - (over_2c(theta) * delta_p);
  else mass_flux = base - (over_2c(Chi) * delta_p);
  R mfL = R(0.5) * (mass_flux + absd(mass_flux));
  R mfR = R(0.5) * (mass_flux - absd(mass_flux));
  R gg = gas.gg;
  R H_L = gg * over_rho(p_L, ir_L, dr_L) + R(0.5) * (sq(u_L) + sq(v_L));
  R H_R = gg * over_rho(p_R, ir_R, dr_R) + R(0.5) * (sq(u_R) + sq(v_R));
  out[0] = mfL + mfR;
  out[1] = (mfL * u_L) + (mfR * u_R) + p_tilde * n_x;
  out[2] = (mfL * v_L) + (mfR * v_R) + p_tilde * n_y;
  out[3] = (mfL * H_L) + (mfR * H_R);
}

// AUSMPW+: ausmpw_plus_solver.fypp:341-458 and split functions 486-568 (alpha = 1/8)
template <typename R> __device__ inline R pw_P_plus(R M) {
  const R alpha = R(1) / R(8);
  if (absd(M) > R(1)) return R(0.5) * (R(1) + copysignr(R(1), M));
  return R(0.25) * sq(M + R(1)) * (R(2) - M) + alpha * M * sq(sq(M) - R(1));
}
template <typename R> __device__ inline R pw_P_minus(R M) {
  const R alpha = R(1) / R(8);
  if (absd(M) > R(1)) return R(0.5) * (R(1) - copysignr(R(1), M));
  return R(0.25) * sq(M - R(1)) * (R(2) + M) - alpha * M * sq(sq(M) - R(1));
}
template <typename R> __device__ inline R pw_M_plus(R M) {
  if (absd(M) > R(1)) return R(0.5) * (M + absd(M));
  return R(0.25) * sq(M + R(1));
}
template <typename R> __device__ inline R pw_M_minus(R M) {
  if (absd(M) > R(1)) return R(0.5) * (M - absd(M));
  return -(R(0.25) * sq(M - R(1)));
}

template <typename R, int DIR>
__device__ inline void flux_ausmpw(const Gas<R> &gas, int low_mach_fix, const EdgeState<R> &s, const PressNbr<R> &pn,
                                   R out[4]) {
  const R n_x = edge_nx<R, DIR>(s), n_y = edge_ny<R, DIR>(s);
  const R LOW_MACH = R(0.1);  // globals.f90:32
  R rho_L = s.rL, u_LHS = s.uL, v_LHS = s.vL, p_L = s.pL;
  R rho_R = s.rR, u_RHS = s.uR, v_RHS = s.vR, p_R = s.pR;
  R U_L = edge_ax0<R, DIR>(s) ? u_LHS : v_LHS, U_R = edge_ax0<R, DIR>(s) ? u_RHS : v_RHS;
  R V_L = edge_ax0<R, DIR>(s) ? v_LHS : u_LHS, V_R = edge_ax0<R, DIR>(s) ? v_RHS : u_RHS;
  const R gamma = gas.gamma;
  R cs_L = sqrtr(gamma * p_L / rho_L);
  R cs_R = sqrtr(gamma * p_R / rho_R);
  R M_LHS = sqrtr(sq(u_LHS) + sq(v_LHS)) / cs_L;
  R M_RHS = sqrtr(sq(u_RHS) + sq(v_RHS)) / cs_R;
  if (low_mach_fix && (absd(M_LHS) < LOW_MACH || absd(M_RHS) < LOW_MACH)) {
    R f_M = fmin2(R(1), fmax2(M_LHS, M_RHS));
    R ulf = R(0.5) * (u_LHS + u_RHS) + f_M * R(0.5) * (u_LHS - u_RHS);
    R urf = R(0.5) * (u_LHS + u_RHS) - f_M * R(0.5) * (u_LHS - u_RHS);
    R vlf = R(0.5) * (v_LHS + v_RHS) + f_M * R(0.5) * (v_LHS - v_RHS);
    R vrf = R(0.5) * (v_LHS + v_RHS) - f_M * R(0.5) * (v_LHS - v_RHS);
    u_LHS = ulf; u_RHS = urf; v_LHS = vlf; v_RHS = vrf;
  }
  R gg = gas.gg;
  R H_L = gg * (p_L / rho_L) + R(0.5) * (sq(u_LHS) + sq(v_LHS));
  R H_R = gg * (p_R / rho_R) + R(0.5) * (sq(u_RHS) + sq(v_RHS));
  R H_normal = fmin2(H_L - R(0.5) * sq(V_L), H_R - R(0.5) * sq(V_R));
  R c_s = sqrtr(gas.cr2 * H_normal);
  R c_half = (R(0.5) * (U_L + U_R) < R(0)) ? sq(c_s) / fmax2(absd(U_R), c_s) : sq(c_s) / fmax2(absd(U_L), c_s);
  R M_L = U_L / c_half;
  R M_R = U_R / c_half;
  R M_L_plus = pw_M_plus(M_L);
  R M_R_minus = pw_M_minus(M_R);
  R P_R_minus = pw_P_minus(M_R);
  R P_L_plus = pw_P_plus(M_L);
  R p_s = p_L * P_L_plus + p_R * P_R_minus;
  R mt = fmin2((p_L / p_R), (p_R / p_L));
  R w = R(1) - mt * mt * mt;
  // get_f neighbours (ausmpw_plus_solver.fypp:433-441): i-dir p1=(0,-1),(1,-1) p2=(0,+1),(1,+1);
  // j-dir p1=(i+1,j),(i+1,j+1) p2=(i-1,j),(i-1,j+1), i.e. transverse +1 then -1.
  R p1L = edge_ax0<R, DIR>(s) ? pn.p0m : pn.p0p, p1R = edge_ax0<R, DIR>(s) ? pn.p1m : pn.p1p;
  R p2L = edge_ax0<R, DIR>(s) ? pn.p0p : pn.p0m, p2R = edge_ax0<R, DIR>(s) ? pn.p1p : pn.p1m;
  R f_L, f_R;
  if (absd(p_s) < Lim<R>::tiny) {
    f_L = R(0);
    f_R = R(0);
  } else {
    R mn4 = fmin2(fmin2(fmin2(p1L, p1R), p2L), p2R);
    R m1 = fmin2(R(1), mn4 / fmin2(p_L, p_R));
    R min_term = sq(m1);
    f_L = ((p_L / p_s) - R(1)) * min_term;
    f_R = ((p_R / p_s) - R(1)) * min_term;
  }
  R Mb_L, Mb_R;
  if (M_L_plus + M_R_minus < R(0)) {
    Mb_L = M_L_plus * w * (R(1) + f_L);
    Mb_R = M_R_minus + M_L_plus * ((R(1) - w) * (R(1) + f_L) - f_R);
  } else {
    Mb_L = M_L_plus + M_R_minus * ((R(1) - w) * (R(1) + f_R) - f_L);
    Mb_R = M_R_minus * w * (R(1) + f_R);
  }
  R mfL = Mb_L * c_half * rho_L;
  R mfR = Mb_R * c_half * rho_R;
  out[0] = mfL + mfR;
  out[1] = (mfL * u_LHS) + (mfR * u_RHS) + ((P_L_plus * n_x * p_L) + (P_R_minus * n_x * p_R));
  out[2] = (mfL * v_LHS) + (mfR * v_RHS) + ((P_L_plus * n_y * p_L) + (P_R_minus * n_y * p_R));
  out[3] = (mfL * H_L) + (mfR * H_R);
}

// Roe (reference variant): roe_solver.fypp:420-511, f_h 519-527
template <typename R> __device__ inline R roe_f_h(R pL, R pR) {
  R ratio = absd(pR - pL) / (fmin2(pR, pL) + R(1e-16));
  return (ratio < R(1)) ? R(0) : R(1);
}

template <typename R, int DIR>
__device__ inline void flux_roe(const Gas<R> &gas, const EdgeState<R> &s, const PressNbr<R> &pn, R out[4]) {
  const R nx = edge_nx<R, DIR>(s), ny = edge_ny<R, DIR>(s);
  R rl = absd(s.rL), ul = s.uL, vl = s.vL, pl = absd(s.pL);
  const R gamma = gas.gamma;
  R gg = gas.gg;
  R hl = gg * (pl / rl) + R(0.5) * (sq(ul) + sq(vl));
  R rr = absd(s.rR), ur = s.uR, vr = s.vR, pr = absd(s.pR);
  R hr = gg * (pr / rr) + R(0.5) * (sq(ur) + sq(vr));
  R rav = sqrtr(rl * rr);
  R gam1 = gamma - R(1);
  R dd = rav / rl;
  R dd1 = R(1) / (R(1) + dd);
  R uav = (ul + dd * ur) * dd1;
  R vav = (vl + dd * vr) * dd1;
  R hav = (hl + dd * hr) * dd1;
  R q2a = R(0.5) * (uav * uav + vav * vav);
  R c2a = gam1 * (hav - q2a);
  R cav = sqrtr(c2a);
  R uv = uav * nx + vav * ny;
  R du = (ur - ul) * nx + (vr - vl) * ny;
  R du_trans = -((ur - ul) * ny) + (vr - vl) * nx;
  R cs_R = sqrtr(gamma * pr / rr);
  R cs_L = sqrtr(gamma * pl / rl);
  R eabs1 = absd(fmin2(uv - cav, (ul * nx + vl * ny) - cs_L));
  R eabs2 = absd(uv);
  R eabs4 = absd(fmax2(uv + cav, (ur * nx + vr * ny) + cs_R));
  R h1 = rav * cav * du;
  R h2 = eabs1 * (pr - pl - h1) / (R(2) * c2a);
  R h3 = eabs2 * (rr - rl - (pr - pl) / c2a);
  R h4 = eabs2 * rav;
  R h5 = eabs4 * (pr - pl + h1) / (R(2) * c2a);
  R psi = absd(cs_R - cs_L) / fmax2(cs_R, cs_L);
  R FR0 = rr * uv, FR1 = rr * ur * uv + nx * pr, FR2 = rr * vr * uv + ny * pr, FR3 = rr * hr * uv;
  R FL0 = rl * uv, FL1 = rl * ul * uv + nx * pl, FL2 = rl * vl * uv + ny * pl, FL3 = rl * hl * uv;
  R rm0 = (h2 + h3 + h5);
  R rm1 = (h2 * (uav - cav * nx) + h3 * uav + h4 * (ur - ul - du * nx) + h5 * (uav + cav * nx));
  R rm2 = (h2 * (vav - cav * ny) + h3 * vav + h4 * (vr - vl - du * ny) + h5 * (vav + cav * ny));
  R rm3 = (h2 * (hav - cav * uv) + h3 * q2a + h4 * (uav * (ur - ul) + vav * (vr - vl) - uv * du) +
           h5 * (hav + cav * uv));
  out[0] = R(0.5) * (FR0 + FL0 - rm0);
  out[1] = R(0.5) * (FR1 + FL1 - rm1);
  out[2] = R(0.5) * (FR2 + FL2 - rm2);
  out[3] = R(0.5) * (FR3 + FL3 - rm3);
  // f_p (roe_solver.fypp:496-507); f_h is symmetric in its two pressures.
  R fp;
  if (edge_ax0<R, DIR>(s)) {
    fp = roe_f_h(pn.p00, pn.p10);            // (i,j)   (i+1,j)
    fp = fmax2(fp, roe_f_h(pn.p00, pn.p0p)); // (i,j)   (i,j+1)
    fp = fmax2(fp, roe_f_h(pn.p00, pn.p0m)); // (i,j)   (i,j-1)
    fp = fmax2(fp, roe_f_h(pn.p0p, pn.p1p)); // (i,j+1) (i+1,j+1)
    fp = fmax2(fp, roe_f_h(pn.p0m, pn.p1m)); // (i,j-1) (i+1,j-1)
  } else {
    fp = roe_f_h(pn.p00, pn.p10);            // (i,j)   (i,j+1)
    fp = fmax2(fp, roe_f_h(pn.p00, pn.p0m)); // (i,j)   (i-1,j)
    fp = fmax2(fp, roe_f_h(pn.p00, pn.p0p)); // (i,j)   (i+1,j)
    fp = fmax2(fp, roe_f_h(pn.p10, pn.p1m)); // (i,j+1) (i-1,j+1)
    fp = fmax2(fp, roe_f_h(pn.p10, pn.p1p)); // (i,j+1) (i+1,j+1)
  }
  out[1] = out[1] - R(0.5) * fp * rav * cav * (-(du_trans * ny) + psi * du * nx);
  out[2] = out[2] - R(0.5) * fp * rav * cav * (du_trans * nx + psi * du * ny);
}

// AUSM+ from interface_state (ausm_plus_solver.f90:499-643) with the AUSM+-u flags
// (M_p = p_u = 0), beta = 1/8, alpha = 3/16, upwinded as in the commented solve (312-325).
template <typename R, int DIR>
__device__ inline void flux_ausm_plus(const Gas<R> &gas, const EdgeState<R> &s, R out[4]) {
  const R n_x = edge_nx<R, DIR>(s), n_y = edge_ny<R, DIR>(s);
  const R beta = R(1) / R(8), alpha = R(3) / R(16);
  R rho_L = s.rL, u_L = s.uL, v_L = s.vL, p_L = s.pL;
  R rho_R = s.rR, u_R = s.uR, v_R = s.vR, p_R = s.pR;
  const R gamma = gas.gamma;
  R gamma_factor = gas.cr2;
  R vel_L = u_L * n_x + v_L * n_y;
  R vel_R = u_R * n_x + v_R * n_y;
  R gg = gas.gg;
  R H_L = (p_L / rho_L) * gg + R(0.5) * (u_L * u_L + v_L * v_L);
  R H_R = (p_R / rho_R) * gg + R(0.5) * (u_R * u_R + v_R * v_R);
  R a_crit_L = sqrtr(gamma_factor * absd(H_L));
  R a_crit_R = sqrtr(gamma_factor * absd(H_R));
  R a_circ_L = sq(a_crit_L) / fmax2(a_crit_L, vel_L);
  R a_circ_R = sq(a_crit_R) / fmax2(a_crit_R, -vel_R);
  R a_half = fmin2(a_circ_L, a_circ_R);
  R M_L = vel_L / a_half;
  R M_R = vel_R / a_half;
  R M2p_L = R(0.25) * sq(M_L + R(1)), M2m_L = -(R(0.25) * sq(M_L - R(1)));
  R M2p_R = R(0.25) * sq(M_R + R(1)), M2m_R = -(R(0.25) * sq(M_R - R(1)));
  R M_plus = (absd(M_L) >= R(1)) ? R(0.5) * (M_L + absd(M_L)) : M2p_L * (R(1) - R(16) * beta * M2m_L);
  R M_minus = (absd(M_R) >= R(1)) ? R(0.5) * (M_R - absd(M_R)) : M2m_R * (R(1) + R(16) * beta * M2p_R);
  R M_sum = M_plus + M_minus;
  if (absd(M_sum) < R(1e-13)) M_sum = R(0);
  R M_half = M_sum + R(0);
  if (absd(M_half) < R(1e-13)) M_half = R(0);
  R P_plus = (absd(M_L) >= R(1)) ? (R(0.5) * (M_L + absd(M_L))) / M_L
                                  : M2p_L * ((R(2) - M_L) - R(16) * alpha * M_L * M2m_L);
  R P_minus = (absd(M_R) >= R(1)) ? (R(0.5) * (M_R - absd(M_R))) / M_R
                                   : M2m_R * ((R(-2) - M_R) + R(16) * alpha * M_R * M2p_R);
  R p_half = P_plus * p_L + P_minus * p_R + R(0);
  if (M_half > R(0)) {
    R mass_flux = M_half * a_half * rho_L;
    out[0] = mass_flux;
    out[1] = mass_flux * u_L + (n_x * p_half);
    out[2] = mass_flux * v_L + (n_y * p_half);
    out[3] = mass_flux * H_L;
  } else {
    R mass_flux = M_half * a_half * rho_R;
    out[0] = mass_flux;
    out[1] = mass_flux * u_R + (n_x * p_half);
    out[2] = mass_flux * v_R + (n_y * p_half);
    out[3] = mass_flux * H_R;
  }
}

// HLLC (Toro 2009, sec. 10.4-10.6): PVRS pressure, pressure-based wave speeds,
// contact speed eq. 10.37, star states eq. 10.39.  D = FastOps shares the
// reciprocals of rho_L, rho_R and gamma - 1.
template <typename R, int DIR, class D = IeeeOps>
__device__ inline void flux_hllc(const Gas<R> &gas, const EdgeState<R> &s, R out[4]) {
  const R n_x = edge_nx<R, DIR>(s), n_y = edge_ny<R, DIR>(s);
  R rL = s.rL, uL = s.uL, vL = s.vL, pL = s.pL;
  R rR = s.rR, uR = s.uR, vR = s.vR, pR = s.pR;
  const R gamma = gas.gamma;
  R gm1 = gamma - R(1);
  R qL = uL * n_x + vL * n_y;
  R qR = uR * n_x + vR * n_y;
  const auto d_rL = D::rcp(rL), d_rR = D::rcp(rR);
  R cL = D::sqrt_pos(D::div(gamma * pL, d_rL));
  R cR = D::sqrt_pos(D::div(gamma * pR, d_rR));
  R rho_bar = R(0.5) * (rL + rR);
  R cs_bar = R(0.5) * (cL + cR);
  R p_pvrs = R(0.5) * (pL + pR) - R(0.5) * (qR - qL) * rho_bar * cs_bar;
  R p_star = D::max(R(0), p_pvrs);
  R gfac = gas.hf;
  R wL = (p_star <= pL) ? R(1) : D::sqrt_pos(R(1) + gfac * (D::div(p_star, pL) - R(1)));
  R wR = (p_star <= pR) ? R(1) : D::sqrt_pos(R(1) + gfac * (D::div(p_star, pR) - R(1)));
  R SL = qL - cL * wL;
  R SR = qR + cR * wR;
  R mL = rL * (SL - qL);
  R mR = rR * (SR - qR);
  R S_star = D::div(pR - pL + qL * mL - qR * mR, mL - mR);
  const auto d_gm1 = D::rcp(gm1);
  R EL = D::div(pL, d_gm1) + R(0.5) * rL * (uL * uL + vL * vL);
  R ER = D::div(pR, d_gm1) + R(0.5) * rR * (uR * uR + vR * vR);
  if (R(0) <= SL) {
    out[0] = rL * qL;
    out[1] = rL * uL * qL + pL * n_x;
    out[2] = rL * vL * qL + pL * n_y;
    out[3] = qL * (EL + pL);
  } else if (SR <= R(0)) {
    out[0] = rR * qR;
    out[1] = rR * uR * qR + pR * n_x;
    out[2] = rR * vR * qR + pR * n_y;
    out[3] = qR * (ER + pR);
  } else if (S_star >= R(0)) {
    R fac = D::div(mL, SL - S_star);
    R Us1 = fac * (uL + (S_star - qL) * n_x);
    R Us2 = fac * (vL + (S_star - qL) * n_y);
    R Us3 = fac * (D::div(EL, d_rL) + (S_star - qL) * (S_star + D::div(pL, mL)));
    out[0] = rL * qL + SL * (fac - rL);
    out[1] = (rL * uL * qL + pL * n_x) + SL * (Us1 - rL * uL);
    out[2] = (rL * vL * qL + pL * n_y) + SL * (Us2 - rL * vL);
    out[3] = qL * (EL + pL) + SL * (Us3 - EL);
  } else {
    R fac = D::div(mR, SR - S_star);
    R Us1 = fac * (uR + (S_star - qR) * n_x);
    R Us2 = fac * (vR + (S_star - qR) * n_y);
    R Us3 = fac * (D::div(ER, d_rR) + (S_star - qR) * (S_star + D::div(pR, mR)));
    out[0] = rR * qR + SR * (fac - rR);
    out[1] = (rR * uR * qR + pR * n_x) + SR * (Us1 - rR * uR);
    out[2] = (rR * vR * qR + pR * n_y) + SR * (Us2 - rR * vR);
    out[3] = qR * (ER + pR) + SR * (Us3 - ER);
  }
}

// M-AUSMPW+: m_ausmpw_plus_solver.fypp:437-738 (the path the reference runs: AUSMPW+ pressure
// splitting without the alpha term, w from the normal pressure ratio), get_f 829-862 with its
// 1e-10 flushes, phi_L_half / phi_R_half 790-827 blending in the superbee states (sb), and the
// 1e3*epsilon flushes of the assembled flux 715-735.
template <typename R> __device__ inline R m_phi_L_half(R phi_L, R phi_R, R phi_L_sb, R a) {
  if (absd(phi_R - phi_L) < R(1) || absd(phi_L_sb - phi_L) < Lim<R>::eps) return phi_L;
  return phi_L + fmax2(R(0), (phi_R - phi_L) * (phi_L_sb - phi_L)) / ((phi_R - phi_L) * absd(phi_L_sb - phi_L)) *
                     fmin2(fmin2(a, R(0.5) * absd(phi_R - phi_L)), absd(phi_L_sb - phi_L));
}
template <typename R> __device__ inline R m_phi_R_half(R phi_L, R phi_R, R phi_R_sb, R a) {
  if (absd(phi_L - phi_R) < Lim<R>::eps || absd(phi_R_sb - phi_R) < R(1)) return phi_R;
  return phi_R + fmax2(R(0), (phi_L - phi_R) * (phi_R_sb - phi_R)) / ((phi_L - phi_R) * absd(phi_R_sb - phi_R)) *
                     fmin2(fmin2(a, R(0.5) * absd(phi_L - phi_R)), absd(phi_R_sb - phi_R));
}

template <typename R, int DIR>
__device__ inline void flux_m_ausmpw(const Gas<R> &gas, int low_mach_fix, const EdgeState<R> &s, const EdgeState<R> &sb,
                                     const PressNbr<R> &pn, R out[4]) {
  const R n_x = edge_nx<R, DIR>(s), n_y = edge_ny<R, DIR>(s);
  const R LOW_MACH = R(0.1);
  const R FEPS = R(Lim<double>::eps * 1e3);
  R rho_L = s.rL, u_LHS = s.uL, v_LHS = s.vL, p_L = s.pL;
  R rho_R = s.rR, u_RHS = s.uR, v_RHS = s.vR, p_R = s.pR;
  R U_L = edge_ax0<R, DIR>(s) ? u_LHS : v_LHS, U_R = edge_ax0<R, DIR>(s) ? u_RHS : v_RHS;
  R V_L = edge_ax0<R, DIR>(s) ? v_LHS : u_LHS, V_R = edge_ax0<R, DIR>(s) ? v_RHS : u_RHS;
  const R gamma = gas.gamma;
  R cs_L = sqrtr(gamma * p_L / rho_L);
  R cs_R = sqrtr(gamma * p_R / rho_R);
  R M_LHS = sqrtr(sq(u_LHS) + sq(v_LHS)) / cs_L;
  R M_RHS = sqrtr(sq(u_RHS) + sq(v_RHS)) / cs_R;
  if (low_mach_fix && (absd(M_LHS) < LOW_MACH || absd(M_RHS) < LOW_MACH)) {
    R f_M = fmin2(R(1), fmax2(M_LHS, M_RHS));
    R ulf = R(0.5) * (u_LHS + u_RHS) + f_M * R(0.5) * (u_LHS - u_RHS);
    R urf = R(0.5) * (u_LHS + u_RHS) - f_M * R(0.5) * (u_LHS - u_RHS);
    R vlf = R(0.5) * (v_LHS + v_RHS) + f_M * R(0.5) * (v_LHS - v_RHS);
    R vrf = R(0.5) * (v_LHS + v_RHS) - f_M * R(0.5) * (v_LHS - v_RHS);
    u_LHS = ulf; u_RHS = urf; v_LHS = vlf; v_RHS = vrf;
  }
  R H_L = gas.gg * (p_L / rho_L) + R(0.5) * (sq(u_LHS) + sq(v_LHS));
  R H_R = gas.gg * (p_R / rho_R) + R(0.5) * (sq(u_RHS) + sq(v_RHS));
  R H_normal = fmin2(H_L - R(0.5) * sq(V_L), H_R - R(0.5) * sq(V_R));
  R c_s = sqrtr(gas.cr2 * H_normal);
  R c_half = (R(0.5) * (U_L + U_R) < R(0)) ? sq(c_s) / fmax2(absd(U_R), c_s) : sq(c_s) / fmax2(absd(U_L), c_s);
  R M_L = U_L / c_half;
  R M_R = U_R / c_half;
  R M_L_plus = (absd(M_L) > R(1)) ? R(0.5) * (M_L + absd(M_L)) : R(0.25) * sq(M_L + R(1));
  R M_R_minus = (absd(M_R) > R(1)) ? R(0.5) * (M_R - absd(M_R)) : -(R(0.25) * sq(M_R - R(1)));
  R P_R_minus = (absd(M_R) > R(1)) ? R(0.5) * (R(1) - copysignr(R(1), M_R)) : R(0.25) * sq(M_R - R(1)) * (R(2) + M_R);
  R P_L_plus = (absd(M_L) > R(1)) ? R(0.5) * (R(1) + copysignr(R(1), M_L)) : R(0.25) * sq(M_L + R(1)) * (R(2) - M_L);
  R p_s = p_L * P_L_plus + p_R * P_R_minus;
  R mt = fmin2((p_L / p_R), (p_R / p_L));
  R w = R(1) - (mt * mt * mt);
  // get_f neighbours as AUSMPW+: i-dir p1 = (0,-1),(1,-1), p2 = (0,+1),(1,+1); j-dir the transverse +1 then -1
  R p1L = edge_ax0<R, DIR>(s) ? pn.p0m : pn.p0p, p1R = edge_ax0<R, DIR>(s) ? pn.p1m : pn.p1p;
  R p2L = edge_ax0<R, DIR>(s) ? pn.p0p : pn.p0m, p2R = edge_ax0<R, DIR>(s) ? pn.p1p : pn.p1m;
  R f_L, f_R;
  if (absd(p_s) < Lim<R>::tiny) {
    f_L = R(0);
    f_R = R(0);
  } else {
    const R FE = R(1e-10);
    R mn4 = fmin2(fmin2(fmin2(p1L, p1R), p2L), p2R);
    R min_term = sq(fmin2(R(1), mn4 / fmin2(p_L, p_R)));
    if (absd(R(1) - min_term) < FE) min_term = R(1);
    R fl = (p_L / p_s) - R(1);
    if (absd(fl) < FE) fl = R(0);
    f_L = fl * min_term;
    R fr = (p_R / p_s) - R(1);
    if (absd(fr) < FE) fr = R(0);
    f_R = fr * min_term;
  }
  R Mb_L, Mb_R;
  if (M_L_plus + M_R_minus < R(0)) {
    Mb_L = M_L_plus * w * (R(1) + f_L);
    Mb_R = M_R_minus + M_L_plus * ((R(1) - w) * (R(1) + f_L) - f_R);
  } else {
    Mb_L = M_L_plus + M_R_minus * ((R(1) - w) * (R(1) + f_R) - f_L);
    Mb_R = M_R_minus * w * (R(1) + f_R);
  }
  R a = R(1) - sq(fmin2(R(1), fmax2(absd(M_L), absd(M_R))));
  R rho_L_half = m_phi_L_half(rho_L, rho_R, sb.rL, a);
  R rho_R_half = m_phi_R_half(rho_L, rho_R, sb.rR, a);
  R u_L_half = m_phi_L_half(u_LHS, u_RHS, sb.uL, a);
  R u_R_half = m_phi_R_half(u_LHS, u_RHS, sb.uR, a);
  R v_L_half = m_phi_L_half(v_LHS, v_RHS, sb.vL, a);
  R v_R_half = m_phi_R_half(v_LHS, v_RHS, sb.vR, a);
  R p_L_half = m_phi_L_half(p_L, p_R, sb.pL, a);
  R p_R_half = m_phi_R_half(p_L, p_R, sb.pR, a);
  R H_L_half = gas.gg * (p_L_half / rho_L_half) + R(0.5) * (sq(u_L_half) + sq(v_L_half));
  R H_R_half = gas.gg * (p_R_half / rho_R_half) + R(0.5) * (sq(u_R_half) + sq(v_R_half));
  R mfL = Mb_L * c_half * rho_L_half;
  R mfR = Mb_R * c_half * rho_R_half;
  if (absd(mfL + mfR) < FEPS) {
    mfL = R(0);
    mfR = R(0);
  }
  out[0] = mfL + mfR;
  R pu = (P_L_plus * n_x * p_L) + (P_R_minus * n_x * p_R);
  R mu = (mfL * u_L_half) + (mfR * u_R_half);
  out[1] = (absd(mu) < FEPS) ? pu : mu + pu;
  R pv = (P_L_plus * n_y * p_L) + (P_R_minus * n_y * p_R);
  R mv = (mfL * v_L_half) + (mfR * v_R_half);
  out[2] = (absd(mv) < FEPS) ? pv : mv + pv;
  R e = (mfL * H_L_half) + (mfR * H_R_half);
  out[3] = (absd(e) < FEPS) ? R(0) : e;
}

template <int FLUX> constexpr bool needs_superbee_states() { return FLUX == FLUX_M_AUSMPW_PLUS; }

template <typename R, int FLUX, int DIR, class D = IeeeOps>
__device__ inline void edge_flux(const Gas<R> &gas, int low_mach_fix, const EdgeState<R> &s, const PressNbr<R> &pn,
                                 R out[4], const EdgeState<R> *sb = nullptr) {
  if constexpr (FLUX == FLUX_AUSMPW_PLUS) flux_ausmpw<R, DIR>(gas, low_mach_fix, s, pn, out);
  else if constexpr (FLUX == FLUX_M_AUSMPW_PLUS) flux_m_ausmpw<R, DIR>(gas, low_mach_fix, s, *sb, pn, out);
  else if constexpr (FLUX == FLUX_SLAU || FLUX == FLUX_SLAU2 || FLUX == FLUX_SD_SLAU || FLUX == FLUX_SD_SLAU2)
    flux_slau<R, FLUX, DIR, D>(gas, s, pn, out);
  else if constexpr (FLUX == FLUX_ROE) flux_roe<R, DIR>(gas, s, pn, out);
  else if constexpr (FLUX == FLUX_AUSM_PLUS) flux_ausm_plus<R, DIR>(gas, s, out);
  else flux_hllc<R, DIR, D>(gas, s, out);
}

// FP64 fluxes with a FastOps path (the SLAU family and HLLC; every limiter has
// one), and the bound B of fast_ops_cell that keeps a scheme's operations in
// range: 60 for TVD2 with the SLAU family, 20 otherwise.
template <int FLUX> constexpr bool fast_ops_flux() {
  return FLUX == FLUX_SLAU || FLUX == FLUX_SLAU2 || FLUX == FLUX_SD_SLAU || FLUX == FLUX_SD_SLAU2 ||
         FLUX == FLUX_HLLC;
}
template <int FLUX, int LIMC> constexpr int fast_ops_bound() {
  return (FLUX != FLUX_HLLC && fast_ops_flux<FLUX>() && LIMC != LIM_MLP3 && LIMC != LIM_MLP5) ? 60 : 20;
}

template <int FLUX> constexpr bool needs_pressure_nbrs() {
  return FLUX == FLUX_AUSMPW_PLUS || FLUX == FLUX_SD_SLAU || FLUX == FLUX_SD_SLAU2 || FLUX == FLUX_ROE ||
         FLUX == FLUX_M_AUSMPW_PLUS;
}

// qa = a / m and qb = b / m for m = max(|a|, |b|, 1) != 1.  A finite m is |a|
// or |b| itself, and x / |x| is exactly copysign(1, x), so one IEEE division
// gives both quotients bit for bit; inf / NaN operands keep both divisions.
__device__ inline void scaled_pair(double a, double b, double m, double &qa, double &qb) {
  const bool a_big = absd(a) == m;
  const double q = (a_big ? b : a) / m;
  qa = a_big ? copysignr(1.0, a) : q;
  qb = a_big ? q : copysignr(1.0, b);
  if (!(m <= 1.7976931348623157e308)) {
    qa = a / m;
    qb = b / m;
  }
}

// ---------------------------------------------------------------------------
// Divergence for one conserved variable of one cell: flux_split_edges
// (flux_solver.f90:252-277 and 395-425).  l1..l4 = bottom/right/top/left lengths.
// ---------------------------------------------------------------------------
__device__ inline double divergence(double Fi, double Fim, double Gj, double Gjm, double l1, double l2, double l3,
                                    double l4, double vol, bool vol_pow2, double inv_vol) {
  const double REL_TOL = 1e-14;
  const double NEAR_ZERO = 5.0 * Lim<double>::tiny;
  // m = max(|F|, |Fm|, 1) by the hardware max (one instruction instead of a
  // compare and two selects): for finite fluxes the same value; where a flux is
  // NaN, the reference's max gives m = NaN and v_max_f64 the other operand, and
  // both paths below then return a NaN balance (the NaN flux enters the
  // quotient or the product), so the state and its error flag are the same
  double mi = __builtin_fmax(__builtin_fmax(absd(Fi), absd(Fim)), 1.0);
  double mj = __builtin_fmax(__builtin_fmax(absd(Gj), absd(Gjm)), 1.0);
  double ri, rj;
  if (mi == 1.0) {
    ri = (Fi * l2) + (-(Fim * l4));  // x/1 and x*1 are exact
  } else {
    double qa, qb;
    scaled_pair(Fi, Fim, mi, qa, qb);
    ri = ((qa * l2) + (-(qb * l4))) * mi;
  }
  if (mj == 1.0) {
    rj = (Gj * l3) + (-(Gjm * l1));
  } else {
    double qa, qb;
    scaled_pair(Gj, Gjm, mj, qa, qb);
    rj = ((qa * l3) + (-(qb * l1))) * mj;
  }
  if (absd(ri) < mi * REL_TOL) ri = 0.0;
  if (absd(rj) < mj * REL_TOL) rj = 0.0;
  double fl = ri + rj;
  if (absd(fl) < NEAR_ZERO) fl = 0.0;
  double d = -fl;
  // x / 2^k is exactly x * 2^-k: skip the IEEE division for power-of-two volumes
  return vol_pow2 ? d * inv_vol : d / vol;
}

}  // namespace cato
