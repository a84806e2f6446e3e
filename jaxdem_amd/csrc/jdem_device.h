// jdem_device.h -- device arithmetic of the DEM time step for gfx950.
//
// Every formula here follows the reference file:line cited beside it with an
// explicit association of its floating-point operations.  The library is
// compiled with -ffp-contract=off and IEEE float32 division and square root,
// so each expression rounds exactly as the same expression evaluated op by op
// in numpy float32; ``rsqrt(x)`` is evaluated as ``1/sqrt(x)``.
//
// The arithmetic is written once over the scalar type ``Real``: float (the
// reference's default precision, and every kernel of the resident and batch
// paths) and double (the float64 form of the plain step, which the reference
// runs under jax_enable_x64).  Constants are written as Real(c), so each is
// the literal rounded once to the scalar type, as a Python float is in a
// jnp expression of that dtype.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jdem.h"

#define JD __device__ __forceinline__

namespace jdem {

// Math of one scalar type: the float32 functions for float, the float64
// ones for double (overloads, so that a mixed call does not compile).
JD float sqrt_(float x) { return sqrtf(x); }
JD double sqrt_(double x) { return sqrt(x); }
JD float fmax_(float a, float b) { return fmaxf(a, b); }
JD double fmax_(double a, double b) { return fmax(a, b); }
JD float fmin_(float a, float b) { return fminf(a, b); }
JD double fmin_(double a, double b) { return fmin(a, b); }
JD float fabs_(float x) { return fabsf(x); }
JD double fabs_(double x) { return fabs(x); }
JD float rint_(float x) { return rintf(x); }
JD double rint_(double x) { return rint(x); }
JD float floor_(float x) { return floorf(x); }
JD double floor_(double x) { return floor(x); }

// ---------------------------------------------------------------------------
// small vectors: Vec<3>, Vec<2>; angular quantities Ang<3> = Vec<3>, Ang<2>
// is a scalar (the z component), as State.ang_vel (N,1) in 2D.
// ---------------------------------------------------------------------------
template <int D, class Real = float>
struct Vec;
template <class Real>
struct Vec<3, Real> {
  Real x, y, z;
};
template <class Real>
struct Vec<2, Real> {
  Real x, y;
};
template <class Real = float>
struct Ang1T {
  Real z;
};
using Ang1 = Ang1T<float>;
template <int D, class Real = float>
struct AngSel {
  using T = Vec<3, Real>;
};
template <class Real>
struct AngSel<2, Real> {
  using T = Ang1T<Real>;
};
template <int D, class Real = float>
using Ang = typename AngSel<D, Real>::T;

template <class Real>
JD Vec<3, Real> operator+(Vec<3, Real> a, Vec<3, Real> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class Real>
JD Vec<3, Real> operator-(Vec<3, Real> a, Vec<3, Real> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class Real>
JD Vec<3, Real> operator*(Vec<3, Real> a, Real s) {
  return {a.x * s, a.y * s, a.z * s};
}
template <class Real>
JD Vec<3, Real> operator*(Vec<3, Real> a, Vec<3, Real> b) {
  return {a.x * b.x, a.y * b.y, a.z * b.z};
}
template <class Real>
JD Vec<3, Real> operator/(Vec<3, Real> a, Real s) {
  return {a.x / s, a.y / s, a.z / s};
}
template <class Real>
JD Vec<3, Real> operator-(Vec<3, Real> a) {
  return {-a.x, -a.y, -a.z};
}
template <class Real>
JD Vec<2, Real> operator+(Vec<2, Real> a, Vec<2, Real> b) {
  return {a.x + b.x, a.y + b.y};
}
template <class Real>
JD Vec<2, Real> operator-(Vec<2, Real> a, Vec<2, Real> b) {
  return {a.x - b.x, a.y - b.y};
}
template <class Real>
JD Vec<2, Real> operator*(Vec<2, Real> a, Real s) {
  return {a.x * s, a.y * s};
}
template <class Real>
JD Vec<2, Real> operator*(Vec<2, Real> a, Vec<2, Real> b) {
  return {a.x * b.x, a.y * b.y};
}
template <class Real>
JD Vec<2, Real> operator/(Vec<2, Real> a, Real s) {
  return {a.x / s, a.y / s};
}
template <class Real>
JD Vec<2, Real> operator-(Vec<2, Real> a) {
  return {-a.x, -a.y};
}
template <class Real>
JD Ang1T<Real> operator+(Ang1T<Real> a, Ang1T<Real> b) {
  return {a.z + b.z};
}
template <class Real>
JD Ang1T<Real> operator-(Ang1T<Real> a, Ang1T<Real> b) {
  return {a.z - b.z};
}
template <class Real>
JD Ang1T<Real> operator*(Ang1T<Real> a, Real s) {
  return {a.z * s};
}
template <class Real>
JD Ang1T<Real> operator*(Ang1T<Real> a, Ang1T<Real> b) {
  return {a.z * b.z};
}
template <class Real>
JD Ang1T<Real> operator/(Ang1T<Real> a, Real s) {
  return {a.z / s};
}

// jnp.vecdot with left-to-right association (linalg.py:71)
template <class Real>
JD Real dot(Vec<3, Real> a, Vec<3, Real> b) {
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
template <class Real>
JD Real dot(Vec<2, Real> a, Vec<2, Real> b) {
  return a.x * b.x + a.y * b.y;
}
template <class Real>
JD Real dot(Ang1T<Real> a, Ang1T<Real> b) {
  return a.z * b.z;
}

// linalg.py:15-60
template <class Real>
JD Vec<3, Real> cross(Vec<3, Real> a, Vec<3, Real> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// 2D cross product as the (N,1) torque
template <class Real>
JD Ang1T<Real> cross_t(Vec<2, Real> a, Vec<2, Real> b) {
  return {a.x * b.y - a.y * b.x};
}
template <class Real>
JD Vec<3, Real> cross_t(Vec<3, Real> a, Vec<3, Real> b) {
  return cross(a, b);
}
// cross_3X3D_1X2D (linalg.py:186-231): w x r
template <class Real>
JD Vec<3, Real> cross_wr(Vec<3, Real> w, Vec<3, Real> r) {
  return cross(w, r);
}
template <class Real>
JD Vec<2, Real> cross_wr(Ang1T<Real> w, Vec<2, Real> r) {
  return {-w.z * r.y, w.z * r.x};
}

template <class Real>
JD Real rsq(Real x) {
  return Real(1) / sqrt_(x);
}

template <int D, class Real = float>
JD Vec<D, Real> zero_vec() {
  if constexpr (D == 3)
    return {Real(0), Real(0), Real(0)};
  else
    return {Real(0), Real(0)};
}
template <int D, class Real = float>
JD Ang<D, Real> zero_ang() {
  if constexpr (D == 3)
    return {Real(0), Real(0), Real(0)};
  else
    return {Real(0)};
}

template <int D, class Real>
JD Vec<D, Real> ld(const Real* p, int64_t i) {
  if constexpr (D == 3)
    return {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
  else
    return {p[2 * i], p[2 * i + 1]};
}
template <class Real>
JD void st(Real* p, int64_t i, Vec<3, Real> v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
template <class Real>
JD void st(Real* p, int64_t i, Vec<2, Real> v) {
  p[2 * i] = v.x;
  p[2 * i + 1] = v.y;
}
template <int D, class Real>
JD Ang<D, Real> lda(const Real* p, int64_t i) {
  if constexpr (D == 3)
    return ld<3>(p, i);
  else
    return {p[i]};
}
template <class Real>
JD void sta(Real* p, int64_t i, Vec<3, Real> v) {
  st(p, i, v);
}
template <class Real>
JD void sta(Real* p, int64_t i, Ang1T<Real> v) {
  p[i] = v.z;
}

// ---------------------------------------------------------------------------
// quaternions [w, x, y, z] (utils/quaternion.py)
// ---------------------------------------------------------------------------
template <class Real = float>
struct QuatT {
  Real w, x, y, z;
};
using Quat = QuatT<float>;
template <class Real>
JD Vec<3, Real> qv(QuatT<Real> q) {
  return {q.x, q.y, q.z};
}
JD Quat ldq(const float* p, int64_t i) {
  const float4 v = reinterpret_cast<const float4*>(p)[i];
  return {v.x, v.y, v.z, v.w};
}
JD QuatT<double> ldq(const double* p, int64_t i) {
  const double2 a = reinterpret_cast<const double2*>(p)[2 * i];
  const double2 b = reinterpret_cast<const double2*>(p)[2 * i + 1];
  return {a.x, a.y, b.x, b.y};
}
JD void stq(float* p, int64_t i, Quat q) { reinterpret_cast<float4*>(p)[i] = make_float4(q.w, q.x, q.y, q.z); }
JD void stq(double* p, int64_t i, QuatT<double> q) {
  reinterpret_cast<double2*>(p)[2 * i] = make_double2(q.w, q.x);
  reinterpret_cast<double2*>(p)[2 * i + 1] = make_double2(q.y, q.z);
}

// quaternion.py:347-367
template <class Real>
JD QuatT<Real> q_mul(QuatT<Real> a, QuatT<Real> b) {
  Vec<3, Real> av = qv(a), bv = qv(b);
  Real w = a.w * b.w - dot(av, bv);
  Vec<3, Real> xyz = (bv * a.w + av * b.w) + cross(av, bv);
  return {w, xyz.x, xyz.y, xyz.z};
}
// quaternion.py:75-94
template <class Real>
JD QuatT<Real> q_unit(QuatT<Real> q) {
  Vec<3, Real> v = qv(q);
  Real n2 = q.w * q.w + dot(v, v);
  Real safe = (n2 == Real(0)) ? Real(1) : n2;
  Real inv = rsq(safe);
  return {q.w * inv, q.x * inv, q.y * inv, q.z * inv};
}
// Quaternion.conj (quaternion.py:143-162): (w, -xyz), what rotate_back
// applies (quaternion.py:267-268)
template <class Real>
JD QuatT<Real> q_conj(QuatT<Real> q) {
  return {q.w, -q.x, -q.y, -q.z};
}
// quaternion.py:193-242 (3D)
template <class Real>
JD Vec<3, Real> q_rotate(QuatT<Real> q, Vec<3, Real> v) {
  Vec<3, Real> t = cross(qv(q), v) * Real(2);
  return (v + t * q.w) + cross(qv(q), t);
}
// quaternion.py:226-236: a 2D vector turns about z
template <class Real>
JD Vec<2, Real> q_rotate(QuatT<Real> q, Vec<2, Real> v) {
  Real c = q.w * q.w - q.z * q.z;
  Real s = (Real(2) * q.w) * q.z;
  return {c * v.x - s * v.y, s * v.x + c * v.y};
}
// utils.norm (linalg.py:113-133)
template <class V>
JD auto norm_(V v) {
  auto n2 = dot(v, v);
  using Real = decltype(n2);
  return (n2 == Real(0)) ? Real(0) : sqrt_(fmax_(n2, Real(1e-16)));
}
// quaternion.py:247-268
template <class Real>
JD Vec<3, Real> q_rotate_back(QuatT<Real> q, Vec<3, Real> v) {
  return q_rotate(q_conj(q), v);
}
// quaternion.py:132-141
template <class Real>
JD QuatT<Real> q_from_small_rotvec(Vec<3, Real> r) {
  Real n2 = dot(r, r);
  Real ch = Real(1) - n2 / Real(8);
  Real sc = Real(0.5) - n2 / Real(48);
  return {ch, r.x * sc, r.y * sc, r.z * sc};
}
template <class Real>
JD Vec<3, Real> to_rotvec3(Vec<3, Real> v) {
  return v;
}
// velocity_verlet_spiral.py:108: [0,0,1] * rotvec in 2D
template <class Real>
JD Vec<3, Real> to_rotvec3(Ang1T<Real> v) {
  return {Real(0) * v.z, Real(0) * v.z, v.z};
}

// ---------------------------------------------------------------------------
// domain helpers (domains/__init__.py:128-160, periodic.py:39-79)
// ---------------------------------------------------------------------------
template <class Real = float>
struct DomPT {
  int kind;
  Real box[3], inv_box[3], anchor[3];
  Real restitution;
};
using DomP = DomPT<float>;

// Domain._displacement: periodic uses r - B*round(r*invB)
template <bool PER, class Real>
JD Real disp1_(Real a, Real b, Real B, Real iB) {
  Real r = a - b;
  if (PER) r = r - B * rint_(r * iB);
  return r;
}
// Domain.displacement: periodic uses r - B*round(r/B)
template <bool PER, class Real>
JD Real disp1(Real a, Real b, Real B) {
  Real r = a - b;
  if (PER) r = r - B * rint_(r / B);
  return r;
}
// Domain.displacement for squared distances only.  When |r| <= B/4, r/B
// rounds to a value of magnitude <= 1/4, round(r/B) is +-0 and the result is
// r up to the sign of a zero, which a square does not see; the division is
// needed only for pairs across the periodic boundary.
template <bool PER, class Real>
JD Real disp1_sq(Real a, Real b, Real B) {
  Real r = a - b;
  if (PER && !(fabs_(r) <= Real(0.25) * B)) r = r - B * rint_(r / B);
  return r;
}
template <bool PER, class Real>
JD Real dist_sq(Vec<3, Real> a, Vec<3, Real> b, const DomPT<Real>& d) {
  const Vec<3, Real> r = {disp1_sq<PER>(a.x, b.x, d.box[0]), disp1_sq<PER>(a.y, b.y, d.box[1]),
                          disp1_sq<PER>(a.z, b.z, d.box[2])};
  return dot(r, r);
}
template <bool PER, class Real>
JD Real dist_sq(Vec<2, Real> a, Vec<2, Real> b, const DomPT<Real>& d) {
  const Vec<2, Real> r = {disp1_sq<PER>(a.x, b.x, d.box[0]), disp1_sq<PER>(a.y, b.y, d.box[1])};
  return dot(r, r);
}
template <bool PER, class Real>
JD Vec<3, Real> disp_(Vec<3, Real> a, Vec<3, Real> b, const DomPT<Real>& d) {
  return {disp1_<PER>(a.x, b.x, d.box[0], d.inv_box[0]), disp1_<PER>(a.y, b.y, d.box[1], d.inv_box[1]),
          disp1_<PER>(a.z, b.z, d.box[2], d.inv_box[2])};
}
template <bool PER, class Real>
JD Vec<2, Real> disp_(Vec<2, Real> a, Vec<2, Real> b, const DomPT<Real>& d) {
  return {disp1_<PER>(a.x, b.x, d.box[0], d.inv_box[0]), disp1_<PER>(a.y, b.y, d.box[1], d.inv_box[1])};
}
template <bool PER, class Real>
JD Vec<3, Real> disp(Vec<3, Real> a, Vec<3, Real> b, const DomPT<Real>& d) {
  return {disp1<PER>(a.x, b.x, d.box[0]), disp1<PER>(a.y, b.y, d.box[1]), disp1<PER>(a.z, b.z, d.box[2])};
}
template <bool PER, class Real>
JD Vec<2, Real> disp(Vec<2, Real> a, Vec<2, Real> b, const DomPT<Real>& d) {
  return {disp1<PER>(a.x, b.x, d.box[0]), disp1<PER>(a.y, b.y, d.box[1])};
}

// ---------------------------------------------------------------------------
// pair laws
// ---------------------------------------------------------------------------
template <class Real = float>
struct LawPT {
  int kind, n_mat;
  const Real *t0, *t1, *t2, *m0, *m1, *m2, *m3;
  Real p0, p1;
  // routed laws: the terms of ordered species pair p are
  // route_terms[route_off[p] .. route_off[p + 1]) (device arrays)
  int n_species;
  const int32_t* route_off;
  const LawPT* route_terms;
  const int32_t* species;
};
using LawP = LawPT<float>;

// Kernel kinds of a routed law (jdem_law kind ROUTED): without and with a
// history-carrying term.
constexpr int kLawRouted = JDEM_LAW_ROUTED;
constexpr int kLawRoutedHist = JDEM_LAW_ROUTED + 1;
// Material and species ids share the per-particle word of the resident
// layout: fixed | mat << 1 | species << 16.
constexpr int kMatMask = 0x7fff;

// Data of one side of a pair.
template <int D, class Real = float>
struct Side {
  Vec<D, Real> x;
  Real r;
  int mat;
  int sp;  // species (routed laws)
  Vec<D, Real> v;
  Ang<D, Real> w;
  Real m;
};

template <int LAW>
struct LawTraits {
  static constexpr bool routed = (LAW == kLawRouted || LAW == kLawRoutedHist);
  // Does the law read velocities / masses of the partner?
  static constexpr bool kin = (LAW == JDEM_LAW_CUNDALL_STRACK || LAW == JDEM_LAW_SPRING_FRICTION ||
                               LAW == JDEM_LAW_HERTZ_FRICTION || routed);
  static constexpr bool hist = (LAW == JDEM_LAW_SPRING_FRICTION || LAW == JDEM_LAW_HERTZ_FRICTION ||
                                LAW == kLawRoutedHist);
  static constexpr bool torque = kin;
};

template <class Real>
JD Real tab(const Real* t, int n_mat, int a, int b) {
  return t[a * n_mat + b];
}

// Result of one pair evaluation.
template <int D, class Real = float>
struct PairOut {
  Vec<D, Real> F;
  Ang<D, Real> T;
  Vec<D, Real> xi;  // updated tangential spring (history laws)
  bool contact;
};

// Force on i from j.  rij = _displacement(x_i, x_j); d_sq = |rij|^2.
// spring.py:97-108
template <int D, class Real>
JD PairOut<D, Real> law_spring(const LawPT<Real>& L, const Side<D, Real>& a, const Side<D, Real>& b,
                               Vec<D, Real> rij, Real d_sq) {
  PairOut<D, Real> o;
  Real R = a.r + b.r;
  Real k = (L.n_mat == 1) ? L.t0[0] : tab(L.t0, L.n_mat, a.mat, b.mat);
  Real inv = (d_sq == Real(0)) ? Real(0) : rsq(fmax_(d_sq, Real(1e-16)));
  Real r = d_sq * inv;
  Real delta = fmax_(Real(0), R - r);
  Real mag = (k * delta) * inv;
  o.F = rij * mag;
  o.T = zero_ang<D, Real>();
  o.contact = delta > Real(0);
  return o;
}

// hertz.py:98-117 with pow(delta, 1.5) = delta*sqrt(delta)
template <int D, class Real>
JD PairOut<D, Real> law_hertz(const LawPT<Real>& L, const Side<D, Real>& a, const Side<D, Real>& b,
                              Vec<D, Real> rij, Real d_sq) {
  PairOut<D, Real> o;
  Real kE = (L.n_mat == 1) ? L.t0[0] : tab(L.t0, L.n_mat, a.mat, b.mat);
  Real R_star = (a.r * b.r) / (a.r + b.r);
  Real k = kE * sqrt_(R_star);
  Real inv = (d_sq == Real(0)) ? Real(0) : rsq(fmax_(d_sq, Real(1e-16)));
  Real r = d_sq * inv;
  Vec<D, Real> n = rij * inv;
  Real delta = fmax_(Real(0), (a.r + b.r) - r);
  Real mag = k * (delta * sqrt_(delta));
  o.F = n * mag;
  o.T = zero_ang<D, Real>();
  o.contact = delta > Real(0);
  return o;
}

// cundall_strack.py:123-193
template <int D, class Real>
JD PairOut<D, Real> law_cundall(const LawPT<Real>& L, const Side<D, Real>& a, const Side<D, Real>& b,
                                Vec<D, Real> rij, Real d_sq) {
  PairOut<D, Real> o;
  const int M = L.n_mat;
  Real Ei = L.m0[a.mat], Ej = L.m0[b.mat];
  Real Gi = L.m1[a.mat], Gj = L.m1[b.mat];
  Real kn = ((((Real(2) * Ei) * a.r) * Ej) * b.r) / (Ei * a.r + Ej * b.r);
  Real kt = ((((Real(2) * Gi) * a.r) * Gj) * b.r) / (Gi * a.r + Gj * b.r);
  Real m_eff = (a.m * b.m) / (a.m + b.m);
  Real beta = tab(L.t0, M, a.mat, b.mat);
  Real mu_eff = fmin_(L.m2[a.mat], L.m2[b.mat]);
  Real gamma_n = (Real(2) * beta) * sqrt_(kn * m_eff);
  Real gamma_t = (Real(2) * beta) * sqrt_(kt * m_eff);
  Real inv = (d_sq == Real(0)) ? Real(0) : rsq(fmax_(d_sq, Real(1e-16)));
  Real r = d_sq * inv;
  Vec<D, Real> n = rij * inv;
  Real delta = (a.r + b.r) - r;
  Real isc = (delta > Real(0)) ? Real(1) : Real(0);
  delta = delta * isc;
  Vec<D, Real> r_ci = n * (-a.r);
  Vec<D, Real> r_cj = n * b.r;
  Vec<D, Real> v_ci = a.v + cross_wr(a.w, r_ci);
  Vec<D, Real> v_cj = b.v + cross_wr(b.w, r_cj);
  Vec<D, Real> v_rel = v_ci - v_cj;
  Real vn = dot(v_rel, n);
  Vec<D, Real> vt_vec = v_rel - n * vn;
  Real vt2 = dot(vt_vec, vt_vec);
  Real tinv = (vt2 == Real(0)) ? Real(0) : rsq(fmax_(vt2, Real(1e-16)));
  Vec<D, Real> t = vt_vec * tinv;
  Real vt = vt2 * tinv;
  Real Fn = fmax_(Real(0), kn * delta - gamma_n * vn) * isc;
  Real Ft = gamma_t * vt;
  Real Ft_max = mu_eff * Fn;
  Ft = fmin_(Ft, Ft_max) * isc;
  o.F = n * Fn - t * Ft;
  Ang<D, Real> tq = cross_t(r_ci, o.F);
  Real mu_r_eff = fmin_(L.m3[a.mat], L.m3[b.mat]);
  Real R_eff = (a.r * b.r) / (a.r + b.r);
  Ang<D, Real> w_rel = a.w - b.w;
  Real w2 = dot(w_rel, w_rel);
  Real wsafe = (w2 == Real(0)) ? Real(1) : w2;
  Ang<D, Real> w_hat = w_rel * rsq(wsafe);
  o.T = tq - w_hat * ((mu_r_eff * R_eff) * Fn);
  o.contact = isc > Real(0);
  return o;
}

// Tangential spring with Coulomb cap and history (no reference counterpart;
// Cundall & Strack 1979 in the form of Luding 2008).  HERTZ=true uses the
// Hertz-Mindlin stiffnesses kn*sqrt(R* d), kt*sqrt(R* d).
template <int D, bool HERTZ, class Real>
JD PairOut<D, Real> law_friction(const LawPT<Real>& L, const Side<D, Real>& a, const Side<D, Real>& b,
                                 Vec<D, Real> rij, Real d_sq, Vec<D, Real> xi_old, Real dt) {
  PairOut<D, Real> o;
  const int M = L.n_mat;
  Real kn = (M == 1) ? L.t0[0] : tab(L.t0, M, a.mat, b.mat);
  Real kt = (M == 1) ? L.t1[0] : tab(L.t1, M, a.mat, b.mat);
  Real mu = (M == 1) ? L.t2[0] : tab(L.t2, M, a.mat, b.mat);
  Real inv = (d_sq == Real(0)) ? Real(0) : rsq(fmax_(d_sq, Real(1e-16)));
  Real r = d_sq * inv;
  Vec<D, Real> n = rij * inv;
  Real delta = fmax_(Real(0), (a.r + b.r) - r);
  o.contact = delta > Real(0);
  if (!o.contact) {
    o.F = zero_vec<D, Real>();
    o.T = zero_ang<D, Real>();
    o.xi = zero_vec<D, Real>();
    return o;
  }
  if (HERTZ) {
    Real R_star = (a.r * b.r) / (a.r + b.r);
    Real s = sqrt_(R_star * delta);
    kn = kn * s;
    kt = kt * s;
  }
  Real m_eff = (a.m * b.m) / (a.m + b.m);
  Real gamma_n = L.p0 * sqrt_(kn * m_eff);
  Vec<D, Real> r_ci = n * (-a.r);
  Vec<D, Real> r_cj = n * b.r;
  Vec<D, Real> v_ci = a.v + cross_wr(a.w, r_ci);
  Vec<D, Real> v_cj = b.v + cross_wr(b.w, r_cj);
  Vec<D, Real> v_rel = v_ci - v_cj;
  Real vn = dot(v_rel, n);
  Vec<D, Real> vt = v_rel - n * vn;
  Real Fn = fmax_(Real(0), kn * delta - gamma_n * vn);
  Vec<D, Real> xi = xi_old - n * dot(xi_old, n);
  xi = xi + vt * dt;
  Vec<D, Real> Ft = xi * (-kt);
  Real ft2 = dot(Ft, Ft);
  Real fmax = mu * Fn;
  if (ft2 > fmax * fmax) {
    Real scale = fmax * rsq(ft2);
    Ft = Ft * scale;
    if (kt > Real(0)) xi = Ft / (-kt);
  }
  o.F = n * Fn + Ft;
  o.T = cross_t(r_ci, o.F);
  o.xi = xi;
  return o;
}

// Routed laws (router.py:70-160, law_combiner.py:60-93): the terms of the
// pair's species entry, in order; the first gives the force and torque, each
// later one is added to them.  A LawCombiner entry starts with a NONE term
// (the combiner's zeros); the entry's history term, if any, gives the
// tangential spring and the contact flag.
template <class Real>
JD int route_of(const LawPT<Real>& L, int sa, int sb) {
  return L.n_species > 1 ? sa * L.n_species + sb : 0;
}

template <int D, bool H, class Real>
JD PairOut<D, Real> law_routed(const LawPT<Real>& L, const Side<D, Real>& a, const Side<D, Real>& b,
                               Vec<D, Real> rij, Real d_sq, Vec<D, Real> xi_old, Real dt) {
  const int p = route_of(L, a.sp, b.sp);
  const int k0 = L.route_off[p], k1 = L.route_off[p + 1];
  PairOut<D, Real> o;
  o.F = zero_vec<D, Real>();
  o.T = zero_ang<D, Real>();
  o.xi = zero_vec<D, Real>();
  o.contact = false;
  for (int k = k0; k < k1; ++k) {
    const LawPT<Real>& T = L.route_terms[k];
    PairOut<D, Real> r;
    r.F = zero_vec<D, Real>();
    r.T = zero_ang<D, Real>();
    switch (T.kind) {
      case JDEM_LAW_SPRING:
        r = law_spring<D>(T, a, b, rij, d_sq);
        break;
      case JDEM_LAW_HERTZ:
        r = law_hertz<D>(T, a, b, rij, d_sq);
        break;
      case JDEM_LAW_CUNDALL_STRACK:
        r = law_cundall<D>(T, a, b, rij, d_sq);
        break;
      case JDEM_LAW_SPRING_FRICTION:
        if constexpr (H) {
          r = law_friction<D, false>(T, a, b, rij, d_sq, xi_old, dt);
          o.xi = r.xi;
          o.contact = r.contact;
        }
        break;
      case JDEM_LAW_HERTZ_FRICTION:
        if constexpr (H) {
          r = law_friction<D, true>(T, a, b, rij, d_sq, xi_old, dt);
          o.xi = r.xi;
          o.contact = r.contact;
        }
        break;
      default:
        break;
    }
    if (k == k0) {
      o.F = r.F;
      o.T = r.T;
    } else {
      o.F = o.F + r.F;
      o.T = o.T + r.T;
    }
  }
  return o;
}

template <int D, int LAW, class Real>
JD PairOut<D, Real> law_eval(const LawPT<Real>& L, const Side<D, Real>& a, const Side<D, Real>& b,
                             Vec<D, Real> rij, Real d_sq, Vec<D, Real> xi_old, Real dt) {
  if constexpr (LAW == JDEM_LAW_SPRING) return law_spring<D>(L, a, b, rij, d_sq);
  if constexpr (LAW == JDEM_LAW_HERTZ) return law_hertz<D>(L, a, b, rij, d_sq);
  if constexpr (LAW == JDEM_LAW_CUNDALL_STRACK) return law_cundall<D>(L, a, b, rij, d_sq);
  if constexpr (LAW == JDEM_LAW_SPRING_FRICTION) return law_friction<D, false>(L, a, b, rij, d_sq, xi_old, dt);
  if constexpr (LAW == JDEM_LAW_HERTZ_FRICTION) return law_friction<D, true>(L, a, b, rij, d_sq, xi_old, dt);
  if constexpr (LAW == kLawRouted) return law_routed<D, false>(L, a, b, rij, d_sq, xi_old, dt);
  if constexpr (LAW == kLawRoutedHist) return law_routed<D, true>(L, a, b, rij, d_sq, xi_old, dt);
  PairOut<D, Real> o;
  o.F = zero_vec<D, Real>();
  o.T = zero_ang<D, Real>();
  o.contact = false;
  return o;
}

// Pair energy (spring.py:138-147, hertz.py:147-162, cundall_strack.py:220-233);
// the history laws report their elastic normal part.
template <int D, int LAW, class Real>
JD Real law_energy(const LawPT<Real>& L, const Side<D, Real>& a, const Side<D, Real>& b, Real d_sq) {
  const int M = L.n_mat;
  if constexpr (LAW == JDEM_LAW_SPRING || LAW == JDEM_LAW_SPRING_FRICTION) {
    Real k = (M == 1) ? L.t0[0] : tab(L.t0, M, a.mat, b.mat);
    Real inv = (d_sq == Real(0)) ? Real(0) : rsq(fmax_(d_sq, Real(1e-16)));
    Real r = d_sq * inv;
    Real s = fmax_(Real(0), (a.r + b.r) - r);
    return (Real(0.5) * k) * (s * s);
  }
  Real r = (d_sq == Real(0)) ? Real(0) : sqrt_(fmax_(d_sq, Real(1e-16)));
  Real delta = (a.r + b.r) - r;
  delta = delta * ((delta > Real(0)) ? Real(1) : Real(0));
  if constexpr (LAW == JDEM_LAW_HERTZ || LAW == JDEM_LAW_HERTZ_FRICTION) {
    Real kE = (M == 1) ? L.t0[0] : tab(L.t0, M, a.mat, b.mat);
    Real R_star = (a.r * b.r) / (a.r + b.r);
    Real k = kE * sqrt_(R_star);
    return (Real(0.4) * k) * ((delta * delta) * sqrt_(delta));
  }
  if constexpr (LAW == JDEM_LAW_CUNDALL_STRACK) {
    Real Ei = L.m0[a.mat], Ej = L.m0[b.mat];
    Real kn = ((((Real(2) * Ei) * a.r) * Ej) * b.r) / (Ei * a.r + Ej * b.r);
    return (Real(0.5) * kn) * (delta * delta);
  }
  if constexpr (LAW == kLawRouted || LAW == kLawRoutedHist) {
    // router.py:262-330 selects the entry; law_combiner.py:180-188 sums it
    const int p = route_of(L, a.sp, b.sp);
    const int k0 = L.route_off[p], k1 = L.route_off[p + 1];
    Real e = Real(0);
    for (int k = k0; k < k1; ++k) {
      const LawPT<Real>& T = L.route_terms[k];
      Real ek = Real(0);
      switch (T.kind) {
        case JDEM_LAW_SPRING:
        case JDEM_LAW_SPRING_FRICTION:
          ek = law_energy<D, JDEM_LAW_SPRING>(T, a, b, d_sq);
          break;
        case JDEM_LAW_HERTZ:
        case JDEM_LAW_HERTZ_FRICTION:
          ek = law_energy<D, JDEM_LAW_HERTZ>(T, a, b, d_sq);
          break;
        case JDEM_LAW_CUNDALL_STRACK:
          ek = law_energy<D, JDEM_LAW_CUNDALL_STRACK>(T, a, b, d_sq);
          break;
        default:
          break;
      }
      e = (k == k0) ? ek : e + ek;
    }
    return e;
  }
  return Real(0);
}

// ---------------------------------------------------------------------------
// integrators
// ---------------------------------------------------------------------------
// spiral.py:24-40
template <class Real>
JD Vec<3, Real> omega_dot(Vec<3, Real> w, Vec<3, Real> tq, Vec<3, Real> I, Vec<3, Real> invI) {
  Real cx = w.y * (I.z * w.z) - w.z * (I.y * w.y);
  Real cy = w.z * (I.x * w.x) - w.x * (I.z * w.z);
  Real cz = w.x * (I.y * w.y) - w.y * (I.x * w.x);
  return (tq - Vec<3, Real>{cx, cy, cz}) * invI;
}
template <class Real>
JD Ang1T<Real> omega_dot(Ang1T<Real> w, Ang1T<Real> tq, Ang1T<Real> I, Ang1T<Real> invI) {
  return tq * invI;
}

template <class Real>
JD Vec<3, Real> inv_of(Vec<3, Real> I) {
  return {Real(1) / I.x, Real(1) / I.y, Real(1) / I.z};
}
template <class Real>
JD Ang1T<Real> inv_of(Ang1T<Real> I) {
  return {Real(1) / I.z};
}

template <class Real>
JD Vec<3, Real> body_of(QuatT<Real> q, Vec<3, Real> v) {
  return q_rotate_back(q, v);
}
template <class Real>
JD Ang1T<Real> body_of(QuatT<Real>, Ang1T<Real> v) {
  return v;
}
template <class Real>
JD Vec<3, Real> lab_of(QuatT<Real> q, Vec<3, Real> v) {
  return q_rotate(q, v);
}
template <class Real>
JD Ang1T<Real> lab_of(QuatT<Real>, Ang1T<Real> v) {
  return v;
}

// SSPRK3 kick of velocity_verlet_spiral.py:100-106 (also 168-174)
template <class A, class Real>
JD A ssprk3_half(A w, A tq, A I, A invI, Real dt_2, Real mask) {
  A k1 = omega_dot(w, tq, I, invI) * dt_2;
  A k2 = omega_dot(w + k1, tq, I, invI) * dt_2;
  A k3 = omega_dot(w + (k1 + k2) * Real(0.25), tq, I, invI) * dt_2;
  return w + (((k1 + k2) + k3 * Real(4)) * mask) / Real(6);
}

// VelocityVerletSpiral.step_before_force (velocity_verlet_spiral.py:83-116)
template <int D, class Real>
JD void vvs_before(QuatT<Real>& q, Ang<D, Real>& w_lab, Ang<D, Real> tq_lab, Ang<D, Real> I, Real dt, Real mask) {
  Real dt_2 = dt / Real(2);
  Ang<D, Real> invI = inv_of(I);
  Ang<D, Real> w = body_of(q, w_lab);
  Ang<D, Real> tq = body_of(q, tq_lab);
  w = ssprk3_half(w, tq, I, invI, dt_2, mask);
  Vec<3, Real> rv = to_rotvec3(w * (Real(2) * dt_2));
  QuatT<Real> dq = q_from_small_rotvec(rv);
  q = q_unit(q_mul(q, dq));
  w_lab = lab_of(q, w);
}
// VelocityVerletSpiral.step_after_force (velocity_verlet_spiral.py:156-180)
template <int D, class Real>
JD void vvs_after(QuatT<Real> q, Ang<D, Real>& w_lab, Ang<D, Real> tq_lab, Ang<D, Real> I, Real dt, Real mask) {
  Real dt_2 = dt / Real(2);
  Ang<D, Real> invI = inv_of(I);
  Ang<D, Real> w = body_of(q, w_lab);
  Ang<D, Real> tq = body_of(q, tq_lab);
  w = ssprk3_half(w, tq, I, invI, dt_2, mask);
  w_lab = lab_of(q, w);
}
// Spiral.step_after_force (spiral.py:60-144)
template <int D, class Real>
JD void spiral_after(QuatT<Real>& q, Ang<D, Real>& w_lab, Ang<D, Real> tq_lab, Ang<D, Real> I, Real dt, Real mask) {
  Real dt_2 = dt / Real(2);
  Ang<D, Real> invI = inv_of(I);
  Ang<D, Real> w = body_of(q, w_lab);
  Ang<D, Real> tq = body_of(q, tq_lab);
  Ang<D, Real> w_dot = omega_dot(w, tq, I, invI);
  Vec<3, Real> rv1 = to_rotvec3(w * (Real(2) * dt_2));
  Vec<3, Real> rv2 = to_rotvec3(w_dot * ((Real(2) * dt_2) * dt_2));
  QuatT<Real> dq = q_mul(q_from_small_rotvec(rv1), q_from_small_rotvec(rv2));
  q = q_unit(q_mul(q, dq));
  Ang<D, Real> k1 = w_dot * dt;
  Ang<D, Real> k2 = omega_dot(w + k1, tq, I, invI) * dt;
  Ang<D, Real> k3 = omega_dot(w + (k1 + k2) * Real(0.25), tq, I, invI) * dt;
  w = w + (((k1 + k2) + k3 * Real(4)) * mask) / Real(6);
  w_lab = lab_of(q, w);
}

// ---------------------------------------------------------------------------
// reflect domain (reflect.py:165-297, _toc.py:11-94), one sphere
// ---------------------------------------------------------------------------
template <class Real>
JD Real toc_alpha(Real v, Real acc, Real delta, Real ws, Real dt) {
  Real v0 = v - dt * acc;
  Real vmid = v - (Real(0.5) * dt) * acc;
  Real v0n = v0 * ws;
  Real vmidn = vmid * ws;
  Real accn = acc * ws;
  Real A = ((Real(0.5) * accn) * dt) * dt;
  Real B = v0n * dt;
  Real dw = delta + vmidn * dt;
  Real disc = B * B + (Real(4) * A) * dw;
  disc = fmax_(Real(0), disc);
  Real sq = (disc > Real(0)) ? sqrt_(disc) : Real(0);
  Real alpha;
  if (B < Real(0)) {
    Real den = B - sq;
    alpha = (Real(2) * dw) / ((den < Real(-1e-10)) ? den : Real(-1));
  } else {
    Real twoA = Real(2) * A;
    alpha = ((-B) - sq) / ((fabs_(twoA) > Real(1e-10)) ? twoA : Real(1));
  }
  // jnp.clip(alpha, 0, 1) (NaN-free inputs)
  return fmin_(fmax_(alpha, Real(0)), Real(1));
}

template <class Real>
JD void to_arr(Vec<3, Real> v, Real* a) {
  a[0] = v.x;
  a[1] = v.y;
  a[2] = v.z;
}
template <class Real>
JD void to_arr(Vec<2, Real> v, Real* a) {
  a[0] = v.x;
  a[1] = v.y;
}
template <class Real>
JD void from_arr(Vec<3, Real>& v, const Real* a) {
  v = {a[0], a[1], a[2]};
}
template <class Real>
JD void from_arr(Vec<2, Real>& v, const Real* a) {
  v = {a[0], a[1]};
}

// Register arrays with fully unrolled component loops (no stack traffic).
template <int D, class Real>
JD void reflect_apply(const DomPT<Real>& dom, Vec<D, Real>& x, Vec<D, Real>& v, QuatT<Real>& q, Vec<D, Real> F,
                      Real m, Real rad, bool fixed, Real dt) {
  Real xs[D], vs[D], fs[D];
  to_arr(x, xs);
  to_arr(v, vs);
  to_arr(F, fs);
  const Real e = dom.restitution;
  const Real inv_mass = Real(1) / m;
  Real denom = inv_mass + Real(0);
  denom = (denom == Real(0)) ? Real(1) : denom;
  Real ws[D], act[D], acc[D];
  Real a_min = Real(1);
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const Real lo = dom.anchor[c] + rad;
    const Real hi = (dom.anchor[c] + dom.box[c]) - rad;
    const Real over_lo = fmax_(Real(0), lo - xs[c]);
    const Real over_hi = fmax_(Real(0), xs[c] - hi);
    const Real delta = fmax_(over_lo, over_hi);
    ws[c] = ((over_lo > Real(0)) ? Real(1) : Real(0)) - ((over_hi > Real(0)) ? Real(1) : Real(0));
    act[c] = fabs_(ws[c]);
    acc[c] = fs[c] * inv_mass;
    const Real alpha = toc_alpha(vs[c], acc[c], delta, ws[c], dt);
    if (act[c] > Real(0)) a_min = fmin_(a_min, alpha);
  }
  const Real dt_factor = (a_min - Real(1)) * dt;
  const Real dt_rem = (Real(1) - a_min) * dt;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const Real v_contact = vs[c] + dt_factor * acc[c];
    Real j = ((-(Real(1) + e)) * v_contact) / denom;
    const Real closing = ((v_contact * ws[c]) < Real(0)) ? Real(1) : Real(0);
    const Real weight = act[c] / fmax_(act[c], Real(1));
    j = j * (weight * closing);
    const Real dv = fixed ? Real(0) : j * inv_mass;
    vs[c] = vs[c] + dv;
    xs[c] = xs[c] + dv * dt_rem;
  }
  from_arr(v, vs);
  from_arr(x, xs);
  q = q_unit(q);
}

}  // namespace jdem
