// Reverse mode of one polarized DELO-Bezier depth step (delo_step in
// rhg_delo.hpp, FP64): given the adjoints of what the step produces (the
// Stokes vector at depth k and the carried derivatives and optical-depth
// interval), the adjoints of what it consumes (the carried state entering
// the step, the points k+1, k, k-1, k-2 and the path lengths).  One call
// handles NO cotangent vectors at once (the four emergent Stokes
// parameters); the forward intermediates are recomputed once and shared.
//
// The step, in the order of the code it reverses (bezier_1D.c:169-264):
//   dsdn = |dz_k| / mu, dchi_dn = FB(P0.chi, Pd.chi, Pn.chi) (or one-sided),
//   c2 = max(P0.chi + dsdn/3 dchi_c, 0), c1 = max(Pd.chi - dsdn/3 dchi_dn, 0),
//   dtau_dw = dsdn (P0.chi + Pd.chi + c1 + c2) / 4,
//   (alpha, beta, gamma, theta, eps) = Bezier3(dtau_uw), dt03 = dtau_uw / 3,
//   dS0, dq0 = FB on (dtau_uw, dtau_dw) of S and q at k+1, k, k-1,
//   V0 = eps Iv + ca Ku Iv + gu (Ku^2 + dKu) Iv + cb0 Su - gu Ku Su + cc0 S0
//        + g0 K0 S0 + dt03 (gamma dS0 - theta dSu),
//   Md = I + cM K0 + gM (K0^2 + dK0),  Iv' = Md^-1 V0,
// with ca = theta (dt03 - 1) - beta, gu = theta dt03, cb0 = beta + theta,
// cc0 = alpha + gamma, g0 = gM = gamma dt03, cM = alpha + gamma (1 + dt03),
// FB the Fritsch-Butland centred derivative and K(q) = [[0, q^T], [q, 0]].
// K is symmetric, so is Md, and the adjoint of the solve is a second solve
// with the same matrix.
#pragma once

#include "rhg_delo.hpp"
#include "rhg_dual.hpp"

namespace rhg {

// The carried state entering a depth step: what delo_step reads of
// DeloState besides the points.
struct StepCarry {
  double Iv[4], dSu[4], dqu[3], dtau_uw, dchi_c;
};
constexpr int NCARRY_STEP = RF_NCARRY;

__device__ __forceinline__ void carry_to_array(const DeloState<double> &s, double *c) {
#pragma unroll
  for (int i = 0; i < 4; i++) c[i] = s.Iv[i];
#pragma unroll
  for (int i = 0; i < 4; i++) c[4 + i] = s.dSu[i];
#pragma unroll
  for (int i = 0; i < 3; i++) c[8 + i] = s.dqu[i];
  c[11] = s.dtau_uw;
  c[12] = s.dchi_c;
}

// adjoint of one point (chi, S[4], q[3]), accumulated in A
template <typename A>
struct PointBar {
  A chi, S[4], q[3];
};

template <typename A>
__device__ __forceinline__ void pbar_zero(PointBar<A> &p) {
  p.chi = A(0);
#pragma unroll
  for (int i = 0; i < 4; i++) p.S[i] = A(0);
#pragma unroll
  for (int i = 0; i < 3; i++) p.q[i] = A(0);
}

// Adjoint of the Fritsch-Butland derivative D = FB(up, c, dn) on the
// intervals (hu, hd) (rh/bezier_aux.c:34-51): fim1 = (c - up)/hu,
// fi = (dn - c)/hd, al = (1 + hd/(hd + hu))/3, D = fim1 fi / ((1 - al) fim1
// + al fi) where fim1 fi > 0, else 0 (no gradient).  Adds Db times the
// partials to upb, cb, dnb, hub, hdb.
__device__ __forceinline__ void fb_vjp(double hu, double hd, double up, double c, double dn,
                                       double Db, double &upb, double &cb, double &dnb,
                                       double &hub, double &hdb) {
  const double idu = 1.0 / hu, idd = 1.0 / hd;
  const double fim1 = (c - up) * idu, fi = (dn - c) * idd;
  if (!(fim1 * fi > 0.0)) return;
  const double s = hd + hu;
  const double al = (1.0 / 3.0) * (1.0 + hd / s);
  const double den = (1.0 - al) * fim1 + al * fi;
  const double iden2 = 1.0 / (den * den);
  const double dfim1 = al * fi * fi * iden2;
  const double dfi = (1.0 - al) * fim1 * fim1 * iden2;
  const double dal = fim1 * fi * (fim1 - fi) * iden2;
  const double fim1b = Db * dfim1, fib = Db * dfi, alb = Db * dal;
  cb += fim1b * idu - fib * idd;
  upb -= fim1b * idu;
  dnb += fib * idd;
  hub += -fim1b * fim1 * idu - alb * (1.0 / 3.0) * hd / (s * s);
  hdb += -fib * fi * idd + alb * (1.0 / 3.0) * hu / (s * s);
}

// Point-adjoint accumulators: add(slot, o, m, v) adds v to component m
// (0 chi, 1-4 S, 5-7 q) of output o's adjoint of the point in slot (0: the
// upwind point Pu, 1: P0, 2: Pd, 3: Pn).  PbarRegs keeps them in registers.
template <typename A>
struct PbarRegs {
  PointBar<A> *p[4];
  __device__ __forceinline__ void add(int slot, int o, int m, double v) const {
    PointBar<A> &b = p[slot][o];
    if (m == 0) b.chi += A(v);
    else if (m < 5) b.S[m - 1] += A(v);
    else b.q[m - 5] += A(v);
  }
};

// u^T E_n v + v^T E_n u pattern: for K(q) = sum_n q_n E_n (E_n symmetric
// with ones at [0][n+1] and [n+1][0]), d/dq_n (a^T K(q) b) = a0 b_{n+1} +
// a_{n+1} b0.
__device__ __forceinline__ double kpair(const double *a, const double *b, int n) {
  return a[0] * b[n + 1] + a[n + 1] * b[0];
}

// Reverse of delo_step.  In: C (carried state entering the step), points
// Pu (k+1), P0 (k), Pd (k-1), Pn (k-2, when has_pn), |dz_k|, |dz_{k-1}|,
// imu = 1/mu; Cb[o] holds on entry the adjoints of the step's outputs in
// StepCarry order (Iv' = Stokes at k, dSu' = dS0, dqu' = dq0, dtau_uw' =
// dtau_dw, dchi_c' = dchi_dn) and on exit those of its inputs.  The point
// and path-length adjoints are added (in A) to Pub..Pnb, dzb, dz1b.
template <int NO, typename Acc, typename A>
__device__ __forceinline__ void delo_step_vjp(const StepCarry &C, const Point<double> &Pu,
                                              const Point<double> &P0, const Point<double> &Pd,
                                              const Point<double> &Pn, bool has_pn, double dzk,
                                              double dzk1, double imu, StepCarry *Cb,
                                              const Acc &pb, A *dzb, A *dz1b) {
  // ---- forward intermediates (as delo_step computes them)
  const double dsdn = dzk * imu;
  const double dsdn2 = has_pn ? dzk1 * imu : 0.0;
  double dchi_dn;
  if (has_pn) dchi_dn = CentDeriv<double>(dsdn, dsdn2)(P0.chi, Pd.chi, Pn.chi);
  else dchi_dn = (Pd.chi - P0.chi) / dsdn;
  const double a2 = P0.chi + (dsdn * (1.0 / 3.0)) * C.dchi_c;
  const double a1 = Pd.chi - (dsdn * (1.0 / 3.0)) * dchi_dn;
  const double c2 = a2 > 0.0 ? a2 : 0.0, c1 = a1 > 0.0 ? a1 : 0.0;
  const double dtau_dw = 0.25 * dsdn * (P0.chi + Pd.chi + c1 + c2);
  const Bezier3<Dual<double>> bzd = bezier3(Dual<double>(C.dtau_uw, 1.0));
  const double al = bzd.alpha.v, be = bzd.beta.v, ga = bzd.gamma.v, th = bzd.theta.v,
               ep = bzd.eps.v;
  const double dt03 = C.dtau_uw * (1.0 / 3.0);
  double dS0[4], dq0[3];
  {
    const CentDeriv<double> cd(C.dtau_uw, dtau_dw);
#pragma unroll
    for (int i = 0; i < 4; i++) dS0[i] = cd(Pu.S[i], P0.S[i], Pd.S[i]);
#pragma unroll
    for (int i = 0; i < 3; i++) dq0[i] = cd(Pu.q[i], P0.q[i], Pd.q[i]);
  }
  const double ca = th * (dt03 - 1.0) - be, gu = th * dt03, cb0 = be + th, cc0 = al + ga,
               g0 = ga * dt03;
  const double cM = al + ga * (1.0 + dt03);
  double KI[4], KKI[4], dKI[4], KSu[4], KS0[4], V0[4], x[4];
  kmul(Pu.q, C.Iv, KI);
  kmul(Pu.q, KI, KKI);
  kmul(C.dqu, C.Iv, dKI);
  kmul(Pu.q, Pu.S, KSu);
  kmul(P0.q, P0.S, KS0);
#pragma unroll
  for (int i = 0; i < 4; i++)
    V0[i] = ep * C.Iv[i] + ca * KI[i] + gu * (KKI[i] + dKI[i]) + cb0 * Pu.S[i] - gu * KSu[i] +
            cc0 * P0.S[i] + g0 * KS0[i] + dt03 * (ga * dS0[i] - th * C.dSu[i]);
  solve_md<double>(cM, g0, P0.q, dq0, V0, x);
  double Kx[4], KKx[4], dKx[4];
  kmul(P0.q, x, Kx);
  kmul(P0.q, Kx, KKx);
  kmul(dq0, x, dKx);

#pragma unroll
  for (int o = 0; o < NO; o++) {
    StepCarry &B = Cb[o];
    // outputs' adjoints
    double xb[4], dS0b[4], dq0b[3];
#pragma unroll
    for (int i = 0; i < 4; i++) { xb[i] = B.Iv[i]; dS0b[i] = B.dSu[i]; }
#pragma unroll
    for (int i = 0; i < 3; i++) dq0b[i] = B.dqu[i];
    double dtdwb = B.dtau_uw, dchidnb = B.dchi_c;

    // R10: x = Md^-1 V0, Md symmetric
    double u[4];
    solve_md<double>(cM, g0, P0.q, dq0, xb, u);
    double Ku[4];
    kmul(P0.q, u, Ku);
    double cMb = 0.0, gMb = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      cMb -= u[i] * Kx[i];
      gMb -= u[i] * (KKx[i] + dKx[i]);
    }
    double q0b[3];
#pragma unroll
    for (int n = 0; n < 3; n++) {
      q0b[n] = -cM * kpair(u, x, n) - g0 * (kpair(u, Kx, n) + kpair(Ku, x, n));
      dq0b[n] -= g0 * kpair(u, x, n);
    }

    // R9: V0 (u = V0b)
    double Kuu[4], KKuu[4], dKuu[4], ivb[4];
    kmul(Pu.q, u, Kuu);
    kmul(Pu.q, Kuu, KKuu);
    kmul(C.dqu, u, dKuu);
#pragma unroll
    for (int i = 0; i < 4; i++) ivb[i] = ep * u[i] + ca * Kuu[i] + gu * (KKuu[i] + dKuu[i]);
    double epb = 0.0, cab = 0.0, gub = 0.0, cb0b = 0.0, cc0b = 0.0, g0b = gMb, dt03b = 0.0,
           gab = 0.0, thb = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      epb += u[i] * C.Iv[i];
      cab += u[i] * KI[i];
      gub += u[i] * (KKI[i] + dKI[i] - KSu[i]);
      cb0b += u[i] * Pu.S[i];
      cc0b += u[i] * P0.S[i];
      g0b += u[i] * KS0[i];
      dt03b += u[i] * (ga * dS0[i] - th * C.dSu[i]);
      gab += dt03 * u[i] * dS0[i];
      thb -= dt03 * u[i] * C.dSu[i];
    }
    double K0u[4], puSb[4], p0Sb[4], pdSb[4], puqb[3], pdqb[3];
    kmul(P0.q, u, K0u);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      puSb[i] = cb0 * u[i] - gu * Kuu[i];
      p0Sb[i] = cc0 * u[i] + g0 * K0u[i];
      pdSb[i] = 0.0;
      dS0b[i] += dt03 * ga * u[i];
      B.dSu[i] = -dt03 * th * u[i];
    }
#pragma unroll
    for (int n = 0; n < 3; n++) {
      puqb[n] = ca * kpair(u, C.Iv, n) + gu * (kpair(u, KI, n) + kpair(Kuu, C.Iv, n)) -
                gu * kpair(u, Pu.S, n);
      pdqb[n] = 0.0;
      B.dqu[n] = gu * kpair(u, C.Iv, n);
      q0b[n] += g0 * kpair(u, P0.S, n);
    }
    // coefficients
    thb += cab * (dt03 - 1.0) + gub * dt03 + cb0b;
    dt03b += cab * th + gub * th + g0b * ga + cMb * ga;
    double beb = -cab + cb0b;
    double alb = cc0b + cMb;
    gab += cc0b + g0b * dt03 + cMb * (1.0 + dt03);

    // R8: centred derivatives of S and q on (dtau_uw, dtau_dw)
    double dtuwb = 0.0;
#pragma unroll
    for (int i = 0; i < 4; i++)
      fb_vjp(C.dtau_uw, dtau_dw, Pu.S[i], P0.S[i], Pd.S[i], dS0b[i], puSb[i], p0Sb[i], pdSb[i],
             dtuwb, dtdwb);
#pragma unroll
    for (int n = 0; n < 3; n++)
      fb_vjp(C.dtau_uw, dtau_dw, Pu.q[n], P0.q[n], Pd.q[n], dq0b[n], puqb[n], q0b[n], pdqb[n],
             dtuwb, dtdwb);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      pb.add(0, o, 1 + i, puSb[i]);
      pb.add(1, o, 1 + i, p0Sb[i]);
      pb.add(2, o, 1 + i, pdSb[i]);
    }
#pragma unroll
    for (int n = 0; n < 3; n++) {
      pb.add(0, o, 5 + n, puqb[n]);
      pb.add(1, o, 5 + n, q0b[n]);
      pb.add(2, o, 5 + n, pdqb[n]);
    }

    // R7, R6: dt03 and the Bezier weights as functions of dtau_uw
    dtuwb += dt03b * (1.0 / 3.0) + alb * bzd.alpha.d + beb * bzd.beta.d + gab * bzd.gamma.d +
             thb * bzd.theta.d + epb * bzd.eps.d;

    // R5: dtau_dw = 0.25 dsdn (P0.chi + Pd.chi + c1 + c2)
    double dsdnb = 0.25 * (P0.chi + Pd.chi + c1 + c2) * dtdwb;
    const double sb = 0.25 * dsdn * dtdwb;
    double p0chib = sb, pdchib = sb;
    // R4: c1 = max(Pd.chi - dsdn/3 dchi_dn, 0)
    if (a1 > 0.0) {
      pdchib += sb;
      dsdnb -= (1.0 / 3.0) * dchi_dn * sb;
      dchidnb -= (dsdn * (1.0 / 3.0)) * sb;
    }
    // R3: c2 = max(P0.chi + dsdn/3 dchi_c, 0)
    double dchicb = 0.0;
    if (a2 > 0.0) {
      p0chib += sb;
      dsdnb += (1.0 / 3.0) * C.dchi_c * sb;
      dchicb = (dsdn * (1.0 / 3.0)) * sb;
    }
    // R2: dchi_dn
    double dsdn2b = 0.0;
    if (has_pn) {
      double pnchib = 0.0;
      fb_vjp(dsdn, dsdn2, P0.chi, Pd.chi, Pn.chi, dchidnb, p0chib, pdchib, pnchib, dsdnb,
             dsdn2b);
      pb.add(3, o, 0, pnchib);
    } else {
      const double g = dchidnb / dsdn;
      pdchib += g;
      p0chib -= g;
      dsdnb -= dchi_dn * g;
    }
    pb.add(1, o, 0, p0chib);
    pb.add(2, o, 0, pdchib);
    // R1
    dzb[o] += A(imu * dsdnb);
    dz1b[o] += A(imu * dsdn2b);

    // the carried inputs' adjoints
#pragma unroll
    for (int i = 0; i < 4; i++) B.Iv[i] = ivb[i];
    B.dtau_uw = dtuwb;
    B.dchi_c = dchicb;
  }
}

}  // namespace rhg

namespace rhg {

// Reverse of delo_final (the last, linear interval, bezier_1D.c:264-299):
//   dtl = (P0.chi + Pu.chi) |dz| / (2 mu), (w0, w1) = w3(dtl), w1t = w1 / dtl,
//   V0 = w0 S0 - w1 dSu + (1 - w0) Iv - w1t K(qu) Iv,  e = w0 - w1t,
//   x0 = (V0_0 - e q0.V0_123) / (1 - e^2 |q0|^2),  x_123 = V0_123 - e q0 x0.
// xb[o] are the adjoints of the emergent Stokes vector; the adjoints of the
// carried Iv and dSu go to Cb[o] (the other carried values are not read),
// those of the points (Pu = depth 1, P0 = depth 0) and |dz(1)| are added.
// Point slots of pb: 0 = Pu, 1 = P0.
template <int NO, typename Acc, typename A>
__device__ __forceinline__ void delo_final_vjp(const StepCarry &C, const Point<double> &Pu,
                                               const Point<double> &P0, double dz, double imu,
                                               const double (*xb)[4], StepCarry *Cb,
                                               const Acc &pb, A *dzb) {
  const double dtl = 0.5 * imu * (P0.chi + Pu.chi) * dz;
  double w[3], dw[2];
  w3d(dtl, w, dw);
  const double w0 = w[0], w1 = w[1], w1t = w[1] / dtl;
  double KI[4], V0[4];
  kmul(Pu.q, C.Iv, KI);
#pragma unroll
  for (int n = 0; n < 4; n++)
    V0[n] = w0 * P0.S[n] - w1 * C.dSu[n] + (1.0 - w0) * C.Iv[n] - w1t * KI[n];
  const double e = w0 - w1t;
  const double q0V = P0.q[0] * V0[1] + P0.q[1] * V0[2] + P0.q[2] * V0[3];
  const double qq = P0.q[0] * P0.q[0] + P0.q[1] * P0.q[1] + P0.q[2] * P0.q[2];
  const double den = 1.0 - e * e * qq;
  const double x0 = (V0[0] - e * q0V) / den;
#pragma unroll
  for (int o = 0; o < NO; o++) {
    double V0b[4], q0b[3], eb = 0.0;
    double x0b = xb[o][0];
#pragma unroll
    for (int n = 0; n < 3; n++) {
      V0b[n + 1] = xb[o][n + 1];
      eb -= P0.q[n] * x0 * xb[o][n + 1];
      q0b[n] = -e * x0 * xb[o][n + 1];
      x0b -= e * P0.q[n] * xb[o][n + 1];
    }
    const double numb = x0b / den, denb = -x0 * x0b / den;
    V0b[0] = numb;
    eb += -q0V * numb - 2.0 * e * qq * denb;
    const double q0Vb = -e * numb, qqb = -e * e * denb;
#pragma unroll
    for (int n = 0; n < 3; n++) {
      q0b[n] += q0Vb * V0[n + 1] + 2.0 * P0.q[n] * qqb;
      V0b[n + 1] += q0Vb * P0.q[n];
    }
    double w0b = eb, w1tb = -eb, w1b = 0.0;
    double KV[4];
    kmul(Pu.q, V0b, KV);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      w0b += (P0.S[n] - C.Iv[n]) * V0b[n];
      w1b -= C.dSu[n] * V0b[n];
      w1tb -= KI[n] * V0b[n];
      pb.add(1, o, 1 + n, w0 * V0b[n]);
      Cb[o].dSu[n] = -w1 * V0b[n];
      Cb[o].Iv[n] = (1.0 - w0) * V0b[n] - w1t * KV[n];
    }
#pragma unroll
    for (int n = 0; n < 3; n++) {
      pb.add(0, o, 5 + n, -w1t * kpair(V0b, C.Iv, n));
      pb.add(1, o, 5 + n, q0b[n]);
    }
    w1b += w1tb / dtl;
    const double dtlb = -w1t * w1tb / dtl + w0b * dw[0] + w1b * dw[1];
    const double chib = 0.5 * imu * dz * dtlb;
    pb.add(1, o, 0, chib);
    pb.add(0, o, 0, chib);
    dzb[o] += A(0.5 * imu * (P0.chi + Pu.chi) * dtlb);
#pragma unroll
    for (int n = 0; n < 3; n++) Cb[o].dqu[n] = 0.0;
    Cb[o].dtau_uw = 0.0;
    Cb[o].dchi_c = 0.0;
  }
}

// Reverse of delo_prologue (bezier_1D.c:93-101, 169-190 at the deepest
// step): from the points N-1 (Pu), N-2 (P0), N-3 (Pd, its chi only), the path
// lengths dz1 = |dz(N-1)|, dz2 = |dz(N-2)| and the Planck values B2a =
// B(N-2), B2b = B(N-1) it forms the carried state entering the step at
// N-2 (Iv = (B2b - (B2a - B2b)/dtau0, 0, 0, 0), dSu, dqu, dtau_uw, dchi_c),
// whose adjoints are Cb[o].  Adds the adjoints of its inputs.
// Point slots of pb: 0 = Pu, 1 = P0, 2 = Pd.
template <int NO, typename Acc, typename A>
__device__ __forceinline__ void delo_prologue_vjp(const Point<double> &Pu, const Point<double> &P0,
                                                  const Point<double> &Pd, double dz1, double dz2,
                                                  double imu, double B2a, double B2b,
                                                  const StepCarry *Cb, const Acc &pb, A *dz1b,
                                                  A *dz2b, A *B2ab, A *B2bb) {
  const double dtau0 = 0.5 * imu * (Pu.chi + P0.chi) * dz1;
  const double dsup = dz1 * imu, dsdn = dz2 * imu;
  const double dchi_up = (P0.chi - Pu.chi) / dsup;
  const double dchi_c = CentDeriv<double>(dsup, dsdn)(Pu.chi, P0.chi, Pd.chi);
  const double a2 = P0.chi - (dsup * (1.0 / 3.0)) * dchi_c;
  const double a1 = Pu.chi + (dsup * (1.0 / 3.0)) * dchi_up;
  const double c2 = a2 > 0.0 ? a2 : 0.0, c1 = a1 > 0.0 ? a1 : 0.0;
  const double dtau_uw = 0.25 * dsup * (P0.chi + Pu.chi + c1 + c2);
  const double idt = 1.0 / dtau_uw;
#pragma unroll
  for (int o = 0; o < NO; o++) {
    const StepCarry &B = Cb[o];
    double idtb = 0.0;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      idtb += (P0.S[n] - Pu.S[n]) * B.dSu[n];
      pb.add(1, o, 1 + n, idt * B.dSu[n]);
      pb.add(0, o, 1 + n, -idt * B.dSu[n]);
    }
#pragma unroll
    for (int n = 0; n < 3; n++) {
      idtb += (P0.q[n] - Pu.q[n]) * B.dqu[n];
      pb.add(1, o, 5 + n, idt * B.dqu[n]);
      pb.add(0, o, 5 + n, -idt * B.dqu[n]);
    }
    const double dtub = B.dtau_uw - idt * idt * idtb;
    double dsupb = 0.25 * (P0.chi + Pu.chi + c1 + c2) * dtub;
    const double sb = 0.25 * dsup * dtub;
    double p0chib = sb, puchib = sb, pdchib = 0.0, dsdnb = 0.0, dchiupb = 0.0;
    double dchicb = B.dchi_c;
    if (a1 > 0.0) {
      puchib += sb;
      dsupb += (1.0 / 3.0) * dchi_up * sb;
      dchiupb += (dsup * (1.0 / 3.0)) * sb;
    }
    if (a2 > 0.0) {
      p0chib += sb;
      dsupb -= (1.0 / 3.0) * dchi_c * sb;
      dchicb -= (dsup * (1.0 / 3.0)) * sb;
    }
    fb_vjp(dsup, dsdn, Pu.chi, P0.chi, Pd.chi, dchicb, puchib, p0chib, pdchib, dsupb, dsdnb);
    p0chib += dchiupb / dsup;
    puchib -= dchiupb / dsup;
    dsupb -= dchi_up * dchiupb / dsup;
    const double ivb = B.Iv[0];
    const double dtau0b = ivb * (B2a - B2b) / (dtau0 * dtau0);
    B2bb[o] += A(ivb * (1.0 + 1.0 / dtau0));
    B2ab[o] += A(-ivb / dtau0);
    const double chi0b = 0.5 * imu * dz1 * dtau0b;
    p0chib += chi0b;
    puchib += chi0b;
    dz1b[o] += A(0.5 * imu * (Pu.chi + P0.chi) * dtau0b + imu * dsupb);
    dz2b[o] += A(imu * dsdnb);
    pb.add(1, o, 0, p0chib);
    pb.add(0, o, 0, puchib);
    pb.add(2, o, 0, pdchib);
  }
}

}  // namespace rhg
