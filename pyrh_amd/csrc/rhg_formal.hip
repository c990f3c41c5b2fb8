// Stage 4: emergent-ray formal solution, one thread per (column,
// wavelength).  Opacity and source functions are formed depth by depth,
// two points ahead of the recursion, from the stage-3 continuum and the
// per-depth line constants of the column state; nothing per depth is kept
// except the rolling window the recursion needs.
#include "rhg_kernels.hpp"

namespace rhg {

namespace {

// Opacity, source functions and reduced propagation-matrix parameters of
// the emergent ray at depth k (rh/background.c:519-546 sums, formal.c:178-
// 208 source functions, stokesopac.c:28-88 K').
struct Point {
  double chi;     // chi_I
  double S[4];    // eta_n / chi_I
  double q[3];    // chi_Q,U,V / chi_I  (K'[0][1..3])
};

__device__ __forceinline__ void form_point(const DevModel &M, int ls, int has_line,
                                           const double *st, int64_t fs, double mu, int moving,
                                           double cchi, double ceta, Point &p) {
  double chi4[4], eta4[4];
  chi4[0] = cchi; eta4[0] = ceta;
  chi4[1] = chi4[2] = chi4[3] = 0.0;
  eta4[1] = eta4[2] = eta4[3] = 0.0;
  if (has_line) {
    double lI, lE, lq[3], le[3];
    line_point(M, ls, st, fs, mu, lI, lq, lE, le, nullptr, moving);
    chi4[0] += lI; eta4[0] += lE;
    for (int n = 0; n < 3; n++) { chi4[n + 1] += lq[n]; eta4[n + 1] += le[n]; }
  }
  p.chi = chi4[0];
  for (int n = 0; n < 4; n++) p.S[n] = (0.0 + eta4[n]) / chi4[0];
  for (int n = 0; n < 3; n++) p.q[n] = chi4[n + 1] / p.chi;
}

__device__ __forceinline__ void make_K(const Point &p, double K[4][4]) {
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int i = 0; i < 4; i++) K[j][i] = 0.0;
  K[0][1] = p.q[0]; K[0][2] = p.q[1]; K[0][3] = p.q[2];
  K[1][0] = K[0][1]; K[2][0] = K[0][2]; K[3][0] = K[0][3];
}

// 4x4 inverse by cofactors with the pairing of rh/bezier_aux.c:235-328
// (FP64); mat is row-major and replaced by its inverse.
__device__ __forceinline__ void matinv4(double *mat) {
  double tmp[12], src[16], dst[16], det;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    src[i] = mat[i * 4]; src[i + 4] = mat[i * 4 + 1];
    src[i + 8] = mat[i * 4 + 2]; src[i + 12] = mat[i * 4 + 3];
  }
  tmp[0] = src[10] * src[15]; tmp[1] = src[11] * src[14];
  tmp[2] = src[9] * src[15];  tmp[3] = src[11] * src[13];
  tmp[4] = src[9] * src[14];  tmp[5] = src[10] * src[13];
  tmp[6] = src[8] * src[15];  tmp[7] = src[11] * src[12];
  tmp[8] = src[8] * src[14];  tmp[9] = src[10] * src[12];
  tmp[10] = src[8] * src[13]; tmp[11] = src[9] * src[12];
  dst[0] = tmp[0]*src[5] + tmp[3]*src[6] + tmp[4]*src[7];
  dst[0] -= tmp[1]*src[5] + tmp[2]*src[6] + tmp[5]*src[7];
  dst[1] = tmp[1]*src[4] + tmp[6]*src[6] + tmp[9]*src[7];
  dst[1] -= tmp[0]*src[4] + tmp[7]*src[6] + tmp[8]*src[7];
  dst[2] = tmp[2]*src[4] + tmp[7]*src[5] + tmp[10]*src[7];
  dst[2] -= tmp[3]*src[4] + tmp[6]*src[5] + tmp[11]*src[7];
  dst[3] = tmp[5]*src[4] + tmp[8]*src[5] + tmp[11]*src[6];
  dst[3] -= tmp[4]*src[4] + tmp[9]*src[5] + tmp[10]*src[6];
  dst[4] = tmp[1]*src[1] + tmp[2]*src[2] + tmp[5]*src[3];
  dst[4] -= tmp[0]*src[1] + tmp[3]*src[2] + tmp[4]*src[3];
  dst[5] = tmp[0]*src[0] + tmp[7]*src[2] + tmp[8]*src[3];
  dst[5] -= tmp[1]*src[0] + tmp[6]*src[2] + tmp[9]*src[3];
  dst[6] = tmp[3]*src[0] + tmp[6]*src[1] + tmp[11]*src[3];
  dst[6] -= tmp[2]*src[0] + tmp[7]*src[1] + tmp[10]*src[3];
  dst[7] = tmp[4]*src[0] + tmp[9]*src[1] + tmp[10]*src[2];
  dst[7] -= tmp[5]*src[0] + tmp[8]*src[1] + tmp[11]*src[2];
  tmp[0] = src[2]*src[7];  tmp[1] = src[3]*src[6];
  tmp[2] = src[1]*src[7];  tmp[3] = src[3]*src[5];
  tmp[4] = src[1]*src[6];  tmp[5] = src[2]*src[5];
  tmp[6] = src[0]*src[7];  tmp[7] = src[3]*src[4];
  tmp[8] = src[0]*src[6];  tmp[9] = src[2]*src[4];
  tmp[10] = src[0]*src[5]; tmp[11] = src[1]*src[4];
  dst[8] = tmp[0]*src[13] + tmp[3]*src[14] + tmp[4]*src[15];
  dst[8] -= tmp[1]*src[13] + tmp[2]*src[14] + tmp[5]*src[15];
  dst[9] = tmp[1]*src[12] + tmp[6]*src[14] + tmp[9]*src[15];
  dst[9] -= tmp[0]*src[12] + tmp[7]*src[14] + tmp[8]*src[15];
  dst[10] = tmp[2]*src[12] + tmp[7]*src[13] + tmp[10]*src[15];
  dst[10] -= tmp[3]*src[12] + tmp[6]*src[13] + tmp[11]*src[15];
  dst[11] = tmp[5]*src[12] + tmp[8]*src[13] + tmp[11]*src[14];
  dst[11] -= tmp[4]*src[12] + tmp[9]*src[13] + tmp[10]*src[14];
  dst[12] = tmp[2]*src[10] + tmp[5]*src[11] + tmp[1]*src[9];
  dst[12] -= tmp[4]*src[11] + tmp[0]*src[9] + tmp[3]*src[10];
  dst[13] = tmp[8]*src[11] + tmp[0]*src[8] + tmp[7]*src[10];
  dst[13] -= tmp[6]*src[10] + tmp[9]*src[11] + tmp[1]*src[8];
  dst[14] = tmp[6]*src[9] + tmp[11]*src[11] + tmp[3]*src[8];
  dst[14] -= tmp[10]*src[11] + tmp[2]*src[8] + tmp[7]*src[9];
  dst[15] = tmp[10]*src[10] + tmp[4]*src[8] + tmp[9]*src[9];
  dst[15] -= tmp[8]*src[9] + tmp[11]*src[10] + tmp[5]*src[8];
  det = src[0]*dst[0] + src[1]*dst[1] + src[2]*dst[2] + src[3]*dst[3];
  det = 1.0 / det;
#pragma unroll
  for (int i = 0; i < 16; i++) mat[i] = dst[i] * det;
}

}  // namespace

__global__ void __launch_bounds__(64)
k_formal(DevModel M, const int32_t *__restrict__ ndep, const int32_t *__restrict__ moving_arr,
         int ncol, int Nmax, const double *__restrict__ state, int nfield,
         const int32_t *__restrict__ out_ls, int nout, const double *__restrict__ cchi,
         const double *__restrict__ ceta, double mu, double *__restrict__ stokes) {
  const int io = blockIdx.x * blockDim.x + threadIdx.x;
  const int ic = blockIdx.y;
  if (io >= nout) return;
  const int N = ndep[ic];
  const int ls = out_ls[io];
  const int moving = moving_arr[ic];
  const double *st = state + (int64_t) ic * nfield * Nmax;
  const double *z = st + (int64_t) RHG_CF_HEIGHT * Nmax;
  const double *T = st + (int64_t) RHG_CF_T * Nmax;
  const double *cc = cchi + (int64_t) ic * Nmax * nout + io;
  const double *ce = ceta + (int64_t) ic * Nmax * nout + io;
  const int has_line = M.lam_line_ptr[ls + 1] > M.lam_line_ptr[ls];
  int solver = M.solver[ls];
  if (solver == RHG_SOLVER_BEZIER && !moving) solver = RHG_SOLVER_FEAUTRIER;
  const double *lc = M.lamconst + (int64_t) ls * RHG_LC_NCONST;
  double B2[2];
  B2[0] = planck(T[N - 2], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
  B2[1] = planck(T[N - 1], lc[RHG_LC_HC_KLA], lc[RHG_LC_TWOHNU3_C2]);
  double *out = stokes + (int64_t) ic * 4 * nout;

#define POINT(kk, P) form_point(M, ls, has_line, st + (kk), Nmax, mu, moving, \
                                cc[(int64_t) (kk) * nout], ce[(int64_t) (kk) * nout], P)

  if (solver == RHG_SOLVER_STOKES) {
    // cubic DELO-Bezier, rh/rhf1d/bezier_1D.c:52-300 (to_obs: k from N-1 to 0)
    const double imu = 1.0 / mu;
    const double ident[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    Point Pu, P0, Pd, Pdd;
    double Ku[4][4], K0[4][4], Kd[4][4], dKu[4][4], dK0[4][4];
    double dSu[4], dS0[4], Iv[4], A[4][4], Ma[4][4], Mb[4][4], Mc[4][4], Md[16], V0[4];
    POINT(N - 1, Pu);
    POINT(N - 2, P0);
    POINT(N - 3, Pd);
    double dtau_uw = 0.5 * imu * (Pu.chi + P0.chi) * fabs(z[N - 1] - z[N - 2]);
    Iv[0] = B2[1] - (B2[0] - B2[1]) / dtau_uw;
    Iv[1] = Iv[2] = Iv[3] = 0.0;
    int k = N - 2;
    double dsup = fabs(z[k] - z[k + 1]) * imu;
    double dsdn = fabs(z[k - 1] - z[k]) * imu;
    double dchi_up = (P0.chi - Pu.chi) / dsup;
    double dchi_c = cent_deriv(dsup, dsdn, Pu.chi, P0.chi, Pd.chi);
    double c2 = fmax(P0.chi - (dsup / 3.0) * dchi_c, 0.0);
    double c1 = fmax(Pu.chi + (dsup / 3.0) * dchi_up, 0.0);
    dtau_uw = 0.25 * dsup * (P0.chi + Pu.chi + c1 + c2);
    make_K(Pu, Ku);
    make_K(P0, K0);
    for (int n = 0; n < 4; n++) {
      dSu[n] = (P0.S[n] - Pu.S[n]) / dtau_uw;
      for (int m = 0; m < 4; m++) dKu[n][m] = (K0[n][m] - Ku[n][m]) / dtau_uw;
    }
    for (k = N - 2; k != 0; k--) {
      double dchi_dn;
      dsdn = fabs(z[k - 1] - z[k]) * imu;
      if (k > 1) {
        POINT(k - 2, Pdd);
        double dsdn2 = fabs(z[k - 2] - z[k - 1]) * imu;
        dchi_dn = cent_deriv(dsdn, dsdn2, P0.chi, Pd.chi, Pdd.chi);
      } else {
        dchi_dn = (Pd.chi - P0.chi) / dsdn;
      }
      c2 = fmax(P0.chi + (dsdn / 3.0) * dchi_c, 0.0);
      c1 = fmax(Pd.chi - (dsdn / 3.0) * dchi_dn, 0.0);
      double dtau_dw = 0.25 * dsdn * (P0.chi + Pd.chi + c1 + c2);
      double dt = dtau_uw, dt03 = dt / 3.0, alpha, beta, gamma, theta, eps;
      bezier3_coeffs(dt, alpha, beta, gamma, theta, eps);
      make_K(Pd, Kd);
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) dK0[j][i] = cent_deriv(dtau_uw, dtau_dw, Ku[j][i], K0[j][i], Kd[j][i]);
#pragma unroll
      for (int i = 0; i < 4; i++) dS0[i] = cent_deriv(dtau_uw, dtau_dw, Pu.S[i], P0.S[i], Pd.S[i]);
      // Ma = Ku.Ku, A = K0.K0 (rh/bezier_aux.c:94-103 summation order)
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          double sa = 0.0, sb = 0.0;
#pragma unroll
          for (int kk = 0; kk < 4; kk++) {
            sa += Ku[kk][i] * Ku[j][kk];
            sb += K0[kk][i] * K0[j][kk];
          }
          Ma[j][i] = sa;
          A[j][i] = sb;
        }
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          Md[j * 4 + i] = ident[j][i] + alpha * K0[j][i] - gamma *
                          -(dt03 * (A[j][i] + dK0[j][i] + K0[j][i]) + K0[j][i]);
          Ma[j][i] = eps * ident[j][i] - beta * Ku[j][i] + theta *
                     (dt03 * (Ma[j][i] + dKu[j][i] + Ku[j][i]) - Ku[j][i]);
          Mb[j][i] = beta * ident[j][i] + theta * (ident[j][i] - dt03 * Ku[j][i]);
          Mc[j][i] = alpha * ident[j][i] + gamma * (ident[j][i] + dt03 * K0[j][i]);
        }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < 4; j++) v += Ma[i][j] * Iv[j] + Mb[i][j] * Pu.S[j] + Mc[i][j] * P0.S[j];
        v += dt03 * (gamma * dS0[i] - theta * dSu[i]);
        V0[i] = v;
      }
      matinv4(Md);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        double v = 0.0;
#pragma unroll
        for (int kk = 0; kk < 4; kk++) v += Md[i * 4 + kk] * V0[kk];
        Iv[i] = v;
      }
      Pu = P0; P0 = Pd; if (k > 1) Pd = Pdd;
#pragma unroll
      for (int n = 0; n < 4; n++) dSu[n] = dS0[n];
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) { Ku[j][i] = K0[j][i]; K0[j][i] = Kd[j][i]; dKu[j][i] = dK0[j][i]; }
      dtau_uw = dtau_dw;
      dsup = dsdn;
      dchi_up = dchi_c;
      dchi_c = dchi_dn;
    }
    // last interval: linear, rh/rhf1d/bezier_1D.c:264-299 (P0 = depth 0)
    {
      double w[3];
      dtau_uw = 0.5 * imu * (P0.chi + Pu.chi) * fabs(z[0] - z[1]);
      w3(dtau_uw, w);
      for (int n = 0; n < 4; n++) V0[n] = w[0] * P0.S[n] + w[1] * -dSu[n];
      for (int n = 0; n < 4; n++) {
        for (int m = 0; m < 4; m++) {
          A[n][m] = -w[1] / dtau_uw * Ku[n][m];
          Md[n * 4 + m] = (w[0] - w[1] / dtau_uw) * K0[n][m];
        }
        A[n][n] = 1.0 - w[0];
        Md[n * 4 + n] = 1.0;
      }
      for (int n = 0; n < 4; n++)
        for (int m = 0; m < 4; m++) V0[n] += A[n][m] * Iv[m];
      matinv4(Md);
      for (int i = 0; i < 4; i++) {
        double v = 0.0;
        for (int kk = 0; kk < 4; kk++) v += Md[i * 4 + kk] * V0[kk];
        out[i * nout + io] = v;
      }
    }
  } else if (solver == RHG_SOLVER_BEZIER) {
    // unpolarized cubic Bezier, rh/rhf1d/bezier_1D.c:306-541
    const double zmu = 1.0 / mu;
    Point Pu, P0, Pd, Pdd;
    POINT(N - 1, Pu);
    POINT(N - 2, P0);
    POINT(N - 3, Pd);
    double dtau_uw = 0.5 * zmu * (Pu.chi + P0.chi) * fabs(z[N - 1] - z[N - 2]);
    double I_upw = B2[1] - (B2[0] - B2[1]) / dtau_uw;
    int k = N - 2;
    double dsup = fabs(z[k] - z[k + 1]) * zmu;
    double dsdn = fabs(z[k - 1] - z[k]) * zmu;
    double dchi_up = (P0.chi - Pu.chi) / dsup;
    double dchi_c = cent_deriv(dsup, dsdn, Pu.chi, P0.chi, Pd.chi);
    double c1 = fmax(P0.chi - (dsup / 3.0) * dchi_c, 0.0);
    double c2 = fmax(Pu.chi + (dsup / 3.0) * dchi_up, 0.0);
    dtau_uw = dsup * (P0.chi + Pu.chi + c1 + c2) * 0.25;
    double dS_up = (P0.S[0] - Pu.S[0]) / dtau_uw;
    double I = I_upw;
    for (k = N - 2; k != 0; k--) {
      double dchi_dn;
      dsdn = fabs(z[k - 1] - z[k]) * zmu;
      if (k > 1) {
        POINT(k - 2, Pdd);
        double dsdn2 = fabs(z[k - 2] - z[k - 1]) * zmu;
        dchi_dn = cent_deriv(dsdn, dsdn2, P0.chi, Pd.chi, Pdd.chi);
      } else dchi_dn = (Pd.chi - P0.chi) / dsdn;
      c1 = fmax(P0.chi + (dsdn / 3.0) * dchi_c, 0.0);
      c2 = fmax(Pd.chi - (dsdn / 3.0) * dchi_dn, 0.0);
      double dtau_dw = dsdn * (P0.chi + Pd.chi + c1 + c2) * 0.25;
      double dt03 = dtau_uw / 3.0, alpha, beta, gamma, theta, eps;
      bezier3_coeffs(dtau_uw, alpha, beta, gamma, theta, eps);
      double dS_c = cent_deriv(dtau_uw, dtau_dw, Pu.S[0], P0.S[0], Pd.S[0]);
      c1 = fmax(P0.S[0] - dt03 * dS_c, 0.0);
      c2 = fmax(Pu.S[0] + dt03 * dS_up, 0.0);
      I = I_upw * eps + alpha * P0.S[0] + beta * Pu.S[0] + gamma * c1 + theta * c2;
      I_upw = I;
      Pu = P0; P0 = Pd; if (k > 1) Pd = Pdd;
      dsup = dsdn;
      dchi_up = dchi_c;
      dchi_c = dchi_dn;
      dtau_uw = dtau_dw;
      dS_up = dS_c;
    }
    {
      double w[3];
      dtau_uw = 0.5 * zmu * (P0.chi + Pu.chi) * fabs(z[0] - z[1]);
      double dS_uw = -(P0.S[0] - Pu.S[0]) / dtau_uw;
      w3(dtau_uw, w);
      I = (1.0 - w[0]) * I_upw + w[0] * P0.S[0] + w[1] * dS_uw;
    }
    out[io] = I;
    out[nout + io] = 0.0;
    out[2 * nout + io] = 0.0;
    out[3 * nout + io] = 0.0;
  } else {
    // Feautrier, rh/rhf1d/feautrier.c:56-202 (STANDARD order, zero
    // incident intensity at the top, thermalised bottom).  The back
    // substitution for P[0] is accumulated in the forward sweep:
    // P[0] = sum_k ztmp[k] prod_{j<k} 1/(1+F[j]) + P[N-1] prod_{j<N-1} 1/(1+F[j]).
    const double zmu = 0.5 / mu;
    Point Pk, Pn;
    POINT(0, Pk);
    POINT(1, Pn);
    double dtau_prev = zmu * (Pk.chi + Pn.chi) * (z[0] - z[1]);
    double abc0 = 1.0 + 2.0 * 1.0 / dtau_prev;
    double C10 = 2.0 / (dtau_prev * dtau_prev);
    double Stmp = Pk.S[0] + 2.0 * 0.0 / ((1.0 + 0.0) * dtau_prev);
    double F = abc0 / C10;
    double ztmp = Stmp / (abc0 + C10);
    double P0acc = ztmp, prod = 1.0 / (1.0 + F);
    for (int k = 1; k < N - 1; k++) {
      Pk = Pn;
      POINT(k + 1, Pn);
      double dtau_k = zmu * (Pk.chi + Pn.chi) * (z[k] - z[k + 1]);
      double dtau_mid = 0.5 * (dtau_k + dtau_prev);
      double A1 = 1.0 / (dtau_mid * dtau_prev);
      double C1 = 1.0 / (dtau_mid * dtau_k);
      double Fn = (1.0 + A1 * F / (1.0 + F)) / C1;
      ztmp = (Pk.S[0] + A1 * ztmp) / (C1 * (1.0 + Fn));
      F = Fn;
      P0acc += ztmp * prod;
      prod *= 1.0 / (1.0 + F);
      dtau_prev = dtau_k;
    }
    double hN = B2[1] - (B2[0] - B2[1]) / dtau_prev;
    double abcN = 1.0 + 2.0 * 1.0 / dtau_prev;
    double A1N = 2.0 / (dtau_prev * dtau_prev);
    double StmpN = Pn.S[0] + 2.0 * hN / ((1.0 + 0.0) * dtau_prev);
    double PN = (StmpN + A1N * ztmp) / (abcN + A1N * (F / (1.0 + F)));
    double P0 = P0acc + PN * prod;
    out[io] = (1.0 + 1.0) * P0 - 0.0 / (1.0 + 0.0);
    out[nout + io] = 0.0;
    out[2 * nout + io] = 0.0;
    out[3 * nout + io] = 0.0;
  }
#undef POINT
}

// Work census: one thread per (column, output wavelength), counts per ray in
// census[(ic * nout + io) * RHG_NWORK + slot].
__global__ void __launch_bounds__(64)
k_census(DevModel M, const int32_t *__restrict__ ndep, const int32_t *__restrict__ moving_arr,
         int ncol, int Nmax, const double *__restrict__ state, int nfield,
         const int32_t *__restrict__ out_ls, int nout, double mu, int32_t *__restrict__ census) {
  const int io = blockIdx.x * blockDim.x + threadIdx.x;
  const int ic = blockIdx.y;
  if (io >= nout || ic >= ncol) return;
  const int N = ndep[ic], ls = out_ls[io], moving = moving_arr[ic];
  const double *st = state + (int64_t) ic * nfield * Nmax;
  int32_t cnt[RHG_NWORK];
  for (int i = 0; i < RHG_NWORK; i++) cnt[i] = 0;
  int solver = M.solver[ls];
  if (solver == RHG_SOLVER_BEZIER && !moving) solver = RHG_SOLVER_FEAUTRIER;
  const int slot = solver == RHG_SOLVER_STOKES ? RHG_WORK_STOKES
                 : solver == RHG_SOLVER_BEZIER ? RHG_WORK_BEZIER : RHG_WORK_FEAUTRIER;
  const int has_line = M.lam_line_ptr[ls + 1] > M.lam_line_ptr[ls];
  for (int k = 1; k < N; k++) {
    cnt[slot]++;
    if (has_line) line_census(M, ls, st + k, Nmax, mu, moving, cnt);
  }
  int32_t *o = census + ((int64_t) ic * nout + io) * RHG_NWORK;
  for (int i = 0; i < RHG_NWORK; i++) o[i] = cnt[i];
}

// Per-column sums of the census: thread per column.
__global__ void k_census_sum(const int32_t *__restrict__ census, int ncol, int nout,
                             int64_t *__restrict__ col_sum) {
  const int ic = blockIdx.x * blockDim.x + threadIdx.x;
  if (ic >= ncol) return;
  int64_t acc[RHG_NWORK];
  for (int i = 0; i < RHG_NWORK; i++) acc[i] = 0;
  for (int io = 0; io < nout; io++)
    for (int i = 0; i < RHG_NWORK; i++) acc[i] += census[((int64_t) ic * nout + io) * RHG_NWORK + i];
  for (int i = 0; i < RHG_NWORK; i++) col_sum[(int64_t) ic * RHG_NWORK + i] = acc[i];
}

}  // namespace rhg
