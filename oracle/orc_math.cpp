// TEST INFRASTRUCTURE ONLY -- CPU oracle, expansion algebra and quadrature.
// See orc_math.h for the reference lines each function restates.
#include "orc_math.h"

#include <algorithm>

namespace orc {

const ATable &Atab() {
  static ATable t;
  return t;
}

static inline int isgn(int x) { return x < 0 ? -1 : (x > 0 ? 1 : 0); }

// multipole_expansion.cc:93-122
void me_add_point(Exp &M, double q, const double y[3]) {
  if (q == 0) return;
  M.is_zero = false;
  double d[3] = {y[0] - M.c[0], y[1] - M.c[1], y[2] - M.c[2]};
  double rho, ca, beta;
  to_sph(d, rho, ca, beta);
  LegTable L;
  L.fill(M.p, ca, false);
  for (int n = 0; n <= M.p; ++n)
    for (int m = 0; m <= n; ++m) {
      double rf = L.p(n, m) * std::pow(rho, double(n)) * q;
      M.a[tri(n, m)] += std::exp(cd(0.0, -m * beta)) * rf;
    }
}

// multipole_expansion.cc:124-181 (chain rule through rho, alpha, beta)
void me_add_normder(Exp &M, double q, const double y[3], const double nrm[3]) {
  if (q == 0) return;
  M.is_zero = false;
  double d[3] = {y[0] - M.c[0], y[1] - M.c[1], y[2] - M.c[2]};
  double nl = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  double nv[3] = {nrm[0] / nl, nrm[1] / nl, nrm[2] / nl};
  double rho = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double drho = (d[0] / rho) * nv[0] + (d[1] / rho) * nv[1] + (d[2] / rho) * nv[2];
  double beta = std::atan2(d[1], d[0]);
  double rxy2 = d[0] * d[0] + d[1] * d[1];
  double dbeta = (-d[1] / rxy2) * nv[0] + (d[0] / rxy2) * nv[1];
  double ca = d[2] / rho;
  double sa = std::sqrt(1.0 - ca * ca);
  double dalpha = (ca * std::cos(beta) / rho) * nv[0] + (ca * std::sin(beta) / rho) * nv[1] +
                  (-sa / rho) * nv[2];
  LegTable L;
  L.fill(M.p, ca, true);
  for (int n = 0; n <= M.p; ++n)
    for (int m = 0; m <= n; ++m) {
      double P = L.p(n, m);
      double dPs = L.dp(n, m) * sa;
      double rn = std::pow(rho, double(n));
      cd z = std::exp(cd(0.0, -double(m) * beta));
      z *= cd(double(n) * std::pow(rho, double(n) - 1.0) * P * drho - rn * dPs * dalpha,
              -double(m) * rn * P * dbeta);
      z *= q;
      M.a[tri(n, m)] += z;
    }
}

// multipole_expansion.cc:74-91
void me_add_scaled(Exp &M, const Exp &src, double s) {
  if (src.is_zero || s == 0) return;
  M.is_zero = false;
  for (int i = 0; i < M.kt(); ++i) M.a[i] += src.a[i] * s;
}

// multipole_expansion.cc:184-261
void me_add_child(Exp &M, const Exp &o) {
  if (o.is_zero) return;
  M.is_zero = false;
  const ATable &A = Atab();
  double d[3] = {o.c[0] - M.c[0], o.c[1] - M.c[1], o.c[2] - M.c[2]};
  double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (!(dist > 1e-7)) {
    for (int i = 0; i < M.kt(); ++i) M.a[i] += o.a[i];
    return;
  }
  double rho, ca, beta;
  to_sph(d, rho, ca, beta);
  LegTable L;
  L.fill(M.p, ca, false);
  for (int n = 0; n <= M.p; ++n)
    for (int m = 0; m <= n; ++m) {
      cd z(0, 0);
      for (int nn = 0; nn <= n; ++nn)
        for (int mm = -nn; mm <= nn; ++mm) {
          if (std::abs(m - mm) > n - nn) continue;
          const cd &src = o.a[tri(n - nn, std::abs(m - mm))];
          cd a(src.real(), isgn(m - mm) * src.imag());
          double rf = L.p(nn, std::abs(mm)) * std::pow(rho, double(nn)) * A(nn, std::abs(mm)) *
                      A(n - nn, std::abs(m - mm)) / A(n, m);
          rf *= ipow_re(std::abs(m) - std::abs(mm) - std::abs(m - mm));
          z += rf * (a * std::exp(cd(0.0, -mm * beta)));
        }
      M.a[tri(n, m)] += z;
    }
}

// multipole_expansion.cc:264-298
double me_evaluate(const Exp &M, const double x[3]) {
  if (M.is_zero) return 0.0;
  double d[3] = {x[0] - M.c[0], x[1] - M.c[1], x[2] - M.c[2]};
  double rho, ca, beta;
  to_sph(d, rho, ca, beta);
  LegTable L;
  L.fill(M.p, ca, false);
  cd f(0, 0);
  for (int n = 0; n <= M.p; ++n) {
    double rf = L.p(n, 0) * std::pow(rho, double(-n - 1));
    f += M.a[tri(n, 0)] * rf;
    for (int m = 1; m <= n; ++m) {
      double r2 = L.p(n, m) * std::pow(rho, double(-n - 1));
      f += M.a[tri(n, m)] * std::exp(cd(0.0, m * beta)) * 2.0 * r2;
    }
  }
  return f.real();
}

// local_expansion.h:145-180
double m2l_coeff(int n, int m, int nn, int mm) {
  const ATable &A = Atab();
  double rf = A(nn, std::abs(mm)) / A(n + nn, std::abs(m - mm)) * A(n, std::abs(m));
  rf *= ipow_re(std::abs(m - mm) - std::abs(m) - std::abs(mm)) / ((nn % 2) ? -1.0 : 1.0);
  return rf;
}

// local_expansion.h:182-215
double l2l_coeff(int n, int m, int nn, int mm) {
  const ATable &A = Atab();
  double rf = A(nn - n, std::abs(mm - m)) / A(nn, std::abs(mm)) * A(n, std::abs(m));
  rf *= ipow_re(std::abs(mm) - std::abs(mm - m) - std::abs(m)) * (((nn + n) % 2) ? -1.0 : 1.0);
  return rf;
}

// local_expansion.cc:179-248
void le_add_multipole(Exp &L, const Exp &M) {
  if (M.is_zero) return;
  double d[3] = {M.c[0] - L.c[0], M.c[1] - L.c[1], M.c[2] - L.c[2]};
  double rho, ca, beta;
  to_sph(d, rho, ca, beta);
  LegTable T;
  T.fill(2 * L.p, ca, false);
  for (int n = 0; n <= L.p; ++n)
    for (int m = 0; m <= n; ++m) {
      cd z(0, 0);
      for (int nn = 0; nn <= L.p; ++nn) {
        double rhoFact = std::pow(rho, double(-n - nn - 1));
        for (int mm = -nn; mm < 0; ++mm) {
          cd a = std::conj(M.a[tri(nn, -mm)]);
          double rf = T.p(nn + n, std::abs(mm - m)) * rhoFact * m2l_coeff(n, m, nn, mm);
          z += a * cd(std::cos((mm - m) * beta), std::sin((mm - m) * beta)) * rf;
        }
        for (int mm = 0; mm <= nn; ++mm) {
          cd a = M.a[tri(nn, mm)];
          double rf = T.p(nn + n, std::abs(mm - m)) * rhoFact * m2l_coeff(n, m, nn, mm);
          z += a * cd(std::cos((mm - m) * beta), std::sin((mm - m) * beta)) * rf;
        }
      }
      L.a[tri(n, m)] += z;
    }
  L.is_zero = false;
}

// local_expansion.cc:107-176
void le_add_parent(Exp &L, const Exp &o) {
  if (o.is_zero) return;
  double d[3] = {o.c[0] - L.c[0], o.c[1] - L.c[1], o.c[2] - L.c[2]};
  double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (dist > 1e-7) {
    double rho, ca, beta;
    to_sph(d, rho, ca, beta);
    LegTable T;
    T.fill(L.p, ca, false);
    for (int n = 0; n <= L.p; ++n)
      for (int m = 0; m <= n; ++m) {
        cd z(0, 0);
        for (int nn = n; nn <= L.p; ++nn) {
          double rhoFact = std::pow(rho, double(nn - n));
          for (int mm = -nn; mm <= nn; ++mm) {
            if (std::abs(mm - m) > nn - n) continue;
            const cd &src = o.a[tri(nn, std::abs(mm))];
            cd a(src.real(), isgn(mm) * src.imag());
            double rf = T.p(nn - n, std::abs(mm - m)) * rhoFact * l2l_coeff(n, m, nn, mm);
            z += a * cd(std::cos((mm - m) * beta), std::sin((mm - m) * beta)) * rf;
          }
        }
        L.a[tri(n, m)] += z;
      }
  } else {
    for (int i = 0; i < L.kt(); ++i) L.a[i] += o.a[i];
  }
  L.is_zero = false;
}

// local_expansion.cc:251-287
double le_evaluate(const Exp &L, const double x[3]) {
  if (L.is_zero) return 0.0;
  double d[3] = {x[0] - L.c[0], x[1] - L.c[1], x[2] - L.c[2]};
  double rho, ca, beta;
  to_sph(d, rho, ca, beta);
  LegTable T;
  T.fill(L.p, ca, false);
  cd f(0, 0);
  for (int n = 0; n <= L.p; ++n) {
    double rf = T.p(n, 0) * std::pow(rho, double(n));
    f += L.a[tri(n, 0)] * rf;
    for (int m = 1; m <= n; ++m) {
      double r2 = T.p(n, m) * std::pow(rho, double(n));
      cd cf = cd(std::cos(m * beta), std::sin(m * beta)) * 2.0 * r2;
      f += L.a[tri(n, m)] * cf;
    }
  }
  return f.real();
}

// ---------------------------------------------------------------------------
// Quadrature
// ---------------------------------------------------------------------------
void gauss01(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    // Newton iteration on P_n from the Chebyshev-like initial guess.
    long double z = std::cos(PI * (i + 0.75) / (n + 0.5)), dp = 0;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1, p1 = 0;
      for (int k = 1; k <= n; ++k) {
        long double p2 = p1;
        p1 = p0;
        p0 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p2) / k;
      }
      dp = n * (z * p0 - p1) / (z * z - 1.0L);
      long double dz = p0 / dp;
      z -= dz;
      if (std::fabs((double)dz) < 1e-19) break;
    }
    long double wi = 2.0L / ((1.0L - z * z) * dp * dp);
    // ascending order on [0,1]
    x[n - 1 - i] = (double)((1.0L + z) / 2.0L);
    w[n - 1 - i] = (double)(wi / 2.0L);
  }
}

void telles01(int n, double x0, std::vector<double> &x, std::vector<double> &w) {
  std::vector<double> gx, gw;
  gauss01(n, gx, gw);
  const double tol = 1e-10;
  const double eb = 2.0 * x0 - 1.0;  // singular point on [-1,1]
  const double es = eb * eb - 1.0;
  const double gb = std::cbrt(eb * es + std::fabs(es)) + std::cbrt(eb * es - std::fabs(es)) + eb;
  x.clear();
  w.clear();
  for (int i = 0; i < n; ++i) {
    if (std::fabs(gx[i] - x0) <= tol) continue;
    const double g = 2.0 * gx[i] - 1.0;
    const double den = 1.0 + 3.0 * gb * gb;
    const double eta = (std::pow(g - gb, 3.0) + gb * (gb * gb + 3.0)) / den;
    const double J = 3.0 * (g - gb) * (g - gb) / den;
    x.push_back((eta + 1.0) / 2.0);
    w.push_back(gw[i] * J);
  }
}

void tensor2(const std::vector<double> &x1, const std::vector<double> &w1,
             const std::vector<double> &x2, const std::vector<double> &w2,
             std::vector<double> &pts, std::vector<double> &wts) {
  pts.clear();
  wts.clear();
  for (size_t j = 0; j < x2.size(); ++j)
    for (size_t i = 0; i < x1.size(); ++i) {
      pts.push_back(x1[i]);
      pts.push_back(x2[j]);
      wts.push_back(w1[i] * w2[j]);
    }
}

// QSplit<2>(QDuffy(n, 1), s) (bem_problem.cc:213-229), restated from
// deal.II's construction (Duffy 1982): for each face f of the unit square in
// deal.II order (x=0, x=1, y=0, y=1) the triangle (s, v0, v1) of its two
// vertices, skipped when it has no area (s on the face); on it the Duffy
// rule of the n x n Gauss product (radial q0 fastest): reference point
// (q0 (1 - q1), q0 q1), weight q0 w0 w1, mapped affinely by
// x = s + (v0 - s) xk + (v1 - s) yk with weight |det| times it.
void duffy_split(int n, double sx, double sy, std::vector<double> &pts, std::vector<double> &wts) {
  std::vector<double> gx, gw;
  gauss01(n, gx, gw);
  static const double face[4][2][2] = {{{0, 0}, {0, 1}}, {{1, 0}, {1, 1}}, {{0, 0}, {1, 0}}, {{0, 1}, {1, 1}}};
  pts.clear();
  wts.clear();
  for (const auto &f : face) {
    const double a0 = f[0][0] - sx, a1 = f[0][1] - sy, b0 = f[1][0] - sx, b1 = f[1][1] - sy;
    const double J = std::fabs(a0 * b1 - b0 * a1);
    if (J < 1e-12) continue;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double q0 = gx[i], q1 = gx[j];
        const double xk = q0 * (1.0 - q1), yk = q0 * q1;
        const double wd = q0 * (gw[i] * gw[j]);
        pts.push_back(sx + (a0 * xk + b0 * yk));
        pts.push_back(sy + (a1 * xk + b1 * yk));
        wts.push_back(J * wd);
      }
  }
}

// ---------------------------------------------------------------------------
// FE_Q on the unit square
// ---------------------------------------------------------------------------
static const int q2_map[9][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}, {0, 2},
                                 {1, 2}, {2, 0}, {2, 1}, {2, 2}};

FEQ::FEQ(int dpc_) : dpc(dpc_) {
  deg = (dpc == 4) ? 1 : 2;
  const double c[3] = {0.0, 1.0, 0.5};
  for (int j = 0; j < dpc; ++j) {
    usp[j][0] = c[q2_map[j][0]];
    usp[j][1] = c[q2_map[j][1]];
  }
}

static inline void lag1d(int deg, double t, double *L, double *dL) {
  if (deg == 1) {
    L[0] = 1.0 - t; L[1] = t; L[2] = 0.0;
    dL[0] = -1.0; dL[1] = 1.0; dL[2] = 0.0;
  } else {
    L[0] = (1.0 - t) * (1.0 - 2.0 * t);
    L[1] = t * (2.0 * t - 1.0);
    L[2] = 4.0 * t * (1.0 - t);
    dL[0] = 4.0 * t - 3.0;
    dL[1] = 4.0 * t - 1.0;
    dL[2] = 4.0 - 8.0 * t;
  }
}

void FEQ::shape(double xi, double eta, double *N) const {
  double Lx[3], Ly[3], dLx[3], dLy[3];
  lag1d(deg, xi, Lx, dLx);
  lag1d(deg, eta, Ly, dLy);
  for (int j = 0; j < dpc; ++j) N[j] = Lx[q2_map[j][0]] * Ly[q2_map[j][1]];
}

void FEQ::grad(double xi, double eta, double *dNx, double *dNy) const {
  double Lx[3], Ly[3], dLx[3], dLy[3];
  lag1d(deg, xi, Lx, dLx);
  lag1d(deg, eta, Ly, dLy);
  for (int j = 0; j < dpc; ++j) {
    dNx[j] = dLx[q2_map[j][0]] * Ly[q2_map[j][1]];
    dNy[j] = Lx[q2_map[j][0]] * dLy[q2_map[j][1]];
  }
}

}  // namespace orc
