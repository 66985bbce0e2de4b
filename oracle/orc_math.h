// TEST INFRASTRUCTURE ONLY -- part of the CPU oracle.
//
// This header restates, in plain FP64 C++, the expansion algebra of pi-BEM's
// fast multipole method.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load the oracle, and only as a checker.  The product
// (pi-bem_amd/) never includes or links anything from oracle/.
//
// Parity status of each function is stated next to it.  "pinned" names the
// check in tests/ that would fail on a dropped term, a wrong sign or index.
#pragma once
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

namespace orc {

using cd = std::complex<double>;
constexpr double PI = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Normalised associated Legendre functions
//   Pt(n,m,x) = sqrt((n-m)!/(n+m)!) P_n^m(x),  Condon-Shortley phase included,
// and d/dx Pt(n,m,x).  Reference: source/ass_leg_function.cc:521-536 (lookup),
// closed forms :538-14587 (e.g. P_1_1 = -(sqrt2/2) sqrt(1-x^2) at :563-572).
// Restated with the three-term recurrence in n; pinned in
// tests/test_oracle_math.py against the reference's closed forms for low
// degrees and against scipy.special.lpmv for all n <= 20.
// ---------------------------------------------------------------------------
struct LegTable {
  int nmax = -1;
  std::vector<double> P;   // (nmax+1)^2, index n*(nmax+1)+m
  std::vector<double> dP;  // d/dx, same layout (only when requested)
  inline double p(int n, int m) const { return P[n * (nmax + 1) + m]; }
  inline double dp(int n, int m) const { return dP[n * (nmax + 1) + m]; }

  void fill(int nm, double x, bool deriv) {
    nmax = nm;
    const int ld = nmax + 1;
    P.assign(ld * ld, 0.0);
    const double s = std::sqrt(std::max(0.0, 1.0 - x * x));
    double pmm = 1.0;  // Pt(m,m)
    for (int m = 0; m <= nmax; ++m) {
      if (m > 0) pmm *= -std::sqrt((2.0 * m - 1.0) / (2.0 * m)) * s;
      P[m * ld + m] = pmm;
      if (m + 1 <= nmax) P[(m + 1) * ld + m] = x * std::sqrt(2.0 * m + 1.0) * pmm;
      for (int n = m + 2; n <= nmax; ++n) {
        const double a = (2.0 * n - 1.0) * x * P[(n - 1) * ld + m];
        const double b = std::sqrt(double((n - 1) * (n - 1) - m * m)) * P[(n - 2) * ld + m];
        P[n * ld + m] = (a - b) / std::sqrt(double(n * n - m * m));
      }
    }
    if (!deriv) return;
    // (x^2-1) dPt_n^m/dx = n x Pt_n^m - sqrt(n^2-m^2) Pt_{n-1}^m
    dP.assign(ld * ld, 0.0);
    const double xm1 = x * x - 1.0;
    for (int n = 0; n <= nmax; ++n)
      for (int m = 0; m <= n; ++m) {
        const double prev = (n - 1 >= m) ? P[(n - 1) * ld + m] : 0.0;
        dP[n * ld + m] = (n * x * P[n * ld + m] - std::sqrt(double(n * n - m * m)) * prev) / xm1;
      }
  }
};

// A_n^m = (-1)^n / sqrt((n-m)!(n+m)!), 0<=m<=n<=20.
// Reference: include/multipole_expansion.h:123-147 (A_n_m_Matrix(20)).
struct ATable {
  double a[21][21];
  ATable() {
    for (int n = 0; n <= 20; ++n)
      for (int m = 0; m <= 20; ++m) a[n][m] = 0.0;
    for (int n = 0; n <= 20; ++n)
      for (int m = 0; m <= n; ++m) {
        double f1 = 1.0, f2 = 1.0;
        for (int i = n - m; i > 0; --i) f1 *= i;
        for (int i = n + m; i > 0; --i) f2 *= i;
        a[n][m] = ((n % 2) ? -1.0 : 1.0) / std::sqrt(f1 * f2);
      }
  }
  double operator()(int n, int m) const { return a[n][m]; }
};
const ATable &Atab();

// Real part of i^k for the even integers k that occur in the translation
// coefficients (the reference evaluates pow(complex(0,1), k).real()).
inline double ipow_re(int k) {
  int r = ((k % 4) + 4) % 4;
  return r == 0 ? 1.0 : (r == 2 ? -1.0 : 0.0);
}

inline int tri(int n, int m) { return n * (n + 1) / 2 + m; }

// ---------------------------------------------------------------------------
// Multipole / local expansion in the reference storage: Kt=(p+1)(p+2)/2
// complex coefficients at index n(n+1)/2+m, m>=0 only.
// Reference: include/multipole_expansion.h:31-148, include/local_expansion.h.
// ---------------------------------------------------------------------------
struct Exp {
  int p = 0;
  double c[3] = {0, 0, 0};
  std::vector<cd> a;
  bool is_zero = true;
  Exp() = default;
  Exp(int p_, const double *cen) : p(p_), a((p_ + 1) * (p_ + 2) / 2, cd(0, 0)) {
    c[0] = cen[0]; c[1] = cen[1]; c[2] = cen[2];
  }
  int kt() const { return (p + 1) * (p + 2) / 2; }
};

inline void to_sph(const double d[3], double &rho, double &cosa, double &beta) {
  rho = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  cosa = d[2] / rho;
  beta = std::atan2(d[1], d[0]);
}

// P2M, monopole.  multipole_expansion.cc:93-122.  Pinned by the reference
// golden multipole_expansion_creation_and_evaluation.output.
void me_add_point(Exp &M, double q, const double y[3]);
// P2M, normal derivative (dipole) with respect to the source point.
// multipole_expansion.cc:124-181.  Pinned by the reference golden
// multipole_expansion_creation_and_evaluation_neumann_kernel.output and by
// finite differences of me_add_point.
void me_add_normder(Exp &M, double q, const double y[3], const double nrm[3]);
// this += s*M.  multipole_expansion.cc:74-91.
void me_add_scaled(Exp &M, const Exp &src, double s);
// M2M (child -> this, the parent).  multipole_expansion.cc:184-261.  Pinned by
// multipole_expansion_translation.output and by the exact-translation identity.
void me_add_child(Exp &M, const Exp &child);
// M2P.  multipole_expansion.cc:264-298.
double me_evaluate(const Exp &M, const double x[3]);
// M2L.  local_expansion.cc:179-248 with the table at local_expansion.h:145-180.
// Pinned by local_expansion_creation_and_evaluation.output.
void le_add_multipole(Exp &L, const Exp &M);
// L2L.  local_expansion.cc:107-176 with the table at local_expansion.h:182-215.
// Pinned by local_expansion_translation.output (L2L of a degree-p local
// expansion is exact, so the value must not move).
void le_add_parent(Exp &L, const Exp &parent);
// L2P.  local_expansion.cc:251-287.
double le_evaluate(const Exp &L, const double x[3]);

// M2L coefficient C(n,m,nn,mm) (local_expansion.h:145-180) and L2L
// coefficient (local_expansion.h:182-215).
double m2l_coeff(int n, int m, int nn, int mm);
double l2l_coeff(int n, int m, int nn, int mm);

// ---------------------------------------------------------------------------
// Quadrature on [0,1] (deal.II QGauss / QTelles / QDuffy+QSplit are not
// vendored in the reference: restated from the published methods).
// ---------------------------------------------------------------------------
// Gauss-Legendre n-point rule on [0,1] (Newton on P_n).  Pinned by exactness
// on polynomials of degree 2n-1.
void gauss01(int n, std::vector<double> &x, std::vector<double> &w);
// Telles (1987, IJNME 24:959) cubic transformation of the n-point Gauss rule
// towards the singular coordinate x0 in [0,1]; a base point coinciding with
// x0 (Jacobian 0) is dropped.  Pinned by the corner-singular closed form
// int_0^1 int_0^1 dx dy / sqrt(x^2+y^2) = 2 ln(1+sqrt2).
void telles01(int n, double x0, std::vector<double> &x, std::vector<double> &w);
// 2-D tensor rules on the unit square, x-coordinate fastest.
void tensor2(const std::vector<double> &x1, const std::vector<double> &w1,
             const std::vector<double> &x2, const std::vector<double> &w2,
             std::vector<double> &pts, std::vector<double> &wts);
// Duffy (1982) rule: the unit square is split into the triangles that share
// the singular point as a vertex, each integrated by an n x n Gauss rule
// collapsed onto that vertex.
void duffy_split(int n, double sx, double sy, std::vector<double> &pts, std::vector<double> &wts);

// ---------------------------------------------------------------------------
// Lagrange Q1 / Q2 elements on the unit square in deal.II's FE_Q numbering
// (vertices, then lines x=0, x=1, y=0, y=1, then the interior).
// ---------------------------------------------------------------------------
struct FEQ {
  int deg = 1, dpc = 4;
  double usp[9][2];
  explicit FEQ(int dpc_);
  void shape(double xi, double eta, double *N) const;
  void grad(double xi, double eta, double *dNx, double *dNy) const;
};

}  // namespace orc
