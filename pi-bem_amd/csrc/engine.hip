// HIP kernels of the BEM-FMM path for gfx950 (CDNA4) and their host-side
// orchestration.  One wavefront is 64 lanes; every launch below sizes its grid
// from the plan arrays built in setup.cpp and checks its indices against them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>

#include "engine.h"

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      throw pbfmm::HipError(std::string(#x) + ": " + hipGetErrorString(e_));             \
  } while (0)

namespace pbfmm {

// ---------------------------------------------------------------------------
// Legendre recurrence tables (degree <= 20):
//   Pt(m,m) = pmm(m) (-s)^m,  pmm(m) = prod_k sqrt((2k-1)/(2k))
//   Pt(n,m) = ra(n,m) x Pt(n-1,m) - rb(n,m) Pt(n-2,m)
//   ra = (2n-1)/sqrt(n^2-m^2), rb = sqrt((n-1)^2-m^2)/sqrt(n^2-m^2), rs = sqrt(n^2-m^2)
// ---------------------------------------------------------------------------
constexpr int kLd = 21;
__constant__ double c_ra_d[kLd * kLd], c_rb_d[kLd * kLd], c_rs_d[kLd * kLd], c_pmm_d[kLd], c_pmf_d[kLd];
__constant__ float c_ra_f[kLd * kLd], c_rb_f[kLd * kLd], c_rs_f[kLd * kLd], c_pmm_f[kLd], c_pmf_f[kLd];

template <typename T> __device__ __forceinline__ T RA(int n, int m);
template <> __device__ __forceinline__ double RA<double>(int n, int m) { return c_ra_d[n * kLd + m]; }
template <> __device__ __forceinline__ float RA<float>(int n, int m) { return c_ra_f[n * kLd + m]; }
template <typename T> __device__ __forceinline__ T RB(int n, int m);
template <> __device__ __forceinline__ double RB<double>(int n, int m) { return c_rb_d[n * kLd + m]; }
template <> __device__ __forceinline__ float RB<float>(int n, int m) { return c_rb_f[n * kLd + m]; }
template <typename T> __device__ __forceinline__ T RS(int n, int m);
template <> __device__ __forceinline__ double RS<double>(int n, int m) { return c_rs_d[n * kLd + m]; }
template <> __device__ __forceinline__ float RS<float>(int n, int m) { return c_rs_f[n * kLd + m]; }
template <typename T> __device__ __forceinline__ T PMM(int m);
template <> __device__ __forceinline__ double PMM<double>(int m) { return c_pmm_d[m]; }
template <> __device__ __forceinline__ float PMM<float>(int m) { return c_pmm_f[m]; }
// single-step factor sqrt((2m-1)/(2m))
template <typename T> __device__ __forceinline__ T PMF(int m);
template <> __device__ __forceinline__ double PMF<double>(int m) { return c_pmf_d[m]; }
template <> __device__ __forceinline__ float PMF<float>(int m) { return c_pmf_f[m]; }

__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float dsqrt(float a) { return sqrtf(a); }

static void upload_legendre_tables() {
  double ra[kLd * kLd] = {0}, rb[kLd * kLd] = {0}, rs[kLd * kLd] = {0}, pmm[kLd], pmf[kLd];
  float raf[kLd * kLd], rbf[kLd * kLd], rsf[kLd * kLd], pmmf[kLd], pmff[kLd];
  for (int n = 0; n < kLd; ++n)
    for (int m = 0; m <= n; ++m) {
      const double q = std::sqrt(double(n * n - m * m));
      rs[n * kLd + m] = q;
      if (n > m) {
        ra[n * kLd + m] = (2.0 * n - 1.0) / q;
        rb[n * kLd + m] = std::sqrt(double((n - 1) * (n - 1) - m * m)) / q;
      }
    }
  pmm[0] = 1.0;
  pmf[0] = 1.0;
  for (int m = 1; m < kLd; ++m) {
    pmf[m] = std::sqrt((2.0 * m - 1.0) / (2.0 * m));
    pmm[m] = pmm[m - 1] * pmf[m];
  }
  for (int i = 0; i < kLd * kLd; ++i) {
    raf[i] = (float)ra[i];
    rbf[i] = (float)rb[i];
    rsf[i] = (float)rs[i];
  }
  for (int i = 0; i < kLd; ++i) {
    pmmf[i] = (float)pmm[i];
    pmff[i] = (float)pmf[i];
  }
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_ra_d), ra, sizeof(ra)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_rb_d), rb, sizeof(rb)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_rs_d), rs, sizeof(rs)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_pmm_d), pmm, sizeof(pmm)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_pmf_d), pmf, sizeof(pmf)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_ra_f), raf, sizeof(raf)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_rb_f), rbf, sizeof(rbf)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_rs_f), rsf, sizeof(rsf)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_pmm_f), pmmf, sizeof(pmmf)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_pmf_f), pmff, sizeof(pmff)));
}

// Wavefront sums through DPP lane moves (VALU operand modifiers, no LDS
// crossbar round trips): quad_perm [1,0,3,2] and [2,3,0,1] sum each quad,
// row_half_mirror pairs lane i with 7-i and row_mirror lane i with 15-i, so
// after four adds every lane holds its 16-lane row's sum; the four row sums
// are read into scalars and added.  Every lane of the wavefront must be
// active; the sum is returned in every lane.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float read_lane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  v += dpp_move<0x140>(v);  // row_mirror
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// ---------------------------------------------------------------------------
// Near-field assembly (BEMFMA::direct_integrals, bem_fma.cc:670-913 regular
// and singular, 997-1126 bigger blocks).  An item is a (row, entry) pair: the
// regular rule over the entry's quad points, or the singular rule (Telles or
// Duffy split, chosen at set-up) over the whole cell when the cell holds a
// double of the row's node, both in FP64.  One workgroup per row; G lanes
// (G >= quad points per cell) share an item, a lane per quad point (or every
// G-th singular point), so that a load of the cell's quad data is one
// contiguous run.  The group sums the (N, D) parts of the cell's shape
// functions and adds them, a lane per value, into the row's (N, D) pairs in
// LDS (columns in tiles of kAsmTile; a row with more columns computes its
// items once per tile), and the row is stored once, whole.  The FP32
// operator is this block rounded once (k_round_f32), since in FP32 the
// difference R = y - x loses ~|x|/|R| ulps and R.n cancels again, which at a
// radius-4 surface costs 1e-4 of the double-layer product.
// ---------------------------------------------------------------------------
// Contraction is off in this kernel: near a singular vertex R = X - x_i is
// ~1e-5 of |X| and R.n cancels to O(r^2), so the singular double-layer
// entries follow the last bit of X; evaluating X and R exactly as the
// reference's FEValues path does (separately rounded products and sums) keeps
// FP64 entries reproducible to rounding.
template <int G>
__device__ __forceinline__ double group_sum(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  v += dpp_move<0x140>(v);  // row_mirror: every lane holds its 16-lane row's sum
  if constexpr (G >= 32) v += __shfl_xor(v, 16);
  if constexpr (G >= 64) v += __shfl_xor(v, 32);
  return v;
}
constexpr int kAsmTile = 2048;
template <int G>
__global__ void __launch_bounds__(256) k_assemble_near(
    int nleaf, const int *__restrict__ row_beg,
    const int *__restrict__ rows, const int *__restrict__ col_beg, const int *__restrict__ ent_beg,
    const int *__restrict__ ent_cell, const uint64_t *__restrict__ ent_mask,
    const uint8_t *__restrict__ ent_near, const int *__restrict__ ent_colpos,
    const int64_t *__restrict__ val_beg, const int *__restrict__ cdofs, int dpc, int nq,
    const int *__restrict__ dbl_ptr, const int *__restrict__ dbl_idx, const double *__restrict__ xyz,
    const double *__restrict__ qx, const double *__restrict__ qn, const double *__restrict__ qw,
    const double *__restrict__ qsh, const double *__restrict__ geom, const double *__restrict__ sing,
    const int *__restrict__ sing_cnt, int sing_stride, double *__restrict__ vals) {
#pragma clang fp contract(off)
  using T = double;
  constexpr int IPB = 256 / G;  // items per workgroup pass
  __shared__ double acc[2 * kAsmTile];
  const int lane = (int)(threadIdx.x & (G - 1));
  const int g = (int)blockIdx.x;  // row g of the leaf-ordered rows
  int lo = 0, hi = nleaf - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (row_beg[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const int leaf = lo;
  const int r = g - row_beg[leaf], node = rows[g];
  const int e0 = ent_beg[leaf], e1 = ent_beg[leaf + 1];
  const int C = col_beg[leaf + 1] - col_beg[leaf];
  T *out = vals + 2 * (val_beg[leaf] + (int64_t)r * C);
  const T x0 = (T)xyz[3 * node], x1 = (T)xyz[3 * node + 1], x2 = (T)xyz[3 * node + 2];
  for (int c0 = 0; c0 < C; c0 += kAsmTile) {
    const int cw = min(kAsmTile, C - c0);
    for (int i = (int)threadIdx.x; i < 2 * cw; i += 256) acc[i] = 0.0;
    __syncthreads();
    // every lane runs every pass (the group sums need the whole wavefront)
    for (int eb = e0; eb < e1; eb += IPB) {
      const int e = eb + (int)(threadIdx.x / G);
      const bool valid = e < e1;
      T Nj[9], Dj[9];
      for (int j = 0; j < 9; ++j) Nj[j] = Dj[j] = T(0);
      const int *cp = nullptr;
      if (valid) {
        const int cell = ent_cell[e];
        const int *cd = cdofs + (int64_t)cell * dpc;
        int js = -1;
        if (ent_near[e]) {
          for (int j = 0; j < dpc && js < 0; ++j)
            for (int k = dbl_ptr[node]; k < dbl_ptr[node + 1]; ++k)
              if (dbl_idx[k] == cd[j]) {
                js = j;
                break;
              }
        }
        cp = ent_colpos + (int64_t)e * dpc;
        if (js < 0) {
          const int q = lane;
          if (q < nq && ((ent_mask[e] >> q) & 1)) {
            const int64_t g = (int64_t)cell * nq + q;
            const T R0 = qx[3 * g] - x0, R1 = qx[3 * g + 1] - x1, R2 = qx[3 * g + 2] - x2;
            const T r = dsqrt(R0 * R0 + R1 * R1 + R2 * R2);
            const T r2 = r * r;
            const T s = T(1) / (r * T(4 * kPi));
            const T den = T(-4 * kPi) * r2 * r;
            const T dn = (R0 / den) * qn[3 * g] + (R1 / den) * qn[3 * g + 1] + (R2 / den) * qn[3 * g + 2];
            const T wq = qw[g];
            for (int j = 0; j < dpc; ++j) {
              const T sh = qsh[g * dpc + j];
              Nj[j] = dn * sh * wq;
              Dj[j] = s * sh * wq;
            }
          }
        } else {
          const double *Gm = geom + (int64_t)cell * dpc * 3;
          const double *rule = sing + (int64_t)js * sing_stride;
          const int nsq = sing_cnt[js];
          const int ps = 1 + 3 * dpc;
          for (int q = lane; q < nsq; q += G) {
            const double *pt = rule + q * ps;
            double X[3] = {0, 0, 0}, a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
            for (int j = 0; j < dpc; ++j)
              for (int k = 0; k < 3; ++k) {
                X[k] += pt[1 + j] * Gm[3 * j + k];
                a[k] += pt[1 + dpc + j] * Gm[3 * j + k];
                b[k] += pt[1 + 2 * dpc + j] * Gm[3 * j + k];
              }
            const double cr0 = a[1] * b[2] - a[2] * b[1], cr1 = a[2] * b[0] - a[0] * b[2],
                         cr2 = a[0] * b[1] - a[1] * b[0];
            const double J = sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
            const double R0 = X[0] - x0, R1 = X[1] - x1, R2 = X[2] - x2;
            const double r = sqrt(R0 * R0 + R1 * R1 + R2 * R2);
            const double r2 = r * r;
            const double s = 1. / (r * 4 * kPi);
            const double den = -4 * kPi * r2 * r;
            const double dn = (R0 / den) * (cr0 / J) + (R1 / den) * (cr1 / J) + (R2 / den) * (cr2 / J);
            const double wq = J * pt[0];
            for (int j = 0; j < dpc; ++j) {
              Nj[j] += dn * pt[1 + j] * wq;
              Dj[j] += s * pt[1 + j] * wq;
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        Nj[j] = group_sum<G>(Nj[j]);
        Dj[j] = group_sum<G>(Dj[j]);
      }
      if (valid) {
        // lane t adds value t of the item's 2 dpc: N of shape function t, or D of t - dpc
        for (int t = lane; t < 2 * dpc; t += G) {
          const int j = t < dpc ? t : t - dpc;
          T v = T(0);
#pragma unroll
          for (int k = 0; k < 9; ++k)
            if (k == j) v = t < dpc ? Nj[k] : Dj[k];
          const int c = cp[j] - c0;
          if (c >= 0 && c < cw) atomicAdd(&acc[2 * c + (t < dpc ? 0 : 1)], v);
        }
      }
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < 2 * cw; i += 256) out[2 * c0 + i] = acc[i];
    __syncthreads();
  }
}

__global__ void k_round_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// The FP32 operator's single-value near field in row pairs: leaf l's R x C
// block (row-major in the FP64 copy at val_beg[l]) goes to pval_beg[l] (an
// even offset: 8-byte aligned) with rows 2p and 2p + 1 interleaved column by
// column, (2p + s, c) at 2 p C + 2 c + s, and an odd last row as it is, at
// (R - 1) C + c.  A lane then loads both rows of a pair at a column with one
// 8-byte load.  One workgroup per leaf.
__global__ void __launch_bounds__(256) k_round_f32_pairs(int nleaf, const int *__restrict__ row_beg,
                                                         const int *__restrict__ col_beg,
                                                         const int64_t *__restrict__ val_beg,
                                                         const int64_t *__restrict__ pval_beg,
                                                         const double *__restrict__ a, float *__restrict__ b) {
  const int l = blockIdx.x;
  if (l >= nleaf) return;
  const int R = row_beg[l + 1] - row_beg[l], C = col_beg[l + 1] - col_beg[l], Re = R & ~1;
  const double *src = a + val_beg[l];
  float *dst = b + pval_beg[l];
  for (int64_t i = threadIdx.x; i < (int64_t)R * C; i += blockDim.x) {
    const int r = (int)(i / C), c = (int)(i - (int64_t)r * C);
    const int64_t o = r < Re ? (int64_t)(r & ~1) * C + 2 * c + (r & 1) : i;
    dst[o] = (float)src[i];
  }
}

// The operator's single values in column vectors of 16 bytes (CW = 4 FP32
// or 2 FP64 columns): leaf l's block goes to qval_beg[l] with (r, c) at
// ((c / CW) R + r) CW + c % CW, its columns zero-padded to a multiple of CW
// (16-byte aligned).  A lane (a row) then loads CW columns of its row with
// one 16-byte load, and a wavefront's lanes read consecutive addresses
// (k_spmv_cols).  From the FP64 row-major copy; one workgroup per leaf.
template <typename T>
__global__ void __launch_bounds__(256) k_round_cols(int nleaf, const int *__restrict__ row_beg,
                                                    const int *__restrict__ col_beg,
                                                    const int64_t *__restrict__ val_beg,
                                                    const int64_t *__restrict__ qval_beg, const double *__restrict__ a,
                                                    T *__restrict__ b) {
  constexpr int CW = 16 / sizeof(T);
  const int l = blockIdx.x;
  if (l >= nleaf) return;
  const int R = row_beg[l + 1] - row_beg[l], C = col_beg[l + 1] - col_beg[l];
  const int64_t nv = (int64_t)R * ((C + CW - 1) / CW * CW);
  const double *src = a + val_beg[l];
  T *dst = b + qval_beg[l];
  for (int64_t o = threadIdx.x; o < nv; o += blockDim.x) {
    const int64_t q = o / (CW * (int64_t)R);
    const int rem = (int)(o - q * CW * R), r = rem / CW, c = CW * (int)q + rem % CW;
    dst[o] = c < C ? (T)src[(int64_t)r * C + c] : T(0);
  }
}

// ---------------------------------------------------------------------------
// Near-field SpMV, fused N.phi and D.dphi on the shared pattern
// (prec_neumann_matrix.vmult / prec_dirichlet_matrix.vmult, bem_fma.cc:2408-2409),
// on the leaf-blocked layout.  Three kernels (a workgroup per leaf; a
// wavefront per leaf with 2 or 4 rows in flight); set-up times them on the
// assembled operator and keeps the fastest per precision and storage.
//
// k_spmv_lds: one workgroup per leaf.  All rows of a leaf share
// its column set, so the workgroup gathers (phi, dphi) at those columns once
// into LDS (tiles of kSpmvTile columns) and each wavefront then streams whole
// rows of (N, D) pairs against the LDS copy: the value block is read once,
// coalesced, and the gathers are paid per leaf instead of per row.  Leaves are
// dealt out so that each XCD (workgroups are placed round-robin over the 8
// XCDs) takes a contiguous range of Morton-ordered leaves, whose columns
// overlap, into its own L2.
// ---------------------------------------------------------------------------
template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

constexpr int kSpmvTile = 2048;
constexpr int kXcd = 8;

// NK = 2: (N, D) pairs per entry, outN = N.phi and outD = D.dphi.  NK = 1:
// one value per entry (the operator's matrix: N on Neumann columns, -D on
// Dirichlet columns), outN = M.phi (dphi and outD unused).
// (N, D) pairs are clang's native two-lane vectors: register-resident arrays
// of them stay out of scratch memory.
template <typename T> struct NativePair;
template <> struct NativePair<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct NativePair<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <typename T, int NK> struct NearEntry { using type = typename NativePair<T>::type; };
template <typename T> struct NearEntry<T, 1> { using type = T; };
// The FP32 operator's single values are kept twice: row-major like every
// other near-field storage, and in row pairs (k_round_f32_pairs) read two
// rows per 8-byte load by the kernels instantiated with PR (SpMV variants
// 3..5); set-up's tuner times both.
template <typename T, int NK>
constexpr bool kPairedF32 = sizeof(T) == 4 && NK == 1;
template <typename T, int NK>
__device__ __forceinline__ typename NearEntry<T, NK>::type near_x(const double *phi, const double *dphi, int col) {
  if constexpr (NK == 2) return typename NearEntry<T, NK>::type{(T)phi[col], (T)dphi[col]};
  else return (T)phi[col];
}
template <typename T, int NK>
__device__ __forceinline__ void near_fma(T &aN, T &aD, const typename NearEntry<T, NK>::type &a,
                                         const typename NearEntry<T, NK>::type &x) {
  if constexpr (NK == 2) {
    aN += a.x * x.x;
    aD += a.y * x.y;
  } else {
    aN += a * x;
  }
}
template <int NK>
__device__ __forceinline__ void near_store(double *outN, double *outD, int node, bool first, double aN, double aD) {
  outN[node] = (first ? 0.0 : outN[node]) + aN;
  if constexpr (NK == 2) outD[node] = (first ? 0.0 : outD[node]) + aD;
}

// The near field's per-leaf blocks and the vectors of one SpMV.
template <typename T>
struct NearArgs {
  const int *row_beg, *rows, *col_beg, *cols;
  const int64_t *val_beg;
  const T *vals;
  const double *phi, *dphi;
  double *outN, *outD;
  int64_t stride;  // between right-hand sides (blockIdx.y) of the vectors
};
// The vectors of right-hand side blockIdx.y.
template <typename T>
__device__ __forceinline__ NearArgs<T> near_rhs(NearArgs<T> a) {
  const int64_t o = (int64_t)blockIdx.y * a.stride;
  a.phi += o;
  if (a.dphi) a.dphi += o;
  a.outN += o;
  if (a.outD) a.outD += o;
  return a;
}

// One leaf by a workgroup of NT threads: its inputs gathered once into LDS
// (xs, kSpmvTile entries), each wavefront streaming whole rows.
template <typename T, int NK, int NT, bool PR = false>
__device__ __forceinline__ void spmv_lds_leaf(int leaf, typename NearEntry<T, NK>::type *__restrict__ xs,
                                              const NearArgs<T> &a) {
  static_assert(!PR || kPairedF32<T, NK>, "row pairs hold the FP32 single values only");
  using E = typename NearEntry<T, NK>::type;
  constexpr int NW = NT / 64;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int r0 = a.row_beg[leaf], R = a.row_beg[leaf + 1] - r0;
  const int cb = a.col_beg[leaf], C = a.col_beg[leaf + 1] - cb;
  const E *V = reinterpret_cast<const E *>(a.vals) + a.val_beg[leaf];
  if constexpr (PR) {
    // FP32 single values in row pairs (k_round_f32_pairs): a wavefront takes
    // rows 2p, 2p + 1 with one 8-byte load per column; an odd last row goes
    // to wavefront (R / 2) % NW, as a single row
    using P2 = typename NativePair<float>::type;
    const int Re = R & ~1;
    const P2 *V2 = reinterpret_cast<const P2 *>(V);
    for (int c0 = 0; c0 < C; c0 += kSpmvTile) {
      const int CT = min(kSpmvTile, C - c0);
      if (c0 > 0) __syncthreads();
      for (int c = t; c < CT; c += NT) xs[c] = near_x<T, NK>(a.phi, a.dphi, a.cols[cb + c0 + c]);
      __syncthreads();
      // two row pairs per wavefront at a time (pa and pa + NW): twice the
      // loads in flight
      const int NP = Re / 2;
      for (int pa = w; pa < NP; pa += 2 * NW) {
        const int pb = pa + NW;
        const bool hb = pb < NP;
        const P2 *va = V2 + (int64_t)pa * C + c0, *vb = V2 + (int64_t)(hb ? pb : pa) * C + c0;
        T a0 = 0, a1 = 0, b0 = 0, b1 = 0;
#pragma unroll 4
        for (int c = lane; c < CT; c += 64) {
          const T xv = xs[c];
          const P2 u = va[c], z = vb[c];
          a0 += u.x * xv;
          a1 += u.y * xv;
          b0 += z.x * xv;
          b1 += z.y * xv;
        }
        a0 = wave_sum(a0);
        a1 = wave_sum(a1);
        b0 = wave_sum(b0);
        b1 = wave_sum(b1);
        if (lane == 0) {
          near_store<1>(a.outN, nullptr, a.rows[r0 + 2 * pa], c0 == 0, (double)a0, 0.0);
          near_store<1>(a.outN, nullptr, a.rows[r0 + 2 * pa + 1], c0 == 0, (double)a1, 0.0);
          if (hb) {
            near_store<1>(a.outN, nullptr, a.rows[r0 + 2 * pb], c0 == 0, (double)b0, 0.0);
            near_store<1>(a.outN, nullptr, a.rows[r0 + 2 * pb + 1], c0 == 0, (double)b1, 0.0);
          }
        }
      }
      // an odd last row: the last wavefront, in every tile
      if ((R & 1) && w == NW - 1) {
        const T *v = V + (int64_t)Re * C + c0;
        T s = 0;
#pragma unroll 4
        for (int c = lane; c < CT; c += 64) s += v[c] * xs[c];
        s = wave_sum(s);
        if (lane == 0) near_store<1>(a.outN, nullptr, a.rows[r0 + R - 1], c0 == 0, (double)s, 0.0);
      }
    }
    return;
  }
  for (int c0 = 0; c0 < C; c0 += kSpmvTile) {
    const int CT = min(kSpmvTile, C - c0);
    if (c0 > 0) __syncthreads();  // previous tile fully consumed
    for (int c = t; c < CT; c += NT) xs[c] = near_x<T, NK>(a.phi, a.dphi, a.cols[cb + c0 + c]);
    __syncthreads();
    for (int r = w; r < R; r += NW) {
      const E *v = V + (int64_t)r * C + c0;
      T aN = 0, aD = 0;
#pragma unroll 4
      for (int c = lane; c < CT; c += 64) near_fma<T, NK>(aN, aD, v[c], xs[c]);
      aN = wave_sum(aN);
      if constexpr (NK == 2) aD = wave_sum(aD);
      // the wave that owns row r owns it in every tile: plain read-modify-write
      if (lane == 0) near_store<NK>(a.outN, a.outD, a.rows[r0 + r], c0 == 0, (double)aN, (double)aD);
    }
  }
}

// Leaves are dealt out so that each XCD takes a contiguous range of per_xcd
// leaves (workgroups go round-robin over the 8 XCDs): neighbouring leaves,
// whose columns overlap, share an L2.
template <typename T, int NK, bool PR = false>
__global__ void __launch_bounds__(256) k_spmv_lds(int nleaf, int per_xcd, NearArgs<T> a) {
  using E = typename NearEntry<T, NK>::type;
  __shared__ E xs[kSpmvTile];
  const int leaf = (int)(blockIdx.x % kXcd) * per_xcd + (int)(blockIdx.x / kXcd);
  if (leaf >= nleaf) return;
  spmv_lds_leaf<T, NK, 256, PR>(leaf, xs, near_rhs(a));
}

// The operator's single values in column vectors (k_round_cols: CW = 4 FP32
// or 2 FP64 columns per 16 bytes), one workgroup per leaf: the leaf's inputs
// gathered once into LDS as in k_spmv_lds, then lane = row and the four
// wavefronts split the column vectors, each lane streaming its row with
// 16-byte loads against the vectors' inputs broadcast from LDS; no wavefront
// sums.  A leaf of at most 32 (16) rows runs two (four) lane groups per
// wavefront on alternate vectors; one of more than 64 rows runs in blocks of
// 64.  The 4 x groups partial sums of a row meet in LDS.  Leaves dealt out
// per XCD as k_spmv_lds.
__device__ __forceinline__ float vdot16(const float4 &m, const float4 &x) {
  return (m.x * x.x + m.y * x.y) + (m.z * x.z + m.w * x.w);
}
__device__ __forceinline__ double vdot16(const double2 &m, const double2 &x) { return m.x * x.x + m.y * x.y; }
template <typename T, int UNR>
__global__ void __launch_bounds__(256) k_spmv_cols(int nleaf, int per_xcd, NearArgs<T> a) {
  using TV = std::conditional_t<sizeof(T) == 4, float4, double2>;
  constexpr int CW = 16 / sizeof(T);
  constexpr int QTILE = kSpmvTile / CW;  // column vectors per LDS tile
  __shared__ TV xsv[QTILE];
  __shared__ T part[16][64];
  const int leaf = (int)(blockIdx.x % kXcd) * per_xcd + (int)(blockIdx.x / kXcd);
  if (leaf >= nleaf) return;
  a = near_rhs(a);
  const int t = (int)threadIdx.x, w = t >> 6, lane = t & 63;
  const int r0 = a.row_beg[leaf], R = a.row_beg[leaf + 1] - r0;
  const int cb = a.col_beg[leaf], C = a.col_beg[leaf + 1] - cb;
  const int NQ = (C + CW - 1) / CW;
  const TV *VV = reinterpret_cast<const TV *>(a.vals + a.val_beg[leaf]);
  T *xs = reinterpret_cast<T *>(xsv);
  const int H = R <= 16 ? 4 : (R <= 32 ? 2 : 1), RL = 64 / H;
  const int h = lane / RL, rl = lane - h * RL;
  const int S = 4 * H, s = w * H + h;  // vector streams of the workgroup, this lane's
  auto gather = [&](int q0, int QT) {
    for (int c = t; c < CW * QT; c += 256) {
      const int cc = CW * q0 + c;
      xs[c] = cc < C ? near_x<T, 1>(a.phi, nullptr, a.cols[cb + cc]) : T(0);
    }
  };
  const bool one_tile = NQ <= QTILE;
  if (one_tile) {
    gather(0, NQ);
    __syncthreads();
  }
  for (int rb = 0; rb < R; rb += 64) {
    const int r = rb + rl;
    T acc = 0;
    for (int q0 = 0; q0 < NQ; q0 += QTILE) {
      const int QT = min(QTILE, NQ - q0);
      if (!one_tile) {
        __syncthreads();  // the previous tile consumed
        gather(q0, QT);
        __syncthreads();
      }
      if (r < R) {
        const TV *v = VV + (int64_t)q0 * R + r;
#pragma unroll UNR
        for (int q = s; q < QT; q += S) acc += vdot16(v[(int64_t)q * R], xsv[q]);
      }
    }
    part[s][lane] = acc;
    __syncthreads();
    if (t < min(RL, R - rb)) {
      T z = 0;
      for (int ww = 0; ww < 4; ++ww)
        for (int g = 0; g < H; ++g) z += part[ww * H + g][g * RL + t];
      a.outN[a.rows[r0 + rb + t]] = (double)z;
    }
    __syncthreads();  // part reused by the next row block
  }
}

// One leaf by one wavefront, no LDS and no barrier, RPI rows per iteration.
// For each tile of 64 x CPL leaf columns (64 bytes of values per lane and
// row: CPL = 8 / 4 for FP32 / FP64 pairs, 16 / 8 for single values) every
// lane holds the inputs of its CPL columns in registers, and the wave streams
// RPI rows at a time against them (RPI CPL coalesced loads per lane in
// flight).  Outputs as spmv_lds_leaf.
template <typename T, int NK, int RPI, bool PR = false>
__device__ __forceinline__ void spmv_wave_leaf(int leaf, int lane, const NearArgs<T> &a) {
  static_assert(!PR || kPairedF32<T, NK>, "row pairs hold the FP32 single values only");
  using E = typename NearEntry<T, NK>::type;
  constexpr int CPL = 64 / (int)sizeof(E);
  constexpr int TILE = 64 * CPL;
  const int r0 = a.row_beg[leaf], R = a.row_beg[leaf + 1] - r0;
  const int cb = a.col_beg[leaf], C = a.col_beg[leaf + 1] - cb;
  const E *V = reinterpret_cast<const E *>(a.vals) + a.val_beg[leaf];
  const E zero{};
  if constexpr (PR) {
    // FP32 single values in row pairs (k_round_f32_pairs): RPI row pairs per
    // iteration, one 8-byte load per (pair, column); an odd last row alone
    using P2 = typename NativePair<float>::type;
    const int Re = R & ~1, NP = Re / 2;
    const P2 *V2 = reinterpret_cast<const P2 *>(V);
    const P2 zero2{};
    for (int c0 = 0; c0 < C; c0 += TILE) {
      T x[CPL];
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        const int c = c0 + lane + 64 * j;
        x[j] = c < C ? near_x<T, NK>(a.phi, a.dphi, a.cols[cb + c]) : zero;
      }
      for (int p = 0; p < NP; p += RPI) {
        P2 v[RPI][CPL];
#pragma unroll
        for (int q = 0; q < RPI; ++q) {
          const P2 *vq = V2 + (int64_t)(p + q < NP ? p + q : p) * C;
#pragma unroll
          for (int j = 0; j < CPL; ++j) {
            const int c = c0 + lane + 64 * j;
            v[q][j] = c < C ? vq[c] : zero2;
          }
        }
        T s0[RPI], s1[RPI];
#pragma unroll
        for (int q = 0; q < RPI; ++q) {
          s0[q] = 0;
          s1[q] = 0;
#pragma unroll
          for (int j = 0; j < CPL; ++j) {
            s0[q] += v[q][j].x * x[j];
            s1[q] += v[q][j].y * x[j];
          }
          s0[q] = wave_sum(s0[q]);
          s1[q] = wave_sum(s1[q]);
        }
        if (lane == 0)
#pragma unroll
          for (int q = 0; q < RPI; ++q)
            if (p + q < NP) {
              near_store<1>(a.outN, nullptr, a.rows[r0 + 2 * (p + q)], c0 == 0, (double)s0[q], 0.0);
              near_store<1>(a.outN, nullptr, a.rows[r0 + 2 * (p + q) + 1], c0 == 0, (double)s1[q], 0.0);
            }
      }
      if (R & 1) {
        const T *vl = V + (int64_t)Re * C;
        T s = 0;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          const int c = c0 + lane + 64 * j;
          if (c < C) s += vl[c] * x[j];
        }
        s = wave_sum(s);
        if (lane == 0) near_store<1>(a.outN, nullptr, a.rows[r0 + R - 1], c0 == 0, (double)s, 0.0);
      }
    }
    return;
  }
  for (int c0 = 0; c0 < C; c0 += TILE) {
    E x[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int c = c0 + lane + 64 * j;
      x[j] = c < C ? near_x<T, NK>(a.phi, a.dphi, a.cols[cb + c]) : zero;
    }
    // RPI rows per iteration: their loads are in flight together
    for (int r = 0; r < R; r += RPI) {
      E v[RPI][CPL];
#pragma unroll
      for (int q = 0; q < RPI; ++q) {
        const E *vq = V + (int64_t)(r + q < R ? r + q : r) * C;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          const int c = c0 + lane + 64 * j;
          v[q][j] = c < C ? vq[c] : zero;
        }
      }
      T aN[RPI], aD[RPI];
#pragma unroll
      for (int q = 0; q < RPI; ++q) {
        aN[q] = 0;
        aD[q] = 0;
#pragma unroll
        for (int j = 0; j < CPL; ++j) near_fma<T, NK>(aN[q], aD[q], v[q][j], x[j]);
        aN[q] = wave_sum(aN[q]);
        if constexpr (NK == 2) aD[q] = wave_sum(aD[q]);
      }
      // the wave owns the leaf's rows in every tile: plain read-modify-write
      if (lane == 0)
#pragma unroll
        for (int q = 0; q < RPI; ++q)
          if (r + q < R) near_store<NK>(a.outN, a.outD, a.rows[r0 + r + q], c0 == 0, (double)aN[q], (double)aD[q]);
    }
  }
}

template <typename T, int NK, int RPI, bool PR = false>
__global__ void __launch_bounds__(256) k_spmv_wave(int nleaf, NearArgs<T> a) {
  const int leaf = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (leaf >= nleaf) return;
  spmv_wave_leaf<T, NK, RPI, PR>(leaf, threadIdx.x & 63, near_rhs(a));
}

// ---------------------------------------------------------------------------
// P2M at the leaves (generate_multipole_expansions, bem_fma.cc:2064-2114) as a
// product with unit multipoles precomputed at set-up (multipole_integrals,
// bem_fma.cc:1675-1737).  A quad point y with weight w and normal n adds, for
// each DoF j of its cell, N_j(y) w / 4pi times
//   Ker1 (MultipoleExpansion::AddNormDer, multipole_expansion.cc:124-181):
//        (n . grad) R_n^m(y - c)                              [applied to phi]
//   Ker2 (MultipoleExpansion::Add, multipole_expansion.cc:93-122):
//        R_n^m(y - c)                                         [applied to dphi]
// with the regular solid harmonic R_n^m(d) = rho^n Pt_n^m(cos a) e^{-i m b}.
// Both come from trig-free Cartesian recurrences in d = y - c:
//   R_m^m = -sqrt((2m-1)/2m) (x - i y) R_{m-1}^{m-1}
//   R_n^m = ra(n,m) z R_{n-1}^m - rb(n,m) |d|^2 R_{n-2}^m
// and their directional derivatives (product rule, D(x - i y) = n_x - i n_y,
// D z = n_z, D |d|^2 = 2 d.n).  Summed per (leaf, DoF) these give the leaf's
// unit-multipole block U[column][kernel][kp]; per matvec
//   ME_leaf = sum_columns (phi, dphi)[dof] * U[column]
// is a dense, coalesced stream of U.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_allsum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Set-up, FP64: one thread per quad point of the leaves, atomics into U
// (for leaves whose unit multipoles do not fit in LDS).
__global__ void __launch_bounds__(256) k_p2m_unit(
    int npts, int p, int kt, int kp, const int *__restrict__ qpt, const int *__restrict__ qleaf,
    const int *__restrict__ leaves, const int *__restrict__ qcol, int dpc, const double *__restrict__ qx,
    const double *__restrict__ qn, const double *__restrict__ qw, const double *__restrict__ qsh,
    const double *__restrict__ pmin, const double *__restrict__ delta, double *__restrict__ U) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const int64_t g = qpt[i];
  const int b = leaves[qleaf[i]];
  const double hd = delta[b] * 0.5;
  const double x = qx[3 * g] - (pmin[3 * b] + hd), y = qx[3 * g + 1] - (pmin[3 * b + 1] + hd),
               z = qx[3 * g + 2] - (pmin[3 * b + 2] + hd);
  const double nx = qn[3 * g], ny = qn[3 * g + 1], nz = qn[3 * g + 2];
  const double r2 = x * x + y * y + z * z;
  const double dn2 = 2.0 * (x * nx + y * ny + z * nz);
  double wj[9];
  double *col[9];
  for (int j = 0; j < dpc; ++j) {
    wj[j] = qsh[g * dpc + j] * qw[g] / (4 * kPi);
    col[j] = U + (int64_t)qcol[(int64_t)i * dpc + j] * 2 * kp;
  }
  // level-scaled multipoles: coefficient of degree n divided by delta^n
  const double inv_delta = 1.0 / delta[b];
  auto emit = [&](int k, int n, double Rr, double Ri, double Gr, double Gi) {
    double sc = 1.0;
    for (int q = 0; q < n; ++q) sc *= inv_delta;
    Rr *= sc; Ri *= sc; Gr *= sc; Gi *= sc;
    for (int j = 0; j < dpc; ++j) {
      atomicAdd(col[j] + k, wj[j] * Gr);
      atomicAdd(col[j] + kt + k, wj[j] * Gi);
      atomicAdd(col[j] + kp + k, wj[j] * Rr);
      atomicAdd(col[j] + kp + kt + k, wj[j] * Ri);
    }
  };
  double Rmr = 1, Rmi = 0, Gmr = 0, Gmi = 0;  // R_m^m, D R_m^m
  for (int m = 0; m <= p; ++m) {
    if (m > 0) {
      const double f = -PMF<double>(m);
      const double nr = f * (x * Rmr + y * Rmi), ni = f * (x * Rmi - y * Rmr);
      const double gr = f * ((nx * Rmr + ny * Rmi) + (x * Gmr + y * Gmi));
      const double gi = f * ((nx * Rmi - ny * Rmr) + (x * Gmi - y * Gmr));
      Rmr = nr; Rmi = ni; Gmr = gr; Gmi = gi;
    }
    double Rcr = Rmr, Rci = Rmi, Gcr = Gmr, Gci = Gmi, Rpr = 0, Rpi = 0, Gpr = 0, Gpi = 0;
    emit(m * (m + 1) / 2 + m, m, Rcr, Rci, Gcr, Gci);
    for (int n = m + 1; n <= p; ++n) {
      const double a = RA<double>(n, m), bb = RB<double>(n, m);
      const double Rnr = a * z * Rcr - bb * r2 * Rpr, Rni = a * z * Rci - bb * r2 * Rpi;
      const double Gnr = a * (nz * Rcr + z * Gcr) - bb * (dn2 * Rpr + r2 * Gpr);
      const double Gni = a * (nz * Rci + z * Gci) - bb * (dn2 * Rpi + r2 * Gpi);
      emit(n * (n + 1) / 2 + m, n, Rnr, Rni, Gnr, Gni);
      Rpr = Rcr; Rpi = Rci; Gpr = Gcr; Gpi = Gci;
      Rcr = Rnr; Rci = Rni; Gcr = Gnr; Gci = Gni;
    }
  }
}

// Set-up, FP64, one workgroup per leaf whose unit multipoles fit in LDS: the
// block [column][2 kp] of the leaf is summed in LDS (atomic adds there) and
// written once with plain stores, since no other leaf owns those columns.
// half 0 / 1: only the first / second kernel's kp values of every column
// ([column][kp] in LDS, for a leaf whose whole block does not fit), written
// into their half of the [column][2 kp] block; -1: both.
// Work items are (quad point, m): the item runs the m-recurrence to R_m^m
// and then the n-recurrence of column m of the harmonics, as k_p2m_unit
// does for all m of its point.
__global__ void __launch_bounds__(256) k_p2m_unit_leaf(
    int p, int kt, int kp, const int *__restrict__ leaf_ids, const int *__restrict__ leaves,
    const int *__restrict__ qbeg, const int *__restrict__ col_beg, const int *__restrict__ qpt,
    const int *__restrict__ qcol, int dpc, const double *__restrict__ qx, const double *__restrict__ qn,
    const double *__restrict__ qw, const double *__restrict__ qsh, const double *__restrict__ pmin,
    const double *__restrict__ delta, double *__restrict__ U, int half) {
  extern __shared__ double sU[];
  const int li = leaf_ids[blockIdx.x];
  const int b = leaves[li];
  const int c0 = col_beg[li], nc = col_beg[li + 1] - c0;
  const int q0 = qbeg[li], nq = qbeg[li + 1] - q0;
  const int K2 = 2 * kp;
  const int S = half < 0 ? K2 : kp;  // LDS values per column
  for (int i = threadIdx.x; i < nc * S; i += blockDim.x) sU[i] = 0.0;
  __syncthreads();
  const double hd = delta[b] * 0.5, inv_delta = 1.0 / delta[b];
  const double cx = pmin[3 * b] + hd, cy = pmin[3 * b + 1] + hd, cz = pmin[3 * b + 2] + hd;
  for (int item = threadIdx.x; item < nq * (p + 1); item += blockDim.x) {
    const int i = q0 + item / (p + 1), m = item % (p + 1);
    const int64_t g = qpt[i];
    const double x = qx[3 * g] - cx, y = qx[3 * g + 1] - cy, z = qx[3 * g + 2] - cz;
    const double nx = qn[3 * g], ny = qn[3 * g + 1], nz = qn[3 * g + 2];
    const double r2 = x * x + y * y + z * z;
    const double dn2 = 2.0 * (x * nx + y * ny + z * nz);
    double wj[9];
    int cj[9];
    for (int j = 0; j < dpc; ++j) {
      wj[j] = qsh[g * dpc + j] * qw[g] / (4 * kPi);
      cj[j] = (qcol[(int64_t)i * dpc + j] - c0) * S;
    }
    double Rmr = 1, Rmi = 0, Gmr = 0, Gmi = 0;  // R_m^m, D R_m^m
    double sc = 1.0;                            // 1 / delta^n
    for (int mm = 1; mm <= m; ++mm) {
      const double f = -PMF<double>(mm);
      const double nr = f * (x * Rmr + y * Rmi), ni = f * (x * Rmi - y * Rmr);
      const double gr = f * ((nx * Rmr + ny * Rmi) + (x * Gmr + y * Gmi));
      const double gi = f * ((nx * Rmi - ny * Rmr) + (x * Gmi - y * Gmr));
      Rmr = nr; Rmi = ni; Gmr = gr; Gmi = gi;
      sc *= inv_delta;
    }
    auto emit = [&](int k, double s, double Rr, double Ri, double Gr, double Gi) {
      Rr *= s; Ri *= s; Gr *= s; Gi *= s;
      for (int j = 0; j < dpc; ++j) {
        double *col = sU + cj[j];
        if (half <= 0) {
          atomicAdd(col + k, wj[j] * Gr);
          atomicAdd(col + kt + k, wj[j] * Gi);
        }
        if (half < 0) {
          atomicAdd(col + kp + k, wj[j] * Rr);
          atomicAdd(col + kp + kt + k, wj[j] * Ri);
        } else if (half == 1) {
          atomicAdd(col + k, wj[j] * Rr);
          atomicAdd(col + kt + k, wj[j] * Ri);
        }
      }
    };
    double Rcr = Rmr, Rci = Rmi, Gcr = Gmr, Gci = Gmi, Rpr = 0, Rpi = 0, Gpr = 0, Gpi = 0;
    emit(m * (m + 1) / 2 + m, sc, Rcr, Rci, Gcr, Gci);
    for (int n = m + 1; n <= p; ++n) {
      const double a = RA<double>(n, m), bb = RB<double>(n, m);
      const double Rnr = a * z * Rcr - bb * r2 * Rpr, Rni = a * z * Rci - bb * r2 * Rpi;
      const double Gnr = a * (nz * Rcr + z * Gcr) - bb * (dn2 * Rpr + r2 * Gpr);
      const double Gni = a * (nz * Rci + z * Gci) - bb * (dn2 * Rpi + r2 * Gpi);
      sc *= inv_delta;
      emit(n * (n + 1) / 2 + m, sc, Rnr, Rni, Gnr, Gni);
      Rpr = Rcr; Rpi = Rci; Gpr = Gcr; Gpi = Gci;
      Rcr = Rnr; Rci = Rni; Gcr = Gnr; Gci = Gni;
    }
  }
  __syncthreads();
  double *out = U + (int64_t)c0 * K2;
  if (half < 0) {
    for (int i = threadIdx.x; i < nc * K2; i += blockDim.x) out[i] = sU[i];
  } else {
    for (int i = threadIdx.x; i < nc * kp; i += blockDim.x) {
      const int c = i / kp;
      out[(int64_t)c * K2 + half * kp + (i - c * kp)] = sU[i];
    }
  }
}

// Per matvec: one workgroup per leaf.  The leaf's 2 kp outputs are read as
// NOV 16-byte vectors per column; thread t owns vector o = t % NOV over the
// columns ph, ph + nph, ... (ph = t / NOV, nph = 256 / NOV column phases), and
// the phases are summed through LDS at the end.  The leaf's (phi, dphi) values
// are gathered into LDS once per tile of columns.
constexpr int kP2mTile = 1024;
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };
__device__ __forceinline__ void axpy16(float4 &a, float x, const float4 &u) {
  a.x += x * u.x; a.y += x * u.y; a.z += x * u.z; a.w += x * u.w;
}
__device__ __forceinline__ void axpy16(double2 &a, double x, const double2 &u) {
  a.x += x * u.x; a.y += x * u.y;
}
__device__ __forceinline__ void add16(float4 &a, const float4 &u) { a.x += u.x; a.y += u.y; a.z += u.z; a.w += u.w; }
// lanes 0..31 add the values of lanes 32..63
__device__ __forceinline__ void add_upper_half(float4 &a) {
  a.x += __shfl_down(a.x, 32);
  a.y += __shfl_down(a.y, 32);
  a.z += __shfl_down(a.z, 32);
  a.w += __shfl_down(a.w, 32);
}
__device__ __forceinline__ void add_upper_half(double2 &a) {
  a.x += __shfl_down(a.x, 32);
  a.y += __shfl_down(a.y, 32);
}
__device__ __forceinline__ void add16(double2 &a, const double2 &u) { a.x += u.x; a.y += u.y; }

// NK = 2: U is [column][Ker1 kp][Ker2 kp] and the leaf gets both kernels'
// expansions from (phi, dphi).  NK = 1: U is the operator's [column][kp]
// (U1 on Neumann DoFs, -U2 on Dirichlet DoFs) and the leaf gets one combined
// expansion from phi (dphi unused).  ME is [leaf][NK kp].
// Right-hand side blockIdx.y reads phi (dphi) + y vs and writes ME + y es.
// Workgroups past the last leaf (fused pipelines, cl.leaf set) clear, for
// the same application, the MEs of the blocks that are not leaves, every LE
// and the M2P accumulator (if any); the leaves' MEs are stored whole by their
// own workgroups.
template <typename T>
struct P2mClear {
  const uint8_t *leaf = nullptr;  // per block: 1 for a leaf
  T *le = nullptr;
  double *acc = nullptr;
  int64_t ne = 0, na = 0;  // expansion values (blocks x kp), accumulator length
};
// NK = 1 reads the compact unit multipoles: kpc values per column, compact
// row c holding full row cmap[c] (-1: padding), without the imaginary parts
// of the m = 0 coefficients (zero for real sources) and without the
// expansion's padding; the leaf's ME is expanded to kp rows in LDS.
constexpr int kMaxKp = 144;  // kp at order 10
template <typename T, int NK>
__global__ void __launch_bounds__(256) k_p2m(int nleaf, const int *__restrict__ leaves,
                                             const int *__restrict__ col_beg, const int *__restrict__ col_dof,
                                             const T *__restrict__ U, int kp, int kpc, const int *__restrict__ cmap,
                                             const double *__restrict__ phi, const double *__restrict__ dphi,
                                             T *__restrict__ ME, int64_t vs, int64_t es,
                                             P2mClear<T> cl = P2mClear<T>()) {
  if ((int)blockIdx.x >= nleaf) {  // clearing workgroups (gridDim.y is 1 here)
    const int64_t nt = (int64_t)(gridDim.x - nleaf) * blockDim.x;
    for (int64_t i = (int64_t)(blockIdx.x - nleaf) * blockDim.x + threadIdx.x; i < 2 * cl.ne + cl.na; i += nt) {
      if (i < cl.ne) {
        if (!cl.leaf[i / (NK * kp)]) ME[i] = T(0);
      } else if (i < 2 * cl.ne) {
        cl.le[i - cl.ne] = T(0);
      } else {
        cl.acc[i - 2 * cl.ne] = 0.0;
      }
    }
    return;
  }
  phi += (int64_t)blockIdx.y * vs;
  if (dphi) dphi += (int64_t)blockIdx.y * vs;
  ME += (int64_t)blockIdx.y * es;
  using T2 = typename Vec2<T>::type;
  using TV = typename Vec16<T>::type;
  constexpr int VW = 16 / sizeof(T);
  __shared__ T2 xs[kP2mTile];
  __shared__ TV part[256];
  const int li = blockIdx.x;
  const int t = threadIdx.x;
  // values per column: 2 kp (both kernels) or kpc (a multiple of 4): NOV <= 144
  const int NO = NK == 2 ? 2 * kp : kpc, NOV = NO / VW;
  const int nph = 256 / NOV;
  const int o = t % NOV, ph = t / NOV;
  const bool active = ph < nph;
  const bool k2 = NK == 2 && o * VW >= kp;  // a vector never straddles the two kernels
  const int c0 = col_beg[li], C = col_beg[li + 1] - c0;
  const TV *Ub = reinterpret_cast<const TV *>(U + (int64_t)c0 * NO) + o;
  TV acc{};
  for (int cb = 0; cb < C; cb += kP2mTile) {
    const int CT = min(kP2mTile, C - cb);
    if (cb > 0) __syncthreads();
    for (int c = t; c < CT; c += 256) {
      const int d = col_dof[c0 + cb + c];
      xs[c] = T2{(T)phi[d], NK == 2 ? (T)dphi[d] : T(0)};
    }
    __syncthreads();
    if (active) {
#pragma unroll 4
      for (int c = ph; c < CT; c += nph) {
        const T2 xv = xs[c];
        axpy16(acc, k2 ? xv.y : xv.x, Ub[(int64_t)(cb + c) * NOV]);
      }
    }
  }
  part[t] = acc;
  __syncthreads();
  if constexpr (NK == 2) {
    if (t < NOV) {
      TV a = part[t];
      for (int q = 1; q < nph; ++q) add16(a, part[q * NOV + t]);
      reinterpret_cast<TV *>(ME + (int64_t)leaves[li] * NO)[t] = a;
    }
  } else {
    // the compact sums into their rows of the leaf's kp-row ME (zeros
    // elsewhere), then stored whole
    __shared__ T sme[kMaxKp];
    for (int i = t; i < kp; i += 256) sme[i] = T(0);
    __syncthreads();
    if (t < NOV) {
      TV a = part[t];
      for (int q = 1; q < nph; ++q) add16(a, part[q * NOV + t]);
      const T *av = reinterpret_cast<const T *>(&a);
#pragma unroll
      for (int k = 0; k < VW; ++k) {
        const int r = cmap[t * VW + k];
        if (r >= 0) sme[r] = av[k];
      }
    }
    __syncthreads();
    T *me = ME + (int64_t)leaves[li] * kp;
    for (int i = t; i < kp; i += 256) me[i] = sme[i];
  }
}

// The operator's single-value data from the two-kernel data (set-up, FP64):
// near-field M = N on Neumann columns, -D on Dirichlet columns; unit
// multipoles U1 on Neumann DoFs, -U2 on Dirichlet DoFs.
__global__ void k_op_near(int nleaf, const int *__restrict__ row_beg, const int *__restrict__ col_beg,
                          const int *__restrict__ cols, const int64_t *__restrict__ val_beg,
                          const uint8_t *__restrict__ dir, const double *__restrict__ vals,
                          double *__restrict__ op) {
  const int leaf = blockIdx.x;
  const int R = row_beg[leaf + 1] - row_beg[leaf];
  const int cb = col_beg[leaf], C = col_beg[leaf + 1] - cb;
  const int64_t vb = val_beg[leaf];
  for (int64_t i = threadIdx.x; i < (int64_t)R * C; i += blockDim.x) {
    const int c = (int)(i % C);
    const bool d = dir[cols[cb + c]];
    op[vb + i] = d ? -vals[2 * (vb + i) + 1] : vals[2 * (vb + i)];
  }
}
// (compact: kpc values per column, value o of row cmap[o], zero for padding)
__global__ void k_op_unit(int64_t ncol, int kp, int kpc, const int *__restrict__ cmap,
                          const int *__restrict__ col_dof, const uint8_t *__restrict__ dir,
                          const double *__restrict__ U, double *__restrict__ op) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ncol * kpc;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = i / kpc;
    const int r = cmap[(int)(i - c * kpc)];
    op[i] = r < 0 ? 0.0 : (dir[col_dof[c]] ? -U[c * 2 * kp + kp + r] : U[c * 2 * kp + r]);
  }
}

// Padded length of one real expansion vector at order P.
template <int P>
struct Kp {
  static constexpr int kt = (P + 1) * (P + 2) / 2;
  static constexpr int value = ((2 * kt + 15) / 16) * 16;
};

// ---------------------------------------------------------------------------
// Translations as batched GEMMs on the matrix cores: M2M
// (MultipoleExpansion::Add(child), multipole_expansion.cc:184-261, driven by
// bem_fma.cc:2274-2339), M2L (LocalExpansion::Add(multipole),
// local_expansion.cc:179-248, driven by bem_fma.cc:2622-2652) and L2L
// (LocalExpansion::Add(parent), local_expansion.cc:107-176, driven by
// bem_fma.cc:2600-2602) are all OUT[tgt] += T_class . IN[src] with one R-linear
// KP x KP operator per class (M2M / L2L: level and child octant; M2L:
// integer offset, all levels).  Pairs are grouped by class into chunks of 32; a block
// computes C (KP x 64) = A (KP x KP) . B (KP x 64 source columns, both
// kernels) with 16x16x4 MFMA tiles (KP/16 wavefronts, wave w: row tile w,
// all 64 columns), stages C through LDS and adds it to the targets with
// coalesced atomics (targets are distinct within a chunk, not across chunks).
//   f32 (v_mfma_f32_16x16x4f32): lane l holds A[l%16][l/16], B[l/16][l%16],
//       D[4(l/16)+j][l%16], j < 4
//   f64 (v_mfma_f64_16x16x4f64): same inputs, D[l/16+4j][l%16]
// (both layouts checked on gfx950 with an in-bounds probe).
// ---------------------------------------------------------------------------
typedef float v4f32 __attribute__((ext_vector_type(4)));
typedef double v4f64 __attribute__((ext_vector_type(4)));
template <typename T> struct Mfma;
template <> struct Mfma<float> {
  using acc = v4f32;
  static __device__ __forceinline__ acc mac(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int j) { return 4 * (l >> 4) + j; }
};
template <> struct Mfma<double> {
  using acc = v4f64;
  static __device__ __forceinline__ acc mac(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int j) { return (l >> 4) + 4 * j; }
};

static_assert(kM2LChunk == 32 && kChunk1 == 64, "k_translate maps 2 x 32 or 1 x 64 pairs onto 64 GEMM columns");

// LDS tile (elements of T).  B is [64 columns][BS] (one source vector per
// column), C is [64 columns][CS], both column-major, sharing one buffer.
//   FP32: BS = KP + 2, so that the MFMA reads (16 columns x 2 k-rows per half
//   wavefront) hit 32 distinct banks; filled with 8-byte stores.  CS = KP + 4:
//   an FP32 MFMA lane holds 4 consecutive rows of one column, stored as one
//   16-byte write, and 8 such writes cover the 32 banks once.
//   FP64: BS = KP + 2 (16-byte fill stores), CS = KP + 1.
template <int P, typename T>
struct TrLayout {
  static constexpr int KP = Kp<P>::value;
  static constexpr int BS = KP + 2;
  static constexpr int CS = sizeof(T) == 4 ? KP + 4 : KP + 1;
  static constexpr int NR = KP / 16;        // 16-row tiles = wavefronts per block
  static constexpr int NT = 64 * NR;        // threads per block
  static constexpr int elems = 64 * (BS > CS ? BS : CS);
  static constexpr size_t bytes = (size_t)elems * sizeof(T);
};

// Wavefront w owns output rows 16w..16w+15 for all 64 columns (4 column
// groups), so each operator fragment it loads from L2 feeds 4 MFMAs.  The
// source vectors of the chunk (NK kp contiguous values per pair) are gathered
// with 16-byte loads, all issued before the LDS stores.  NK = 2: both
// kernels' expansions, 32 pairs x 2 columns; NK = 1: one combined expansion,
// 64 pairs.
// KIND (0 M2M, 1 M2L, 2 L2L) only names the instantiation, so that traces and
// profiles tell the three passes apart; KIND 3 is M2L that stores each pair's
// result into its own slot instead of adding it to the target (summed per
// target by k_m2l_reduce).
template <typename T>
struct TrArgs {
  const int *chunk_op, *chunk_beg, *chunk_cnt, *tgt, *src;
  const T *ops, *IN;
  T *OUT;
};

template <typename T, int P, int KIND, int NK>
__device__ __forceinline__ void translate_chunk(int c, T *__restrict__ sB, int *__restrict__ sT,
                                                int *__restrict__ sS, const TrArgs<T> &args) {
  const int *__restrict__ chunk_op = args.chunk_op;
  const int *__restrict__ chunk_beg = args.chunk_beg;
  const int *__restrict__ chunk_cnt = args.chunk_cnt;
  const int *__restrict__ tgt = args.tgt;
  const int *__restrict__ src = args.src;
  const T *__restrict__ ops = args.ops;
  const T *__restrict__ IN = args.IN;
  T *__restrict__ OUT = args.OUT;
  using Lay = TrLayout<P, T>;
  constexpr int KP = Lay::KP, K2 = NK * KP, BS = Lay::BS, CS = Lay::CS, NT = Lay::NT;
  constexpr int PC = 64 / NK;                // pairs per chunk
  static_assert(NT == 4 * KP, "launch bound");
  static_assert(sizeof(T) == 8 || (BS % 2 == 0 && CS % 4 == 0), "LDS vector alignment");
  using M = Mfma<T>;
  using TV = typename Vec16<T>::type;
  constexpr int V = 16 / sizeof(T);          // elements per 16-byte vector
  constexpr int NVP = K2 / V;                // vectors per pair (both kernels)
  constexpr int NV = PC * NVP;               // vectors per chunk
  constexpr int PER = (NV + NT - 1) / NT;    // vectors per thread
  constexpr int KT2 = 2 * Kp<P>::kt;
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const T *A = ops + (int64_t)chunk_op[c] * KP * KP;  // column-major, read from L2
  const int beg = chunk_beg[c], cnt = chunk_cnt[c];
  if (t < PC) {
    sT[t] = t < cnt ? tgt[beg + t] : -1;
    sS[t] = t < cnt ? src[beg + t] : -1;
  }
  __syncthreads();
  TV buf[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = t + j * NT;
    const int pr = i / NVP;
    buf[j] = TV{};
    if (i < NV && pr < cnt)
      buf[j] = *reinterpret_cast<const TV *>(IN + (int64_t)sS[pr] * K2 + (i - pr * NVP) * V);
  }
  // M2M and L2L launches hold few chunks per level and wait on latency: each
  // lane loads its operator values for the whole K loop at once after the
  // gather (row 16 w + (l & 15), columns k0 + (l >> 4) for k0 = 0, 4, ..).
  // M2L has tens of thousands of chunks in flight and keeps its registers
  // for occupancy, loading operator values inside the K loop.
  constexpr bool PRE = KIND != 1;
  constexpr int NA = PRE ? KP / 4 : 1;
  T areg[NA];
  const T *ak = A + (l >> 4) * KP + 16 * w + (l & 15);
  if constexpr (PRE) {
#pragma unroll
    for (int j = 0; j < NA; ++j) areg[j] = ak[(int64_t)(4 * j) * KP];
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = t + j * NT;
    if (i < NV) {
      const int pr = i / NVP, off = (i - pr * NVP) * V;
      const int col = NK * pr + (off >= KP), k = off >= KP ? off - KP : off;
      T *d = sB + col * BS + k;
      if constexpr (sizeof(T) == 4) {  // 8-byte aligned: two 8-byte stores
        reinterpret_cast<float2 *>(d)[0] = float2{buf[j].x, buf[j].y};
        reinterpret_cast<float2 *>(d)[1] = float2{buf[j].z, buf[j].w};
      } else {
        *reinterpret_cast<TV *>(d) = buf[j];
      }
    }
  }
  __syncthreads();
  typename M::acc acc[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) acc[g] = typename M::acc{0, 0, 0, 0};
  const T *bk = sB + (l & 15) * BS + (l >> 4);
  if constexpr (PRE) {
#pragma unroll
    for (int k0 = 0; k0 < KP; k0 += 4) {
      const T a = areg[k0 / 4];
#pragma unroll
      for (int g = 0; g < 4; ++g) acc[g] = M::mac(a, bk[g * 16 * BS + k0], acc[g]);
    }
  } else {
#pragma unroll 4
    for (int k0 = 0; k0 < KP; k0 += 4) {
      const T a = ak[k0 * KP];
#pragma unroll
      for (int g = 0; g < 4; ++g) acc[g] = M::mac(a, bk[g * 16 * BS + k0], acc[g]);
    }
  }
  __syncthreads();  // B tile no longer needed: reuse the buffer for C
  T *sC = sB;
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int g = 0; g < 4; ++g)
      *reinterpret_cast<float4 *>(sC + (16 * g + (l & 15)) * CS + 16 * w + 4 * (l >> 4)) =
          float4{acc[g][0], acc[g][1], acc[g][2], acc[g][3]};
  } else {
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int j = 0; j < 4; ++j) sC[(16 * g + (l & 15)) * CS + 16 * w + M::row(l, j)] = acc[g][j];
  }
  __syncthreads();
  if constexpr (KIND == 3) {
    // M2L into per-pair slots: pair beg + cc into its slot sT[cc] (tgt holds
    // the slot positions here) at OUT + sT[cc] * KP, whole: the padding rows
    // are zero, so every slot is whole, aligned cache lines, written by
    // consecutive lanes
    static_assert(NK == 1, "slot M2L is the single-expansion operator's");
    for (int i = t; i < cnt * KP; i += NT) {
      const int cc = i / KP, row = i - cc * KP;
      OUT[(int64_t)sT[cc] * KP + row] = sC[cc * CS + row];
    }
  } else {
    // coalesced accumulation: consecutive lanes add consecutive rows of one
    // target (the Im rows of m = 0, zero for real densities, included:
    // skipping them scatters the lanes and measured slower)
    for (int i = t; i < NK * cnt * KT2; i += NT) {
      const int cc = i / KT2, row = i - cc * KT2;
      atomicAdd(OUT + (int64_t)sT[cc / NK] * K2 + (cc % NK) * KP + row, sC[cc * CS + row]);
    }
  }
}

// The slot M2L's sums (KIND 3): one wavefront per target adds its slots,
// consecutive and kp values apart (aligned, whole cache lines), and stores
// the target's local expansion, which nothing else has written since it was
// cleared.  A lane loads 16 bytes of a slot, so a slot takes kp / V lanes (V
// values per 16 bytes): SP = 2 slots at a time in the two half-waves when
// that fits (FP32, kp <= 128), else one.  Each half adds every SP-th slot
// from its first, eight loads in flight; the halves meet at the end.  The
// order of the additions is fixed, so the sums do not vary from run to run.
template <typename T, int SP>
__global__ void __launch_bounds__(256) k_m2l_reduce(int ntgt, const int *__restrict__ tgt,
                                                    const int *__restrict__ beg, int kp,
                                                    const T *__restrict__ slots, T *__restrict__ LE) {
  using TV = typename Vec16<T>::type;
  constexpr int V = 16 / sizeof(T), HS = 64 / SP, U = 8;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (w >= ntgt) return;  // uniform per wavefront
  const int h = l / HS, li = l - h * HS;
  const bool act = li * V < kp;
  const int k1 = beg[w + 1];
  const T *s = slots + li * V;
  TV a = {};
  int k = beg[w] + h;
  for (; k + (U - 1) * SP < k1; k += U * SP) {
    TV v[U];
#pragma unroll
    for (int q = 0; q < U; ++q)
      if (act) v[q] = *reinterpret_cast<const TV *>(s + (int64_t)(k + q * SP) * kp);
#pragma unroll
    for (int q = 0; q < U; ++q)
      if (act) add16(a, v[q]);
  }
  for (; k < k1; k += SP)
    if (act) add16(a, *reinterpret_cast<const TV *>(s + (int64_t)k * kp));
  if constexpr (SP == 2) add_upper_half(a);
  if (h == 0 && act) *reinterpret_cast<TV *>(LE + (int64_t)tgt[w] * kp + li * V) = a;
}

// Chunks of one class are consecutive in a chunk list and share their
// operator.  Workgroups go round-robin over the 8 XCDs, so workgroup w takes
// chunk (w % 8) * per + w / 8: each XCD works through a contiguous run of
// chunks and finds their operators in its own L2.  The launch has 8 * per
// workgroups (per = ceil(nch / 8)); those past the last chunk return.
constexpr bool kXcdChunks = true;
__device__ __forceinline__ int xcd_chunk(int w, int nch) {
  if (!kXcdChunks) return w;
  const int per = (nch + kXcd - 1) / kXcd;
  return (w % kXcd) * per + w / kXcd;
}
static inline int xcd_blocks(int nch) { return kXcdChunks ? (nch + kXcd - 1) / kXcd * kXcd : nch; }

template <typename T, int P, int KIND, int NK>
__global__ void __launch_bounds__(4 * Kp<P>::value) k_translate(TrArgs<T> a, int nch) {
  __shared__ __align__(16) T sB[TrLayout<P, T>::elems];
  __shared__ int sT[64 / NK], sS[64 / NK];
  const int c = xcd_chunk(blockIdx.x, nch);
  if (c >= nch) return;
  translate_chunk<T, P, KIND, NK>(c, sB, sT, sS, a);
}

// Evaluation of both kernels' expansions at a point: M2P
// (MultipoleExpansion::Evaluate, multipole_expansion.cc:264-298, rho^(-n-1))
// or L2P (LocalExpansion::Evaluate, local_expansion.cc:251-287, rho^n).
// NK = 1: the single expansion E (v2 unused).  E is level-scaled by the box
// size sc of its block (ME_n / sc^n, LE_n sc^(n+1)), so the radial factors
// are (sc/rho)^n / rho and (rho/sc)^n / sc.
// P > 0 fixes the order at compile time (loops unrolled, the expansion's
// loads scheduled ahead of the recurrences); P = 0 takes p, kt, kp at run time.
template <typename T, bool LOCAL, int NK, int P = 0>
__device__ __forceinline__ void eval2(const T *__restrict__ E, int p_rt, int kt_rt, int kp_rt, T d0, T d1, T d2,
                                      T sc, T &v1, T &v2) {
  const int p = P > 0 ? P : p_rt;
  const int kt = P > 0 ? (P + 1) * (P + 2) / 2 : kt_rt;
  const int kp = P > 0 ? Kp<P>::value : kp_rt;
  // spherical angles of d without trigonometric calls: cos(alpha) = z/rho,
  // sin(alpha) = rho_xy/rho, e^{i beta} = (x + i y)/rho_xy (beta = 0 on the axis)
  const T rxy2 = d0 * d0 + d1 * d1;
  const T rho = dsqrt(rxy2 + d2 * d2);
  const T rxy = dsqrt(rxy2);
  const T x = d2 / rho;
  const T s = rxy / rho;
  const T cb = rxy > T(0) ? d0 / rxy : T(1), sb = rxy > T(0) ? d1 / rxy : T(0);
  const T step = LOCAL ? rho / sc : sc / rho;
  T rstart = T(1) / (LOCAL ? sc : rho);
  T pmm = 1, cm = 1, sm = 0;
  v1 = 0;
  v2 = 0;
#pragma unroll
  for (int m = 0; m <= p; ++m) {
    if (m > 0) {
      pmm *= -PMF<T>(m) * s;
      const T c2 = cm * cb - sm * sb;
      sm = sm * cb + cm * sb;
      cm = c2;
      rstart *= step;
    }
    T pp = 0, pc = pmm, r = rstart;
    const T f2 = m ? T(2) : T(1);
#pragma unroll
    for (int n = m; n <= p; ++n) {
      if (n > m) {
        const T pn = RA<T>(n, m) * x * pc - RB<T>(n, m) * pp;
        pp = pc;
        pc = pn;
        r *= step;
      }
      const int idx = n * (n + 1) / 2 + m;
      const T f = pc * r * f2;
      v1 += (E[idx] * cm - E[kt + idx] * sm) * f;
      if constexpr (NK == 2) v2 += (E[kp + idx] * cm - E[kp + kt + idx] * sm) * f;
    }
  }
}

// M2P for smaller far blocks (bem_fma.cc:2664-2704): one thread per
// (source block, target node) evaluation, listed at set-up.
template <typename T, int NK, int P = 0>
__device__ __forceinline__ void m2p_item(int64_t w, const int *__restrict__ item_node,
                                         const int *__restrict__ item_src, const double *__restrict__ xyz,
                                         const double *__restrict__ pmin, const double *__restrict__ delta, int p,
                                         int kt, int kp, const T *__restrict__ ME, double *__restrict__ outN,
                                         double *__restrict__ outD) {
  const int node = item_node[w], sblk = item_src[w];
  const double hd = delta[sblk] * 0.5;
  const T d0 = (T)(xyz[3 * node] - (pmin[3 * sblk] + hd));
  const T d1 = (T)(xyz[3 * node + 1] - (pmin[3 * sblk + 1] + hd));
  const T d2 = (T)(xyz[3 * node + 2] - (pmin[3 * sblk + 2] + hd));
  T v1, v2;
  eval2<T, false, NK, P>(ME + (int64_t)sblk * NK * kp, p, kt, kp, d0, d1, d2, (T)delta[sblk], v1, v2);
  atomicAdd(outN + node, (double)v1);
  if constexpr (NK == 2) atomicAdd(outD + node, (double)v2);
}

// Right-hand side blockIdx.y reads ME + y es and adds into outN (outD) + y os.
template <typename T, int NK, int P>
__global__ void __launch_bounds__(256) k_m2p(int64_t nitems, const int *__restrict__ item_node,
                                             const int *__restrict__ item_src, const double *__restrict__ xyz,
                                             const double *__restrict__ pmin, const double *__restrict__ delta, int p,
                                             int kt, int kp, const T *__restrict__ ME, double *__restrict__ outN,
                                             double *__restrict__ outD, int64_t es, int64_t os) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= nitems) return;
  const int64_t y = blockIdx.y;
  m2p_item<T, NK, P>(w, item_node, item_src, xyz, pmin, delta, p, kt, kp, ME + y * es, outN + y * os,
                     outD ? outD + y * os : nullptr);
}

// L2P at every node of every leaf (bem_fma.cc:2959-2986): one thread per node.
template <typename T, int NK, int P>
__global__ void __launch_bounds__(256) k_l2p(int nrows, const int *__restrict__ rows,
                                             const int *__restrict__ row_leaf, const int *__restrict__ leaf_block,
                                             const double *__restrict__ xyz, const double *__restrict__ pmin,
                                             const double *__restrict__ delta, int p, int kt, int kp,
                                             const T *__restrict__ LE, double *__restrict__ outN,
                                             double *__restrict__ outD, int64_t es, int64_t os) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  // right-hand side blockIdx.y: LE + y es, outN (outD) + y os
  LE += (int64_t)blockIdx.y * es;
  outN += (int64_t)blockIdx.y * os;
  if (outD) outD += (int64_t)blockIdx.y * os;
  const int node = rows[r];
  const int b = leaf_block[row_leaf[r]];
  const double hd = delta[b] * 0.5;
  const T d0 = (T)(xyz[3 * node] - (pmin[3 * b] + hd));
  const T d1 = (T)(xyz[3 * node + 1] - (pmin[3 * b + 1] + hd));
  const T d2 = (T)(xyz[3 * node + 2] - (pmin[3 * b + 2] + hd));
  T v1, v2;
  eval2<T, true, NK, P>(LE + (int64_t)b * NK * kp, p, kt, kp, d0, d1, d2, (T)delta[b], v1, v2);
  outN[node] += (double)v1;
  if constexpr (NK == 2) outD[node] += (double)v2;
}

// L2P of the operator pipeline with BEMProblem::vmult's epilogue
// (k_vmult_post) folded in: dst = out + m2p + L2P + alpha x on Neumann DoFs
// (out the near field, m2p the M2P sums).
template <typename T, int P>
__global__ void __launch_bounds__(256) k_l2p_post(int nrows, const int *__restrict__ rows,
                                                  const int *__restrict__ row_leaf, const int *__restrict__ leaf_block,
                                                  const double *__restrict__ xyz, const double *__restrict__ pmin,
                                                  const double *__restrict__ delta, int p, int kt, int kp,
                                                  const T *__restrict__ LE, const double *__restrict__ out,
                                                  const double *__restrict__ m2p, const double *__restrict__ alpha,
                                                  const uint8_t *__restrict__ dir, const double *__restrict__ x,
                                                  double *__restrict__ dst) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int node = rows[r];
  const int b = leaf_block[row_leaf[r]];
  const double hd = delta[b] * 0.5;
  const T d0 = (T)(xyz[3 * node] - (pmin[3 * b] + hd));
  const T d1 = (T)(xyz[3 * node + 1] - (pmin[3 * b + 1] + hd));
  const T d2 = (T)(xyz[3 * node + 2] - (pmin[3 * b + 2] + hd));
  T v1, v2;
  eval2<T, true, 1, P>(LE + (int64_t)b * kp, p, kt, kp, d0, d1, d2, (T)delta[b], v1, v2);
  const double nd = (out[node] + m2p[node]) + (double)v1;
  dst[node] = nd + (dir[node] ? 0.0 : alpha[node] * x[node]);
}

template <typename T>
__global__ void k_fill(T *__restrict__ a, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// Clears the operator pipeline's expansions (ME and LE, n_e entries each)
// and its M2P accumulator (n_a doubles) in one launch.
template <typename T>
__global__ void k_clear_operator(T *__restrict__ me, T *__restrict__ le, int64_t ne, double *__restrict__ acc,
                                 int64_t na) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 2 * ne + na;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < ne)
      me[i] = T(0);
    else if (i < 2 * ne)
      le[i - ne] = T(0);
    else
      acc[i - 2 * ne] = 0.0;
  }
}

// Work that shares a launch with M2M, M2L or L2L chunks in the operator
// (single-value) pipeline: the near-field SpMV of leaves [lo, lo + count) in
// spmv_blocks workgroups, and/or the M2P evaluations [m2p_lo, m2p_hi), added
// into their own accumulator m2p_out.
template <typename T>
struct SideArgs {
  NearArgs<T> near;
  int lo, count, per_xcd, spmv_blocks;
  int64_t m2p_lo, m2p_hi;
  const int *item_node, *item_src;
  const double *xyz, *pmin, *delta;
  int p, kt, kp;
  const T *ME;
  double *m2p_out;
};

// Horizontal fusion.  M2M and L2L level launches hold few chunks (one
// workgroup each) and wait on latency, leaving most CUs idle, and M2L's
// workgroups are latency-bound too.  The workgroups after a launch's nch
// chunks run independent work meanwhile: the SpMV (SPMV 0 workgroup per
// leaf, 1 and 2 wavefront per leaf with 2 or 4 rows in flight, -1 none) and
// then, with M2P set, the M2P evaluations.  The SpMV writes the rows of out
// and M2P adds into a separate accumulator, so neither waits for the other;
// both need only the final MEs.  Each goes whole into one launch: its
// workgroups are latency-bound themselves, and shares spread over several
// launches would pay that latency once per launch.  One LDS buffer serves
// every role.
template <typename T, int P, int KIND, int SPMV, bool M2P>
__global__ void __launch_bounds__(4 * Kp<P>::value) k_translate_side(int nch, TrArgs<T> tr, SideArgs<T> sd) {
  constexpr int NT = 4 * Kp<P>::value;
  using E = typename NearEntry<T, 1>::type;
  constexpr size_t TB = TrLayout<P, T>::bytes;
  constexpr size_t SB = SPMV == 0 ? kSpmvTile * sizeof(E) : 0;
  __shared__ __align__(16) unsigned char smem[TB > SB ? TB : SB];
  __shared__ int sT[64], sS[64];
  const int npad = kXcdChunks ? (nch + kXcd - 1) / kXcd * kXcd : nch;  // xcd_blocks(nch)
  if ((int)blockIdx.x < npad) {
    const int c = xcd_chunk(blockIdx.x, nch);
    if (c < nch) translate_chunk<T, P, KIND, 1>(c, reinterpret_cast<T *>(smem), sT, sS, tr);
    return;
  }
  int b = (int)blockIdx.x - npad;
  if constexpr (SPMV >= 0) {
    if (b < sd.spmv_blocks) {
      if constexpr (SPMV == 0) {
        // leaves dealt out per XCD as in k_spmv_lds (b and blockIdx.x differ by a
        // multiple of 8, so b % 8 is the XCD)
        const int leaf = (b % kXcd) * sd.per_xcd + b / kXcd;
        if (leaf < sd.count) spmv_lds_leaf<T, 1, NT>(sd.lo + leaf, reinterpret_cast<E *>(smem), sd.near);
      } else {
        const int leaf = b * (NT / 64) + (int)(threadIdx.x >> 6);
        if (leaf < sd.count) spmv_wave_leaf<T, 1, SPMV == 1 ? 2 : 4>(sd.lo + leaf, threadIdx.x & 63, sd.near);
      }
      return;
    }
    b -= sd.spmv_blocks;
  }
  if constexpr (M2P) {
    const int64_t w = sd.m2p_lo + (int64_t)b * NT + threadIdx.x;
    if (w < sd.m2p_hi)
      m2p_item<T, 1, P>(w, sd.item_node, sd.item_src, sd.xyz, sd.pmin, sd.delta, sd.p, sd.kt, sd.kp, sd.ME,
                     sd.m2p_out, nullptr);
  }
}

// ---------------------------------------------------------------------------
// BEMProblem operator pieces (bem_problem.cc:1164-1260) and the constrained
// rows of ConstrainedOperator::vmult (constrained_matrix.h:91-135).
// ---------------------------------------------------------------------------
// The operator's input: serv_phi (Neumann DoFs, shifted in the pure-Neumann
// case) and serv_dphi_dn (Dirichlet DoFs) share one vector, since the
// operator's near field and unit multipoles already pick N or -D per column.
__global__ void k_vmult_pre(int64_t n, const double *__restrict__ src, const uint8_t *__restrict__ dir,
                            const double *__restrict__ shift, double *__restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double sh = shift ? *shift : 0.0;
  x[i] = dir[i] ? src[i] : (src[i] - sh);
}
// ND = N.sp - D.sd from the operator's FMM, x its input
// (right-hand side blockIdx.y: ND, x and dst + y stride)
__global__ void k_vmult_post(int64_t n, const double *__restrict__ ND, const double *__restrict__ alpha,
                             const uint8_t *__restrict__ dir, const double *__restrict__ x,
                             double *__restrict__ dst, int64_t stride) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t o = (int64_t)blockIdx.y * stride;
  ND += o;
  x += o;
  dst += o;
  dst[i] = ND[i] + (dir[i] ? 0.0 : alpha[i] * x[i]);
}
__global__ void k_rhs_pre(int64_t n, const double *__restrict__ g, const uint8_t *__restrict__ dir,
                          double *__restrict__ sp, double *__restrict__ sd) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  sp[i] = dir[i] ? g[i] : 0.0;
  sd[i] = dir[i] ? 0.0 : g[i];
}
__global__ void k_rhs_post(int64_t n, const double *__restrict__ N, const double *__restrict__ D,
                           const double *__restrict__ alpha, const double *__restrict__ sp, double *__restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[i] = -(N[i] + alpha[i] * sp[i]) + D[i];
}
__global__ void k_shift(int64_t n, double *__restrict__ v, const double *__restrict__ s) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] -= *s;
}
// v += (*sum) * scale
// v += sum scale; right-hand side blockIdx.y: v + y stride, sum[y]
__global__ void k_add_scaled_scalar(int64_t n, double *__restrict__ v, const double *__restrict__ sum, double scale,
                                    int64_t stride) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) v[(int64_t)blockIdx.y * stride + i] += sum[blockIdx.y] * scale;
}
// (right-hand side blockIdx.y: src and dst + y stride)
__global__ void k_constrain(int64_t n, const uint8_t *__restrict__ is_c, const int *__restrict__ cptr,
                            const int *__restrict__ cidx, const double *__restrict__ cval,
                            const double *__restrict__ src, double *__restrict__ dst, int64_t stride) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n || !is_c[i]) return;
  src += (int64_t)blockIdx.y * stride;
  dst += (int64_t)blockIdx.y * stride;
  double v = src[i];
  for (int k = cptr[i]; k < cptr[i + 1]; ++k) v -= cval[k] * src[cidx[k]];
  dst[i] = v;
}
__global__ void k_set_constrained(int64_t n, const uint8_t *__restrict__ is_c, const double *__restrict__ inh,
                                  double *__restrict__ rhs) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n && is_c[i]) rhs[i] = inh[i];
}

// ---------------------------------------------------------------------------
// Near-field block-Jacobi preconditioner (FMA_preconditioner,
// bem_fma.cc:3078-3411, restricted to blocks of the near leaves that share a
// parent block): extraction of the rows x rows blocks of the preconditioner
// matrix and their application.
// ---------------------------------------------------------------------------
// Block g of the preconditioner (several leaves): row r of block g is
// the leaf-ordered near row pos[r] (its leaf row_leaf[pos]); an entry is the
// operator's near value when the column is in that leaf's column list, else
// zero.
__global__ void k_extract_groups(const int *__restrict__ grp_beg, const int *__restrict__ pos,
                                 const int *__restrict__ row_leaf, const int *__restrict__ row_beg,
                                 const int *__restrict__ rows, const int *__restrict__ col_beg,
                                 const int *__restrict__ cols, const int64_t *__restrict__ val_beg,
                                 const double *__restrict__ vals, const uint8_t *__restrict__ dir,
                                 const double *__restrict__ alpha, const int64_t *__restrict__ blk_beg,
                                 double *__restrict__ out) {
  const int g = blockIdx.x;
  const int g0 = grp_beg[g], R = grp_beg[g + 1] - g0;
  for (int idx = threadIdx.x; idx < R * R; idx += blockDim.x) {
    const int r = idx / R, c = idx - r * R;
    const int pr = pos[g0 + r], leaf = row_leaf[pr];
    const int node_r = rows[pr], node_c = rows[pos[g0 + c]];
    const int cb = col_beg[leaf], C = col_beg[leaf + 1] - cb;
    int lo = 0, hi = C - 1, at = -1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int v = cols[cb + mid];
      if (v == node_c) { at = mid; break; }
      if (v < node_c) lo = mid + 1;
      else hi = mid - 1;
    }
    double v = 0.0;
    if (at >= 0) {
      const double *q = vals + 2 * (val_beg[leaf] + (int64_t)(pr - row_beg[leaf]) * C + at);
      v = dir[node_c] ? -q[1] : q[0];
    }
    if (r == c && !dir[node_r]) v += alpha[node_r];
    out[blk_beg[g] + idx] = v;
  }
}

// The inverse blocks are stored in column vectors of 16 bytes (PT = double:
// two columns, the FP64 solves' copy; PT = float: four, the mixed solves'
// copy, a fixed linear preconditioner within FP32 rounding of the inverse,
// which GMRES with right preconditioning applies to its basis and to the
// update alike) and applied in FP64: entry (r, c) of an R-row block at
// vec_beg + ((c / CW) R + r) CW + c % CW, columns zero-padded to a multiple
// of CW, so that a lane reads CW columns of its row in one 16-byte load and
// a wavefront's lanes (rows) read consecutive addresses.
template <typename PT>
__host__ __device__ inline int64_t pinv_vec_index(int r, int c, int R) {
  constexpr int CW = 16 / sizeof(PT);
  return ((int64_t)(c / CW) * R + r) * CW + c % CW;
}
__device__ __forceinline__ void pinv_load(const float *p, double *a) {
  const float4 u = *reinterpret_cast<const float4 *>(p);
  a[0] = u.x, a[1] = u.y, a[2] = u.z, a[3] = u.w;
}
__device__ __forceinline__ void pinv_load(const double *p, double *a) {
  const double2 u = *reinterpret_cast<const double2 *>(p);
  a[0] = u.x, a[1] = u.y;
}

// z[rows of block] = Pinv_block . v[rows of block]; one thread per block-ordered
// row (row_leaf: its block)
template <typename PT>
__global__ void __launch_bounds__(256) k_bjacobi(int nrows, const int *__restrict__ row_leaf,
                                                 const int *__restrict__ row_beg, const int *__restrict__ rows,
                                                 const int64_t *__restrict__ blk_beg, const PT *__restrict__ pinv,
                                                 const double *__restrict__ v, double *__restrict__ z) {
  constexpr int CW = 16 / sizeof(PT);
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows) return;
  const int leaf = row_leaf[g];
  const int r0 = row_beg[leaf], R = row_beg[leaf + 1] - r0, r = g - r0;
  const PT *P = pinv + blk_beg[leaf] + (int64_t)r * CW;
  double acc = 0.0;
  for (int c = 0; c < R; c += CW) {
    double a[CW];
    pinv_load(P + (int64_t)(c / CW) * R * CW, a);
#pragma unroll
    for (int k = 0; k < CW; ++k)
      if (c + k < R) acc += a[k] * v[rows[r0 + c + k]];
  }
  z[rows[g]] = acc;
}
// The same with one workgroup of four wavefronts per block (blocks of at
// most 64 RPL rows, RPL rows per lane): wavefront w sums the column vectors
// w, w + 4, ... of all rows (lane r: row r).  It loads all of them first,
// since their addresses need only the block's offsets, so that they are in
// flight while the block's inputs are gathered (lane c: v at row c + 64 q);
// the inputs are then broadcast from their lanes, and the four partial sums
// meet in LDS.  (right-hand side blockIdx.y: v and z + y stride)
template <typename PT> struct PinvVec;
template <> struct PinvVec<float> {
  using type = float4;
  static __device__ __forceinline__ double at(const float4 &u, int k) {
    return k == 0 ? u.x : k == 1 ? u.y : k == 2 ? u.z : u.w;
  }
};
template <> struct PinvVec<double> {
  using type = double2;
  static __device__ __forceinline__ double at(const double2 &u, int k) { return k == 0 ? u.x : u.y; }
};
template <int RPL, typename PT>
__global__ void __launch_bounds__(256) k_bjacobi_wg(const int *__restrict__ row_beg, const int *__restrict__ rows,
                                                    const int64_t *__restrict__ blk_beg,
                                                    const PT *__restrict__ pinv, const double *__restrict__ v,
                                                    double *__restrict__ z, int64_t stride) {
  constexpr int CW = 16 / sizeof(PT);
  constexpr int NI = 64 * RPL / CW / 4;  // column vectors per wavefront in a block of 64 RPL rows
  using VT = typename PinvVec<PT>::type;
  __shared__ double part[4][64 * RPL];
  const int blk = blockIdx.x;
  v += (int64_t)blockIdx.y * stride;
  z += (int64_t)blockIdx.y * stride;
  const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int r0 = row_beg[blk], R = row_beg[blk + 1] - r0, nq = (R + CW - 1) / CW;
  const PT *P = pinv + blk_beg[blk];
  VT a[NI][RPL];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int cq = w + 4 * i;
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int row = lane + 64 * q;
      a[i][q] = cq < nq && row < R ? *reinterpret_cast<const VT *>(P + ((int64_t)cq * R + row) * CW) : VT{};
    }
  }
  double xv[RPL], acc[RPL];
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int row = lane + 64 * q;
    xv[q] = row < R ? v[rows[r0 + row]] : 0.0;  // zero past the block: the padded columns
    acc[q] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int cq = w + 4 * i;
    if (cq >= nq) break;  // the same for the whole wavefront
    double xc[CW];
#pragma unroll
    for (int k = 0; k < CW; ++k) {
      const int c = CW * cq + k;
      xc[k] = RPL == 1 || c < 64 ? read_lane(xv[0], c & 63) : read_lane(xv[RPL - 1], c & 63);
    }
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < CW; ++k) s += PinvVec<PT>::at(a[i][q], k) * xc[k];
      acc[q] += s;
    }
  }
#pragma unroll
  for (int q = 0; q < RPL; ++q) part[w][lane + 64 * q] = acc[q];
  __syncthreads();
  if (w < RPL) {
    const int row = lane + 64 * w;
    if (row < R) z[rows[r0 + row]] = (part[0][row] + part[1][row]) + (part[2][row] + part[3][row]);
  }
}

// ---------------------------------------------------------------------------
// GMRES building blocks (FP64).  Dots of w against nv basis vectors: blocks
// (chunk, j) write partial sums; a second kernel finishes them.
// ---------------------------------------------------------------------------
constexpr int kRedThreads = 256;
constexpr int kRedChunk = kRedThreads * 8;

__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  return r;
}

__global__ void __launch_bounds__(kRedThreads) k_mdot(int64_t n, int64_t ldv, const double *__restrict__ V,
                                                      const double *__restrict__ w, double *__restrict__ part) {
  __shared__ double sh[kRedThreads / 64];
  const int j = blockIdx.y;
  const double *v = V + (int64_t)j * ldv;
  const int64_t beg = (int64_t)blockIdx.x * kRedChunk;
  const int64_t end = min(n, beg + (int64_t)kRedChunk);
  double acc = 0;
  for (int64_t i = beg + threadIdx.x; i < end; i += kRedThreads) acc += v[i] * w[i];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[(int64_t)j * gridDim.x + blockIdx.x] = r;
}
__global__ void k_finish(int nchunk, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double sh[kRedThreads / 64];
  const int j = blockIdx.x;
  double acc = 0;
  for (int i = threadIdx.x; i < nchunk; i += blockDim.x) acc += part[(int64_t)j * nchunk + i];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) out[j] = r;
}
__global__ void k_sqrt_inplace(double *v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) v[0] = sqrt(v[0]);
}
// w -= sum_j V_j h_j
__global__ void k_gemv_sub(int64_t n, int64_t ldv, int nv, const double *__restrict__ V,
                           const double *__restrict__ h, double *__restrict__ w) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = w[i];
  for (int j = 0; j < nv; ++j) acc -= V[(int64_t)j * ldv + i] * h[j];
  w[i] = acc;
}
// x += sum_j V_j y_j (basis vectors stored as VT: double, or float for the
// refinement's inner solves; arithmetic in FP64)
template <typename VT>
__global__ void k_gemv_add(int64_t n, int64_t ldv, int nv, const VT *__restrict__ V, const double *__restrict__ y,
                           double *__restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = x[i];
  for (int j = 0; j < nv; ++j) acc += (double)V[(int64_t)j * ldv + i] * y[j];
  x[i] = acc;
}
// out_j = sum_l V_l Q[j nv + l] for 8 outputs j = 8 blockIdx.y .. per thread
// (vectors ld apart): the recycled Krylov space's C = V Q
template <typename VT>
__global__ void __launch_bounds__(256) k_tsgemm8(int64_t n, int64_t ld, int nv, const VT *__restrict__ V,
                                                 const double *__restrict__ Q, int nout, VT *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j0 = 8 * blockIdx.y;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int l = 0; l < nv; ++l) {
    const double v = (double)V[(int64_t)l * ld + i];
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (j0 + q < nout) acc[q] += v * Q[(int64_t)(j0 + q) * nv + l];
  }
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (j0 + q < nout) out[(int64_t)(j0 + q) * ld + i] = (VT)acc[q];
}
__global__ void k_axpby(int64_t n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}
__global__ void k_scale_to(int64_t n, const double *__restrict__ x, double s, double *__restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = s * x[i];
}
// y = s x stored as VT, and the stored value also into y2 (FP64) when given
template <typename VT>
__global__ void k_scale_to_basis(int64_t n, const double *__restrict__ x, double s, VT *__restrict__ y,
                                 double *__restrict__ y2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const VT v = (VT)(s * x[i]);
  y[i] = v;
  if (y2) y2[i] = (double)v;
}
// v /= *nrm when *nrm > 0 (the Arnoldi normalisation, norm kept on the device)
__global__ void k_scale_inv(int64_t n, double *__restrict__ v, const double *__restrict__ nrm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const double s = *nrm;
  if (i < n && s > 0) v[i] /= s;
}

// The Arnoldi step's dot products with a fixed, small number of partial sums
// per column (P row ranges), which lets the kernels that use the sums finish
// them themselves instead of in launches of their own: V_j . w for j < nv
// into part[j * P + range].
// Enough row ranges that each thread streams only a few dozen entries of a
// column: with 16 the projection's dots waited on their load chains (about
// 11 us for any number of columns up to 16 at 393k rows).
constexpr int kDotParts = 64;    // per column of V . w (a wavefront finishes a column: one partial per lane)
constexpr int kNormParts = 256;  // of |w|^2 (each scaling workgroup finishes them)
// (skip: when set and *skip != 0, the launch does nothing: the second
// projection of a step that the first left without severe cancellation)
template <typename VT>
__global__ void __launch_bounds__(kRedThreads) k_mdot_parts(int64_t n, int64_t ldv, const VT *__restrict__ V,
                                                            const double *__restrict__ w, double *__restrict__ part,
                                                            const int *__restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sh[kRedThreads / 64];
  const int j = blockIdx.y, P = (int)gridDim.x;
  const VT *v = V + (int64_t)j * ldv;
  const int64_t len = (n + P - 1) / P;
  const int64_t beg = (int64_t)blockIdx.x * len, end = min(n, beg + len);
  double acc = 0;
#pragma unroll 4
  for (int64_t i = beg + threadIdx.x; i < end; i += kRedThreads) acc += (double)v[i] * w[i];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[(int64_t)j * P + blockIdx.x] = r;
}
// w -= sum_j V_j h_j, h_j the sum of column j's P <= 64 partials, summed
// by every block in the same order (wavefront j mod 4, which loads all of
// its columns' partials first); block 0 also stores h (a Hessenberg
// column), and every block the sum of squares of its rows of the new w in
// npart_out[block] when npart_out is given.  nv <= 128 (GMRES restart <=
// 127).
template <typename VT>
__global__ void __launch_bounds__(256) k_gemv_sub_fin(int64_t n, int64_t ldv, int nv, const VT *__restrict__ V,
                                                      const double *__restrict__ part, int P,
                                                      double *__restrict__ h_out, double *__restrict__ w,
                                                      const int *__restrict__ skip, double *__restrict__ npart_out) {
  if (skip && *skip) {  // a skipped second projection: coefficients zero, w as it is
    if (blockIdx.x == 0)
      for (int j = threadIdx.x; j < nv; j += blockDim.x) h_out[j] = 0.0;
    return;
  }
  constexpr int kMaxJ = 128 / 4;  // columns per wavefront
  __shared__ double hs[128];
  {
    const int wf = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
    double pv[kMaxJ];
#pragma unroll
    for (int q = 0; q < kMaxJ; ++q) {
      const int j = wf + 4 * q;
      pv[q] = j < nv && l < P ? part[(int64_t)j * P + l] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kMaxJ; ++q) {
      const int j = wf + 4 * q;
      if (j >= nv) break;  // the same for the whole wavefront
      const double s = wave_sum(pv[q]);
      if (l == 0) {
        hs[j] = s;
        if (blockIdx.x == 0) h_out[j] = s;
      }
    }
  }
  __syncthreads();
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (i < n) {
    acc = w[i];
    for (int j = 0; j < nv; ++j) acc -= (double)V[(int64_t)j * ldv + i] * hs[j];
    w[i] = acc;
  }
  if (npart_out) {  // the same for the whole block
    __shared__ double sh[4];
    const double r = block_sum(acc * acc, sh);
    if (threadIdx.x == 0) npart_out[blockIdx.x] = r;
  }
}
// Selective reorthogonalisation (Daniel, Gragg, Kaufman & Stewart 1976, with
// a smaller threshold): after the first projection w1 = w0 - V h1, with
// |w0|^2 = |h1|^2 + |w1|^2, the second one is needed only when |w1| <
// eta |w0| (cancellation has eaten all but a fraction eta of w0).  One
// workgroup sums the nv coefficients h1 and the P norm partials of w1 (the
// first projection's blocks), sets skip[0] for the second projection's
// launches and stores |w1|^2 in nrm1_sq for the scaling that follows.
__global__ void __launch_bounds__(256) k_reorth_flag(const double *__restrict__ h1, int nv,
                                                     const double *__restrict__ npart, int P, double eta,
                                                     int *__restrict__ skip, double *__restrict__ nrm1_sq) {
  __shared__ double sh[4];
  double a = 0.0, b = 0.0;
  for (int j = (int)threadIdx.x; j < nv; j += 256) a += h1[j] * h1[j];
  for (int c = (int)threadIdx.x; c < P; c += 256) b += npart[c];
  const double shh = block_sum(a, sh);
  __syncthreads();
  const double sw = block_sum(b, sh);
  if (threadIdx.x == 0) {
    skip[0] = sw >= eta * eta * (shh + sw) ? 1 : 0;
    *nrm1_sq = sw;
  }
}
// |w|^2 for the scaling after the projections: nrm1_sq when the second
// projection was skipped (skip set and *skip != 0), else the sum of the P
// partials of the last projection's blocks (summed by the whole block, in
// the same order in every block).  The result is returned to every thread.
__device__ __forceinline__ double proj_norm_sq(const int *__restrict__ skip, const double *__restrict__ nrm1_sq,
                                               const double *__restrict__ part, int P) {
  __shared__ double sh[4], q;
  if (skip && *skip) return *nrm1_sq;  // the same for the whole block
  double a = 0.0;
  for (int c = (int)threadIdx.x; c < P; c += 256) a += part[c];
  const double r = block_sum(a, sh);
  if (threadIdx.x == 0) q = r;
  __syncthreads();
  return q;
}
// v /= |v| with |v|^2 from proj_norm_sq; block 0 stores |v|.
__global__ void __launch_bounds__(256) k_scale_fin(int64_t n, double *__restrict__ v, const int *__restrict__ skip,
                                                   const double *__restrict__ nrm1_sq,
                                                   const double *__restrict__ part, int P,
                                                   double *__restrict__ nrm_out) {
  const double s = sqrt(proj_norm_sq(skip, nrm1_sq, part, P));
  if (blockIdx.x == 0 && threadIdx.x == 0) *nrm_out = s;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n && s > 0) v[i] /= s;
}
// k_scale_fin from w into two places: the new basis vector (stored as VT)
// and, as stored, the next step's fixed input (the operator's, FP64)
template <typename VT>
__global__ void __launch_bounds__(256) k_scale_fin2(int64_t n, const double *__restrict__ w, VT *__restrict__ v,
                                                    double *__restrict__ v2, const int *__restrict__ skip,
                                                    const double *__restrict__ nrm1_sq,
                                                    const double *__restrict__ part, int P,
                                                    double *__restrict__ nrm_out) {
  const double s = sqrt(proj_norm_sq(skip, nrm1_sq, part, P));
  if (blockIdx.x == 0 && threadIdx.x == 0) *nrm_out = s;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    const VT x = (VT)(s > 0 ? w[i] / s : w[i]);
    v[i] = x;
    v2[i] = (double)x;
  }
}

// Forest segments: sum of squares of each part (one block per part) and a
// per-part scaling y_i = f[part(i)] x_i.
__global__ void __launch_bounds__(kRedThreads) k_seg_sumsq(const int64_t *__restrict__ off,
                                                           const double *__restrict__ v, double *__restrict__ out) {
  __shared__ double sh[kRedThreads / 64];
  const int s = blockIdx.x;
  double acc = 0;
  for (int64_t i = off[s] + threadIdx.x; i < off[s + 1]; i += kRedThreads) acc += v[i] * v[i];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) out[s] = r;
}
// out[b] = sum of v[b n .. b n + n), one workgroup per right-hand side b
__global__ void __launch_bounds__(kRedThreads) k_rhs_sums(int64_t n, const double *__restrict__ v,
                                                          double *__restrict__ out) {
  __shared__ double sh[kRedThreads / 64];
  const double *vb = v + (int64_t)blockIdx.x * n;
  double acc = 0;
  for (int64_t i = threadIdx.x; i < n; i += kRedThreads) acc += vb[i];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}
__global__ void k_seg_scale(int64_t n, const int *__restrict__ seg, const double *__restrict__ f,
                            const double *__restrict__ x, double *__restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = f[seg[i]] * x[i];
}

// Segmented GMRES building blocks: every part of a forest runs its own
// Arnoldi process inside shared basis vectors.  Dots over part-aligned chunks
// (part[j * nch + c]) finished per (part, vector): out[s * ldo + j].
__global__ void __launch_bounds__(kRedThreads) k_mdot_seg(const int64_t *__restrict__ cbeg, int64_t ldv,
                                                          const double *__restrict__ V, const double *__restrict__ w,
                                                          double *__restrict__ part) {
  __shared__ double sh[kRedThreads / 64];
  const int c = blockIdx.x, j = blockIdx.y;
  const double *v = V + (int64_t)j * ldv;
  double acc = 0;
  for (int64_t i = cbeg[c] + threadIdx.x; i < cbeg[c + 1]; i += kRedThreads) acc += v[i] * w[i];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[(int64_t)j * gridDim.x + blockIdx.x] = r;
}
__global__ void __launch_bounds__(kRedThreads) k_finish_seg(int nch, const int *__restrict__ sch_off, int ldo,
                                                            const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double sh[kRedThreads / 64];
  const int s = blockIdx.x, j = blockIdx.y;
  double acc = 0;
  for (int c = sch_off[s] + threadIdx.x; c < sch_off[s + 1]; c += kRedThreads) acc += part[(int64_t)j * nch + c];
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) out[(int64_t)s * ldo + j] = r;
}
// w -= sum_j V_j h[part][j]
__global__ void k_gemv_sub_seg(int64_t n, int64_t ldv, int nv, const double *__restrict__ V,
                               const double *__restrict__ h, int ldh, const int *__restrict__ seg,
                               double *__restrict__ w) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *hs = h + (int64_t)seg[i] * ldh;
  double acc = w[i];
  for (int j = 0; j < nv; ++j) acc -= V[(int64_t)j * ldv + i] * hs[j];
  w[i] = acc;
}
// x += sum_j V_j y[part][j]
__global__ void k_gemv_add_seg(int64_t n, int64_t ldv, int nv, const double *__restrict__ V,
                               const double *__restrict__ y, int ldy, const int *__restrict__ seg,
                               double *__restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *ys = y + (int64_t)seg[i] * ldy;
  double acc = x[i];
  for (int j = 0; j < nv; ++j) acc += V[(int64_t)j * ldv + i] * ys[j];
  x[i] = acc;
}
// Arnoldi normalisation per part (nrm2 holds squared norms); parts that are
// no longer active get zero vectors, so they stay inert
__global__ void k_scale_inv_seg(int64_t n, const int *__restrict__ seg, const double *__restrict__ nrm2,
                                const uint8_t *__restrict__ active, double *__restrict__ v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = seg[i];
  const double q = nrm2[s];
  v[i] = (active[s] && q > 0) ? v[i] / sqrt(q) : 0.0;
}

// ===========================================================================
// Host orchestration
// ===========================================================================
static inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// Device allocations made so far by this library (a captured operator graph
// holds the addresses it was captured with: one captured before the last
// allocation is captured again).
static int64_t g_alloc_epoch = 0;
template <typename T>
static void up(DevBuf &b, const std::vector<T> &v) {
  b.release();
  const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
  HIPCHK(hipMalloc(&b.p, bytes));
  b.bytes = bytes;
  ++g_alloc_epoch;
  if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}
static void alloc(DevBuf &b, size_t bytes) {
  b.release();
  bytes = std::max<size_t>(bytes, 16);
  HIPCHK(hipMalloc(&b.p, bytes));
  b.bytes = bytes;
  ++g_alloc_epoch;
}
template <typename D, typename S>
static std::vector<D> cast_vec(const std::vector<S> &v) {
  std::vector<D> o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = (D)v[i];
  return o;
}

void Engine::init_device(int dev, void *st) {
  device = dev;
  HIPCHK(hipSetDevice(dev));
  if (st) {
    stream = (hipStream_t)st;
    own_stream = false;
  } else {
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    own_stream = true;
  }
  upload_legendre_tables();
  ps_scratch_init(dev);
}

Engine::~Engine() {
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (h_ready) (void)hipEventDestroy(h_ready);
  if (h_pinned) (void)hipHostFree(h_pinned);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
  drop_graphs();
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side_stream) (void)hipStreamDestroy(side_stream);
}

void Engine::drop_graphs() {
  for (int f = 0; f < 2; ++f) {
    if (op_graph[f]) (void)hipGraphExecDestroy(op_graph[f]);
    op_graph[f] = nullptr;
    op_graph_key[f].clear();
  }
}

// Phase events on the handle's stream (no synchronisation between kernels):
// 10 clear ME 0 P2M 1 M2M 2, 3 SpMV 4 clear LE 5 M2L [11 slot sum] 6 M2P 7
// L2L 8 L2P 9.
void Engine::mark(int i) {
  if (!time_phases) return;
  if (!ev[i]) HIPCHK(hipEventCreate(&ev[i]));
  HIPCHK(hipEventRecord(ev[i], stream));
}

void Engine::collect_phase_times() {
  if (!time_phases) return;
  HIPCHK(hipEventSynchronize(ev[9]));
  auto el = [&](int a, int b) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev[a], ev[b]));
    return (double)ms;
  };
  last_ms[0] = el(3, 4);  // spmv
  last_ms[1] = el(0, 1);  // p2m
  last_ms[2] = el(1, 2);  // m2m
  last_ms[3] = el(7, 8);  // l2l
  last_ms[4] = slot_marked ? el(5, 11) : el(5, 6);  // m2l (translations)
  last_ms[8] = slot_marked ? el(11, 6) : 0.0;       // m2l slot sums
  slot_marked = false;
  last_ms[5] = el(6, 7);  // m2p
  last_ms[6] = el(8, 9);  // l2p
  last_ms[7] = el(10, 9);  // whole matvec
}

void Engine::sync() { HIPCHK(hipStreamSynchronize(stream)); }

void Engine::upload_mesh_and_tree() {
  HIPCHK(hipSetDevice(device));
  up(d_xyz, mesh.xyz);
  up(d_geom, mesh.geom);
  up(d_cdofs, mesh.cdofs);
  up(d_dbl_ptr, mesh.dbl_ptr);
  up(d_dbl_idx, mesh.dbl_idx);
  up(d_dir, mesh.dir);
  up(d_pmin, tree.pmin);
  up(d_delta, tree.delta);
  up(d_nchild, tree.nchild);
  up(d_child, tree.child);
  up(d_octant, tree.octant);
  up(d_node_beg, tree.node_beg);
  up(d_node_cnt, tree.node_cnt);
  up(d_nperm, tree.nperm);
  up(d_quad_beg, tree.quad_beg);
  up(d_quad_cnt, tree.quad_cnt);
  up(d_qperm, tree.qperm);
  up(d_parent, tree.parent);
  up(d_alpha, std::vector<double>(n, 0.0));
  alloc(d_sp, n * 8);
  alloc(d_sd, n * 8);
  alloc(d_outN, n * 8);
  alloc(d_outD, n * 8);
  alloc(d_in, n * 8);
  alloc(d_out, n * 8);
  alloc(d_red, (size_t)(nblk(n, kRedChunk) * 128 + 256) * 8);
}

void Engine::upload_near_plan() {
  up(d_row_leaf, near.row_leaf);
  up(d_row_beg, near.row_beg);
  up(d_rows, near.rows);
  up(d_col_beg, near.col_beg);
  up(d_cols, near.cols);
  up(d_val_beg, near.val_beg);
  up(d_ent_beg, near.ent_beg);
  up(d_ent_cell, near.ent_cell);
  up(d_ent_mask, near.ent_mask);
  up(d_ent_near, near.ent_near);
  up(d_ent_colpos, near.ent_colpos);
  up(d_leaf_block, near.leaf_block);
}

template <typename T>
static void upload_far_t(Engine &e, FarDev<T> &F) {
  up(F.m2m, cast_vec<T>(e.far.m2m_ops));
  up(F.l2l, cast_vec<T>(e.far.l2l_ops));
  up(F.m2l, cast_vec<T>(e.far.m2l_ops));
  alloc(F.me, (size_t)e.tree.nb * 2 * e.kp * sizeof(T));
  alloc(F.le, (size_t)e.tree.nb * 2 * e.kp * sizeof(T));
  HIPCHK(hipMemset(F.me.p, 0, F.me.bytes));
  HIPCHK(hipMemset(F.le.p, 0, F.le.bytes));
}

void Engine::upload_quad(bool want_f32) {
  up(f64.qx, qd.x);
  up(f64.qn, qd.n);
  up(f64.qw, qd.jxw);
  up(f64.qsh, qd.sh);
  have_f32 = want_f32;
}

static void upload_chunks(ChunkDev &d, const ChunkList &c) {
  up(d.op, c.op);
  up(d.beg, c.beg);
  up(d.cnt, c.cnt);
  up(d.op1, c.op1);
  up(d.beg1, c.beg1);
  up(d.cnt1, c.cnt1);
  up(d.tgt, c.tgt);
  up(d.src, c.src);
}

void Engine::upload_far(bool want_f32) {
  {
    std::vector<uint8_t> is_leaf(std::max(tree.nb, 1), 0);
    for (int b : far.p2m_leaf) is_leaf[b] = 1;
    up(d_is_leaf, is_leaf);
  }
  me_lo = le_lo = -1;  // recomputed from this plan on first use
  up(d_p2m_leaf, far.p2m_leaf);
  up(d_p2m_col_beg, far.p2m_col_beg);
  up(d_p2m_col_dof, far.p2m_col_dof);
  upload_chunks(m2m_cl, far.m2m);
  upload_chunks(m2m_g_cl, far.m2m_g);
  upload_chunks(l2l_g_cl, far.l2l_g);
  upload_chunks(m2l_cl, far.m2l);
  upload_chunks(l2l_cl, far.l2l);
  up(d_m2p_node, far.m2p_item_node);
  up(d_m2p_isrc, far.m2p_item_src);
  upload_far_t<double>(*this, f64);
  if (want_f32) upload_far_t<float>(*this, f32);
}

// Unit multipoles of every leaf (FP64, rounded once for the FP32 operator).
void Engine::build_unit_multipoles(bool want_f32) {
  const int64_t ncol = (int64_t)far.p2m_col_dof.size();
  const int64_t nu = std::max<int64_t>(ncol, 1) * 2 * kp;
  alloc(f64.unit, (size_t)nu * sizeof(double));
  HIPCHK(hipMemsetAsync(f64.unit.p, 0, f64.unit.bytes, stream));
  const int npts = (int)far.p2m_qpt.size();
  if (npts > 0) {
    // leaves whose [column][2 kp] block fits in LDS: one workgroup each, sums
    // in LDS; those whose [column][kp] half fits: two workgroups each, one
    // per kernel; the others by global atomics
    const int nleaf = (int)far.p2m_leaf.size();
    const size_t col_bytes = (size_t)2 * kp * sizeof(double);
    constexpr size_t kLdsMax = 160 * 1024 - 1024;
    std::vector<int32_t> qbeg(nleaf + 1, 0), fit, fit_half, big_qpt, big_qleaf, big_qcol;
    for (int i = 0; i < npts; ++i) qbeg[far.p2m_qleaf[i] + 1]++;
    for (int li = 0; li < nleaf; ++li) qbeg[li + 1] += qbeg[li];
    size_t max_bytes = 0, max_half = 0;
    for (int li = 0; li < nleaf; ++li) {
      const size_t bytes = (size_t)(far.p2m_col_beg[li + 1] - far.p2m_col_beg[li]) * col_bytes;
      if (bytes <= kLdsMax) {
        if (qbeg[li + 1] > qbeg[li]) fit.push_back(li);
        max_bytes = std::max(max_bytes, bytes);
      } else if (bytes / 2 <= kLdsMax) {
        if (qbeg[li + 1] > qbeg[li]) fit_half.push_back(li);
        max_half = std::max(max_half, bytes / 2);
      } else {
        for (int i = qbeg[li]; i < qbeg[li + 1]; ++i) {
          big_qpt.push_back(far.p2m_qpt[i]);
          big_qleaf.push_back(li);
          for (int j = 0; j < dpc; ++j) big_qcol.push_back(far.p2m_qcol[(size_t)i * dpc + j]);
        }
      }
    }
    DevBuf qpt, qcol, dqbeg, dfit, dfit_half;
    up(qpt, far.p2m_qpt);
    up(qcol, far.p2m_qcol);
    up(dqbeg, qbeg);
    up(dfit, fit);
    up(dfit_half, fit_half);
    if (!fit.empty() || !fit_half.empty())
      HIPCHK(hipFuncSetAttribute((const void *)k_p2m_unit_leaf, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)std::max<size_t>(std::max(max_bytes, max_half), 16)));
    if (!fit.empty()) {
      hipLaunchKernelGGL(k_p2m_unit_leaf, dim3((unsigned)fit.size()), dim3(256), max_bytes, stream, p, kt, kp,
                         dfit.as<int>(), d_p2m_leaf.as<int>(), dqbeg.as<int>(), d_p2m_col_beg.as<int>(),
                         qpt.as<int>(), qcol.as<int>(), dpc, f64.qx.as<double>(), f64.qn.as<double>(),
                         f64.qw.as<double>(), f64.qsh.as<double>(), d_pmin.as<double>(), d_delta.as<double>(),
                         f64.unit.as<double>(), -1);
      HIPCHK(hipGetLastError());
    }
    for (int half = 0; half < 2 && !fit_half.empty(); ++half) {
      hipLaunchKernelGGL(k_p2m_unit_leaf, dim3((unsigned)fit_half.size()), dim3(256), max_half, stream, p, kt, kp,
                         dfit_half.as<int>(), d_p2m_leaf.as<int>(), dqbeg.as<int>(), d_p2m_col_beg.as<int>(),
                         qpt.as<int>(), qcol.as<int>(), dpc, f64.qx.as<double>(), f64.qn.as<double>(),
                         f64.qw.as<double>(), f64.qsh.as<double>(), d_pmin.as<double>(), d_delta.as<double>(),
                         f64.unit.as<double>(), half);
      HIPCHK(hipGetLastError());
    }
    if (!big_qpt.empty()) {
      DevBuf bqpt, bqleaf, bqcol;
      up(bqpt, big_qpt);
      up(bqleaf, big_qleaf);
      up(bqcol, big_qcol);
      const int nb = (int)big_qpt.size();
      hipLaunchKernelGGL(k_p2m_unit, dim3(nblk(nb, 256)), dim3(256), 0, stream, nb, p, kt, kp, bqpt.as<int>(),
                         bqleaf.as<int>(), d_p2m_leaf.as<int>(), bqcol.as<int>(), dpc, f64.qx.as<double>(),
                         f64.qn.as<double>(), f64.qw.as<double>(), f64.qsh.as<double>(), d_pmin.as<double>(),
                         d_delta.as<double>(), f64.unit.as<double>());
      HIPCHK(hipGetLastError());
      sync();  // the index copies are freed on return
    }
    sync();  // qpt, qcol, dqbeg, dfit, dfit_half are freed on return
  }
  // the operator's unit multipoles (U1 on Neumann, -U2 on Dirichlet DoFs),
  // compact: the real parts, then the imaginary parts of m >= 1 (those of
  // m = 0 vanish for real sources), padded to a multiple of 4
  {
    std::vector<int32_t> cmap;
    for (int i = 0; i < kt; ++i) cmap.push_back(i);
    for (int nn = 0; nn <= p; ++nn)
      for (int m = 1; m <= nn; ++m) cmap.push_back(kt + nn * (nn + 1) / 2 + m);
    while (cmap.size() % 4) cmap.push_back(-1);
    kpc = (int)cmap.size();
    up(d_cmap, cmap);
  }
  const int64_t nou = std::max<int64_t>(ncol, 1) * kpc;
  alloc(f64.opunit, (size_t)nou * sizeof(double));
  if (ncol > 0) {
    hipLaunchKernelGGL(k_op_unit, dim3((unsigned)std::min<int64_t>(nblk(ncol * kpc, 256), 8192)), dim3(256), 0,
                       stream, ncol, kp, kpc, d_cmap.as<int>(), d_p2m_col_dof.as<int>(), d_dir.as<uint8_t>(),
                       f64.unit.as<double>(), f64.opunit.as<double>());
    HIPCHK(hipGetLastError());
  }
  if (want_f32) {
    alloc(f32.unit, (size_t)nu * sizeof(float));
    hipLaunchKernelGGL(k_round_f32, dim3((unsigned)std::min<int64_t>(nblk(nu, 256), 8192)), dim3(256), 0, stream, nu,
                       f64.unit.as<double>(), f32.unit.as<float>());
    alloc(f32.opunit, (size_t)nou * sizeof(float));
    hipLaunchKernelGGL(k_round_f32, dim3((unsigned)std::min<int64_t>(nblk(nou, 256), 8192)), dim3(256), 0, stream,
                       nou, f64.opunit.as<double>(), f32.opunit.as<float>());
    HIPCHK(hipGetLastError());
  }
  sync();
}

// Singular rules for every local support point js, with the shape values and
// gradients at the rule points: Telles in each direction
// (QTelles<2>(sing_quad_order, support_point), bem_fma.cc:127-131) or the
// Duffy split (QSplit<2>(QDuffy(sing_quad_order, 1), support_point),
// bem_problem.cc:213-229).
void Engine::assemble_near(bool want_f32) {
  const int ps = 1 + 3 * dpc;
  std::vector<std::vector<double>> rules(dpc);
  int maxq = 0;
  std::vector<int32_t> cnt(dpc);
  for (int js = 0; js < dpc; ++js) {
    double sx, sy;
    unit_support_point(dpc, js, sx, sy);
    Quadrature2D r;
    if (sing_rule == 1) {
      r = duffy_split_rule(sing_order, sx, sy);
    } else {
      std::vector<double> x1, w1, x2, w2;
      telles_rule01(sing_order, sx, x1, w1);
      telles_rule01(sing_order, sy, x2, w2);
      r = tensor_rule(x1, w1, x2, w2);
    }
    const int nqs = (int)r.w.size();
    cnt[js] = nqs;
    maxq = std::max(maxq, nqs);
    rules[js].resize((size_t)nqs * ps);
    for (int q = 0; q < nqs; ++q) {
      double *pt = &rules[js][(size_t)q * ps];
      pt[0] = r.w[q];
      lagrange_q(dpc, r.pts[2 * q], r.pts[2 * q + 1], pt + 1, pt + 1 + dpc, pt + 1 + 2 * dpc);
    }
  }
  sing_nq = maxq;
  std::vector<double> flat((size_t)dpc * maxq * ps, 0.0);
  for (int js = 0; js < dpc; ++js) std::copy(rules[js].begin(), rules[js].end(), flat.begin() + (size_t)js * maxq * ps);
  up(d_sing, flat);
  DevBuf d_scnt;
  up(d_scnt, cnt);

  const int64_t nv = std::max<int64_t>(near.nnz, 1) * 2;
  alloc(f64.vals, (size_t)nv * sizeof(double));
  HIPCHK(hipMemsetAsync(f64.vals.p, 0, f64.vals.bytes, stream));
  const int64_t total = near.work_beg.back();
  if (near.nleaf > 0 && total > 0) {
    // G lanes per item: the quad points per cell (at most 64, the entry masks' width), rounded up
    if (nq > 64 || dpc > 9) throw std::runtime_error("near assembly: more than 64 quad points or 9 DoFs per cell");
    const int G = nq <= 16 ? 16 : nq <= 32 ? 32 : 64;
    const unsigned grid = (unsigned)near.row_beg[near.nleaf];  // a workgroup per row
#define PBFMM_ASM(GG)                                                                                              \
  hipLaunchKernelGGL(k_assemble_near<GG>, dim3(grid), dim3(256), 0, stream, near.nleaf,                            \
                     d_row_beg.as<int>(), d_rows.as<int>(), d_col_beg.as<int>(), d_ent_beg.as<int>(),                \
                     d_ent_cell.as<int>(), d_ent_mask.as<uint64_t>(), d_ent_near.as<uint8_t>(),                    \
                     d_ent_colpos.as<int>(), d_val_beg.as<int64_t>(), d_cdofs.as<int>(), dpc, nq,                  \
                     d_dbl_ptr.as<int>(), d_dbl_idx.as<int>(), d_xyz.as<double>(), f64.qx.as<double>(),            \
                     f64.qn.as<double>(), f64.qw.as<double>(), f64.qsh.as<double>(), d_geom.as<double>(),          \
                     d_sing.as<double>(), d_scnt.as<int>(), maxq * ps, f64.vals.as<double>())
    if (G == 16) PBFMM_ASM(16);
    else if (G == 32) PBFMM_ASM(32);
    else PBFMM_ASM(64);
#undef PBFMM_ASM
    HIPCHK(hipGetLastError());
  }
  // the operator's single-value near field (N on Neumann, -D on Dirichlet columns)
  const int64_t nop = std::max<int64_t>(near.nnz, 1);
  alloc(f64.opvals, (size_t)nop * sizeof(double));
  if (near.nleaf > 0) {
    hipLaunchKernelGGL(k_op_near, dim3(near.nleaf), dim3(256), 0, stream, near.nleaf, d_row_beg.as<int>(),
                       d_col_beg.as<int>(), d_cols.as<int>(), d_val_beg.as<int64_t>(), d_dir.as<uint8_t>(),
                       f64.vals.as<double>(), f64.opvals.as<double>());
    HIPCHK(hipGetLastError());
  }
  if (want_f32) {
    alloc(f32.vals, (size_t)nv * sizeof(float));
    hipLaunchKernelGGL(k_round_f32, dim3((unsigned)std::min<int64_t>(nblk(nv, 256), 8192)), dim3(256), 0, stream, nv,
                       f64.vals.as<double>(), f32.vals.as<float>());
    alloc(f32.opvals, (size_t)nop * sizeof(float));
    hipLaunchKernelGGL(k_round_f32, dim3((unsigned)std::min<int64_t>(nblk(nop, 256), 8192)), dim3(256), 0, stream,
                       nop, f64.opvals.as<double>(), f32.opvals.as<float>());
    // and in row pairs, each leaf at an even (8-byte aligned) offset
    std::vector<int64_t> pvb(near.nleaf + 1, 0);
    for (int l = 0; l < near.nleaf; ++l) {
      const int64_t sz = (int64_t)(near.row_beg[l + 1] - near.row_beg[l]) * (near.col_beg[l + 1] - near.col_beg[l]);
      pvb[l + 1] = pvb[l] + ((sz + 1) & ~(int64_t)1);
    }
    up(d_pval_beg, pvb);
    alloc(f32.opvals_p, (size_t)std::max<int64_t>(pvb[near.nleaf], 2) * sizeof(float));
    if (near.nleaf > 0)
      hipLaunchKernelGGL(k_round_f32_pairs, dim3(near.nleaf), dim3(256), 0, stream, near.nleaf, d_row_beg.as<int>(),
                         d_col_beg.as<int>(), d_val_beg.as<int64_t>(), d_pval_beg.as<int64_t>(),
                         f64.opvals.as<double>(), f32.opvals_p.as<float>());
    HIPCHK(hipGetLastError());
  }
  // and in column vectors, per precision
  auto cols_copy = [&](auto &F) {
    using T = typename std::remove_reference_t<decltype(F)>::value_type;
    constexpr int CW = 16 / sizeof(T);
    std::vector<int64_t> qvb(near.nleaf + 1, 0);
    for (int l = 0; l < near.nleaf; ++l)
      qvb[l + 1] = qvb[l] + (int64_t)(near.row_beg[l + 1] - near.row_beg[l]) *
                                ((near.col_beg[l + 1] - near.col_beg[l] + CW - 1) / CW * CW);
    up(F.qval_beg, qvb);
    alloc(F.opvals_q, (size_t)std::max<int64_t>(qvb[near.nleaf], CW) * sizeof(T));
    if (near.nleaf > 0)
      hipLaunchKernelGGL(k_round_cols<T>, dim3(near.nleaf), dim3(256), 0, stream, near.nleaf, d_row_beg.as<int>(),
                         d_col_beg.as<int>(), d_val_beg.as<int64_t>(), F.qval_beg.template as<int64_t>(),
                         f64.opvals.as<double>(), F.opvals_q.template as<T>());
    HIPCHK(hipGetLastError());
  };
  cols_copy(f64);
  if (want_f32) cols_copy(f32);
  tune_spmv<double, 2>();
  tune_spmv<double, 1>();
  if (want_f32) {
    tune_spmv<float, 2>();
    tune_spmv<float, 1>();
  }
  sync();
}

template <typename T>
static FarDev<T> &far_of(Engine &e);
template <>
FarDev<double> &far_of<double>(Engine &e) { return e.f64; }
template <>
FarDev<float> &far_of<float>(Engine &e) { return e.f32; }

// Runs STMT with PP the truncation order p (1..10) as a compile-time constant.
#define PBFMM_BY_ORDER(p, STMT)                                              \
  switch (p) {                                                               \
    case 1: { constexpr int PP = 1; STMT; } break;                           \
    case 2: { constexpr int PP = 2; STMT; } break;                           \
    case 3: { constexpr int PP = 3; STMT; } break;                           \
    case 4: { constexpr int PP = 4; STMT; } break;                           \
    case 5: { constexpr int PP = 5; STMT; } break;                           \
    case 6: { constexpr int PP = 6; STMT; } break;                           \
    case 7: { constexpr int PP = 7; STMT; } break;                           \
    case 8: { constexpr int PP = 8; STMT; } break;                           \
    case 9: { constexpr int PP = 9; STMT; } break;                           \
    case 10: { constexpr int PP = 10; STMT; } break;                         \
    default: throw std::runtime_error("trunc_order outside 1..10");          \
  }

// OUT[tgt] += ops[class] . IN[src] for the chunks [c0, c1) of a chunk list
// (chunks of 32 pairs for NK = 2, of 64 pairs for NK = 1).
template <typename T, int KIND, int NK>
static void launch_translate(Engine &e, hipStream_t st, const ChunkDev &cl, int c0, int c1, const T *ops,
                             const T *IN, T *OUT, const int *slot_pos = nullptr) {
  const int n = c1 - c0;
  if (n <= 0) return;
  const DevBuf &op = NK == 2 ? cl.op : cl.op1, &beg = NK == 2 ? cl.beg : cl.beg1, &cnt = NK == 2 ? cl.cnt : cl.cnt1;
  // the slot M2L (KIND 3) reads each pair's slot where the others read its target
  const int *tgt = KIND == 3 ? slot_pos : cl.tgt.as<int>();
  if (!tgt) throw std::runtime_error("slot M2L without slot positions");
  const TrArgs<T> tr{op.as<int>() + c0, beg.as<int>() + c0, cnt.as<int>() + c0, tgt, cl.src.as<int>(),
                     ops, IN, OUT};
#define PBFMM_TR(PP)                                                                                         \
  case PP:                                                                                                   \
    hipLaunchKernelGGL((k_translate<T, PP, KIND, NK>), dim3(xcd_blocks(n)), dim3(TrLayout<PP, T>::NT), 0, st, tr, \
                       n);                                                                                   \
    break;
  switch (e.p) {
    PBFMM_TR(1) PBFMM_TR(2) PBFMM_TR(3) PBFMM_TR(4) PBFMM_TR(5)
    PBFMM_TR(6) PBFMM_TR(7) PBFMM_TR(8) PBFMM_TR(9) PBFMM_TR(10)
    default: throw std::runtime_error("trunc_order outside 1..10");
  }
#undef PBFMM_TR
}

// Workgroups of the SpMV side work for a launch of NT threads.
template <int SPMV, typename T>
static int spmv_side_blocks(const SideArgs<T> &sd, int nt) {
  if (SPMV < 0 || sd.count <= 0) return 0;
  if (SPMV == 0) return sd.per_xcd * kXcd;
  return (sd.count + nt / 64 - 1) / (nt / 64);
}

// Chunks [c0, c1) of an M2M (KIND 0), M2L (1) or L2L (2) chunk list of the
// single-value pipeline, with side work sd in the same launch
// (k_translate_side).
template <typename T, int KIND, int SPMV, bool M2P>
static void launch_translate_side(Engine &e, const ChunkDev &cl, int c0, int c1, const T *ops, const T *IN, T *OUT,
                                  SideArgs<T> sd) {
  const int nch = std::max(0, c1 - c0);
  TrArgs<T> tr{};
  if (nch > 0)
    tr = TrArgs<T>{cl.op1.as<int>() + c0, cl.beg1.as<int>() + c0, cl.cnt1.as<int>() + c0, cl.tgt.as<int>(),
                   cl.src.as<int>(), ops, IN, OUT};
#define PBFMM_TRS(PP)                                                                                         \
  case PP: {                                                                                                  \
    constexpr int NT = TrLayout<PP, T>::NT;                                                                   \
    sd.spmv_blocks = spmv_side_blocks<SPMV>(sd, NT);                                                          \
    const int64_t mb = M2P ? (sd.m2p_hi - sd.m2p_lo + NT - 1) / NT : 0;                                      \
    const int64_t nb = xcd_blocks(nch) + sd.spmv_blocks + mb;                                                 \
    if (nb > 0)                                                                                               \
      hipLaunchKernelGGL((k_translate_side<T, PP, KIND, SPMV, M2P>), dim3((unsigned)nb), dim3(NT), 0, e.stream, \
                         nch, tr, sd);                                                                        \
    break;                                                                                                    \
  }
  switch (e.p) {
    PBFMM_TRS(1) PBFMM_TRS(2) PBFMM_TRS(3) PBFMM_TRS(4) PBFMM_TRS(5)
    PBFMM_TRS(6) PBFMM_TRS(7) PBFMM_TRS(8) PBFMM_TRS(9) PBFMM_TRS(10)
    default: throw std::runtime_error("trunc_order outside 1..10");
  }
#undef PBFMM_TRS
}

// The SpMV variant on the operator's single values in column vectors
// (k_spmv_cols): 6 for FP32 (after the row-pair forms 3..5), 3 for FP64.
template <typename T>
constexpr int kColsVariant = sizeof(T) == 4 ? 6 : 3;
// variant 3..5 (FP32): the operator's values in row pairs, at their own
// (even) per-leaf offsets; kColsVariant: in column vectors
template <typename T, int NK>
static NearArgs<T> near_args(Engine &e, const double *phi, const double *dphi, double *outN, double *outD,
                             int variant) {
  const bool cols = NK == 1 && variant == kColsVariant<T>;
  const bool paired = kPairedF32<T, NK> && variant >= 3 && !cols;
  const FarDev<T> &F = far_of<T>(e);
  const T *vals = (NK == 2 ? F.vals : cols ? F.opvals_q : paired ? F.opvals_p : F.opvals).template as<T>();
  const int64_t *vb = cols     ? F.qval_beg.template as<int64_t>()
                      : paired ? e.d_pval_beg.as<int64_t>()
                               : e.d_val_beg.as<int64_t>();
  return NearArgs<T>{e.d_row_beg.as<int>(), e.d_rows.as<int>(), e.d_col_beg.as<int>(), e.d_cols.as<int>(),
                     vb, vals, phi, dphi, outN, outD, 0};
}

template <typename T, int NK>
void Engine::launch_spmv(int variant, const double *phi, const double *dphi, double *outN, double *outD, int R,
                         hipStream_t st) {
  NearArgs<T> a = near_args<T, NK>(*this, phi, dphi, outN, outD, variant);
  a.stride = n;  // R right-hand sides, n apart (blockIdx.y)
  if (!st) st = stream;
  if constexpr (NK == 1) {  // column vectors (k_spmv_cols)
    if (variant == kColsVariant<T>) {
      const int per_xcd = (near.nleaf + kXcd - 1) / kXcd;
      hipLaunchKernelGGL((k_spmv_cols<T, 8>), dim3(per_xcd * kXcd, R), dim3(256), 0, st, near.nleaf, per_xcd, a);
      return;
    }
  }
  if constexpr (kPairedF32<T, NK>) {  // 3..5: the kernels of 0..2 on the row pairs
    if (variant == 3) {
      const int per_xcd = (near.nleaf + kXcd - 1) / kXcd;
      hipLaunchKernelGGL((k_spmv_lds<T, NK, true>), dim3(per_xcd * kXcd, R), dim3(256), 0, st, near.nleaf, per_xcd,
                         a);
      return;
    }
    if (variant == 4 || variant == 5) {
      if (variant == 4)
        hipLaunchKernelGGL((k_spmv_wave<T, NK, 2, true>), dim3((near.nleaf + 3) / 4, R), dim3(256), 0, st,
                           near.nleaf, a);
      else
        hipLaunchKernelGGL((k_spmv_wave<T, NK, 4, true>), dim3((near.nleaf + 3) / 4, R), dim3(256), 0, st,
                           near.nleaf, a);
      return;
    }
  }
  variant %= 3;
  if (variant == 0) {
    const int per_xcd = (near.nleaf + kXcd - 1) / kXcd;
    hipLaunchKernelGGL((k_spmv_lds<T, NK>), dim3(per_xcd * kXcd, R), dim3(256), 0, st, near.nleaf, per_xcd, a);
  } else if (variant == 1) {
    hipLaunchKernelGGL((k_spmv_wave<T, NK, 2>), dim3((near.nleaf + 3) / 4, R), dim3(256), 0, st, near.nleaf, a);
  } else {
    hipLaunchKernelGGL((k_spmv_wave<T, NK, 4>), dim3((near.nleaf + 3) / 4, R), dim3(256), 0, st, near.nleaf, a);
  }
}

// Times the SpMV kernels (workgroup per leaf; wavefront per leaf with 2 or 4
// rows in flight) on the assembled near field (HIP events, scratch vectors)
// and keeps the fastest for this precision and storage.
template <typename T, int NK>
void Engine::tune_spmv() {
  if (near.nleaf == 0) return;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  float best = 0;
  int pick = 0;
  // one warm-up and three timed launches per variant (the variants differ by
  // 5-30% at the sizes measured, the repetitions by ~1%)
  for (int v = 0; v < (NK == 2 ? 3 : kColsVariant<T> + 1); ++v) {
    launch_spmv<T, NK>(v, d_sp.as<double>(), d_sd.as<double>(), d_outN.as<double>(), d_outD.as<double>());
    HIPCHK(hipEventRecord(e0, stream));
    for (int rep = 0; rep < 3; ++rep)
      launch_spmv<T, NK>(v, d_sp.as<double>(), d_sd.as<double>(), d_outN.as<double>(), d_outD.as<double>());
    HIPCHK(hipEventRecord(e1, stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (v == 0 || ms < best) {
      best = ms;
      pick = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  spmv_variant[sizeof(T) == 4 ? 1 : 0][NK - 1] = pick;
}

// Chunk range of level l in a chunk list, for the NK-expansion chunking.
template <int NK>
static inline int lvl(const ChunkList &c, int l) { return NK == 2 ? c.level_off[l] : c.level_off1[l]; }
template <int NK>
static inline int nchunks(const ChunkList &c) { return (int)(NK == 2 ? c.op.size() : c.op1.size()); }

// Upward pass: leaf P2M, then M2M level by level.  NK = 2 keeps the two
// kernels' expansions ([block][2][kp]); NK = 1 one combined expansion
// ME1(phi) - ME2(dphi) per block ([block][kp]).
template <typename T, int NK>
void Engine::generate(const double *phi, const double *dphi) {
  FarDev<T> &F = far_of<T>(*this);
  const int nleaf = (int)far.p2m_leaf.size();
  mark(10);
  const int64_t nme = (int64_t)tree.nb * NK * kp;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)std::min<int64_t>(nblk(nme, 256), 4096)), dim3(256), 0, stream,
                     F.me.template as<T>(), nme, T(0));
  mark(0);
  if (nleaf > 0)
    hipLaunchKernelGGL((k_p2m<T, NK>), dim3(nleaf), dim3(256), 0, stream, nleaf, d_p2m_leaf.as<int>(),
                       d_p2m_col_beg.as<int>(), d_p2m_col_dof.as<int>(),
                       (NK == 2 ? F.unit : F.opunit).template as<T>(), kp, kpc, d_cmap.as<int>(), phi, dphi,
                       F.me.template as<T>(), (int64_t)0, (int64_t)0);
  mark(1);
  for (int l = L; l >= 1; --l)
    launch_translate<T, 0, NK>(*this, stream, m2m_cl, lvl<NK>(far.m2m, l), lvl<NK>(far.m2m, l + 1),
                               F.m2m.template as<T>(), F.me.template as<T>(), F.me.template as<T>());
  mark(2);
  HIPCHK(hipGetLastError());
}

// Downward pass and near field.  NK = 2: outN = N.phi, outD = D.dphi;
// NK = 1: outN = N.phi - D.dphi (outD unused).
template <typename T, int NK>
void Engine::products(const double *phi, const double *dphi, double *outN, double *outD) {
  FarDev<T> &F = far_of<T>(*this);
  const int nrows = (int)near.rows.size();
  mark(3);
  if (near.nleaf > 0) launch_spmv<T, NK>(spmv_variant[sizeof(T) == 4 ? 1 : 0][NK - 1], phi, dphi, outN, outD);
  mark(4);
  const int64_t nle = (int64_t)tree.nb * NK * kp;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)std::min<int64_t>(nblk(nle, 256), 4096)), dim3(256), 0, stream,
                     F.le.template as<T>(), nle, T(0));
  mark(5);
  if (NK == 1 && side_fusion[sizeof(T) == 4 ? 1 : 0] >= 0 && (side_fusion[sizeof(T) == 4 ? 1 : 0] & 4))
    launch_m2l_slots<T>(F.m2l.template as<T>(), F.me.template as<T>(), F.le.template as<T>());
  else
    launch_translate<T, 1, NK>(*this, stream, m2l_cl, 0, nchunks<NK>(far.m2l), F.m2l.template as<T>(),
                               F.me.template as<T>(), F.le.template as<T>());
  mark(6);
  const int64_t nitems = (int64_t)far.m2p_item_node.size();
  if (nitems > 0)
    PBFMM_BY_ORDER(p, hipLaunchKernelGGL((k_m2p<T, NK, PP>), dim3((unsigned)nblk(nitems, 256)), dim3(256), 0, stream, nitems,
                       d_m2p_node.as<int>(), d_m2p_isrc.as<int>(), d_xyz.as<double>(), d_pmin.as<double>(),
                       d_delta.as<double>(), p, kt, kp, F.me.template as<T>(), outN, outD, (int64_t)0, (int64_t)0));
  mark(7);
  for (int l = 2; l <= L; ++l)
    launch_translate<T, 2, NK>(*this, stream, l2l_cl, lvl<NK>(far.l2l, l), lvl<NK>(far.l2l, l + 1),
                               F.l2l.template as<T>(), F.le.template as<T>(), F.le.template as<T>());
  mark(8);
  if (nrows > 0)
    PBFMM_BY_ORDER(p, hipLaunchKernelGGL((k_l2p<T, NK, PP>), dim3(nblk(nrows, 256)), dim3(256), 0, stream, nrows, d_rows.as<int>(),
                       d_row_leaf.as<int>(), d_leaf_block.as<int>(), d_xyz.as<double>(), d_pmin.as<double>(),
                       d_delta.as<double>(), p, kt, kp, F.le.template as<T>(), outN, outD, (int64_t)0, (int64_t)0));
  mark(9);
  HIPCHK(hipGetLastError());
}

// The two-output FMM for device vectors, alpha and the right-hand side.
// While phase timing is off it runs in fewer launches: ME and LE cleared in
// one launch, and M2M / L2L skipped on the levels whose expansions nothing
// reads (consumer_levels).  The reference API's generate + products calls
// keep every level, since pbfmm_get_expansions exports them all.
template <typename T>
void Engine::matvec(const double *phi, const double *dphi, double *outN, double *outD) {
  if (time_phases) {
    generate<T, 2>(phi, dphi);
    products<T, 2>(phi, dphi, outN, outD);
    collect_phase_times();
    return;
  }
  FarDev<T> &F = far_of<T>(*this);
  const int64_t ne = (int64_t)tree.nb * 2 * kp;
  T *me = F.me.template as<T>(), *le = F.le.template as<T>();
  const int nleaf = (int)far.p2m_leaf.size();
  if (nleaf > 0) {  // P2M, with clearing workgroups after the leaves' ones
    P2mClear<T> cl;
    cl.leaf = d_is_leaf.as<uint8_t>();
    cl.le = le;
    cl.ne = ne;
    const unsigned nclear = (unsigned)std::min<int64_t>(nblk(2 * ne, 256), 1024);
    hipLaunchKernelGGL((k_p2m<T, 2>), dim3(nleaf + nclear), dim3(256), 0, stream, nleaf, d_p2m_leaf.as<int>(),
                       d_p2m_col_beg.as<int>(), d_p2m_col_dof.as<int>(), F.unit.template as<T>(), kp, kpc,
                       nullptr, phi, dphi, me, (int64_t)0, (int64_t)0, cl);
  } else {
    hipLaunchKernelGGL(k_clear_operator<T>, dim3((unsigned)std::min<int64_t>(nblk(2 * ne, 256), 4096)), dim3(256),
                       0, stream, me, le, ne, nullptr, (int64_t)0);
  }
  if (me_lo < 0) consumer_levels();
  for (int l = L; l > me_lo; --l)
    launch_translate<T, 0, 2>(*this, stream, m2m_cl, lvl<2>(far.m2m, l), lvl<2>(far.m2m, l + 1),
                              F.m2m.template as<T>(), me, me);
  if (near.nleaf > 0) launch_spmv<T, 2>(spmv_variant[sizeof(T) == 4 ? 1 : 0][1], phi, dphi, outN, outD);
  launch_translate<T, 1, 2>(*this, stream, m2l_cl, 0, nchunks<2>(far.m2l), F.m2l.template as<T>(), me, le);
  const int64_t ni = (int64_t)far.m2p_item_node.size();
  if (ni > 0)
    PBFMM_BY_ORDER(p, hipLaunchKernelGGL((k_m2p<T, 2, PP>), dim3((unsigned)nblk(ni, 256)), dim3(256), 0, stream, ni, d_m2p_node.as<int>(),
                       d_m2p_isrc.as<int>(), d_xyz.as<double>(), d_pmin.as<double>(), d_delta.as<double>(), p, kt,
                       kp, me, outN, outD, (int64_t)0, (int64_t)0));
  for (int l = std::max(2, le_lo + 1); l <= L; ++l)
    launch_translate<T, 2, 2>(*this, stream, l2l_cl, lvl<2>(far.l2l, l), lvl<2>(far.l2l, l + 1),
                              F.l2l.template as<T>(), le, le);
  const int nrows = (int)near.rows.size();
  if (nrows > 0)
    PBFMM_BY_ORDER(p, hipLaunchKernelGGL((k_l2p<T, 2, PP>), dim3(nblk(nrows, 256)), dim3(256), 0, stream, nrows, d_rows.as<int>(),
                       d_row_leaf.as<int>(), d_leaf_block.as<int>(), d_xyz.as<double>(), d_pmin.as<double>(),
                       d_delta.as<double>(), p, kt, kp, le, outN, outD, (int64_t)0, (int64_t)0));
  HIPCHK(hipGetLastError());
}

template <typename T>
void Engine::matvec_combined(const double *x, double *out) {
  generate<T, 1>(x, nullptr);
  products<T, 1>(x, nullptr, out, nullptr);
  collect_phase_times();
}

// The coarsest tree levels whose expansions anything reads.  MEs at level k
// are read by M2L (pairs of level k) and M2P (source blocks of level k); LEs
// at level k are nonzero only from the coarsest M2L level down.  M2M into a
// level above me_lo computes MEs nobody reads, and L2L out of a level above
// le_lo adds zeros (levels 0 and 1 on closed surfaces: their blocks all
// touch).
void Engine::consumer_levels() {
  me_lo = le_lo = L + 1;
  for (int l = 1; l <= L; ++l)
    if (far.m2l_per_level[l] > 0) {
      me_lo = std::min(me_lo, l);
      le_lo = std::min(le_lo, l);
    }
  for (int b : far.m2p_src) me_lo = std::min(me_lo, (int)tree.level[b]);
}

// The slot M2L's buffers on first use: per target the pair indices of the
// M2L chunk list (ascending), and [pair][2 Kt] slots of the precision.
template <typename T>
T *Engine::m2l_slots() {
  FarDev<T> &F = far_of<T>(*this);
  const int64_t npair = (int64_t)far.m2l.tgt.size();
  if (!d_red_beg.p) {
    std::vector<int32_t> cnt(std::max(tree.nb, 1), 0);
    for (int32_t t : far.m2l.tgt) ++cnt[t];
    std::vector<int32_t> tg, beg{0};
    std::vector<int32_t> pos(cnt.size(), -1);
    for (int b = 0; b < (int)cnt.size(); ++b)
      if (cnt[b] > 0) {
        pos[b] = (int32_t)tg.size();
        tg.push_back(b);
        beg.push_back(beg.back() + cnt[b]);
      }
    std::vector<int32_t> slot(npair), fill(beg.begin(), beg.end() - 1);
    for (int64_t i = 0; i < npair; ++i) slot[i] = fill[pos[far.m2l.tgt[i]]]++;
    up(d_red_tgt, tg);
    up(d_red_beg, beg);
    up(d_slot_pos, slot);
    n_red_tgt = (int)tg.size();
  }
  const size_t bytes = (size_t)std::max<int64_t>(npair, 1) * kp * sizeof(T);
  if (F.slots.bytes < bytes) alloc(F.slots, bytes);
  return F.slots.template as<T>();
}

template <typename T>
void Engine::launch_m2l_slots(const T *ops, const T *IN, T *LE) {
  T *slots = m2l_slots<T>();
  launch_translate<T, 3, 1>(*this, stream, m2l_cl, 0, nchunks<1>(far.m2l), ops, IN, slots, d_slot_pos.as<int>());
  mark(11);
  slot_marked = time_phases;
  if (n_red_tgt > 0) {
    const dim3 grid((unsigned)((n_red_tgt + 3) / 4));
    if (sizeof(T) == 4 && kp <= 128)
      hipLaunchKernelGGL((k_m2l_reduce<T, 2>), grid, dim3(256), 0, stream, n_red_tgt, d_red_tgt.as<int>(),
                         d_red_beg.as<int>(), kp, (const T *)slots, LE);
    else
      hipLaunchKernelGGL((k_m2l_reduce<T, 1>), grid, dim3(256), 0, stream, n_red_tgt, d_red_tgt.as<int>(),
                         d_red_beg.as<int>(), kp, (const T *)slots, LE);
  }
  HIPCHK(hipGetLastError());
}

// The single-value pipeline in fewer launches, for vmult: P2M, whose extra
// workgroups clear the other MEs, LE and the M2P accumulator; M2M levels
// L..me_lo + 1; M2L; L2L
// levels le_lo + 1..L (at least 2); L2P, which also adds the M2P
// accumulator and applies vmult's epilogue, dst = out + m2p + L2P + alpha x
// on Neumann DoFs (every DoF is a row of one leaf).  The SpMV writes out and
// M2P adds into the accumulator (d_outD, unused by this pipeline), so the two
// can run beside any launch after P2M (SpMV) or after M2M (M2P), as side
// work (k_translate_side):
//   side 0  both in launches of their own
//   side 1  SpMV beside the deepest M2M level, M2P beside the deepest L2L level
//   side 2  SpMV beside the deepest M2M level, M2P beside M2L
//   side 3  both beside M2L
// The same arithmetic as generate + products + k_vmult_post, with M2P's sum
// added to the SpMV's once instead of term by term.  Phase timing runs that
// unfused sequence instead, so that each phase can be timed.
template <typename T>
void Engine::matvec_fused(const double *x, double *out, double *dst, int side) {
  FarDev<T> &F = far_of<T>(*this);
  const bool async = (side & 8) != 0;  // SpMV on the second stream
  const bool late = (side & 16) != 0;  // ... forked after the M2M levels, beside M2L onwards
  const bool slot = (side & 4) != 0;   // M2L into per-pair slots, summed per target
  const bool m2p_async = (side & 32) != 0;  // M2P after the SpMV on the second stream (late fork)
  const bool m2p_first = (side & 64) != 0;  // ... before it instead, beside M2L
  side &= 3;
  const int nleaf = (int)far.p2m_leaf.size();
  const int64_t ne = (int64_t)tree.nb * kp;
  T *me = F.me.template as<T>(), *le = F.le.template as<T>();
  double *acc = d_outD.as<double>();
  if (nleaf > 0) {  // P2M, with clearing workgroups after the leaves' ones
    P2mClear<T> cl;
    cl.leaf = d_is_leaf.as<uint8_t>();
    cl.le = le;
    cl.acc = acc;
    cl.ne = ne;
    cl.na = n;
    const unsigned nclear = (unsigned)std::min<int64_t>(nblk(2 * ne + n, 256), 1024);
    hipLaunchKernelGGL((k_p2m<T, 1>), dim3(nleaf + nclear), dim3(256), 0, stream, nleaf, d_p2m_leaf.as<int>(),
                       d_p2m_col_beg.as<int>(), d_p2m_col_dof.as<int>(), F.opunit.template as<T>(), kp, kpc,
                       d_cmap.as<int>(), x, nullptr, me, (int64_t)0, (int64_t)0, cl);
  } else {
    hipLaunchKernelGGL(k_clear_operator<T>, dim3((unsigned)std::min<int64_t>(nblk(2 * ne + n, 256), 4096)),
                       dim3(256), 0, stream, me, le, ne, acc, n);
  }
  if (me_lo < 0) consumer_levels();
  const int var_full = spmv_variant[sizeof(T) == 4 ? 1 : 0][0];
  const int var = var_full % 3;  // the kernel form; var_full >= 3 reads the row pairs
  const int nl = near.nleaf;
  const int64_t ni = (int64_t)far.m2p_item_node.size();
  SideArgs<T> sd{};
  // side-fused SpMV work reads the row-major values (its kernels' form)
  sd.near = near_args<T, 1>(*this, x, nullptr, out, nullptr, var);
  sd.lo = 0;
  sd.count = nl;
  sd.per_xcd = (nl + kXcd - 1) / kXcd;
  sd.m2p_lo = 0;
  sd.m2p_hi = ni;
  sd.item_node = d_m2p_node.as<int>();
  sd.item_src = d_m2p_isrc.as<int>();
  sd.xyz = d_xyz.as<double>();
  sd.pmin = d_pmin.as<double>();
  sd.delta = d_delta.as<double>();
  sd.p = p;
  sd.kt = kt;
  sd.kp = kp;
  sd.ME = me;
  sd.m2p_out = acc;
  bool near_done = nl == 0, m2p_done = ni == 0;
  bool join = false;
  auto launch_m2p = [&](hipStream_t st) {
    PBFMM_BY_ORDER(p, hipLaunchKernelGGL((k_m2p<T, 1, PP>), dim3((unsigned)nblk(ni, 256)), dim3(256), 0, st, ni,
                                         d_m2p_node.as<int>(), d_m2p_isrc.as<int>(), d_xyz.as<double>(),
                                         d_pmin.as<double>(), d_delta.as<double>(), p, kt, kp, me, acc, nullptr,
                                         (int64_t)0, (int64_t)0));
    m2p_done = true;
  };
  // the SpMV reads only x, ready when this application starts, and writes
  // out, which only the final L2P reads: on the second stream it runs beside
  // the FMM, from the start or (late) from M2L on, leaving the M2M levels'
  // latency-bound launches the whole chip
  auto fork_spmv = [&]() {
    if (!side_stream) {
      HIPCHK(hipStreamCreateWithFlags(&side_stream, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    HIPCHK(hipEventRecord(ev_fork, stream));
    HIPCHK(hipStreamWaitEvent(side_stream, ev_fork, 0));
    // M2P first (plans 88, 92): compute-bound, beside the M2L's atomics
    if (late && m2p_first && !m2p_done) launch_m2p(side_stream);
    launch_spmv<T, 1>(var_full, x, nullptr, out, nullptr, 1, side_stream);
    // forked after the M2M levels, the multipole expansions are final: M2P
    // (which adds into the accumulator only) follows the SpMV there
    if (late && m2p_async && !m2p_done) launch_m2p(side_stream);
    HIPCHK(hipEventRecord(ev_join, side_stream));
    near_done = join = true;
  };
  if (async && !late && !near_done) fork_spmv();
  const bool fork_later = async && late && !near_done;
  if (side == 0 && !near_done && !fork_later) {
    launch_spmv<T, 1>(var_full, x, nullptr, out, nullptr);
    near_done = true;
  }
  // M2M, the deepest launched level (or group of two levels) carrying the
  // SpMV (sides 1, 2)
  const bool grouped = !far.m2m_g.op1.empty();
  const ChunkDev &mcl = grouped ? m2m_g_cl : m2m_cl;
  const int nlaunch = grouped ? (int)far.m2m_g.level_off1.size() - 1 : L - me_lo;
  for (int s = 0; s < nlaunch; ++s) {  // group s, or M2M from level L - s into L - s - 1
    const int c0 = grouped ? far.m2m_g.level_off1[s] : lvl<1>(far.m2m, L - s);
    const int c1 = grouped ? far.m2m_g.level_off1[s + 1] : lvl<1>(far.m2m, L - s + 1);
    if (c1 <= c0) continue;
    const T *ops = F.m2m.template as<T>();
    if (near_done || side == 3 || fork_later) {
      launch_translate<T, 0, 1>(*this, stream, mcl, c0, c1, ops, me, me);
      continue;
    }
    if (var == 0)
      launch_translate_side<T, 0, 0, false>(*this, mcl, c0, c1, ops, me, me, sd);
    else if (var == 1)
      launch_translate_side<T, 0, 1, false>(*this, mcl, c0, c1, ops, me, me, sd);
    else
      launch_translate_side<T, 0, 2, false>(*this, mcl, c0, c1, ops, me, me, sd);
    near_done = true;
  }
  if (fork_later) fork_spmv();
  // M2L, carrying M2P (sides 2, 3) and the SpMV (side 3, or whatever no M2M
  // launch took)
  const int c1 = nchunks<1>(far.m2l);
  const T *mops = F.m2l.template as<T>();
  const bool m2p_here = !m2p_done && side >= 2, near_here = !near_done;
  if (slot) {
    if (near_here) launch_spmv<T, 1>(var_full, x, nullptr, out, nullptr);
    near_done = true;
    launch_m2l_slots<T>(mops, me, le);
  } else if (near_here && var == 0)
    m2p_here ? launch_translate_side<T, 1, 0, true>(*this, m2l_cl, 0, c1, mops, me, le, sd)
             : launch_translate_side<T, 1, 0, false>(*this, m2l_cl, 0, c1, mops, me, le, sd);
  else if (near_here && var == 1)
    m2p_here ? launch_translate_side<T, 1, 1, true>(*this, m2l_cl, 0, c1, mops, me, le, sd)
             : launch_translate_side<T, 1, 1, false>(*this, m2l_cl, 0, c1, mops, me, le, sd);
  else if (near_here)
    m2p_here ? launch_translate_side<T, 1, 2, true>(*this, m2l_cl, 0, c1, mops, me, le, sd)
             : launch_translate_side<T, 1, 2, false>(*this, m2l_cl, 0, c1, mops, me, le, sd);
  else if (m2p_here)
    launch_translate_side<T, 1, -1, true>(*this, m2l_cl, 0, c1, mops, me, le, sd);
  else
    launch_translate<T, 1, 1>(*this, stream, m2l_cl, 0, c1, mops, me, le);
  m2p_done = m2p_done || m2p_here;
  // L2L, the deepest launched level (or group of two levels) carrying M2P
  // (side 1)
  const bool lgrouped = !far.l2l_g.op1.empty();
  const ChunkDev &lcl = lgrouped ? l2l_g_cl : l2l_cl;
  const int l2l_lo = std::max(2, le_lo + 1);  // L2L from level l - 1 into l
  const int nlg = lgrouped ? (int)far.l2l_g.level_off1.size() - 1 : std::max(0, L + 1 - l2l_lo);
  auto lrange = [&](int s, int &c0, int &cc1) {  // group s, or target level l2l_lo + s
    c0 = lgrouped ? far.l2l_g.level_off1[s] : lvl<1>(far.l2l, l2l_lo + s);
    cc1 = lgrouped ? far.l2l_g.level_off1[s + 1] : lvl<1>(far.l2l, l2l_lo + s + 1);
  };
  int deepest = -1;
  for (int s = 0; !m2p_done && side == 1 && s < nlg; ++s) {
    int c0, cc1;
    lrange(s, c0, cc1);
    if (cc1 > c0) deepest = s;
  }
  for (int s = 0; s < nlg; ++s) {
    int c0, cc1;
    lrange(s, c0, cc1);
    if (cc1 <= c0) continue;
    if (s != deepest) {
      launch_translate<T, 2, 1>(*this, stream, lcl, c0, cc1, F.l2l.template as<T>(), le, le);
      continue;
    }
    launch_translate_side<T, 2, -1, true>(*this, lcl, c0, cc1, F.l2l.template as<T>(), le, le, sd);
    m2p_done = true;
  }
  if (!m2p_done) launch_m2p(stream);
  if (join) HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
  const int nrows = (int)near.rows.size();
  if (nrows > 0)
    PBFMM_BY_ORDER(p, hipLaunchKernelGGL((k_l2p_post<T, PP>), dim3(nblk(nrows, 256)), dim3(256), 0, stream, nrows, d_rows.as<int>(),
                       d_row_leaf.as<int>(), d_leaf_block.as<int>(), d_xyz.as<double>(), d_pmin.as<double>(),
                       d_delta.as<double>(), p, kt, kp, le, out, acc, d_alpha.as<double>(), d_dir.as<uint8_t>(), x,
                       dst));
  HIPCHK(hipGetLastError());
}

template void Engine::matvec<double>(const double *, const double *, double *, double *);
template void Engine::matvec<float>(const double *, const double *, double *, double *);
template void Engine::matvec_combined<double>(const double *, double *);
template void Engine::matvec_combined<float>(const double *, double *);
template void Engine::generate<double, 2>(const double *, const double *);
template void Engine::products<double, 2>(const double *, const double *, double *, double *);

double Engine::norm2(const double *v) {
  const int nchunk = (int)nblk(n, kRedChunk);
  double *part = d_red.as<double>();
  double *res = part + (size_t)nchunk * 128;
  hipLaunchKernelGGL(k_mdot, dim3(nchunk, 1), dim3(kRedThreads), 0, stream, n, n, v, v, part);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRedThreads), 0, stream, nchunk, part, res);
  double h = 0;
  HIPCHK(hipMemcpyAsync(&h, res, 8, hipMemcpyDeviceToHost, stream));
  sync();
  return std::sqrt(h);
}

// Side fusion pays while the M2M / L2L level launches are latency-bound and
// the SpMV and M2P are short beside them (tens of thousands of DoFs); on
// large trees the fused workgroups only crowd each other.  The modes are
// matvec_fused's side 0..3.  Timed once per precision, on the first operator
// application: six applications each way on scratch vectors, after one
// warm-up; a fused mode is kept only if it beats the unfused one by 2%.
int Engine::use_side_fusion(bool f32) {
  int &mode = side_fusion[f32 ? 1 : 0];
  if (mode >= 0) return mode;
  const unsigned g = nblk(n, 256);
  if (d_cost.bytes < (size_t)2 * n * 8) alloc(d_cost, (size_t)2 * n * 8);
  double *u = d_cost.as<double>(), *v = d_cost.as<double>() + n;
  hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, stream, u, n, 1.0);
  hipEvent_t t0, t1;
  HIPCHK(hipEventCreate(&t0));
  HIPCHK(hipEventCreate(&t1));
  constexpr int np = (int)(sizeof(kPlans) / sizeof(kPlans[0]));
  float ms[np] = {};
  for (int s = 0; s < np; ++s)
    for (int rep = 0; rep < 7; ++rep) {
      if (rep == 1) HIPCHK(hipEventRecord(t0, stream));
      if (f32)
        matvec_fused<float>(u, d_outN.as<double>(), v, kPlans[s]);
      else
        matvec_fused<double>(u, d_outN.as<double>(), v, kPlans[s]);
      if (rep == 6) {
        HIPCHK(hipEventRecord(t1, stream));
        HIPCHK(hipEventSynchronize(t1));
        HIPCHK(hipEventElapsedTime(&ms[s], t0, t1));
      }
    }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  int best = 0;
  for (int s = 1; s < np; ++s)
    if (ms[s] < ms[best]) best = s;
  mode = ms[best] < 0.98f * ms[0] ? kPlans[best] : 0;
  return mode;
}

// BEMProblem::vmult.  Without Dirichlet DoFs the reference shifts src and dst
// by their l2 norms (bem_problem.cc:1169-1172, 1214-1215), which is not a linear
// map; vmult does the same.  The GMRES solves (neumann_mean) use instead the
// linear operator A0 x + mean(x) 1, A0 the unshifted one: A0 is singular on
// constants (alpha = -N.1), the rank-one term removes that and makes the solution
// the zero-mean one, which is what the reference's comment asks of the shift.
void Engine::vmult(const double *src, double *dst, int precision) {
  const unsigned g = nblk(n, 256);
  double *shift = nullptr;
  double *res = d_red.as<double>() + (size_t)nblk(n, kRedChunk) * 128;
  const int nchunk = (int)nblk(n, kRedChunk);
  if (!have_dirichlet && !neumann_mean) {
    // serv_phi -= ||src||_2 (bem_problem.cc:1169-1172)
    hipLaunchKernelGGL(k_mdot, dim3(nchunk, 1), dim3(kRedThreads), 0, stream, n, n, src, src, d_red.as<double>());
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRedThreads), 0, stream, nchunk, d_red.as<double>(), res);
    hipLaunchKernelGGL(k_sqrt_inplace, dim3(1), dim3(64), 0, stream, res);
    shift = res;
  }
  // without the shift the operator's input is src itself
  const double *xin = src;
  if (shift) {
    hipLaunchKernelGGL(k_vmult_pre, dim3(g), dim3(256), 0, stream, n, src, d_dir.as<uint8_t>(), shift,
                       d_sp.as<double>());
    xin = d_sp.as<double>();
  }
  const bool f32 = precision == PBFMM_MIXED && have_f32;
  if (!time_phases && (int64_t)near.rows.size() == n) {
    // fewer launches, L2P applying the epilogue (every DoF is one leaf row)
    const int side = use_side_fusion(f32);
    if (f32)
      matvec_fused<float>(xin, d_outN.as<double>(), dst, side);
    else
      matvec_fused<double>(xin, d_outN.as<double>(), dst, side);
  } else {
    if (f32)
      matvec_combined<float>(xin, d_outN.as<double>());
    else
      matvec_combined<double>(xin, d_outN.as<double>());
    hipLaunchKernelGGL(k_vmult_post, dim3(g), dim3(256), 0, stream, n, d_outN.as<double>(), d_alpha.as<double>(),
                       d_dir.as<uint8_t>(), xin, dst, (int64_t)0);
  }
  if (!have_dirichlet) {
    if (neumann_mean) {
      // dst += (sum src / n) 1
      if (!d_ones.p) {
        alloc(d_ones, n * 8);
        hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, stream, d_ones.as<double>(), n, 1.0);
      }
      hipLaunchKernelGGL(k_mdot, dim3(nchunk, 1), dim3(kRedThreads), 0, stream, n, n, src, d_ones.as<double>(),
                         d_red.as<double>());
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRedThreads), 0, stream, nchunk, d_red.as<double>(), res);
      hipLaunchKernelGGL(k_add_scaled_scalar, dim3(g), dim3(256), 0, stream, n, dst, res, 1.0 / (double)n,
                         (int64_t)0);
    } else {
      // dst -= ||dst||_2 (bem_problem.cc:1214-1215)
      hipLaunchKernelGGL(k_mdot, dim3(nchunk, 1), dim3(kRedThreads), 0, stream, n, n, dst, dst, d_red.as<double>());
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRedThreads), 0, stream, nchunk, d_red.as<double>(), res);
      hipLaunchKernelGGL(k_sqrt_inplace, dim3(1), dim3(64), 0, stream, res);
      hipLaunchKernelGGL(k_shift, dim3(g), dim3(256), 0, stream, n, dst, res);
    }
  }
  if (have_constraints)
    hipLaunchKernelGGL(k_constrain, dim3(g), dim3(256), 0, stream, n, d_c_is.as<uint8_t>(), d_c_ptr.as<int>(),
                       d_c_idx.as<int>(), d_c_val.as<double>(), src, dst, (int64_t)0);
  HIPCHK(hipGetLastError());
}

void Engine::compute_alpha() {
  // alpha = N.(-1) through the FP64 FMM (bem_problem.cc:1138-1162)
  const unsigned g = nblk(n, 256);
  hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, stream, d_sp.as<double>(), n, -1.0);
  hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, stream, d_sd.as<double>(), n, 0.0);
  matvec<double>(d_sp.as<double>(), d_sd.as<double>(), d_alpha.as<double>(), d_outD.as<double>());
  HIPCHK(hipGetLastError());
}

void Engine::compute_rhs(const double *gvec, double *rhs) {
  const unsigned g = nblk(n, 256);
  hipLaunchKernelGGL(k_rhs_pre, dim3(g), dim3(256), 0, stream, n, gvec, d_dir.as<uint8_t>(), d_sp.as<double>(),
                     d_sd.as<double>());
  matvec<double>(d_sp.as<double>(), d_sd.as<double>(), d_outN.as<double>(), d_outD.as<double>());
  hipLaunchKernelGGL(k_rhs_post, dim3(g), dim3(256), 0, stream, n, d_outN.as<double>(), d_outD.as<double>(),
                     d_alpha.as<double>(), d_sp.as<double>(), rhs);
  HIPCHK(hipGetLastError());
}

// Gauss-Jordan inverse with partial pivoting of a row-major R x R block in
// place; returns false when a pivot vanishes.
static bool invert_block(std::vector<double> &A, int R) {
  std::vector<double> I((size_t)R * R, 0.0);
  for (int i = 0; i < R; ++i) I[(size_t)i * R + i] = 1.0;
  double scale = 0;
  for (double x : A) scale = std::max(scale, std::fabs(x));
  for (int c = 0; c < R; ++c) {
    int piv = c;
    for (int r = c + 1; r < R; ++r)
      if (std::fabs(A[(size_t)r * R + c]) > std::fabs(A[(size_t)piv * R + c])) piv = r;
    if (!(std::fabs(A[(size_t)piv * R + c]) > 1e-13 * scale)) return false;
    if (piv != c)
      for (int k = 0; k < R; ++k) {
        std::swap(A[(size_t)piv * R + k], A[(size_t)c * R + k]);
        std::swap(I[(size_t)piv * R + k], I[(size_t)c * R + k]);
      }
    const double d = 1.0 / A[(size_t)c * R + c];
    for (int k = 0; k < R; ++k) {
      A[(size_t)c * R + k] *= d;
      I[(size_t)c * R + k] *= d;
    }
    for (int r = 0; r < R; ++r) {
      if (r == c) continue;
      const double f = A[(size_t)r * R + c];
      if (f == 0.0) continue;
      for (int k = 0; k < R; ++k) {
        A[(size_t)r * R + k] -= f * A[(size_t)c * R + k];
        I[(size_t)r * R + k] -= f * I[(size_t)c * R + k];
      }
    }
  }
  A.swap(I);
  return true;
}

// The preconditioner's blocks before inversion, row-major, block g at
// pinv_beg[g] over the rows prows[prow_beg[g] ..): the near field with -D in
// Dirichlet and N in Neumann columns, alpha on the diagonal of Neumann rows
// (k_extract_groups, on the GPU from the assembled FP64 near field), and
// constrained rows replaced by their constraint row src(i) - sum c_ij src(j)
// (constrained_matrix.h:109-121), on the host.
void Engine::prec_blocks(std::vector<double> &blk) {
  blk.assign(pinv_beg.empty() ? 0 : pinv_beg[nprec], 0.0);
  if (nprec <= 0) return;
  DevBuf d_pos, d_raw;
  up(d_pos, prec_pos);
  alloc(d_raw, std::max<size_t>(blk.size(), 1) * 8);
  hipLaunchKernelGGL(k_extract_groups, dim3(nprec), dim3(256), 0, stream, d_prow_beg.as<int>(), d_pos.as<int>(),
                     d_row_leaf.as<int>(), d_row_beg.as<int>(), d_rows.as<int>(), d_col_beg.as<int>(),
                     d_cols.as<int>(), d_val_beg.as<int64_t>(), f64.vals.as<double>(), d_dir.as<uint8_t>(),
                     d_alpha.as<double>(), d_pinv_beg.as<int64_t>(), d_raw.as<double>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(blk.data(), d_raw.p, blk.size() * 8, hipMemcpyDeviceToHost, stream));
  sync();  // d_pos and d_raw are freed on return
  if (!have_constraints) return;
  for (int g = 0; g < nprec; ++g) {
    const int r0 = prow_beg[g], R = prow_beg[g + 1] - r0;
    double *A = blk.data() + pinv_beg[g];
    for (int r = 0; r < R; ++r) {
      const int i = prows[r0 + r];
      if (!hc_is[i]) continue;
      for (int c = 0; c < R; ++c) A[(size_t)r * R + c] = (c == r) ? 1.0 : 0.0;
      for (int k = hc_ptr[i]; k < hc_ptr[i + 1]; ++k) {
        const int j = hc_idx[k];
        for (int c = 0; c < R; ++c)
          if (prows[r0 + c] == j) A[(size_t)r * R + c] -= hc_val[k];
      }
    }
  }
}

void Engine::build_preconditioner() {
  // blocks: near leaves with the same parent block, in near-leaf order; a
  // parent of more than prec_rows rows is cut into ceil(rows / prec_rows)
  // blocks of about equal rows
  const int nl = near.nleaf;
  std::vector<int> order(nl);
  for (int l = 0; l < nl; ++l) order[l] = l;
  auto parent_of = [&](int l) { return tree.parent[near.leaf_block[l]]; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return parent_of(a) < parent_of(b); });
  std::vector<int32_t> pos;  // block-ordered row -> leaf-ordered near row
  prow_beg.assign(1, 0);
  prows.clear();
  for (int i = 0; i < nl;) {
    const int par = parent_of(order[i]);
    int total = 0;
    for (int k = i; k < nl && parent_of(order[k]) == par; ++k)
      total += near.row_beg[order[k] + 1] - near.row_beg[order[k]];
    const int parts = (total + prec_rows - 1) / prec_rows, target = (total + parts - 1) / parts;
    int rows_in = 0;
    for (; i < nl && parent_of(order[i]) == par; ++i) {
      const int l = order[i], R = near.row_beg[l + 1] - near.row_beg[l];
      if (rows_in > 0 && (rows_in + R > prec_rows || rows_in >= target)) {
        prow_beg.push_back((int32_t)pos.size());
        rows_in = 0;
      }
      for (int r = near.row_beg[l]; r < near.row_beg[l + 1]; ++r) {
        pos.push_back(r);
        prows.push_back(near.rows[r]);
      }
      rows_in += R;
    }
    if (rows_in > 0) prow_beg.push_back((int32_t)pos.size());
  }
  nprec = (int)prow_beg.size() - 1;
  pinv_beg.assign(nprec + 1, 0);
  max_leaf_rows = 0;
  std::vector<int32_t> prow_blk(prows.size());
  for (int g = 0; g < nprec; ++g) {
    const int64_t R = prow_beg[g + 1] - prow_beg[g];
    pinv_beg[g + 1] = pinv_beg[g] + R * R;
    max_leaf_rows = std::max(max_leaf_rows, (int)R);
    for (int r = prow_beg[g]; r < prow_beg[g + 1]; ++r) prow_blk[r] = g;
  }
  up(d_pinv_beg, pinv_beg);
  up(d_prow_beg, prow_beg);
  up(d_prows, prows);
  up(d_prow_blk, prow_blk);
  prec_pos = pos;
  // per copy f (0 FP64 pairs, 1 FP32 quads) the blocks' offsets
  std::vector<int64_t> pvb[2];
  for (int f = 0; f < 2; ++f) {
    const int CW = f ? 4 : 2;
    pvb[f].assign(nprec + 1, 0);
    for (int g = 0; g < nprec; ++g) {
      const int64_t R = prow_beg[g + 1] - prow_beg[g];
      pvb[f][g + 1] = pvb[f][g] + R * ((R + CW - 1) / CW * CW);
    }
    up(d_pvec_beg[f], pvb[f]);
  }
  std::vector<double> b64((size_t)pvb[0][nprec], 0.0);
  std::vector<float> b32((size_t)pvb[1][nprec], 0.0f);
  std::vector<double> blk;
  prec_blocks(blk);
  // invert the blocks on host threads
  auto invert_range = [&](int g_lo, int g_hi) {
    std::vector<double> A;
    for (int g = g_lo; g < g_hi; ++g) {
      const int R = prow_beg[g + 1] - prow_beg[g];
      A.assign(blk.begin() + pinv_beg[g], blk.begin() + pinv_beg[g + 1]);
      std::vector<double> keep = A;
      if (!invert_block(A, R)) {
        // singular block (e.g. unconstrained double nodes): diagonal scaling
        for (int r = 0; r < R; ++r)
          for (int c = 0; c < R; ++c) {
            const double d = keep[(size_t)r * R + r];
            A[(size_t)r * R + c] = (r == c) ? (d != 0.0 ? 1.0 / d : 1.0) : 0.0;
          }
      }
      for (int r = 0; r < R; ++r)
        for (int c = 0; c < R; ++c) {
          const double a = A[(size_t)r * R + c];
          b64[pvb[0][g] + pinv_vec_index<double>(r, c, R)] = a;
          b32[pvb[1][g] + pinv_vec_index<float>(r, c, R)] = (float)a;
        }
    }
  };
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (nprec < 64 || nth == 1) {
    invert_range(0, nprec);
  } else {
    std::vector<std::thread> th;
    for (int k = 0; k < nth; ++k)
      th.emplace_back(invert_range, (int)((int64_t)nprec * k / nth), (int)((int64_t)nprec * (k + 1) / nth));
    for (auto &x : th) x.join();
  }
  up(d_pinv[0], b64);
  up(d_pinv[1], b32);
  alloc(d_ptmp, n * 8);
  alloc(d_ptmp2, n * 8);
  have_prec = true;
}

template <typename PT>
static void apply_prec_copy(Engine &e, const double *v, double *z, int R) {
  const int f = sizeof(PT) == 4 ? 1 : 0;
  const int64_t *vb = e.d_pvec_beg[f].as<int64_t>();
  const PT *pinv = e.d_pinv[f].as<PT>();
  const int64_t n = e.n;
  if (e.max_leaf_rows <= 128) {  // R right-hand sides, n apart, in one launch
    if (e.nprec > 0 && e.max_leaf_rows <= 64)
      hipLaunchKernelGGL((k_bjacobi_wg<1, PT>), dim3(e.nprec, R), dim3(256), 0, e.stream, e.d_prow_beg.as<int>(),
                         e.d_prows.as<int>(), vb, pinv, v, z, n);
    else if (e.nprec > 0)
      hipLaunchKernelGGL((k_bjacobi_wg<2, PT>), dim3(e.nprec, R), dim3(256), 0, e.stream, e.d_prow_beg.as<int>(),
                         e.d_prows.as<int>(), vb, pinv, v, z, n);
    return;
  }
  const int nrows = (int)e.prows.size();
  for (int r = 0; r < R; ++r)
    hipLaunchKernelGGL(k_bjacobi<PT>, dim3(nblk(nrows, 256)), dim3(256), 0, e.stream, nrows,
                       e.d_prow_blk.as<int>(), e.d_prow_beg.as<int>(), e.d_prows.as<int>(), vb, pinv,
                       v + (int64_t)r * n, z + (int64_t)r * n);
}
// The FP32 copy for mixed solves (prec_f32, set per solve), else the FP64 one
void Engine::apply_prec(const double *v, double *z, int R) {
  if (prec_f32)
    apply_prec_copy<float>(*this, v, z, R);
  else
    apply_prec_copy<double>(*this, v, z, R);
}

// One Arnoldi step's operator: w = A M^-1 vin (right preconditioning).
void Engine::step_op(const double *vin, double *w, int precision) {
  if (have_prec) {
    apply_prec(vin, d_ptmp.as<double>());
    vmult(d_ptmp.as<double>(), w, precision);
  } else {
    vmult(vin, w, precision);
  }
}

// What a captured step holds: the step precision, the tuned choices its
// launches follow, the operator's form and every device address (through the
// allocation count).
std::vector<int64_t> Engine::graph_key(int precision) const {
  const int f = precision == PBFMM_MIXED && have_f32 ? 1 : 0;
  return {precision, side_fusion[f], spmv_variant[f][0], have_prec ? 1 : 0, prec_f32 ? 1 : 0, nprec,
          have_constraints ? 1 : 0,
          constraint_version, neumann_mean ? 1 : 0, have_dirichlet ? 1 : 0, g_alloc_epoch};
}

// step_op(d_vn -> d_wk) as a graph replay when that was measured faster.
// The first call per step precision runs it directly (the side-fusion and
// SpMV choices time themselves there), the next captures it after one more
// direct run (allocations on first use happen outside the capture) and times
// three direct runs against three replays; from then on the faster one runs,
// recaptured whenever graph_key changes.
void Engine::step_op_graph(int precision) {
  const int f = precision == PBFMM_MIXED && have_f32 ? 1 : 0;
  double *vin = d_vn.as<double>(), *w = d_wk.as<double>();
  if (!graph_enabled || graph_use[f] == 0 || side_fusion[f] < 0) {
    step_op(vin, w, precision);
    return;
  }
  if (!op_graph[f] || op_graph_key[f] != graph_key(precision)) {
    if (op_graph[f]) (void)hipGraphExecDestroy(op_graph[f]);
    op_graph[f] = nullptr;
    step_op(vin, w, precision);  // allocations on first use, outside the capture
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      step_op(vin, w, precision);
    } catch (...) {
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(stream, &g));
    const hipError_t ie = hipGraphInstantiate(&op_graph[f], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      op_graph[f] = nullptr;
      throw HipError("operator graph instantiation failed");
    }
    op_graph_key[f] = graph_key(precision);
  }
  if (graph_use[f] < 0) {
    hipEvent_t t0, t1, t2;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    HIPCHK(hipEventCreate(&t2));
    HIPCHK(hipGraphLaunch(op_graph[f], stream));  // warm
    HIPCHK(hipEventRecord(t0, stream));
    for (int r = 0; r < 3; ++r) step_op(vin, w, precision);
    HIPCHK(hipEventRecord(t1, stream));
    for (int r = 0; r < 3; ++r) HIPCHK(hipGraphLaunch(op_graph[f], stream));
    HIPCHK(hipEventRecord(t2, stream));
    HIPCHK(hipEventSynchronize(t2));
    float md = 0, mg = 0;
    HIPCHK(hipEventElapsedTime(&md, t0, t1));
    HIPCHK(hipEventElapsedTime(&mg, t1, t2));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipEventDestroy(t2);
    graph_use[f] = mg < 0.98f * md ? 1 : 0;
    graph_ms[f][0] = md / 3;
    graph_ms[f][1] = mg / 3;
  }
  if (graph_use[f] == 1)
    HIPCHK(hipGraphLaunch(op_graph[f], stream));
  else
    step_op(vin, w, precision);
}

// Restarted GMRES with classical Gram-Schmidt and one reorthogonalisation
// (CGS2), FP64 basis and Hessenberg, operator precision as requested.  Solves
// A x = b from the given x until the Arnoldi residual estimate reaches
// tol_abs or maxit operator applications have been spent.
// A step is enqueued ahead of a host round trip only while Arnoldi steps
// take less than this (ms): the idle time a round trip leaves is tens of
// microseconds, an unused step costs a whole operator application.
constexpr double kSpeculateMs = 0.5;

// Krylov space recycled from one inner solve of the refinement into the next
// (GCRO, de Sturler 1996; Parks et al. 2006, without the deflation step: the
// whole previous space is kept).  The previous solve's last cycle satisfies
// B V_k = V_k+1 Hbar (B the preconditioned FP32 operator, Hbar (k+1) x k);
// with Hbar = Q R, C = V_k+1 Q is orthonormal and B (V_k R^-1) = C.  The next
// solve projects its residual onto C, runs its Arnoldi process orthogonal to
// C (C sits in front of the new basis, so the Gram-Schmidt kernels treat it
// as k more basis vectors) and adds V_k R^-1 (C^T r - E y) to the correction,
// E the new steps' coefficients on C: its residual is the minimum over the
// old and the new space together, without applying the operator to the old
// one again.
struct Recycle {
  bool record = false;        // keep the last cycle's Hessenberg (k, Hbar)
  int k = 0;
  std::vector<double> Hbar;   // column j at (m + 1) j, zero below row j + 1
  int kc = 0;                 // > 0: C in d_V[0 .. kc), V_kc in d_rec, R
  std::vector<double> R;      // kc x kc upper triangular, row-major
};

// r0 (device) with norm r0n, when given, is the residual b - A x of the
// starting x: the first cycle then skips its operator application.
// VT: the basis vectors' storage, double, or float for the refinement's
// inner solves when asked (mixed scheme 4: half the bytes the Gram-Schmidt
// kernels stream, which read the basis four times per step; orthogonality
// to FP32 rounding is far below an inner solve's stopping reduction).  Every dot, projection and
// update runs in FP64, and the operator reads each new basis vector as
// stored, so the Arnoldi relation holds for the stored basis.
template <typename VT>
static int gmres_core(Engine &e, const double *b, double *x, double tol_abs, int m, int maxit, int precision,
                      int &it, double &est, double switch_abs = -1.0, const double *r0 = nullptr, double r0n = 0,
                      Recycle *rc = nullptr) {
  const int64_t n = e.n;
  const unsigned g = nblk(n, 256);
  VT *V = e.d_V.as<VT>();
  double *w = e.d_w.as<double>(), *hd = e.d_h.as<double>();
  // recycled vectors in front of the basis (first cycle only)
  int kc = rc && rc->kc > 0 && rc->kc < m ? rc->kc : 0;
  std::vector<double> Ecol, cproj;  // coefficients on C of each step; C^T r
  if (rc && rc->record) rc->k = 0;
  // dot partials (kDotParts x 128), the norm partials of the two
  // projections (one per block of k_gemv_sub_fin each), |w1|^2 and the
  // second projection's skip flag
  const size_t cg_doubles = (size_t)kDotParts * 128 + 2 * (size_t)g + 4;
  if (e.d_cgs.bytes < cg_doubles * 8) alloc(e.d_cgs, cg_doubles * 8);
  double *cg = e.d_cgs.as<double>();
  double *npart1 = cg + (size_t)kDotParts * 128, *npart2 = npart1 + g, *nrm1_sq = npart2 + g;
  int *skip2 = reinterpret_cast<int *>(nrm1_sq + 1);
  static_assert(kNormParts <= kDotParts * 128, "the recycled vectors' norm partials share the dot partials");
  // An FP32 basis always takes the second projection: with the selective
  // test its steps lose orthogonality (config 3: 113 Arnoldi steps against
  // 53, where the FP64 basis needs 54 either way)
  const bool selective = e.reorth_eta < 1.0 && sizeof(VT) == 8;
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), gv(m + 1), h(m + 2), y(m);
  // Hessenberg slots on the device, one per Arnoldi step: [h1 (m+2)][h2 (m+2)][|w|]
  const int S = 2 * (m + 2) + 1;
  double *ybuf = hd + (size_t)m * S;
  if (e.h_pinned_n < (size_t)m * S) {
    if (e.h_pinned) HIPCHK(hipHostFree(e.h_pinned));
    e.h_pinned = nullptr;
    e.h_pinned_n = 0;
    HIPCHK(hipHostMalloc((void **)&e.h_pinned, (size_t)m * S * 8, hipHostMallocDefault));
    e.h_pinned_n = (size_t)m * S;
  }
  if (!e.h_ready) HIPCHK(hipEventCreateWithFlags(&e.h_ready, hipEventDisableTiming));
  double *hbuf = e.h_pinned;
  // relaxed mixed precision (switch_abs > 0): FP64 operator while the
  // residual estimate is above switch_abs, the FP32 operator below it, in one
  // Krylov space; the cycle's starting residual always in FP64
  const bool relaxed = switch_abs > 0 && precision == PBFMM_MIXED;
  auto step_prec = [&]() { return relaxed ? (est <= switch_abs ? PBFMM_MIXED : PBFMM_FP64) : precision; };
  // the step's operator from a fixed input (a copy of the last basis vector)
  // into a fixed output, so that it can be replayed as a graph
  // (Engine::step_op_graph); the fused vmult is the one a graph can hold
  const bool graphs = e.graph_enabled && !e.time_phases && (int64_t)e.near.rows.size() == n;
  // the operator from d_vn into d_wk: for graphs, and for an FP32 basis
  // (the operator reads FP64 vectors)
  constexpr bool f32_basis = sizeof(VT) == 4;
  const bool fixed_io = graphs || f32_basis;
  if (fixed_io) {
    if (e.d_vn.bytes < (size_t)n * 8) alloc(e.d_vn, (size_t)n * 8);
    if (e.d_wk.bytes < (size_t)n * 8) alloc(e.d_wk, (size_t)n * 8);
  }
  // Arnoldi step k: V_k+1 from V_k (the basis after the kc recycled vectors),
  // its Hessenberg column (the coefficients on C first) into slot k
  auto enqueue_step = [&](int k) {
    VT *Vb = V + (int64_t)kc * n;
    VT *vk1 = Vb + (int64_t)(k + 1) * n;
    const int pk = step_prec();
    double *wv = e.d_wk.as<double>();  // the new vector before its normalisation
    if (graphs) {
      e.step_op_graph(pk);  // d_vn (= V_k) -> d_wk
    } else if constexpr (f32_basis) {
      e.step_op(e.d_vn.as<double>(), wv, pk);
    } else {
      wv = vk1;
      e.step_op(Vb + (int64_t)k * n, vk1, pk);
    }
    ++it;
    // CGS with a second projection when needed, all on the device, six
    // launches: the projections and the scaling finish the dot products they
    // use, and the projections sum the squares of their results by block
    double *slot = hd + (size_t)k * S;
    double *h1 = slot, *h2 = slot + (m + 2), *hn_d = slot + 2 * (m + 2);
    const int nv = kc + k + 1;
    // CGS with the second projection when the first cancelled severely
    // (selective, k_reorth_flag), or always (reorth_eta >= 1)
    hipLaunchKernelGGL(k_mdot_parts<VT>, dim3(kDotParts, nv), dim3(kRedThreads), 0, e.stream, n, n, V, wv, cg,
                       nullptr);
    hipLaunchKernelGGL(k_gemv_sub_fin<VT>, dim3(g), dim3(256), 0, e.stream, n, n, nv, V, cg, kDotParts, h1, wv,
                       nullptr, selective ? npart1 : nullptr);
    const int *sk = nullptr;
    if (selective) {
      hipLaunchKernelGGL(k_reorth_flag, dim3(1), dim3(256), 0, e.stream, h1, nv, npart1, (int)g, e.reorth_eta, skip2,
                         nrm1_sq);
      sk = skip2;
    }
    hipLaunchKernelGGL(k_mdot_parts<VT>, dim3(kDotParts, nv), dim3(kRedThreads), 0, e.stream, n, n, V, wv, cg, sk);
    // the norm after the second projection, or w1's when it was skipped
    hipLaunchKernelGGL(k_gemv_sub_fin<VT>, dim3(g), dim3(256), 0, e.stream, n, n, nv, V, cg, kDotParts, h2, wv, sk,
                       npart2);
    if constexpr (f32_basis) {
      hipLaunchKernelGGL(k_scale_fin2<VT>, dim3(g), dim3(256), 0, e.stream, n, wv, vk1, e.d_vn.as<double>(), sk,
                         nrm1_sq, npart2, (int)g, hn_d);
    } else {
      if (graphs)
        hipLaunchKernelGGL(k_scale_fin2<double>, dim3(g), dim3(256), 0, e.stream, n, wv, vk1, e.d_vn.as<double>(), sk,
                           nrm1_sq, npart2, (int)g, hn_d);
      else
        hipLaunchKernelGGL(k_scale_fin, dim3(g), dim3(256), 0, e.stream, n, vk1, sk, nrm1_sq, npart2, (int)g, hn_d);
    }
  };
  while (it < maxit) {
    // r = b - A x into V_0
    const double *rv = w;
    double beta = r0n;
    if (r0) {
      rv = r0;
      r0 = nullptr;
    } else {
      e.vmult(x, w, relaxed ? PBFMM_FP64 : precision);
      hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, e.stream, n, 1.0, b, -1.0, w);
      beta = e.norm2(w);
    }
    est = beta;
    if (beta <= tol_abs) return 0;
    if (kc > 0) {
      // r - C (C^T r), twice, then normalised into the basis' first vector
      // (and the graph's input); C^T r and the norm in the last step slot,
      // which this cycle's kc + steps <= m never reaches
      if (e.d_vn.bytes < (size_t)n * 8) alloc(e.d_vn, (size_t)n * 8);
      double *ps = hd + (size_t)(m - 1) * S;
      HIPCHK(hipMemcpyAsync(w, rv, (size_t)n * 8, hipMemcpyDeviceToDevice, e.stream));
      for (double *pj : {ps, ps + (m + 2)}) {
        hipLaunchKernelGGL(k_mdot_parts<VT>, dim3(kDotParts, kc), dim3(kRedThreads), 0, e.stream, n, n, V, w, cg,
                           nullptr);
        hipLaunchKernelGGL(k_gemv_sub_fin<VT>, dim3(g), dim3(256), 0, e.stream, n, n, kc, V, cg, kDotParts, pj, w,
                           nullptr, nullptr);
      }
      hipLaunchKernelGGL(k_mdot_parts<double>, dim3(kNormParts, 1), dim3(kRedThreads), 0, e.stream, n, n, w, w, cg,
                         nullptr);
      hipLaunchKernelGGL(k_scale_fin2<VT>, dim3(g), dim3(256), 0, e.stream, n, w, V + (int64_t)kc * n,
                         e.d_vn.as<double>(), nullptr, nullptr, cg, kNormParts, ps + 2 * (m + 2));
      std::vector<double> pr((size_t)S);
      HIPCHK(hipMemcpyAsync(pr.data(), ps, (size_t)S * 8, hipMemcpyDeviceToHost, e.stream));
      e.sync();
      cproj.assign(kc, 0.0);
      for (int i = 0; i < kc; ++i) cproj[i] = pr[i] + pr[(m + 2) + i];
      beta = pr[2 * (m + 2)];
      est = beta;
      Ecol.clear();
    } else {
      hipLaunchKernelGGL(k_scale_to_basis<VT>, dim3(g), dim3(256), 0, e.stream, n, rv, 1.0 / beta, V,
                         fixed_io ? e.d_vn.as<double>() : nullptr);
    }
    if (rc && rc->record) rc->Hbar.clear();
    const int mk = m - kc;  // steps this cycle
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = beta;
    // kq: steps enqueued; k_done: columns reduced on the host; the next host
    // round trip comes after step k_sync.  A round trip copies the new columns
    // to pinned memory and, unless convergence is predicted for them, enqueues
    // the next step before it waits, so that the GPU works while the host
    // reduces (the step is unused if they do converge).
    int kq = 0, k_done = 0, k_sync = 0, k = 0;
    bool done = false, speculate = true;
    auto t_trip = std::chrono::steady_clock::now();
    while (true) {
      if (kq < mk && it < maxit) enqueue_step(kq++);
      const bool more = kq < mk && it < maxit;
      if (kq - 1 < k_sync && more) continue;
      if (k_done == kq) break;  // every enqueued column reduced and no step left
      // host round trip: the columns k_done..kq-1, Givens rotations, residual estimate
      const int kk = kq;
      HIPCHK(hipMemcpyAsync(hbuf + (size_t)k_done * S, hd + (size_t)k_done * S, (size_t)(kk - k_done) * S * 8,
                            hipMemcpyDeviceToHost, e.stream));
      HIPCHK(hipEventRecord(e.h_ready, e.stream));
      if (more && speculate) enqueue_step(kq++);
      HIPCHK(hipEventSynchronize(e.h_ready));
      const double est0 = est;
      const int j0 = k_done;
      for (int j = k_done; j < kk; ++j) {
        const double *hb = hbuf + (size_t)j * S;
        for (int i = 0; i < kc; ++i) Ecol.push_back(hb[i] + hb[(m + 2) + i]);
        for (int i = 0; i <= j; ++i) h[i] = hb[kc + i] + hb[(m + 2) + kc + i];
        const double hn = hb[2 * (m + 2)];
        h[j + 1] = hn;
        if (rc && rc->record) {  // the column before the rotations, padded to m + 1
          for (int i = 0; i <= j + 1; ++i) rc->Hbar.push_back(h[i]);
          for (int i = j + 2; i <= m; ++i) rc->Hbar.push_back(0.0);
        }
        for (int i = 0; i < j; ++i) {
          const double a = h[i], bb = h[i + 1];
          h[i] = cs[i] * a + sn[i] * bb;
          h[i + 1] = -sn[i] * a + cs[i] * bb;
        }
        const double den = std::hypot(h[j], h[j + 1]);
        cs[j] = h[j] / den;
        sn[j] = h[j + 1] / den;
        h[j] = den;
        for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = h[i];
        gv[j + 1] = -sn[j] * gv[j];
        gv[j] = cs[j] * gv[j];
        est = std::fabs(gv[j + 1]);
        if (est <= tol_abs || hn == 0) {
          k = j + 1;  // the Arnoldi steps after j are not used
          done = true;
          break;
        }
      }
      if (done) break;
      // host time per step since the previous round trip
      const auto now = std::chrono::steady_clock::now();
      const double step_ms = std::chrono::duration<double, std::milli>(now - t_trip).count() / (kk - j0);
      t_trip = now;
      k_done = kk;
      k = kk;
      // next round trip: half the steps the observed decay predicts to the
      // next event (the precision switch, then convergence), 1..8 ahead.  A
      // step is enqueued ahead of a round trip only when convergence is not
      // expected there and steps are short (an unused step of a large
      // problem costs more than the round trip's idle time saves)
      int ahead = 1;
      speculate = step_ms < kSpeculateMs;
      if (est < est0 && est0 > 0) {
        const double rate = std::log(est / est0) / (kk - j0);  // per step, < 0
        const double goal = (relaxed && est > switch_abs) ? switch_abs : tol_abs;
        const double left = std::log(goal / est) / rate;
        ahead = (int)std::max(1.0, std::min(8.0, 0.5 * left));
        // long steps too while convergence is predicted well past the next
        // round trip (an unused step would take a misprediction by 4 steps)
        speculate = (speculate && left > ahead + 1.5) || left > ahead + 4.0;
      }
      k_sync = kk - 1 + ahead;
    }
    // k: the columns that enter the update (all reduced ones, or up to the
    // converged step)
    for (int i = k - 1; i >= 0; --i) {
      double s = gv[i];
      for (int j = i + 1; j < k; ++j) s -= H[(size_t)i * m + j] * y[j];
      y[i] = s / H[(size_t)i * m + i];
    }
    HIPCHK(hipMemcpyAsync(ybuf, y.data(), k * 8, hipMemcpyHostToDevice, e.stream));
    // recycled part: V_kc R^-1 (C^T r - E y)
    std::vector<double> zr;
    if (kc > 0) {
      std::vector<double> a(cproj);
      for (int j = 0; j < k; ++j)
        for (int i = 0; i < kc; ++i) a[i] -= Ecol[(size_t)j * kc + i] * y[j];
      zr.assign(kc, 0.0);
      for (int i = kc - 1; i >= 0; --i) {
        double s = a[i];
        for (int j = i + 1; j < kc; ++j) s -= rc->R[(size_t)i * kc + j] * zr[j];
        zr[i] = s / rc->R[(size_t)i * kc + i];
      }
      if (e.d_zr.bytes < (size_t)kc * 8) alloc(e.d_zr, (size_t)std::max(kc, 16) * 8);
      HIPCHK(hipMemcpyAsync(e.d_zr.p, zr.data(), (size_t)kc * 8, hipMemcpyHostToDevice, e.stream));
    }
    const VT *Vb = V + (int64_t)kc * n;
    if (e.have_prec) {
      // x += M^-1 (V y [+ V_kc z])
      double *t = e.d_ptmp2.as<double>(), *z = e.d_ptmp.as<double>();
      hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, e.stream, t, n, 0.0);
      hipLaunchKernelGGL(k_gemv_add<VT>, dim3(g), dim3(256), 0, e.stream, n, n, k, Vb, ybuf, t);
      if (kc > 0)
        hipLaunchKernelGGL(k_gemv_add<VT>, dim3(g), dim3(256), 0, e.stream, n, n, kc, e.d_rec.as<VT>(),
                           e.d_zr.as<double>(), t);
      e.apply_prec(t, z);
      hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, e.stream, n, 1.0, z, 1.0, x);
    } else {
      hipLaunchKernelGGL(k_gemv_add<VT>, dim3(g), dim3(256), 0, e.stream, n, n, k, Vb, ybuf, x);
      if (kc > 0)
        hipLaunchKernelGGL(k_gemv_add<VT>, dim3(g), dim3(256), 0, e.stream, n, n, kc, e.d_rec.as<VT>(),
                           e.d_zr.as<double>(), x);
    }
    if (rc && rc->record) rc->k = k;
    e.sync();
    if (done) return 0;
    kc = 0;  // later cycles restart plainly
  }
  return 1;
}

// Sets up the recycled space for the next inner solve from the one just
// finished (rec.k steps, Hbar, its basis still in d_V): Hbar = Q R by
// Gram-Schmidt with reorthogonalisation on the host, the old V_k copied to
// d_rec, C = V_k+1 Q formed beside it and moved to the front of d_V.  Returns
// false (no recycling) when R is numerically singular.
template <typename VT>
static bool prepare_recycle(Engine &e, Recycle &rec, int m) {
  const int k = rec.k;
  if (k <= 0 || k >= m) return false;
  const int64_t n = e.n;
  std::vector<double> Q((size_t)(k + 1) * k, 0.0), R((size_t)k * k, 0.0);
  double rmax = 0;
  for (int j = 0; j < k; ++j) {
    double *q = Q.data() + (size_t)j * (k + 1);
    for (int i = 0; i <= k; ++i) q[i] = rec.Hbar[(size_t)j * (m + 1) + i];
    for (int pass = 0; pass < 2; ++pass)
      for (int l = 0; l < j; ++l) {
        const double *ql = Q.data() + (size_t)l * (k + 1);
        double dot = 0;
        for (int i = 0; i <= k; ++i) dot += ql[i] * q[i];
        for (int i = 0; i <= k; ++i) q[i] -= dot * ql[i];
        R[(size_t)l * k + j] += dot;
      }
    double nr = 0;
    for (int i = 0; i <= k; ++i) nr += q[i] * q[i];
    nr = std::sqrt(nr);
    R[(size_t)j * k + j] = nr;
    rmax = std::max(rmax, nr);
    if (!(nr > 0)) return false;
    for (int i = 0; i <= k; ++i) q[i] /= nr;
  }
  for (int j = 0; j < k; ++j)
    if (!(R[(size_t)j * k + j] > 1e-12 * rmax)) return false;
  if (e.d_rec.bytes < (size_t)2 * k * n * sizeof(VT)) alloc(e.d_rec, (size_t)2 * k * n * sizeof(VT));
  if (e.d_qr.bytes < Q.size() * 8) alloc(e.d_qr, Q.size() * 8);
  VT *V = e.d_V.as<VT>(), *U = e.d_rec.as<VT>(), *Ct = U + (int64_t)k * n;
  HIPCHK(hipMemcpyAsync(U, V, (size_t)k * n * sizeof(VT), hipMemcpyDeviceToDevice, e.stream));
  HIPCHK(hipMemcpyAsync(e.d_qr.p, Q.data(), Q.size() * 8, hipMemcpyHostToDevice, e.stream));
  hipLaunchKernelGGL(k_tsgemm8<VT>, dim3(nblk(n, 256), (k + 7) / 8), dim3(256), 0, e.stream, n, n, k + 1, V,
                     e.d_qr.as<double>(), k, Ct);
  HIPCHK(hipMemcpyAsync(V, Ct, (size_t)k * n * sizeof(VT), hipMemcpyDeviceToDevice, e.stream));
  HIPCHK(hipGetLastError());
  rec.kc = k;
  rec.R = std::move(R);
  rec.record = false;
  return true;
}

// Largest residual reduction asked of one FP32-operator inner solve of the
// iterative refinement: the FP32 operator differs from the FP64 one by ~1e-6
// relative, so an inner solve cannot lower the true residual much further.
constexpr double kInnerReduction = 1e-5;
// Relaxed mixed precision: the FP32 operator differs from the FP64 one by
// ~3e-7..7e-7 relative (measured on the 24,578- and 393,218-DoF test
// problems); inexact-Krylov theory lets step k carry an operator error up to
// ~tol ||b|| / ||r_k||, so FP32 steps are safe once ||r_k|| <= tol ||b|| / eps,
// taken here with a factor-5 margin on eps = 1e-6
constexpr double kRelaxSwitch = 2e5;

// Stopping residual of the next inner solve of the iterative refinement.  The
// reduction still needed, q = (tol ||b|| / 2) / ||r||, is split into the
// fewest equal stages of at most kInnerReduction each, and this solve takes
// one stage: for tol = 1e-8 two solves of ~7e-5 each.  Later steps of one
// Krylov space converge more slowly than the first steps of a restarted one,
// so equal stages need the fewest steps: 67 at the 393,218-DoF Q2 problem,
// against 72 with a first solve to 1e-5 (38 -> 42 at the 24,578-DoF sphere).
// An inner_reduction set through the API replaces the split by that fixed
// reduction per solve.
double Engine::inner_target(double tol, double bn, double rn) const {
  const double goal = tol * bn * 0.5;
  if (inner_reduction > 0) return std::max(goal, inner_reduction * rn);
  if (!(rn > goal)) return goal;
  const double q = goal / rn;
  const int stages = std::max(1, (int)std::ceil(std::log(q) / std::log(kInnerReduction) - 1e-9));
  return std::max(goal, rn * std::pow(q, 1.0 / stages));
}
double Engine::relax_sw() const { return relax_switch > 0 ? relax_switch : kRelaxSwitch; }

bool Engine::mixed_relaxed() {
  if (mixed_scheme != 0) return mixed_scheme == 1;
  if (op_cost_ratio <= 0) {
    // one timed application of each operator on scratch vectors of its own
    const unsigned g = nblk(n, 256);
    alloc(d_cost, (size_t)2 * n * 8);
    double *u = d_cost.as<double>(), *v = d_cost.as<double>() + n;
    hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, stream, u, n, 1.0);
    hipEvent_t t0, t1, t2;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    HIPCHK(hipEventCreate(&t2));
    vmult(u, v, PBFMM_MIXED);  // warm both paths once
    vmult(u, v, PBFMM_FP64);
    HIPCHK(hipEventRecord(t0, stream));
    vmult(u, v, PBFMM_MIXED);
    HIPCHK(hipEventRecord(t1, stream));
    vmult(u, v, PBFMM_FP64);
    HIPCHK(hipEventRecord(t2, stream));
    HIPCHK(hipEventSynchronize(t2));
    float m32 = 0, m64 = 0;
    HIPCHK(hipEventElapsedTime(&m32, t0, t1));
    HIPCHK(hipEventElapsedTime(&m64, t1, t2));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipEventDestroy(t2);
    op_cost_ratio = m32 > 0 ? (double)m64 / m32 : 1.0;
    d_cost.release();
  }
  return op_cost_ratio < kRelaxCostRatio;
}

int Engine::gmres(const double *b, double *x, double tol, int restart, int maxit, int precision, int *iters,
                  double *relres) {
  if (restart != gm_restart) {
    alloc(d_V, (size_t)(restart + 1) * n * 8);
    alloc(d_w, n * 8);
    alloc(d_r, n * 8);
    alloc(d_x, n * 8);
    alloc(d_h, ((size_t)restart * (2 * (restart + 2) + 1) + restart + 2) * 8);
    gm_restart = restart;
  }
  const unsigned g = nblk(n, 256);
  // |b| and |x| in one round trip: a zero starting guess needs no operator
  // application for its residual
  double sq[2] = {0, 0};
  {
    const int nchunk = (int)nblk(n, kRedChunk);
    double *part = d_red.as<double>(), *res = part + (size_t)nchunk * 128;
    hipLaunchKernelGGL(k_mdot, dim3(nchunk, 1), dim3(kRedThreads), 0, stream, n, n, b, b, part);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRedThreads), 0, stream, nchunk, part, res);
    hipLaunchKernelGGL(k_mdot, dim3(nchunk, 1), dim3(kRedThreads), 0, stream, n, n, x, x, part);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRedThreads), 0, stream, nchunk, part, res + 1);
    HIPCHK(hipMemcpyAsync(sq, res, 16, hipMemcpyDeviceToHost, stream));
    sync();
  }
  const double bn = std::max(std::sqrt(sq[0]), 1e-300);
  bool x_zero = sq[1] == 0.0;
  struct MeanMode {  // linear pure-Neumann operator for the duration of the solve
    bool &flag;
    bool old;
    MeanMode(bool &f, bool v) : flag(f), old(f) { flag = v; }
    ~MeanMode() { flag = old; }
  } mean_mode(neumann_mean, !have_dirichlet);
  // the preconditioner's FP32 copy for a mixed solve, its FP64 copy else
  MeanMode prec_mode(prec_f32, precision == PBFMM_MIXED && have_f32);
  int it = 0;
  double est = 0, rr = 1;
  double *r = d_r.as<double>(), *d = d_x.as<double>();
  Recycle rec;  // the refinement's first inner solve, recycled into the second
  int stage = 0;
  for (int outer = 0; outer < 100 && it < maxit; ++outer) {
    // true FP64 residual (b itself for a zero starting guess)
    double rn = bn;
    if (x_zero) {
      HIPCHK(hipMemcpyAsync(r, b, (size_t)n * 8, hipMemcpyDeviceToDevice, stream));
      x_zero = false;
    } else {
      vmult(x, r, PBFMM_FP64);
      hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, stream, n, 1.0, b, -1.0, r);
      rn = norm2(r);
    }
    rr = rn / bn;
    if (rr <= tol) break;
    if (precision == PBFMM_MIXED && have_f32) {
      hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, stream, d, n, 0.0);
      if (mixed_relaxed()) {
        // correction equation A d = r to the final tolerance in one Krylov
        // space: FP64 operator while the residual is large, FP32 once it is
        // below kRelaxSwitch tol ||b|| (inexact Krylov: the operator error a
        // step may carry grows as the residual shrinks); the outer FP64
        // residual catches what the FP32 steps leave
        gmres_core<double>(*this, r, d, tol * bn * 0.5, restart, maxit, PBFMM_MIXED, it, est,
                           std::min(rn, relax_sw() * tol * bn), r, rn);
      } else {
        // iterative refinement: FP32-operator GMRES on the FP64 residual;
        // the second inner solve recycles the first one's Krylov space
        const double tin = inner_target(tol, bn, rn);
        Recycle *rp = nullptr;
        const bool recyc = recycle && mixed_scheme != 3;  // 3: refinement without recycling
        const bool f32b = mixed_scheme == 4;              // 4: refinement with an FP32 basis
        if (recyc && stage == 0) {
          rec = Recycle();
          rec.record = true;
          rp = &rec;
        } else if (recyc && stage == 1 &&
                   (f32b ? prepare_recycle<float>(*this, rec, restart) : prepare_recycle<double>(*this, rec, restart))) {
          rp = &rec;
        }
        if (f32b)
          gmres_core<float>(*this, r, d, tin, restart, maxit, PBFMM_MIXED, it, est, -1.0, r, rn, rp);
        else
          gmres_core<double>(*this, r, d, tin, restart, maxit, PBFMM_MIXED, it, est, -1.0, r, rn, rp);
        ++stage;
      }
      hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, stream, n, 1.0, d, 1.0, x);
    } else {
      gmres_core<double>(*this, b, x, tol * bn * 0.5, restart, maxit, PBFMM_FP64, it, est, -1.0, r, rn);
    }
  }
  sync();
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return rr <= tol ? 0 : 1;
}

void Engine::forest_segments() {
  const int S = (int)seg_off.size() - 1;
  std::vector<int32_t> seg(n);
  for (int s = 0; s < S; ++s)
    for (int64_t i = seg_off[s]; i < seg_off[s + 1]; ++i) seg[i] = s;
  up(d_seg_off, seg_off);
  up(d_seg, seg);
  alloc(d_segv, (size_t)std::max(S, 1) * 8);
  alloc(d_fb, (size_t)n * 8);
  std::vector<int64_t> rch{0};
  std::vector<int32_t> sch{0};
  for (int s = 0; s < S; ++s) {
    for (int64_t b = seg_off[s]; b < seg_off[s + 1]; b += kRedChunk) rch.push_back(std::min(b + kRedChunk, seg_off[s + 1]));
    sch.push_back((int32_t)rch.size() - 1);
  }
  n_rch = (int)rch.size() - 1;
  up(d_rch_beg, rch);
  up(d_sch_off, sch);
  alloc(d_seg_active, (size_t)std::max(S, 1));
}

void Engine::seg_norms(const double *v, std::vector<double> &out) {
  const int S = (int)seg_off.size() - 1;
  hipLaunchKernelGGL(k_seg_sumsq, dim3(S), dim3(kRedThreads), 0, stream, d_seg_off.as<int64_t>(), v,
                     d_segv.as<double>());
  out.assign(S, 0.0);
  HIPCHK(hipMemcpyAsync(out.data(), d_segv.p, (size_t)S * 8, hipMemcpyDeviceToHost, stream));
  sync();
  for (double &x : out) x = std::sqrt(x);
}

// Segmented restarted GMRES (CGS2, FP64 basis): part s of a vector of
// several independent parts (problems of a forest, or right-hand sides of one
// operator) runs its own Arnoldi process and stops when its residual estimate
// reaches tol_abs[s]; the operator is applied to all parts at once.
struct SegCtx {
  int64_t N = 0;                 // total vector length
  int S = 0, nch = 0;            // parts; part-aligned reduction chunks
  const int *seg = nullptr;      // part of each entry
  const int64_t *seg_off = nullptr, *rch = nullptr;
  const int *sch = nullptr;
  uint8_t *act = nullptr;        // per part: still iterating
  double *segv = nullptr;        // per-part scratch
  double *V = nullptr, *w = nullptr, *t1 = nullptr, *t2 = nullptr, *part = nullptr, *hd = nullptr,
         *yd = nullptr;
  std::function<void(const double *, double *)> op64, op, prec;  // op64: FP64 operator; op: inner
  std::vector<int> *part_it = nullptr;  // optional: Arnoldi steps each part used
};

static void seg_norms_ctx(Engine &e, const SegCtx &c, const double *v, std::vector<double> &out) {
  hipLaunchKernelGGL(k_seg_sumsq, dim3(c.S), dim3(kRedThreads), 0, e.stream, c.seg_off, v, c.segv);
  out.assign(c.S, 0.0);
  HIPCHK(hipMemcpyAsync(out.data(), c.segv, (size_t)c.S * 8, hipMemcpyDeviceToHost, e.stream));
  e.sync();
  for (double &x : out) x = std::sqrt(x);
}

// r0 (device) with per-part norms r0n, when given, is the residual b - A x
// of the starting x: the first cycle then skips its operator application.
static int gmres_core_seg(Engine &e, const SegCtx &c, const double *b, double *x, const std::vector<double> &tol_abs,
                          int m, int maxit, int &it, const std::vector<double> *sw = nullptr,
                          const double *r0 = nullptr, const std::vector<double> *r0n = nullptr) {
  const int64_t n = c.N;
  const int S = c.S, nch = c.nch;
  const unsigned g = nblk(n, 256);
  double *V = c.V, *w = c.w, *part = c.part, *hd = c.hd, *yd = c.yd;
  const int *seg = c.seg, *sch = c.sch;
  const int64_t *rch = c.rch;
  uint8_t *act_d = c.act;
  // Hessenberg slots per step: [h1: S x (m+2)][h2: S x (m+2)][|w|^2: S]
  const int ldh = m + 2;
  const size_t SL = (size_t)2 * S * ldh + S;
  std::vector<double> hbuf(SL * m);
  std::vector<double> H((size_t)S * (m + 1) * m), cs((size_t)S * m), sn((size_t)S * m), gv((size_t)S * (m + 1));
  std::vector<double> est(S), beta, y((size_t)S * m), hcol(m + 2), fsc(S);
  std::vector<int> ks(S);
  std::vector<uint8_t> active(S), done(S);
  auto dots = [&](int nv, const double *Vs, const double *vec, double *out, int ldo) {
    hipLaunchKernelGGL(k_mdot_seg, dim3(nch, nv), dim3(kRedThreads), 0, e.stream, rch, n, Vs, vec, part);
    hipLaunchKernelGGL(k_finish_seg, dim3(S, nv), dim3(kRedThreads), 0, e.stream, nch, sch, ldo, part, out);
  };
  // relaxed mixed precision (sw given): the FP64 operator until every
  // active part's estimate is below its switch level sw[s], then the inner
  // (FP32) operator; each cycle's starting residual in FP64
  auto low_ok = [&]() {
    if (!sw) return true;
    for (int s = 0; s < S; ++s)
      if (!done[s] && est[s] > (*sw)[s]) return false;
    return true;
  };
  while (it < maxit) {
    const double *rv = w;
    if (r0) {
      rv = r0;
      beta = *r0n;
      r0 = nullptr;
    } else {
      (sw ? c.op64 : c.op)(x, w);
      hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, e.stream, n, 1.0, b, -1.0, w);
      seg_norms_ctx(e, c, w, beta);
    }
    int nact = 0;
    for (int s = 0; s < S; ++s) {
      active[s] = beta[s] > tol_abs[s];
      done[s] = !active[s];
      nact += active[s];
      fsc[s] = active[s] ? 1.0 / beta[s] : 0.0;
      est[s] = beta[s];
      ks[s] = 0;
      std::fill(gv.begin() + (size_t)s * (m + 1), gv.begin() + (size_t)(s + 1) * (m + 1), 0.0);
      gv[(size_t)s * (m + 1)] = beta[s];
    }
    if (nact == 0) return 0;
    HIPCHK(hipMemcpyAsync(c.segv, fsc.data(), (size_t)S * 8, hipMemcpyHostToDevice, e.stream));
    HIPCHK(hipMemcpyAsync(act_d, active.data(), (size_t)S, hipMemcpyHostToDevice, e.stream));
    hipLaunchKernelGGL(k_seg_scale, dim3(g), dim3(256), 0, e.stream, n, seg, c.segv, rv, V);
    e.sync();
    int k = 0, k_done = 0, k_sync = 0;
    for (; k < m && it < maxit; ++k) {
      double *vk1 = V + (int64_t)(k + 1) * n;
      const auto &opk = low_ok() ? c.op : c.op64;
      if (c.prec) {
        c.prec(V + (int64_t)k * n, c.t1);
        opk(c.t1, vk1);
      } else {
        opk(V + (int64_t)k * n, vk1);
      }
      ++it;
      double *slot = hd + (size_t)k * SL;
      double *h1 = slot, *h2 = slot + (size_t)S * ldh, *hn = slot + (size_t)2 * S * ldh;
      dots(k + 1, V, vk1, h1, ldh);
      hipLaunchKernelGGL(k_gemv_sub_seg, dim3(g), dim3(256), 0, e.stream, n, n, k + 1, V, h1, ldh, seg, vk1);
      dots(k + 1, V, vk1, h2, ldh);
      hipLaunchKernelGGL(k_gemv_sub_seg, dim3(g), dim3(256), 0, e.stream, n, n, k + 1, V, h2, ldh, seg, vk1);
      dots(1, vk1, vk1, hn, 1);
      hipLaunchKernelGGL(k_scale_inv_seg, dim3(g), dim3(256), 0, e.stream, n, seg, hn, act_d, vk1);
      if (k < k_sync && k + 1 < m && it < maxit) continue;
      HIPCHK(hipMemcpyAsync(hbuf.data() + (size_t)k_done * SL, hd + (size_t)k_done * SL,
                            (size_t)(k + 1 - k_done) * SL * 8, hipMemcpyDeviceToHost, e.stream));
      e.sync();
      bool changed = false;
      double worst_left = 0;
      int nleft = 0;
      for (int s = 0; s < S; ++s) {
        if (done[s]) continue;
        const double est0 = est[s];
        double *Hs = &H[(size_t)s * (m + 1) * m], *c_ = &cs[(size_t)s * m], *s_ = &sn[(size_t)s * m];
        double *g_ = &gv[(size_t)s * (m + 1)];
        for (int j = k_done; j <= k; ++j) {
          const double *hb = hbuf.data() + (size_t)j * SL;
          for (int i = 0; i <= j; ++i) hcol[i] = hb[(size_t)s * ldh + i] + hb[(size_t)S * ldh + (size_t)s * ldh + i];
          const double hnv = std::sqrt(std::max(hb[(size_t)2 * S * ldh + s], 0.0));
          hcol[j + 1] = hnv;
          for (int i = 0; i < j; ++i) {
            const double a = hcol[i], bb = hcol[i + 1];
            hcol[i] = c_[i] * a + s_[i] * bb;
            hcol[i + 1] = -s_[i] * a + c_[i] * bb;
          }
          const double den = std::hypot(hcol[j], hcol[j + 1]);
          c_[j] = den > 0 ? hcol[j] / den : 1.0;
          s_[j] = den > 0 ? hcol[j + 1] / den : 0.0;
          hcol[j] = den;
          for (int i = 0; i <= j; ++i) Hs[(size_t)i * m + j] = hcol[i];
          g_[j + 1] = -s_[j] * g_[j];
          g_[j] = c_[j] * g_[j];
          est[s] = std::fabs(g_[j + 1]);
          ks[s] = j + 1;
          if (est[s] <= tol_abs[s] || hnv == 0) {
            done[s] = 1;
            active[s] = 0;
            changed = true;
            break;
          }
        }
        if (!done[s]) {
          ++nleft;
          if (est[s] < est0 && est0 > 0) {
            const double rate = std::log(est[s] / est0) / (k + 1 - k_done);
            const double goal = (sw && est[s] > (*sw)[s]) ? (*sw)[s] : tol_abs[s];
            const double left = std::log(goal / est[s]) / rate;
            worst_left = std::max(worst_left, left);
          } else {
            worst_left = std::max(worst_left, 2.0);  // no decay observed yet: look again soon
          }
        }
      }
      if (nleft == 0) {
        ++k;
        break;
      }
      if (changed) HIPCHK(hipMemcpyAsync(act_d, active.data(), (size_t)S, hipMemcpyHostToDevice, e.stream));
      k_done = k + 1;
      k_sync = k + (int)std::max(1.0, std::min(8.0, 0.5 * worst_left));
    }
    // columns of each part: those processed up to its convergence, or all
    // of this cycle's (the last step always synchronises); none for parts
    // that were converged when the cycle began
    int kmax = 0;
    for (int s = 0; s < S; ++s) {
      const int ku = beta[s] > tol_abs[s] ? ks[s] : 0;
      const double *Hs = &H[(size_t)s * (m + 1) * m], *g_ = &gv[(size_t)s * (m + 1)];
      double *ys = &y[(size_t)s * m];
      std::fill(ys, ys + m, 0.0);
      for (int i = ku - 1; i >= 0; --i) {
        double t = g_[i];
        for (int j = i + 1; j < ku; ++j) t -= Hs[(size_t)i * m + j] * ys[j];
        ys[i] = Hs[(size_t)i * m + i] != 0 ? t / Hs[(size_t)i * m + i] : 0.0;
      }
      kmax = std::max(kmax, ku);
      if (c.part_it) (*c.part_it)[s] += ku;
    }
    HIPCHK(hipMemcpyAsync(yd, y.data(), y.size() * 8, hipMemcpyHostToDevice, e.stream));
    if (c.prec) {
      hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, e.stream, c.t2, n, 0.0);
      hipLaunchKernelGGL(k_gemv_add_seg, dim3(g), dim3(256), 0, e.stream, n, n, kmax, V, yd, m, seg, c.t2);
      c.prec(c.t2, c.t1);
      hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, e.stream, n, 1.0, c.t1, 1.0, x);
    } else {
      hipLaunchKernelGGL(k_gemv_add_seg, dim3(g), dim3(256), 0, e.stream, n, n, kmax, V, yd, m, seg, x);
    }
    e.sync();
    bool all = true;
    for (int s = 0; s < S; ++s) all = all && done[s];
    if (all) return 0;
  }
  return 1;
}

// Iterative refinement around the segmented core: FP64 residual per part;
// mixed precision solves for the correction with the inner (FP32) operator,
// each part to Engine::inner_target of its own residual.
static int gmres_seg_outer(Engine &e, const SegCtx &c, const double *b, double *x, double *r, double *d, double tol,
                           int restart, int maxit, bool mixed, int *iters, double *relres) {
  const int S = c.S;
  const unsigned g = nblk(c.N, 256);
  std::vector<double> bn, rn, xn, tin(S), rr(S, 1.0);
  seg_norms_ctx(e, c, b, bn);
  // a zero starting guess needs no operator application for its residual
  seg_norms_ctx(e, c, x, xn);
  bool x_zero = true;
  for (double v : xn) x_zero = x_zero && v == 0.0;
  int it = 0;
  for (int outer = 0; outer < 100 && it < maxit; ++outer) {
    if (x_zero) {
      HIPCHK(hipMemcpyAsync(r, b, (size_t)c.N * 8, hipMemcpyDeviceToDevice, e.stream));
      rn = bn;
      x_zero = false;
    } else {
      c.op64(x, r);
      hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, e.stream, c.N, 1.0, b, -1.0, r);
      seg_norms_ctx(e, c, r, rn);
    }
    bool ok = true;
    for (int s = 0; s < S; ++s) {
      rr[s] = bn[s] > 0 ? rn[s] / bn[s] : rn[s];
      ok = ok && rr[s] <= tol;
    }
    if (ok) break;
    // parts already converged stay inactive (a margin over the FMM's
    // atomic-order rounding when the core recomputes their residual)
    if (mixed) {
      // the relaxed scheme or refinement, chosen as for one system
      const bool relaxed = e.mixed_relaxed();
      std::vector<double> sw(S);
      for (int s = 0; s < S; ++s) {
        tin[s] = rr[s] <= tol ? 2.0 * rn[s]
                 : relaxed    ? tol * bn[s] * 0.5
                              : e.inner_target(tol, bn[s], rn[s]);
        sw[s] = std::min(rn[s], e.relax_sw() * tol * bn[s]);
      }
      hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, e.stream, d, c.N, 0.0);
      gmres_core_seg(e, c, r, d, tin, restart, maxit, it, relaxed ? &sw : nullptr, r, &rn);
      hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, e.stream, c.N, 1.0, d, 1.0, x);
    } else {
      for (int s = 0; s < S; ++s) tin[s] = rr[s] <= tol ? 2.0 * rn[s] : tol * bn[s] * 0.5;
      gmres_core_seg(e, c, b, x, tin, restart, maxit, it, nullptr, r, &rn);
    }
  }
  e.sync();
  bool ok = true;
  for (int s = 0; s < S; ++s) {
    if (relres) relres[s] = rr[s];
    ok = ok && rr[s] <= tol;
  }
  if (iters) *iters = it;
  return ok ? 0 : 1;
}

int Engine::gmres_forest(const double *b, double *x, double tol, int restart, int maxit, int precision,
                         int *iters, double *relres) {
  const int S = (int)seg_off.size() - 1;
  if (restart != gm_restart) {
    alloc(d_V, (size_t)(restart + 1) * n * 8);
    alloc(d_w, n * 8);
    alloc(d_r, n * 8);
    alloc(d_x, n * 8);
    alloc(d_h, ((size_t)restart * (2 * (restart + 2) + 1) + restart + 2) * 8);
    gm_restart = restart;
  }
  const size_t SL = (size_t)2 * S * (restart + 2) + S;
  if (d_hs.bytes < SL * restart * 8) alloc(d_hs, SL * restart * 8);
  if (d_ys.bytes < (size_t)S * restart * 8) alloc(d_ys, (size_t)S * restart * 8);
  const size_t red = (size_t)n_rch * (restart + 2) * 8;
  if (d_red.bytes < red) alloc(d_red, red);
  if (!d_ptmp.p) alloc(d_ptmp, n * 8);
  if (!d_ptmp2.p) alloc(d_ptmp2, n * 8);
  const bool mixed = precision == PBFMM_MIXED && have_f32;
  struct PrecMode {  // the preconditioner's FP32 copy for a mixed solve
    bool &flag;
    bool old;
    PrecMode(bool &f, bool v) : flag(f), old(f) { flag = v; }
    ~PrecMode() { flag = old; }
  } prec_mode(prec_f32, mixed);
  SegCtx c;
  c.N = n;
  c.S = S;
  c.nch = n_rch;
  c.seg = d_seg.as<int>();
  c.seg_off = d_seg_off.as<int64_t>();
  c.rch = d_rch_beg.as<int64_t>();
  c.sch = d_sch_off.as<int>();
  c.act = d_seg_active.as<uint8_t>();
  c.segv = d_segv.as<double>();
  c.V = d_V.as<double>();
  c.w = d_w.as<double>();
  c.t1 = d_ptmp.as<double>();
  c.t2 = d_ptmp2.as<double>();
  c.part = d_red.as<double>();
  c.hd = d_hs.as<double>();
  c.yd = d_ys.as<double>();
  c.op64 = [this](const double *u, double *v) { vmult(u, v, PBFMM_FP64); };
  c.op = [this, mixed](const double *u, double *v) { vmult(u, v, mixed ? PBFMM_MIXED : PBFMM_FP64); };
  if (have_prec) c.prec = [this](const double *u, double *v) { apply_prec(u, v); };
  return gmres_seg_outer(*this, c, b, x, d_r.as<double>(), d_x.as<double>(), tol, restart, maxit, mixed, iters,
                         relres);
}

// ---------------------------------------------------------------------------
// Several right-hand sides of one operator in lockstep
// ---------------------------------------------------------------------------
void Engine::multi_setup(int R) {
  if (R == mr_R) return;
  if (R < 1) throw std::runtime_error("multi_setup: no right-hand sides");
  const int nb = tree.nb;
  if ((int64_t)R * nb >= (int64_t(1) << 31)) throw std::runtime_error("too many right-hand sides for this tree");
  // pairs of each chunk for every rhs (block ids rhs * nb + block), cut
  // again into chunks of <= kChunk1 pairs of one operator class
  auto expand = [&](const ChunkList &c, ChunkList &o) {
    o = ChunkList();
    const size_t L2 = c.level_off1.size();
    o.level_off1.assign(L2, 0);
    o.level_off.assign(L2, 0);
    for (size_t l = 0; l + 1 < L2; ++l) {
      o.level_off1[l] = (int)o.op1.size();
      for (int k = c.level_off1[l]; k < c.level_off1[l + 1]; ++k) {
        const int cnt = c.cnt1[k], total = cnt * R;
        for (int placed = 0; placed < total; placed += kChunk1) {
          const int m = std::min(kChunk1, total - placed);
          o.op1.push_back(c.op1[k]);
          o.beg1.push_back((int)o.tgt.size());
          o.cnt1.push_back(m);
          for (int q = placed; q < placed + m; ++q) {
            const int r = q / cnt, pr = c.beg1[k] + q % cnt;
            o.tgt.push_back(r * nb + c.tgt[pr]);
            o.src.push_back(r * nb + c.src[pr]);
          }
        }
      }
    }
    o.level_off1[L2 - 1] = (int)o.op1.size();
  };
  expand(far.m2m, mr_m2m_h);
  expand(far.m2l, mr_m2l_h);
  expand(far.l2l, mr_l2l_h);
  upload_chunks(mr_m2m, mr_m2m_h);
  upload_chunks(mr_m2l, mr_m2l_h);
  upload_chunks(mr_l2l, mr_l2l_h);
  const size_t ne = (size_t)R * nb * kp;
  alloc(mr_me64, ne * 8);
  alloc(mr_le64, ne * 8);
  if (have_f32) {
    alloc(mr_me32, ne * 4);
    alloc(mr_le32, ne * 4);
  }
  const int64_t N = (int64_t)R * n;
  alloc(mr_out, N * 8);
  // parts of the lockstep solve: rhs r is entries r * n .. (r + 1) * n
  std::vector<int64_t> off(R + 1), rch{0};
  std::vector<int32_t> sg(N), sch{0};
  for (int r = 0; r <= R; ++r) off[r] = (int64_t)r * n;
  for (int r = 0; r < R; ++r) {
    for (int64_t i = off[r]; i < off[r + 1]; ++i) sg[i] = r;
    for (int64_t b = off[r]; b < off[r + 1]; b += kRedChunk) rch.push_back(std::min(b + kRedChunk, off[r + 1]));
    sch.push_back((int32_t)rch.size() - 1);
  }
  mr_nch = (int)rch.size() - 1;
  up(mr_seg_off, off);
  up(mr_seg, sg);
  up(mr_rch, rch);
  up(mr_sch, sch);
  alloc(mr_act, (size_t)R);
  alloc(mr_segv, (size_t)R * 8);
  mr_restart = 0;
  mr_R = R;
}

template <typename T>
void Engine::matvec_combined_multi(const double *x, double *out, int R) {
  FarDev<T> &F = far_of<T>(*this);
  T *ME = (sizeof(T) == 4 ? mr_me32 : mr_me64).template as<T>();
  T *LE = (sizeof(T) == 4 ? mr_le32 : mr_le64).template as<T>();
  const int nb = tree.nb;
  const int64_t ne = (int64_t)R * nb * kp, sb = (int64_t)nb * kp;
  const int nleaf = (int)far.p2m_leaf.size();
  const int nrows = (int)near.rows.size();
  const int64_t nitems = (int64_t)far.m2p_item_node.size();
  // upward pass
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)std::min<int64_t>(nblk(ne, 256), 4096)), dim3(256), 0, stream, ME,
                     ne, T(0));
  if (nleaf > 0)  // every right-hand side in one launch (blockIdx.y)
    hipLaunchKernelGGL((k_p2m<T, 1>), dim3(nleaf, R), dim3(256), 0, stream, nleaf, d_p2m_leaf.as<int>(),
                       d_p2m_col_beg.as<int>(), d_p2m_col_dof.as<int>(), F.opunit.template as<T>(), kp, kpc,
                       d_cmap.as<int>(), x, nullptr, ME, (int64_t)n, sb);
  if (me_lo < 0) consumer_levels();
  for (int l = L; l > me_lo; --l)
    launch_translate<T, 0, 1>(*this, stream, mr_m2m, lvl<1>(mr_m2m_h, l), lvl<1>(mr_m2m_h, l + 1),
                              F.m2m.template as<T>(), ME, ME);
  // near field, then the far field accumulated onto it
  if (near.nleaf > 0) launch_spmv<T, 1>(spmv_variant[sizeof(T) == 4 ? 1 : 0][0], x, nullptr, out, nullptr, R);
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)std::min<int64_t>(nblk(ne, 256), 4096)), dim3(256), 0, stream, LE,
                     ne, T(0));
  launch_translate<T, 1, 1>(*this, stream, mr_m2l, 0, nchunks<1>(mr_m2l_h), F.m2l.template as<T>(), ME, LE);
  if (nitems > 0)  // every right-hand side in one launch (blockIdx.y)
    PBFMM_BY_ORDER(p, hipLaunchKernelGGL((k_m2p<T, 1, PP>), dim3((unsigned)nblk(nitems, 256), R), dim3(256), 0, stream,
                                         nitems, d_m2p_node.as<int>(), d_m2p_isrc.as<int>(), d_xyz.as<double>(),
                                         d_pmin.as<double>(), d_delta.as<double>(), p, kt, kp, ME, out, nullptr, sb,
                                         (int64_t)n));
  for (int l = std::max(2, le_lo + 1); l <= L; ++l)
    launch_translate<T, 2, 1>(*this, stream, mr_l2l, lvl<1>(mr_l2l_h, l), lvl<1>(mr_l2l_h, l + 1),
                              F.l2l.template as<T>(), LE, LE);
  if (nrows > 0)  // every right-hand side in one launch (blockIdx.y)
    PBFMM_BY_ORDER(p, hipLaunchKernelGGL((k_l2p<T, 1, PP>), dim3(nblk(nrows, 256), R), dim3(256), 0, stream, nrows,
                                         d_rows.as<int>(), d_row_leaf.as<int>(), d_leaf_block.as<int>(),
                                         d_xyz.as<double>(), d_pmin.as<double>(), d_delta.as<double>(), p, kt, kp,
                                         LE, out, nullptr, sb, (int64_t)n));
  HIPCHK(hipGetLastError());
}

// BEMProblem::vmult for R vectors at once; the pure-Neumann operator is the
// linear A0 x + mean(x) 1 of each vector (the form GMRES uses).
void Engine::vmult_multi(const double *src, double *dst, int R, int precision) {
  const unsigned g = nblk(n, 256);
  // no shift here: the operator's input is src itself
  const double *sp = src;
  double *o = mr_out.as<double>();
  if (precision == PBFMM_MIXED && have_f32)
    matvec_combined_multi<float>(sp, o, R);
  else
    matvec_combined_multi<double>(sp, o, R);
  // every right-hand side in one launch per step (blockIdx.y)
  hipLaunchKernelGGL(k_vmult_post, dim3(g, R), dim3(256), 0, stream, n, o, d_alpha.as<double>(), d_dir.as<uint8_t>(),
                     sp, dst, (int64_t)n);
  if (!have_dirichlet) {
    if (d_rhs_sum.bytes < (size_t)R * 8) alloc(d_rhs_sum, (size_t)R * 8);
    hipLaunchKernelGGL(k_rhs_sums, dim3(R), dim3(kRedThreads), 0, stream, n, src, d_rhs_sum.as<double>());
    hipLaunchKernelGGL(k_add_scaled_scalar, dim3(g, R), dim3(256), 0, stream, n, dst, d_rhs_sum.as<double>(),
                       1.0 / (double)n, (int64_t)n);
  }
  if (have_constraints)
    hipLaunchKernelGGL(k_constrain, dim3(g, R), dim3(256), 0, stream, n, d_c_is.as<uint8_t>(), d_c_ptr.as<int>(),
                       d_c_idx.as<int>(), d_c_val.as<double>(), src, dst, (int64_t)n);
  HIPCHK(hipGetLastError());
}

void Engine::apply_prec_multi(const double *v, double *z, int R) { apply_prec(v, z, R); }

int Engine::gmres_multi(const double *b, double *x, int R, double tol, int restart, int maxit, int precision,
                        int *iters, double *relres) {
  multi_setup(R);
  const int64_t N = (int64_t)R * n;
  if (restart != mr_restart) {
    alloc(mr_V, (size_t)(restart + 1) * N * 8);
    alloc(mr_w, N * 8);
    alloc(mr_r, N * 8);
    alloc(mr_d, N * 8);
    alloc(mr_t1, N * 8);
    alloc(mr_t2, N * 8);
    const size_t SL = (size_t)2 * R * (restart + 2) + R;
    alloc(mr_hs, SL * restart * 8);
    alloc(mr_ys, (size_t)R * restart * 8);
    alloc(mr_red, (size_t)mr_nch * (restart + 2) * 8);
    mr_restart = restart;
  }
  struct MeanMode {  // linear pure-Neumann operator for the duration of the solve
    bool &flag;
    bool old;
    MeanMode(bool &f, bool v) : flag(f), old(f) { flag = v; }
    ~MeanMode() { flag = old; }
  } mean_mode(neumann_mean, !have_dirichlet);
  const bool mixed = precision == PBFMM_MIXED && have_f32;
  struct PrecMode {  // the preconditioner's FP32 copy for a mixed solve
    bool &flag;
    bool old;
    PrecMode(bool &f, bool v) : flag(f), old(f) { flag = v; }
    ~PrecMode() { flag = old; }
  } prec_mode(prec_f32, mixed);
  std::vector<int> pit(R, 0);
  SegCtx c;
  c.N = N;
  c.S = R;
  c.nch = mr_nch;
  c.seg = mr_seg.as<int>();
  c.seg_off = mr_seg_off.as<int64_t>();
  c.rch = mr_rch.as<int64_t>();
  c.sch = mr_sch.as<int>();
  c.act = mr_act.as<uint8_t>();
  c.segv = mr_segv.as<double>();
  c.V = mr_V.as<double>();
  c.w = mr_w.as<double>();
  c.t1 = mr_t1.as<double>();
  c.t2 = mr_t2.as<double>();
  c.part = mr_red.as<double>();
  c.hd = mr_hs.as<double>();
  c.yd = mr_ys.as<double>();
  c.part_it = &pit;
  c.op64 = [this, R](const double *u, double *v) { vmult_multi(u, v, R, PBFMM_FP64); };
  c.op = [this, R, mixed](const double *u, double *v) { vmult_multi(u, v, R, mixed ? PBFMM_MIXED : PBFMM_FP64); };
  if (have_prec) c.prec = [this, R](const double *u, double *v) { apply_prec_multi(u, v, R); };
  int steps = 0;
  const int rc = gmres_seg_outer(*this, c, b, x, mr_r.as<double>(), mr_d.as<double>(), tol, restart, maxit, mixed,
                                 &steps, relres);
  if (iters)
    for (int r = 0; r < R; ++r) iters[r] = pit[r];
  return rc;
}

template void Engine::matvec_combined_multi<double>(const double *, double *, int);
template void Engine::matvec_combined_multi<float>(const double *, double *, int);

}  // namespace pbfmm
