// FP64 host oracle for the Gaussian-basis density-descriptor path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain CPU restatement of the
// reference algorithms; it is the checker for the HIP product path, and the
// CPU baseline that bench.py reports beside the GPU number.  Nothing in the
// product (chemtoolscuda_amd/) links, loads or calls it.
//
// Every function below cites the reference file:line whose behaviour it
// restates (paths relative to the qtchem/ChemToolsCUDA repository root).
// The algorithms are deliberately the textbook ones, written differently from
// the GPU code so that a shared mistake is unlikely:
//   * AO values and derivatives: per primitive, per Cartesian monomial, with
//     the 1-D derivative factors of x^a exp(-a x^2) multiplied out (the GPU
//     evaluates radial sums once per shell and applies generated polynomial
//     derivative code instead).
//   * Pure (solid-harmonic) functions: the explicit Helgaker table formulas,
//     expanded into monomials by a tiny polynomial algebra (the GPU tables
//     come from the general closed form for real solid harmonics).
//   * Contraction: explicit rho = sum_ij P_ij phi_i phi_j style loops (the GPU
//     uses MFMA tiles with fused epilogues).
//   * ESP: per-point nuclear-attraction matrix V_ij(C) by McMurchie-Davidson,
//     then sum_ij P_ij V_ij (the GPU pre-contracts P into Hermite densities).
//   * Boys function: power series + downward recursion / erf + upward
//     recursion (the GPU uses a Taylor table).
//
// Build: oracle/Makefile (g++ -O3 -fopenmp).  No Eigen, no BLAS.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <array>
#include <omp.h>

extern "C" {

// Flat basis description shared with oracle/oracle.py.
//   shell_l: 0,1,2,3,4 Cartesian;  -2,-3,-4 pure (reference convention,
//            src/iodata.cpp:76-87 stores pure shells with negative L).
//   coefs:   contraction coefficients of L2-normalised primitives
//            (src/iodata.cpp:106-117 reads iodata's coeffs, which refer to
//            normalised primitives).
typedef struct {
  int nshell;
  int natom;
  const int* shell_atom;
  const int* shell_l;
  const int* shell_nprim;
  const int* shell_prim_off;
  const double* exps;
  const double* coefs;
  const double* atom_xyz;  // natom x 3, row-major
} or_basis;

}  // extern "C"

namespace {

const double kPi = 3.14159265358979323846264338327950288;

// ---------------------------------------------------------------------------
// Tiny polynomial algebra: sum of c * x^a y^b z^c.
struct Mono {
  double c;
  int a, b, z;
};
using Poly = std::vector<Mono>;

Poly pmono(double c, int a, int b, int z) { return Poly{Mono{c, a, b, z}}; }

Poly padd(const Poly& p, const Poly& q) {
  Poly r = p;
  for (const Mono& m : q) {
    bool merged = false;
    for (Mono& s : r) {
      if (s.a == m.a && s.b == m.b && s.z == m.z) {
        s.c += m.c;
        merged = true;
        break;
      }
    }
    if (!merged) r.push_back(m);
  }
  return r;
}

Poly pscale(const Poly& p, double s) {
  Poly r = p;
  for (Mono& m : r) m.c *= s;
  return r;
}

Poly pmul(const Poly& p, const Poly& q) {
  Poly r;
  for (const Mono& m : p)
    for (const Mono& n : q) r = padd(r, pmono(m.c * n.c, m.a + n.a, m.b + n.b, m.z + n.z));
  return r;
}

Poly psub(const Poly& p, const Poly& q) { return padd(p, pscale(q, -1.0)); }

// Solid harmonics, Helgaker/Jorgensen/Olsen table 6.3 (p.211) as restated in
// src/cuda_basis_utils.cu:111-224; component order c0,c1,s1,c2,s2,... as in
// src/evaluate_density.cu:114,196,314.
std::vector<Poly> pure_components(int L) {
  const Poly X = pmono(1, 1, 0, 0), Y = pmono(1, 0, 1, 0), Z = pmono(1, 0, 0, 1);
  const Poly X2 = pmul(X, X), Y2 = pmul(Y, Y), Z2 = pmul(Z, Z);
  const Poly R2 = padd(padd(X2, Y2), Z2);
  std::vector<Poly> out;
  if (L == 2) {
    Poly c0 = pscale(psub(psub(pscale(Z2, 2.0), X2), Y2), 0.5);
    Poly c1 = pscale(pmul(X, Z), std::sqrt(3.0));
    Poly s1 = pscale(pmul(Y, Z), std::sqrt(3.0));
    Poly c2 = pscale(psub(X2, Y2), std::sqrt(3.0) / 2.0);
    Poly s2 = pscale(pmul(X, Y), std::sqrt(3.0));
    out = {c0, c1, s1, c2, s2};
  } else if (L == 3) {
    Poly c0 = pscale(pmul(psub(pscale(Z2, 2.0), pscale(padd(X2, Y2), 3.0)), Z), 0.5);
    Poly c1 = pscale(pmul(psub(pscale(Z2, 4.0), padd(X2, Y2)), X), std::sqrt(1.5) / 2.0);
    Poly s1 = pscale(pmul(psub(pscale(Z2, 4.0), padd(X2, Y2)), Y), std::sqrt(1.5) / 2.0);
    Poly c2 = pscale(pmul(psub(X2, Y2), Z), std::sqrt(15.0) / 2.0);
    Poly s2 = pscale(pmul(pmul(X, Y), Z), std::sqrt(15.0));
    Poly c3 = pscale(pmul(psub(X2, pscale(Y2, 3.0)), X), std::sqrt(2.5) / 2.0);
    Poly s3 = pscale(pmul(psub(pscale(X2, 3.0), Y2), Y), std::sqrt(2.5) / 2.0);
    out = {c0, c1, s1, c2, s2, c3, s3};
  } else if (L == 4) {
    Poly Z4 = pmul(Z2, Z2), R4 = pmul(R2, R2);
    Poly c0 = pscale(padd(psub(pscale(Z4, 35.0), pscale(pmul(Z2, R2), 30.0)), pscale(R4, 3.0)), 1.0 / 8.0);
    Poly c1 = pscale(pmul(psub(pscale(Z2, 7.0), pscale(R2, 3.0)), pmul(X, Z)), std::sqrt(2.5) / 2.0);
    Poly s1 = pscale(pmul(psub(pscale(Z2, 7.0), pscale(R2, 3.0)), pmul(Y, Z)), std::sqrt(2.5) / 2.0);
    Poly c2 = pscale(pmul(psub(pscale(Z2, 7.0), R2), psub(X2, Y2)), std::sqrt(5.0) / 4.0);
    Poly s2 = pscale(pmul(psub(pscale(Z2, 7.0), R2), pmul(X, Y)), std::sqrt(5.0) / 2.0);
    Poly c3 = pscale(pmul(psub(X2, pscale(Y2, 3.0)), pmul(X, Z)), std::sqrt(35.0 / 2.0) / 2.0);
    Poly s3 = pscale(pmul(psub(pscale(X2, 3.0), Y2), pmul(Y, Z)), std::sqrt(35.0 / 2.0) / 2.0);
    Poly c4 = pscale(padd(psub(pmul(X2, X2), pscale(pmul(X2, Y2), 6.0)), pmul(Y2, Y2)), std::sqrt(35.0) / 8.0);
    Poly s4 = pscale(pmul(psub(X2, Y2), pmul(X, Y)), std::sqrt(35.0) / 2.0);
    out = {c0, c1, s1, c2, s2, c3, s3, c4, s4};
  }
  return out;
}

// Cartesian exponent triples in the fchk order used by the reference:
// src/evaluate_density.cu:80 (d), :143 (f), :235-236 (g).
std::vector<std::array<int, 3>> cart_exponents(int L) {
  switch (L) {
    case 0: return {{0, 0, 0}};
    case 1: return {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    case 2: return {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    case 3:
      return {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {1, 2, 0}, {2, 1, 0},
              {2, 0, 1}, {1, 0, 2}, {0, 1, 2}, {0, 2, 1}, {1, 1, 1}};
    case 4:
      return {{0, 0, 4}, {0, 1, 3}, {0, 2, 2}, {0, 3, 1}, {0, 4, 0},
              {1, 0, 3}, {1, 1, 2}, {1, 2, 1}, {1, 3, 0}, {2, 0, 2},
              {2, 1, 1}, {2, 2, 0}, {3, 0, 1}, {3, 1, 0}, {4, 0, 0}};
  }
  return {};
}

double dfact(int n) {  // n!! with (-1)!! = 1
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// One shell component: polynomial (unnormalised) plus the factor that turns
// the primitive radial normalisation sqrt((2a/pi)^1.5 (4a)^L) into the full
// L2 normalisation: 1/sqrt(prod (2n_k-1)!!) for Cartesian
// (src/cuda_basis_utils.cu:7-83), 1/sqrt((2L-1)!!) for pure (:86-108).
struct Component {
  Poly poly;
  double norm_factor;
};

std::vector<Component> shell_components(int l) {
  std::vector<Component> out;
  if (l >= 0) {
    for (auto& e : cart_exponents(l)) {
      double f = 1.0 / std::sqrt(dfact(2 * e[0] - 1) * dfact(2 * e[1] - 1) * dfact(2 * e[2] - 1));
      out.push_back({pmono(1.0, e[0], e[1], e[2]), f});
    }
  } else {
    int L = -l;
    double f = 1.0 / std::sqrt(dfact(2 * L - 1));
    for (auto& p : pure_components(L)) out.push_back({p, f});
  }
  return out;
}

int shell_size(int l) {
  switch (l) {
    case 0: return 1;
    case 1: return 3;
    case 2: return 6;
    case -2: return 5;
    case 3: return 10;
    case -3: return 7;
    case 4: return 15;
    case -4: return 9;
  }
  return 0;
}

inline double ipow(double x, int n) {
  if (n < 0) return 0.0;
  double r = 1.0;
  for (int k = 0; k < n; ++k) r *= x;
  return r;
}

struct ShellCache {
  std::vector<std::vector<Component>> comps;  // per shell
  std::vector<int> first;                     // first basis index per shell
  int nbasis = 0;
};

ShellCache build_cache(const or_basis* b) {
  ShellCache c;
  c.comps.resize(b->nshell);
  c.first.resize(b->nshell);
  int idx = 0;
  for (int s = 0; s < b->nshell; ++s) {
    c.comps[s] = shell_components(b->shell_l[s]);
    c.first[s] = idx;
    idx += shell_size(b->shell_l[s]);
  }
  c.nbasis = idx;
  return c;
}

// AO values / derivatives at one point.
// out[d * M + i], d = 0 value; 1..3 d/dx,d/dy,d/dz; 4..9 xx,xy,xz,yy,yz,zz
// (second-derivative order of include/evaluate_hessian.cuh:28).
// Restates src/evaluate_density.cu:25-393, src/evaluate_gradient.cu:18-1126,
// src/evaluate_hessian.cu:17-1836 (sum over primitives of c_p N_p poly e^{-a r^2}).
void ao_point(const or_basis* b, const ShellCache& sc, const double* r, int nder, double* out) {
  const int M = sc.nbasis;
  for (int k = 0; k < nder * M; ++k) out[k] = 0.0;
  for (int s = 0; s < b->nshell; ++s) {
    const double* A = b->atom_xyz + 3 * b->shell_atom[s];
    const double X = r[0] - A[0], Y = r[1] - A[1], Z = r[2] - A[2];
    const double r2 = X * X + Y * Y + Z * Z;
    const int l = b->shell_l[s];
    const int L = l < 0 ? -l : l;
    const auto& comps = sc.comps[s];
    for (int p = 0; p < b->shell_nprim[s]; ++p) {
      const double al = b->exps[b->shell_prim_off[s] + p];
      const double cp = b->coefs[b->shell_prim_off[s] + p];
      const double radial_norm = std::sqrt(std::pow(2.0 * al / kPi, 1.5) * std::pow(4.0 * al, L));
      const double e = std::exp(-al * r2);
      for (size_t ic = 0; ic < comps.size(); ++ic) {
        const double pref = cp * radial_norm * comps[ic].norm_factor * e;
        const int i = sc.first[s] + (int)ic;
        for (const Mono& m : comps[ic].poly) {
          const int ex[3] = {m.a, m.b, m.z};
          const double q[3] = {X, Y, Z};
          double f0[3], f1[3], f2[3];
          for (int k = 0; k < 3; ++k) {
            const int a = ex[k];
            const double x = q[k];
            f0[k] = ipow(x, a);
            f1[k] = a * ipow(x, a - 1) - 2.0 * al * ipow(x, a + 1);
            f2[k] = a * (a - 1) * ipow(x, a - 2) - 2.0 * al * (2 * a + 1) * ipow(x, a) +
                    4.0 * al * al * ipow(x, a + 2);
          }
          const double c = pref * m.c;
          out[i] += c * f0[0] * f0[1] * f0[2];
          if (nder >= 4) {
            out[1 * M + i] += c * f1[0] * f0[1] * f0[2];
            out[2 * M + i] += c * f0[0] * f1[1] * f0[2];
            out[3 * M + i] += c * f0[0] * f0[1] * f1[2];
          }
          if (nder >= 10) {
            out[4 * M + i] += c * f2[0] * f0[1] * f0[2];
            out[5 * M + i] += c * f1[0] * f1[1] * f0[2];
            out[6 * M + i] += c * f1[0] * f0[1] * f1[2];
            out[7 * M + i] += c * f0[0] * f2[1] * f0[2];
            out[8 * M + i] += c * f0[0] * f1[1] * f1[2];
            out[9 * M + i] += c * f0[0] * f0[1] * f2[2];
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Boys function F_m(T), m = 0..mmax.
// The reference (include/boys_functions.cuh:54-72) uses erf + upward
// recursion; here: series for F_mmax with downward recursion for T <= 40,
// erf + upward recursion (stable for large T) above.
void boys_oracle(int mmax, double T, double* F) {
  if (T < 1e-14) {
    // F_m(T) = sum_k (-T)^k / (k! (2m+2k+1)); two terms are exact to 1e-28.
    for (int m = 0; m <= mmax; ++m) F[m] = 1.0 / (2 * m + 1) - T / (2 * m + 3);
    return;
  }
  const double eT = std::exp(-T);
  if (T > 40.0) {
    F[0] = 0.5 * std::sqrt(kPi / T) * std::erf(std::sqrt(T));
    for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - eT) / (2.0 * T);
    return;
  }
  // F_m(T) = e^{-T} sum_{i>=0} (2T)^i / ((2m+1)(2m+3)...(2m+2i+1))
  double term = 1.0 / (2 * mmax + 1);
  double sum = term;
  for (int i = 1; i < 2000; ++i) {
    term *= 2.0 * T / (2 * mmax + 2 * i + 1);
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  F[mmax] = eT * sum;
  for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + eT) / (2 * m + 1);
}

// Hermite expansion coefficients E^{ij}_t (McMurchie-Davidson), restating
// include/integral_coeffs.cuh:15-55.  E has shape [(la+1)][(lb+1)][(la+lb+1)].
void hermite_E(int la, int lb, double a, double b, double XAB, double* E) {
  const double p = a + b, mu = a * b / p;
  const double XPA = -b / p * XAB, XPB = a / p * XAB;
  const int T = la + lb + 1;
  auto at = [&](int i, int j, int t) -> double& { return E[(i * (lb + 1) + j) * T + t]; };
  for (int k = 0; k < (la + 1) * (lb + 1) * T; ++k) E[k] = 0.0;
  at(0, 0, 0) = std::exp(-mu * XAB * XAB);
  auto get = [&](int i, int j, int t) -> double {
    if (t < 0 || t > i + j) return 0.0;
    return at(i, j, t);
  };
  for (int i = 0; i <= la; ++i) {
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0)
          v = get(i - 1, j, t - 1) / (2 * p) + XPA * get(i - 1, j, t) + (t + 1) * get(i - 1, j, t + 1);
        else
          v = get(i, j - 1, t - 1) / (2 * p) + XPB * get(i, j - 1, t) + (t + 1) * get(i, j - 1, t + 1);
        at(i, j, t) = v;
      }
    }
  }
}

// Hermite Coulomb integrals R_tuv = R^0_tuv(p, P-C), t+u+v <= L, restating
// include/integral_coeffs.cuh:175-237.  R indexed [t][u][v] with dim L+1.
void hermite_R(int L, double p, double X, double Y, double Z, double* R0) {
  const int D = L + 1;
  std::vector<double> R((size_t)(L + 2) * D * D * D, 0.0);
  auto at = [&](int n, int t, int u, int v) -> double& { return R[(((size_t)n * D + t) * D + u) * D + v]; };
  double F[32];
  boys_oracle(L, p * (X * X + Y * Y + Z * Z), F);
  for (int n = L; n >= 0; --n) {
    at(n, 0, 0, 0) = std::pow(-2.0 * p, n) * F[n];
    for (int k = 1; k <= L - n; ++k) {
      for (int t = 0; t <= k; ++t) {
        for (int u = 0; u <= k - t; ++u) {
          const int v = k - t - u;
          double val;
          if (t > 0)
            val = (t > 1 ? (t - 1) * at(n + 1, t - 2, u, v) : 0.0) + X * at(n + 1, t - 1, u, v);
          else if (u > 0)
            val = (u > 1 ? (u - 1) * at(n + 1, t, u - 2, v) : 0.0) + Y * at(n + 1, t, u - 1, v);
          else
            val = (v > 1 ? (v - 1) * at(n + 1, t, u, v - 2) : 0.0) + Z * at(n + 1, t, u, v - 1);
          at(n, t, u, v) = val;
        }
      }
    }
  }
  for (int t = 0; t < D; ++t)
    for (int u = 0; u < D; ++u)
      for (int v = 0; v < D; ++v) R0[(t * D + u) * D + v] = (t + u + v <= L) ? at(0, t, u, v) : 0.0;
}

}  // namespace

extern "C" {

int or_nbasis(const or_basis* b) { return build_cache(b).nbasis; }

void or_boys(int mmax, const double* T, long n, double* F) {
  for (long i = 0; i < n; ++i) boys_oracle(mmax, T[i], F + i * (mmax + 1));
}

// AO arrays, out[(d * M + i) * n + g]; deriv = 0 (1 array), 1 (4), 2 (10).
void or_ao(const or_basis* b, const double* pts, long n, int deriv, double* out) {
  const ShellCache sc = build_cache(b);
  const int M = sc.nbasis;
  const int nder = deriv == 0 ? 1 : (deriv == 1 ? 4 : 10);
#pragma omp parallel
  {
    std::vector<double> buf((size_t)nder * M);
#pragma omp for schedule(static)
    for (long g = 0; g < n; ++g) {
      ao_point(b, sc, pts + 3 * g, nder, buf.data());
      for (int k = 0; k < nder * M; ++k) out[(size_t)k * n + g] = buf[k];
    }
  }
}

// Density and derivatives from P (row-major M x M, symmetric).
//   rho   = sum_ij P_ij phi_i phi_j                  (src/evaluate_density.cu:780-823)
//   grad  = 2 sum_ij P_ij d_k phi_i phi_j            (src/evaluate_gradient.cu:1318-1375)
//   hess  = 2 sum_ij P_ij (d_kl phi_i phi_j + d_k phi_i d_l phi_j)
//                                                    (src/evaluate_hessian.cu:1906-2230)
//   lap   = trace of hess                            (src/evaluate_laplacian.cu:479-821)
//   tau   = 1/2 sum_k sum_ij P_ij d_k phi_i d_k phi_j (src/evaluate_kinetic_dens.cu:123-162)
// Any output pointer may be NULL.  hess is (n,3,3) row-major.
// Points are processed in blocks of 32 so the P contraction vectorises; this
// is also the CPU baseline bench.py times.
void or_density(const or_basis* b, const double* P, const double* pts, long n, double* rho, double* grad,
                double* hess, double* lap, double* tau) {
  const ShellCache sc = build_cache(b);
  const int M = sc.nbasis;
  const bool need2 = (hess != nullptr) || (lap != nullptr);
  const bool need1 = need2 || grad != nullptr || tau != nullptr;
  const int nder = need2 ? 10 : (need1 ? 4 : 1);
  const int ncol = need1 ? 4 : 1;  // columns contracted with P: phi, d_x, d_y, d_z
  const long BS = 32;
  const long nblk = (n + BS - 1) / BS;
#pragma omp parallel
  {
    std::vector<double> ao((size_t)nder * M * BS);  // [d][i][b]
    std::vector<double> Y((size_t)ncol * M * BS);   // [c][i][b]
    std::vector<double> tmp((size_t)nder * M);
#pragma omp for schedule(dynamic, 1)
    for (long blk = 0; blk < nblk; ++blk) {
      const long g0 = blk * BS;
      const long nb = std::min(BS, n - g0);
      std::fill(ao.begin(), ao.end(), 0.0);
      for (long bb = 0; bb < nb; ++bb) {
        ao_point(b, sc, pts + 3 * (g0 + bb), nder, tmp.data());
        for (int d = 0; d < nder; ++d)
          for (int i = 0; i < M; ++i) ao[((size_t)d * M + i) * BS + bb] = tmp[(size_t)d * M + i];
      }
      // Y[c] = P * ao[c]
      std::fill(Y.begin(), Y.end(), 0.0);
      for (int c = 0; c < ncol; ++c) {
        for (int i = 0; i < M; ++i) {
          double* yi = &Y[((size_t)c * M + i) * BS];
          for (int j = 0; j < M; ++j) {
            const double pij = P[(size_t)i * M + j];
            const double* aj = &ao[((size_t)c * M + j) * BS];
            for (long bb = 0; bb < BS; ++bb) yi[bb] += pij * aj[bb];
          }
        }
      }
      for (long bb = 0; bb < nb; ++bb) {
        const long g = g0 + bb;
        auto A = [&](int d, int i) { return ao[((size_t)d * M + i) * BS + bb]; };
        auto Yv = [&](int c, int i) { return Y[((size_t)c * M + i) * BS + bb]; };
        double r = 0, gk[3] = {0, 0, 0}, dd[3][3] = {{0}}, h2[6] = {0};
        for (int i = 0; i < M; ++i) {
          const double y0 = Yv(0, i);
          r += A(0, i) * y0;
          if (need1) {
            for (int k = 0; k < 3; ++k) {
              gk[k] += A(1 + k, i) * y0;
              for (int l = 0; l < 3; ++l) dd[k][l] += A(1 + k, i) * Yv(1 + l, i);
            }
          }
          if (need2)
            for (int q = 0; q < 6; ++q) h2[q] += A(4 + q, i) * y0;
        }
        if (rho) rho[g] = r;
        if (grad)
          for (int k = 0; k < 3; ++k) grad[3 * g + k] = 2.0 * gk[k];
        if (hess) {
          const int qi[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) hess[9 * g + 3 * k + l] = 2.0 * (h2[qi[k][l]] + dd[k][l]);
        }
        if (lap) lap[g] = 2.0 * (h2[0] + h2[3] + h2[5] + dd[0][0] + dd[1][1] + dd[2][2]);
        if (tau) tau[g] = 0.5 * (dd[0][0] + dd[1][1] + dd[2][2]);
      }
    }
  }
}

// Overlap matrix S (M x M, row-major) by McMurchie-Davidson:
// S_ab = E^x_0 E^y_0 E^z_0 (pi/p)^{3/2} per primitive pair.  Used only to pin
// normalisation / ordering / pure transforms through C^T S C = I.
void or_overlap(const or_basis* b, double* S) {
  const ShellCache sc = build_cache(b);
  const int M = sc.nbasis;
  for (long k = 0; k < (long)M * M; ++k) S[k] = 0.0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int s1 = 0; s1 < b->nshell; ++s1) {
    for (int s2 = 0; s2 < b->nshell; ++s2) {
      const double* A = b->atom_xyz + 3 * b->shell_atom[s1];
      const double* B = b->atom_xyz + 3 * b->shell_atom[s2];
      const int l1 = b->shell_l[s1], l2 = b->shell_l[s2];
      const int L1 = std::abs(l1), L2 = std::abs(l2);
      const auto& c1 = sc.comps[s1];
      const auto& c2 = sc.comps[s2];
      std::vector<double> Ex((L1 + 1) * (L2 + 1) * (L1 + L2 + 1)), Ey(Ex.size()), Ez(Ex.size());
      for (int p1 = 0; p1 < b->shell_nprim[s1]; ++p1) {
        const double a = b->exps[b->shell_prim_off[s1] + p1];
        const double ca = b->coefs[b->shell_prim_off[s1] + p1] *
                          std::sqrt(std::pow(2.0 * a / kPi, 1.5) * std::pow(4.0 * a, L1));
        for (int p2 = 0; p2 < b->shell_nprim[s2]; ++p2) {
          const double bb = b->exps[b->shell_prim_off[s2] + p2];
          const double cb = b->coefs[b->shell_prim_off[s2] + p2] *
                            std::sqrt(std::pow(2.0 * bb / kPi, 1.5) * std::pow(4.0 * bb, L2));
          const double p = a + bb;
          hermite_E(L1, L2, a, bb, A[0] - B[0], Ex.data());
          hermite_E(L1, L2, a, bb, A[1] - B[1], Ey.data());
          hermite_E(L1, L2, a, bb, A[2] - B[2], Ez.data());
          const int T = L1 + L2 + 1;
          auto E0 = [&](const std::vector<double>& E, int i, int j) { return E[(i * (L2 + 1) + j) * T]; };
          const double pre = ca * cb * std::pow(kPi / p, 1.5);
          for (size_t i1 = 0; i1 < c1.size(); ++i1) {
            for (size_t i2 = 0; i2 < c2.size(); ++i2) {
              double v = 0.0;
              for (const Mono& m : c1[i1].poly)
                for (const Mono& n : c2[i2].poly)
                  v += m.c * n.c * E0(Ex, m.a, n.a) * E0(Ey, m.b, n.b) * E0(Ez, m.z, n.z);
              S[(size_t)(sc.first[s1] + i1) * M + sc.first[s2] + i2] +=
                  pre * c1[i1].norm_factor * c2[i2].norm_factor * v;
            }
          }
        }
      }
    }
  }
}

// Molecular electrostatic potential, restating
// src/evaluate_electrostatic.cu:10076-10372:
//   ESP(C) = sum_A Z_A / |C - R_A|  -  sum_ij P_ij V_ij(C),
//   V_ij(C) = <phi_i | 1/|r - C| | phi_j>  (McMurchie-Davidson, 2 pi / p prefactor,
//   src/integral_coeffs.cu:21).
void or_esp(const or_basis* b, const double* P, const double* Zc, const double* pts, long n, double* out) {
  const ShellCache sc = build_cache(b);
  const int M = sc.nbasis;
#pragma omp parallel
  {
    std::vector<double> V((size_t)M * M);
    std::vector<double> Ex, Ey, Ez, R;
#pragma omp for schedule(dynamic, 1)
    for (long g = 0; g < n; ++g) {
      const double* C = pts + 3 * g;
      std::fill(V.begin(), V.end(), 0.0);
      for (int s1 = 0; s1 < b->nshell; ++s1) {
        for (int s2 = 0; s2 < b->nshell; ++s2) {
          const double* A = b->atom_xyz + 3 * b->shell_atom[s1];
          const double* B = b->atom_xyz + 3 * b->shell_atom[s2];
          const int L1 = std::abs(b->shell_l[s1]), L2 = std::abs(b->shell_l[s2]);
          const int Lt = L1 + L2, D = Lt + 1, T = Lt + 1;
          const auto& c1 = sc.comps[s1];
          const auto& c2 = sc.comps[s2];
          Ex.assign((L1 + 1) * (L2 + 1) * T, 0.0);
          Ey = Ex;
          Ez = Ex;
          R.assign((size_t)D * D * D, 0.0);
          for (int p1 = 0; p1 < b->shell_nprim[s1]; ++p1) {
            const double a = b->exps[b->shell_prim_off[s1] + p1];
            const double ca = b->coefs[b->shell_prim_off[s1] + p1] *
                              std::sqrt(std::pow(2.0 * a / kPi, 1.5) * std::pow(4.0 * a, L1));
            for (int p2 = 0; p2 < b->shell_nprim[s2]; ++p2) {
              const double bb = b->exps[b->shell_prim_off[s2] + p2];
              const double cb = b->coefs[b->shell_prim_off[s2] + p2] *
                                std::sqrt(std::pow(2.0 * bb / kPi, 1.5) * std::pow(4.0 * bb, L2));
              const double p = a + bb;
              hermite_E(L1, L2, a, bb, A[0] - B[0], Ex.data());
              hermite_E(L1, L2, a, bb, A[1] - B[1], Ey.data());
              hermite_E(L1, L2, a, bb, A[2] - B[2], Ez.data());
              double Pc[3];
              for (int k = 0; k < 3; ++k) Pc[k] = (a * A[k] + bb * B[k]) / p;
              hermite_R(Lt, p, Pc[0] - C[0], Pc[1] - C[1], Pc[2] - C[2], R.data());
              const double pre = ca * cb * 2.0 * kPi / p;
              auto Et = [&](const std::vector<double>& E, int i, int j, int t) {
                return E[(i * (L2 + 1) + j) * T + t];
              };
              for (size_t i1 = 0; i1 < c1.size(); ++i1) {
                for (size_t i2 = 0; i2 < c2.size(); ++i2) {
                  double v = 0.0;
                  for (const Mono& m : c1[i1].poly) {
                    for (const Mono& q : c2[i2].poly) {
                      double acc = 0.0;
                      for (int t = 0; t <= m.a + q.a; ++t)
                        for (int u = 0; u <= m.b + q.b; ++u)
                          for (int w = 0; w <= m.z + q.z; ++w)
                            acc += Et(Ex, m.a, q.a, t) * Et(Ey, m.b, q.b, u) * Et(Ez, m.z, q.z, w) *
                                   R[(t * D + u) * D + w];
                      v += m.c * q.c * acc;
                    }
                  }
                  V[(size_t)(sc.first[s1] + i1) * M + sc.first[s2] + i2] +=
                      pre * c1[i1].norm_factor * c2[i2].norm_factor * v;
                }
              }
            }
          }
        }
      }
      double e = 0.0;
      for (long k = 0; k < (long)M * M; ++k) e += P[k] * V[k];
      double nuc = 0.0;
      for (int A = 0; A < b->natom; ++A) {
        const double* RA = b->atom_xyz + 3 * A;
        const double dx = C[0] - RA[0], dy = C[1] - RA[1], dz = C[2] - RA[2];
        nuc += Zc[A] / std::sqrt(dx * dx + dy * dy + dz * dz);
      }
      out[g] = nuc - e;
    }
  }
}

// Adjoint of rho with respect to P:  dL/dP_ij = sum_g w_g phi_i(g) phi_j(g).
void or_drho_dP(const or_basis* b, const double* pts, long n, const double* w, double* dP) {
  const ShellCache sc = build_cache(b);
  const int M = sc.nbasis;
  for (long k = 0; k < (long)M * M; ++k) dP[k] = 0.0;
  const int nth = omp_get_max_threads();
  std::vector<std::vector<double>> part(nth, std::vector<double>((size_t)M * M, 0.0));
#pragma omp parallel
  {
    std::vector<double> phi(M);
    std::vector<double>& acc = part[omp_get_thread_num()];
#pragma omp for schedule(static)
    for (long g = 0; g < n; ++g) {
      ao_point(b, sc, pts + 3 * g, 1, phi.data());
      for (int i = 0; i < M; ++i) {
        const double wi = w[g] * phi[i];
        for (int j = 0; j < M; ++j) acc[(size_t)i * M + j] += wi * phi[j];
      }
    }
  }
  for (int t = 0; t < nth; ++t)
    for (long k = 0; k < (long)M * M; ++k) dP[k] += part[t][k];
}

// Adjoint of rho with respect to the atomic centres:
//   dL/dR_A = -2 sum_g w_g sum_{i on A} grad phi_i(g) (P phi(g))_i
// (phi_i depends on R_A only through r - R_A).
void or_drho_dR(const or_basis* b, const double* P, const double* pts, long n, const double* w, double* dR) {
  const ShellCache sc = build_cache(b);
  const int M = sc.nbasis;
  std::vector<int> atom_of(M);
  for (int s = 0; s < b->nshell; ++s)
    for (int k = 0; k < shell_size(b->shell_l[s]); ++k) atom_of[sc.first[s] + k] = b->shell_atom[s];
  for (int k = 0; k < 3 * b->natom; ++k) dR[k] = 0.0;
  const int nth = omp_get_max_threads();
  std::vector<std::vector<double>> part(nth, std::vector<double>(3 * b->natom, 0.0));
#pragma omp parallel
  {
    std::vector<double> ao(4 * (size_t)M), y(M);
    std::vector<double>& acc = part[omp_get_thread_num()];
#pragma omp for schedule(static)
    for (long g = 0; g < n; ++g) {
      ao_point(b, sc, pts + 3 * g, 4, ao.data());
      for (int i = 0; i < M; ++i) {
        double s = 0.0;
        for (int j = 0; j < M; ++j) s += P[(size_t)i * M + j] * ao[j];
        y[i] = s;
      }
      for (int i = 0; i < M; ++i)
        for (int k = 0; k < 3; ++k) acc[3 * atom_of[i] + k] += -2.0 * w[g] * ao[(size_t)(1 + k) * M + i] * y[i];
    }
  }
  for (int t = 0; t < nth; ++t)
    for (int k = 0; k < 3 * b->natom; ++k) dR[k] += part[t][k];
}

int or_max_threads() { return omp_get_max_threads(); }

}  // extern "C"
