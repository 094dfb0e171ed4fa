// Spatial screening of the mixed factor route, host side: rigorous cutoff
// radii per shell, the shell order of the screened AO layout and the table of
// shells per 32-function k-step that the tile-activity kernel tests.
//
// The reference evaluates every basis function at every point
// (src/evaluate_density.cu:25-411); screening is this library's own.
//
// Bound.  A shell on centre A at r (d = r - A, s = |d|) is
// phi_mu = q_mu(d) S0 with q_mu homogeneous of degree L and
// S0 = sum_p c_p N_p exp(-a_p s^2); with S1 = sum_p c_p N_p (-2 a_p) e_p and
// S2 = sum_p c_p N_p (4 a_p^2) e_p the AO kernels evaluate
//   d_k phi  = q_k S0 + d_k q S1
//   d_kl phi = q_kl S0 + (d_l q_k + d_k q_l + delta_kl q) S1 + d_k d_l q S2
//   lap phi  = (lap q) S0 + (2L + 3) q S1 + s^2 q S2.
// Every monomial x^a y^b z^c of degree n is at most s^n in magnitude, so with
// K0 = max_mu sum |coef|, K1 = max_mu,k sum |coef| a_k,
// K2 = max_mu,k,l sum |coef| a_k (a_l - delta_kl), KL = max_mu sum |coef|
// sum_k a_k (a_k - 1) over the monomial table of the shell type and
// T_n = sum_p |c_p N_p| (2 a_p)^n e_p:
//   |phi|      <= K0 s^L T0
//   |d_k phi|  <= K1 s^(L-1) T0 + K0 s^(L+1) T1
//   |d_kl phi| <= K2 s^(L-2) T0 + (2 K1 + K0) s^L T1 + K0 s^(L+2) T2
//   |lap phi|  <= KL s^(L-2) T0 + (2L + 3) K0 s^L T1 + K0 s^(L+2) T2.
// Each term is C s^n exp(-a s^2) with n <= L + 2, decreasing for
// s >= sqrt((L + 2) / (2 a_min)); beyond that point the largest of the four
// bounds decreases too, and the radius is where it falls to eps (bisection).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "ct_internal.h"
#include "generated/shell_monomials.h"

namespace ct {

namespace {

struct AngularConsts {
  double k0 = 0, k1 = 0, k2 = 0, kl = 0;
};

AngularConsts angular_consts(int code) {
  AngularConsts out;
  const int lo = kShellMonoOffset[code], hi = kShellMonoOffset[code + 1];
  int ncomp = 0;
  for (int t = lo; t < hi; ++t) ncomp = std::max(ncomp, kShellMonos[t].comp + 1);
  for (int comp = 0; comp < ncomp; ++comp) {
    double s0 = 0, s1[3] = {0, 0, 0}, s2[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, sl = 0;
    for (int t = lo; t < hi; ++t) {
      const MonoTerm& m = kShellMonos[t];
      if (m.comp != comp) continue;
      const double c = std::fabs(m.c);
      const int e[3] = {m.a, m.b, m.z};
      s0 += c;
      for (int k = 0; k < 3; ++k) {
        s1[k] += c * e[k];
        sl += c * e[k] * (e[k] - 1);
        for (int l = 0; l < 3; ++l) s2[k][l] += c * e[k] * (e[l] - (k == l ? 1 : 0));
      }
    }
    out.k0 = std::max(out.k0, s0);
    out.kl = std::max(out.kl, sl);
    for (int k = 0; k < 3; ++k) {
      out.k1 = std::max(out.k1, s1[k]);
      for (int l = 0; l < 3; ++l) out.k2 = std::max(out.k2, s2[k][l]);
    }
  }
  return out;
}

// Largest of the four bounds at distance s (cn = c_p N_p).
double shell_bound(const AngularConsts& K, int L, int nprim, const double* a, const double* cn, double s) {
  double t0 = 0, t1 = 0, t2 = 0;
  for (int p = 0; p < nprim; ++p) {
    const double e = std::fabs(cn[p]) * std::exp(-a[p] * s * s);
    t0 += e;
    t1 += 2 * a[p] * e;
    t2 += 4 * a[p] * a[p] * e;
  }
  auto pw = [&](int n) { return n <= 0 ? 1.0 : std::pow(s, n); };
  const double v = K.k0 * pw(L) * t0;
  const double g = (K.k1 > 0 ? K.k1 * pw(L - 1) * t0 : 0.0) + K.k0 * pw(L + 1) * t1;
  const double h = (K.k2 > 0 ? K.k2 * pw(L - 2) * t0 : 0.0) + (2 * K.k1 + K.k0) * pw(L) * t1 + K.k0 * pw(L + 2) * t2;
  const double lp = (K.kl > 0 ? K.kl * pw(L - 2) * t0 : 0.0) + (2 * L + 3) * K.k0 * pw(L) * t1 + K.k0 * pw(L + 2) * t2;
  return std::max(std::max(v, g), std::max(h, lp));
}

}  // namespace

double prim_radial_norm(double a, int L) {
  constexpr double pi = 3.14159265358979323846264338327950288;
  return std::sqrt(std::pow(2.0 * a / pi, 1.5) * std::pow(4.0 * a, L));
}

double shell_screen_radius(int l, int nprim, const double* a, const double* cn, double eps) {
  CT_REQUIRE(eps > 0 && std::isfinite(eps), "screening threshold must be positive and finite");
  const int L = std::abs(l);
  const AngularConsts K = angular_consts(shell_code(l));
  double amin = 1e300;
  for (int p = 0; p < nprim; ++p) amin = std::min(amin, a[p]);
  const double s_mono = std::sqrt((L + 2) / (2.0 * amin));  // all terms decrease beyond
  if (shell_bound(K, L, nprim, a, cn, s_mono) <= eps) return s_mono;
  double lo = s_mono, hi = std::max(2.0 * s_mono, 1.0);
  while (shell_bound(K, L, nprim, a, cn, hi) > eps) {
    lo = hi;
    hi *= 2.0;
    CT_REQUIRE(hi < 1e8, "screening radius does not converge");
  }
  for (int it = 0; it < 100 && hi - lo > 1e-9 * hi; ++it) {
    const double mid = 0.5 * (lo + hi);
    (shell_bound(K, L, nprim, a, cn, mid) > eps ? lo : hi) = mid;
  }
  return hi;
}

ScreenLayout build_screen_layout(const std::vector<int>& shell_l, const std::vector<int>& shell_atom,
                                 const std::vector<double>& radius, const double* xyz, int nbasis_pad) {
  const int nshell = (int)shell_l.size();
  ScreenLayout out;
  // Order: three classes of cutoff radius (largest first, split at the
  // radius tertiles), within a class by the Morton code of the shell centre
  // on a 2-bohr lattice, then by the original index.  K-steps then hold
  // shells of similar reach from one region of space, so a compact point
  // tile is out of reach of whole k-steps.
  std::vector<double> sorted_r(radius);
  std::sort(sorted_r.begin(), sorted_r.end());
  const double t1 = sorted_r[nshell / 3], t2 = sorted_r[(2 * nshell) / 3];
  double lo[3] = {1e300, 1e300, 1e300};
  for (int s = 0; s < nshell; ++s)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], xyz[3 * shell_atom[s] + k]);
  auto spread = [](uint64_t v) {
    uint64_t out = 0;
    for (int b = 0; b < 21; ++b) out |= ((v >> b) & 1ull) << (3 * b);
    return out;
  };
  std::vector<int> cls(nshell);
  std::vector<uint64_t> key(nshell);
  for (int s = 0; s < nshell; ++s) {
    cls[s] = radius[s] > t2 ? 0 : radius[s] > t1 ? 1 : 2;
    uint64_t q[3];
    for (int k = 0; k < 3; ++k) q[k] = (uint64_t)std::floor((xyz[3 * shell_atom[s] + k] - lo[k]) / 2.0);
    key[s] = (spread(q[0]) << 2) | (spread(q[1]) << 1) | spread(q[2]);
  }
  out.order.resize(nshell);
  std::iota(out.order.begin(), out.order.end(), 0);
  std::stable_sort(out.order.begin(), out.order.end(), [&](int x, int y) {
    if (cls[x] != cls[y]) return cls[x] < cls[y];
    return key[x] < key[y];
  });
  // function map and per-k-step spans
  std::vector<int> nat_first(nshell, 0);
  for (int s = 1; s < nshell; ++s) nat_first[s] = nat_first[s - 1] + shell_nfunc(shell_l[s - 1]);
  out.first.assign(nshell, 0);
  int f = 0;
  for (int i = 0; i < nshell; ++i) {
    const int s = out.order[i];
    out.first[s] = f;
    for (int c = 0; c < shell_nfunc(shell_l[s]); ++c) out.func_src.push_back(nat_first[s] + c);
    f += shell_nfunc(shell_l[s]);
  }
  const int nk = nbasis_pad / 32;
  out.span_off.assign(nk + 1, 0);
  std::vector<std::vector<int>> per_k(nk);
  for (int s = 0; s < nshell; ++s) {
    const int k0 = out.first[s] / 32, k1 = (out.first[s] + shell_nfunc(shell_l[s]) - 1) / 32;
    for (int k = k0; k <= k1; ++k) per_k[k].push_back(s);
  }
  for (int k = 0; k < nk; ++k) {
    out.span_off[k + 1] = out.span_off[k] + (int)per_k[k].size();
    for (int s : per_k[k]) {
      const double* c = xyz + 3 * shell_atom[s];
      out.spans.push_back(ScreenSpan{c[0], c[1], c[2], radius[s] * radius[s]});
    }
  }
  return out;
}

}  // namespace ct
