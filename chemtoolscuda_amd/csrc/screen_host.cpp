// Spatial screening of the factor route, host side: rigorous cutoff
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

namespace {

// Reach of every shell on a lattice of cubic cells around the molecule: bit
// (s, c) is set when cell c comes closer to the centre of shell s than its
// cutoff radius.  A stand-in for the point tiles, which the layout cannot
// know in advance.
struct Reach {
  int words = 0;
  std::vector<uint64_t> bits;  // nshell x words
  std::vector<int> count;      // cells reached per shell
  const uint64_t* row(int s) const { return bits.data() + (size_t)s * words; }
};

Reach shell_reach(const std::vector<int>& shell_atom, const std::vector<double>& radius, const double* xyz) {
  const int nshell = (int)shell_atom.size();
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int s = 0; s < nshell; ++s)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], xyz[3 * shell_atom[s] + k]);
      hi[k] = std::max(hi[k], xyz[3 * shell_atom[s] + k]);
    }
  // the atoms' box and 6 bohr around it (where grids put their points), in
  // cells of 1.5 bohr, at most 24 per axis for large molecules
  constexpr double pad = 6.0;
  double span = 0.0;
  for (int k = 0; k < 3; ++k) {
    lo[k] -= pad;
    span = std::max(span, hi[k] + pad - lo[k]);
  }
  const double cell = std::max(1.5, span / 24.0);
  int n[3];
  for (int k = 0; k < 3; ++k) n[k] = std::max(1, (int)std::ceil((hi[k] + pad - lo[k]) / cell));
  const int ncell = n[0] * n[1] * n[2];
  Reach r;
  r.words = (ncell + 63) / 64;
  r.bits.assign((size_t)nshell * r.words, 0);
  r.count.assign(nshell, 0);
  for (int s = 0; s < nshell; ++s) {
    const double* c = xyz + 3 * shell_atom[s];
    const double r2 = radius[s] * radius[s];
    uint64_t* row = r.bits.data() + (size_t)s * r.words;
    int b0[3], b1[3];  // cells whose box can come within the radius
    for (int k = 0; k < 3; ++k) {
      b0[k] = std::max(0, (int)std::floor((c[k] - radius[s] - lo[k]) / cell));
      b1[k] = std::min(n[k] - 1, (int)std::floor((c[k] + radius[s] - lo[k]) / cell));
    }
    for (int i = b0[0]; i <= b1[0]; ++i)
      for (int j = b0[1]; j <= b1[1]; ++j)
        for (int l = b0[2]; l <= b1[2]; ++l) {
          const int idx[3] = {i, j, l};
          double d2 = 0.0;
          for (int k = 0; k < 3; ++k) {
            const double a = lo[k] + idx[k] * cell, b = a + cell;
            const double d = std::max(0.0, std::max(a - c[k], c[k] - b));
            d2 += d * d;
          }
          if (d2 < r2) {
            const int cid = (i * n[1] + j) * n[2] + l;
            row[cid >> 6] |= 1ull << (cid & 63);
            ++r.count[s];
          }
        }
  }
  return r;
}

}  // namespace

ScreenLayout build_screen_layout(const std::vector<int>& shell_l, const std::vector<int>& shell_atom,
                                 const std::vector<double>& radius, const double* xyz, int nbasis_pad) {
  const int nshell = (int)shell_l.size();
  ScreenLayout out;
  // Order.  A k-step of 32 functions is needed by a point tile when any of
  // its shells reaches the tile, so the order should put shells with the
  // same reach side by side.  On the lattice stand-in (shell_reach):
  //   1. a greedy chain from the shell of widest reach, each next shell the
  //      remaining one whose reach differs from the last one's in the fewest
  //      cells;
  //   2. simulated annealing over swaps of two shells of the same size (the
  //      function offsets stay put) in different k-steps, on the sum over
  //      k-steps of the cells their shells reach: a swap is kept when that
  //      sum does not grow, or with probability exp(-growth / T), T falling
  //      linearly from 0.75 % of the lattice to 0 (fixed pseudo-random
  //      sequence: the same order on every run).
  // organic60 (288 shells): 46 % of the (tile, k-step) blocks of its grid
  // active (46-49 % over other seeds of the sequence), against 51 % for the
  // chain with a plain descent and 57 % for classes of cutoff radius in
  // Morton order; 0.17 s on the host (large300, 1125 shells: 0.7 s).
  const Reach reach = shell_reach(shell_atom, radius, xyz);
  const int W = reach.words;
  auto popc = [](uint64_t v) { return __builtin_popcountll(v); };
  out.order.clear();
  {
    std::vector<char> used(nshell, 0);
    int cur = (int)(std::max_element(reach.count.begin(), reach.count.end()) - reach.count.begin());
    for (int step = 0; step < nshell; ++step) {
      out.order.push_back(cur);
      used[cur] = 1;
      int best = -1;
      long best_d = 0;
      const uint64_t* a = reach.row(cur);
      for (int s = 0; s < nshell; ++s) {
        if (used[s]) continue;
        const uint64_t* b = reach.row(s);
        long inter = 0;
        for (int w = 0; w < W; ++w) inter += popc(a[w] & b[w]);
        const long d = reach.count[cur] + reach.count[s] - 2 * inter;
        if (best < 0 || d < best_d) {
          best = s;
          best_d = d;
        }
      }
      cur = best;
    }
  }
  {
    // positions -> k-steps they touch, k-steps -> positions
    std::vector<int> off(nshell + 1, 0);
    for (int i = 0; i < nshell; ++i) off[i + 1] = off[i] + shell_nfunc(shell_l[out.order[i]]);
    const int nkb = (off[nshell] + 31) / 32;
    std::vector<std::vector<int>> at(nkb);
    std::vector<int> k0(nshell), k1(nshell);
    for (int i = 0; i < nshell; ++i) {
      k0[i] = off[i] / 32;
      k1[i] = (off[i + 1] - 1) / 32;
      for (int k = k0[i]; k <= k1[i]; ++k) at[k].push_back(i);
    }
    std::vector<uint64_t> u(W);
    auto cost = [&](int k) {
      std::fill(u.begin(), u.end(), 0ull);
      for (int i : at[k]) {
        const uint64_t* r = reach.row(out.order[i]);
        for (int w = 0; w < W; ++w) u[w] |= r[w];
      }
      long c = 0;
      for (int w = 0; w < W; ++w) c += popc(u[w]);
      return c;
    };
    std::vector<long> kc(nkb);
    for (int k = 0; k < nkb; ++k) kc[k] = cost(k);
    uint64_t seed = 0x9E3779B97F4A7C15ull;  // fixed: the same order on every run
    auto next = [&]() {
      seed ^= seed << 13;
      seed ^= seed >> 7;
      seed ^= seed << 17;
      return seed;
    };
    const int iters = std::min(200000, 600 * nshell);
    const double t0 = 0.0075 * 64.0 * W;
    for (int it = 0; it < iters; ++it) {
      const int i = (int)(next() % nshell), j = (int)(next() % nshell);
      if (shell_nfunc(shell_l[out.order[i]]) != shell_nfunc(shell_l[out.order[j]])) continue;
      if (k0[i] == k0[j] && k1[i] == k1[j]) continue;
      int aff[4], na = 0;
      for (int k : {k0[i], k1[i], k0[j], k1[j]})
        if (std::find(aff, aff + na, k) == aff + na) aff[na++] = k;
      std::swap(out.order[i], out.order[j]);
      long delta = 0, nc[4];
      for (int a = 0; a < na; ++a) {
        nc[a] = cost(aff[a]);
        delta += nc[a] - kc[aff[a]];
      }
      const double temp = t0 * (1.0 - (double)it / iters);
      if (delta <= 0 || (temp > 0.0 && (double)(next() % 1000000) * 1e-6 < std::exp(-(double)delta / temp))) {
        for (int a = 0; a < na; ++a) kc[aff[a]] = nc[a];
      } else {
        std::swap(out.order[i], out.order[j]);
      }
    }
  }
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
  screen_layout_spans(out, shell_l, shell_atom, radius, xyz, nbasis_pad);
  return out;
}

void screen_layout_spans(ScreenLayout& out, const std::vector<int>& shell_l, const std::vector<int>& shell_atom,
                         const std::vector<double>& radius, const double* xyz, int nbasis_pad) {
  const int nshell = (int)shell_l.size();
  const int nk = nbasis_pad / 32;
  out.span_off.assign(nk + 1, 0);
  out.spans.clear();
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
}

}  // namespace ct
