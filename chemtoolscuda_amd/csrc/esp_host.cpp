// Host side of the ESP path: the Boys-function Taylor table and the
// per-primitive-pair polynomial coefficients (see esp.h for the algebra).

#include <algorithm>
#include <cmath>
#include <map>

#include "esp.h"
#include "generated/shell_monomials.h"

namespace ct {

namespace {

const double kPi = 3.14159265358979323846264338327950288;

// Hermite expansion coefficients of a 1-D Gaussian product (McMurchie &
// Davidson; Helgaker, Jorgensen, Olsen eqs. 9.5.6-9.5.7):
//   E^{i+1,j}_t = E^{ij}_{t-1} / (2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}
//   E^{i,j+1}_t = E^{ij}_{t-1} / (2p) + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1}
//   E^{00}_0 = exp(-mu X_AB^2)
struct HermiteE {
  int la, lb;
  std::vector<double> v;  // [(la+1)][(lb+1)][(la+lb+1)]
  double at(int i, int j, int t) const {
    if (t < 0 || t > i + j) return 0.0;
    return v[((size_t)i * (lb + 1) + j) * (la + lb + 1) + t];
  }
  HermiteE(int la_, int lb_, double a, double b, double XAB) : la(la_), lb(lb_) {
    const int T = la + lb + 1;
    v.assign((size_t)(la + 1) * (lb + 1) * T, 0.0);
    const double p = a + b, mu = a * b / p;
    const double XPA = -b / p * XAB, XPB = a / p * XAB;
    auto ref = [&](int i, int j, int t) -> double& { return v[((size_t)i * (lb + 1) + j) * T + t]; };
    ref(0, 0, 0) = std::exp(-mu * XAB * XAB);
    for (int i = 0; i <= la; ++i)
      for (int j = 0; j <= lb; ++j) {
        if (i == 0 && j == 0) continue;
        for (int t = 0; t <= i + j; ++t) {
          if (j == 0)
            ref(i, j, t) = at(i - 1, j, t - 1) / (2 * p) + XPA * at(i - 1, j, t) + (t + 1) * at(i - 1, j, t + 1);
          else
            ref(i, j, t) = at(i, j - 1, t - 1) / (2 * p) + XPB * at(i, j - 1, t) + (t + 1) * at(i, j - 1, t + 1);
        }
      }
  }
};

// Physicists' Hermite polynomial coefficients h[t][i] (H_t(x) = sum_i h_ti x^i).
std::vector<std::vector<double>> hermite_poly(int L) {
  std::vector<std::vector<double>> h(L + 1, std::vector<double>(L + 1, 0.0));
  h[0][0] = 1.0;
  if (L >= 1) h[1][1] = 2.0;
  for (int n = 1; n < L; ++n)
    for (int i = 0; i <= L; ++i)
      h[n + 1][i] = (i > 0 ? 2.0 * h[n][i - 1] : 0.0) - 2.0 * n * h[n - 1][i];
  return h;
}

}  // namespace

namespace {

// F_0..F_{mmax}(T) and e^{-T} in long double: the convergent series
// F_m(T) = e^{-T} sum_i (2T)^i / ((2m+1)(2m+3)...(2m+2i+1)) for the top order,
// then the downward recursion F_m = (2T F_{m+1} + e^{-T}) / (2m+1).
long double boys_row(long double T, int mmax, long double* F) {
  long double term = 1.0L / (2 * mmax + 1), sum = term;
  for (int i = 1; i < 4000; ++i) {
    term *= 2.0L * T / (2 * mmax + 2 * i + 1);
    sum += term;
    if (term < 1e-22L * sum) break;
  }
  const long double eT = std::exp(-T);
  F[mmax] = eT * sum;
  for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0L * T * F[m + 1] + eT) / (2 * m + 1);
  return eT;
}

}  // namespace

namespace {

template <int LMAX>
void fill_boys_grid(double* tab) {
  using G = BoysGrid<LMAX>;
  long double F[G::ncol];
  for (int g = 0; g < G::ng; ++g) {
    const long double eT = boys_row((long double)g * G::step, G::ncol - 1, F);
    for (int m = 0; m < G::ncol; ++m) tab[(size_t)g * G::row + m] = (double)F[m];
    tab[(size_t)g * G::row + G::ncol] = (double)eT;
  }
}

}  // namespace

std::vector<double> build_boys_tables() {
  std::vector<double> tab((size_t)kBoysTableDoubles);
  fill_boys_grid<4>(tab.data());
  fill_boys_grid<8>(tab.data() + BoysGrid<4>::size);
  return tab;
}

EspData build_esp_pairs(const HostBasis& b, const std::vector<double>& xyz, const std::vector<double>& P,
                        double screen) {
  const int nshell = (int)b.shell_l.size();
  const int M = b.nbasis;
  std::vector<int> first(nshell);
  for (int s = 0, f = 0; s < nshell; ++s) {
    first[s] = f;
    f += shell_nfunc(b.shell_l[s]);
  }
  // records grouped by L
  std::map<int, std::vector<double>> byL;
  std::map<int, int64_t> countL;
  EspData out;
  const auto h = hermite_poly(2 * 4);
  for (int sI = 0; sI < nshell; ++sI) {
    for (int sJ = sI; sJ < nshell; ++sJ) {
      const int lI = b.shell_l[sI], lJ = b.shell_l[sJ];
      const int LI = std::abs(lI), LJ = std::abs(lJ), L = LI + LJ;
      if (L > kEspLmax) throw Error(CT_E_UNSUPPORTED, "ESP: pair angular momentum above 8");
      const int cI = shell_code(lI), cJ = shell_code(lJ);
      const double* A = &xyz[3 * b.shell_atom[sI]];
      const double* B = &xyz[3 * b.shell_atom[sJ]];
      const int nI = shell_nfunc(lI), nJ = shell_nfunc(lJ);
      // P block of the shell pair (symmetric part), doubled off the diagonal
      std::vector<double> Pb((size_t)nI * nJ);
      double pmax = 0.0;
      for (int i = 0; i < nI; ++i)
        for (int j = 0; j < nJ; ++j) {
          const size_t r = first[sI] + i, c = first[sJ] + j;
          double v = 0.5 * (P[r * M + c] + P[c * M + r]);
          if (sI != sJ) v *= 2.0;
          Pb[(size_t)i * nJ + j] = v;
          pmax = std::max(pmax, std::fabs(v));
        }
      if (pmax == 0.0) continue;
      const int ncoef = esp_ncoef(L);
      // coefficient slot of (N, i, j, k)
      std::vector<int> slot((size_t)(L + 1) * (L + 1) * (L + 1) * (L + 1), -1);
      auto sidx = [&](int N, int i, int j, int k) -> int& {
        return slot[(((size_t)N * (L + 1) + i) * (L + 1) + j) * (L + 1) + k];
      };
      {
        int idx = 0;
        for (int N = 0; N <= L; ++N)
          for (int tot = std::max(0, 2 * N - L); tot <= N; ++tot)
            for (int i = tot; i >= 0; --i)
              for (int j = tot - i; j >= 0; --j) sidx(N, i, j, tot - i - j) = idx++;
      }
      for (int pa = 0; pa < b.shell_nprim[sI]; ++pa) {
        const double a = b.exps[b.shell_prim_off[sI] + pa];
        const double ca = b.coefs[b.shell_prim_off[sI] + pa] * std::sqrt(std::pow(2 * a / kPi, 1.5) * std::pow(4 * a, LI));
        for (int pb = 0; pb < b.shell_nprim[sJ]; ++pb) {
          const double bb = b.exps[b.shell_prim_off[sJ] + pb];
          const double cb =
              b.coefs[b.shell_prim_off[sJ] + pb] * std::sqrt(std::pow(2 * bb / kPi, 1.5) * std::pow(4 * bb, LJ));
          const double p = a + bb;
          const HermiteE Ex(LI, LJ, a, bb, A[0] - B[0]);
          const HermiteE Ey(LI, LJ, a, bb, A[1] - B[1]);
          const HermiteE Ez(LI, LJ, a, bb, A[2] - B[2]);
          // D_tuv, packed in a cube of side L+1
          const int D1 = L + 1;
          std::vector<double> D((size_t)D1 * D1 * D1, 0.0);
          const double pre = 2.0 * kPi / p * ca * cb;
          for (int m = kShellMonoOffset[cI]; m < kShellMonoOffset[cI + 1]; ++m) {
            const MonoTerm& tm = kShellMonos[m];
            for (int n = kShellMonoOffset[cJ]; n < kShellMonoOffset[cJ + 1]; ++n) {
              const MonoTerm& tn = kShellMonos[n];
              const double w = pre * tm.c * tn.c * Pb[(size_t)tm.comp * nJ + tn.comp];
              if (w == 0.0) continue;
              for (int t = 0; t <= tm.a + tn.a; ++t) {
                const double ex = Ex.at(tm.a, tn.a, t);
                if (ex == 0.0) continue;
                for (int u = 0; u <= tm.b + tn.b; ++u) {
                  const double ey = Ey.at(tm.b, tn.b, u);
                  if (ey == 0.0) continue;
                  for (int v = 0; v <= tm.z + tn.z; ++v)
                    D[((size_t)t * D1 + u) * D1 + v] += w * ex * ey * Ez.at(tm.z, tn.z, v);
                }
              }
            }
          }
          double dsum = 0.0;
          for (double d : D) dsum += std::fabs(d);
          ++out.npairs;
          if (dsum < screen) {
            ++out.screened;
            continue;
          }
          std::vector<double> g(ncoef, 0.0);
          for (int t = 0; t <= L; ++t)
            for (int u = 0; u + t <= L; ++u)
              for (int v = 0; v + u + t <= L; ++v) {
                const double d = D[((size_t)t * D1 + u) * D1 + v];
                if (d == 0.0) continue;
                const double sgn = ((t + u + v) & 1) ? -1.0 : 1.0;
                for (int i = t & 1; i <= t; i += 2)
                  for (int j = u & 1; j <= u; j += 2)
                    for (int k = v & 1; k <= v; k += 2) {
                      const int N = (t + u + v + i + j + k) / 2;
                      g[sidx(N, i, j, k)] += d * sgn * h[t][i] * h[u][j] * h[v][k] * std::pow(p, N);
                    }
              }
          auto& rec = byL[L];
          rec.push_back((a * A[0] + bb * B[0]) / p);
          rec.push_back((a * A[1] + bb * B[1]) / p);
          rec.push_back((a * A[2] + bb * B[2]) / p);
          rec.push_back(p);
          rec.insert(rec.end(), g.begin(), g.end());
          countL[L] += 1;
        }
      }
    }
  }
  for (auto& kv : byL) {
    EspSegment sg;
    sg.L = kv.first;
    sg.count = (int32_t)countL[kv.first];
    sg.off = (int64_t)out.records.size();
    sg.off32 = 0;
    out.records.insert(out.records.end(), kv.second.begin(), kv.second.end());
    out.segs.push_back(sg);
  }
  return out;
}

std::vector<float> esp_coefficients_f32(const std::vector<double>& records, std::vector<EspSegment>& segs) {
  std::vector<float> out;
  for (auto& sg : segs) {
    const int nc = esp_ncoef(sg.L);
    sg.off32 = (int64_t)out.size();
    for (int64_t k = 0; k < sg.count; ++k) {
      const double* g = records.data() + sg.off + k * (4 + nc) + 4;
      for (int c = 0; c < nc; ++c) out.push_back((float)g[c]);
    }
  }
  return out;
}

}  // namespace ct
