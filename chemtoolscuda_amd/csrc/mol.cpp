// Host orchestration of the density-descriptor path (replaces the reference's
// chunked evaluate_* host functions, e.g. src/evaluate_density.cu:663-840,
// and the constant-memory staging of src/basis_to_gpu.cu:16-220).
//
// Per call: points are split into chunks sized from the workspace budget
// (the reference sizes chunks from free memory minus 0.5 GB,
// src/evaluate_density.cu:690-705); per chunk the AO kernel, the fused
// contraction and the FP64 epilogue run back to back on the handle's stream.
// P and the shell table are uploaded once per handle (not per call).

#include "mol.h"

#include "esp.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

namespace ct {

namespace {

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

int select_mode(uint32_t q) {
  if (q & CT_Q_HESS) return MODE_HESS;
  if (q & (CT_Q_LAP | CT_Q_TAU_PD | CT_Q_GGE | CT_Q_GENERAL_KED | CT_Q_KED_ALPHA)) return MODE_LAP;
  if (q & (CT_Q_GRAD | CT_Q_RDG | CT_Q_WEIZSACKER)) return MODE_GRAD;
  return MODE_RHO;
}

struct OutSlot {
  uint32_t bit;
  double* ct_outputs::*field;
  int width;
};

const OutSlot kSlots[] = {
    {CT_Q_RHO, &ct_outputs::rho, 1},
    {CT_Q_GRAD, &ct_outputs::grad, 3},
    {CT_Q_HESS, &ct_outputs::hess, 9},
    {CT_Q_LAP, &ct_outputs::lap, 1},
    {CT_Q_TAU_PD, &ct_outputs::tau_pd, 1},
    {CT_Q_RDG, &ct_outputs::rdg, 1},
    {CT_Q_SHANNON, &ct_outputs::shannon, 1},
    {CT_Q_WEIZSACKER, &ct_outputs::weizsacker, 1},
    {CT_Q_THOMAS_FERMI, &ct_outputs::thomas_fermi, 1},
    {CT_Q_GGE, &ct_outputs::gge, 1},
    {CT_Q_GENERAL_KED, &ct_outputs::general_ked, 1},
    {CT_Q_KED_ALPHA, &ct_outputs::ked_alpha, 1},
};

}  // namespace

int shell_code(int l) {
  switch (l) {
    case 0: return 0;
    case 1: return 1;
    case 2: return 2;
    case 3: return 3;
    case 4: return 4;
    case -2: return 5;
    case -3: return 6;
    case -4: return 7;
  }
  throw Error(CT_E_INVALID, "unsupported shell angular momentum " + std::to_string(l));
}

int shell_nfunc(int l) {
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
  throw Error(CT_E_INVALID, "unsupported shell angular momentum " + std::to_string(l));
}

bool is_device_pointer(const void* ptr) {
  if (ptr == nullptr) return false;
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  hipError_t e = hipPointerGetAttributes(&attr, ptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

namespace {
// Process-wide cap on the device memory held by the library's buffers
// (ct_set_device_allocation_limit; -1: none) and what they hold now.
std::atomic<int64_t> g_alloc_cap{-1};
std::atomic<int64_t> g_alloc_held{0};
}  // namespace

// Set while an evaluation runs inside a caller's region (ct_eval_noalloc):
// any buffer growth then is a bug and fails loudly instead of allocating.
thread_local bool t_no_alloc = false;

struct NoAllocScope {
  bool prev;
  explicit NoAllocScope(bool on) : prev(t_no_alloc) { t_no_alloc = prev || on; }
  ~NoAllocScope() { t_no_alloc = prev; }
};

void set_device_allocation_limit(int64_t bytes) { g_alloc_cap.store(bytes); }
int64_t device_allocation_held() { return g_alloc_held.load(); }

bool DeviceBuffer::reserve(size_t bytes) {
  if (bytes <= bytes_) return true;
  if (t_no_alloc) throw Error(CT_E_INVALID, "internal error: device allocation inside the allocation-free path");
  release();
  const int64_t cap = g_alloc_cap.load();
  if (cap >= 0 && g_alloc_held.load() + (int64_t)bytes > cap) return false;  // as if hipMalloc had failed
  if (hipMalloc(&ptr_, bytes) != hipSuccess) {
    (void)hipGetLastError();
    ptr_ = nullptr;
    return false;
  }
  bytes_ = bytes;
  g_alloc_held += (int64_t)bytes;
  return true;
}

void DeviceBuffer::release() {
  if (ptr_) {
    (void)hipFree(ptr_);
    g_alloc_held -= (int64_t)bytes_;
  }
  ptr_ = nullptr;
  bytes_ = 0;
}

void PinnedBuffer::reserve(size_t bytes) {
  if (bytes <= bytes_) return;
  if (t_no_alloc) throw Error(CT_E_INVALID, "internal error: host allocation inside the allocation-free path");
  release();
  CT_HIP_CHECK(hipHostMalloc(&ptr_, bytes, hipHostMallocDefault));
  bytes_ = bytes;
}

void PinnedBuffer::release() {
  if (ptr_) (void)hipHostFree(ptr_);
  ptr_ = nullptr;
  bytes_ = 0;
}

KernelTimer::~KernelTimer() {
  for (auto& p : pending_) {
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t KernelTimer::take() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  CT_HIP_CHECK(hipEventCreate(&e));
  return e;
}

void KernelTimer::enable(bool on) {
  on_ = on;
  for (int k = 0; k < NKIND; ++k) {
    ms_[k] = 0.0;
    calls_[k] = 0;
  }
}

void KernelTimer::begin(Kind k, hipStream_t s) {
  if (!on_) return;
  Pending p{k, take(), take()};
  CT_HIP_CHECK(hipEventRecord(p.a, s));
  pending_.push_back(p);
}

void KernelTimer::end(hipStream_t s) {
  if (!on_ || pending_.empty()) return;
  CT_HIP_CHECK(hipEventRecord(pending_.back().b, s));
}

void KernelTimer::collect() {
  for (auto& p : pending_) {
    float ms = 0.f;
    CT_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    ms_[p.kind] += ms;
    calls_[p.kind] += 1;
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  pending_.clear();
}

ct_kernel_stats KernelTimer::stats() const {
  ct_kernel_stats st;
  st.ao_ms = ms_[AO];
  st.contract_ms = ms_[CONTRACT];
  st.finalize_ms = ms_[FINALIZE];
  st.esp_ms = ms_[ESP];
  st.backward_ms = ms_[BACKWARD];
  st.reeval_ms = ms_[REEVAL];
  st.ao_calls = calls_[AO];
  st.contract_calls = calls_[CONTRACT];
  st.finalize_calls = calls_[FINALIZE];
  st.esp_calls = calls_[ESP];
  st.backward_calls = calls_[BACKWARD];
  st.reeval_calls = calls_[REEVAL];
  st.screen_ms = ms_[SCREEN];
  st.screen_calls = calls_[SCREEN];
  return st;
}

Mol::Mol(const HostBasis& basis, const double* atom_xyz, const double* charges, const int32_t* numbers, int natom,
         const double* P)
    : basis_(basis), natom_(natom) {
  CT_REQUIRE(natom > 0, "molecule needs at least one atom");
  nshell_ = (int)basis.shell_l.size();
  CT_REQUIRE(nshell_ > 0, "basis has no shells");
  CT_REQUIRE(basis.shell_atom.size() == (size_t)nshell_ && basis.shell_nprim.size() == (size_t)nshell_ &&
                 basis.shell_prim_off.size() == (size_t)nshell_,
             "inconsistent shell arrays");
  M_ = 0;
  meta_.resize(nshell_);
  for (int s = 0; s < nshell_; ++s) {
    const int l = basis.shell_l[s];
    CT_REQUIRE(basis.shell_atom[s] >= 0 && basis.shell_atom[s] < natom, "shell centre index out of range");
    CT_REQUIRE(basis.shell_nprim[s] > 0, "shell without primitives");
    CT_REQUIRE(basis.shell_prim_off[s] >= 0 &&
                   (size_t)basis.shell_prim_off[s] + basis.shell_nprim[s] <= basis.exps.size() &&
                   basis.exps.size() == basis.coefs.size(),
               "primitive offsets out of range");
    meta_[s] = ShellMeta{shell_code(l), basis.shell_nprim[s], basis.shell_prim_off[s], M_};
    lmax_ = std::max(lmax_, std::abs(l));
    M_ += shell_nfunc(l);
  }
  basis_.nbasis = M_;
  Mp_ = (int)round_up(M_, kRowBlock);
  prim_.resize(basis.exps.size());
  for (int s = 0; s < nshell_; ++s) {
    const int L = std::abs(basis.shell_l[s]);
    for (int k = 0; k < basis.shell_nprim[s]; ++k) {
      const int j = basis.shell_prim_off[s] + k;
      const double a = basis.exps[j];
      CT_REQUIRE(a > 0.0 && std::isfinite(a), "exponents must be positive and finite");
      prim_[j] = PrimData{a, basis.coefs[j] * prim_radial_norm(a, L)};
    }
  }
  xyz_.assign(atom_xyz, atom_xyz + 3 * (size_t)natom);
  charges_.assign(natom, 0.0);
  if (charges) charges_.assign(charges, charges + natom);
  numbers_.assign(natom, 0);
  if (numbers) numbers_.assign(numbers, numbers + natom);
  P_.assign(P, P + (size_t)M_ * M_);

  CT_HIP_CHECK(hipGetDevice(&device_));
  CT_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  // default workspace budget: a quarter of the device's memory, at most
  // 48 GiB (one chunk of 2.7M points for rho, grad, lap and tau of a
  // 624-function basis on a 288 GB device; fewer, larger chunks run fewer
  // sorts and launches per point)
  {
    size_t free_b = 0, total_b = 0;
    CT_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    ws_limit_ = std::max<int64_t>(int64_t(1) << 30, std::min<int64_t>(int64_t(48) << 30, (int64_t)(total_b / 4)));
  }
  if (!d_meta_.reserve(sizeof(ShellMeta) * nshell_) || !d_prim_.reserve(sizeof(PrimData) * prim_.size()) ||
      !d_geom_.reserve(sizeof(ShellGeom) * nshell_) || !d_P_.reserve(sizeof(double) * (size_t)Mp_ * Mp_))
    throw Error(CT_E_NOMEM, "cannot allocate the device shell table / density matrix");
  CT_HIP_CHECK(hipMemcpy(d_meta_.get(), meta_.data(), sizeof(ShellMeta) * nshell_, hipMemcpyHostToDevice));
  CT_HIP_CHECK(hipMemcpy(d_prim_.get(), prim_.data(), sizeof(PrimData) * prim_.size(), hipMemcpyHostToDevice));
  upload_geometry();
  upload_P();
}

Mol::~Mol() {
  for (hipStream_t s : {stream_, h2d_stream_, d2h_stream_, re_stream_})
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  for (auto& slot : pipe_ev_)
    for (hipEvent_t e : slot)
      if (e) (void)hipEventDestroy(e);
}

void Mol::ensure_pipeline() {
  if (!h2d_stream_) CT_HIP_CHECK(hipStreamCreateWithFlags(&h2d_stream_, hipStreamNonBlocking));
  if (!d2h_stream_) CT_HIP_CHECK(hipStreamCreateWithFlags(&d2h_stream_, hipStreamNonBlocking));
  if (!re_stream_) CT_HIP_CHECK(hipStreamCreateWithFlags(&re_stream_, hipStreamNonBlocking));
  for (auto& slot : pipe_ev_)
    for (hipEvent_t& e : slot)
      if (!e) CT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

void Mol::upload_geometry() {
  std::vector<ShellGeom> geom(nshell_);
  for (int s = 0; s < nshell_; ++s) {
    const int a = basis_.shell_atom[s];
    double amin = 1e300;
    for (int k = 0; k < basis_.shell_nprim[s]; ++k) amin = std::min(amin, basis_.exps[basis_.shell_prim_off[s] + k]);
    geom[s] = ShellGeom{xyz_[3 * a], xyz_[3 * a + 1], xyz_[3 * a + 2], amin};
  }
  CT_HIP_CHECK(hipMemcpy(d_geom_.get(), geom.data(), sizeof(ShellGeom) * nshell_, hipMemcpyHostToDevice));
  // the screened shell order follows the centres: rebuilt (annealing, a
  // fraction of a second) once some centre has moved by more than
  // kLayoutMove since it was built; smaller moves keep the order and refresh
  // the k-step tables with the new centres
  double moved = 1e300;
  if (xyz_layout_.size() == xyz_.size()) {
    moved = 0.0;
    for (int a = 0; a < natom_; ++a) {
      double d2 = 0.0;
      for (int k = 0; k < 3; ++k) d2 += (xyz_[3 * a + k] - xyz_layout_[3 * a + k]) * (xyz_[3 * a + k] - xyz_layout_[3 * a + k]);
      moved = std::max(moved, std::sqrt(d2));
    }
  }
  if (moved > kLayoutMove) layout_stale_ = true;
  if (screen_built_) build_screen();
}

namespace {
uint16_t bf16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf16_float(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}
}  // namespace

bool pivoted_cholesky(const double* A, int n, int lda, int rmax, double rel_tol, std::vector<double>& L, int& rank) {
  L.clear();
  rank = 0;
  std::vector<double> d(n);
  double dmax = 0.0;
  for (int i = 0; i < n; ++i) {
    d[i] = A[(size_t)i * lda + i];
    dmax = std::max(dmax, std::fabs(d[i]));
  }
  if (dmax == 0.0) return true;
  const double tol = rel_tol * dmax;
  std::vector<char> used(n, 0);
  for (;;) {
    int p = -1;
    for (int i = 0; i < n; ++i)
      if (!used[i] && (p < 0 || d[i] > d[p])) p = i;
    if (p < 0 || d[p] <= tol) break;
    if (rank >= rmax) return false;
    // column rank: (A[:, p] - L L[p, :]^T) / sqrt(d[p]); rows pivoted earlier are 0
    const double piv = std::sqrt(d[p]);
    std::vector<double> col(n, 0.0);
    for (int i = 0; i < n; ++i) {
      if (used[i]) continue;
      double v = A[(size_t)i * lda + p];
      for (int k = 0; k < rank; ++k) v -= L[(size_t)k * n + i] * L[(size_t)k * n + p];
      col[i] = v / piv;
    }
    col[p] = piv;
    for (int i = 0; i < n; ++i)
      if (!used[i]) d[i] -= col[i] * col[i];
    used[p] = 1;
    L.insert(L.end(), col.begin(), col.end());
    ++rank;
  }
  // positive semidefinite: what is left of the diagonal must not be negative
  for (int i = 0; i < n; ++i)
    if (!used[i] && d[i] < -tol) return false;
  return true;
}

void Mol::upload_P() {
  // symmetrise: the descriptors depend on the symmetric part of P only
  std::vector<double> Pp((size_t)Mp_ * Mp_, 0.0);
  for (int i = 0; i < M_; ++i)
    for (int j = 0; j < M_; ++j)
      Pp[(size_t)i * Mp_ + j] = 0.5 * (P_[(size_t)i * M_ + j] + P_[(size_t)j * M_ + i]);
  CT_HIP_CHECK(hipMemcpy(d_P_.get(), Pp.data(), sizeof(double) * Pp.size(), hipMemcpyHostToDevice));
  // mixed path: FP32 P split exactly into three bf16 pieces
  const size_t plane = (size_t)Mp_ * Mp_;
  std::vector<uint16_t> Pb(3 * plane);
  for (size_t k = 0; k < plane; ++k) {
    const float x = (float)Pp[k];
    const uint16_t h = bf16_bits(x);
    const float r1 = x - bf16_float(h);
    const uint16_t m = bf16_bits(r1);
    const uint16_t l = bf16_bits(r1 - bf16_float(m));
    Pb[k] = h;
    Pb[plane + k] = m;
    Pb[2 * plane + k] = l;
  }
  std::vector<uint16_t> Pf(Pb.size());
  mixed_fragment_order(Pb.data(), Mp_, Mp_, Pf.data());
  if (!d_Pb_.reserve(sizeof(uint16_t) * Pf.size())) throw Error(CT_E_NOMEM, "cannot allocate the bf16 density matrix");
  CT_HIP_CHECK(hipMemcpy(d_Pb_.get(), Pf.data(), sizeof(uint16_t) * Pf.size(), hipMemcpyHostToDevice));

  // mixed path, factor route: P = L L^T (pivoted Cholesky, to 1e-12 of the
  // largest diagonal element), kept when P is positive semidefinite and its
  // rank is at most about half the basis (beyond that the density route needs
  // fewer products in every mode); L^T padded to a multiple of 128 rows, in
  // FP32 split into three bf16 pieces like P.
  factor_rank_ = -1;
  std::vector<double> L;
  int rank = 0;
  if (pivoted_cholesky(Pp.data(), M_, Mp_, Mp_ / 2 + 1, 1e-12, L, rank)) {
    const int rows = std::max(kRowBlock, (rank + kRowBlock - 1) / kRowBlock * kRowBlock);
    // FP64 copy for the FP64 factor route (re-evaluation of the mixed factor
    // route's flagged points; FP64 evaluation on request): L k-major,
    // L64[k * rows + j], padded with zeros
    std::vector<double> L64((size_t)Mp_ * rows, 0.0);
    for (int j = 0; j < rank; ++j)
      for (int i = 0; i < M_; ++i) L64[(size_t)i * rows + j] = L[(size_t)j * M_ + i];
    if (!d_L64_.reserve(sizeof(double) * L64.size())) throw Error(CT_E_NOMEM, "cannot allocate the factor of P");
    CT_HIP_CHECK(hipMemcpy(d_L64_.get(), L64.data(), sizeof(double) * L64.size(), hipMemcpyHostToDevice));
    // the remainder R = P - L L^T must itself be positive semidefinite: with
    // its diagonal at most tol that means |R_ij| <= sqrt(R_ii R_jj) for all
    // i, j, up to rounding (the pivoted Cholesky only looked at the diagonal)
    double dmax = 0.0;
    std::vector<double> rdiag(M_);
    for (int i = 0; i < M_; ++i) {
      double r = Pp[(size_t)i * Mp_ + i];
      dmax = std::max(dmax, std::fabs(r));
      for (int j = 0; j < rank; ++j) r -= L[(size_t)j * M_ + i] * L[(size_t)j * M_ + i];
      rdiag[i] = r;
    }
    DeviceBuffer chk;
    if (!chk.reserve(sizeof(double) * (M_ + 1))) throw Error(CT_E_NOMEM, "cannot allocate the factor check");
    double* d_diag = static_cast<double*>(chk.get());
    auto* d_viol = reinterpret_cast<unsigned long long*>(d_diag + M_);
    CT_HIP_CHECK(hipMemcpyAsync(d_diag, rdiag.data(), sizeof(double) * M_, hipMemcpyHostToDevice, stream_));
    launch_fill_zero(d_viol, 1, stream_);
    launch_factor_residual(static_cast<const double*>(d_P_.get()), Mp_, static_cast<const double*>(d_L64_.get()),
                           rows, rank, M_, d_diag, d_viol, stream_);
    unsigned long long vbits = 0;
    CT_HIP_CHECK(hipMemcpyAsync(&vbits, d_viol, sizeof(vbits), hipMemcpyDeviceToHost, stream_));
    CT_HIP_CHECK(hipStreamSynchronize(stream_));
    double viol;
    std::memcpy(&viol, &vbits, sizeof(viol));
    const double slack = 1e-12 * dmax + 4.0 * (rank + 2) * 2.220446049250313e-16 * dmax;
    if (viol <= slack) {
      factor_rank_ = rank;
      factor_rows_ = rows;
    }
  }
  L_.swap(L);
  if (factor_rank_ < 0) L_.clear();
  build_screen();
}

void Mol::set_screening(double eps) {
  CT_REQUIRE(eps >= 0.0 && std::isfinite(eps), "screening threshold must be finite and non-negative");
  CT_HIP_CHECK(hipStreamSynchronize(stream_));
  screen_eps_ = eps;
  layout_stale_ = true;
  build_screen();
}

void Mol::set_morton_frame() {
  // Morton lattice: the atoms' box with a margin of 12 bohr, 256 cells per
  // axis, cells of at least 0.25 bohr (points outside clamp to the border;
  // order within a cell is the input order, which only matters below the
  // size of a point tile)
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int a = 0; a < natom_; ++a)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], xyz_[3 * a + k]);
      hi[k] = std::max(hi[k], xyz_[3 * a + k]);
    }
  double span = 0.0;
  for (int k = 0; k < 3; ++k) {
    frame_.origin[k] = lo[k] - 12.0;
    span = std::max(span, hi[k] - lo[k] + 24.0);
  }
  frame_.inv_cell = 1.0 / std::max(0.25, span / 256.0);
}

void Mol::build_screen() {
  // the shell order and the radii depend on the centres and the threshold
  // only: a new P keeps them
  screen_built_ = true;
  // shell order: the screened one, or the natural one when screening is off
  std::vector<int> shell_l(basis_.shell_l.begin(), basis_.shell_l.end());
  std::vector<int> shell_atom(basis_.shell_atom.begin(), basis_.shell_atom.end());
  std::vector<double>& radius = shell_radius_;
  if (screen_eps_ > 0.0 && layout_stale_) {
    radius.assign(nshell_, 0.0);
    for (int s = 0; s < nshell_; ++s) {
      const ShellMeta& m = meta_[s];
      std::vector<double> a(m.nprim), cn(m.nprim);
      for (int k = 0; k < m.nprim; ++k) {
        a[k] = prim_[m.prim_off + k].a;
        cn[k] = prim_[m.prim_off + k].c;
      }
      radius[s] = shell_screen_radius(shell_l[s], m.nprim, a.data(), cn.data(), screen_eps_);
    }
    layout_ = build_screen_layout(shell_l, shell_atom, radius, xyz_.data(), Mp_);
    xyz_layout_ = xyz_;
    set_morton_frame();
    // the FP64 path's cutoff radii, at its tighter threshold
    shell_radius64_.assign(nshell_, 0.0);
    for (int s = 0; s < nshell_; ++s) {
      const ShellMeta& m = meta_[s];
      std::vector<double> a(m.nprim), cn(m.nprim);
      for (int k = 0; k < m.nprim; ++k) {
        a[k] = prim_[m.prim_off + k].a;
        cn[k] = prim_[m.prim_off + k].c;
      }
      shell_radius64_[s] = shell_screen_radius(shell_l[s], m.nprim, a.data(), cn.data(), screen_eps_ * kFp64ScreenFactor);
    }
  } else if (screen_eps_ > 0.0) {
    // same order, current centres (small moves since the order was built)
    screen_layout_spans(layout_, shell_l, shell_atom, radius, xyz_.data(), Mp_);
    set_morton_frame();
  } else {
    layout_ = ScreenLayout{};
    layout_.order.resize(nshell_);
    layout_.first.resize(nshell_);
    for (int s = 0; s < nshell_; ++s) {
      layout_.order[s] = s;
      layout_.first[s] = meta_[s].first;
    }
    layout_.func_src.resize(M_);
    for (int i = 0; i < M_; ++i) layout_.func_src[i] = i;
  }
  layout_stale_ = false;
  // shell table in that order
  std::vector<ShellMeta> sm(nshell_);
  std::vector<ShellGeom> sg(nshell_);
  for (int i = 0; i < nshell_; ++i) {
    const int s = layout_.order[i];
    sm[i] = ShellMeta{meta_[s].code, meta_[s].nprim, meta_[s].prim_off, layout_.first[s]};
    const int a = basis_.shell_atom[s];
    double amin = 1e300;
    for (int k = 0; k < meta_[s].nprim; ++k) amin = std::min(amin, prim_[meta_[s].prim_off + k].a);
    // screened: the AO kernel writes a shell as zeros, without its
    // exponentials, when every point of a wave has amin r^2 > 50; with
    // amin >= 50 / R_s^2 that holds beyond the cutoff radius too, where every
    // value is at most eps (covered by the same truncation bound as a
    // skipped k-step: each function is either computed or below eps)
    if (screen_eps_ > 0.0) amin = std::max(amin, 50.0 / (radius[s] * radius[s]));
    sg[i] = ShellGeom{xyz_[3 * a], xyz_[3 * a + 1], xyz_[3 * a + 2], amin};
  }
  if (!d_smeta_.reserve(sizeof(ShellMeta) * nshell_) || !d_sgeom_.reserve(sizeof(ShellGeom) * nshell_))
    throw Error(CT_E_NOMEM, "cannot allocate the screened shell table");
  CT_HIP_CHECK(hipMemcpy(d_smeta_.get(), sm.data(), sizeof(ShellMeta) * nshell_, hipMemcpyHostToDevice));
  CT_HIP_CHECK(hipMemcpy(d_sgeom_.get(), sg.data(), sizeof(ShellGeom) * nshell_, hipMemcpyHostToDevice));
  if (screen_eps_ > 0.0) {
    if (!d_span_off_.reserve(sizeof(int32_t) * layout_.span_off.size()) ||
        !d_spans_.reserve(sizeof(ScreenSpan) * std::max<size_t>(1, layout_.spans.size())))
      throw Error(CT_E_NOMEM, "cannot allocate the screening tables");
    CT_HIP_CHECK(hipMemcpy(d_span_off_.get(), layout_.span_off.data(), sizeof(int32_t) * layout_.span_off.size(),
                           hipMemcpyHostToDevice));
    CT_HIP_CHECK(hipMemcpy(d_spans_.get(), layout_.spans.data(), sizeof(ScreenSpan) * layout_.spans.size(),
                           hipMemcpyHostToDevice));
    // the same k-steps with the FP64 path's radii
    ScreenLayout l64;
    l64.first = layout_.first;
    screen_layout_spans(l64, shell_l, shell_atom, shell_radius64_, xyz_.data(), Mp_);
    if (!d_span_off64_.reserve(sizeof(int32_t) * l64.span_off.size()) ||
        !d_spans64_.reserve(sizeof(ScreenSpan) * std::max<size_t>(1, l64.spans.size())))
      throw Error(CT_E_NOMEM, "cannot allocate the screening tables");
    CT_HIP_CHECK(hipMemcpy(d_span_off64_.get(), l64.span_off.data(), sizeof(int32_t) * l64.span_off.size(),
                           hipMemcpyHostToDevice));
    CT_HIP_CHECK(hipMemcpy(d_spans64_.get(), l64.spans.data(), sizeof(ScreenSpan) * l64.spans.size(),
                           hipMemcpyHostToDevice));
  }
  if (factor_rank_ < 0) return;
  // L^T with its columns (basis functions) in that order, FP32 split into
  // three bf16 pieces, in MFMA fragment order
  const int rank = factor_rank_;
  const size_t fplane = (size_t)factor_rows_ * Mp_;
  std::vector<uint16_t> Lb(3 * fplane, 0);
  for (int j = 0; j < rank; ++j)
    for (int i = 0; i < M_; ++i) {
      const float x = (float)L_[(size_t)j * M_ + layout_.func_src[i]];  // row j of L^T
      const uint16_t h = bf16_bits(x);
      const float r1 = x - bf16_float(h);
      const uint16_t m = bf16_bits(r1);
      const uint16_t l = bf16_bits(r1 - bf16_float(m));
      const size_t k = (size_t)j * Mp_ + i;
      Lb[k] = h;
      Lb[fplane + k] = m;
      Lb[2 * fplane + k] = l;
    }
  std::vector<uint16_t> Lf(Lb.size());
  mixed_fragment_order(Lb.data(), factor_rows_, Mp_, Lf.data());
  if (!d_Lb_.reserve(sizeof(uint16_t) * Lf.size())) throw Error(CT_E_NOMEM, "cannot allocate the bf16 factor of P");
  CT_HIP_CHECK(hipMemcpy(d_Lb_.get(), Lf.data(), sizeof(uint16_t) * Lf.size(), hipMemcpyHostToDevice));
  // FP64 copy in the same function order, k-major (L64s[i * rows + j] =
  // L_ij for screened function i): the FP64 factor route and the
  // FP64 factor route in the screened order (also the re-evaluation of the
  // mixed factor route's flagged points)
  std::vector<double> L64s((size_t)Mp_ * factor_rows_, 0.0);
  for (int j = 0; j < rank; ++j)
    for (int i = 0; i < M_; ++i) L64s[(size_t)i * factor_rows_ + j] = L_[(size_t)j * M_ + layout_.func_src[i]];
  if (!d_L64s_.reserve(sizeof(double) * L64s.size())) throw Error(CT_E_NOMEM, "cannot allocate the factor of P");
  CT_HIP_CHECK(hipMemcpy(d_L64s_.get(), L64s.data(), sizeof(double) * L64s.size(), hipMemcpyHostToDevice));
  // Truncation bound: a skipped AO value is at most eps, so the skipped part
  // of psi_j is at most t_j = eps ||L_j||_1 and rho moves by at most
  // sum_j t_j (2 |psi_j| + t_j) <= 2 eps ||c||_2 sqrt(rho) + O(eps^2)
  // (Cauchy-Schwarz, c_j = ||L_j||_1); the finaliser adds twice that.
  double c2 = 0.0;
  for (int j = 0; j < rank; ++j) {
    double c = 0.0;
    for (int i = 0; i < M_; ++i) c += std::fabs(L_[(size_t)j * M_ + i]);
    c2 += c * c;
  }
  screen_trunc_ = 4.0 * screen_eps_ * std::sqrt(c2);
  // the k-step list of unscreened tiles: every k-step
  std::vector<int32_t> all(Mp_ / 32);
  for (size_t k = 0; k < all.size(); ++k) all[k] = (int32_t)k;
  if (!d_klist_all_.reserve(sizeof(int32_t) * all.size())) throw Error(CT_E_NOMEM, "cannot allocate the k-step list");
  CT_HIP_CHECK(hipMemcpy(d_klist_all_.get(), all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice));
}

bool Mol::factor_route(int mode, int route, bool /*mixed*/) const {
  if (route == CT_ROUTE_DENSITY) return false;
  const bool avail = factor_rank_ >= 0;
  if (route == CT_ROUTE_FACTOR) {
    if (!avail)
      throw Error(CT_E_INVALID,
                  "factor route unavailable: P is not positive semidefinite with rank at most half the basis");
    return true;
  }
  if (!avail) return false;
  // both precisions (the FP64 kernels follow the same column counts):
  // matrix-core products per point, in units of Mp^2 x (the six split
  // products): density route rho 0.525 (triangle), grad 1, lap 1 + 3 x 0.525,
  // Hessian 4; factor route 1, 4, 5 and 4 + 2 x 4 columns of rows / Mp.  The
  // Hessian mode's factor route flags more points for the FP64 re-evaluation
  // (12 % against 3 % on the 5400-function workload), but with spatial
  // screening, which only the factor route has, it still runs 4.7x faster
  // there (3.2x on the 624-function one).
  const double f = (double)factor_rows_ / Mp_;
  switch (mode) {
    case MODE_RHO: return f < 0.525;
    case MODE_GRAD: return 4 * f < 1.0;
    case MODE_LAP: return 5 * f < 2.575;
    default: return 12 * f < 4.0;
  }
}

// Both setters compare the new values with the current ones first: setting
// the same centres or the same P again (an autograd backward after its
// forward, a loop that resets them) costs a copy and a comparison, not the
// factorisation, the bf16 splits and the screening tables.
void Mol::set_coordinates(const double* xyz) {
  CT_REQUIRE(xyz != nullptr, "null coordinates");
  CT_HIP_CHECK(hipStreamSynchronize(stream_));
  std::vector<double> v(3 * (size_t)natom_);
  if (is_device_pointer(xyz))
    CT_HIP_CHECK(hipMemcpy(v.data(), xyz, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
  else
    std::memcpy(v.data(), xyz, sizeof(double) * v.size());
  if (std::memcmp(v.data(), xyz_.data(), sizeof(double) * v.size()) == 0) return;
  for (double c : v) CT_REQUIRE(std::isfinite(c), "coordinates must be finite");
  xyz_.swap(v);
  upload_geometry();
  esp_valid_ = false;
}

void Mol::set_density_matrix(const double* P) {
  CT_REQUIRE(P != nullptr, "null density matrix");
  CT_HIP_CHECK(hipStreamSynchronize(stream_));
  std::vector<double> v((size_t)M_ * M_);
  if (is_device_pointer(P))
    CT_HIP_CHECK(hipMemcpy(v.data(), P, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
  else
    std::memcpy(v.data(), P, sizeof(double) * v.size());
  if (std::memcmp(v.data(), P_.data(), sizeof(double) * v.size()) == 0) return;
  P_.swap(v);
  upload_P();
  esp_valid_ = false;
}

void Mol::prepare_esp() {
  if (esp_valid_) return;
  EspData d = build_esp_pairs(basis_, xyz_, P_, 1e-18);
  const std::vector<float> c32 = esp_coefficients_f32(d.records, d.segs);
  std::vector<double> atoms(4 * (size_t)natom_);
  for (int a = 0; a < natom_; ++a) {
    for (int k = 0; k < 3; ++k) atoms[4 * a + k] = xyz_[3 * a + k];
    atoms[4 * a + 3] = charges_[a];
  }
  if (!d_boys_.get()) {
    const std::vector<double> tab = build_boys_tables();
    if (!d_boys_.reserve(sizeof(double) * tab.size())) throw Error(CT_E_NOMEM, "cannot allocate the Boys table");
    CT_HIP_CHECK(hipMemcpy(d_boys_.get(), tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  }
  if (!d_esp_rec_.reserve(sizeof(double) * std::max<size_t>(1, d.records.size())) ||
      !d_esp_seg_.reserve(sizeof(EspSegment) * std::max<size_t>(1, d.segs.size())) ||
      !d_esp_atoms_.reserve(sizeof(double) * atoms.size()) ||
      !d_esp_rec32_.reserve(sizeof(float) * std::max<size_t>(1, c32.size())))
    throw Error(CT_E_NOMEM, "cannot allocate the ESP pair table");
  if (!c32.empty())
    CT_HIP_CHECK(hipMemcpy(d_esp_rec32_.get(), c32.data(), sizeof(float) * c32.size(), hipMemcpyHostToDevice));
  if (!d.records.empty())
    CT_HIP_CHECK(hipMemcpy(d_esp_rec_.get(), d.records.data(), sizeof(double) * d.records.size(), hipMemcpyHostToDevice));
  if (!d.segs.empty())
    CT_HIP_CHECK(hipMemcpy(d_esp_seg_.get(), d.segs.data(), sizeof(EspSegment) * d.segs.size(), hipMemcpyHostToDevice));
  CT_HIP_CHECK(hipMemcpy(d_esp_atoms_.get(), atoms.data(), sizeof(double) * atoms.size(), hipMemcpyHostToDevice));
  esp_nseg_ = (int)d.segs.size();
  esp_lmax_ = 0;
  for (int64_t& c : esp_pairs_by_L_) c = 0;
  for (const EspSegment& sg : d.segs) {
    esp_lmax_ = std::max(esp_lmax_, (int)sg.L);
    esp_pairs_by_L_[sg.L] += sg.count;
  }
  esp_npairs_ = d.npairs - d.screened;
  esp_valid_ = true;
}

void Mol::backward(const double* points, int64_t n, int layout, const double* w, int precision, double* dP,
                   double* dR) {
  CT_REQUIRE(n >= 0, "negative number of points");
  CT_REQUIRE(layout == CT_LAYOUT_ROW || layout == CT_LAYOUT_COL, "unknown points layout");
  CT_REQUIRE(precision == CT_PREC_FP64 || precision == CT_PREC_MIXED, "unknown precision");
  CT_REQUIRE(n == 0 || (points != nullptr && w != nullptr), "null points or weights");
  CT_HIP_CHECK(hipStreamSynchronize(stream_));
  const int mode = MODE_GRAD;
  const int nder = mode_nder(mode);
  const bool mixed = precision == CT_PREC_MIXED;
  const bool col = layout == CT_LAYOUT_COL;
  // AO block, staged points, staged weights (+ host columns)
  const int64_t per_point = (int64_t)nder * Mp_ * (mixed ? 4 : 8) + 24 + 8 + 24;
  int64_t Nc = n > 0 ? chunk_points(per_point, n) : kPointAlign;
  while (!d_ws_.reserve((size_t)Nc * per_point)) {
    if (Nc <= kPointAlign) throw Error(CT_E_NOMEM, "cannot allocate the adjoint workspace");
    Nc = std::max<int64_t>(kPointAlign, Nc / 2 / kPointAlign * kPointAlign);
  }
  int nslice_max = gram_slices(Nc);
  int64_t slice_len = 0;
  if (mixed) gram_mixed_plan(Nc, Mp_, &nslice_max, &slice_len);
  const size_t gsize = (size_t)Mp_ * Mp_;
  const size_t adj_bytes = sizeof(double) * (gsize + (size_t)Mp_ * 3 + (size_t)nslice_max * gsize);
  if (!d_adj_.reserve(adj_bytes)) throw Error(CT_E_NOMEM, "cannot allocate the adjoint accumulators");
  double* G = static_cast<double*>(d_adj_.get());
  double* T = G + gsize;
  double* part = T + (size_t)Mp_ * 3;
  launch_fill_zero(reinterpret_cast<unsigned long long*>(G), (int64_t)(gsize + (size_t)Mp_ * 3), stream_);
  const int32_t* tiles = nullptr;
  if (mixed) {
    // upper-triangle 128 x 128 tiles of the Gram matrix, (I, J) with I <= J
    // (fixed by the basis: uploaded on the first mixed call)
    if (!tiles_ready_) {
      const int nb = Mp_ / kRowBlock;
      std::vector<int32_t> tl;
      for (int I = 0; I < nb; ++I)
        for (int J = I; J < nb; ++J) {
          tl.push_back(I);
          tl.push_back(J);
        }
      if (!d_tiles_.reserve(sizeof(int32_t) * tl.size())) throw Error(CT_E_NOMEM, "cannot allocate the tile list");
      CT_HIP_CHECK(hipMemcpy(d_tiles_.get(), tl.data(), sizeof(int32_t) * tl.size(), hipMemcpyHostToDevice));
      tiles_ready_ = true;
    }
    tiles = static_cast<const int32_t*>(d_tiles_.get());
  }
  const bool dP_dev = dP != nullptr && is_device_pointer(dP);
  const bool dR_dev = dR != nullptr && is_device_pointer(dR);
  if (dR_dev && !func_atom_ready_) {
    std::vector<int32_t> fa((size_t)M_, 0);
    for (int s = 0; s < nshell_; ++s)
      for (int f = 0; f < shell_nfunc(basis_.shell_l[s]); ++f) fa[(size_t)meta_[s].first + f] = basis_.shell_atom[s];
    if (!d_func_atom_.reserve(sizeof(int32_t) * fa.size()))
      throw Error(CT_E_NOMEM, "cannot allocate the function-to-atom map");
    CT_HIP_CHECK(hipMemcpy(d_func_atom_.get(), fa.data(), sizeof(int32_t) * fa.size(), hipMemcpyHostToDevice));
    func_atom_ready_ = true;
  }
  char* ws = static_cast<char*>(d_ws_.get());
  double* pts_stage = reinterpret_cast<double*>(ws + (size_t)nder * Mp_ * Nc * (mixed ? 4 : 8));
  double* w_stage = pts_stage + 3 * (size_t)Nc;
  double* col_stage = w_stage + (size_t)Nc;
  const bool pd = is_device_pointer(points), wd = is_device_pointer(w);
  const auto* meta = static_cast<const ShellMeta*>(d_meta_.get());
  const auto* geom = static_cast<const ShellGeom*>(d_geom_.get());
  const auto* prim = static_cast<const PrimData*>(d_prim_.get());
  for (int64_t start = 0; start < n; start += Nc) {
    const int64_t nc = std::min(Nc, n - start);
    const int64_t ld = round_up(nc, kPointAlign);
    const double* pts = points + 3 * start;
    const double* wc = w + start;
    if (col) {
      const double* cols = points + start;
      int64_t stride = n;
      if (!pd) {
        CT_HIP_CHECK(hipMemcpy2DAsync(col_stage, sizeof(double) * nc, points + start, sizeof(double) * n,
                                      sizeof(double) * nc, 3, hipMemcpyHostToDevice, stream_));
        cols = col_stage;
        stride = nc;
      }
      launch_gather_points(cols, nullptr, nc, pts_stage, stream_, stride);
      pts = pts_stage;
    } else if (!pd) {
      CT_HIP_CHECK(hipMemcpyAsync(pts_stage, pts, sizeof(double) * 3 * nc, hipMemcpyHostToDevice, stream_));
      pts = pts_stage;
    }
    if (!wd) {
      CT_HIP_CHECK(hipMemcpyAsync(w_stage, wc, sizeof(double) * nc, hipMemcpyHostToDevice, stream_));
      wc = w_stage;
    }
    timer_.begin(KernelTimer::BACKWARD, stream_);
    if (!mixed) {
      double* ao = reinterpret_cast<double*>(ws);
      // rows up to the 64-row blocks of the FP64 Gram and adjoint (zeros past M)
      launch_ao_f64(mode, lmax_, meta, geom, prim, nshell_, M_, Mp_, pts, nc, ld, ao, 700.0, stream_, nullptr, 0, 64,
                    std::min(Mp_, (M_ + 63) / 64 * 64));
      launch_gram_f64(ao, Mp_, M_, ld, wc, nc, part, stream_);
      launch_gram_reduce(part, gram_slices(nc), (int64_t)gsize, G, stream_);
      launch_contract_adjoint_f64(static_cast<const double*>(d_P_.get()), Mp_, M_, ao, ld, wc, nc, T, stream_);
    } else {
      float* ao = reinterpret_cast<float*>(ws);
      // rows up to the last k-step with a real function (zeros past M); the
      // Gram stages zeros for the rest of its last row block itself
      launch_ao_f32(mode, lmax_, meta, geom, prim, nshell_, M_, Mp_, pts, nc, ld, ao, 50.0, nullptr, 0, 64, stream_,
                    std::min(Mp_, (M_ + 31) / 32 * 32));
      int nslice = 0;
      int64_t len = 0;
      gram_mixed_plan(nc, Mp_, &nslice, &len, nslice_max);
      launch_gram_mixed(ao, Mp_, M_, ld, wc, nc, tiles, nslice, len, part, stream_);
      launch_gram_reduce(part, nslice, (int64_t)gsize, G, stream_);
      launch_contract_adjoint_mixed(static_cast<const uint16_t*>(d_Pb_.get()), Mp_, M_, ao, ld, wc, nc, T, stream_);
    }
    timer_.end(stream_);
  }
  // device outputs are finished on the device (dP: the real M x M block of
  // G; dR: the per-atom sums of T, in the host loop's order); host outputs
  // from host copies of G and T
  if (dP_dev)
    CT_HIP_CHECK(hipMemcpy2DAsync(dP, sizeof(double) * M_, G, sizeof(double) * Mp_, sizeof(double) * M_, M_,
                                  hipMemcpyDeviceToDevice, stream_));
  if (dR_dev)
    launch_atom_gradient(T, static_cast<const int32_t*>(d_func_atom_.get()), M_, natom_, dR, stream_);
  std::vector<double> Gh, Th;
  if (dP && !dP_dev) {
    Gh.resize(gsize);
    CT_HIP_CHECK(hipMemcpyAsync(Gh.data(), G, sizeof(double) * gsize, hipMemcpyDeviceToHost, stream_));
  }
  if (dR && !dR_dev) {
    Th.resize((size_t)Mp_ * 3);
    CT_HIP_CHECK(hipMemcpyAsync(Th.data(), T, sizeof(double) * Th.size(), hipMemcpyDeviceToHost, stream_));
  }
  CT_HIP_CHECK(hipStreamSynchronize(stream_));
  timer_.collect();
  if (dP && !dP_dev)
    for (int i = 0; i < M_; ++i) std::memcpy(dP + (size_t)i * M_, Gh.data() + (size_t)i * Mp_, sizeof(double) * M_);
  if (dR && !dR_dev) {
    std::fill(dR, dR + 3 * (size_t)natom_, 0.0);
    for (int s = 0; s < nshell_; ++s) {
      const int a = basis_.shell_atom[s];
      for (int f = 0; f < shell_nfunc(basis_.shell_l[s]); ++f) {
        const int i = meta_[s].first + f;
        for (int k = 0; k < 3; ++k) dR[3 * a + k] -= 2.0 * Th[(size_t)i * 3 + k];
      }
    }
  }
}

void Mol::esp(const double* points, int64_t n, int layout, int precision, double* out) {
  CT_REQUIRE(n >= 0, "negative number of points");
  CT_REQUIRE(precision == CT_PREC_FP64 || precision == CT_PREC_MIXED, "unknown precision");
  CT_REQUIRE(layout == CT_LAYOUT_ROW || layout == CT_LAYOUT_COL, "unknown points layout");
  if (n == 0) return;
  CT_REQUIRE(points != nullptr && out != nullptr, "null pointer");
  CT_HIP_CHECK(hipStreamSynchronize(stream_));
  prepare_esp();
  const bool pd = is_device_pointer(points), od = is_device_pointer(out);
  const bool col = layout == CT_LAYOUT_COL;
  const double* dp = points;
  double* dout = out;
  if (!pd || !od || col) {
    // [0, 3n) row-major points, [3n, 4n) output, [4n, 7n) host columns
    const size_t nd = (size_t)n * (col && !pd ? 7 : 4);
    if (!d_esp_io_.reserve(sizeof(double) * nd)) throw Error(CT_E_NOMEM, "cannot allocate ESP staging");
    double* io = static_cast<double*>(d_esp_io_.get());
    if (col) {
      const double* cols = points;
      if (!pd) {
        CT_HIP_CHECK(hipMemcpyAsync(io + 4 * n, points, sizeof(double) * 3 * n, hipMemcpyHostToDevice, stream_));
        cols = io + 4 * n;
      }
      launch_gather_points(cols, nullptr, n, io, stream_, n);
      dp = io;
    } else if (!pd) {
      CT_HIP_CHECK(hipMemcpyAsync(io, points, sizeof(double) * 3 * n, hipMemcpyHostToDevice, stream_));
      dp = io;
    }
    if (!od) dout = io + 3 * n;
  }
  // 512-thread workgroups (the Boys table, 45 or 69 KB of LDS, allows three
  // or two per CU) and the pair list split over enough slices for about
  // 12 288 workgroups, at least 256 pairs per slice, at most 32 slices:
  // esp30 (2e5 points, 512-point blocks) 1.30e7 points/s unsplit at 256
  // threads, 2.27e7 with 8 slices, 2.64e7 with 32
  const int block = 512;
  const int64_t nbx = (n + block - 1) / block;
  const int nslice = (int)std::max<int64_t>(
      1, std::min<int64_t>({32, (12288 + nbx - 1) / nbx, std::max<int64_t>(1, esp_npairs_ / 256)}));
  double* part = nullptr;
  if (nslice > 1) {
    if (!d_esp_part_.reserve(sizeof(double) * (size_t)nslice * n)) throw Error(CT_E_NOMEM, "cannot allocate ESP partials");
    part = static_cast<double*>(d_esp_part_.get());
  }
  timer_.begin(KernelTimer::ESP, stream_);
  const float* rec32 = precision == CT_PREC_MIXED ? static_cast<const float*>(d_esp_rec32_.get()) : nullptr;
  launch_esp(dp, n, static_cast<const double*>(d_esp_rec_.get()), rec32,
             static_cast<const EspSegment*>(d_esp_seg_.get()), esp_nseg_, esp_npairs_, esp_lmax_,
             static_cast<const double*>(d_esp_atoms_.get()), natom_, static_cast<const double*>(d_boys_.get()), dout,
             part, nslice, block, stream_);
  timer_.end(stream_);
  if (!od) CT_HIP_CHECK(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, stream_));
  CT_HIP_CHECK(hipStreamSynchronize(stream_));
  timer_.collect();
}

void Mol::esp_pairs_by_L(int64_t* out) {
  CT_HIP_CHECK(hipStreamSynchronize(stream_));
  prepare_esp();
  for (int L = 0; L <= kEspLmax; ++L) out[L] = esp_pairs_by_L_[L];
}

void Mol::synchronize() { CT_HIP_CHECK(hipStreamSynchronize(stream_)); }

// Chunk sizes of the host-array pipeline for n points, at most C each
// (C a multiple of kPointAlign): sizes growing fourfold from C/16 at the
// start and shrinking to C/16 at the end, so that the first upload and the
// last download, which nothing overlaps, are short; the middle in near-equal
// chunks of at most C.  Every size but the last is a multiple of kPointAlign
// (the chunks start at aligned points).
static std::vector<int64_t> pipeline_chunks(int64_t n, int64_t C) {
  std::vector<int64_t> head, mid, tail;
  const int64_t N = round_up(n, kPointAlign);
  int64_t rem = N;  // a multiple of kPointAlign throughout
  for (int64_t c = std::max<int64_t>(kPointAlign, C / 16 / kPointAlign * kPointAlign);; c = std::min(C, 4 * c)) {
    if (rem <= 2 * c) {
      const int64_t a = round_up(rem / 2, kPointAlign);
      head.push_back(a);
      if (rem > a) tail.push_back(rem - a);
      break;
    }
    head.push_back(c);
    tail.push_back(c);
    rem -= 2 * c;
    if (c == C) {
      const int64_t k = (rem + C - 1) / C;
      const int64_t each = round_up((rem + k - 1) / k, kPointAlign);
      for (; rem > 0; rem -= std::min(each, rem)) mid.push_back(std::min(each, rem));
      break;
    }
  }
  head.insert(head.end(), mid.begin(), mid.end());
  head.insert(head.end(), tail.rbegin(), tail.rend());
  head.back() -= N - n;  // the last chunk >= kPointAlign > N - n
  return head;
}

int64_t Mol::chunk_points(int64_t per_point, int64_t n) const {
  int64_t nc = ws_limit_ / per_point / kPointAlign * kPointAlign;
  nc = std::max<int64_t>(nc, kPointAlign);
  return std::min<int64_t>(nc, round_up(n, kPointAlign));
}

size_t Mol::reeval_bytes(int mode, int64_t count, bool factor) const {
  // AO arrays, dots, gathered points, the split contraction's partials and
  // the factor route's psi (lap and Hessian modes) (an upper bound for every
  // slice of at most count points); the screened
  // factor route: the active k-step counters, lists, counts and masks of the
  // slice's point tiles
  const int nder = mode_nder(mode), ndot = mode_ndot(mode);
  const int64_t ld = round_up(count, kPointAlign);
  const int blocks = factor ? factor_rows_ / 64 : Mp_ / 64;
  const size_t psi = factor && (mode == MODE_LAP || mode == MODE_HESS) ? (size_t)factor_rows_ * 8 : 0;
  size_t bytes = (size_t)ld * ((size_t)nder * Mp_ * 8 + (size_t)ndot * 8 + 24 + (size_t)blocks * ndot * 8 + psi);
  if (factor && screen_eps_ > 0.0) {
    const int nk = Mp_ / 32, nwords = (nk + 31) / 32;
    bytes += 128 * kScreenCounters + (size_t)(ld / factor_tile_points(mode)) * (4 * nk + 4 + 4 * nwords);
  }
  return bytes;
}

// FP64 evaluation of `count` points (index: their positions in pts), written
// into fin's outputs through the index map, in slices that fit the FP64
// workspace: the caller's region ws when given, else d_ws64_, grown to a
// slice sized like an FP64 chunk.
void Mol::reevaluate(int mode, const double* pts, int64_t col_stride, int64_t count, const int32_t* index,
                     const FinalizeArgs& fin, bool factor, char* ws, size_t ws_bytes, hipStream_t rs) {
  if (count <= 0) return;
  if (rs == nullptr) rs = stream_;
  const int nder = mode_nder(mode), ndot = mode_ndot(mode);
  int64_t S = round_up(count, kPointAlign);
  if (ws) {
    while (S > kPointAlign && reeval_bytes(mode, S, factor) > ws_bytes)
      S = std::max<int64_t>(kPointAlign, S / 2 / kPointAlign * kPointAlign);
    if (reeval_bytes(mode, S, factor) > ws_bytes)
      throw Error(CT_E_NOMEM, "the FP64 re-evaluation needs " + std::to_string(reeval_bytes(mode, S, factor)) +
                                  " bytes for one slice of 256 points; the workspace holds " + std::to_string(ws_bytes));
  } else {
    S = chunk_points((int64_t)(reeval_bytes(mode, kPointAlign, factor) / kPointAlign), count);
    while (!d_ws64_.reserve(reeval_bytes(mode, S, factor))) {
      if (S <= kPointAlign) throw Error(CT_E_NOMEM, "cannot allocate the FP64 re-evaluation workspace");
      S = std::max<int64_t>(kPointAlign, S / 2 / kPointAlign * kPointAlign);
    }
    ws = static_cast<char*>(d_ws64_.get());
  }
  for (int64_t s0 = 0; s0 < count; s0 += S) {
    const int64_t c = std::min(S, count - s0);
    const int32_t* idx = index + s0;
    const int64_t ld = round_up(c, kPointAlign);
    // few points: one wave per (16 or 64 points, 64-row block of P, or
    // 64-column pass of the factor) so the GPU fills; the factor route with
    // three or more passes at any size (the passes of a point group share
    // its AO columns in one XCD's L2).  Not for the density
    // route's gradient, lap and Hessian modes: their point-set kernel has
    // four waves per point set already, and keeping it for every slice gives
    // a point the same bits whatever slice it falls in.
    const bool split = ld < 65536 || (factor && factor_rows_ / 64 >= 3);
    double* ao = reinterpret_cast<double*>(ws);
    double* dots = ao + (size_t)nder * Mp_ * ld;
    double* gp = dots + (size_t)ndot * ld;
    double* part = gp + 3 * (size_t)ld;
    double* psi = part + (size_t)(factor ? factor_rows_ / 64 : Mp_ / 64) * ndot * ld;
    const size_t psi_len = factor && (mode == MODE_LAP || mode == MODE_HESS) ? (size_t)factor_rows_ * ld : 0;
    launch_gather_points(pts, idx, c, gp, rs, col_stride);
    // the factor route in the order of its factor (the screened shell
    // order), screened as the FP64 path screens it: per tile of the slice's
    // points (in index order), the k-steps of its shells within the FP64
    // cutoff radii
    const bool scr = factor && screen_eps_ > 0.0;
    const int tp = scr ? factor_tile_points(mode) : 64;
    const int nk = Mp_ / 32, nwords = (nk + 31) / 32;
    uint32_t* kmask = nullptr;
    if (scr) {
      char* sp = reinterpret_cast<char*>(psi + psi_len);
      auto* counters = reinterpret_cast<unsigned long long*>(sp);  // (unused here)
      const int64_t ntile = ld / tp;
      int32_t* klist = reinterpret_cast<int32_t*>(sp + 128 * kScreenCounters);
      int32_t* kcount = klist + ntile * nk;
      kmask = reinterpret_cast<uint32_t*>(kcount + ntile);
      launch_tile_activity(gp, c, tp, ntile, static_cast<const int32_t*>(d_span_off64_.get()),
                           static_cast<const ScreenSpan*>(d_spans64_.get()), nk, klist, kcount, kmask, nwords,
                           counters, rs);
    }
    launch_ao_f64(mode, lmax_, static_cast<const ShellMeta*>(factor ? d_smeta_.get() : d_meta_.get()),
                  static_cast<const ShellGeom*>(factor ? d_sgeom_.get() : d_geom_.get()),
                  static_cast<const PrimData*>(d_prim_.get()), nshell_, M_, Mp_, gp, c, ld, ao, 700.0, rs, kmask,
                  nwords, tp);
    if (factor && split)
      launch_contract_f64_factor_split(mode, static_cast<const double*>(d_L64s_.get()), factor_rows_, Mp_, ao, ld,
                                       part, dots, rs, kmask, nwords, tp, psi_len ? psi : nullptr);
    else if (factor)
      launch_contract_f64_factor(mode, static_cast<const double*>(d_L64s_.get()), factor_rows_, Mp_, ao, ld, dots,
                                 rs, kmask, nwords, tp, psi_len ? psi : nullptr);
    else if (split && mode == MODE_RHO)
      launch_contract_f64_split(mode, static_cast<const double*>(d_P_.get()), Mp_, ao, ld, part, dots, rs);
    else
      launch_contract_f64(mode, static_cast<const double*>(d_P_.get()), Mp_, ao, ld, dots, rs);
    FinalizeArgs f = fin;
    f.dots = dots;
    f.index = idx;
    f.gather = nullptr;  // the slice's dots are in index order; outputs go to index[p]
    f.nrt = 0;
    f.ld = ld;
    f.n = c;
    launch_finalize(f, rs);
  }
}

void Mol::eval(const double* points, int64_t n, int layout, uint32_t q, int precision, const ct_eval_params& par,
               const ct_outputs& out, ct_eval_report* rep, const CubicGrid* cube, const Workspace* fixed) {
  CT_REQUIRE(n >= 0, "negative number of points");
  CT_REQUIRE(q != 0, "no quantity requested");
  CT_REQUIRE(precision == CT_PREC_FP64 || precision == CT_PREC_MIXED, "unknown precision");
  CT_REQUIRE(layout == CT_LAYOUT_ROW || layout == CT_LAYOUT_COL, "unknown points layout");
  CT_REQUIRE(par.out_layout == CT_LAYOUT_ROW || par.out_layout == CT_LAYOUT_COL, "unknown output layout");
  for (const OutSlot& s : kSlots)
    if ((q & s.bit) && out.*(s.field) == nullptr) throw Error(CT_E_INVALID, "output array missing for a requested quantity");
  if (rep) *rep = ct_eval_report{n, 0, 0, 0.0, -1, 1.0};
  if (n == 0) return;
  CT_REQUIRE(points != nullptr || cube != nullptr, "null points");
  const NoAllocScope no_alloc(fixed != nullptr);

  const bool mixed = precision == CT_PREC_MIXED;
  const int mode = select_mode(q);
  const int nder = mode_nder(mode), ndot = mode_ndot(mode);
  const bool use_factor = factor_route(mode, par.route, mixed);
  const int nrt = (use_factor ? factor_rows_ : Mp_) / kRowBlock;  // row blocks of the mixed partials
  // screened layout (factor route): points in Morton order, per-tile
  // lists of the active k-steps
  const int nk = Mp_ / 32, nwords = (nk + 31) / 32;
  const bool screened = use_factor && screen_eps_ > 0.0;
  const int tp = screened ? factor_tile_points(mode) : 64;
  const bool pts_dev = cube == nullptr && is_device_pointer(points);
  // column-major points: (3, n), x, y, z of point p at points[p], points[n + p],
  // points[2n + p]; each chunk is gathered into row-major staging
  const bool col_in = cube == nullptr && layout == CT_LAYOUT_COL;
  const bool col_out = par.out_layout == CT_LAYOUT_COL;
  constexpr size_t NS = sizeof(kSlots) / sizeof(kSlots[0]);
  bool out_dev[NS];
  int stage_width = 0;
  for (size_t k = 0; k < NS; ++k) {
    out_dev[k] = (q & kSlots[k].bit) ? is_device_pointer(out.*(kSlots[k].field)) : true;
    if ((q & kSlots[k].bit) && !out_dev[k]) stage_width += kSlots[k].width;
  }

  // With device-resident points and outputs, flagged points of all chunks are
  // collected in one list and re-evaluated in a single FP64 pass at the end
  // (one large launch instead of one small launch per chunk).
  bool all_dev = pts_dev && n < (int64_t(1) << 31);  // (grid points are not materialised: per-chunk re-evaluation)
  for (size_t k = 0; k < NS; ++k) all_dev = all_dev && out_dev[k];
  // a per-point target density (mixed path): chunk slices of the caller's
  // device array, or staged from host memory
  const bool has_target = mixed && par.rho_target != nullptr;
  const bool tgt_dev = has_target && is_device_pointer(par.rho_target);
  if (fixed)
    CT_REQUIRE(cube == nullptr && all_dev && (!has_target || tgt_dev),
               "allocation-free evaluation: points, outputs and targets must be device memory, fewer than 2^31 "
               "points");
  // (the call-wide flag list is an allocation: the allocation-free path
  // re-evaluates chunk by chunk)
  // Host arrays (points, outputs or targets in host memory): the chunks run as
  // a pipeline through kPipeSlots slots of pinned and device staging.  Per
  // chunk i, in slot i % kPipeSlots: the caller's inputs are copied into the
  // slot's pinned region (host threads) and uploaded on the upload stream;
  // the kernels run on the handle's stream; once chunk i + 1 is queued behind
  // them, chunk i's flag count is read, its flagged points re-evaluated and
  // its outputs downloaded on the download stream; the host copies them into
  // the caller's arrays when the slot comes round again.  The uploads and
  // downloads of neighbouring chunks overlap the kernels.
  const bool host_in = cube == nullptr && !pts_dev;
  const bool host_tgt = has_target && !tgt_dev;
  const bool piped = host_in || host_tgt || stage_width > 0;
  const int nslot = piped ? kPipeSlots : 1;
  // (the call-wide flag list is an allocation: the allocation-free path
  // re-evaluates chunk by chunk)
  const bool batch = mixed && all_dev && !fixed && !piped;
  const int64_t ao_bytes_pt = (int64_t)nder * Mp_ * (mixed ? 4 : 8);
  // per slot: points in row layout (24 B), the uploaded inputs (points as
  // given, 24 B; targets, 8 B), staged outputs, flag index (4 B)
  const int64_t in_pt = (host_in ? 24 : 0) + (host_tgt ? 8 : 0);
  const int64_t slot_pt = 24 + in_pt + stage_width * 8 + 4;
  // AO block, dots (mixed: per row block, with the bound), the slots: the
  // pieces carved below, in that order
  // the factor route's psi of a mode's first launch (factor_rows_ doubles per
  // point), read by its other launches: mixed Hessian mode, FP64 lap and
  // Hessian modes
  const bool psi_used = use_factor && (mode == MODE_HESS || (!mixed && mode == MODE_LAP));
  const int64_t psi_pt = psi_used ? (int64_t)factor_rows_ * 8 : 0;
  // the FP64 factor route with three or more 64-row passes of L^T runs one
  // workgroup per (point group, pass), the passes of a group back to back on
  // one XCD, sharing its AO columns in that L2 (one pass per wave streamed
  // the AOs from HBM once per pass); partial dots per pass, summed in order
  const bool pass_split = use_factor && !mixed && factor_rows_ / 64 >= 2;
  const int64_t part_pt = pass_split ? (int64_t)(factor_rows_ / 64) * ndot * 8 : 0;
  const int64_t per_point =
      (mixed ? ao_bytes_pt + (int64_t)nrt * (ndot + 1) * 8 : ao_bytes_pt + (int64_t)ndot * 8) + psi_pt + part_pt +
      nslot * slot_pt;
  // screened: keys (2 x 4 B), indices (2 x 4 B), sorted points (24 B) per
  // point; k-step list (2 nk B), count and mask (4 + 4 nwords B) per tile
  const int64_t per_point_s = screened ? 40 + (4 * nk + 4 + 4 * nwords + tp - 1) / tp : 0;
  // flag counters (one per slot, 16 bytes apart), max estimate (bytes 8-15);
  // from byte 256 the active k-step counters (one per 128-byte line)
  const size_t head = 256 + 128 * kScreenCounters;
  size_t sort_bytes = 0;
  int64_t Nc = 0;
  char* ws = nullptr;
  size_t ws_size = 0;
  char* ws64 = nullptr;  // caller's region: the FP64 re-evaluation slice
  size_t ws64_bytes = 0;
  if (!fixed) {
    Nc = chunk_points(per_point + per_point_s, n);
    if (piped) Nc = std::min<int64_t>(Nc, kHostChunk);
    for (;;) {
      sort_bytes = screened ? (screen_sort_workspace(Nc) + 255) / 256 * 256 : 0;
      if (d_ws_.reserve(head + (size_t)Nc * (per_point + per_point_s) + sort_bytes + 4096)) break;
      if (Nc <= kPointAlign) throw Error(CT_E_NOMEM, "cannot allocate the evaluation workspace");
      Nc = std::max<int64_t>(kPointAlign, Nc / 2 / kPointAlign * kPointAlign);
    }
    ws = static_cast<char*>(d_ws_.get());
    ws_size = d_ws_.size();
  } else {
    // the re-evaluation slice (mixed) at the end of the region, at most a
    // quarter of it (4096 points or fewer), the chunk workspace before it
    if (mixed) {
      int64_t S = 4096;
      while (S > kPointAlign && reeval_bytes(mode, S, use_factor) > fixed->bytes / 4) S /= 2;
      ws64_bytes = (reeval_bytes(mode, S, use_factor) + 255) / 256 * 256;
    }
    const size_t main = fixed->bytes > ws64_bytes ? fixed->bytes - ws64_bytes : 0;
    const size_t base_cost = head + 4096;
    Nc = main > base_cost ? (int64_t)((main - base_cost) / (size_t)(per_point + per_point_s)) : 0;
    Nc = std::min<int64_t>(Nc / kPointAlign * kPointAlign, round_up(n, kPointAlign));
    for (;;) {
      if (Nc < kPointAlign || ws64_bytes > fixed->bytes)
        throw Error(CT_E_NOMEM, "allocation-free evaluation: the scratch (" + std::to_string(fixed->bytes) +
                                    " bytes) cannot hold one 256-point chunk of this call (" +
                                    std::to_string(per_point + per_point_s) + " bytes per point, " +
                                    std::to_string(ws64_bytes) + " for the FP64 re-evaluation); use ct_eval");
      sort_bytes = screened ? (screen_sort_workspace(Nc) + 255) / 256 * 256 : 0;
      if (base_cost + (size_t)Nc * (per_point + per_point_s) + sort_bytes <= main) break;
      Nc = (Nc - Nc / 8) / kPointAlign * kPointAlign;
    }
    ws = fixed->base;
    ws_size = main;
    ws64 = fixed->base + main;
  }
  unsigned long long* d_maxest = reinterpret_cast<unsigned long long*>(ws + 8);
  unsigned long long* d_active = reinterpret_cast<unsigned long long*>(ws + 256);
  char* cur = ws + head;
  void* ao = cur;
  cur += (size_t)Nc * ao_bytes_pt;
  double* dots = reinterpret_cast<double*>(cur);
  cur += (size_t)Nc * (mixed ? nrt * (ndot + 1) : ndot) * 8;
  double* psi = psi_pt ? reinterpret_cast<double*>(cur) : nullptr;
  cur += (size_t)Nc * psi_pt;
  double* part = part_pt ? reinterpret_cast<double*>(cur) : nullptr;
  cur += (size_t)Nc * part_pt;

  // a pipeline slot: its staging and the chunk it holds
  struct Slot {
    double* pts_stage;  // device: the chunk's points, row layout
    char* in_dev;       // device: uploaded points as given ([0, 24 Nc)), targets (after them)
    double* out_stage;  // device: outputs bound for host arrays, each quantity w x nc
    int32_t* flag_idx;  // device: the chunk's flagged points
    int32_t* count;     // device: their number
    char* in_host;      // pinned: inputs to upload
    char* out_host;     // pinned: downloaded outputs
    int64_t* cnt_host;  // pinned: the count read back
    hipEvent_t* ev;     // PipeEvent events
    bool used[NPIPEEV];  // which events this call has recorded
    // the chunk in the slot
    int64_t start, nc;
    const double* pts;
    FinalizeArgs fa;
    bool out_pending;  // downloaded (or downloading) outputs not yet copied to the caller
  };
  Slot slots[kPipeSlots];
  // readbacks of the counters: pinned, or (allocation-free) pageable memory
  // on this stack, read after the stream synchronises
  int64_t h_local[8 + 16 * kScreenCounters];
  int64_t* h_small = h_local;
  if (!fixed) {
    h_small_.reserve(sizeof(h_local));
    h_small = static_cast<int64_t*>(h_small_.get());
  }
  const size_t hslot = (size_t)Nc * (size_t)(in_pt + stage_width * 8);
  if (piped) {
    ensure_pipeline();
    // (idle unless an earlier call ended in an error with copies in flight)
    CT_HIP_CHECK(hipStreamSynchronize(h2d_stream_));
    CT_HIP_CHECK(hipStreamSynchronize(d2h_stream_));
    h_stage_.reserve(hslot * nslot);
  }
  for (int s = 0; s < nslot; ++s) {
    Slot& sl = slots[s];
    std::memset(&sl, 0, sizeof(sl));
    sl.pts_stage = reinterpret_cast<double*>(cur);
    cur += (size_t)Nc * 24;
    sl.in_dev = cur;
    cur += (size_t)Nc * in_pt;
    sl.out_stage = reinterpret_cast<double*>(cur);
    cur += (size_t)Nc * stage_width * 8;
    sl.flag_idx = reinterpret_cast<int32_t*>(cur);
    cur += (size_t)Nc * 4;
    sl.count = reinterpret_cast<int32_t*>(ws + 16 * s);
    sl.cnt_host = h_small + 2 + s;
    if (piped) {
      sl.in_host = static_cast<char*>(h_stage_.get()) + hslot * s;
      sl.out_host = sl.in_host + (size_t)Nc * in_pt;
      sl.ev = pipe_ev_[s];
    }
  }
  // screened layout buffers (256-byte aligned pieces)
  auto take = [&](size_t bytes) {
    char* p = cur;
    cur += (bytes + 255) / 256 * 256;
    return p;
  };
  uint32_t *keys_in = nullptr, *keys_out = nullptr, *kmask = nullptr;
  int32_t *idx_in = nullptr, *perm = nullptr, *kcount = nullptr;
  int32_t* klist = nullptr;
  double* spts = nullptr;
  void* sort_tmp = nullptr;
  if (screened) {
    const int64_t ntile_max = Nc / tp;
    spts = reinterpret_cast<double*>(take((size_t)Nc * 24));
    keys_in = reinterpret_cast<uint32_t*>(take((size_t)Nc * 4));
    keys_out = reinterpret_cast<uint32_t*>(take((size_t)Nc * 4));
    idx_in = reinterpret_cast<int32_t*>(take((size_t)Nc * 4));
    perm = reinterpret_cast<int32_t*>(take((size_t)Nc * 4));
    klist = reinterpret_cast<int32_t*>(take((size_t)ntile_max * nk * 4));
    kcount = reinterpret_cast<int32_t*>(take((size_t)ntile_max * 4));
    kmask = reinterpret_cast<uint32_t*>(take((size_t)ntile_max * nwords * 4));
    sort_tmp = take(sort_bytes);
  }
  CT_REQUIRE(cur <= ws + ws_size, "internal error: evaluation workspace overflow");
  int64_t valid_tiles = 0;

  const auto* meta = static_cast<const ShellMeta*>(d_meta_.get());
  const auto* geom = static_cast<const ShellGeom*>(d_geom_.get());
  const auto* prim = static_cast<const PrimData*>(d_prim_.get());
  int32_t* list = slots[0].flag_idx;
  if (batch) {
    if (!d_flags_.reserve(sizeof(int32_t) * (size_t)n)) throw Error(CT_E_NOMEM, "cannot allocate the flag list");
    list = static_cast<int32_t*>(d_flags_.get());
  }
  // the flag counters (bytes 0-3, 16-19, 32-35), the largest estimate (8-15)
  // and the active k-step counters start at zero
  if (mixed || screened) launch_fill_zero(reinterpret_cast<unsigned long long*>(ws), (int64_t)(head / 8), stream_);

  int64_t chunks = 0, reevaluated = 0;
  // Queues chunk [start, start + nc) in slot sl: upload, kernels, finaliser
  // and (mixed, per chunk) the read-back of its flag count.
  auto enqueue_chunk = [&](Slot& sl, int64_t start, int64_t nc) {
    const int64_t ld = round_up(nc, kPointAlign);
    sl.start = start;
    sl.nc = nc;
    if (piped) {
      // the slot's device buffers are free once its previous chunk is final
      // (re-evaluated) and downloaded (the host has copied that chunk out
      // before this call)
      if (sl.used[EV_READY]) {
        CT_HIP_CHECK(hipStreamWaitEvent(h2d_stream_, sl.ev[EV_READY], 0));
        CT_HIP_CHECK(hipStreamWaitEvent(stream_, sl.ev[EV_READY], 0));
      }
      if (sl.used[EV_D2H]) CT_HIP_CHECK(hipStreamWaitEvent(stream_, sl.ev[EV_D2H], 0));
      if (in_pt > 0) {
        // the pinned region is free once its previous upload is done
        if (sl.used[EV_H2D]) CT_HIP_CHECK(hipEventSynchronize(sl.ev[EV_H2D]));
        HostCopy jobs[4];
        int nj = 0;
        if (host_in) {
          if (col_in)
            for (int c = 0; c < 3; ++c)
              jobs[nj++] = HostCopy{sl.in_host + (size_t)c * nc * 8, points + (size_t)c * n + start, (size_t)nc * 8};
          else
            jobs[nj++] = HostCopy{sl.in_host, points + 3 * start, (size_t)nc * 24};
        }
        const size_t tgt_off = host_in ? (size_t)Nc * 24 : 0;
        if (host_tgt) jobs[nj++] = HostCopy{sl.in_host + tgt_off, par.rho_target + start, (size_t)nc * 8};
        host_copy(jobs, nj);
        if (host_in)
          CT_HIP_CHECK(hipMemcpyAsync(sl.in_dev, sl.in_host, (size_t)nc * 24, hipMemcpyHostToDevice, h2d_stream_));
        if (host_tgt)
          CT_HIP_CHECK(hipMemcpyAsync(sl.in_dev + tgt_off, sl.in_host + tgt_off, (size_t)nc * 8,
                                      hipMemcpyHostToDevice, h2d_stream_));
        CT_HIP_CHECK(hipEventRecord(sl.ev[EV_H2D], h2d_stream_));
        sl.used[EV_H2D] = true;
        CT_HIP_CHECK(hipStreamWaitEvent(stream_, sl.ev[EV_H2D], 0));
      }
    }
    const double* pts = nullptr;
    if (cube) {
      launch_cubic_points(*cube, start, nc, sl.pts_stage, stream_);
      pts = sl.pts_stage;
    } else if (col_in) {
      // column-major points: (3, n), x, y, z of point p at points[p],
      // points[n + p], points[2n + p]; gathered into row layout (uploaded
      // columns: stride nc)
      launch_gather_points(host_in ? reinterpret_cast<const double*>(sl.in_dev) : points + start, nullptr, nc,
                           sl.pts_stage, stream_, host_in ? nc : n);
      pts = sl.pts_stage;
    } else {
      pts = host_in ? reinterpret_cast<const double*>(sl.in_dev) : points + 3 * start;
    }
    sl.pts = pts;
    FinalizeArgs& fa = sl.fa;
    std::memset(&fa, 0, sizeof(fa));
    fa.dots = dots;
    fa.index = nullptr;
    fa.nrt = nrt;
    fa.ld = ld;
    fa.n = nc;
    fa.out_off = 0;
    fa.mode = mode;
    fa.q = q;
    fa.par = par;
    double* stage_cursor = sl.out_stage;
    for (size_t k = 0; k < NS; ++k) {
      if (!(q & kSlots[k].bit)) continue;
      double* user = out.*(kSlots[k].field);
      const int w = kSlots[k].width;
      // column layout of grad / hess: component c of the chunk's point p at
      // c * ld + p (ld = n in the caller's array, nc in staging)
      const bool cols = col_out && w > 1;
      if (out_dev[k]) {
        fa.out.*(kSlots[k].field) = user + start * (cols ? 1 : w);
      } else {
        fa.out.*(kSlots[k].field) = stage_cursor;
        stage_cursor += (size_t)w * nc;
      }
      if (cols) (kSlots[k].bit == CT_Q_GRAD ? fa.grad_ld : fa.hess_ld) = out_dev[k] ? n : nc;
    }

    // the factor route reads the AOs in the order of its factor (screened
    // shell order, or the natural one when screening is off)
    const auto* ameta = use_factor ? static_cast<const ShellMeta*>(d_smeta_.get()) : meta;
    const auto* ageom = use_factor ? static_cast<const ShellGeom*>(d_sgeom_.get()) : geom;
    const double* apts = pts;  // the points the AO kernel and the contraction see
    if (screened) {
      timer_.begin(KernelTimer::SCREEN, stream_);
      launch_screen_sort(pts, nc, frame_, keys_in, keys_out, idx_in, perm, sort_tmp, sort_bytes, stream_);
      launch_gather_points(pts, perm, nc, spts, stream_);
      // the sort's output keys are spent: their space holds the inverse permutation
      launch_invert_permutation(perm, nc, reinterpret_cast<int32_t*>(keys_out), stream_);
      // the FP64 path's k-step tables hold the cutoff radii of its tighter
      // threshold
      launch_tile_activity(spts, nc, tp, ld / tp,
                           static_cast<const int32_t*>(mixed ? d_span_off_.get() : d_span_off64_.get()),
                           static_cast<const ScreenSpan*>(mixed ? d_spans_.get() : d_spans64_.get()), nk, klist,
                           kcount, kmask, nwords, d_active, stream_);
      timer_.end(stream_);
      apts = spts;
      // the finaliser runs over the chunk's points in their own order and
      // reads each one's results at its sorted position
      fa.gather = reinterpret_cast<const int32_t*>(keys_out);
      valid_tiles += (nc + tp - 1) / tp;
    }

    if (!mixed) {
      timer_.begin(KernelTimer::AO, stream_);
      launch_ao_f64(mode, lmax_, ameta, ageom, prim, nshell_, M_, Mp_, apts, nc, ld, static_cast<double*>(ao), 700.0,
                    stream_, screened ? kmask : nullptr, nwords, tp);
      timer_.end(stream_);
      timer_.begin(KernelTimer::CONTRACT, stream_);
      if (use_factor && pass_split)
        launch_contract_f64_factor_split(mode, static_cast<const double*>(d_L64s_.get()), factor_rows_, Mp_,
                                         static_cast<const double*>(ao), ld, part, dots, stream_,
                                         screened ? kmask : nullptr, nwords, tp, psi);
      else if (use_factor)
        launch_contract_f64_factor(mode, static_cast<const double*>(d_L64s_.get()), factor_rows_, Mp_,
                                   static_cast<const double*>(ao), ld, dots, stream_, screened ? kmask : nullptr,
                                   nwords, tp, psi);
      else
        launch_contract_f64(mode, static_cast<const double*>(d_P_.get()), Mp_, static_cast<const double*>(ao), ld,
                            dots, stream_);
      timer_.end(stream_);
      timer_.begin(KernelTimer::FINALIZE, stream_);
      launch_finalize(fa, stream_);
      timer_.end(stream_);
    } else {
      float* aoc = static_cast<float*>(ao);
      timer_.begin(KernelTimer::AO, stream_);
      launch_ao_f32(mode, lmax_, ameta, ageom, prim, nshell_, M_, Mp_, apts, nc, ld, aoc, 50.0,
                    screened ? kmask : nullptr, nwords, tp, stream_);
      timer_.end(stream_);
      timer_.begin(KernelTimer::CONTRACT, stream_);
      if (use_factor)
        launch_contract_mixed_factor(mode, static_cast<const uint16_t*>(d_Lb_.get()), factor_rows_, Mp_, aoc, ld,
                                     dots, screened ? klist : static_cast<const int32_t*>(d_klist_all_.get()),
                                     screened ? kcount : nullptr, stream_, psi);
      else
        launch_contract_mixed(mode, static_cast<const uint16_t*>(d_Pb_.get()), Mp_, aoc, ld, dots, stream_);
      timer_.end(stream_);
      FlagArgs fl;
      fl.gamma = use_factor ? kMixedGammaFactor : mixed_gamma(mode);
      fl.tol = par.mixed_tol;
      fl.rho_floor = par.rho_floor;
      fl.count = batch ? slots[0].count : sl.count;
      fl.index = batch ? list : sl.flag_idx;
      fl.index_base = batch ? start : 0;
      fl.max_est = d_maxest;
      fl.trunc = screened ? screen_trunc_ : 0.0;
      fl.target = nullptr;
      if (has_target)
        fl.target = host_tgt ? reinterpret_cast<const double*>(sl.in_dev + (host_in ? (size_t)Nc * 24 : 0))
                             : par.rho_target + start;
      if (!batch) launch_fill_zero(reinterpret_cast<unsigned long long*>(sl.count), 1, stream_);
      timer_.begin(KernelTimer::FINALIZE, stream_);
      launch_finalize_mixed(fa, fl, stream_);
      timer_.end(stream_);
      if (!batch) {
        CT_HIP_CHECK(hipMemcpyAsync(sl.cnt_host, sl.count, 4, hipMemcpyDeviceToHost, stream_));
        if (piped) {
          CT_HIP_CHECK(hipEventRecord(sl.ev[EV_COUNT], stream_));
          sl.used[EV_COUNT] = true;
        }
      }
    }
    ++chunks;
  };
  // Pipelined, the FP64 re-evaluation of a chunk's flagged points runs on
  // its own stream, beside the next chunk's kernels, when its workspace (for
  // the largest chunk) can be held for the whole call: growing it would free
  // memory that a re-evaluation in flight may still use.
  const bool re_side = piped && mixed &&
                       d_ws64_.reserve(reeval_bytes(
                           mode, chunk_points((int64_t)(reeval_bytes(mode, kPointAlign, use_factor) / kPointAlign), Nc),
                           use_factor));
  hipStream_t rs = re_side ? re_stream_ : stream_;
  // Completes the chunk in slot sl: (mixed, per chunk) waits for its flag
  // count and queues the FP64 re-evaluation of its flagged points; then
  // (piped) marks the slot's results final and queues their download.
  auto finish_chunk = [&](Slot& sl) {
    if (mixed && !batch) {
      if (piped)
        CT_HIP_CHECK(hipEventSynchronize(sl.ev[EV_COUNT]));
      else
        CT_HIP_CHECK(hipStreamSynchronize(stream_));
      const int64_t count = (int64_t) * reinterpret_cast<const int32_t*>(sl.cnt_host);
      // (after the chunk's finaliser)
      if (rs != stream_) CT_HIP_CHECK(hipStreamWaitEvent(rs, sl.ev[EV_COUNT], 0));
      if (count > 0) {
        timer_.begin(KernelTimer::REEVAL, rs);
        reevaluate(mode, sl.pts, 0, count, sl.flag_idx, sl.fa, use_factor, ws64, ws64_bytes, rs);
        timer_.end(rs);
        reevaluated += count;
      }
    }
    if (!piped) return;
    CT_HIP_CHECK(hipEventRecord(sl.ev[EV_READY], rs));
    sl.used[EV_READY] = true;
    if (stage_width > 0) {
      CT_HIP_CHECK(hipStreamWaitEvent(d2h_stream_, sl.ev[EV_READY], 0));
      CT_HIP_CHECK(hipMemcpyAsync(sl.out_host, sl.out_stage, (size_t)sl.nc * stage_width * 8, hipMemcpyDeviceToHost,
                                  d2h_stream_));
      CT_HIP_CHECK(hipEventRecord(sl.ev[EV_D2H], d2h_stream_));
      sl.used[EV_D2H] = true;
      sl.out_pending = true;
    }
  };
  // Copies the downloaded outputs of the chunk in slot sl into the caller's
  // host arrays.
  auto drain_slot = [&](Slot& sl) {
    if (!sl.out_pending) return;
    CT_HIP_CHECK(hipEventSynchronize(sl.ev[EV_D2H]));
    std::vector<HostCopy> jobs;
    const double* src = reinterpret_cast<const double*>(sl.out_host);
    for (size_t k = 0; k < NS; ++k) {
      if (!(q & kSlots[k].bit) || out_dev[k]) continue;
      const int w = kSlots[k].width;
      double* user = out.*(kSlots[k].field);
      if (col_out && w > 1)
        for (int c = 0; c < w; ++c)
          jobs.push_back(HostCopy{user + (size_t)c * n + sl.start, src + (size_t)c * sl.nc, (size_t)sl.nc * 8});
      else
        jobs.push_back(HostCopy{user + sl.start * w, src, (size_t)sl.nc * w * 8});
      src += (size_t)w * sl.nc;
    }
    host_copy(jobs.data(), (int)jobs.size());
    sl.out_pending = false;
  };

  if (piped) {
    // slot i % kPipeSlots; chunk i - 1 is completed once chunk i is queued
    Slot* prev = nullptr;
    int64_t i = 0, start = 0;
    for (int64_t nc : pipeline_chunks(n, Nc)) {
      Slot& sl = slots[i % kPipeSlots];
      drain_slot(sl);
      enqueue_chunk(sl, start, nc);
      if (prev) finish_chunk(*prev);
      prev = &sl;
      start += nc;
      ++i;
    }
    if (prev) finish_chunk(*prev);
    for (int s = 0; s < kPipeSlots; ++s) drain_slot(slots[(i + s) % kPipeSlots]);
    // (outputs in device memory: nothing above waited for the last
    // re-evaluation)
    CT_HIP_CHECK(hipStreamSynchronize(rs));
  } else {
    for (int64_t start = 0; start < n; start += Nc) {
      enqueue_chunk(slots[0], start, std::min(Nc, n - start));
      finish_chunk(slots[0]);
    }
  }
  if (batch) {
    CT_HIP_CHECK(hipMemcpyAsync(h_small, slots[0].count, 4, hipMemcpyDeviceToHost, stream_));
    CT_HIP_CHECK(hipStreamSynchronize(stream_));
    const int64_t count = (int64_t) * reinterpret_cast<int32_t*>(h_small);
    if (count > 0) {
      FinalizeArgs fa;
      std::memset(&fa, 0, sizeof(fa));
      fa.mode = mode;
      fa.q = q;
      fa.par = par;
      for (size_t k = 0; k < NS; ++k)
        if (q & kSlots[k].bit) fa.out.*(kSlots[k].field) = out.*(kSlots[k].field);
      if (col_out) fa.grad_ld = fa.hess_ld = n;
      // the FP64 re-evaluation runs in slices sized like the FP64 path's
      // chunks, on the flagged points in index order (appended in any order;
      // in order, points given in a spatial order stay so for the AO kernel's
      // waves)
      int bits = 1;  // the indices lie below n
      while (bits < 31 && (int64_t{1} << bits) < n) ++bits;
      const size_t tmp = (sort_indices_workspace(count, bits) + 255) / 256 * 256;
      const int32_t* sorted = list;
      if (d_flags_sorted_.reserve(tmp + sizeof(int32_t) * (size_t)count)) {
        char* base = static_cast<char*>(d_flags_sorted_.get());
        launch_sort_indices(list, reinterpret_cast<int32_t*>(base + tmp), count, bits, base, tmp, stream_);
        sorted = reinterpret_cast<const int32_t*>(base + tmp);
      }
      timer_.begin(KernelTimer::REEVAL, stream_);
      reevaluate(mode, points, col_in ? n : 0, count, sorted, fa, use_factor);
      timer_.end(stream_);
      reevaluated = count;
    }
  }
  if (mixed) CT_HIP_CHECK(hipMemcpyAsync(h_small + 1, d_maxest, 8, hipMemcpyDeviceToHost, stream_));
  if (screened)
    CT_HIP_CHECK(hipMemcpyAsync(h_small + 8, d_active, 128 * kScreenCounters, hipMemcpyDeviceToHost, stream_));
  CT_HIP_CHECK(hipStreamSynchronize(stream_));
  timer_.collect();
  if (rep) {
    rep->n_chunks = chunks;
    rep->n_reevaluated = reevaluated;
    if (mixed) std::memcpy(&rep->max_error_estimate, h_small + 1, 8);
    rep->factor_rank = use_factor ? factor_rank_ : -1;
    // k-steps holding basis functions (the padding beyond them is never active)
    const int64_t nk_basis = (M_ + 31) / 32;
    if (screened && valid_tiles > 0) {
      uint64_t active = 0;
      for (int c = 0; c < kScreenCounters; ++c) active += (uint64_t)h_small[8 + 16 * c];
      rep->active_fraction = (double)active / ((double)valid_tiles * (double)nk_basis);
    }
  }
}

}  // namespace ct
