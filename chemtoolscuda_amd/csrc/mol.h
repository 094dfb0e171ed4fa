// Host-side molecule handle: owns the device shell table, P, a HIP stream and
// the chunk workspace.
#pragma once

#include <memory>

#include "ct_internal.h"

namespace ct {

struct HostBasis {
  std::vector<int32_t> shell_atom, shell_l, shell_nprim, shell_prim_off;
  std::vector<double> exps, coefs;
  int nbasis = 0;
};

struct FchkData {
  HostBasis basis;
  std::vector<double> atom_xyz;   // natom x 3 (bohr)
  std::vector<double> charges;    // nuclear (core) charges
  std::vector<int32_t> numbers;   // atomic numbers
  std::vector<double> P;          // nbasis x nbasis, row-major
};

// Native reader for Gaussian formatted checkpoint files (fchk.cpp).
FchkData load_fchk(const std::string& path);
// AIM .wfn / .wfx and Molden readers (wavefunction.cpp), and the dispatch on
// the file extension that also covers .fchk.
FchkData load_wfn(const std::string& path);
FchkData load_wfx(const std::string& path);
FchkData load_molden(const std::string& path);
FchkData load_wavefunction(const std::string& path);

// Caps the device memory all DeviceBuffers together may hold (bytes; -1:
// no cap); a reserve() beyond it fails as a failed hipMalloc does.
void set_device_allocation_limit(int64_t bytes);
int64_t device_allocation_held();

// Device address of the allocation-free entry point's scratch on a device
// (CT_NOALLOC_SCRATCH_BYTES; looked up on the first call per device).
void* noalloc_scratch(int device);

// Host memory copies (dst, src, bytes), split over a few threads when large
// (host_copy.cpp); returns when all are done.
struct HostCopy {
  void* dst;
  const void* src;
  size_t bytes;
};
void host_copy(const HostCopy* jobs, int njobs);

// A caller-provided device region that an evaluation uses for all of its
// scratch instead of the handle's growing buffers.
struct Workspace {
  char* base;
  size_t bytes;
};

class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  void* get() const { return ptr_; }
  size_t size() const { return bytes_; }
  // grows (never shrinks) the allocation; returns false if hipMalloc fails
  bool reserve(size_t bytes);
  void release();

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

class PinnedBuffer {
 public:
  PinnedBuffer() = default;
  ~PinnedBuffer() { release(); }
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  void* get() const { return ptr_; }
  void reserve(size_t bytes);
  void release();

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

// Records HIP events around kernel launches on one stream and accumulates
// their elapsed times when collected.
class KernelTimer {
 public:
  enum Kind { AO = 0, CONTRACT, FINALIZE, ESP, BACKWARD, REEVAL, SCREEN, NKIND };
  ~KernelTimer();
  void enable(bool on);
  bool enabled() const { return on_; }
  void begin(Kind k, hipStream_t s);
  void end(hipStream_t s);
  void collect();  // after a stream synchronisation
  ct_kernel_stats stats() const;

 private:
  struct Pending {
    Kind kind;
    hipEvent_t a, b;
  };
  bool on_ = false;
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  double ms_[NKIND] = {0};
  int64_t calls_[NKIND] = {0};
  hipEvent_t take();
};

class Mol {
 public:
  Mol(const HostBasis& basis, const double* atom_xyz, const double* charges, const int32_t* numbers, int natom,
      const double* P);
  ~Mol();

  int nbasis() const { return M_; }
  int device() const { return device_; }
  int nbasis_pad() const { return Mp_; }
  int natom() const { return natom_; }
  const std::vector<double>& coordinates() const { return xyz_; }
  const std::vector<double>& charges() const { return charges_; }
  const std::vector<int32_t>& numbers() const { return numbers_; }
  const std::vector<double>& density_matrix() const { return P_; }
  const HostBasis& basis() const { return basis_; }

  void set_coordinates(const double* xyz);  // host or device
  void set_density_matrix(const double* P);  // host or device
  void set_workspace_limit(int64_t bytes) { ws_limit_ = bytes; }
  // Spatial screening of the factor route: k-steps of basis functions whose
  // values and derivatives are at most eps (FP64 path: eps x
  // kFp64ScreenFactor) at every point of a tile are skipped (mixed path: their
  // bound enters the error estimate); 0 turns it off.
  void set_screening(double eps);
  double screening() const { return screen_eps_; }
  hipStream_t stream() const { return stream_; }
  void synchronize();
  KernelTimer& timer() { return timer_; }

  // points: host or device, (n, 3) row-major (layout CT_LAYOUT_ROW) or (3, n)
  // (CT_LAYOUT_COL); or points == nullptr and cube != nullptr: the points of
  // a uniform grid, generated on the device chunk by chunk.
  // fixed: evaluate inside that region only, allocating nothing (device
  // points and outputs; flagged points re-evaluated chunk by chunk).
  void eval(const double* points, int64_t n, int layout, uint32_t q, int precision, const ct_eval_params& par,
            const ct_outputs& out, ct_eval_report* rep, const CubicGrid* cube = nullptr,
            const Workspace* fixed = nullptr);
  void esp(const double* points, int64_t n, int layout, int precision, double* out);
  void backward(const double* points, int64_t n, int layout, const double* w, int precision, double* dP, double* dR);
  int64_t esp_pairs() const { return esp_npairs_; }
  // primitive pairs of the ESP pre-contraction by total angular momentum 0..8
  // (after screening; builds the pre-contraction if needed)
  void esp_pairs_by_L(int64_t* out);

 private:
  void upload_geometry();
  void upload_P();
  // screened layout of the factor route (shell order, bf16 and FP64 factor in
  // that order, k-step spans, truncation bounds); after P, the geometry or
  // the threshold change
  void build_screen();
  void set_morton_frame();  // the Morton lattice of the point sort, from the centres
  int64_t chunk_points(int64_t per_point, int64_t n) const;
  // FP64 re-evaluation of `count` points (pts: (m, 3) row-major, col_stride
  // 0, or column-major with that stride; index: positions in pts), in slices
  // that fit ws (ws_bytes) or, ws null, the handle's growing FP64 workspace;
  // on stream rs (null: the handle's stream)
  void reevaluate(int mode, const double* pts, int64_t col_stride, int64_t count, const int32_t* index,
                  const FinalizeArgs& fin, bool factor, char* ws = nullptr, size_t ws_bytes = 0,
                  hipStream_t rs = nullptr);
  // workspace bytes of one re-evaluation slice of at most count points
  size_t reeval_bytes(int mode, int64_t count, bool factor) const;
  // whether this mode runs on the factor of P (ct_eval_params.route, the
  // same decision in both precisions)
  bool factor_route(int mode, int route, bool mixed) const;

  HostBasis basis_;
  int M_ = 0, Mp_ = 0, natom_ = 0, nshell_ = 0;
  int lmax_ = 0;  // largest |l| of the basis: selects the AO kernel instantiation
  std::vector<double> xyz_, charges_, P_;
  std::vector<int32_t> numbers_;
  std::vector<ShellMeta> meta_;
  std::vector<PrimData> prim_;
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  DeviceBuffer d_meta_, d_geom_, d_prim_, d_P_;
  DeviceBuffer d_Pb_;      // mixed path: bf16 pieces of P (hi, mid, lo) in MFMA fragment order
  DeviceBuffer d_Lb_;      // mixed path: bf16 pieces of L^T (P = L L^T) in MFMA fragment order
  DeviceBuffer d_L64_;     // L in FP64, k-major (nbasis_pad x factor rows), natural order (residual check)
  DeviceBuffer d_L64s_;    // the same in the screened function order (FP64 factor route, re-evaluation)
  int factor_rank_ = -1;   // rank of the factor, -1 when P has none (not PSD or rank too high)
  int factor_rows_ = 0;    // rows of L^T padded to a multiple of 128
  std::vector<double> L_;  // the factor, rank x M (row j = column j of L), natural order
  // spatial screening of the factor route
  double screen_eps_ = 1e-15;
  bool screen_built_ = false;
  bool layout_stale_ = true;           // the shell order must be rebuilt (threshold changed, centres moved far)
  std::vector<double> xyz_layout_;     // centres the screened shell order was built for
  static constexpr double kLayoutMove = 1.0;  // bohr: a larger move of any centre rebuilds the order
  std::vector<double> shell_radius_;   // cutoff radii at screen_eps_
  ScreenLayout layout_;    // identity order (and no spans) when screening is off
  MortonFrame frame_{};
  DeviceBuffer d_smeta_, d_sgeom_;     // shell table in the screened order
  DeviceBuffer d_span_off_, d_spans_;  // per k-step shells and cutoff radii
  // The FP64 path screens the factor route at screen_eps_ * kFp64ScreenFactor
  // (1e-22 by default): a skipped AO value is at most that, and rho moves by
  // at most 2 eps64 ||c||_2 sqrt(rho) (c_j = ||L_j||_1), below rounding for
  // FP64 parity wherever rho is above ~1e-12 (||c||_2 of a few thousand).
  static constexpr double kFp64ScreenFactor = 1e-7;
  std::vector<double> shell_radius64_;  // cutoff radii at the FP64 threshold
  DeviceBuffer d_span_off64_, d_spans64_;  // per k-step shells and the FP64 cutoff radii
  DeviceBuffer d_klist_all_;           // 0 .. nbasis_pad / 32 - 1: the k-steps of an unscreened tile
  double screen_trunc_ = 0.0;          // 4 eps ||c||_2, c_j = ||L_j||_1: rho bound of the skipped blocks / sqrt(rho)
  DeviceBuffer d_ws_;      // AO + dots + staged points / outputs
  DeviceBuffer d_ws64_;    // mixed path: FP64 re-evaluation of flagged points
  DeviceBuffer d_flags_;   // mixed path: flagged point list of a whole call (device outputs)
  DeviceBuffer d_flags_sorted_;  // the same in index order, and the sort's scratch
  PinnedBuffer h_small_;   // pinned readback of the flag counter / error estimate
  PinnedBuffer h_stage_;   // pinned staging of host-array inputs/outputs, one region per pipeline slot
  // Host-array pipeline (eval with points, outputs or targets in host
  // memory): chunks pass through kPipeSlots slots of pinned and device
  // staging; uploads, downloads and the FP64 re-evaluation of flagged points
  // run on their own streams, overlapping the kernels of the neighbouring
  // chunks.  Created on first use.
  static constexpr int kPipeSlots = 3;
  static constexpr int64_t kHostChunk = int64_t(1) << 20;  // largest pipelined chunk (points)
  enum PipeEvent { EV_H2D = 0, EV_COUNT, EV_READY, EV_D2H, NPIPEEV };
  hipStream_t h2d_stream_ = nullptr, d2h_stream_ = nullptr, re_stream_ = nullptr;
  hipEvent_t pipe_ev_[kPipeSlots][NPIPEEV] = {};
  void ensure_pipeline();
  int64_t ws_limit_ = int64_t(8) << 30;  // set from the device's memory at construction
  KernelTimer timer_;
  // ESP: Hermite pre-contraction of P, rebuilt when P or the centres change
  bool esp_valid_ = false;
  int esp_nseg_ = 0;
  int esp_lmax_ = 0;  // largest pair L of the ESP segments (selects the Boys grid)
  int64_t esp_pairs_by_L_[9] = {0};
  int64_t esp_npairs_ = 0;
  DeviceBuffer d_tiles_;    // mixed adjoint: upper-triangle Gram tile list
  bool tiles_ready_ = false;
  DeviceBuffer d_func_atom_;  // adjoint: atom of each basis function (original order)
  bool func_atom_ready_ = false;
  DeviceBuffer d_esp_rec_, d_esp_rec32_, d_esp_seg_, d_esp_atoms_, d_boys_, d_esp_io_, d_esp_part_;
  DeviceBuffer d_adj_;     // adjoint: G, T, slice partials
  void prepare_esp();
};

// true when ptr is device (HBM) memory usable directly by kernels
bool is_device_pointer(const void* ptr);

}  // namespace ct
