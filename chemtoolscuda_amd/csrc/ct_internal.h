// Internal declarations shared by the host orchestration and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "chemtools_amd.h"

namespace ct {

// ---------------------------------------------------------------------------
// Errors: every failure becomes a ct::Error carrying a CT_E_* status; the C
// ABI converts it to a return code and a thread-local message.
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define CT_HIP_CHECK(expr)                                                                     \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      throw ::ct::Error(CT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) + " (" + \
                                      __FILE__ + ":" + std::to_string(__LINE__) + ")");        \
  } while (0)

#define CT_REQUIRE(cond, msg)                                  \
  do {                                                         \
    if (!(cond)) throw ::ct::Error(CT_E_INVALID, (msg));       \
  } while (0)

// ---------------------------------------------------------------------------
// Device shell table.  Shell type codes index the generated SHELL_TYPES list
// (0,1,2,3,4 Cartesian; 5,6,7 = pure d,f,g).
int shell_code(int l);
int shell_nfunc(int l);

struct ShellMeta {  // read with wave-uniform (scalar) loads
  int32_t code;
  int32_t nprim;
  int32_t prim_off;
  int32_t first;  // first basis function of the shell
};

struct ShellGeom {
  double cx, cy, cz;
  double amin;  // smallest exponent of the shell (screening)
};

// Point-independent primitive data: exponent a and c_p * N_p(L), where
// N_p(L) = sqrt((2a/pi)^{3/2} (4a)^L) is the radial part of the L2
// normalisation (the angular part is folded into the generated polynomials).
struct PrimData {
  double a;
  double c;
};

// Granularities.
constexpr int kRowBlock = 64;      // contraction I-block (rows of P per pass)
constexpr int kPointAlign = 256;   // chunk stride alignment (points)

// Evaluation modes of the fused contraction (which GEMM columns and dots).
enum Mode : int { MODE_RHO = 0, MODE_GRAD = 1, MODE_LAP = 2, MODE_HESS = 3 };
// AO arrays evaluated per mode: val; val+grad; val+grad+lap; val+grad+hess.
constexpr int mode_nder(int m) { return m == MODE_RHO ? 1 : m == MODE_GRAD ? 4 : m == MODE_LAP ? 5 : 10; }
// Per-point dot products produced by the contraction epilogue.
//   RHO : [rho]
//   GRAD: [rho, gx, gy, gz]                          g_k = d_k phi . P phi
//   LAP : [rho, gx, gy, gz, L0, txx, tyy, tzz]       L0 = lap phi . P phi,  t_kk = d_k phi . P d_k phi
//   HESS: [rho, gx, gy, gz, hxx,hxy,hxz,hyy,hyz,hzz, txx,txy,tyy,txz,tyz,tzz]
//         h_kl = d_kl phi . P phi,  t_kl = d_k phi . P d_l phi
constexpr int mode_ndot(int m) { return m == MODE_RHO ? 1 : m == MODE_GRAD ? 4 : m == MODE_LAP ? 8 : 16; }

// ---------------------------------------------------------------------------
// Kernel launchers (defined in the .hip files).
struct FinalizeArgs {
  const double* dots;
  int64_t ld;
  int64_t n;        // points in this chunk
  int64_t out_off;  // index of the chunk's first point in the output arrays
  int mode;
  uint32_t q;
  ct_eval_params par;
  ct_outputs out;   // device pointers (full-length arrays, indexed with out_off)
};

void launch_ao_f64(int mode, const ShellMeta* meta, const ShellGeom* geom, const PrimData* prim, int nshell,
                   int nbasis, int nbasis_pad, const double* pts, int64_t npts, int64_t ld, double* ao,
                   double screen, hipStream_t s);
void launch_contract_f64(int mode, const double* P, int nbasis_pad, const double* ao, int64_t ld, double* dots,
                         hipStream_t s);
void launch_finalize(const FinalizeArgs& a, hipStream_t s);
void launch_norm_rows(const double* v, int64_t n, double* out, hipStream_t s);

}  // namespace ct
