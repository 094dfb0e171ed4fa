// Peak-rate probes: what this device sustains for the instruction mixes the
// evaluator's kernels are bound by, measured with the same clocks, power and
// launch shape as a real run (the published peaks assume the top clock).
//   * bf16 matrix core: v_mfma_f32_16x16x32_bf16 on eight independent
//     accumulators per wave, two waves per SIMD;
//   * FP64 matrix core: v_mfma_f64_16x16x4f64, likewise;
//   * FP64 vector: eight independent FMA chains per lane;
//   * HBM: a streaming copy (read + write).
// Each kernel writes one value per lane so that nothing is optimised away.
#include <chrono>

#include "ct_internal.h"

namespace ct {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kAcc = 8;

__global__ __launch_bounds__(256, 2) void peak_mfma_bf16_kernel(int iters, float seed, float* __restrict__ out) {
  bf16x8 a, b;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    a[e] = (__bf16)(seed * (float)(threadIdx.x + e));
    b[e] = (__bf16)(seed * (float)(e + 1));
  }
  f4 acc[kAcc];
#pragma unroll
  for (int k = 0; k < kAcc; ++k) acc[k] = f4{0.f, 0.f, 0.f, 0.f};
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int k = 0; k < kAcc; ++k) acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[k], 0, 0, 0);
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kAcc; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256, 2) void peak_mfma_f64_kernel(int iters, double seed, double* __restrict__ out) {
  const double a = seed * (double)(threadIdx.x + 1), b = seed * 0.5;
  d4 acc[kAcc];
#pragma unroll
  for (int k = 0; k < kAcc; ++k) acc[k] = d4{0.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int k = 0; k < kAcc; ++k) acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[k], 0, 0, 0);
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kAcc; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void peak_valu_f64_kernel(int iters, double seed, double* __restrict__ out) {
  double x[kAcc];
#pragma unroll
  for (int k = 0; k < kAcc; ++k) x[k] = seed * (double)(threadIdx.x + k);
  const double m = 0.999999, c = 1e-7;
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int k = 0; k < kAcc; ++k) x[k] = fma(x[k], m, c);
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kAcc; ++k) s += x[k];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// n is a multiple of 4 x the grid's thread count: four 16-byte loads in flight per lane
__global__ __launch_bounds__(256) void peak_copy_kernel(const f4* __restrict__ src, f4* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += 4 * stride) {
    const f4 v0 = __builtin_nontemporal_load(src + k), v1 = __builtin_nontemporal_load(src + k + stride);
    const f4 v2 = __builtin_nontemporal_load(src + k + 2 * stride), v3 = __builtin_nontemporal_load(src + k + 3 * stride);
    __builtin_nontemporal_store(v0, dst + k);
    __builtin_nontemporal_store(v1, dst + k + stride);
    __builtin_nontemporal_store(v2, dst + k + 2 * stride);
    __builtin_nontemporal_store(v3, dst + k + 3 * stride);
  }
}

// Milliseconds of the fastest of three launches of `launch`, after two
// warm-up launches (the first launches run while the clocks ramp up).
template <class F>
float time_launch(F&& launch, hipStream_t s) {
  hipEvent_t a, b;
  CT_HIP_CHECK(hipEventCreate(&a));
  CT_HIP_CHECK(hipEventCreate(&b));
  launch();
  launch();
  float best = 0.f;
  for (int r = 0; r < 3; ++r) {
    CT_HIP_CHECK(hipEventRecord(a, s));
    launch();
    CT_HIP_CHECK(hipEventRecord(b, s));
    CT_HIP_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    CT_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    if (r == 0 || ms < best) best = ms;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

}  // namespace

// Device buffers and the stream of one measurement, released on every exit.
struct ProbeResources {
  void* out = nullptr;
  void* src = nullptr;
  void* dst = nullptr;
  hipStream_t s = nullptr;
  ~ProbeResources() {
    if (s) (void)hipStreamSynchronize(s);
    if (out) (void)hipFree(out);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (s) (void)hipStreamDestroy(s);
  }
};

void measure_peaks(double* bf16_tflops, double* f64_mfma_tflops, double* f64_valu_tflops, double* hbm_gbs) {
  int dev = 0, cu = 0;
  CT_HIP_CHECK(hipGetDevice(&dev));
  CT_HIP_CHECK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
  // two 256-thread workgroups per CU, eight rounds of them
  const int nblk = 16 * cu;
  const int64_t nthr = (int64_t)nblk * 256;
  const size_t copy_bytes = size_t(2) << 30;  // 2 GiB read, 2 GiB written
  // 2 GiB = 2^27 16-byte elements; the copy grid (a power-of-two count of
  // 256-lane workgroups) divides it into whole rounds of four loads per lane
  const int64_t n4 = (int64_t)(copy_bytes / 16);
  int copy_blk = 1024;
  while (copy_blk < 16 * cu) copy_blk *= 2;
  if (n4 % (4 * (int64_t)copy_blk * 256) != 0) throw Error(CT_E_INVALID, "peak copy: grid does not divide the buffer");

  ProbeResources r;
  CT_HIP_CHECK(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
  CT_HIP_CHECK(hipMalloc(&r.out, sizeof(double) * nthr));
  if (hipMalloc(&r.src, copy_bytes) != hipSuccess || hipMalloc(&r.dst, copy_bytes) != hipSuccess) {
    (void)hipGetLastError();
    throw Error(CT_E_NOMEM, "cannot allocate the peak probes' buffers");
  }
  hipStream_t s = r.s;
  CT_HIP_CHECK(hipMemsetAsync(r.src, 0, copy_bytes, s));
  const int it_mm = 2048, it_f64 = 512, it_valu = 4096;
  const float t_bf16 = time_launch(
      [&] {
        hipLaunchKernelGGL(peak_mfma_bf16_kernel, dim3(nblk), dim3(256), 0, s, it_mm, 1e-3f, static_cast<float*>(r.out));
      },
      s);
  const float t_f64 = time_launch(
      [&] {
        hipLaunchKernelGGL(peak_mfma_f64_kernel, dim3(nblk), dim3(256), 0, s, it_f64, 1e-3,
                           static_cast<double*>(r.out));
      },
      s);
  const float t_valu = time_launch(
      [&] {
        hipLaunchKernelGGL(peak_valu_f64_kernel, dim3(nblk), dim3(256), 0, s, it_valu, 1e-3,
                           static_cast<double*>(r.out));
      },
      s);
  const float t_copy = time_launch(
      [&] {
        hipLaunchKernelGGL(peak_copy_kernel, dim3(copy_blk), dim3(256), 0, s, static_cast<const f4*>(r.src),
                           static_cast<f4*>(r.dst), n4);
      },
      s);
  CT_HIP_CHECK(hipGetLastError());
  const double waves = (double)nthr / 64.0;
  // flops per MFMA: 2 x 16 x 16 x 32 (bf16), 2 x 16 x 16 x 4 (FP64); FMA = 2 flops per lane
  *bf16_tflops = waves * it_mm * kAcc * 16384.0 / (t_bf16 * 1e-3) / 1e12;
  *f64_mfma_tflops = waves * it_f64 * kAcc * 2048.0 / (t_f64 * 1e-3) / 1e12;
  *f64_valu_tflops = (double)nthr * it_valu * kAcc * 2.0 / (t_valu * 1e-3) / 1e12;
  *hbm_gbs = 2.0 * (double)copy_bytes / (t_copy * 1e-3) / 1e9;
}

}  // namespace ct
