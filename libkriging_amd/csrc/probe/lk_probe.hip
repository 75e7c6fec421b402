// Measurement probes for the benchmark (not part of the Kriging library):
// a plain device copy -- the HBM bandwidth and traffic-counter reference of
// the roofline figures -- and an FP64 MFMA issue-rate probe with the shader
// clock it ran at.  Built into liblkprobe.so, loaded by bench.py only.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace {

constexpr int TPB = 256;
typedef double d4 __attribute__((ext_vector_type(4)));

thread_local std::string g_err;

// plain device copy (16-byte vectors, grid-stride)
__global__ __launch_bounds__(TPB) void k_copy(const double2* __restrict__ src, double2* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) dst[i] = src[i];
}

// Every wave runs `iters` rounds of CH independent V_MFMA_F64_16X16X4_F64
// accumulation chains on register operands (no memory traffic in the loop);
// each thread stores one sum so the chains stay live.  Thread 0 of each
// workgroup records the shader-clock and constant-rate counters around the
// loop: stamps[4 b ..] = clock start, clock end, wall start, wall end.
template <int CH>
__global__ __launch_bounds__(TPB) void k_mfma_probe(int iters, double* __restrict__ out,
                                                    unsigned long long* __restrict__ stamps) {
  const unsigned long long c0 = clock64(), w0 = wall_clock64();
  d4 acc[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
  const double a = 1.0 + 1e-12 * threadIdx.x, b = 1.0 - 1e-12 * threadIdx.x;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < CH; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[(size_t)blockIdx.x * TPB + threadIdx.x] = s;
  const unsigned long long c1 = clock64(), w1 = wall_clock64();
  if (threadIdx.x == 0) {
    stamps[4 * (size_t)blockIdx.x + 0] = c0;
    stamps[4 * (size_t)blockIdx.x + 1] = c1;
    stamps[4 * (size_t)blockIdx.x + 2] = w0;
    stamps[4 * (size_t)blockIdx.x + 3] = w1;
  }
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
  }
  return -1;
}

struct Timer {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Timer() {
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
  }
  ~Timer() {
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
  float ms() {
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    float v = 0.f;
    check(hipEventElapsedTime(&v, e0, e1), "hipEventElapsedTime");
    return v;
  }
};

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { check(hipMalloc(&p, bytes), "hipMalloc"); }
  ~DevBuf() { hipFree(p); }
};

}  // namespace

extern "C" {

const char* lkprobe_last_error(void) { return g_err.c_str(); }

// mean time of a plain copy kernel over `bytes` (a multiple of 16)
int lkprobe_device_copy(int device, size_t bytes, int reps, double* ms_out) {
  return guarded([&] {
    if (bytes == 0 || bytes % 16 != 0 || reps <= 0) throw std::invalid_argument("bytes: positive multiple of 16");
    check(hipSetDevice(device), "hipSetDevice");
    DevBuf src(bytes), dst(bytes);
    Timer t;
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(TPB), 0, 0, (const double2*)src.p, (double2*)dst.p, n);
    check(hipEventRecord(t.e0, 0), "hipEventRecord");
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_copy, dim3(4096), dim3(TPB), 0, 0, (const double2*)dst.p, (double2*)src.p, n);
    check(hipGetLastError(), "k_copy");
    check(hipEventRecord(t.e1, 0), "hipEventRecord");
    *ms_out = t.ms() / reps;
  });
}

// FP64 MFMA probe: kernel time, achieved TFLOP/s and the mean shader clock
// (MHz) of the workgroups while they ran, from the shader-clock counter
// against the constant-rate counter of hipDeviceAttributeWallClockRate
int lkprobe_mfma(int device, int blocks, int iters, int chains, double* ms_out, double* tflops_out,
                 double* clock_mhz_out) {
  return guarded([&] {
    if (blocks <= 0 || iters <= 0 || !(chains == 1 || chains == 2 || chains == 4 || chains == 8))
      throw std::invalid_argument("blocks, iters > 0 and chains in {1, 2, 4, 8}");
    check(hipSetDevice(device), "hipSetDevice");
    int wall_khz = 0;
    check(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device), "wall clock rate");
    DevBuf out(sizeof(double) * (size_t)blocks * TPB), st(sizeof(unsigned long long) * 4 * (size_t)blocks);
    auto launch = [&] {
      double* o = (double*)out.p;
      unsigned long long* s = (unsigned long long*)st.p;
      switch (chains) {
        case 1: hipLaunchKernelGGL(k_mfma_probe<1>, dim3(blocks), dim3(TPB), 0, 0, iters, o, s); break;
        case 2: hipLaunchKernelGGL(k_mfma_probe<2>, dim3(blocks), dim3(TPB), 0, 0, iters, o, s); break;
        case 4: hipLaunchKernelGGL(k_mfma_probe<4>, dim3(blocks), dim3(TPB), 0, 0, iters, o, s); break;
        default: hipLaunchKernelGGL(k_mfma_probe<8>, dim3(blocks), dim3(TPB), 0, 0, iters, o, s); break;
      }
      check(hipGetLastError(), "k_mfma_probe");
    };
    launch();  // warm
    Timer t;
    check(hipEventRecord(t.e0, 0), "hipEventRecord");
    launch();
    check(hipEventRecord(t.e1, 0), "hipEventRecord");
    const double ms = t.ms();
    std::vector<unsigned long long> h(4 * (size_t)blocks);
    check(hipMemcpy(h.data(), st.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost), "download");
    double sum = 0.0;
    int cnt = 0;
    for (int b = 0; b < blocks; ++b) {
      const double dc = (double)(h[4 * b + 1] - h[4 * b]), dw = (double)(h[4 * b + 3] - h[4 * b + 2]);
      if (dw > 0) {
        sum += dc / dw * (wall_khz * 1e-3);
        ++cnt;
      }
    }
    const double flop = (double)blocks * 4.0 * iters * chains * 2.0 * 16 * 16 * 4;
    *ms_out = ms;
    *tflops_out = flop / (ms * 1e-3) / 1e12;
    *clock_mhz_out = cnt ? sum / cnt : 0.0;
  });
}

}  // extern "C"
