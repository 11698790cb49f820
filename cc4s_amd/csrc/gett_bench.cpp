// Contraction-kernel benchmark on the CCSD statement shapes.
//
//   gett_bench [generation=2] [repetitions=5] [No Nv NG slice]
//   (generation 1, 2: FP64 four-product kernels; 3: split-bf16 mixed
//    precision; 4, 5, 6: FP64 three-product kernels (contraction-engine
//    generations 3, 4 and 5), each checked against generation 2 on the same
//    operands; 0: a plain 1 GiB device copy kernel, the reference for HBM
//    counters)
//
// Runs the complex GETT on device buffers for the shapes of the ladder
// (formation and contraction, Ccsd.cxx:588-590), the ring and crossed-ring
// terms (Ccsd.cxx:499,502,516-517) and a Gamma-contraction term
// (Ccsd.cxx:489), and prints milliseconds per call and algorithmic TFLOP/s
// (8 flop per complex MAC).  Default sizes: No=32, Nv=256, NG=1500, slice 32.
#include "contract.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace cc4s_amd;

namespace {

__global__ void fill(double *p, size_t n, double s) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull)
    p[i] = s * static_cast<double>((i * 2654435761ull) % 1000) / 1000.0 - 0.5 * s;
}

// reads and writes n doubles once: 8n bytes fetched, 8n bytes written
__global__ void copy_kernel(const double *__restrict__ x, double *__restrict__ y, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull)
    y[i] = x[i];
}

int copy_reference(int reps) {
  const size_t n = size_t(1) << 27;  // 1 GiB of doubles
  double *x = nullptr, *y = nullptr;
  CC4S_HIP(hipMalloc(&x, n * 8));
  CC4S_HIP(hipMalloc(&y, n * 8));
  fill<<<4096, 256>>>(x, n, 1.0);
  hipEvent_t e0, e1;
  CC4S_HIP(hipEventCreate(&e0));
  CC4S_HIP(hipEventCreate(&e1));
  CC4S_HIP(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; ++r) copy_kernel<<<8192, 256>>>(x, y, n);
  CC4S_HIP(hipEventRecord(e1, nullptr));
  CC4S_HIP(hipEventSynchronize(e1));
  float ms = 0;
  CC4S_HIP(hipEventElapsedTime(&ms, e0, e1));
  std::printf("copy 1 GiB: %.3f ms/call, %.1f GB/s (read + write)\n", ms / reps,
              2.0 * n * 8 / (ms / reps * 1e-3) / 1e9);
  (void)hipFree(x);
  (void)hipFree(y);
  return 0;
}

struct Case {
  const char *name;
  std::string ia, ib, ic;
  std::vector<int64_t> la, lb, lc;
};

}  // namespace

int main(int argc, char **argv) {
  const int gen = argc > 1 ? std::atoi(argv[1]) : 2;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
  if (gen == 0) {
    try {
      return copy_reference(reps);
    } catch (const std::exception &e) {
      std::fprintf(stderr, "gett_bench: %s\n", e.what());
      return 1;
    }
  }
  const int64_t o = argc > 3 ? std::atoll(argv[3]) : 32;
  const int64_t v = argc > 4 ? std::atoll(argv[4]) : 256;
  const int64_t G = argc > 5 ? std::atoll(argv[5]) : 1500;
  const int64_t s = argc > 6 ? std::atoll(argv[6]) : o;
  // generation 3: the split-bf16 mixed-precision path (packing included)
  const int engine_gen = gen == 1 ? 1 : gen == 4 ? 3 : gen == 5 ? 4 : gen == 6 ? 5 : 2;
  const bool check = gen >= 4;
  const std::vector<Case> cases = {
      {"ladder-form  Gxc,Gyd->cxdy", "Gxc", "Gyd", "cxdy", {G, s, v}, {G, s, v}, {v, s, v, s}},
      {"ladder-ctr   cxdy,cdij->xyij", "cxdy", "cdij", "xyij", {v, s, v, s}, {v, v, o, o}, {s, s, o, o}},
      {"ring         akic,cbkj->abij", "akic", "cbkj", "abij", {v, o, o, v}, {v, v, o, o}, {v, v, o, o}},
      {"Gterm        Gai,Gbj->abij", "Gai", "Gbj", "abij", {G, v, o}, {G, v, o}, {v, v, o, o}},
      {"crossed      bkci,ackj->abij", "bkci", "ackj", "abij", {v, o, v, o}, {v, v, o, o}, {v, v, o, o}},
      {"Xakic-lhs    dclk,adil->akic", "dclk", "adil", "akic", {v, v, o, o}, {v, v, o, o}, {v, o, o, v}},
  };
  try {
    hipStream_t st;
    CC4S_HIP(hipStreamCreate(&st));
    set_complex_gett_generation(2);
    ContractionEngine Eref(st);
    set_complex_gett_generation(engine_gen);
    ContractionEngine E(st);
    E.set_profiling(true);
    if (gen == 3) {
      E.set_mixed(true);
      E.mixed_min_flops = 0.0;
    }
    for (const auto &c : cases) {
      Tensor A(Field::Complex, c.ia, c.la), B(Field::Complex, c.ib, c.lb),
          C(Field::Complex, c.ic, c.lc);
      fill<<<4096, 256, 0, st>>>(A.buf.as<double>(), A.size() * 2, 1.0);
      fill<<<4096, 256, 0, st>>>(B.buf.as<double>(), B.size() * 2, 1.0);
      fill<<<4096, 256, 0, st>>>(C.buf.as<double>(), C.size() * 2, 1.0);
      double rel = -1.0;
      if (check) {
        // relative Frobenius difference to the four-product kernel
        const size_t n = C.size() * 2;
        std::vector<double> ref(n), got(n);
        set_complex_gett_generation(2);
        Eref.contract(1.0, A.v, B.v, 0.0, C.v);
        CC4S_HIP(hipMemcpyAsync(ref.data(), C.buf.get(), n * 8, hipMemcpyDeviceToHost, st));
        set_complex_gett_generation(engine_gen);
        E.contract(1.0, A.v, B.v, 0.0, C.v);
        CC4S_HIP(hipMemcpyAsync(got.data(), C.buf.get(), n * 8, hipMemcpyDeviceToHost, st));
        CC4S_HIP(hipStreamSynchronize(st));
        double num = 0, den = 0;
        for (size_t i = 0; i < n; ++i) {
          num += (got[i] - ref[i]) * (got[i] - ref[i]);
          den += ref[i] * ref[i];
        }
        rel = den > 0 ? std::sqrt(num / den) : 0.0;
      }
      E.contract(1.0, A.v, B.v, 0.0, C.v);  // plan and warm up
      CC4S_HIP(hipStreamSynchronize(st));
      E.reset_timings();
      hipEvent_t e0, e1;
      CC4S_HIP(hipEventCreate(&e0));
      CC4S_HIP(hipEventCreate(&e1));
      const double f0 = E.stats.flops;
      CC4S_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < reps; ++r) E.contract(1.0, A.v, B.v, 1.0, C.v);
      CC4S_HIP(hipEventRecord(e1, st));
      CC4S_HIP(hipEventSynchronize(e1));
      float ms = 0;
      CC4S_HIP(hipEventElapsedTime(&ms, e0, e1));
      const double fl = E.stats.flops - f0;
      double kms = 0;
      std::string var;
      for (const auto &kv : E.timings()) {
        kms += kv.second.ms;
        var += kv.first + " ";
      }
      std::printf("gen%d %-32s %8.3f ms/call  %6.2f TFLOP/s (kernel only %6.2f)  %s",
                  gen, c.name, ms / reps, fl / (ms * 1e-3) / 1e12,
                  fl / (kms * 1e-3) / 1e12, var.c_str());
      if (check) std::printf(" rel.diff vs gen2 %.2e", rel);
      std::printf("\n");
      std::fflush(stdout);
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
    (void)hipStreamDestroy(st);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "gett_bench: %s\n", e.what());
    return 1;
  }
  return 0;
}
