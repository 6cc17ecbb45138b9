// CPU check of the run-time mixed-radix FFT (holotomocupy_amd/csrc/fft_any.hpp):
// runs its stages on the host, one emulated thread after another (the stage
// bodies are __host__ __device__; a stage only reads the buffer the previous
// stage wrote, so sequential threads see what the barrier would give them),
// and compares forward and inverse transforms with a direct FP64 DFT.
// Usage: fft_any_check N [N ...]; prints "N maxrel_fwd maxrel_inv radices" per
// length and exits non-zero on a factorisation failure.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "fft_any.hpp"

using namespace htamd;

template <bool INV>
static std::vector<float2> run(const AnyFft& f, std::vector<float2> x, int nthr) {
  std::vector<float2> b(x.size());
  float2* a = x.data();
  float2* c = b.data();
  int NS = 1;
  for (int s = 0; s < f.nst; ++s) {
    for (int t = 0; t < nthr; ++t) any_stage<INV>(a, c, f.N, f.radix[s], NS, f.tw, t, nthr);
    NS *= f.radix[s];
    std::swap(a, c);
  }
  return std::vector<float2>(a, a + f.N);
}

int main(int argc, char** argv) {
  int bad = 0;
  for (int i = 1; i < argc; ++i) {
    const int N = std::atoi(argv[i]);
    AnyFft f;
    if (!any_fft_factor(N, &f)) {
      std::printf("%d factor-failed\n", N);
      bad = 1;
      continue;
    }
    std::vector<float2> tw(N);
    for (int m = 0; m < N; ++m) {
      const double ang = -2.0 * M_PI * m / N;
      tw[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    f.tw = tw.data();
    std::mt19937 rng(1234 + N);
    std::normal_distribution<float> g;
    std::vector<float2> x(N);
    for (auto& v : x) v = make_float2(g(rng), g(rng));
    double err[2] = {0, 0};
    for (int inv = 0; inv < 2; ++inv) {
      // several thread counts: butterflies per thread from many to fewer than one
      for (int nthr : {1, 7, 64, 256}) {
        const auto y = inv ? run<true>(f, x, nthr) : run<false>(f, x, nthr);
        double num = 0, den = 0;
        for (int k = 0; k < N; ++k) {
          std::complex<double> s = 0;
          for (int n = 0; n < N; ++n) {
            const double ang = (inv ? 2.0 : -2.0) * M_PI * (double)((int64_t)n * k % N) / N;
            s += std::complex<double>(x[n].x, x[n].y) * std::complex<double>(std::cos(ang), std::sin(ang));
          }
          num += std::norm(std::complex<double>(y[k].x, y[k].y) - s);
          den += std::norm(s);
        }
        err[inv] = std::fmax(err[inv], std::sqrt(num / den));
      }
    }
    std::printf("%d %.3e %.3e", N, err[0], err[1]);
    for (int s = 0; s < f.nst; ++s) std::printf(" %d", f.radix[s]);
    std::printf("\n");
  }
  return bad;
}
