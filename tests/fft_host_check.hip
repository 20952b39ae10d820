// Host-side check of the register FFT (respackdev_amd/csrc/rp_fft.h) against a
// long-double direct DFT, for every length the pair-density kernels use and
// both signs.  Prints one line per length: N, relative max error (sign -1, +1).
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>

#include "../respackdev_amd/csrc/rp_fft.h"

using namespace rp;

template <int N>
double check(int sign) {
  double2 tab[N], x[N], y[N];
  for (int j = 0; j < N; ++j)
    tab[j] = make_double2(std::cos(2.0 * M_PI * j / N), sign * std::sin(2.0 * M_PI * j / N));
  std::mt19937 g(1000 + N);
  std::normal_distribution<double> d;
  for (int j = 0; j < N; ++j) x[j] = make_double2(d(g), d(g));
  fft_reg<N, 1, N>(x, y, tab);
  double err = 0.0, nrm = 0.0;
  for (int k = 0; k < N; ++k) {
    std::complex<long double> acc = 0;
    for (int j = 0; j < N; ++j)
      acc += std::complex<long double>(x[j].x, x[j].y) *
             std::polar((long double)1, (long double)sign * 2 * M_PIl * ((j * k) % N) / N);
    err = std::fmax(err, (double)std::abs(acc - std::complex<long double>(y[k].x, y[k].y)));
    nrm = std::fmax(nrm, (double)std::abs(acc));
  }
  return err / nrm;
}

template <int... Ns>
void run() {
  ((std::printf("%d %.3e %.3e\n", Ns, check<Ns>(-1), check<Ns>(1))), ...);
}

int main() {
  run<1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 18, 20, 24, 25, 27, 30, 32>();
  return 0;
}
