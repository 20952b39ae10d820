// Length-N discrete Fourier transform of values held in registers,
//   y[k] = sum_j x[j] w^(jk),  w = tab[S] (a root of unity of order N),
// for N = 2^a 3^b 5^c, as a mixed-radix decimation-in-time FFT unrolled at
// compile time: N = P M splits into P transforms of length M over the
// decimated inputs x[p + P m], twiddled by w^(p k1), then M butterflies of
// length P (P = 4, 2, 3 or 5, with the odd ones in their symmetric form).
// tab[j] = w0^j (j < N0) for a root w0 of order N0; the transform of length N
// reads tab at steps S = N0 / N, so every root is a table entry at an index
// known at compile time (one table serves a transform and all its parts).
#pragma once

#include <hip/hip_runtime.h>

namespace rp {

__host__ __device__ __forceinline__ double2 fft_cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ double2 fft_add(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__host__ __device__ __forceinline__ double2 fft_sub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
// i s z for a real s
__host__ __device__ __forceinline__ double2 fft_iscale(double s, double2 z) {
  return make_double2(-s * z.y, s * z.x);
}

template <int N>
constexpr int fft_radix() {
  return N % 4 == 0 ? 4 : (N % 2 == 0 ? 2 : (N % 3 == 0 ? 3 : 5));
}

// Butterfly of length P on a[], roots w_P^k = tab[k * T] (T = N0 / P).
template <int P, int T>
__host__ __device__ __forceinline__ void fft_butterfly(const double2 (&a)[P], double2 (&b)[P],
                                                       const double2* tab) {
  if constexpr (P == 1) {
    b[0] = a[0];
  } else if constexpr (P == 2) {
    b[0] = fft_add(a[0], a[1]);
    b[1] = fft_sub(a[0], a[1]);
  } else if constexpr (P == 4) {
    // w_4 = tab[T] = +-i
    const double s = tab[T].y;
    const double2 t0 = fft_add(a[0], a[2]), t1 = fft_sub(a[0], a[2]);
    const double2 t2 = fft_add(a[1], a[3]), t3 = fft_sub(a[1], a[3]);
    b[0] = fft_add(t0, t2);
    b[2] = fft_sub(t0, t2);
    b[1] = fft_add(t1, fft_iscale(s, t3));
    b[3] = fft_sub(t1, fft_iscale(s, t3));
  } else if constexpr (P == 3) {
    // w_3 = c + i s, w_3^2 = conj(w_3)
    const double2 w = tab[T];
    const double2 sm = fft_add(a[1], a[2]), df = fft_sub(a[1], a[2]);
    b[0] = fft_add(a[0], sm);
    const double2 t = make_double2(a[0].x + w.x * sm.x, a[0].y + w.x * sm.y);
    const double2 u = fft_iscale(w.y, df);
    b[1] = fft_add(t, u);
    b[2] = fft_sub(t, u);
  } else {
    static_assert(P == 5, "radix 2, 3, 4 or 5");
    // w_5^k for k = 1, 2; w^3 = conj(w^2), w^4 = conj(w)
    const double2 w1 = tab[T], w2 = tab[2 * T];
    const double2 s1 = fft_add(a[1], a[4]), d1 = fft_sub(a[1], a[4]);
    const double2 s2 = fft_add(a[2], a[3]), d2 = fft_sub(a[2], a[3]);
    b[0] = fft_add(a[0], fft_add(s1, s2));
    const double2 t1 = make_double2(a[0].x + w1.x * s1.x + w2.x * s2.x,
                                    a[0].y + w1.x * s1.y + w2.x * s2.y);
    const double2 u1 = fft_add(fft_iscale(w1.y, d1), fft_iscale(w2.y, d2));
    const double2 t2 = make_double2(a[0].x + w2.x * s1.x + w1.x * s2.x,
                                    a[0].y + w2.x * s1.y + w1.x * s2.y);
    const double2 u2 = fft_sub(fft_iscale(w2.y, d1), fft_iscale(w1.y, d2));
    b[1] = fft_add(t1, u1);
    b[4] = fft_sub(t1, u1);
    b[2] = fft_add(t2, u2);
    b[3] = fft_sub(t2, u2);
  }
}

template <int N, int S, int N0>
__host__ __device__ __forceinline__ void fft_reg(const double2 (&x)[N], double2 (&y)[N],
                                                 const double2* tab) {
  if constexpr (N == 1) {
    y[0] = x[0];
  } else {
    constexpr int P = fft_radix<N>();
    constexpr int M = N / P;
    static_assert(M * P == N, "N must be a product of 2, 3 and 5");
    double2 F[P][M];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      double2 xs[M];
#pragma unroll
      for (int m = 0; m < M; ++m) xs[m] = x[p + P * m];
      fft_reg<M, S * P, N0>(xs, F[p], tab);
    }
#pragma unroll
    for (int k1 = 0; k1 < M; ++k1) {
      double2 a[P], b[P];
      a[0] = F[0][k1];
#pragma unroll
      for (int p = 1; p < P; ++p)
        a[p] = (p * k1) % N == 0 ? F[p][k1] : fft_cmul(F[p][k1], tab[(S * p * k1) % N0]);
      fft_butterfly<P, S * M>(a, b, tab);
#pragma unroll
      for (int k2 = 0; k2 < P; ++k2) y[k1 + M * k2] = b[k2];
    }
  }
}

}  // namespace rp
