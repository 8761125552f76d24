// Host (CPU) build of the device math in mpm_math.h, for unit tests on a
// machine without a GPU.  Arrays are (n, 3, 3) row-major float64; the working
// precision is float (fp32 != 0) or double.
#include "mpm_math.h"

using namespace mpm;

namespace {
template <typename T>
void load(const double* src, M3<T>& m) { for (int i = 0; i < 9; ++i) m.a[i] = (T)src[i]; }
template <typename T>
void store(const M3<T>& m, double* dst) { for (int i = 0; i < 9; ++i) dst[i] = (double)m.a[i]; }

template <typename T>
ModelParams<T> params(int model, const double* prm) {
  ModelParams<T> mp;
  mp.kind = model;
  for (int i = 0; i < 6; ++i) mp.p[i] = (T)prm[i];
  return mp;
}

template <typename T>
void stress_n(int model, const double* prm, long n, const double* F, const double* Fg,
              const double* C, double* tau) {
  const ModelParams<T> mp = params<T>(model, prm);
  for (long p = 0; p < n; ++p) {
    M3<T> f, g, c;
    load(F + 9 * p, f); load(Fg + 9 * p, g); load(C + 9 * p, c);
    store(kirchhoff(mp, f, g, c), tau + 9 * p);
  }
}

template <typename T>
void vjp_n(int model, const double* prm, long n, const double* F, const double* Fg, const double* C,
           const double* tb, double* Fb, double* Fgb, double* Cb, double* th) {
  const ModelParams<T> mp = params<T>(model, prm);
  for (long p = 0; p < n; ++p) {
    M3<T> f, g, c, t, fb, gb, cb;
    load(F + 9 * p, f); load(Fg + 9 * p, g); load(C + 9 * p, c); load(tb + 9 * p, t);
    T theta[2];
    kirchhoff_vjp(mp, f, g, c, t, fb, gb, cb, theta);
    store(fb, Fb + 9 * p); store(gb, Fgb + 9 * p); store(cb, Cb + 9 * p);
    th[2 * p] = (double)theta[0];
    th[2 * p + 1] = (double)theta[1];
  }
}

template <typename T>
void eig_n(long n, const double* S, double* lam, double* V) {
  for (long p = 0; p < n; ++p) {
    M3<T> s, v;
    load(S + 9 * p, s);
    T l[3];
    sym_eig3(s, l, v);
    for (int k = 0; k < 3; ++k) lam[3 * p + k] = (double)l[k];
    store(v, V + 9 * p);
  }
}
}  // namespace

extern "C" {
void mpm_host_stress(int fp32, int model, const double* prm, long n, const double* F,
                     const double* Fg, const double* C, double* tau) {
  if (fp32) stress_n<float>(model, prm, n, F, Fg, C, tau);
  else stress_n<double>(model, prm, n, F, Fg, C, tau);
}
void mpm_host_stress_vjp(int fp32, int model, const double* prm, long n, const double* F,
                         const double* Fg, const double* C, const double* tb, double* Fb,
                         double* Fgb, double* Cb, double* th) {
  if (fp32) vjp_n<float>(model, prm, n, F, Fg, C, tb, Fb, Fgb, Cb, th);
  else vjp_n<double>(model, prm, n, F, Fg, C, tb, Fb, Fgb, Cb, th);
}
void mpm_host_sym_eig(int fp32, long n, const double* S, double* lam, double* V) {
  if (fp32) eig_n<float>(n, S, lam, V);
  else eig_n<double>(n, S, lam, V);
}
}
