// Register-resident 3x3 arithmetic and the constitutive models of the MLS-MPM
// growth step, templated on the working precision T (float or double).
//
// Every function is __host__ __device__: the HIP kernels use them, and the
// host build (g++ on mpm_host.cpp) unit-tests the same code on the CPU.
//
// Matrices are row-major 3x3, element (r, c) at a[3 r + c] -- the component
// order of the reference's (n, 3, 3) arrays (python/morphompm/state.py:12-16).
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MPM_HD __host__ __device__ __forceinline__
#else
#define MPM_HD inline
#endif

namespace mpm {

enum Model : int32_t { NEO_HOOKEAN = 0, HENCKY = 1, HERSCHEL_BULKLEY = 2 };

// Reference thresholds (python/morphompm/constitutive.py:17,
// python/morphompm/transfer.py:57,123).
constexpr double kDegen = 1.0e-9;
constexpr double kMassEps = 1.0e-12;

template <typename T>
struct M3 {
  T a[9];
  MPM_HD T& operator()(int r, int c) { return a[3 * r + c]; }
  MPM_HD const T& operator()(int r, int c) const { return a[3 * r + c]; }
};

template <typename T> MPM_HD M3<T> zero3() { M3<T> m; for (int i = 0; i < 9; ++i) m.a[i] = T(0); return m; }
template <typename T> MPM_HD M3<T> eye3() { M3<T> m = zero3<T>(); m.a[0] = m.a[4] = m.a[8] = T(1); return m; }

template <typename T> MPM_HD M3<T> mul(const M3<T>& A, const M3<T>& B) {
  M3<T> C;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      C(r, c) = A(r, 0) * B(0, c) + A(r, 1) * B(1, c) + A(r, 2) * B(2, c);
  return C;
}
// A^T B
template <typename T> MPM_HD M3<T> mul_tn(const M3<T>& A, const M3<T>& B) {
  M3<T> C;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      C(r, c) = A(0, r) * B(0, c) + A(1, r) * B(1, c) + A(2, r) * B(2, c);
  return C;
}
// A B^T
template <typename T> MPM_HD M3<T> mul_nt(const M3<T>& A, const M3<T>& B) {
  M3<T> C;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      C(r, c) = A(r, 0) * B(c, 0) + A(r, 1) * B(c, 1) + A(r, 2) * B(c, 2);
  return C;
}
template <typename T> MPM_HD M3<T> transpose(const M3<T>& A) {
  M3<T> B;
  for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) B(r, c) = A(c, r);
  return B;
}
template <typename T> MPM_HD M3<T> scaled(const M3<T>& A, T s) {
  M3<T> B; for (int i = 0; i < 9; ++i) B.a[i] = A.a[i] * s; return B;
}
template <typename T> MPM_HD M3<T> sym(const M3<T>& A) {
  M3<T> B;
  for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) B(r, c) = T(0.5) * (A(r, c) + A(c, r));
  return B;
}
template <typename T> MPM_HD T trace(const M3<T>& A) { return A.a[0] + A.a[4] + A.a[8]; }
template <typename T> MPM_HD T frob(const M3<T>& A, const M3<T>& B) {
  T s = T(0); for (int i = 0; i < 9; ++i) s += A.a[i] * B.a[i]; return s;
}
template <typename T> MPM_HD T det(const M3<T>& A) {
  return A.a[0] * (A.a[4] * A.a[8] - A.a[5] * A.a[7])
       - A.a[1] * (A.a[3] * A.a[8] - A.a[5] * A.a[6])
       + A.a[2] * (A.a[3] * A.a[7] - A.a[4] * A.a[6]);
}
// Cofactor matrix cof(A) = det(A) A^{-T}.
template <typename T> MPM_HD M3<T> cofactor(const M3<T>& A) {
  M3<T> K;
  K.a[0] = A.a[4] * A.a[8] - A.a[5] * A.a[7];
  K.a[1] = A.a[5] * A.a[6] - A.a[3] * A.a[8];
  K.a[2] = A.a[3] * A.a[7] - A.a[4] * A.a[6];
  K.a[3] = A.a[2] * A.a[7] - A.a[1] * A.a[8];
  K.a[4] = A.a[0] * A.a[8] - A.a[2] * A.a[6];
  K.a[5] = A.a[1] * A.a[6] - A.a[0] * A.a[7];
  K.a[6] = A.a[1] * A.a[5] - A.a[2] * A.a[4];
  K.a[7] = A.a[2] * A.a[3] - A.a[0] * A.a[5];
  K.a[8] = A.a[0] * A.a[4] - A.a[1] * A.a[3];
  return K;
}
// A^{-T} (no singularity guard: the reference's np.linalg.inv raises on a
// singular matrix; here the result is inf/nan and the caller sees it).
template <typename T> MPM_HD M3<T> inv_t(const M3<T>& A) {
  M3<T> K = cofactor(A);
  T id = T(1) / (A.a[0] * K.a[0] + A.a[1] * K.a[1] + A.a[2] * K.a[2]);
  return scaled(K, id);
}
template <typename T> MPM_HD M3<T> inv(const M3<T>& A) { return transpose(inv_t(A)); }

// ---------------------------------------------------------------------------
// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (the textbook
// method, e.g. Golub & Van Loan, "Matrix Computations", sec. 8.5).  On return
// lam[k] are the eigenvalues and column k of V the matching unit eigenvector,
// S = V diag(lam) V^T.  Converges quadratically; the sweep count bounds the
// work and the loop stops early once the off-diagonal part is negligible.
// ---------------------------------------------------------------------------
template <typename T> struct JacobiSweeps { static constexpr int value = 6; };
template <> struct JacobiSweeps<double> { static constexpr int value = 10; };

template <typename T>
MPM_HD void jacobi_rotate(T* S, M3<T>& V, int p, int q, int r) {
  // S packed as {s00, s11, s22, s01, s02, s12}; rotation in the (p, q) plane.
  auto off = [](int i, int j) { return (i + j == 1) ? 3 : (i + j == 2 ? 4 : 5); };
  const int ipq = off(p, q), ipr = off(p, r), iqr = off(q, r);
  const T apq = S[ipq];
  if (apq == T(0)) return;
  const T app = S[p], aqq = S[q];
  const T theta = (aqq - app) / (T(2) * apq);
  const T t = (theta >= T(0) ? T(1) : T(-1)) / (std::fabs(theta) + std::sqrt(T(1) + theta * theta));
  const T c = T(1) / std::sqrt(T(1) + t * t);
  const T s = t * c;
  S[p] = app - t * apq;
  S[q] = aqq + t * apq;
  S[ipq] = T(0);
  const T arp = S[ipr], arq = S[iqr];
  S[ipr] = c * arp - s * arq;
  S[iqr] = s * arp + c * arq;
  for (int k = 0; k < 3; ++k) {
    const T vkp = V(k, p), vkq = V(k, q);
    V(k, p) = c * vkp - s * vkq;
    V(k, q) = s * vkp + c * vkq;
  }
}

template <typename T>
MPM_HD void sym_eig3(const M3<T>& A, T lam[3], M3<T>& V) {
  T S[6] = {A(0, 0), A(1, 1), A(2, 2), A(0, 1), A(0, 2), A(1, 2)};
  V = eye3<T>();
  const T tiny = std::numeric_limits<T>::epsilon() * std::numeric_limits<T>::epsilon();
  for (int sweep = 0; sweep < JacobiSweeps<T>::value; ++sweep) {
    const T offn = S[3] * S[3] + S[4] * S[4] + S[5] * S[5];
    const T diag = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
    if (offn <= tiny * diag) break;
    jacobi_rotate(S, V, 0, 1, 2);
    jacobi_rotate(S, V, 0, 2, 1);
    jacobi_rotate(S, V, 1, 2, 0);
  }
  lam[0] = S[0]; lam[1] = S[1]; lam[2] = S[2];
}

// ---------------------------------------------------------------------------
// Daleckii-Krein divided difference of log for the Hencky VJP,
// (ln bi - ln bj) / (bi - bj) with the limit 1/bi at repeated eigenvalues
// (python/morphompm/constitutive.py:66-71).  FP64 keeps the reference's
// switch at |bi - bj| <= 1e-9; FP32 uses log1p(r)/r with r = (bi - bj)/bj and
// its series near r = 0, which has no cancellation.
// ---------------------------------------------------------------------------
MPM_HD double dk_logdiff(double bi, double bj, double lbi, double lbj) {
  const double d = bi - bj;
  return std::fabs(d) > kDegen ? (lbi - lbj) / d : 1.0 / bi;
}
MPM_HD float dk_logdiff(float bi, float bj, float, float) {
  const float r = (bi - bj) / bj;
  float q;
  if (std::fabs(r) < 1e-3f) q = 1.0f - r * (0.5f - r * (1.0f / 3.0f - 0.25f * r));
  else q = std::log1p(r) / r;
  return q / bj;
}

// ---------------------------------------------------------------------------
// Model parameters: NH/Hencky use prm[0] = mu, prm[1] = lambda; HB uses
// prm = {kappa, K, n, tau_y, m, eps} (python/morphompm/constitutive.py:115).
// ---------------------------------------------------------------------------
template <typename T>
struct ModelParams {
  int32_t kind;
  T p[6];
};

// Hencky core on Fe: tau and the Fe cotangent.  Templated on the compute
// type so the FP32 path can run the near-repeated branch in FP64.
template <typename T>
MPM_HD M3<T> hencky_tau_core(const M3<T>& Fe, T mu, T lam) {
  T beta[3];
  M3<T> U;
  sym_eig3(mul_nt(Fe, Fe), beta, U);
  T h[3];
  const T S = T(0.5) * lam * (std::log(beta[0]) + std::log(beta[1]) + std::log(beta[2]));
  for (int i = 0; i < 3; ++i) h[i] = mu * std::log(beta[i]) + S;
  M3<T> tau;
  for (int r = 0; r < 3; ++r)
    for (int c = r; c < 3; ++c) {
      const T v = U(r, 0) * h[0] * U(c, 0) + U(r, 1) * h[1] * U(c, 1) + U(r, 2) * h[2] * U(c, 2);
      tau(r, c) = v;
      tau(c, r) = v;
    }
  return tau;
}

// Fe cotangent from a precomputed eigen-decomposition b = U diag(beta) U^T.
template <typename T>
MPM_HD M3<T> hencky_fe_bar_eig(const M3<T>& Fe, const M3<T>& tb, T mu, T lam, const T beta[3],
                               const M3<T>& U, M3<T>* tau_out = nullptr) {
  T lb[3];
  for (int i = 0; i < 3; ++i) lb[i] = std::log(beta[i]);
  if (tau_out) {  // tau = U diag(mu ln b + lam/2 sum ln b) U^T from the same eigenpairs
    const T S = T(0.5) * lam * (lb[0] + lb[1] + lb[2]);
    const T h[3] = {mu * lb[0] + S, mu * lb[1] + S, mu * lb[2] + S};
    for (int r = 0; r < 3; ++r)
      for (int c = r; c < 3; ++c) {
        const T v = U(r, 0) * h[0] * U(c, 0) + U(r, 1) * h[1] * U(c, 1) + U(r, 2) * h[2] * U(c, 2);
        (*tau_out)(r, c) = v;
        (*tau_out)(c, r) = v;
      }
  }
  M3<T> Mm = scaled(tb, mu);
  const T hl = T(0.5) * lam * trace(tb);
  Mm.a[0] += hl; Mm.a[4] += hl; Mm.a[8] += hl;
  M3<T> K = mul(mul_tn(U, Mm), U);                       // U^T M U
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) K(i, j) *= dk_logdiff(beta[i], beta[j], lb[i], lb[j]);
  const M3<T> Bbar = mul_nt(mul(U, K), U);               // U (K o L) U^T
  return scaled(mul(Bbar, Fe), T(2));                    // b = Fe Fe^T -> 2 Bbar Fe
}

template <typename D, typename S> MPM_HD M3<D> cast3(const M3<S>& A) {
  M3<D> B; for (int i = 0; i < 9; ++i) B.a[i] = (D)A.a[i]; return B;
}

// FP32 near-repeated branch: when the smallest eigenvalue gap of b is below
// kNearRepeated times tr b, the FP32 path redoes the eigen-decomposition and
// the divided differences in FP64.  Above it FP32 is accurate: the FP32
// eigenvectors err by ~eps/gap, but U (K o L) U^T depends on them only
// through the variation of L across the near-degenerate block, which is
// ~gap, so the product errs by ~eps |M| for any gap the FP32 Jacobi sweep
// resolves.  The threshold sits at that resolution; a wider one sends whole
// waves (any lane in the branch) down the FP64 path under near-isotropic
// growth.
#ifndef MPM_HENCKY_FP64_DEGEN
#define MPM_HENCKY_FP64_DEGEN 1
#endif
constexpr float kNearRepeated = 1.0e-6f;

MPM_HD bool near_repeated(const float beta[3]) {
  const float s = std::fabs(beta[0]) + std::fabs(beta[1]) + std::fabs(beta[2]);
  const float g01 = std::fabs(beta[0] - beta[1]), g02 = std::fabs(beta[0] - beta[2]),
              g12 = std::fabs(beta[1] - beta[2]);
  const float g = g01 < g02 ? (g01 < g12 ? g01 : g12) : (g02 < g12 ? g02 : g12);
  return g < kNearRepeated * s;
}

// ---------------------------------------------------------------------------
// FP32 Hencky through the matrix logarithm, with no eigen-decomposition.
// tau = mu ln b + (lam/2) tr(ln b) I with b = Fe Fe^T.  With s = tr b / 3 and
// b' = b / s, ln b = ln(s) I + ln b', and ln b' = 2 artanh(Z) with
// Z = (b' - I)(b' + I)^{-1}, summed as Gregory's series
// ln((1 + z) / (1 - z)) = 2 sum_k z^{2k+1} / (2k + 1) through k = kLogTerms,
// by Horner's rule in Y = Z^2.  Every matrix of the series is a function of
// b', so they all commute and are symmetric (packed storage).  The first term
// left out is below |Z|^{2N+2} / (2N+3) relative; with N = 4 and |Z|_F <= 0.2
// (the eigenvalues of b within a factor 1.5 of their mean) that is < 1e-8,
// under FP32 resolution.  Outside that range the eigen path runs.
// The VJP is the reverse-mode derivative of the same steps: the Daleckii-Krein
// derivative of ln without its divided differences, so repeated eigenvalues
// need no separate branch.
// ---------------------------------------------------------------------------
#ifndef MPM_HENCKY_SERIES
#define MPM_HENCKY_SERIES 1
#endif
constexpr int kLogTerms = 4;
constexpr float kLogZ2Max = 0.04f;   // |Z|_F^2 bound of the series

struct S3 {  // packed symmetric 3x3
  float xx, yy, zz, xy, xz, yz;
};

// A B for commuting symmetric A, B (the product is symmetric).
MPM_HD S3 s3_mul(const S3& A, const S3& B) {
  S3 C;
  C.xx = A.xx * B.xx + A.xy * B.xy + A.xz * B.xz;
  C.yy = A.xy * B.xy + A.yy * B.yy + A.yz * B.yz;
  C.zz = A.xz * B.xz + A.yz * B.yz + A.zz * B.zz;
  C.xy = A.xx * B.xy + A.xy * B.yy + A.xz * B.yz;
  C.xz = A.xx * B.xz + A.xy * B.yz + A.xz * B.zz;
  C.yz = A.xy * B.xz + A.yy * B.yz + A.yz * B.zz;
  return C;
}
MPM_HD S3 s3_inv(const S3& A) {
  const float c00 = A.yy * A.zz - A.yz * A.yz, c01 = A.xz * A.yz - A.xy * A.zz,
              c02 = A.xy * A.yz - A.xz * A.yy, c11 = A.xx * A.zz - A.xz * A.xz,
              c12 = A.xy * A.xz - A.xx * A.yz, c22 = A.xx * A.yy - A.xy * A.xy;
  const float id = 1.0f / (A.xx * c00 + A.xy * c01 + A.xz * c02);
  return {c00 * id, c11 * id, c22 * id, c01 * id, c02 * id, c12 * id};
}
MPM_HD M3<float> s3_full(const S3& A) {
  M3<float> M;
  M.a[0] = A.xx; M.a[1] = A.xy; M.a[2] = A.xz;
  M.a[3] = A.xy; M.a[4] = A.yy; M.a[5] = A.yz;
  M.a[6] = A.xz; M.a[7] = A.yz; M.a[8] = A.zz;
  return M;
}

// The series' intermediates: s, W = (I + b')^{-1}, Z, Y = Z^2 and the Horner
// partial sums P[k] = sum_{m=k..N} Y^{m-k} / (2m + 1) for k < N.
struct LogSeries {
  float s;
  S3 W, Z, Y, P[kLogTerms];
};

// ln b = ln(s) I + A; false when b is outside the series' range.
MPM_HD bool log_series(const M3<float>& b, LogSeries& L, S3& A) {
  const float s = (b.a[0] + b.a[4] + b.a[8]) * (1.0f / 3.0f);
  if (!(s > 0.0f && s < 1.0e30f)) return false;
  const float r = 1.0f / s;
  const S3 E = {b.a[0] * r - 1.0f, b.a[4] * r - 1.0f, b.a[8] * r - 1.0f,
                b.a[1] * r, b.a[2] * r, b.a[5] * r};
  L.W = s3_inv(S3{E.xx + 2.0f, E.yy + 2.0f, E.zz + 2.0f, E.xy, E.xz, E.yz});
  L.Z = s3_mul(E, L.W);
  const S3& Z = L.Z;
  const float z2 = Z.xx * Z.xx + Z.yy * Z.yy + Z.zz * Z.zz +
                   2.0f * (Z.xy * Z.xy + Z.xz * Z.xz + Z.yz * Z.yz);
  if (!(z2 <= kLogZ2Max)) return false;  // also NaN
  L.Y = s3_mul(Z, Z);
  const float aN = 1.0f / (2 * kLogTerms + 1);
  S3 P = {aN, aN, aN, 0.0f, 0.0f, 0.0f};
  for (int k = kLogTerms - 1; k >= 0; --k) {
    P = s3_mul(L.Y, P);
    const float ak = 1.0f / (2 * k + 1);
    P.xx += ak;
    P.yy += ak;
    P.zz += ak;
    L.P[k] = P;
  }
  const S3 ZP = s3_mul(Z, P);
  A = {2.0f * ZP.xx, 2.0f * ZP.yy, 2.0f * ZP.zz, 2.0f * ZP.xy, 2.0f * ZP.xz, 2.0f * ZP.yz};
  L.s = s;
  return true;
}

// tau = mu ln b + (lam/2) tr(ln b) I from ln b = ln(s) I + A.
MPM_HD M3<float> tau_from_log(const S3& A, float s, float mu, float lam) {
  const float ls = std::log(s);
  const float d = mu * ls + 0.5f * lam * (3.0f * ls + A.xx + A.yy + A.zz);
  M3<float> t = s3_full(S3{mu * A.xx + d, mu * A.yy + d, mu * A.zz + d, mu * A.xy, mu * A.xz,
                           mu * A.yz});
  return t;
}

// General x symmetric and symmetric x general products for the reverse pass.
MPM_HD M3<float> mul_gs(const M3<float>& A, const S3& B) { return mul(A, s3_full(B)); }
MPM_HD M3<float> mul_sg(const S3& A, const M3<float>& B) { return mul(s3_full(A), B); }

// Fe cotangent through the series: reverse mode over log_series and
// tau_from_log, given tau_bar (symmetrised) tb.
MPM_HD M3<float> hencky_fe_bar_series(const M3<float>& Fe, const M3<float>& b, const M3<float>& tb,
                                      float mu, float lam, const LogSeries& L) {
  // tau = mu Lg + (lam/2) tr(Lg) I  ->  Lg_bar = M
  M3<float> M = scaled(tb, mu);
  const float hl = 0.5f * lam * trace(tb);
  M.a[0] += hl; M.a[4] += hl; M.a[8] += hl;
  // Lg = ln(s) I + A
  float sbar = trace(M) / L.s;
  // A = 2 Z P0
  M3<float> Zb = scaled(mul_gs(M, L.P[0]), 2.0f);
  M3<float> Pb = scaled(mul_sg(L.Z, M), 2.0f);
  // P_k = Y P_{k+1} + I / (2k + 1), P_N = I / (2N + 1)
  M3<float> Yb = zero3<float>();
  for (int k = 0; k < kLogTerms; ++k) {
    if (k + 1 < kLogTerms) {
      const M3<float> t = mul_gs(Pb, L.P[k + 1]);
      for (int i = 0; i < 9; ++i) Yb.a[i] += t.a[i];
    } else {
      const float aN = 1.0f / (2 * kLogTerms + 1);
      for (int i = 0; i < 9; ++i) Yb.a[i] += aN * Pb.a[i];
    }
    Pb = mul_sg(L.Y, Pb);
  }
  // Y = Z Z
  {
    const M3<float> t1 = mul_gs(Yb, L.Z), t2 = mul_sg(L.Z, Yb);
    for (int i = 0; i < 9; ++i) Zb.a[i] += t1.a[i] + t2.a[i];
  }
  // Z = E W with E = b / s - I;  W = (I + b / s)^{-1}
  const float r = 1.0f / L.s;
  const S3 E = {b.a[0] * r - 1.0f, b.a[4] * r - 1.0f, b.a[8] * r - 1.0f,
                b.a[1] * r, b.a[2] * r, b.a[5] * r};
  const M3<float> Eb = mul_gs(Zb, L.W);
  const M3<float> Wb = mul_sg(E, Zb);
  const M3<float> WWbW = mul_gs(mul_sg(L.W, Wb), L.W);
  M3<float> bpb;  // cotangent of b' = b / s
  for (int i = 0; i < 9; ++i) bpb.a[i] = Eb.a[i] - WWbW.a[i];
  // b' = r b, r = 1 / s, s = tr b / 3
  const float rbar = frob(bpb, b);
  sbar -= rbar * r * r;
  M3<float> bb = scaled(bpb, r);
  const float d = sbar * (1.0f / 3.0f);
  bb.a[0] += d; bb.a[4] += d; bb.a[8] += d;
  // b = Fe Fe^T  ->  Fe_bar = (bb + bb^T) Fe
  M3<float> S;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) S(i, j) = bb(i, j) + bb(j, i);
  return mul(S, Fe);
}

MPM_HD M3<double> hencky_fe_bar(const M3<double>& Fe, const M3<double>& tb, double mu, double lam,
                                M3<double>* tau_out = nullptr) {
  double beta[3];
  M3<double> U;
  sym_eig3(mul_nt(Fe, Fe), beta, U);
  return hencky_fe_bar_eig(Fe, tb, mu, lam, beta, U, tau_out);
}
MPM_HD M3<float> hencky_fe_bar(const M3<float>& Fe, const M3<float>& tb, float mu, float lam,
                               M3<float>* tau_out = nullptr) {
#if MPM_HENCKY_SERIES
  {
    const M3<float> b = mul_nt(Fe, Fe);
    LogSeries L;
    S3 A;
    if (log_series(b, L, A)) {
      if (tau_out) *tau_out = tau_from_log(A, L.s, mu, lam);
      return hencky_fe_bar_series(Fe, b, tb, mu, lam, L);
    }
  }
#endif
  float beta[3];
  M3<float> U;
  sym_eig3(mul_nt(Fe, Fe), beta, U);
#if MPM_HENCKY_FP64_DEGEN
  if (near_repeated(beta)) {
    const M3<double> Fed = cast3<double>(Fe);
    double bd[3];
    M3<double> Ud, td;
    sym_eig3(mul_nt(Fed, Fed), bd, Ud);
    const M3<double> r = hencky_fe_bar_eig(Fed, cast3<double>(tb), (double)mu, (double)lam, bd, Ud,
                                           tau_out ? &td : nullptr);
    if (tau_out) *tau_out = cast3<float>(td);
    return cast3<float>(r);
  }
#endif
  return hencky_fe_bar_eig(Fe, tb, mu, lam, beta, U, tau_out);
}

MPM_HD M3<double> hencky_tau(const M3<double>& Fe, double mu, double lam) {
  return hencky_tau_core(Fe, mu, lam);
}
MPM_HD M3<float> hencky_tau(const M3<float>& Fe, float mu, float lam) {
  // tau is smooth through repeated eigenvalues: FP32 throughout.
#if MPM_HENCKY_SERIES
  {
    LogSeries L;
    S3 A;
    if (log_series(mul_nt(Fe, Fe), L, A)) return tau_from_log(A, L.s, mu, lam);
  }
#endif
  return hencky_tau_core(Fe, mu, lam);
}

// Elastic stress tau(F, Fg) for NH / Hencky; HB handled separately.
template <typename T>
MPM_HD M3<T> kirchhoff(const ModelParams<T>& mp, const M3<T>& F, const M3<T>& Fg, const M3<T>& C) {
  if (mp.kind == HERSCHEL_BULKLEY) {
    // tau = kappa ln(det F / det Fg) I + 2 eta(gdot) D  (constitutive.py:126-132)
    const T kappa = mp.p[0], K = mp.p[1], n = mp.p[2], ty = mp.p[3], m = mp.p[4], eps = mp.p[5];
    M3<T> D = sym(C);
    const T gd = std::sqrt(T(2) * frob(D, D) + eps * eps);
    const T one_m_e = -std::expm1(-m * gd);
    const T eta = K * std::pow(gd, n - T(1)) + ty * one_m_e / gd;
    const T vol = kappa * std::log(det(F) / det(Fg));
    M3<T> tau = scaled(D, T(2) * eta);
    tau.a[0] += vol; tau.a[4] += vol; tau.a[8] += vol;
    return tau;
  }
  const T mu = mp.p[0], lam = mp.p[1];
  const M3<T> Fe = mul(F, inv(Fg));
  if (mp.kind == NEO_HOOKEAN) {
    // tau = mu (Fe Fe^T - I) + lam ln(det Fe) I  (constitutive.py:25-28)
    M3<T> b = mul_nt(Fe, Fe);
    const T lj = lam * std::log(det(Fe));
    M3<T> tau;
    for (int i = 0; i < 9; ++i) tau.a[i] = mu * b.a[i];
    tau.a[0] += lj - mu; tau.a[4] += lj - mu; tau.a[8] += lj - mu;
    return tau;
  }
  return hencky_tau(Fe, mu, lam);
}

// HB (1 - e^{-u}) and u e^{-u} - (1 - e^{-u}) with u = m gdot.  The latter
// cancels for small u; FP32 uses its Taylor series -u^2/2 + u^3/3 - u^4/8 +
// u^5/30 - u^6/144 + u^7/840 there.
MPM_HD double hb_dnum(double u) { const double e = std::exp(-u); return u * e - (1.0 - e); }
MPM_HD float hb_dnum(float u) {
  if (u < 0.05f)
    return u * u * (-0.5f + u * (1.0f / 3.0f + u * (-0.125f + u * (1.0f / 30.0f + u * (-1.0f / 144.0f + u * (1.0f / 840.0f))))));
  const float e = std::exp(-u);
  return u * e + std::expm1(-u);
}

// VJP of tau w.r.t. (F, Fg, C) given tau_bar; theta_bar[0..1] get the HB
// (tau_y, K) cotangents (zero for the elastic models).
// NH: constitutive.py:31-39.  Hencky: constitutive.py:50-75.  HB: :134-147.
// tau_out (optional) receives tau itself, from the same intermediates.
template <typename T>
MPM_HD void kirchhoff_vjp(const ModelParams<T>& mp, const M3<T>& F, const M3<T>& Fg, const M3<T>& C,
                          const M3<T>& tau_bar, M3<T>& F_bar, M3<T>& Fg_bar, M3<T>& C_bar,
                          T theta_bar[2], M3<T>* tau_out = nullptr) {
  const M3<T> tb = sym(tau_bar);
  const T trtb = trace(tb);
  theta_bar[0] = T(0);
  theta_bar[1] = T(0);
  if (mp.kind == HERSCHEL_BULKLEY) {
    const T kappa = mp.p[0], K = mp.p[1], n = mp.p[2], ty = mp.p[3], m = mp.p[4], eps = mp.p[5];
    F_bar = scaled(inv_t(F), kappa * trtb);
    Fg_bar = scaled(inv_t(Fg), -kappa * trtb);
    M3<T> D = sym(C);
    const T gd = std::sqrt(T(2) * frob(D, D) + eps * eps);
    const T one_m_e = -std::expm1(-m * gd);
    const T gpow = std::pow(gd, n - T(1));
    const T eta = K * gpow + ty * one_m_e / gd;
    const T deta = K * (n - T(1)) * gpow / gd + ty * hb_dnum(m * gd) / (gd * gd);
    const T tD = frob(tb, D);
    const T cD = T(4) * tD * deta / gd;
    for (int i = 0; i < 9; ++i) C_bar.a[i] = T(2) * eta * tb.a[i] + cD * D.a[i];
    theta_bar[0] = T(2) * one_m_e / gd * tD;
    theta_bar[1] = T(2) * gpow * tD;
    if (tau_out) {
      const T vol = kappa * std::log(det(F) / det(Fg));
      *tau_out = scaled(D, T(2) * eta);
      tau_out->a[0] += vol; tau_out->a[4] += vol; tau_out->a[8] += vol;
    }
    return;
  }
  C_bar = zero3<T>();
  const T mu = mp.p[0], lam = mp.p[1];
  const M3<T> Gi = inv(Fg);
  const M3<T> Fe = mul(F, Gi);
  M3<T> Fe_bar;
  if (mp.kind == NEO_HOOKEAN) {
    const M3<T> tbFe = mul(tb, Fe);
    const M3<T> FeiT = inv_t(Fe);
    for (int i = 0; i < 9; ++i) Fe_bar.a[i] = T(2) * mu * tbFe.a[i] + lam * trtb * FeiT.a[i];
    if (tau_out) {
      const M3<T> b = mul_nt(Fe, Fe);
      const T lj = lam * std::log(det(Fe));
      for (int i = 0; i < 9; ++i) tau_out->a[i] = mu * b.a[i];
      tau_out->a[0] += lj - mu; tau_out->a[4] += lj - mu; tau_out->a[8] += lj - mu;
    }
  } else {
    Fe_bar = hencky_fe_bar(Fe, tb, mu, lam, tau_out);
  }
  F_bar = mul_nt(Fe_bar, Gi);                              // Fe_bar Gi^T
  // Fg_bar = -Gi^T F^T Fe_bar Gi^T
  Fg_bar = scaled(mul_nt(mul_tn(Gi, mul_tn(F, Fe_bar)), Gi), T(-1));
}

// ---------------------------------------------------------------------------
// Quadratic B-spline on a node-centred grid (python/morphompm/transfer.py:18-30).
// base = floor(x/dx - 1/2), fx = x/dx - base in [0.5, 1.5).
// ---------------------------------------------------------------------------
// Contraction is off here so every kernel (sort key, payload, gathers)
// rounds x/dx and x/dx - 1/2 the same way and agrees on the cell.
template <typename T>
MPM_HD int cell_base(T x, T inv_dx, T& fx) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const T X = x * inv_dx;
  if (!(std::fabs(X) < T(1 << 20))) {  // NaN, inf or far off: no grid holds it
    fx = T(1);
    return -(1 << 24);
  }
  const T b = std::floor(X - T(0.5));
  fx = X - b;
  return (int)b;
}
template <typename T>
MPM_HD void bspline_w(T fx, T w[3]) {
  const T a = T(1.5) - fx, c = fx - T(1), d = fx - T(0.5);
  w[0] = T(0.5) * a * a;
  w[1] = T(0.75) - c * c;
  w[2] = T(0.5) * d * d;
}
template <typename T>
MPM_HD void bspline_dw(T fx, T dw[3]) {
  dw[0] = fx - T(1.5);
  dw[1] = T(-2) * (fx - T(1));
  dw[2] = fx - T(0.5);
}

}  // namespace mpm
