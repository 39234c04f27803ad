// CPU oracle for the generalized-harmonic right-hand side, its RK4 step and
// its Kreiss–Oliger term.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library includes, links or
// calls this file: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, as the checker and as the CPU baseline.
//
// It is a restatement, in plain C++ templated on the scalar type T, of the
// reference package's CPU Float64 path (eschnett/TreeGeneralizedHarmonic.jl).
// Each function names the reference file:line it follows. T is instantiated
// with double, with a forward-mode dual number (derivatives for the analytic
// backgrounds and the chain-rule checks) and with an operation-counting double
// (the frozen flop count per point).
//
// Third-party pieces the reference calls but does not vendor are restated from
// their published mathematics and pinned by invariants in the tests:
//   * SpacetimeMetrics 1.6 (GitHub main): Minkowski, GaugeWave and KerrSchild
//     metrics and `dmetric` (forward-mode derivative, dg[a][b][c] = d_c g_ab);
//   * TreeAMR 0.3: vertex positions, closed-block ghost fill on a uniform
//     forest (same-level copies, periodic wrap, Dirichlet hook);
//   * IMEXRungeKutta 1.2: classical explicit RK4.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

constexpr int NC = 10;  // packed components of a symmetric 4x4 (pointwise.jl:53)

// ---------------------------------------------------------------------------
// Scalar helpers that every T provides (double overloads here; the dual and
// counting types define their own).
inline double val(double x) { return x; }

// ---------------------------------------------------------------------------
// Forward-mode dual number with K partials (restates the ForwardDiff passes
// SpacetimeMetrics' `dmetric` makes; nestable for second derivatives).
template <class V, int K>
struct Dual {
  V v{};
  V d[K]{};
  Dual() = default;
  Dual(double x) : v(x) {
    for (int k = 0; k < K; ++k) d[k] = V(0.0);
  }
  Dual(const V& x, int) : v(x) {
    for (int k = 0; k < K; ++k) d[k] = V(0.0);
  }
};
template <class V, int K>
inline Dual<V, K> operator+(const Dual<V, K>& a, const Dual<V, K>& b) {
  Dual<V, K> r;
  r.v = a.v + b.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <class V, int K>
inline Dual<V, K> operator-(const Dual<V, K>& a, const Dual<V, K>& b) {
  Dual<V, K> r;
  r.v = a.v - b.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <class V, int K>
inline Dual<V, K> operator-(const Dual<V, K>& a) {
  Dual<V, K> r;
  r.v = -a.v;
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <class V, int K>
inline Dual<V, K> operator*(const Dual<V, K>& a, const Dual<V, K>& b) {
  Dual<V, K> r;
  r.v = a.v * b.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <class V, int K>
inline Dual<V, K> operator/(const Dual<V, K>& a, const Dual<V, K>& b) {
  Dual<V, K> r;
  r.v = a.v / b.v;
  for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
  return r;
}
template <class V, int K>
inline Dual<V, K> operator+(const Dual<V, K>& a, double b) { return a + Dual<V, K>(b); }
template <class V, int K>
inline Dual<V, K> operator+(double a, const Dual<V, K>& b) { return Dual<V, K>(a) + b; }
template <class V, int K>
inline Dual<V, K> operator-(const Dual<V, K>& a, double b) { return a - Dual<V, K>(b); }
template <class V, int K>
inline Dual<V, K> operator-(double a, const Dual<V, K>& b) { return Dual<V, K>(a) - b; }
template <class V, int K>
inline Dual<V, K> operator*(const Dual<V, K>& a, double b) { return a * Dual<V, K>(b); }
template <class V, int K>
inline Dual<V, K> operator*(double a, const Dual<V, K>& b) { return Dual<V, K>(a) * b; }
template <class V, int K>
inline Dual<V, K> operator/(const Dual<V, K>& a, double b) { return a / Dual<V, K>(b); }
template <class V, int K>
inline Dual<V, K> operator/(double a, const Dual<V, K>& b) { return Dual<V, K>(a) / b; }
template <class V, int K>
inline Dual<V, K>& operator+=(Dual<V, K>& a, const Dual<V, K>& b) { return a = a + b; }
template <class V, int K>
inline Dual<V, K>& operator-=(Dual<V, K>& a, const Dual<V, K>& b) { return a = a - b; }
template <class V, int K>
inline Dual<V, K>& operator*=(Dual<V, K>& a, const Dual<V, K>& b) { return a = a * b; }
template <class V, int K>
inline Dual<V, K> sqrt(const Dual<V, K>& a) {
  using std::sqrt;
  Dual<V, K> r;
  r.v = sqrt(a.v);
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / (2.0 * r.v);
  return r;
}
template <class V, int K>
inline Dual<V, K> exp(const Dual<V, K>& a) {
  using std::exp;
  Dual<V, K> r;
  r.v = exp(a.v);
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * r.v;
  return r;
}
template <class V, int K>
inline Dual<V, K> sin(const Dual<V, K>& a) {
  using std::cos;
  using std::sin;
  Dual<V, K> r;
  r.v = sin(a.v);
  V c = cos(a.v);
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * c;
  return r;
}
template <class V, int K>
inline Dual<V, K> cos(const Dual<V, K>& a) {
  using std::cos;
  using std::sin;
  Dual<V, K> r;
  r.v = cos(a.v);
  V s = sin(a.v);
  for (int k = 0; k < K; ++k) r.d[k] = -(a.d[k] * s);
  return r;
}
template <class V, int K>
inline Dual<V, K> tanh(const Dual<V, K>& a) {
  using std::tanh;
  Dual<V, K> r;
  r.v = tanh(a.v);
  V s = 1.0 - r.v * r.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * s;
  return r;
}
template <class V, int K>
inline double val(const Dual<V, K>& a) { return val(a.v); }
template <class V, int K>
inline bool operator<(const Dual<V, K>& a, const Dual<V, K>& b) { return val(a) < val(b); }
template <class V, int K>
inline bool operator>(const Dual<V, K>& a, const Dual<V, K>& b) { return val(a) > val(b); }
template <class V, int K>
inline bool operator>=(const Dual<V, K>& a, const Dual<V, K>& b) { return val(a) >= val(b); }

// ---------------------------------------------------------------------------
// An operation-counting double: add, sub, mul, div, sqrt, exp each count one
// (an FMA is a mul and an add, i.e. two) -- the weights the roofline uses.
struct Cnt {
  double v = 0.0;
  static inline uint64_t ops = 0;
  Cnt() = default;
  Cnt(double x) : v(x) {}
};
inline Cnt operator+(Cnt a, Cnt b) { ++Cnt::ops; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++Cnt::ops; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++Cnt::ops; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++Cnt::ops; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt operator+(Cnt a, double b) { return a + Cnt(b); }
inline Cnt operator+(double a, Cnt b) { return Cnt(a) + b; }
inline Cnt operator-(Cnt a, double b) { return a - Cnt(b); }
inline Cnt operator-(double a, Cnt b) { return Cnt(a) - b; }
inline Cnt operator*(Cnt a, double b) { return a * Cnt(b); }
inline Cnt operator*(double a, Cnt b) { return Cnt(a) * b; }
inline Cnt operator/(Cnt a, double b) { return a / Cnt(b); }
inline Cnt operator/(double a, Cnt b) { return Cnt(a) / b; }
inline Cnt& operator+=(Cnt& a, Cnt b) { return a = a + b; }
inline Cnt& operator-=(Cnt& a, Cnt b) { return a = a - b; }
inline Cnt sqrt(Cnt a) { ++Cnt::ops; return Cnt(std::sqrt(a.v)); }
inline Cnt exp(Cnt a) { ++Cnt::ops; return Cnt(std::exp(a.v)); }
inline double val(Cnt a) { return a.v; }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline bool operator>=(Cnt a, Cnt b) { return a.v >= b.v; }

using std::exp;
using std::sqrt;

// ---------------------------------------------------------------------------
// Packing (pointwise.jl:83-112): packed order (tt,tx,ty,tz,xx,xy,xz,yy,yz,zz),
// the column-major lower triangle; 0-based indices here.
inline int pairindex(int a, int b) {
  int i = a > b ? a : b, j = a > b ? b : a;
  return j * 4 - (j * (j - 1)) / 2 + (i - j);
}
inline void unpair(int n, int& a, int& b) {  // pointwise.jl:569-574 (0-based)
  if (n <= 3) { a = n; b = 0; return; }
  if (n <= 6) { a = n - 3; b = 1; return; }
  if (n <= 8) { a = n - 5; b = 2; return; }
  a = 3; b = 3;
}
template <class T> using M4 = std::array<std::array<T, 4>, 4>;
template <class T> using V10 = std::array<T, NC>;

template <class T>
inline M4<T> sym4(const V10<T>& v) {
  M4<T> m;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) m[a][b] = v[pairindex(a, b)];
  return m;
}
template <class T>
inline V10<T> pack10(const M4<T>& m) {
  V10<T> v;
  for (int n = 0; n < NC; ++n) {
    int a, b;
    unpair(n, a, b);
    v[n] = m[a][b];
  }
  return v;
}
inline double eta(int a, int b) { return a != b ? 0.0 : (a == 0 ? -1.0 : 1.0); }

template <class T>
inline M4<T> matmul(const M4<T>& A, const M4<T>& B) {
  M4<T> C;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      T s = A[i][0] * B[0][j];
      for (int k = 1; k < 4; ++k) s = s + A[i][k] * B[k][j];
      C[i][j] = s;
    }
  return C;
}
// 4x4 determinant by cofactor expansion along the first row.
template <class T>
inline T det3(T a, T b, T c, T d, T e, T f, T g, T h, T i) {
  return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}
template <class T>
inline T det4(const M4<T>& m) {
  T c0 = det3(m[1][1], m[1][2], m[1][3], m[2][1], m[2][2], m[2][3], m[3][1], m[3][2], m[3][3]);
  T c1 = det3(m[1][0], m[1][2], m[1][3], m[2][0], m[2][2], m[2][3], m[3][0], m[3][2], m[3][3]);
  T c2 = det3(m[1][0], m[1][1], m[1][3], m[2][0], m[2][1], m[2][3], m[3][0], m[3][1], m[3][3]);
  T c3 = det3(m[1][0], m[1][1], m[1][2], m[2][0], m[2][1], m[2][2], m[3][0], m[3][1], m[3][2]);
  return m[0][0] * c0 - m[0][1] * c1 + m[0][2] * c2 - m[0][3] * c3;
}
// 4x4 inverse by the adjugate.
template <class T>
inline M4<T> inv4(const M4<T>& m) {
  M4<T> adj;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      int r[3], c[3], nr = 0, ncol = 0;
      for (int k = 0; k < 4; ++k) {
        if (k != j) r[nr++] = k;  // minor of element (j, i): transpose
        if (k != i) c[ncol++] = k;
      }
      T mn = det3(m[r[0]][c[0]], m[r[0]][c[1]], m[r[0]][c[2]], m[r[1]][c[0]], m[r[1]][c[1]],
                  m[r[1]][c[2]], m[r[2]][c[0]], m[r[2]][c[1]], m[r[2]][c[2]]);
      adj[i][j] = ((i + j) % 2 == 0) ? mn : -mn;
    }
  T d = det4(m);
  M4<T> out;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[i][j] = adj[i][j] / d;
  return out;
}

// ---------------------------------------------------------------------------
// Stencil weights in exact rational arithmetic, rounded once
// (stencils.jl:89-157, 170-174, 215-307).
struct Rat {
  __int128 n, d;
};
inline __int128 gcd128(__int128 a, __int128 b) {
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b) { __int128 t = a % b; a = b; b = t; }
  return a;
}
inline Rat rnorm(Rat r) {
  if (r.d < 0) { r.n = -r.n; r.d = -r.d; }
  __int128 g = gcd128(r.n, r.d);
  if (g > 1) { r.n /= g; r.d /= g; }
  return r;
}
inline Rat radd(Rat a, Rat b) { return rnorm({a.n * b.d + b.n * a.d, a.d * b.d}); }
inline Rat rmul(Rat a, Rat b) { return rnorm({a.n * b.n, a.d * b.d}); }

// stencils.jl:89-110: Lagrange weights of the m-th derivative at 0 on nodes.
inline std::vector<Rat> lagrange_derivative_weights(const std::vector<long>& nodes, int m) {
  int n = (int)nodes.size();
  std::vector<Rat> ws(n);
  for (int i = 0; i < n; ++i) {
    std::vector<Rat> c{{1, 1}};
    Rat den{1, 1};
    for (int k = 0; k < n; ++k) {
      if (k == i) continue;
      // multiply polynomial c by (x - x_k)
      std::vector<Rat> out(c.size() + 1, Rat{0, 1});
      Rat a{-nodes[k], 1};
      out[0] = rmul(a, c[0]);
      for (size_t j = 1; j < c.size(); ++j) out[j] = radd(c[j - 1], rmul(a, c[j]));
      out.back() = c.back();
      c = out;
      den = rmul(den, Rat{nodes[i] - nodes[k], 1});
    }
    __int128 fact = 1;
    for (int j = 2; j <= m; ++j) fact *= j;
    ws[i] = m < n ? rnorm({fact * c[m].n * den.d, c[m].d * den.n}) : Rat{0, 1};
  }
  return ws;
}
// stencils.jl:124-134
inline std::vector<Rat> rational_derivative_weights(int q, int m) {
  if (q < 2 || q % 2) throw std::invalid_argument("q must be even and >= 2");
  if (m != 1 && m != 2) throw std::invalid_argument("m must be 1 or 2");
  int r = q / 2;
  std::vector<long> nodes;
  for (int j = -r; j <= r; ++j) nodes.push_back(j);
  return lagrange_derivative_weights(nodes, m);
}
// stencils.jl:148-157: (-1)^{r+1} 2^{-2r} (D+D-)^r on unit samples.
inline std::vector<Rat> rational_dissipation_weights(int r) {
  if (r < 1) throw std::invalid_argument("r must be >= 1");
  int sgn = ((r + 1) % 2 == 0) ? 1 : -1;
  __int128 den = (__int128)1 << (2 * r);
  std::vector<Rat> w;
  for (int j = -r; j <= r; ++j) {
    int k = r - j;
    __int128 binom = 1;
    for (int i = 1; i <= k; ++i) binom = binom * (2 * r - k + i) / i;
    int s = sgn * (((r - j) % 2 == 0) ? 1 : -1);
    w.push_back(rnorm({s * binom, den}));
  }
  return w;
}
inline std::vector<double> to_double(const std::vector<Rat>& ws) {
  std::vector<double> out;
  for (auto& w : ws) out.push_back((double)(long long)w.n / (double)(long long)w.d);
  return out;
}
inline int dissipation_rank(int q) { return q / 2 + 1; }  // stencils.jl:307

// ---------------------------------------------------------------------------
// metric_quantities (pointwise.jl:145-184).
template <class T>
struct MQ {
  M4<T> g4, gu4, guo;
  T alpha, sqrtg;
  T beta[3];
  T gamu[3][3];
};
template <class T>
inline MQ<T> metric_quantities(const V10<T>& hv) {
  MQ<T> o;
  M4<T> h = sym4(hv), eta4;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      eta4[a][b] = T(eta(a, b));
      o.g4[a][b] = T(eta(a, b)) + h[a][b];
    }
  // guo = -(inv(g4) h) eta
  M4<T> gi = inv4(o.g4);
  M4<T> gih = matmul(gi, h);
  M4<T> t = matmul(gih, eta4);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      o.guo[a][b] = -t[a][b];
      o.gu4[a][b] = T(eta(a, b)) + o.guo[a][b];
    }
  T gutt = o.gu4[0][0];
  T q1 = o.guo[0][0];
  o.alpha = T(1.0) / sqrt(T(1.0) - q1);
  for (int i = 0; i < 3; ++i) o.beta[i] = -o.gu4[0][i + 1] / gutt;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      o.gamu[i][j] = o.gu4[i + 1][j + 1] - o.gu4[0][i + 1] * o.gu4[0][j + 1] / gutt;
  M4<T> B = matmul(eta4, h);
  M4<T> B2 = matmul(B, B);
  M4<T> B3 = matmul(B2, B);
  T trB = B[0][0] + B[1][1] + B[2][2] + B[3][3];
  T trB2 = B2[0][0] + B2[1][1] + B2[2][2] + B2[3][3];
  T trB3 = B3[0][0] + B3[1][1] + B3[2][2] + B3[3][3];
  T e1 = trB;
  T e2 = (trB * trB - trB2) / 2.0;
  T e3 = (trB * trB * trB - 3.0 * trB * trB2 + 2.0 * trB3) / 6.0;
  T e4 = det4(B);
  T d1 = -(e1 + e2 + e3 + e4);
  T detgm1 = -d1 - q1 + d1 * q1;
  o.sqrtg = sqrt(T(1.0) + detgm1);
  return o;
}

// ---------------------------------------------------------------------------
// gh_node_source, the reference spelling (pointwise.jl:346-401).
// dg[a][b][c] = d_a g_bc (derivative axis first); Hl[b]; dHl[a][b] = d_a H_b.
template <class T>
inline V10<T> gh_node_source(const MQ<T>& m, const T dg[4][4][4], const T Hl[4],
                             const T dHl[4][4], T g0, T g2) {
  const auto& gu4 = m.gu4;
  const auto& g4 = m.g4;
  T a_mul = m.alpha * m.sqrtg;
  T Clul[4][4][4], Cluu[4][4][4], Glll[4][4][4], G4[4][4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) {
        T s = gu4[b][0] * dg[a][0][c];
        for (int x = 1; x < 4; ++x) s = s + gu4[b][x] * dg[a][x][c];
        Clul[a][b][c] = s;
      }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) {
        T s = gu4[c][0] * Clul[a][b][0];
        for (int y = 1; y < 4; ++y) s = s + gu4[c][y] * Clul[a][b][y];
        Cluu[a][b][c] = s;
      }
  T C2[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      T s = T(0.0);
      bool first = true;
      for (int y = 0; y < 4; ++y)  // Julia's generator order: x fastest
        for (int x = 0; x < 4; ++x) {
          T t = Cluu[a][x][y] * dg[x][y][b];
          s = first ? t : s + t;
          first = false;
        }
      C2[a][b] = s;
    }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) Glll[a][b][c] = (dg[b][a][c] + dg[c][a][b] - dg[a][b][c]) / 2.0;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) {
        T s = gu4[a][0] * Glll[0][b][c];
        for (int x = 1; x < 4; ++x) s = s + gu4[a][x] * Glll[x][b][c];
        G4[a][b][c] = s;
      }
  T G2[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      T s = T(0.0);
      bool first = true;
      for (int y = 0; y < 4; ++y)
        for (int x = 0; x < 4; ++x) {
          T t = G4[x][y][a] * G4[y][x][b];
          s = first ? t : s + t;
          first = false;
        }
      G2[a][b] = s;
    }
  T Gup[4];
  for (int c = 0; c < 4; ++c) {
    T s = T(0.0);
    bool first = true;
    for (int e = 0; e < 4; ++e)
      for (int d = 0; d < 4; ++d) {
        T t = gu4[d][e] * G4[c][d][e];
        s = first ? t : s + t;
        first = false;
      }
    Gup[c] = s;
  }
  T S0[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      T h4 = G4[0][a][b] * Hl[0];
      for (int c = 1; c < 4; ++c) h4 = h4 + G4[c][a][b] * Hl[c];
      T gd = Gup[0] * dg[0][a][b];
      for (int c = 1; c < 4; ++c) gd = gd + Gup[c] * dg[c][a][b];
      S0[a][b] = (C2[a][b] + C2[b][a]) - 2.0 * G2[a][b] - dHl[a][b] - dHl[b][a] + 2.0 * h4 - gd;
    }
  if (val(g0) != 0.0) {
    T Cup[4], Cl[4], tl[4], tu[4];
    for (int c = 0; c < 4; ++c) {
      T s = gu4[c][0] * Hl[0];
      for (int x = 1; x < 4; ++x) s = s + gu4[c][x] * Hl[x];
      Cup[c] = Gup[c] + s;
    }
    for (int c = 0; c < 4; ++c) {
      T s = g4[c][0] * Cup[0];
      for (int x = 1; x < 4; ++x) s = s + g4[c][x] * Cup[x];
      Cl[c] = s;
    }
    tl[0] = -m.alpha;
    tl[1] = tl[2] = tl[3] = T(0.0);
    for (int c = 0; c < 4; ++c) {
      T s = gu4[c][0] * tl[0];
      for (int x = 1; x < 4; ++x) s = s + gu4[c][x] * tl[x];
      tu[c] = s;
    }
    T tC = tu[0] * Cl[0];
    for (int c = 1; c < 4; ++c) tC = tC + tu[c] * Cl[c];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        S0[a][b] = S0[a][b] + g0 * (tl[a] * Cl[b] + tl[b] * Cl[a] - (1.0 + g2) * g4[a][b] * tC);
  }
  V10<T> out;
  for (int n = 0; n < NC; ++n) {
    int a, b;
    unpair(n, a, b);
    out[n] = -a_mul * S0[a][b];
  }
  return out;
}

// dg[a][b][c] = d_a g_bc from the packed d_t h and spatial gradients
// (pointwise.jl:307-313).
template <class T>
inline void dg4(const V10<T>& dth, const V10<T> dh[3], T dg[4][4][4]) {
  for (int b = 0; b < 4; ++b)
    for (int c = 0; c < 4; ++c) {
      int n = pairindex(b, c);
      dg[0][b][c] = dth[n];
      for (int i = 0; i < 3; ++i) dg[i + 1][b][c] = dh[i][n];
    }
}

// ---------------------------------------------------------------------------
// The lean spelling of the source (pointwise.jl:612-741).
template <class T>
inline T slot(const V10<T>& v, int a, int b) { return v[pairindex(a, b)]; }

template <class T>
inline V10<T> lean_c2(const V10<T>& Gp, const V10<T> D[4], int a) {
  const V10<T>& Da = D[a];
  T E[16];
  for (int n = 0; n < 16; ++n) {
    int r = n % 4, c = n / 4;
    T s = slot(Gp, r, 0) * slot(Da, 0, c);
    for (int x = 1; x < 4; ++x) s = s + slot(Gp, r, x) * slot(Da, x, c);
    E[n] = s;
  }
  V10<T> Cuu;
  for (int n = 0; n < NC; ++n) {
    int i, j;
    unpair(n, i, j);
    T s = E[i] * slot(Gp, 0, j);
    for (int y = 1; y < 4; ++y) s = s + E[i + 4 * y] * slot(Gp, y, j);
    Cuu[n] = s;
  }
  T C2[4];
  for (int b = 0; b < 4; ++b) {
    T s = slot(Cuu, 0, 0) * slot(D[0], 0, b);
    for (int m = 1; m < 16; ++m) s = s + slot(Cuu, m % 4, m / 4) * slot(D[m % 4], m / 4, b);
    C2[b] = s;
  }
  V10<T> out;
  for (int n = 0; n < NC; ++n) {
    int i, j;
    unpair(n, i, j);
    T l = (i == a) ? C2[j] : T(0.0);
    T r = (j == a) ? C2[i] : T(0.0);
    out[n] = (i == a && j == a) ? l + r : ((i == a) ? l : ((j == a) ? r : T(0.0)));
  }
  return out;
}
template <class T>
inline V10<T> lean_christoffel(const V10<T>& Gp, const V10<T> D[4], int a) {
  V10<T> out;
  for (int n = 0; n < NC; ++n) {
    int i, j;
    unpair(n, i, j);
    T s = T(0.0);
    for (int x = 0; x < 4; ++x) {
      T t = slot(Gp, a, x) * ((slot(D[i], x, j) + slot(D[j], x, i) - slot(D[x], i, j)) / 2.0);
      s = (x == 0) ? t : s + t;
    }
    out[n] = s;
  }
  return out;
}
template <class T>
inline V10<T> gh_node_source_lean(const MQ<T>& m, const V10<T>& dth, const V10<T> dh[3],
                                  const T Hl[4], const T dHl[4][4], T g0, T g2) {
  V10<T> Gp = pack10(m.gu4), gp = pack10(m.g4);
  V10<T> D[4] = {dth, dh[0], dh[1], dh[2]};
  V10<T> c1 = lean_c2(Gp, D, 0), c2 = lean_c2(Gp, D, 1), c3 = lean_c2(Gp, D, 2),
         c4 = lean_c2(Gp, D, 3);
  V10<T> S;
  for (int n = 0; n < NC; ++n) S[n] = ((c1[n] + c2[n]) + c3[n]) + c4[n];
  V10<T> G[4];
  for (int a = 0; a < 4; ++a) G[a] = lean_christoffel(Gp, D, a);
  T Gup[4];
  for (int c = 0; c < 4; ++c) {
    T s = slot(Gp, 0, 0) * slot(G[c], 0, 0);
    for (int n = 1; n < 16; ++n) s = s + slot(Gp, n % 4, n / 4) * slot(G[c], n % 4, n / 4);
    Gup[c] = s;
  }
  for (int n = 0; n < NC; ++n) {
    int a, b;
    unpair(n, a, b);
    T t2 = slot(G[0], 0, a) * slot(G[0], 0, b);
    for (int mm = 1; mm < 16; ++mm)
      t2 = t2 + slot(G[mm % 4], mm / 4, a) * slot(G[mm / 4], mm % 4, b);
    T t4 = G[0][n] * Hl[0] + G[1][n] * Hl[1] + G[2][n] * Hl[2] + G[3][n] * Hl[3];
    T t5 = Gup[0] * D[0][n] + Gup[1] * D[1][n] + Gup[2] * D[2][n] + Gup[3] * D[3][n];
    S[n] = S[n] + (-2.0 * t2 - (dHl[a][b] + dHl[b][a]) + 2.0 * t4 - t5);
  }
  if (val(g0) != 0.0) {
    T GH[4], Cup[4], Cl[4];
    for (int c = 0; c < 4; ++c)
      GH[c] = ((slot(Gp, c, 0) * Hl[0] + slot(Gp, c, 1) * Hl[1]) + slot(Gp, c, 2) * Hl[2]) +
              slot(Gp, c, 3) * Hl[3];
    for (int c = 0; c < 4; ++c) Cup[c] = Gup[c] + GH[c];
    for (int c = 0; c < 4; ++c)
      Cl[c] = ((slot(gp, c, 0) * Cup[0] + slot(gp, c, 1) * Cup[1]) + slot(gp, c, 2) * Cup[2]) +
              slot(gp, c, 3) * Cup[3];
    T tC = -m.alpha * (((slot(Gp, 0, 0) * Cl[0] + slot(Gp, 0, 1) * Cl[1]) + slot(Gp, 0, 2) * Cl[2]) +
                       slot(Gp, 0, 3) * Cl[3]);
    for (int n = 0; n < NC; ++n) {
      int a, b;
      unpair(n, a, b);
      T tl1 = -m.alpha;
      T t1 = (a == 0) ? tl1 * Cl[b] : T(0.0);
      T t2 = (b == 0) ? tl1 * Cl[a] : T(0.0);
      T d = (a == 0 && b == 0) ? t1 + t2 : (a == 0 ? t1 : (b == 0 ? t2 : T(0.0)));
      S[n] = S[n] + g0 * (d - (1.0 + g2) * gp[n] * tC);
    }
  }
  T s = -(m.alpha * m.sqrtg);
  for (int n = 0; n < NC; ++n) S[n] = s * S[n];
  return S;
}

// ---------------------------------------------------------------------------
// metric_derivatives (pointwise.jl:462-500): dalpha[i], dbeta[i][j],
// dA[i][j][k] = d_i(alpha sqrtg gamma^{jk}).
template <class T>
struct MDer {
  T dalpha[3];
  T dbeta[3][3];
  T dA[3][3][3];
};
template <class T>
inline MDer<T> metric_derivatives(const MQ<T>& m, const V10<T> dh[3]) {
  MDer<T> o;
  T gutt = m.gu4[0][0];
  M4<T> dgl[3], dgu[3];
  for (int i = 0; i < 3; ++i) {
    dgl[i] = sym4(dh[i]);
    M4<T> t = matmul(matmul(m.gu4, dgl[i]), m.gu4);
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) dgu[i][a][b] = -t[a][b];
  }
  T dsq[3];
  T dgamu[3][3][3];
  for (int i = 0; i < 3; ++i) {
    o.dalpha[i] = m.alpha * m.alpha * m.alpha * dgu[i][0][0] / 2.0;
    for (int j = 0; j < 3; ++j)
      o.dbeta[i][j] = -(dgu[i][0][j + 1] + m.beta[j] * dgu[i][0][0]) / gutt;
    T s = T(0.0);
    bool first = true;
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j) {
        T t = m.gamu[j][k] * dgl[i][j + 1][k + 1];
        s = first ? t : s + t;
        first = false;
      }
    dsq[i] = m.sqrtg * s / 2.0;
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        dgamu[i][j][k] = dgu[i][j + 1][k + 1] + m.beta[k] * dgu[i][0][j + 1] +
                         m.beta[j] * dgu[i][0][k + 1] + m.beta[j] * m.beta[k] * dgu[i][0][0];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        o.dA[i][j][k] = o.dalpha[i] * m.sqrtg * m.gamu[j][k] + m.alpha * dsq[i] * m.gamu[j][k] +
                        m.alpha * m.sqrtg * dgamu[i][j][k];
  return o;
}

// metric_divergences (pointwise.jl:754-778).
template <class T>
inline V10<T> _pad;  // unused helper to keep templates uniform
template <class T>
inline void dgu_row(const V10<T>& Gp, const V10<T>& dg, int a, T row[4]) {
  T v[4];
  for (int d = 0; d < 4; ++d)
    v[d] = ((slot(Gp, a, 0) * slot(dg, 0, d) + slot(Gp, a, 1) * slot(dg, 1, d)) +
            slot(Gp, a, 2) * slot(dg, 2, d)) +
           slot(Gp, a, 3) * slot(dg, 3, d);
  for (int b = 0; b < 4; ++b)
    row[b] = -(((v[0] * slot(Gp, 0, b) + v[1] * slot(Gp, 1, b)) + v[2] * slot(Gp, 2, b)) +
               v[3] * slot(Gp, 3, b));
}
template <class T>
inline void metric_divergences(const MQ<T>& m, const V10<T> dh[3], T& divb, T divA[3]) {
  V10<T> Gp = pack10(m.gu4);
  T gutt = m.gu4[0][0];
  T r1[3][4], ri[3][4];
  for (int i = 0; i < 3; ++i) {
    dgu_row(Gp, dh[i], 0, r1[i]);
    dgu_row(Gp, dh[i], i + 1, ri[i]);
  }
  T t0 = -(r1[0][1] + m.beta[0] * r1[0][0]) / gutt;
  T t1 = -(r1[1][2] + m.beta[1] * r1[1][0]) / gutt;
  T t2 = -(r1[2][3] + m.beta[2] * r1[2][0]) / gutt;
  divb = (t0 + t1) + t2;
  T da[3], ds[3];
  for (int i = 0; i < 3; ++i) {
    da[i] = m.alpha * m.alpha * m.alpha * r1[i][0] / 2.0;
    T s = T(0.0);
    for (int mm = 0; mm < 9; ++mm) {
      T t = m.gamu[mm % 3][mm / 3] * slot(dh[i], mm % 3 + 1, mm / 3 + 1);
      s = (mm == 0) ? t : s + t;
    }
    ds[i] = m.sqrtg * s / 2.0;
  }
  for (int k = 0; k < 3; ++k) {
    T terms[3];
    for (int i = 0; i < 3; ++i)
      terms[i] = da[i] * m.sqrtg * m.gamu[i][k] + m.alpha * ds[i] * m.gamu[i][k] +
                 m.alpha * m.sqrtg *
                     (ri[i][k + 1] + m.beta[k] * r1[i][i + 1] + m.beta[i] * r1[i][k + 1] +
                      m.beta[i] * m.beta[k] * r1[i][0]);
    divA[k] = (terms[0] + terms[1]) + terms[2];
  }
}

// ---------------------------------------------------------------------------
// gh_node_rhs_expanded (pointwise.jl:818-850): the one-expression reference
// of the evolved system. ddh in (xx,xy,xz,yy,yz,zz) order.
template <class T>
inline void gh_node_rhs_expanded(const V10<T>& h, const V10<T>& P, const V10<T> dh[3],
                                 const V10<T> dP[3], const V10<T> ddh[6], const T Hl[4],
                                 const T dHl[4][4], T g0, T g2, V10<T>& dth, V10<T>& dtP) {
  MQ<T> m = metric_quantities(h);
  T a_div = m.alpha / m.sqrtg, a_mul = m.alpha * m.sqrtg;
  T A[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = a_mul * m.gamu[i][j];
  MDer<T> md = metric_derivatives(m, dh);
  T divb = md.dbeta[0][0] + md.dbeta[1][1] + md.dbeta[2][2];
  T divA[3];
  for (int j = 0; j < 3; ++j) divA[j] = md.dA[0][0][j] + md.dA[1][1][j] + md.dA[2][2][j];
  for (int n = 0; n < NC; ++n)
    dth[n] = m.beta[0] * dh[0][n] + m.beta[1] * dh[1][n] + m.beta[2] * dh[2][n] + a_div * P[n];
  for (int n = 0; n < NC; ++n)
    dtP[n] = m.beta[0] * dP[0][n] + m.beta[1] * dP[1][n] + m.beta[2] * dP[2][n] + divb * P[n] +
             A[0][0] * ddh[0][n] + 2.0 * A[0][1] * ddh[1][n] + 2.0 * A[0][2] * ddh[2][n] +
             A[1][1] * ddh[3][n] + 2.0 * A[1][2] * ddh[4][n] + A[2][2] * ddh[5][n] +
             divA[0] * dh[0][n] + divA[1] * dh[1][n] + divA[2] * dh[2][n];
  T dg[4][4][4];
  dg4(dth, dh, dg);
  V10<T> ms = gh_node_source(m, dg, Hl, dHl, g0, g2);
  for (int n = 0; n < NC; ++n) dtP[n] = dtP[n] + ms[n];
}

// gh_node_rhs, GHSO2's flux form (pointwise.jl:229-301): dtg, fluxes F^i, msrc.
template <class T>
inline void gh_node_rhs_flux(const V10<T>& h, const V10<T>& P, const V10<T> dh[3], const T Hl[4],
                             const T dHl[4][4], T g0, T g2, V10<T>& dtg, V10<T> F[3],
                             V10<T>& msrc) {
  MQ<T> m = metric_quantities(h);
  T a_div = m.alpha / m.sqrtg, a_mul = m.alpha * m.sqrtg;
  for (int n = 0; n < NC; ++n)
    dtg[n] = m.beta[0] * dh[0][n] + m.beta[1] * dh[1][n] + m.beta[2] * dh[2][n] + a_div * P[n];
  T dg[4][4][4];
  dg4(dtg, dh, dg);
  msrc = gh_node_source(m, dg, Hl, dHl, g0, g2);
  for (int i = 0; i < 3; ++i)
    for (int n = 0; n < NC; ++n)
      F[i][n] = m.beta[i] * P[n] + a_mul * (m.gamu[i][0] * dh[0][n] + m.gamu[i][1] * dh[1][n] +
                                            m.gamu[i][2] * dh[2][n]);
}

// matter_source (pointwise.jl:667-698).
template <class T>
inline V10<T> matter_source(const MQ<T>& m, T rho, const T S[3], const T Sij[3][3]) {
  const double pi = 3.14159265358979323846;
  double c8 = 8 * pi, c16 = 16 * pi;
  const auto& b = m.beta;
  T Sb1 = Sij[0][0] * b[0] + Sij[1][0] * b[1] + Sij[2][0] * b[2];
  T Sb2 = Sij[1][0] * b[0] + Sij[1][1] * b[1] + Sij[2][1] * b[2];
  T Sb3 = Sij[2][0] * b[0] + Sij[2][1] * b[1] + Sij[2][2] * b[2];
  T bS = b[0] * S[0] + b[1] * S[1] + b[2] * S[2];
  T Ttt = m.alpha * m.alpha * rho - 2.0 * (m.alpha * bS) + (b[0] * Sb1 + b[1] * Sb2 + b[2] * Sb3);
  T Tt[3] = {Sb1 - m.alpha * S[0], Sb2 - m.alpha * S[1], Sb3 - m.alpha * S[2]};
  T trS = m.gamu[0][0] * Sij[0][0] + m.gamu[1][1] * Sij[1][1] + m.gamu[2][2] * Sij[2][2] +
          2.0 * (m.gamu[1][0] * Sij[1][0] + m.gamu[2][0] * Sij[2][0] + m.gamu[2][1] * Sij[2][1]);
  T tr8 = c8 * (trS - rho);
  T a = m.alpha * m.sqrtg;
  V10<T> o;
  o[0] = a * (c16 * Ttt - m.g4[0][0] * tr8);
  for (int i = 0; i < 3; ++i) o[1 + i] = a * (c16 * Tt[i] - m.g4[i + 1][0] * tr8);
  o[4] = a * (c16 * Sij[0][0] - m.g4[1][1] * tr8);
  o[5] = a * (c16 * Sij[1][0] - m.g4[2][1] * tr8);
  o[6] = a * (c16 * Sij[2][0] - m.g4[3][1] * tr8);
  o[7] = a * (c16 * Sij[1][1] - m.g4[2][2] * tr8);
  o[8] = a * (c16 * Sij[2][1] - m.g4[3][2] * tr8);
  o[9] = a * (c16 * Sij[2][2] - m.g4[3][3] * tr8);
  return o;
}

// ---------------------------------------------------------------------------
// The algebraic Kerr-Schild gauge source (gauge.jl:257-326).
template <class T>
inline void ks_algebraic_source(double M, const double S[4], const double u[4], const V10<T>& h,
                                const V10<T>& dth, const V10<T> dh[3], T Hl[4], T dHl[4][4]) {
  M4<T> h4 = sym4(h);
  T w[4];
  for (int a = 0; a < 4; ++a) w[a] = h4[a][0] * u[0] + h4[a][1] * u[1] + h4[a][2] * u[2] + h4[a][3] * u[3];
  T K = u[0] * w[0] + u[1] * w[1] + u[2] * w[2] + u[3] * w[3];
  T sw = S[0] * w[0] + S[1] * w[1] + S[2] * w[2] + S[3] * w[3];
  T rad = M * M - sw * sw;
  T root = val(rad) > 0 ? sqrt(rad) : T(0.0);
  T D = M + root;
  for (int a = 0; a < 4; ++a) Hl[a] = (-K / D) * w[a];
  const V10<T>* Dv[4] = {&dth, &dh[0], &dh[1], &dh[2]};
  T dw[4][4];
  for (int c = 0; c < 4; ++c) {
    M4<T> d4 = sym4(*Dv[c]);
    for (int a = 0; a < 4; ++a)
      dw[c][a] = d4[a][0] * u[0] + d4[a][1] * u[1] + d4[a][2] * u[2] + d4[a][3] * u[3];
  }
  for (int c = 0; c < 4; ++c) {
    T dK = u[0] * dw[c][0] + u[1] * dw[c][1] + u[2] * dw[c][2] + u[3] * dw[c][3];
    T Sdw = S[0] * dw[c][0] + S[1] * dw[c][1] + S[2] * dw[c][2] + S[3] * dw[c][3];
    T dD = val(root) > 0 ? -sw * Sdw / root : T(0.0);
    for (int a = 0; a < 4; ++a) dHl[c][a] = -(dK * w[a] + K * dw[c][a]) / D + K * w[a] * dD / (D * D);
  }
}

// ---------------------------------------------------------------------------
// Analytic backgrounds (SpacetimeMetrics, restated from their closed forms).
enum Background { BG_MINKOWSKI = 0, BG_GAUGE_WAVE = 1, BG_KERR_SCHILD = 2, BG_SHIFTED = 3, BG_HARMONIC = 4 };

struct BgParams {
  int kind = BG_MINKOWSKI;
  double M = 1.0, a = 0.0;        // Kerr-Schild
  double c[3] = {0, 0, 0};        // translation of the hole
  double A = 0.05, d = 1.0;       // gauge wave amplitude and wavelength / shifted A
  double w = 2.0;                 // shifted Minkowski width
  double u[3] = {0, 0, 0};        // boost velocity: the hole moves at -u
};

template <class S>
inline void bg_metric(const BgParams& p, S t, S x, S y, S z, S g[4][4]);

// boost(m, u) and translate (SpacetimeMetrics, unvendored; the convention the
// reference states at src/interior.jl:66-72 and src/initialdata.jl:93-96,
// 625-638): m is evaluated at x' = Lambda (t, x - c), Lambda the symmetric
// Lorentz boost of velocity u,
//   Lambda^0_0 = gamma, Lambda^0_i = Lambda^i_0 = gamma u_i,
//   Lambda^i_j = delta_ij + (gamma - 1) u_i u_j / u^2,
// and pulled back, g_ab(x) = Lambda^c_a Lambda^d_b m_cd(x'): the rest frame's
// origin x' = 0 is x = c - u t, so the hole moves at -u.
inline void lorentz_boost(const double u[3], double L[4][4]) {
  const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double gam = 1.0 / std::sqrt(1.0 - u2);
  const double f = (gam - 1.0) / u2;
  L[0][0] = gam;
  for (int i = 0; i < 3; ++i) {
    L[0][i + 1] = gam * u[i];
    L[i + 1][0] = gam * u[i];
    for (int j = 0; j < 3; ++j) L[i + 1][j + 1] = (i == j ? 1.0 : 0.0) + f * u[i] * u[j];
  }
}
template <class S>
inline void boosted_metric(const BgParams& p, S t, S x, S y, S z, S g[4][4]) {
  double L[4][4];
  lorentz_boost(p.u, L);
  S xr[4] = {t, x - p.c[0], y - p.c[1], z - p.c[2]};
  S xp[4];
  for (int c = 0; c < 4; ++c) xp[c] = L[c][0] * xr[0] + L[c][1] * xr[1] + L[c][2] * xr[2] + L[c][3] * xr[3];
  BgParams q = p;
  for (int i = 0; i < 3; ++i) q.u[i] = q.c[i] = 0.0;
  S gp[4][4];
  bg_metric(q, xp[0], xp[1], xp[2], xp[3], gp);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      S sum(0.0);
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d) sum = sum + (L[c][a] * L[d][b]) * gp[c][d];
      g[a][b] = sum;
    }
}
inline bool bg_boosted(const BgParams& p) {
  return (p.kind == BG_KERR_SCHILD || p.kind == BG_HARMONIC) && (p.u[0] != 0.0 || p.u[1] != 0.0 || p.u[2] != 0.0);
}

// The Kerr metric in harmonic coordinates (SpacetimeMetrics' Harmonic(M, a),
// an unvendored dependency; restated from Cook & Scheel, Phys. Rev. D 56,
// 4775 (1997)), spin along +z, at the hole-relative harmonic position
// (X, Y, Z): the harmonic coordinates are those of ingoing Kerr-Schild with the
// radius lowered by M,
//   X + i Y = (R + i a) e^{i phi} sin(theta),  Z = R cos(theta),  R = r - M,
// phi the ingoing Kerr-Schild angle, and the harmonic time
//   t_H = t_BL + int 2 M r_+ / Delta dr = t_KS - 2 M ln(r - r_-) + const.
// So the Kerr-Schild coordinates of a harmonic point are
//   x + i y = (X + i Y)(r + i a)/(R + i a),  z = r Z / R,  t = t_H + 2 M ln(r - r_-),
// R solving the Kerr-Schild quartic in the harmonic (X, Y, Z), and
// g_H = J^T g_KS J with J the Jacobian of that map (a nested dual pass).
template <class S>
inline void harmonic_kerr_metric(const BgParams& p, S X, S Y, S Z, S g[4][4]) {
  using D3 = Dual<S, 3>;
  const double a = p.a, M = p.M;
  const double rm = M - std::sqrt(M * M - a * a);
  D3 Xd(X, 0), Yd(Y, 0), Zd(Z, 0);
  Xd.d[0] = S(1.0);
  Yd.d[1] = S(1.0);
  Zd.d[2] = S(1.0);
  D3 rho2 = Xd * Xd + Yd * Yd + Zd * Zd;
  D3 hb = (rho2 - a * a) / 2.0;
  D3 R2 = hb + sqrt(hb * hb + (a * a) * Zd * Zd);
  D3 R = sqrt(R2);
  D3 r = R + M;
  D3 den = R2 + a * a;
  D3 c1 = (r * R + a * a) / den, c2 = (-a * M) / den;
  D3 xk = c1 * Xd - c2 * Yd, yk = c2 * Xd + c1 * Yd, zk = r * Zd / R;
  // J[c][b] = d x_KS^c / d x_H^b
  S J[4][4];
  for (int c = 0; c < 4; ++c)
    for (int b = 0; b < 4; ++b) J[c][b] = S(c == 0 && b == 0 ? 1.0 : 0.0);
  const S tfac = (2.0 * M) / (r.v - rm);
  for (int i = 0; i < 3; ++i) {
    J[0][i + 1] = tfac * R.d[i];
    J[1][i + 1] = xk.d[i];
    J[2][i + 1] = yk.d[i];
    J[3][i + 1] = zk.d[i];
  }
  // g_KS at (xk, yk, zk), with the Kerr-Schild radius r = R + M
  S rv = r.v, r2v = rv * rv, zv = zk.v;
  S f = (2.0 * M) * rv * r2v / (r2v * r2v + (a * a) * zv * zv);
  S dl = r2v + a * a;
  S l[4] = {S(1.0), (rv * xk.v + a * yk.v) / dl, (rv * yk.v - a * xk.v) / dl, zv / rv};
  S gk[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) gk[i][j] = S(eta(i, j)) + f * l[i] * l[j];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      S sum(0.0);
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d) sum = sum + J[c][i] * J[d][j] * gk[c][d];
      g[i][j] = sum;
    }
}

// The metric g_ab at (t, x, y, z), generic in the scalar (dual numbers give
// dmetric).
template <class S>
inline void bg_metric(const BgParams& p, S t, S x, S y, S z, S g[4][4]) {
  if (bg_boosted(p)) {
    boosted_metric(p, t, x, y, z, g);
    return;
  }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) g[a][b] = S(eta(a, b));
  if (p.kind == BG_MINKOWSKI) return;
  if (p.kind == BG_GAUGE_WAVE) {
    // ds^2 = -H dt^2 + H dx^2 + dy^2 + dz^2, H = 1 - A sin(2 pi (x - t)/d)
    // (initialdata.jl:897-901).
    const double pi = 3.14159265358979323846;
    S H = S(1.0) - p.A * sin((2.0 * pi / p.d) * (x - t));
    g[0][0] = -H;
    g[1][1] = H;
    return;
  }
  if (p.kind == BG_SHIFTED) {
    // t^ = t + A w tanh(x / w): ds^2 = -(dt + psi' dx)^2 + dx^2 + dy^2 + dz^2,
    // psi' = A sech^2(x / w) (initialdata.jl:917-920).
    S th = tanh(x / p.w);
    S pp = p.A * (S(1.0) - th * th);
    g[0][0] = S(-1.0);
    g[0][1] = g[1][0] = -pp;
    g[1][1] = S(1.0) - pp * pp;
    return;
  }
  if (p.kind == BG_HARMONIC) {
    harmonic_kerr_metric(p, x - p.c[0], y - p.c[1], z - p.c[2], g);
    return;
  }
  // Kerr-Schild, g = eta + f l l with
  //   r^4 - r^2 (rho^2 - a^2) - a^2 z^2 = 0,  f = 2 M r^3 / (r^4 + a^2 z^2),
  //   l = (1, (r x + a y)/(r^2 + a^2), (r y - a x)/(r^2 + a^2), z / r).
  S X = x - p.c[0], Y = y - p.c[1], Z = z - p.c[2];
  double a = p.a, M = p.M;
  S rho2 = X * X + Y * Y + Z * Z;
  S hb = (rho2 - a * a) / 2.0;
  S r2 = hb + sqrt(hb * hb + (a * a) * Z * Z);
  S r = sqrt(r2);
  S f = (2.0 * M) * r * r2 / (r2 * r2 + (a * a) * Z * Z);
  S den = r2 + a * a;
  S l[4] = {S(1.0), (r * X + a * Y) / den, (r * Y - a * X) / den, Z / r};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) g[i][j] = g[i][j] + f * l[i] * l[j];
}

// dmetric: g and dg[a][b][c] = d_c g_ab (SpacetimeMetrics' index order).
inline void dmetric(const BgParams& p, double t, double x, double y, double z, double g[4][4],
                    double dg[4][4][4]) {
  using D4 = Dual<double, 4>;
  D4 X[4];
  double pt[4] = {t, x, y, z};
  for (int k = 0; k < 4; ++k) {
    X[k] = D4(pt[k]);
    X[k].d[k] = 1.0;
  }
  D4 G[4][4];
  bg_metric(p, X[0], X[1], X[2], X[3], G);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      g[a][b] = G[a][b].v;
      for (int c = 0; c < 4; ++c) dg[a][b][c] = G[a][b].d[c];
    }
}

// background_state (initialdata.jl:1674-1694): h, Pi, dh from one dual pass.
inline void background_state(const BgParams& p, double t, const double x[3], V10<double>& h,
                             V10<double>& P, V10<double> dh[3], V10<double>* dtg_out = nullptr) {
  double g[4][4], dg[4][4][4];
  dmetric(p, t, x[0], x[1], x[2], g, dg);
  M4<double> gm;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) gm[a][b] = g[a][b] - eta(a, b);
  h = pack10(gm);
  M4<double> dt4, di4[3];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      dt4[a][b] = dg[a][b][0];
      for (int i = 0; i < 3; ++i) di4[i][a][b] = dg[a][b][i + 1];
    }
  V10<double> dtg = pack10(dt4);
  for (int i = 0; i < 3; ++i) dh[i] = pack10(di4[i]);
  MQ<double> m = metric_quantities(h);
  double s = m.sqrtg / m.alpha;
  for (int n = 0; n < NC; ++n)
    P[n] = s * (dtg[n] - m.beta[0] * dh[0][n] - m.beta[1] * dh[1][n] - m.beta[2] * dh[2][n]);
  if (dtg_out) *dtg_out = dtg;
}

// The sampled gauge source H_b = -Gamma_b and dHl[a][b] = d_a H_b of a
// background (gauge.jl:156-212, SpacetimeMetrics' gauge_source_grad), by a
// nested dual pass for the second derivatives.
inline void gauge_source_grad(const BgParams& p, double t, double x, double y, double z,
                              double Hl[4], double dHl[4][4]) {
  using D1 = Dual<double, 4>;
  using D2 = Dual<D1, 4>;
  D2 X[4];
  double pt[4] = {t, x, y, z};
  for (int k = 0; k < 4; ++k) {
    D1 inner(pt[k]);
    inner.d[k] = 1.0;          // inner partials: d/dx^k
    X[k] = D2(inner, 0);
    X[k].d[k] = D1(1.0);       // outer partials: d/dx^k
  }
  D2 G[4][4];
  bg_metric(p, X[0], X[1], X[2], X[3], G);
  // Work in D1 (value + first derivatives along the outer direction): for each
  // outer direction c, the metric's first derivatives carry their own
  // derivatives in D1.
  // g_ab (as D1 in the outer variable) and d_e g_ab (as D1):
  D1 g[4][4], dg[4][4][4];  // dg[e][a][b] = d_e g_ab, each carrying d/dx^c
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      // G.v is D1 with .v = g, .d[k] = d_k g (inner); G.d[c] is D1 with .v = d_c g.
      // Build D1 objects whose value is the quantity and partials the outer derivative.
      D1 gv;
      gv.v = G[a][b].v.v;
      for (int c = 0; c < 4; ++c) gv.d[c] = G[a][b].d[c].v;
      g[a][b] = gv;
      for (int e = 0; e < 4; ++e) {
        D1 de;
        de.v = G[a][b].v.d[e];
        for (int c = 0; c < 4; ++c) de.d[c] = G[a][b].d[c].d[e];
        dg[e][a][b] = de;
      }
    }
  M4<D1> gm;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) gm[a][b] = g[a][b];
  M4<D1> gu = inv4(gm);
  // Gamma^c = g^{ab} Gamma^c_ab, Gamma^c_ab = g^{cd} (d_a g_db + d_b g_da - d_d g_ab)/2
  D1 Gup[4];
  for (int c = 0; c < 4; ++c) {
    D1 s(0.0);
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        D1 Gc(0.0);
        for (int d = 0; d < 4; ++d)
          Gc = Gc + gu[c][d] * (dg[a][d][b] + dg[b][d][a] - dg[d][a][b]) / 2.0;
        s = s + gu[a][b] * Gc;
      }
    Gup[c] = s;
  }
  for (int b = 0; b < 4; ++b) {
    D1 s(0.0);
    for (int c = 0; c < 4; ++c) s = s + gm[b][c] * Gup[c];
    D1 H = -s;  // H_b = -Gamma_b
    Hl[b] = H.v;
    for (int a = 0; a < 4; ++a) dHl[a][b] = H.d[a];
  }
}

// ---------------------------------------------------------------------------
// Profiles (gauge.jl:1094-1156, interior.jl:117-120, 567-590, 615-625).
inline double smoothstep(double s) {
  double u = std::min(std::max(s, 0.0), 1.0);
  double v = u * u * u * (10.0 + u * (6.0 * u - 15.0));
  return std::min(std::max(v, 0.0), 1.0);
}
inline void layer_profiles(double r, double r0, double r1, double rho_max, double w_ramp,
                           double rho_ramp, double& w, double& rho) {
  double s = (r - r0) / (r1 - r0);
  w = smoothstep(s / w_ramp);
  rho = rho_max * smoothstep((1.0 - s) / rho_ramp);
}

enum Interior { INT_NONE = 0, INT_DAMPED = 1, INT_FROZEN = 2, INT_PASTED = 3, INT_FITTED = 4 };
enum Gauge { GAUGE_NONE = 0, GAUGE_SAMPLED = 1, GAUGE_KS = 2 };
enum Damping { DAMP_CONSTANT = 0, DAMP_GAUSSIAN = 1 };
enum Dissipation { DISS_CONSTANT = 0, DISS_HORIZON = 1 };

struct Case {
  BgParams bg;
  int gauge = GAUGE_NONE;
  double ks_S[4] = {0, 0, 0, 0}, ks_u[4] = {1, 0, 0, 0};
  int damping = DAMP_CONSTANT;
  double g0 = 0, g0_near = 1, g0_far = 0.1, g0_width = 3, g2 = 0;
  double eps_ko = 0;  // the number, or HorizonDissipation's eps_out
  int diss = DISS_CONSTANT;
  double eps_in = 0, eps_r1 = 0, eps_rh = 0;  // HorizonDissipation (gauge.jl:1221-1227)
  int interior = INT_NONE;
  double center[3] = {0, 0, 0};  // the hole's center at t = 0
  double vel[3] = {0, 0, 0};     // and its velocity: c(t) = c0 + v t
  double r0 = 0, r1 = 0, rho_max = 0, w_ramp = 0.5, rho_ramp = 0.5;
  double delta = 0;              // layer_rate_step: cbrt(eps) r_0
  // FittedInterior's tracked geometry (interior.jl:1752-1990)
  bool shaped = false;
  int lmax = 0;
  double shape[81] = {0};
  double r_in = 0, r_out = 0, offset = 0, thickness = 0;
};

// shape_series (interior.jl:1294-1329): sum_lm a_lm y~_lm(n) with the fully
// normalised Legendre recurrences in z and (n_x + i n_y)^m for the azimuth
// (Condon-Shortley phase); _shape_term (:1266-1274) for one term.
inline double shape_term(const double* shape, int l, int m, double q, double C, double S, double s2) {
  if (m == 0) return shape[l * l + l] * q;
  const double ac = shape[l * l + l + m], as = shape[l * l + l - m];
  return s2 * q * (ac * C - as * S);
}
inline double shape_series(const double* shape, int lmax, const double n[3]) {
  const double x = n[0], y = n[1], z = n[2];
  const double s2 = std::sqrt(2.0);
  double acc = 0.0;
  double qmm = 1.0 / std::sqrt(4.0 * M_PI);
  double C = 1.0, S = 0.0;
  for (int m = 0; m <= lmax; ++m) {
    if (m > 0) {
      qmm = -std::sqrt((double)(2 * m + 1) / (double)(2 * m)) * qmm;
      const double Cn = C * x - S * y, Sn = C * y + S * x;
      C = Cn;
      S = Sn;
    }
    acc += shape_term(shape, m, m, qmm, C, S, s2);
    if (m + 1 <= lmax) {
      double qa = qmm;
      double qb = std::sqrt((double)(2 * m + 3)) * z * qmm;
      acc += shape_term(shape, m + 1, m, qb, C, S, s2);
      for (int l = m + 2; l <= lmax; ++l) {
        const double al = std::sqrt((double)(4 * l * l - 1) / (double)(l * l - m * m));
        const double be = std::sqrt((double)((l - 1) * (l - 1) - m * m) / (double)(4 * (l - 1) * (l - 1) - 1));
        const double qc = al * (z * qb - be * qa);
        acc += shape_term(shape, l, m, qc, C, S, s2);
        qa = qb;
        qb = qc;
      }
    }
  }
  return acc;
}
// shape_radius (interior.jl:1846-1853): the series clamped into [r_in, r_out].
inline double shape_radius(const Case& c, const double n[3]) {
  return std::min(std::max(shape_series(c.shape, c.lmax, n), c.r_in), c.r_out);
}
// interior_point (interior.jl:341, 1868-1886): the distance from c(t) and the
// layer's radii along the ray -- the sphere's r_0, r_1, or the tracked
// surface's r_1 = r_h(n) - offset, r_0 = r_1 - thickness, with the two exact
// fast paths outside r_out - offset and inside r_in - offset - thickness.
struct IGeo {
  double r, r0, r1;
};
inline IGeo interior_geo(const Case& c, double t, const double x[3]) {
  double cc[3];
  for (int i = 0; i < 3; ++i) cc[i] = c.center[i] + t * c.vel[i];
  const double d1 = x[0] - cc[0], d2 = x[1] - cc[1], d3 = x[2] - cc[2];
  const double r = std::sqrt(d1 * d1 + d2 * d2 + d3 * d3);
  if (!c.shaped) return {r, c.r0, c.r1};
  const double r1o = c.r_out - c.offset;
  if (r >= r1o) return {r, r1o - c.thickness, r1o};
  const double r1i = c.r_in - c.offset, r0i = r1i - c.thickness;
  if (r < r0i) return {r, r0i, r1i};
  double n[3] = {0.0, 0.0, 1.0};
  if (r != 0.0) {
    n[0] = d1 / r;
    n[1] = d2 / r;
    n[2] = d3 / r;
  }
  const double r1 = shape_radius(c, n) - c.offset;
  return {r, r1 - c.thickness, r1};
}

// center_at (interior.jl:60-93): c(t) = c0 + v t.
inline void center_at(const Case& c, double t, double out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = c.center[i] + t * c.vel[i];
}
inline bool hole_moves(const Case& c) { return c.vel[0] != 0.0 || c.vel[1] != 0.0 || c.vel[2] != 0.0; }
// core_follows (interior.jl:520-523): a moving :damped sphere.
inline bool core_follows(const Case& c) { return c.interior == INT_DAMPED && hole_moves(c); }
// core_rate (interior.jl:480-484): max(rho_max, 12 |v| / r_0).
inline double core_rate(const Case& c) {
  const double s = std::sqrt(c.vel[0] * c.vel[0] + c.vel[1] * c.vel[1] + c.vel[2] * c.vel[2]);
  return std::max(c.rho_max, 12.0 * s / c.r0);
}

inline double damping_rate(const Case& c, double t, const double x[3]) {
  if (c.damping == DAMP_CONSTANT) return c.g0;
  double cc[3];
  center_at(c, t, cc);
  double d1 = x[0] - cc[0], d2 = x[1] - cc[1], d3 = x[2] - cc[2];
  double r2 = d1 * d1 + d2 * d2 + d3 * d3;
  return c.g0_far + (c.g0_near - c.g0_far) * std::exp(-r2 / (2.0 * c.g0_width * c.g0_width));
}
// dissipation_rate (gauge.jl:1180, 1254-1264): the number itself, or the
// HorizonDissipation profile about the (static) center: eps_out at and
// outside r_h, eps_in at and inside r_1, each through a branch, and the
// quintic smoothstep of (r_h - r)/(r_h - r_1) between.
inline double dissipation_rate(const Case& c, double t, const double x[3]) {
  if (c.diss == DISS_CONSTANT) return c.eps_ko;
  double cc[3];
  center_at(c, t, cc);
  double d1 = x[0] - cc[0], d2 = x[1] - cc[1], d3 = x[2] - cc[2];
  double r = std::sqrt(d1 * d1 + d2 * d2 + d3 * d3);
  if (r >= c.eps_rh) return c.eps_ko;
  if (r <= c.eps_r1) return c.eps_in;
  double s = (c.eps_rh - r) / (c.eps_rh - c.eps_r1);
  return c.eps_ko + (c.eps_in - c.eps_ko) * smoothstep(s);
}
// has_dissipation (gauge.jl:1190, 1266).
inline bool has_dissipation(const Case& c) {
  return c.diss == DISS_CONSTANT ? c.eps_ko != 0.0 : !(c.eps_ko == 0.0 && c.eps_in == 0.0);
}
inline double interior_radius(const Case& c, double t, const double x[3]) {
  double cc[3];
  center_at(c, t, cc);
  double d1 = x[0] - cc[0], d2 = x[1] - cc[1], d3 = x[2] - cc[2];
  return std::sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}
// core_position (interior.jl:615-625) about c(t); on the tracked geometry
// (:1950-1965) the core surface's own point on the ray, r_0(n) =
// (r_h(n) - offset) - thickness, the series skipped outside its bounding sphere.
inline void core_position(const Case& c, double t, const double x[3], double out[3]) {
  if (c.interior == INT_NONE) { out[0] = x[0]; out[1] = x[1]; out[2] = x[2]; return; }
  double cc[3];
  center_at(c, t, cc);
  double d[3] = {x[0] - cc[0], x[1] - cc[1], x[2] - cc[2]};
  double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double rc = c.r0;
  if (c.shaped) {
    if (r >= (c.r_out - c.offset) - c.thickness) { out[0] = x[0]; out[1] = x[1]; out[2] = x[2]; return; }
    double n[3] = {0, 0, 1};
    if (r != 0) { n[0] = d[0] / r; n[1] = d[1] / r; n[2] = d[2] / r; }
    rc = (shape_radius(c, n) - c.offset) - c.thickness;
  }
  if (r >= rc) { out[0] = x[0]; out[1] = x[1]; out[2] = x[2]; return; }
  double n[3] = {0, 0, 1};
  if (r != 0) { n[0] = d[0] / r; n[1] = d[1] / r; n[2] = d[2] / r; }
  for (int i = 0; i < 3; ++i) out[i] = cc[i] + rc * n[i];
}
// case_state_tuple (initialdata.jl:1795-1796).
inline void case_state(const Case& c, double t, const double x[3], double out[20]) {
  double xc[3];
  core_position(c, t, x, xc);
  V10<double> h, P, dh[3];
  background_state(c.bg, t, xc, h, P, dh);
  for (int n = 0; n < NC; ++n) { out[n] = h[n]; out[NC + n] = P[n]; }
}

// background_state_rate (initialdata.jl:1697-1729): the state at (t, x), its
// d_t h from the same dual pass, and d_t Pi by the centered difference of Pi
// at t +- delta divided by the difference of the two times as rounded.
inline void background_state_rate(const BgParams& p, double t, const double x[3], double delta,
                                  V10<double>& h, V10<double>& P, V10<double>& dth, V10<double>& dtP) {
  V10<double> dh[3], Pp, Pm, hx;
  background_state(p, t, x, h, P, dh, &dth);
  const double tp = t + delta, tm = t - delta;
  background_state(p, tp, x, hx, Pp, dh);
  background_state(p, tm, x, hx, Pm, dh);
  for (int n = 0; n < NC; ++n) dtP[n] = (Pp[n] - Pm[n]) / (tp - tm);
}
// core_rate_weight (interior.jl:539-548): smoothstep(4 r / r_0 - 1).
inline double core_rate_weight(const Case& c, double r) { return smoothstep(4.0 * r / c.r0 - 1.0); }
// core_target_rate (initialdata.jl:1731-1764): the rate of the core rule's
// target t -> u_exact(t, c(t) + r_0 n(t)) at x, by the centered difference at
// t +- delta, weighted by core_rate_weight; zero (no pass) where that is zero.
inline void core_target_rate(const Case& c, double t, const double x[3], V10<double>& dth, V10<double>& dtP) {
  const double s = core_rate_weight(c, interior_radius(c, t, x));
  for (int n = 0; n < NC; ++n) dth[n] = dtP[n] = 0.0;
  if (s == 0.0) return;
  const double tp = t + c.delta, tm = t - c.delta;
  double xp[3], xm[3];
  core_position(c, tp, x, xp);
  core_position(c, tm, x, xm);
  V10<double> hp, Pp, hm, Pm, dh[3];
  background_state(c.bg, tp, xp, hp, Pp, dh);
  background_state(c.bg, tm, xm, hm, Pm, dh);
  const double f = s / (tp - tm);
  for (int n = 0; n < NC; ++n) {
    dth[n] = f * (hp[n] - hm[n]);
    dtP[n] = f * (Pp[n] - Pm[n]);
  }
}
// _adm_split (interior.jl:505-515): A = alpha^2 = (1 - h_tt) + beta^i beta_i,
// beta_i = h_ti, beta^i = gamma^-1 beta_i, gamma_ij = delta_ij + h_ij.
inline void inv3(const double m[3][3], double o[3][3]) {
  const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1], c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2],
               c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
  o[0][0] = c00 / det;
  o[1][0] = c01 / det;
  o[2][0] = c02 / det;
  o[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
  o[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
  o[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
  o[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
  o[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
  o[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
}
inline void mv3(const double m[3][3], const double v[3], double o[3]) {
  for (int i = 0; i < 3; ++i) o[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
}
inline void adm_split_rate(const V10<double>& h, double& A, double bu[3], double bl[3], double g[3][3]) {
  g[0][0] = 1.0 + h[4]; g[0][1] = h[5]; g[0][2] = h[6];
  g[1][0] = h[5]; g[1][1] = 1.0 + h[7]; g[1][2] = h[8];
  g[2][0] = h[6]; g[2][1] = h[8]; g[2][2] = 1.0 + h[9];
  bl[0] = h[1]; bl[1] = h[2]; bl[2] = h[3];
  double gi[3][3];
  inv3(g, gi);
  mv3(gi, bl, bu);
  A = (1.0 - h[0]) + (bu[0] * bl[0] + bu[1] * bl[1] + bu[2] * bl[2]);
}
// core_relaxation_rate (interior.jl:450-503): the moving core's metric relaxed
// toward h_t at rate rho in the ADM variables (alpha^2, beta^i, gamma_ij),
// plus the target's own ADM rate read off d_t h_t at h_t, mapped back to h by
// the chain rule of g_tt = -alpha^2 + gamma_ij beta^i beta^j, g_ti = gamma_ij beta^j.
inline V10<double> core_relaxation_rate(const V10<double>& h, const V10<double>& ht, double rho,
                                        const V10<double>& dht) {
  double A, bu[3], bl[3], g[3][3], At, but[3], blt[3], gt[3][3];
  adm_split_rate(h, A, bu, bl, g);
  adm_split_rate(ht, At, but, blt, gt);
  const double dgt[3][3] = {{dht[4], dht[5], dht[6]}, {dht[5], dht[7], dht[8]}, {dht[6], dht[8], dht[9]}};
  double gti[3][3], dgbt[3], rhs[3], dbt[3];
  inv3(gt, gti);
  mv3(dgt, but, dgbt);
  for (int i = 0; i < 3; ++i) rhs[i] = dht[1 + i] - dgbt[i];
  mv3(gti, rhs, dbt);
  const double dAt = -dht[0] + (but[0] * dgbt[0] + but[1] * dgbt[1] + but[2] * dgbt[2]) +
                     2.0 * (blt[0] * dbt[0] + blt[1] * dbt[1] + blt[2] * dbt[2]);
  const double dA = dAt - rho * (A - At);
  double db[3];
  for (int i = 0; i < 3; ++i) db[i] = dbt[i] - rho * (bu[i] - but[i]);
  const double dd[3][3] = {{h[4] - ht[4], h[5] - ht[5], h[6] - ht[6]},
                           {h[5] - ht[5], h[7] - ht[7], h[8] - ht[8]},
                           {h[6] - ht[6], h[8] - ht[8], h[9] - ht[9]}};
  double dg[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dg[i][j] = dgt[i][j] - rho * dd[i][j];
  double dgb[3], gdb[3];
  mv3(dg, bu, dgb);
  mv3(g, db, gdb);
  const double dtt = -dA + (bu[0] * dgb[0] + bu[1] * dgb[1] + bu[2] * dgb[2]) +
                     2.0 * (bl[0] * db[0] + bl[1] * db[1] + bl[2] * db[2]);
  V10<double> o;
  o[0] = dtt;
  for (int i = 0; i < 3; ++i) o[1 + i] = dgb[i] + gdb[i];
  o[4] = dg[0][0];
  o[5] = dg[1][0];
  o[6] = dg[2][0];
  o[7] = dg[1][1];
  o[8] = dg[2][1];
  o[9] = dg[2][2];
  return o;
}

// SplitMix64 point draw (initialdata.jl:2004-2013).
inline double point_draw(uint64_t base, const double x[3], int v) {
  uint64_t z = base;
  uint64_t cs[4];
  for (int i = 0; i < 3; ++i) std::memcpy(&cs[i], &x[i], 8);
  cs[3] = (uint64_t)v;
  for (int i = 0; i < 4; ++i) {
    z = (z ^ cs[i]) + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
  }
  return (double)(z >> 11) * 0x1.0p-53;
}

}  // namespace orc
