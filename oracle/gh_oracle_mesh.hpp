// CPU oracle, mesh half: a forest of closed vertex blocks (uniform, or a
// frozen 2:1-balanced hierarchy of levels), the ghost fill, the fused
// right-hand side at every owned point, RK4 and the analysis reductions.
//
// TEST INFRASTRUCTURE ONLY (see gh_oracle.hpp). Restates
// src/evolution.jl:100-435,448-695,932-949,1327-1560, src/stepping.jl:53-125,
// src/constraints.jl:311-396 and TreeAMR 0.3's ghost fill (same-level copies,
// injection from finer leaves, Lagrange prolongation from coarser ones: the
// operator set of test/interface_tests.jl:46-52; TreeAMR's own source is not
// vendored, so the rules below are stated here and pinned by polynomial
// reproduction and by the interface rates of test/interface_tests.jl:64-96).
#pragma once
#include <map>
#include <type_traits>

#include "gh_oracle.hpp"
#include "gh_oracle_bounds.hpp"

namespace orc {

// Morton key of a block's integer coordinates, x lowest.
inline uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  uint64_t k = 0;
  for (int bit = 0; bit < 21; ++bit) {
    k |= (uint64_t)((x >> bit) & 1) << (3 * bit);
    k |= (uint64_t)((y >> bit) & 1) << (3 * bit + 1);
    k |= (uint64_t)((z >> bit) & 1) << (3 * bit + 2);
  }
  return k;
}

// The parity of variable v (h_ab then Pi_ab, packed tt, tx, ty, tz, xx, xy,
// xz, yy, yz, zz) under the mirror x^d -> -x^d (state_parity,
// src/boundaries.jl:84-120): bit d is set when the spatial index d occurs an
// odd number of times in ab, so the component changes sign in that mirror.
inline int parity_odd_mask(int v) {
  static const int odd[10] = {0, 1, 2, 4, 0, 3, 5, 0, 6, 0};
  return odd[v % 10];
}
// The sign of variable v under the mirrors of the axes in mask m.
inline double parity_sign(int v, int m) {
  const int o = parity_odd_mask(v) & m;
  return ((o & 1) ^ ((o >> 1) & 1) ^ ((o >> 2) & 1)) ? -1.0 : 1.0;
}

// The quarter turn about z (state_rotation, src/boundaries.jl:145-205):
// with R (x, y, z) -> (-y, x, z), u(R p) = Q u(p), (Q u)[v] = rot_sign(v) *
// u[rot_src(v)], the covariant tensor turned index by index (t and z stay,
// the index x of the turned tensor is minus the original's y, y is x).
inline int rot_src(int v) {
  static const int src[10] = {0, 2, 1, 3, 7, 5, 8, 4, 6, 9};
  return (v / 10) * 10 + src[v % 10];
}
inline double rot_sign(int v) {
  static const double sg[10] = {1, -1, 1, 1, 1, -1, -1, 1, 1, 1};
  return sg[v % 10];
}

struct Mesh {
  int N = 16, G = 3, q = 4;
  int nb[3] = {1, 1, 1};
  double lo[3] = {0, 0, 0};
  double h = 1.0 / 16;
  bool periodic[3] = {true, true, true};
  // reflecting faces: reflect[d][0] the low face of axis d, [d][1] the high one
  bool reflect[3][2] = {{false, false}, {false, false}, {false, false}};
  // the rotating seam: the low x face glued to the low y face by a quarter
  // turn about z through (lo[0], lo[1])
  bool rotating = false;
  std::vector<std::array<int, 3>> blocks;  // block coordinates, Morton order
  long S = 0, P = 0;                       // work and state extents per axis

  void init() {
    G = q / 2 + 1;
    S = N + 1 + 2 * G;
    P = N + 1;
    blocks.clear();
    for (int k = 0; k < nb[2]; ++k)
      for (int j = 0; j < nb[1]; ++j)
        for (int i = 0; i < nb[0]; ++i) blocks.push_back({i, j, k});
    std::stable_sort(blocks.begin(), blocks.end(), [](const auto& a, const auto& b) {
      return morton3(a[0], a[1], a[2]) < morton3(b[0], b[1], b[2]);
    });
  }
  long nblocks() const { return (long)blocks.size(); }
  long state_size() const { return P * P * P * 20 * nblocks(); }
  long work_size() const { return S * S * S * 20 * nblocks(); }
  long sidx(long b, int v, long i, long j, long k) const {
    return i + P * (j + P * (k + P * (v + 20 * b)));
  }
  // work index of owned-relative l (may be negative, down to -G)
  long widx(long b, int v, long i, long j, long k) const {
    return (i + G) + S * ((j + G) + S * ((k + G) + S * (v + 20 * b)));
  }
  // The spacing of block b: h on a uniform forest, h / 2^level on a forest.
  double hb(long b) const { return level.empty() ? h : h / (double)(1L << level[b]); }
  void position(long b, long i, long j, long k, double x[3]) const {
    long g[3] = {blocks[b][0] * (long)N + i, blocks[b][1] * (long)N + j, blocks[b][2] * (long)N + k};
    const double hh = hb(b);
    for (int d = 0; d < 3; ++d) x[d] = lo[d] + (double)g[d] * hh;
  }
  // Source of a work point: -2 Dirichlet, -1 own primary, else owner block and
  // local index, with *mirror the axes across whose reflecting face the point
  // lies (its value is then the source's times the variables' parities). A
  // point beyond a face that neither wraps nor reflects is the hook's, also
  // where it lies beyond a reflecting face on another axis.
  int classify(long b, const long l[3], long& owner, long loc[3], int* mirror = nullptr,
               int* turn = nullptr) const {
    long ob[3];
    int mm = 0, tn = 0;
    long gs[3];
    for (int d = 0; d < 3; ++d) gs[d] = blocks[b][d] * (long)N + l[d];
    if (rotating) {
      // a point beyond the seam's x face is R p, beyond its y face R^-1 p,
      // beyond both R^2 p, with p the point inside: u = Q^turn u(p)
      const long x = gs[0], y = gs[1];
      if (x < 0 && y < 0) {
        gs[0] = -x;
        gs[1] = -y;
        tn = 2;
      } else if (x < 0) {
        gs[0] = y;
        gs[1] = -x;
        tn = 1;
      } else if (y < 0) {
        gs[0] = -y;
        gs[1] = x;
        tn = 3;
      }
    }
    if (turn) *turn = tn;
    for (int d = 0; d < 3; ++d) {
      long g = gs[d];
      long tot = (long)nb[d] * N;
      if (periodic[d]) {
        long gw = ((g % tot) + tot) % tot;
        ob[d] = gw / N;
        loc[d] = gw - ob[d] * N;
      } else {
        if (g < 0) {
          if (!reflect[d][0]) return -2;
          g = -g;
          mm |= 1 << d;
        } else if (g > tot) {
          if (!reflect[d][1]) return -2;
          g = 2 * tot - g;
          mm |= 1 << d;
        }
        ob[d] = std::min(g / N, (long)nb[d] - 1);
        loc[d] = g - ob[d] * N;
      }
    }
    if (mirror) *mirror = mm;
    // owner block index (linear search over the Morton order is fine here)
    owner = block_index(ob[0], ob[1], ob[2]);
    if (mm == 0 && tn == 0 && owner == b && loc[0] == l[0] && loc[1] == l[1] && loc[2] == l[2]) return -1;
    return 0;
  }
  // The axes whose reflecting wall plane the owned point l of block b lies on.
  int wall_mask(long b, const long l[3]) const {
    int m = 0;
    for (int d = 0; d < 3; ++d) {
      const long g = blocks[b][d] * (long)N + l[d], tot = (long)nb[d] * N;
      if ((reflect[d][0] && g == 0) || (reflect[d][1] && g == tot)) m |= 1 << d;
    }
    return m;
  }
  bool reflects() const {
    for (int d = 0; d < 3; ++d)
      if (reflect[d][0] || reflect[d][1]) return true;
    return false;
  }
  std::vector<long> lookup;
  long block_index(long i, long j, long k) const {
    return lookup[i + nb[0] * (j + nb[1] * k)];
  }
  void build_lookup() {
    lookup.assign((size_t)nb[0] * nb[1] * nb[2], -1);
    for (long b = 0; b < nblocks(); ++b)
      lookup[blocks[b][0] + nb[0] * (blocks[b][1] + nb[1] * blocks[b][2])] = b;
  }
  // Is owned point l of block b a primary (counted once in sums)?
  bool is_primary(long b, long i, long j, long k) const {
    long l[3] = {i, j, k}, owner, loc[3];
    if (forest()) {
      Src s;
      return forest_classify(b, l, s) == F_PRIMARY;
    }
    return classify(b, l, owner, loc) == -1;
  }

  // ---- a frozen hierarchy (gh_forest(; refined), hole_forest:
  // src/initialdata.jl:1370-1515) ----
  // blocks[b] holds leaf b's coordinates in the brick of its level (nb_d 2^l
  // blocks of N cells of spacing h / 2^l per axis); level[b] its level.
  std::vector<int> level;  // empty: a uniform forest
  int Lmax = 0;
  int prolong = 0;         // the prolongation order p (an even number of points)
  std::map<std::array<long, 4>, long> leaf;
  std::vector<double> wmid;  // Lagrange weights at the midpoint, p nodes
  // TreeAMR 0.1's half-open blocks: a leaf owns [0, N) per axis, a point
  // belongs to the leaf whose half-open extent holds it (when it is a vertex
  // of that leaf; else it is prolongated), and the upper plane of a
  // non-periodic axis is a Dirichlet ghost. Off: TreeAMR 0.3's closed blocks.
  bool half_open = false;
  bool forest() const { return !level.empty(); }

  void init_forest(const std::vector<std::array<int, 4>>& leaves, int p) {
    G = q / 2 + 1;
    S = N + 1 + 2 * G;
    P = N + 1;
    prolong = p;
    blocks.clear();
    level.clear();
    leaf.clear();
    Lmax = 0;
    for (const auto& lf : leaves) {
      level.push_back(lf[0]);
      blocks.push_back({lf[1], lf[2], lf[3]});
      leaf[{(long)lf[0], (long)lf[1], (long)lf[2], (long)lf[3]}] = (long)blocks.size() - 1;
      Lmax = std::max(Lmax, lf[0]);
    }
    // w_k = prod_{j != k} (1/2 - x_j) / (x_k - x_j) at nodes x = -p/2+1 .. p/2:
    // the numerator and the denominator are exact, so one rounding.
    wmid.assign(p, 0.0);
    for (int k = 0; k < p; ++k) {
      double num = 1.0, den = 1.0;
      const double xk = k - p / 2 + 1;
      for (int j = 0; j < p; ++j) {
        if (j == k) continue;
        const double xj = j - p / 2 + 1;
        num *= 0.5 - xj;
        den *= xk - xj;
      }
      wmid[k] = num / den;
    }
  }
  enum { F_PRIMARY = 0, F_COPY = 1, F_DIRICHLET = 2, F_PROLONG = 3 };
  struct Src {
    long b = -1;
    long loc[3] = {0, 0, 0};  // COPY: the point; PROLONG: the stencil's first point
    int half[3] = {0, 0, 0};  // PROLONG: the axes where the point is a coarse midpoint
    int mirror = 0;           // the axes whose reflecting face the point lies beyond
    int turn = 0;             // quarter turns across the rotating seam
  };
  // The leaf of level lv holding integer point gv (index in that level's
  // brick) as a vertex: per axis the block where the local index is below N
  // first, then the lower one where it is N (the uniform rule), in the order
  // x, y, z of the alternatives' bits.
  bool find_leaf(int lv, const long gv[3], long& ob, long loc[3]) const {
    long o0[3], alt[3];
    for (int d = 0; d < 3; ++d) {
      const long nbl = (long)nb[d] << lv;
      o0[d] = std::min(gv[d] / N, nbl - 1);
      alt[d] = -1;
      if (half_open) {  // the one leaf whose [0, N) extent holds gv
        // (on the upper plane of a non-periodic axis, a point only a
        // transfer asks for, the closed rule: the last block)
        o0[d] = std::min(gv[d] / N, nbl - 1);
        if (periodic[d] && gv[d] / N >= nbl) o0[d] = gv[d] / N - nbl;
        continue;
      }
      if (gv[d] % N == 0 && o0[d] * N == gv[d]) {
        if (o0[d] > 0) alt[d] = o0[d] - 1;
        else if (periodic[d]) alt[d] = nbl - 1;
      }
    }
    for (int mask = 0; mask < 8; ++mask) {
      long o[3];
      bool ok = true;
      for (int d = 0; d < 3; ++d) {
        if (mask >> d & 1) {
          if (alt[d] < 0) ok = false;
          o[d] = alt[d];
        } else {
          o[d] = o0[d];
        }
      }
      if (!ok) continue;
      auto it = leaf.find({(long)lv, o[0], o[1], o[2]});
      if (it == leaf.end()) continue;
      ob = it->second;
      for (int d = 0; d < 3; ++d) {
        loc[d] = gv[d] - o[d] * N;
        if (loc[d] < 0) loc[d] += ((long)nb[d] << lv) * N;  // the periodic alternative at 0
      }
      return true;
    }
    return false;
  }
  // Where point l of leaf b's work array takes its value from: its own
  // primary; a copy of the primary of the finest leaf holding the point as a
  // vertex (a same-level copy, an injection from a finer leaf, a coincident
  // coarse vertex); the Dirichlet hook outside a non-periodic axis; or the
  // order-p Lagrange prolongation from the leaf one level coarser that holds
  // it, whose midpoint axes take p points and whose other axes one.
  int forest_classify(long b, const long l[3], Src& s) const {
    const int L = level[b];
    long gl[3];
    for (int d = 0; d < 3; ++d) gl[d] = blocks[b][d] * (long)N + l[d];
    // across the rotating seam and the reflecting faces: the point inside the
    // box whose value this one takes, turned and mirrored (as the uniform fill)
    int tn = 0, mm = 0;
    if (rotating) {
      const long x = gl[0], y = gl[1];
      if (x < 0 && y < 0) {
        gl[0] = -x;
        gl[1] = -y;
        tn = 2;
      } else if (x < 0) {
        gl[0] = y;
        gl[1] = -x;
        tn = 1;
      } else if (y < 0) {
        gl[0] = -y;
        gl[1] = x;
        tn = 3;
      }
    }
    for (int d = 0; d < 3; ++d) {
      const long tot = ((long)nb[d] << L) * N;
      if (periodic[d]) {
        gl[d] = ((gl[d] % tot) + tot) % tot;
      } else if (gl[d] < 0) {
        if (!reflect[d][0]) return F_DIRICHLET;
        gl[d] = -gl[d];
        mm |= 1 << d;
      } else if (gl[d] > tot) {
        if (!reflect[d][1]) return F_DIRICHLET;
        gl[d] = 2 * tot - gl[d];
        mm |= 1 << d;
      } else if (half_open && gl[d] == tot) {
        return F_DIRICHLET;
      }
    }
    const int c = classify_point(L, gl, s);
    s.mirror = mm;
    s.turn = tn;
    if (c == F_COPY && mm == 0 && tn == 0 && s.b == b && s.loc[0] == l[0] && s.loc[1] == l[1] && s.loc[2] == l[2])
      return F_PRIMARY;
    return c;
  }
  // The axes whose reflecting wall plane holds owned point l of leaf b.
  int forest_wall_mask(long b, const long l[3]) const {
    int m = 0;
    for (int d = 0; d < 3; ++d) {
      const long g = blocks[b][d] * (long)N + l[d], tot = ((long)nb[d] << level[b]) * N;
      if ((reflect[d][0] && g == 0) || (reflect[d][1] && g == tot)) m |= 1 << d;
    }
    return m;
  }
  // Where the point gl (index in the brick of level L, wrapped) takes its
  // value from in this forest: F_COPY with its owner, or F_PROLONG from the
  // level-(L-1) leaf holding it.
  int classify_point(int L, const long gl[3], Src& s) const {
    for (int lv = Lmax; lv >= 0; --lv) {
      long gv[3];
      bool vertex = true;
      for (int d = 0; d < 3; ++d) {
        if (lv >= L) {
          gv[d] = gl[d] << (lv - L);
        } else {
          const long f = 1L << (L - lv);
          if (gl[d] % f != 0) vertex = false;
          gv[d] = gl[d] / f;
        }
      }
      if (!vertex) continue;
      long ob, loc[3];
      if (!find_leaf(lv, gv, ob, loc)) continue;
      s.b = ob;
      for (int d = 0; d < 3; ++d) s.loc[d] = loc[d];
      return F_COPY;
    }
    if (L == 0) throw std::runtime_error("a level-0 point that no leaf holds: the leaves do not tile the box");
    // the coarse leaf: integer axes by the uniform rule, midpoint axes in
    // the cell [m, m + 1] of level L - 1
    long gc[3];
    int half[3];
    for (int d = 0; d < 3; ++d) {
      half[d] = (int)(gl[d] & 1);
      gc[d] = half[d] ? (gl[d] - 1) / 2 : gl[d] / 2;
    }
    long o0[3], alt[3];
    for (int d = 0; d < 3; ++d) {
      const long nbl = (long)nb[d] << (L - 1);
      o0[d] = std::min(gc[d] / N, nbl - 1);
      alt[d] = -1;
      if (!half[d] && gc[d] % N == 0 && o0[d] * N == gc[d]) {
        if (o0[d] > 0) alt[d] = o0[d] - 1;
        else if (periodic[d]) alt[d] = nbl - 1;
      }
    }
    for (int mask = 0; mask < 8; ++mask) {
      long o[3];
      bool ok = true;
      for (int d = 0; d < 3; ++d) {
        if (mask >> d & 1) {
          if (alt[d] < 0) ok = false;
          o[d] = alt[d];
        } else {
          o[d] = o0[d];
        }
      }
      if (!ok) continue;
      auto it = leaf.find({(long)(L - 1), o[0], o[1], o[2]});
      if (it == leaf.end()) continue;
      s.b = it->second;
      for (int d = 0; d < 3; ++d) {
        long loc = gc[d] - o[d] * N;
        if (loc < 0) loc += ((long)nb[d] << (L - 1)) * N;
        s.half[d] = half[d];
        s.loc[d] = half[d] ? loc - prolong / 2 + 1 : loc;
      }
      return F_PROLONG;
    }
    throw std::runtime_error("a ghost of a level-" + std::to_string(L) +
                             " leaf lies in no leaf of level L - 1: the forest is not 2:1 balanced");
  }
};

struct Problem {
  Mesh mesh;
  Case cs;
  StateBounds bounds;
  std::vector<double> work;
  std::vector<double> Hsrc;  // sampled gauge source at owned points, 20 per point
  // matter's 3+1 projections at owned points (empty: vacuum): rho P^3 x nb,
  // S P^3 x 3 x nb, S_ij P^3 x 6 x nb (xx, xy, xz, yy, yz, zz)
  std::vector<double> mat_rho, mat_S, mat_Sij;
  std::vector<double> w1, w2, wD;

  void setup() {
    if (!mesh.forest()) {
      mesh.init();
      mesh.build_lookup();
    }
    work.assign(mesh.work_size(), 0.0);
    w1 = to_double(rational_derivative_weights(mesh.q, 1));
    w2 = to_double(rational_derivative_weights(mesh.q, 2));
    wD = to_double(rational_dissipation_weights(dissipation_rank(mesh.q)));
    if (cs.gauge == GAUGE_SAMPLED) sample_gauge_source(0.0);
  }
  // sample_gauge_source! (gauge.jl:156-184) with the core rule.
  void sample_gauge_source(double t) {
    const Mesh& m = mesh;
    Hsrc.assign(m.state_size(), 0.0);
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            double x[3], xc[3], Hl[4], dHl[4][4];
            m.position(b, i, j, k, x);
            core_position(cs, t, x, xc);
            gauge_source_grad(cs.bg, t, xc[0], xc[1], xc[2], Hl, dHl);
            for (int a = 0; a < 4; ++a) Hsrc[m.sidx(b, a, i, j, k)] = Hl[a];
            // column-major dHl[c][a] (gauge.jl: entries 5:20 in SMatrix order)
            for (int n = 0; n < 16; ++n) Hsrc[m.sidx(b, 4 + n, i, j, k)] = dHl[n % 4][n / 4];
          }
  }

  // fill_exact! (initialdata.jl:1913-1914): every owned point of the state.
  void fill_exact(double* u, double t) const {
    const Mesh& m = mesh;
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            double x[3], s[20];
            m.position(b, i, j, k, x);
            case_state(cs, t, x, s);
            for (int v = 0; v < 20; ++v) u[m.sidx(b, v, i, j, k)] = s[v];
          }
  }
  // add_noise! (initialdata.jl:1961-2000), with the SplitMix64 draw.
  void add_noise(double* u, uint64_t base, double amp) const {
    const Mesh& m = mesh;
    for (long b = 0; b < m.nblocks(); ++b)
      for (int v = 0; v < 20; ++v)
        for (long k = 0; k <= m.N; ++k)
          for (long j = 0; j <= m.N; ++j)
            for (long i = 0; i <= m.N; ++i) {
              double x[3];
              m.position(b, i, j, k, x);
              // the two seam planes are the same points a quarter turn apart:
              // they keep their data (src/initialdata.jl:1947-1952)
              if (!(m.rotating && ((m.blocks[b][0] == 0 && i == 0) || (m.blocks[b][1] == 0 && j == 0))))
                u[m.sidx(b, v, i, j, k)] += amp * (2.0 * point_draw(base, x, v + 1) - 1.0);
              // on a low reflecting wall the noise is projected onto the
              // parity: the odd variables there are zero (add_noise!,
              // src/initialdata.jl:1938-1990, which projects the low walls)
              for (int d = 0; d < 3; ++d) {
                const long gl = m.blocks[b][d] * (long)m.N + (d == 0 ? i : d == 1 ? j : k);
                if (m.reflect[d][0] && gl == 0 && (parity_odd_mask(v) >> d & 1)) u[m.sidx(b, v, i, j, k)] = 0.0;
              }
            }
  }

  // prepare! (evolution.jl:1349-1357): scatter and fill.
  void prepare(const double* u, double t) {
    const Mesh& m = mesh;
    for (long b = 0; b < m.nblocks(); ++b)
      for (int v = 0; v < 20; ++v)
        for (long k = 0; k <= m.N; ++k)
          for (long j = 0; j <= m.N; ++j)
            for (long i = 0; i <= m.N; ++i) work[m.widx(b, v, i, j, k)] = u[m.sidx(b, v, i, j, k)];
    if (m.forest()) {
      fill_forest(t);
      return;
    }
    // a reflecting wall plane takes each variable's parity: its odd
    // variables are zero (TreeAMR's projection of the wall at every fill of
    // closed blocks, src/constraints.jl:222-225)
    if (m.reflects())
      for (long b = 0; b < m.nblocks(); ++b)
        for (long k = 0; k <= m.N; ++k)
          for (long j = 0; j <= m.N; ++j)
            for (long i = 0; i <= m.N; ++i) {
              const long l[3] = {i, j, k};
              const int wm = m.wall_mask(b, l);
              if (!wm) continue;
              for (int v = 0; v < 20; ++v)
                if (parity_odd_mask(v) & wm) work[m.widx(b, v, i, j, k)] = 0.0;
            }
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = -m.G; k <= m.N + m.G; ++k)
        for (long j = -m.G; j <= m.N + m.G; ++j)
          for (long i = -m.G; i <= m.N + m.G; ++i) {
            long l[3] = {i, j, k}, owner, loc[3];
            int mm = 0, tn = 0;
            int c = m.classify(b, l, owner, loc, &mm, &tn);
            if (c == -1) continue;
            if (c == -2) {
              long g[3] = {m.blocks[b][0] * (long)m.N + i, m.blocks[b][1] * (long)m.N + j,
                           m.blocks[b][2] * (long)m.N + k};
              double x[3], s[20];
              for (int d = 0; d < 3; ++d) x[d] = m.lo[d] + (double)g[d] * m.h;
              case_state(cs, t, x, s);
              for (int v = 0; v < 20; ++v) work[m.widx(b, v, i, j, k)] = s[v];
            } else {
              // Q^tn: (Q w)[v] = rot_sign(v) w[rot_src(v)], applied tn times
              for (int v = 0; v < 20; ++v) {
                int sv = v;
                double sg = parity_sign(v, mm);
                for (int n = 0; n < tn; ++n) {
                  sg *= rot_sign(sv);
                  sv = rot_src(sv);
                }
                work[m.widx(b, v, i, j, k)] = sg * work[m.widx(owner, sv, loc[0], loc[1], loc[2])];
              }
            }
          }
  }

  // The fill of a forest, after the scatter: first every copy (each reads a
  // primary, which no fill writes) and every Dirichlet ghost, then the
  // prolongations level by level from the coarsest, so that the coarse
  // ghosts a stencil reads are final before it runs. A point that is not its
  // block's primary takes its owner's value, the owned face points of a
  // coarse leaf included (the injection of the finer leaf's values).
  void fill_forest(double t) {
    const Mesh& m = mesh;
    // the reflecting walls' odd variables first, so that every copy and
    // prolongation reads the projected values (the uniform fill's order)
    if (m.reflects())
      for (long b = 0; b < m.nblocks(); ++b)
        for (long k = 0; k <= m.N; ++k)
          for (long j = 0; j <= m.N; ++j)
            for (long i = 0; i <= m.N; ++i) {
              const long l[3] = {i, j, k};
              const int wm = m.forest_wall_mask(b, l);
              if (!wm) continue;
              for (int v = 0; v < 20; ++v)
                if (parity_odd_mask(v) & wm) work[m.widx(b, v, i, j, k)] = 0.0;
            }
    // the variable a turned and mirrored entry reads, and its sign
    auto turned = [](const Mesh::Src& s, int v, int& sv) {
      double sg = parity_sign(v, s.mirror);
      sv = v;
      for (int n = 0; n < s.turn; ++n) {
        sg *= rot_sign(sv);
        sv = rot_src(sv);
      }
      return sg;
    };
    for (int pass = 0; pass <= m.Lmax; ++pass)
      for (long b = 0; b < m.nblocks(); ++b) {
        if (pass > 0 && m.level[b] != pass) continue;
        for (long k = -m.G; k <= m.N + m.G; ++k)
          for (long j = -m.G; j <= m.N + m.G; ++j)
            for (long i = -m.G; i <= m.N + m.G; ++i) {
              long l[3] = {i, j, k};
              Mesh::Src s;
              const int c = m.forest_classify(b, l, s);
              if (c == Mesh::F_PRIMARY) continue;
              if (pass == 0 && c == Mesh::F_DIRICHLET) {
                double x[3], st[20];
                m.position(b, i, j, k, x);
                case_state(cs, t, x, st);
                for (int v = 0; v < 20; ++v) work[m.widx(b, v, i, j, k)] = st[v];
              } else if (pass == 0 && c == Mesh::F_COPY) {
                for (int v = 0; v < 20; ++v) {
                  int sv;
                  const double sg = turned(s, v, sv);
                  work[m.widx(b, v, i, j, k)] = sg * work[m.widx(s.b, sv, s.loc[0], s.loc[1], s.loc[2])];
                }
              } else if (pass > 0 && c == Mesh::F_PROLONG) {
                // sum-factorised: x innermost, then y, then z
                const int n[3] = {s.half[0] ? m.prolong : 1, s.half[1] ? m.prolong : 1,
                                  s.half[2] ? m.prolong : 1};
                for (int v = 0; v < 20; ++v) {
                  int sv;
                  const double sg = turned(s, v, sv);
                  const double* base = &work[m.widx(s.b, sv, s.loc[0], s.loc[1], s.loc[2])];
                  double sz = 0.0;
                  for (int c2 = 0; c2 < n[2]; ++c2) {
                    double sy = 0.0;
                    for (int c1 = 0; c1 < n[1]; ++c1) {
                      double sx = 0.0;
                      for (int c0 = 0; c0 < n[0]; ++c0)
                        sx += (n[0] > 1 ? m.wmid[c0] : 1.0) * base[c0 + m.S * (c1 + m.S * c2)];
                      sy += (n[1] > 1 ? m.wmid[c1] : 1.0) * sx;
                    }
                    sz += (n[2] > 1 ? m.wmid[c2] : 1.0) * sy;
                  }
                  work[m.widx(b, v, i, j, k)] = sg * sz;
                }
              }
            }
      }
  }

  // ---- tangents: the reference's CPU path run with a forward-mode dual as
  // its float type (one direction). work_dot holds the tangent of the work
  // array the fill left in work; a load gives the dual (value, tangent).
  using D1 = Dual<double, 1>;
  std::vector<double> work_dot;
  template <class T>
  T ld(long idx) const {
    if constexpr (std::is_same<T, D1>::value) {
      T r(work[idx]);
      r.d[0] = work_dot.empty() ? 0.0 : work_dot[idx];
      return r;
    } else {
      return T(work[idx]);
    }
  }

  // ---- the fused kernel at one owned point (evolution.jl:258-435,448-695) ----
  template <class T>
  T axis_stencil(const std::vector<double>& w, long base, long stride) const {
    int n = (int)w.size(), r = (n - 1) / 2;
    T s = T(0.0);
    bool first = true;
    for (int k = 0; k < n; ++k) {
      if (w[k] == 0.0) continue;  // zero weights are bitwise-neutral (CODE.md:6647-6653)
      T t = w[k] * ld<T>(base + (long)(k - r) * stride);
      s = first ? t : s + t;
      first = false;
    }
    return s;
  }
  template <class T>
  T mixed_stencil(const std::vector<double>& w, long base, long s1, long s2) const {
    int n = (int)w.size(), r = (n - 1) / 2;
    T outer = T(0.0);
    bool fo = true;
    for (int a = 0; a < n; ++a) {
      if (w[a] == 0.0) continue;
      T inner = T(0.0);
      bool fi = true;
      for (int e = 0; e < n; ++e) {
        if (w[e] == 0.0) continue;
        T t = w[e] * ld<T>(base + (long)(a - r) * s1 + (long)(e - r) * s2);
        inner = fi ? t : inner + t;
        fi = false;
      }
      T t = w[a] * inner;
      outer = fo ? t : outer + t;
      fo = false;
    }
    return outer;
  }

  // F(u) at one owned point into dth, dtP (gh_rhs_at_point, evolution.jl:423-435).
  template <class T>
  void rhs_point_F(long b, long i, long j, long k, double tt, T g0, T g2, T epsh, bool diss,
                   V10<T>& dth, V10<T>& dtP, bool head_only = false) const {
    const Mesh& m = mesh;
    long st[3] = {1, m.S, m.S * m.S};
    long sv = m.S * m.S * m.S;
    long var = m.widx(b, 0, i, j, k);
    T inv_h = T(1.0) / T(m.hb(b));
    // (1) the state, the 30 first derivatives, the coefficient set
    V10<T> hv, Pv, dh[3];
    for (int v = 0; v < NC; ++v) {
      hv[v] = ld<T>(var + v * sv);
      Pv[v] = ld<T>(var + (NC + v) * sv);
    }
    for (int d = 0; d < 3; ++d)
      for (int v = 0; v < NC; ++v) dh[d][v] = inv_h * axis_stencil<T>(w1, var + v * sv, st[d]);
    MQ<T> mq = metric_quantities(hv);
    T a_div = mq.alpha / mq.sqrtg;
    // (2) d_t h, complete (evolution.jl:301-307)
    for (int v = 0; v < NC; ++v) {
      T s = mq.beta[0] * dh[0][v] + mq.beta[1] * dh[1][v] + mq.beta[2] * dh[2][v] + a_div * Pv[v];
      if (diss) {
        long bh = var + v * sv;
        s = s + epsh * (axis_stencil<T>(wD, bh, st[0]) + axis_stencil<T>(wD, bh, st[1]) +
                        axis_stencil<T>(wD, bh, st[2]));
      }
      dth[v] = s;
    }
    // (3) the gauge source and the source
    T Hl[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
    T dHl[4][4];
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 4; ++c) dHl[a][c] = T(0.0);
    if (cs.gauge == GAUGE_SAMPLED) {
      for (int a = 0; a < 4; ++a) Hl[a] = T(Hsrc[m.sidx(b, a, i, j, k)]);
      for (int n = 0; n < 16; ++n) dHl[n % 4][n / 4] = T(Hsrc[m.sidx(b, 4 + n, i, j, k)]);
    } else if (cs.gauge == GAUGE_KS) {
      ks_algebraic_source<T>(cs.bg.M, cs.ks_S, cs.ks_u, hv, dth, dh, Hl, dHl);
    }
    V10<T> msrc = gh_node_source_lean(mq, dth, dh, Hl, dHl, g0, g2);
    if (!mat_rho.empty()) {
      // the matter adapter's (rho, S_i, S_ij) at the point (evolution.jl:1295-1311)
      const long P3 = m.P * m.P * m.P, pi = i + m.P * (j + m.P * k);
      const T rho = T(mat_rho[b * P3 + pi]);
      T S[3], Sij[3][3];
      for (int c = 0; c < 3; ++c) S[c] = T(mat_S[(b * 3 + c) * P3 + pi]);
      const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
      for (int n = 0; n < 6; ++n) {
        const T v = T(mat_Sij[(b * 6 + n) * P3 + pi]);
        Sij[ij[n][0]][ij[n][1]] = v;
        Sij[ij[n][1]][ij[n][0]] = v;
      }
      const V10<T> ms = matter_source(mq, rho, S, Sij);
      for (int v = 0; v < NC; ++v) msrc[v] = msrc[v] + ms[v];
    }
    // (4) the coefficients of the Pi components
    T A[3][3];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) A[a][c] = (mq.alpha * mq.sqrtg) * mq.gamu[a][c];
    T divb, divA[3];
    metric_divergences(mq, dh, divb, divA);
    if (head_only) return;
    // gh_rhs_pi per component (evolution.jl:324-354)
    T inv_h2 = inv_h * inv_h;
    for (int v = 0; v < NC; ++v) {
      long bh = var + v * sv, bP = bh + NC * sv;
      T Pi = ld<T>(bP);
      T dh1 = inv_h * axis_stencil<T>(w1, bh, st[0]);
      T dh2 = inv_h * axis_stencil<T>(w1, bh, st[1]);
      T dh3 = inv_h * axis_stencil<T>(w1, bh, st[2]);
      T dP1 = inv_h * axis_stencil<T>(w1, bP, st[0]);
      T dP2 = inv_h * axis_stencil<T>(w1, bP, st[1]);
      T dP3 = inv_h * axis_stencil<T>(w1, bP, st[2]);
      T s = mq.beta[0] * dP1 + mq.beta[1] * dP2 + mq.beta[2] * dP3 + divb * Pi + divA[0] * dh1 +
            divA[1] * dh2 + divA[2] * dh3;
      s = s + (A[0][0] * (inv_h2 * axis_stencil<T>(w2, bh, st[0])) +
               A[1][1] * (inv_h2 * axis_stencil<T>(w2, bh, st[1])) +
               A[2][2] * (inv_h2 * axis_stencil<T>(w2, bh, st[2])));
      T dxy = inv_h2 * mixed_stencil<T>(w1, bh, st[0], st[1]);
      T dxz = inv_h2 * mixed_stencil<T>(w1, bh, st[0], st[2]);
      T dyz = inv_h2 * mixed_stencil<T>(w1, bh, st[1], st[2]);
      s = s + 2.0 * (A[0][1] * dxy + A[0][2] * dxz + A[1][2] * dyz);
      if (diss)
        s = s + epsh * (axis_stencil<T>(wD, bP, st[0]) + axis_stencil<T>(wD, bP, st[1]) +
                        axis_stencil<T>(wD, bP, st[2]));
      dtP[v] = s + msrc[v];
    }
  }

  // gh_rhs_point! (evolution.jl:448-695) for the sphere's variants, at rest
  // or moving (:damped follows a moving hole, :526-583, :652-680).
  void rhs_point(long b, long i, long j, long k, double t, double* du) const {
    const Mesh& m = mesh;
    double x[3];
    m.position(b, i, j, k, x);
    double g0 = damping_rate(cs, t, x);
    double epsh = dissipation_rate(cs, t, x) * (1.0 / m.hb(b));  // evolution.jl:478
    bool diss = has_dissipation(cs);
    V10<double> dth, dtP;
    if (cs.interior != INT_NONE) {
      const IGeo g = interior_geo(cs, t, x);
      const double r = g.r;
      bool frozen = cs.interior == INT_PASTED ? r < g.r1 : r < g.r0;
      const long tp = tidx(b, i, j, k);
      if (frozen) {
        if (cs.interior == INT_FITTED) {
          // the fitted core (evolution.jl:515-546): du = [S] - rho_max (u - (A + (t - t_f) S))
          long var = m.widx(b, 0, i, j, k), sv = m.S * m.S * m.S;
          const long ps = m.P * m.P * m.P;
          for (int v = 0; v < 20; ++v) {
            const double A = tw[tp + v * ps], Sl = tw[tp + (20 + v) * ps];
            du[m.sidx(b, v, i, j, k)] = (rate ? Sl : 0.0) - cs.rho_max * (work[var + v * sv] - (A + (t - t_f) * Sl));
          }
          return;
        }
        if (core_follows(cs)) {
          // the moving :damped core relaxes toward the layer's target at the
          // core rule's point, the metric in the ADM variables
          const double rc = core_rate(cs);
          double xc[3];
          core_position(cs, t, x, xc);
          V10<double> hd, Pd, dhd[3], dthd, dtPd, hu;
          background_state(cs.bg, t, xc, hd, Pd, dhd);
          core_target_rate(cs, t, x, dthd, dtPd);
          long var = m.widx(b, 0, i, j, k), sv = m.S * m.S * m.S;
          for (int v = 0; v < NC; ++v) hu[v] = work[var + v * sv];
          V10<double> dh = core_relaxation_rate(hu, hd, rc, dthd);
          for (int v = 0; v < NC; ++v) {
            du[m.sidx(b, v, i, j, k)] = dh[v];
            du[m.sidx(b, NC + v, i, j, k)] = dtPd[v] - rc * (work[var + (NC + v) * sv] - Pd[v]);
          }
          return;
        }
        for (int v = 0; v < 20; ++v) du[m.sidx(b, v, i, j, k)] = 0.0;
        return;
      }
      rhs_point_F<double>(b, i, j, k, t, g0, cs.g2, epsh, diss, dth, dtP);
      if (r >= g.r1) {
        for (int v = 0; v < NC; ++v) {
          du[m.sidx(b, v, i, j, k)] = dth[v];
          du[m.sidx(b, NC + v, i, j, k)] = dtP[v];
        }
        return;
      }
      double w, rho, rramp = cs.rho_ramp;
      if (cs.interior == INT_FITTED && trail != 0.0) {
        // _trail_profiles (evolution.jl:1256-1272): rho's ramp narrowed on
        // the trailing side by 1 - trail max(0, -n.v/|v|)
        double cc[3];
        center_at(cs, t, cc);
        const double sp = std::sqrt(cs.vel[0] * cs.vel[0] + cs.vel[1] * cs.vel[1] + cs.vel[2] * cs.vel[2]);
        double zeta = 0.0;
        if (!(sp == 0.0 || r == 0.0))
          zeta = std::max(0.0, -((x[0] - cc[0]) * cs.vel[0] + (x[1] - cc[1]) * cs.vel[1] +
                                 (x[2] - cc[2]) * cs.vel[2]) / (r * sp));
        rramp = cs.rho_ramp * (1.0 - trail * zeta);
      }
      layer_profiles(r, g.r0, g.r1, cs.rho_max, cs.w_ramp, rramp, w, rho);
      if (cs.interior == INT_FROZEN) rho = 0.0;
      long var = m.widx(b, 0, i, j, k), sv = m.S * m.S * m.S;
      if (cs.interior == INT_FITTED) {
        // the fitted layer (evolution.jl:600-645): w F + [(1 - w) S] - rho (u - (A + (t - t_f) S))
        const long ps = m.P * m.P * m.P;
        const double om = 1.0 - w;
        for (int v = 0; v < 20; ++v) {
          const double A = tw[tp + v * ps], Sl = tw[tp + (20 + v) * ps];
          const double F = v < NC ? dth[v] : dtP[v - NC];
          du[m.sidx(b, v, i, j, k)] = w * F + (rate ? om * Sl : 0.0) - rho * (work[var + v * sv] - (A + (t - t_f) * Sl));
        }
        return;
      }
      if (core_follows(cs)) {
        // the moving layer: (1 - w) d_t u_exact keeps the moving solution steady
        V10<double> hm, Pm, dthm, dtPm;
        background_state_rate(cs.bg, t, x, cs.delta, hm, Pm, dthm, dtPm);
        const double om = 1.0 - w;
        for (int v = 0; v < NC; ++v) {
          du[m.sidx(b, v, i, j, k)] = w * dth[v] + om * dthm[v] - rho * (work[var + v * sv] - hm[v]);
          du[m.sidx(b, NC + v, i, j, k)] =
              w * dtP[v] + om * dtPm[v] - rho * (work[var + (NC + v) * sv] - Pm[v]);
        }
        return;
      }
      V10<double> he, Pe, dhe[3];
      background_state(cs.bg, t, x, he, Pe, dhe);
      for (int v = 0; v < NC; ++v) {
        du[m.sidx(b, v, i, j, k)] = w * dth[v] - rho * (work[var + v * sv] - he[v]);
        du[m.sidx(b, NC + v, i, j, k)] = w * dtP[v] - rho * (work[var + (NC + v) * sv] - Pe[v]);
      }
      return;
    }
    rhs_point_F<double>(b, i, j, k, t, g0, cs.g2, epsh, diss, dth, dtP);
    for (int v = 0; v < NC; ++v) {
      du[m.sidx(b, v, i, j, k)] = dth[v];
      du[m.sidx(b, NC + v, i, j, k)] = dtP[v];
    }
  }

  // The tangent of one owned point's right-hand side along (work_dot, the
  // parameter tangents pd = d(gamma0_near, gamma0_far, gamma2, rho_max,
  // rho_ramp)) by the dual pass: the branches of a hole at rest (none, the
  // frozen core of :damped / :frozen, the layer w F - rho (u - u_e)).
  void rhs_point_jvp(long b, long i, long j, long k, double t, const double pd[5], double* du,
                     double* du_dot) const {
    const Mesh& m = mesh;
    double x[3];
    m.position(b, i, j, k, x);
    D1 g0;
    if (cs.damping == DAMP_CONSTANT) {
      g0 = D1(cs.g0);
    } else {
      double cc[3];
      center_at(cs, t, cc);
      const double d1 = x[0] - cc[0], d2 = x[1] - cc[1], d3 = x[2] - cc[2];
      const double e = std::exp(-(d1 * d1 + d2 * d2 + d3 * d3) / (2.0 * cs.g0_width * cs.g0_width));
      D1 nr(cs.g0_near), fr(cs.g0_far);
      nr.d[0] = pd[0];
      fr.d[0] = pd[1];
      g0 = fr + (nr - fr) * e;
    }
    D1 g2(cs.g2);
    g2.d[0] = pd[2];
    const D1 epsh(dissipation_rate(cs, t, x) * (1.0 / m.hb(b)));
    const bool diss = has_dissipation(cs);
    auto put = [&](int v, const D1& f) {
      du[m.sidx(b, v, i, j, k)] = f.v;
      du_dot[m.sidx(b, v, i, j, k)] = f.d[0];
    };
    V10<D1> dth, dtP;
    if (cs.interior != INT_NONE) {
      const IGeo g = interior_geo(cs, t, x);
      const double r = g.r;
      if (cs.interior == INT_PASTED || cs.interior == INT_FITTED || core_follows(cs) || cs.shaped)
        throw std::runtime_error("the tangent pass is built for a sphere at rest (:damped, :frozen)");
      if (r < g.r0) {
        for (int v = 0; v < 20; ++v) put(v, D1(0.0));
        return;
      }
      rhs_point_F<D1>(b, i, j, k, t, g0, g2, epsh, diss, dth, dtP);
      if (r >= g.r1) {
        for (int v = 0; v < NC; ++v) {
          put(v, dth[v]);
          put(NC + v, dtP[v]);
        }
        return;
      }
      const double sl = (r - g.r0) / (g.r1 - g.r0);
      const double w = smoothstep(sl / cs.w_ramp);
      D1 rmax(cs.rho_max), rramp(cs.rho_ramp);
      rmax.d[0] = pd[3];
      rramp.d[0] = pd[4];
      // rho = rho_max smoothstep((1 - s) / rho_ramp), the quintic's
      // derivative 30 u^2 (1 - u)^2 inside (0, 1) and zero where it clamps
      const D1 arg = (1.0 - sl) / rramp;
      const double u0 = std::min(std::max(arg.v, 0.0), 1.0);
      D1 ss(smoothstep(arg.v));
      ss.d[0] = (arg.v > 0.0 && arg.v < 1.0) ? 30.0 * u0 * u0 * (1.0 - u0) * (1.0 - u0) * arg.d[0] : 0.0;
      const D1 rho = cs.interior == INT_FROZEN ? D1(0.0) : rmax * ss;
      V10<double> he, Pe, dhe[3];
      background_state(cs.bg, t, x, he, Pe, dhe);
      const long var = m.widx(b, 0, i, j, k), sv = m.S * m.S * m.S;
      for (int v = 0; v < NC; ++v) {
        put(v, w * dth[v] - rho * (ld<D1>(var + v * sv) - he[v]));
        put(NC + v, w * dtP[v] - rho * (ld<D1>(var + (NC + v) * sv) - Pe[v]));
      }
      return;
    }
    rhs_point_F<D1>(b, i, j, k, t, g0, g2, epsh, diss, dth, dtP);
    for (int v = 0; v < NC; ++v) {
      put(v, dth[v]);
      put(NC + v, dtP[v]);
    }
  }
  // prepare! of the tangent: the fill's copies applied to the tangent state;
  // a Dirichlet ghost's tangent is zero (the hook does not depend on the
  // state or on the parameters). Uniform forests.
  void prepare_dot(const double* udot) {
    const Mesh& m = mesh;
    if (m.forest()) throw std::runtime_error("the tangent fill is built for uniform forests");
    work_dot.assign(work.size(), 0.0);
    for (long b = 0; b < m.nblocks(); ++b)
      for (int v = 0; v < 20; ++v)
        for (long k = 0; k <= m.N; ++k)
          for (long j = 0; j <= m.N; ++j)
            for (long i = 0; i <= m.N; ++i) work_dot[m.widx(b, v, i, j, k)] = udot[m.sidx(b, v, i, j, k)];
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = -m.G; k <= m.N + m.G; ++k)
        for (long j = -m.G; j <= m.N + m.G; ++j)
          for (long i = -m.G; i <= m.N + m.G; ++i) {
            long l[3] = {i, j, k}, owner = 0, loc[3] = {0, 0, 0};
            const int c = m.classify(b, l, owner, loc);
            if (c == -1) continue;
            for (int v = 0; v < 20; ++v)
              work_dot[m.widx(b, v, i, j, k)] = c == -2 ? 0.0 : work_dot[m.widx(owner, v, loc[0], loc[1], loc[2])];
          }
  }
  // gh_rhs! and its tangent along (udot, pd) in one dual pass.
  void gh_rhs_jvp(const double* u, const double* udot, double t, const double pd[5], double* du, double* du_dot,
                  int nthreads = 1) {
    if (bounds.on) throw std::runtime_error("the tangent pass is built for cases without range bounds");
    prepare(u, t);
    prepare_dot(udot);
    const Mesh& m = mesh;
    const long nb = m.nblocks();
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (long b = 0; b < nb; ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) rhs_point_jvp(b, i, j, k, t, pd, du, du_dot);
    work_dot.clear();
  }
  // One RK4 step of (u, udot): the tangent linear model of rk4_step.
  void rk4_step_jvp(double* u, double* udot, double t, double dt, const double pd[5], int nthreads = 1) {
    const long n = mesh.state_size();
    std::vector<double> k[4], kd[4], y(n), yd(n);
    for (int s = 0; s < 4; ++s) {
      k[s].assign(n, 0.0);
      kd[s].assign(n, 0.0);
    }
    const double c[4] = {0.0, dt / 2, dt / 2, dt};
    for (int s = 0; s < 4; ++s) {
      if (s == 0) {
        gh_rhs_jvp(u, udot, t, pd, k[0].data(), kd[0].data(), nthreads);
        continue;
      }
      for (long i = 0; i < n; ++i) {
        y[i] = u[i] + c[s] * k[s - 1][i];
        yd[i] = udot[i] + c[s] * kd[s - 1][i];
      }
      gh_rhs_jvp(y.data(), yd.data(), t + c[s], pd, k[s].data(), kd[s].data(), nthreads);
    }
    for (long i = 0; i < n; ++i) {
      u[i] = u[i] + (dt / 6) * (k[0][i] + 2 * k[1][i] + 2 * k[2][i] + k[3][i]);
      udot[i] = udot[i] + (dt / 6) * (kd[0][i] + 2 * kd[1][i] + 2 * kd[2][i] + kd[3][i]);
    }
  }

  // rhs! (evolution.jl:1386-1390): every owned point of every block.
  void rhs(double* du, double t, int nthreads = 1) const {
    const Mesh& m = mesh;
    // (checked here, outside the threads: a :fitted interior reads its cache)
    if (cs.interior == INT_FITTED && tw.empty())
      throw std::runtime_error("a :fitted interior needs its target cache (orc_set_target)");
    long nb = m.nblocks();
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (long b = 0; b < nb; ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) rhs_point(b, i, j, k, t, du);
  }
  void gh_rhs(double* du, const double* u, double t, int nthreads = 1) {
    prepare(u, t);
    rhs(du, t, nthreads);
  }

  // The one-expression assembly the reference's kernel test measures the
  // streaming kernel against (test/evolution_tests.jl:37-89): stencils from
  // apply_stencil and gh_node_rhs_expanded, plus the dissipation and the
  // source's correction for the dissipated d_t h.
  void host_rhs_at(long b, long i, long j, long k, double t, double out[20]) const {
    const Mesh& m = mesh;
    double x[3];
    m.position(b, i, j, k, x);
    double g0 = damping_rate(cs, t, x);
    double hsp = m.hb(b);
    auto at = [&](int v, long ii, long jj, long kk) { return work[m.widx(b, v, ii, jj, kk)]; };
    auto apply = [&](const std::vector<double>& w, int v, int d) {
      int n = (int)w.size(), r = (n - 1) / 2;
      double s = 0;
      for (int kk = 0; kk < n; ++kk) {
        long o = kk - r;
        double val_ = at(v, i + (d == 0) * o, j + (d == 1) * o, k + (d == 2) * o);
        s = (kk == 0) ? w[kk] * val_ : s + w[kk] * val_;
      }
      return s;
    };
    auto apply_mixed = [&](int v, int d1, int d2) {
      int n = (int)w1.size(), r = (n - 1) / 2;
      double outer = 0;
      for (int a = 0; a < n; ++a) {
        double inner = 0;
        for (int e = 0; e < n; ++e) {
          long oa = a - r, oe = e - r;
          long ii = i + (d1 == 0) * oa + (d2 == 0) * oe;
          long jj = j + (d1 == 1) * oa + (d2 == 1) * oe;
          long kk = k + (d1 == 2) * oa + (d2 == 2) * oe;
          double t_ = w1[e] * at(v, ii, jj, kk);
          inner = (e == 0) ? t_ : inner + t_;
        }
        double t_ = w1[a] * inner;
        outer = (a == 0) ? t_ : outer + t_;
      }
      return outer;
    };
    V10<double> hv, Pv, dh[3], dP[3], ddh[6];
    for (int v = 0; v < NC; ++v) {
      hv[v] = at(v, i, j, k);
      Pv[v] = at(NC + v, i, j, k);
      for (int d = 0; d < 3; ++d) {
        dh[d][v] = apply(w1, v, d) / hsp;
        dP[d][v] = apply(w1, NC + v, d) / hsp;
      }
    }
    const int pairs[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int n = 0; n < 6; ++n)
      for (int v = 0; v < NC; ++v) {
        int d1 = pairs[n][0], d2 = pairs[n][1];
        ddh[n][v] = d1 == d2 ? apply(w2, v, d1) / (hsp * hsp) : apply_mixed(v, d1, d2) / (hsp * hsp);
      }
    double Hl[4] = {0, 0, 0, 0}, dHl[4][4] = {{0}};
    if (cs.gauge == GAUGE_SAMPLED) {
      for (int a = 0; a < 4; ++a) Hl[a] = Hsrc[m.sidx(b, a, i, j, k)];
      for (int n = 0; n < 16; ++n) dHl[n % 4][n / 4] = Hsrc[m.sidx(b, 4 + n, i, j, k)];
    }
    V10<double> dth, dtP;
    if (cs.gauge == GAUGE_KS) {
      // the algebraic source reads d_t h, which the expanded form forms inside;
      // evaluate it from the undissipated d_t h first, then correct below
      MQ<double> mq0 = metric_quantities(hv);
      V10<double> dth0;
      for (int n = 0; n < NC; ++n)
        dth0[n] = mq0.beta[0] * dh[0][n] + mq0.beta[1] * dh[1][n] + mq0.beta[2] * dh[2][n] +
                  (mq0.alpha / mq0.sqrtg) * Pv[n];
      ks_algebraic_source<double>(cs.bg.M, cs.ks_S, cs.ks_u, hv, dth0, dh, Hl, dHl);
    }
    gh_node_rhs_expanded(hv, Pv, dh, dP, ddh, Hl, dHl, g0, cs.g2, dth, dtP);
    const double eps = dissipation_rate(cs, t, x);
    const bool diss = has_dissipation(cs);
    if (diss || cs.gauge == GAUGE_KS) {
      V10<double> Qh, QP;
      for (int v = 0; v < NC; ++v) {
        Qh[v] = !diss ? 0.0 : (eps / hsp) * ((apply(wD, v, 0) + apply(wD, v, 1)) + apply(wD, v, 2));
        QP[v] = !diss ? 0.0
                      : (eps / hsp) * ((apply(wD, NC + v, 0) + apply(wD, NC + v, 1)) + apply(wD, NC + v, 2));
      }
      MQ<double> mq = metric_quantities(hv);
      V10<double> dthQ;
      for (int v = 0; v < NC; ++v) dthQ[v] = dth[v] + Qh[v];
      double Hl2[4], dHl2[4][4];
      for (int a = 0; a < 4; ++a) {
        Hl2[a] = Hl[a];
        for (int c = 0; c < 4; ++c) dHl2[a][c] = dHl[a][c];
      }
      if (cs.gauge == GAUGE_KS) ks_algebraic_source<double>(cs.bg.M, cs.ks_S, cs.ks_u, hv, dthQ, dh, Hl2, dHl2);
      double dgA[4][4][4], dgB[4][4][4];
      dg4(dthQ, dh, dgA);
      dg4(dth, dh, dgB);
      V10<double> sA = gh_node_source(mq, dgA, Hl2, dHl2, g0, cs.g2);
      V10<double> sB = gh_node_source(mq, dgB, Hl, dHl, g0, cs.g2);
      for (int v = 0; v < NC; ++v) {
        dth[v] = dthQ[v];
        dtP[v] = dtP[v] + QP[v] + (sA[v] - sB[v]);
      }
    }
    for (int v = 0; v < NC; ++v) {
      out[v] = dth[v];
      out[NC + v] = dtP[v];
    }
  }
  // rhs_scale (test/evolution_tests.jl:97-108).
  double rhs_scale(long b, long i, long j, long k) const {
    const Mesh& m = mesh;
    double s = 0;
    int n = (int)w2.size(), r = (n - 1) / 2;
    for (int v = 0; v < 20; ++v)
      for (int d = 0; d < 3; ++d) {
        double acc = 0;
        for (int kk = 0; kk < n; ++kk) {
          long o = kk - r;
          double t_ = w2[kk] * work[m.widx(b, v, i + (d == 0) * o, j + (d == 1) * o, k + (d == 2) * o)];
          acc = kk == 0 ? t_ : acc + t_;
        }
        s = std::max(s, std::abs(acc) / (m.hb(b) * m.hb(b)));
        s = std::max(s, std::abs(work[m.widx(b, v, i, j, k)]));
      }
    return std::max(s, 1.0);
  }

  // ---- the time step and the analysis reductions ----
  bool is_evolved(double t, const double x[3]) const {
    if (cs.interior == INT_NONE) return true;
    const IGeo g = interior_geo(cs, t, x);
    return g.r >= g.r1;
  }
  // The :fitted target cache (evolution.jl:1007-1021): 40 variables per owned
  // point, A then S, in the state layout; t_f its fill time.
  std::vector<double> tw;
  double t_f = 0.0, trail = 0.0;
  bool rate = false;
  long tidx(long b, long i, long j, long k) const {
    const Mesh& m = mesh;
    return tw.empty() ? 0 : i + m.P * (j + m.P * (k + m.P * (40 * b)));
  }
  // max_speed (evolution.jl:932-949,1539-1560) over evolved owned points.
  double max_speed(const double* u, double t = 0.0) const {
    const Mesh& m = mesh;
    double lam = 0;
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            double x[3];
            m.position(b, i, j, k, x);
            if (!is_evolved(t, x)) continue;
            V10<double> hv;
            for (int v = 0; v < NC; ++v) hv[v] = u[m.sidx(b, v, i, j, k)];
            MQ<double> mq = metric_quantities(hv);
            double l = mq.alpha * std::sqrt(mq.gamu[0][0] + mq.gamu[1][1] + mq.gamu[2][2]) +
                       std::sqrt(mq.beta[0] * mq.beta[0] + mq.beta[1] * mq.beta[1] + mq.beta[2] * mq.beta[2]);
            if (!(l <= lam)) lam = l;  // NaN propagates
          }
    return lam;
  }
  // gh_error_kernel! + error_norms (constraints.jl:311-396, 629-671): at
  // every primary the error e = |u - u_exact(t)|, the 2-norm over the 20
  // variables summed in variable order, where the point is evolved; the
  // masked L2 sqrt(sum w e^2 / sum w) with w the point's volume relative to
  // level 0's (2^(-3 level)), the masked L-inf the largest e; the layer
  // residual, the largest e at r_0 <= r < r_1 (the distance from the target
  // cache A + (t - t_f) S for :fitted); the drift, the largest |h_tt -
  // h_tt,exact| at evolved points with shell_lo <= r <= shell_hi.
  void error_rows(const double* u, double t, double shell_lo, double shell_hi, double out[4]) const {
    const Mesh& m = mesh;
    double num = 0, den = 0, linf = 0, res = 0, drift = 0;
    const bool fitted = cs.interior == INT_FITTED && !tw.empty();
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            if (!m.is_primary(b, i, j, k)) continue;
            double x[3], s[20];
            m.position(b, i, j, k, x);
            const bool keep = is_evolved(t, x);
            bool lay = false;
            double r = 0;
            if (cs.interior != INT_NONE) {
              const IGeo g = interior_geo(cs, t, x);
              lay = g.r0 <= g.r && g.r < g.r1;
              r = g.r;
            }
            const bool inshell = keep && shell_lo <= r && r <= shell_hi;
            if (!(keep || lay)) continue;
            case_state(cs, t, x, s);
            double e2 = 0;
            for (int v = 0; v < 20; ++v) {
              const double e = u[m.sidx(b, v, i, j, k)] - s[v];
              e2 += e * e;
            }
            const double e = std::sqrt(e2);
            if (keep) {
              const double wv = m.forest() ? 1.0 / (double)(1L << (3 * m.level[b])) : 1.0;
              num += m.forest() ? e2 * wv : e2;
              den += wv;
              if (!(e <= linf)) linf = e;  // NaN propagates
            }
            if (lay) {
              double ef = e;
              if (fitted) {
                double f2 = 0;
                for (int v = 0; v < 20; ++v) {
                  const long ti = tidx(b, i, j, k);
                  const double tg = tw[ti + (long)v * m.P * m.P * m.P] + (t - t_f) * tw[ti + (long)(20 + v) * m.P * m.P * m.P];
                  const double d = u[m.sidx(b, v, i, j, k)] - tg;
                  f2 += d * d;
                }
                ef = std::sqrt(f2);
              }
              if (!(ef <= res)) res = ef;
            }
            if (inshell) {
              const double d1 = std::abs(u[m.sidx(b, 0, i, j, k)] - s[0]);
              if (!(d1 <= drift)) drift = d1;
            }
          }
    out[0] = den > 0 ? std::sqrt(num / den) : 0.0;
    out[1] = linf;
    out[2] = res;
    out[3] = drift;
  }
  // The directional derivative of the masked L2 error along udot:
  // dJ = sum w (u - u_exact) . udot / (J sum w) over the evolved primaries.
  double error_tangent(const double* u, const double* udot, double t) const {
    const Mesh& m = mesh;
    double num = 0, den = 0, dnum = 0;
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            if (!m.is_primary(b, i, j, k)) continue;
            double x[3], s[20];
            m.position(b, i, j, k, x);
            if (!is_evolved(t, x)) continue;
            case_state(cs, t, x, s);
            const double wv = m.forest() ? 1.0 / (double)(1L << (3 * m.level[b])) : 1.0;
            for (int v = 0; v < 20; ++v) {
              const double e = u[m.sidx(b, v, i, j, k)] - s[v];
              num += wv * e * e;
              dnum += wv * e * udot[m.sidx(b, v, i, j, k)];
            }
            den += wv;
          }
    const double J = den > 0 ? std::sqrt(num / den) : 0.0;
    return J > 0 ? dnum / (J * den) : 0.0;
  }
  void error_norms(const double* u, double t, double& l2, double& linf) const {
    double r[4];
    error_rows(u, t, 0.0, -1.0, r);
    l2 = r[0];
    linf = r[1];
  }

  // apply_bounds! (bounds.jl:613-660, 733-770): the range projection at every
  // owned point with r < r_gate, written back only where it fired; returns
  // the number of points moved (a no-op without bounds).
  long apply_bounds(double* u, double t = 0.0) const {
    if (!bounds.on) return 0;
    const Mesh& m = mesh;
    long hits = 0;
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            double x[3];
            m.position(b, i, j, k, x);
            if (!(interior_radius(cs, t, x) < bounds.r_gate)) continue;
            double h[10], P[10], ho[10], Po[10];
            for (int v = 0; v < 10; ++v) {
              h[v] = u[m.sidx(b, v, i, j, k)];
              P[v] = u[m.sidx(b, 10 + v, i, j, k)];
            }
            bool nf;
            if (bounds_project(h, P, bounds, ho, Po, nf)) {
              ++hits;
              for (int v = 0; v < 10; ++v) {
                u[m.sidx(b, v, i, j, k)] = ho[v];
                u[m.sidx(b, 10 + v, i, j, k)] = Po[v];
              }
            }
          }
    return hits;
  }

  // gh_limiter! (stepping.jl, bounds.jl:780-800, evolution.jl:888-905): the
  // range projection, then the :pasted overwrite, which has the last word.
  void stage_limit(double* u, double t) const {
    apply_bounds(u, t);
    paste(u, t);
  }
  void paste(double* u, double t) const {
    if (cs.interior != INT_PASTED) return;
    const Mesh& m = mesh;
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            double x[3], s[20];
            m.position(b, i, j, k, x);
            if (interior_radius(cs, t, x) >= cs.r1) continue;
            case_state(cs, t, x, s);
            for (int v = 0; v < 20; ++v) u[m.sidx(b, v, i, j, k)] = s[v];
          }
  }

  // gh_constraint! + constraint_norms' gauge rows (constraints.jl:211-309,
  // 559-566, 848-876): C_a = g_ab (Gamma^b + g^{bc} H_c) at every evolved owned
  // point, Gamma^b = g^{cd} Gamma^b_cd (gauge_constraint_at_node,
  // pointwise.jl:947-960), d_t h from the state without dissipation, H the
  // case's source; masked L2 and L-inf per component over the points
  // error_norms counts.
  void gauge_constraint(const double* u, double t, double l2[4], double linf[4]) {
    prepare(u, t);
    const Mesh& m = mesh;
    double num[4] = {0, 0, 0, 0}, den = 0;
    for (int a = 0; a < 4; ++a) linf[a] = 0;
    const long sv = m.S * m.S * m.S;
    const long st[3] = {1, m.S, m.S * m.S};
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            double x[3];
            m.position(b, i, j, k, x);
            if (!is_evolved(t, x)) continue;
            const bool prim = m.is_primary(b, i, j, k);
            if (m.forest() && !prim) continue;
            const long var = m.widx(b, 0, i, j, k);
            const double inv_h = 1.0 / m.hb(b);
            V10<double> hv, Pv, dh[3], dth;
            for (int v = 0; v < NC; ++v) {
              hv[v] = work[var + v * sv];
              Pv[v] = work[var + (NC + v) * sv];
            }
            for (int d = 0; d < 3; ++d)
              for (int v = 0; v < NC; ++v) dh[d][v] = inv_h * axis_stencil<double>(w1, var + v * sv, st[d]);
            MQ<double> mq = metric_quantities(hv);
            for (int n = 0; n < NC; ++n)
              dth[n] = mq.beta[0] * dh[0][n] + mq.beta[1] * dh[1][n] + mq.beta[2] * dh[2][n] +
                       (mq.alpha / mq.sqrtg) * Pv[n];
            double Hl[4] = {0, 0, 0, 0}, dHl[4][4] = {{0}};
            if (cs.gauge == GAUGE_SAMPLED)
              for (int a = 0; a < 4; ++a) Hl[a] = Hsrc[m.sidx(b, a, i, j, k)];
            else if (cs.gauge == GAUGE_KS)
              ks_algebraic_source<double>(cs.bg.M, cs.ks_S, cs.ks_u, hv, dth, dh, Hl, dHl);
            double dg[4][4][4];  // dg[c][a][b] = d_c g_ab
            dg4(dth, dh, dg);
            // Gamma_{c a b} = (d_b g_ca + d_a g_cb - d_c g_ab) / 2, then
            // Gamma^m_ab = g^{mc} Gamma_cab and Gamma^m = g^{ab} Gamma^m_ab
            double Gl[4][4][4], Gm[4][4][4], Gup[4], Hup[4], Cu[4];
            for (int c = 0; c < 4; ++c)
              for (int a = 0; a < 4; ++a)
                for (int bb = 0; bb < 4; ++bb) Gl[c][a][bb] = (dg[bb][c][a] + dg[a][c][bb] - dg[c][a][bb]) / 2;
            for (int mu = 0; mu < 4; ++mu)
              for (int a = 0; a < 4; ++a)
                for (int bb = 0; bb < 4; ++bb) {
                  double s = 0;
                  for (int c = 0; c < 4; ++c) s += mq.gu4[mu][c] * Gl[c][a][bb];
                  Gm[mu][a][bb] = s;
                }
            for (int mu = 0; mu < 4; ++mu) {
              double s = 0;
              for (int a = 0; a < 4; ++a)
                for (int bb = 0; bb < 4; ++bb) s += mq.gu4[a][bb] * Gm[mu][a][bb];
              Gup[mu] = s;
              double hu = 0;
              for (int c = 0; c < 4; ++c) hu += mq.gu4[mu][c] * Hl[c];
              Hup[mu] = hu;
              Cu[mu] = Gup[mu] + Hup[mu];
            }
            const double wv = m.forest() ? 1.0 / (double)(1L << (3 * m.level[b])) : 1.0;
            for (int a = 0; a < 4; ++a) {
              double C = 0;
              for (int mu = 0; mu < 4; ++mu) C += mq.g4[a][mu] * Cu[mu];
              linf[a] = std::max(linf[a], std::abs(C));
              if (prim) num[a] += m.forest() ? C * C * wv : C * C;
            }
            if (prim) den += wv;
          }
    for (int a = 0; a < 4; ++a) l2[a] = den > 0 ? std::sqrt(num[a] / den) : 0.0;
  }

  // adm_constraint! + constraint_norms' Hamiltonian and momentum rows
  // (constraints.jl:100-210,432-534,584-605,848-876): at every evolved owned
  // point the time derivatives of g from the evolution equations (d_t g and
  // d_t Pi from gh_node_rhs_expanded without dissipation, d_i d_t g and
  // d_t d_t g by differentiating the first equation, metric_derivatives_along,
  // pointwise.jl:530-545), the textbook Ricci tensor of the 4-metric and its
  // projections H = 2 R_nn + R, M_i = -n^a (R_ai - g_ai R / 2). ham = {L2, Linf},
  // mom_l2 / mom_linf per component, counted as error_norms counts.
  static void along(const MQ<double>& m, const V10<double>& dgv, double& da, double db[3], double& dsq) {
    M4<double> dgl = sym4(dgv);
    M4<double> t = matmul(matmul(m.gu4, dgl), m.gu4);
    const double dgu00 = -t[0][0];
    da = m.alpha * m.alpha * m.alpha * dgu00 / 2.0;
    for (int j = 0; j < 3; ++j) db[j] = -(-t[0][j + 1] + m.beta[j] * dgu00) / m.gu4[0][0];
    double s = 0;
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) s += m.gamu[j][k] * dgl[j + 1][k + 1];
    dsq = m.sqrtg * s / 2.0;
  }
  static void adm_at_node(const MQ<double>& m, const double dg[4][4][4], const double ddg[4][4][4][4],
                          double& H, double Mo[3]) {
    // dg[a][b][c] = d_a g_bc, ddg[a][b][c][d] = d_a d_b g_cd
    double dgu[4][4][4];
    for (int e = 0; e < 4; ++e) {
      M4<double> de;
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) de[a][b] = dg[e][a][b];
      M4<double> t = matmul(matmul(m.gu4, de), m.gu4);
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) dgu[e][a][b] = -t[a][b];
    }
    double Gl[4][4][4], G4[4][4][4], Gtr[4];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        for (int c = 0; c < 4; ++c) Gl[a][b][c] = (dg[b][a][c] + dg[c][a][b] - dg[a][b][c]) / 2;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        for (int c = 0; c < 4; ++c) {
          double s = 0;
          for (int x = 0; x < 4; ++x) s += m.gu4[a][x] * Gl[x][b][c];
          G4[a][b][c] = s;
        }
    for (int d = 0; d < 4; ++d) {
      double s = 0;
      for (int c = 0; c < 4; ++c) s += G4[c][c][d];
      Gtr[d] = s;
    }
    double R[4][4];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double A = 0, B = 0, C = 0, D = 0;
        for (int c = 0; c < 4; ++c)
          for (int d = 0; d < 4; ++d) {
            A += dgu[c][c][d] * Gl[d][a][b] + m.gu4[c][d] * (ddg[c][a][d][b] + ddg[c][b][d][a] - ddg[c][d][a][b]) / 2;
            B += dgu[b][c][d] * Gl[d][c][a] + m.gu4[c][d] * (ddg[b][c][d][a] + ddg[b][a][d][c] - ddg[b][d][c][a]) / 2;
            D += G4[c][b][d] * G4[d][c][a];
          }
        for (int d = 0; d < 4; ++d) C += Gtr[d] * G4[d][a][b];
        R[a][b] = A - B + C - D;
      }
    double Rs = 0;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) Rs += m.gu4[a][b] * R[a][b];
    const double n[4] = {1.0 / m.alpha, -m.beta[0] / m.alpha, -m.beta[1] / m.alpha, -m.beta[2] / m.alpha};
    double Rnn = 0;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) Rnn += R[a][b] * n[a] * n[b];
    H = 2 * Rnn + Rs;
    for (int i = 0; i < 3; ++i) {
      double s = 0;
      for (int a = 0; a < 4; ++a) s += n[a] * (R[a][i + 1] - m.g4[a][i + 1] * Rs / 2);
      Mo[i] = -s;
    }
  }
  void adm_constraint(const double* u, double t, double ham[2], double mom_l2[3], double mom_linf[3]) {
    prepare(u, t);
    const Mesh& m = mesh;
    double num[4] = {0, 0, 0, 0}, den = 0, mx[4] = {0, 0, 0, 0};
    const long sv = m.S * m.S * m.S;
    const long st[3] = {1, m.S, m.S * m.S};
    const int pr[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    auto pair3 = [](int i, int j) {
      const int a = i < j ? i : j, b = i < j ? j : i;
      return a == 0 ? b : (a == 1 ? 2 + b : 5);
    };
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            double x[3];
            m.position(b, i, j, k, x);
            if (!is_evolved(t, x)) continue;
            const bool prim = m.is_primary(b, i, j, k);
            if (m.forest() && !prim) continue;
            const long var = m.widx(b, 0, i, j, k);
            const double inv_h = 1.0 / m.hb(b), inv_h2 = inv_h * inv_h;
            V10<double> hv, Pv, dh[3], dP[3], ddh[6];
            for (int v = 0; v < NC; ++v) {
              hv[v] = work[var + v * sv];
              Pv[v] = work[var + (NC + v) * sv];
            }
            for (int d = 0; d < 3; ++d)
              for (int v = 0; v < NC; ++v) {
                dh[d][v] = inv_h * axis_stencil<double>(w1, var + v * sv, st[d]);
                dP[d][v] = inv_h * axis_stencil<double>(w1, var + (NC + v) * sv, st[d]);
              }
            for (int n = 0; n < 6; ++n)
              for (int v = 0; v < NC; ++v) {
                const int d1 = pr[n][0], d2 = pr[n][1];
                ddh[n][v] = inv_h2 * (d1 == d2 ? axis_stencil<double>(w2, var + v * sv, st[d1])
                                               : mixed_stencil<double>(w1, var + v * sv, st[d1], st[d2]));
              }
            MQ<double> mq = metric_quantities(hv);
            V10<double> dth0;
            for (int n = 0; n < NC; ++n)
              dth0[n] = mq.beta[0] * dh[0][n] + mq.beta[1] * dh[1][n] + mq.beta[2] * dh[2][n] +
                        (mq.alpha / mq.sqrtg) * Pv[n];
            double Hl[4] = {0, 0, 0, 0}, dHl[4][4] = {{0}};
            if (cs.gauge == GAUGE_SAMPLED) {
              for (int a = 0; a < 4; ++a) Hl[a] = Hsrc[m.sidx(b, a, i, j, k)];
              for (int n = 0; n < 16; ++n) dHl[n % 4][n / 4] = Hsrc[m.sidx(b, 4 + n, i, j, k)];
            } else if (cs.gauge == GAUGE_KS) {
              ks_algebraic_source<double>(cs.bg.M, cs.ks_S, cs.ks_u, hv, dth0, dh, Hl, dHl);
            }
            V10<double> dth, dtP;
            gh_node_rhs_expanded(hv, Pv, dh, dP, ddh, Hl, dHl, damping_rate(cs, t, x), cs.g2, dth, dtP);
            const double phi = mq.alpha / mq.sqrtg;
            V10<double> dtd[3], dtt;
            for (int ii = 0; ii < 3; ++ii) {
              double da, db[3], dsq;
              along(mq, dh[ii], da, db, dsq);
              const double dphi = (da - phi * dsq) / mq.sqrtg;
              for (int v = 0; v < NC; ++v)
                dtd[ii][v] = (dphi * Pv[v] + phi * dP[ii][v]) + (db[0] * dh[0][v] + db[1] * dh[1][v] + db[2] * dh[2][v]) +
                             (mq.beta[0] * ddh[pair3(ii, 0)][v] + mq.beta[1] * ddh[pair3(ii, 1)][v] +
                              mq.beta[2] * ddh[pair3(ii, 2)][v]);
            }
            {
              double da, db[3], dsq;
              along(mq, dth, da, db, dsq);
              const double dphi = (da - phi * dsq) / mq.sqrtg;
              for (int v = 0; v < NC; ++v)
                dtt[v] = dphi * Pv[v] + phi * dtP[v] + (db[0] * dh[0][v] + db[1] * dh[1][v] + db[2] * dh[2][v]) +
                         (mq.beta[0] * dtd[0][v] + mq.beta[1] * dtd[1][v] + mq.beta[2] * dtd[2][v]);
            }
            double dg[4][4][4], ddg[4][4][4][4];
            dg4(dth, dh, dg);
            for (int mu = 0; mu < 4; ++mu)
              for (int nu = 0; nu < 4; ++nu) {
                const V10<double>* src;
                if (mu == 0 && nu == 0) src = &dtt;
                else if (mu == 0) src = &dtd[nu - 1];
                else if (nu == 0) src = &dtd[mu - 1];
                else src = &ddh[pair3(mu - 1, nu - 1)];
                for (int a = 0; a < 4; ++a)
                  for (int c = 0; c < 4; ++c) ddg[mu][nu][a][c] = (*src)[pairindex(a, c)];
              }
            double H, Mo[3];
            adm_at_node(mq, dg, ddg, H, Mo);
            const double wv = m.forest() ? 1.0 / (double)(1L << (3 * m.level[b])) : 1.0;
            const double vals[4] = {H, Mo[0], Mo[1], Mo[2]};
            for (int q = 0; q < 4; ++q) {
              mx[q] = std::max(mx[q], std::abs(vals[q]));
              if (prim) num[q] += m.forest() ? vals[q] * vals[q] * wv : vals[q] * vals[q];
            }
            if (prim) den += wv;
          }
    ham[0] = den > 0 ? std::sqrt(num[0] / den) : 0.0;
    ham[1] = mx[0];
    for (int q = 0; q < 3; ++q) {
      mom_l2[q] = den > 0 ? std::sqrt(num[1 + q] / den) : 0.0;
      mom_linf[q] = mx[1 + q];
    }
  }

  // The refinement indicator (refinement.jl:152-183 lohner, :184-219
  // cell_tau, :312-365 gh_tau_kernel!, :367-378 tau_max): at every evolved
  // owned point the largest Loehner ratio |u+ - 2u0 + u-| / (|u+ - u0| +
  // |u0 - u-| + 4 eps |scale|) over the ten h components and the three axes;
  // per block the largest over its points (0 where none is evolved).
  // field_scale (refinement.jl:221-300): the largest |h_v| over the evolved
  // owned points.
  static double lohner(double um, double u0, double up, double scale, double eps) {
    const double num = std::abs(up - 2 * u0 + um);
    const double den = std::abs(up - u0) + std::abs(u0 - um) + 4 * eps * std::abs(scale);
    return den == 0.0 ? 0.0 : num / den;
  }
  void tau_blocks(const double* u, double t, double scale, double eps, double* out) {
    prepare(u, t);
    const Mesh& m = mesh;
    const long sv = m.S * m.S * m.S;
    const long st[3] = {1, m.S, m.S * m.S};
    for (long b = 0; b < m.nblocks(); ++b) {
      double tb = 0.0;
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            double x[3];
            m.position(b, i, j, k, x);
            if (!is_evolved(t, x)) continue;
            const long base = m.widx(b, 0, i, j, k);
            double tau = 0.0;
            for (int v = 0; v < NC; ++v) {
              const long i0 = base + v * sv;
              for (int d = 0; d < 3; ++d)
                tau = std::max(tau, lohner(work[i0 - st[d]], work[i0], work[i0 + st[d]], scale, eps));
            }
            tb = std::max(tb, tau);
          }
      out[b] = tb;
    }
  }
  double field_scale(const double* u, double t) const {
    const Mesh& m = mesh;
    double s = 0.0;
    for (long b = 0; b < m.nblocks(); ++b)
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            double x[3];
            m.position(b, i, j, k, x);
            if (!is_evolved(t, x)) continue;
            for (int v = 0; v < NC; ++v) s = std::max(s, std::abs(u[m.sidx(b, v, i, j, k)]));
          }
    return s;
  }

  // The regrid's transfer onto this problem's forest from another forest of
  // the same root brick (refine!/coarsen! then the state carried over,
  // src/refinement.jl:973-1021 with TreeAMR's transfer): every owned point of
  // this forest takes its value as the other forest's fill would give it --
  // a copy of the other forest's owner of the point (an unchanged leaf, or
  // injection from a finer one), or the order-p prolongation from the leaf
  // one level coarser that holds it -- from the other problem's work array
  // prepared at t.
  void transfer_from(Problem& from, const double* u_from, double t, double* u_new) {
    const Mesh& m = mesh;
    const Mesh& o = from.mesh;
    if (!m.forest() || !o.forest()) throw std::runtime_error("the transfer runs between forests with levels");
    from.prepare(u_from, t);
    const long S = o.S;
    for (long b = 0; b < m.nblocks(); ++b) {
      const int L = m.level[b];
      for (long k = 0; k <= m.N; ++k)
        for (long j = 0; j <= m.N; ++j)
          for (long i = 0; i <= m.N; ++i) {
            const long l[3] = {i, j, k};
            long gl[3];
            for (int d = 0; d < 3; ++d) {
              const long tot = ((long)m.nb[d] << L) * m.N;
              gl[d] = m.blocks[b][d] * (long)m.N + l[d];
              if (m.periodic[d]) gl[d] = ((gl[d] % tot) + tot) % tot;
            }
            Mesh::Src sr;
            const int c = o.classify_point(L, gl, sr);
            for (int v = 0; v < 20; ++v) {
              double val;
              if (c == Mesh::F_COPY) {
                val = from.work[o.widx(sr.b, v, sr.loc[0], sr.loc[1], sr.loc[2])];
              } else {
                const int n[3] = {sr.half[0] ? o.prolong : 1, sr.half[1] ? o.prolong : 1,
                                  sr.half[2] ? o.prolong : 1};
                const double* base = &from.work[o.widx(sr.b, v, sr.loc[0], sr.loc[1], sr.loc[2])];
                double sz = 0.0;
                for (int c2 = 0; c2 < n[2]; ++c2) {
                  double sy = 0.0;
                  for (int c1 = 0; c1 < n[1]; ++c1) {
                    double sx = 0.0;
                    for (int c0 = 0; c0 < n[0]; ++c0)
                      sx += (n[0] > 1 ? o.wmid[c0] : 1.0) * base[c0 + S * (c1 + S * c2)];
                    sy += (n[1] > 1 ? o.wmid[c1] : 1.0) * sx;
                  }
                  sz += (n[2] > 1 ? o.wmid[c2] : 1.0) * sy;
                }
                val = sz;
              }
              u_new[m.sidx(b, v, i, j, k)] = val;
            }
          }
    }
  }

  // One classical RK4 step with the limiter on each stage value and on the
  // result (stepping.jl:53-73, IMEXRungeKutta RK4).
  void rk4_step(double* u, double t, double dt, int nthreads = 1) {
    long n = mesh.state_size();
    std::vector<double> k1(n), k2(n), k3(n), k4(n), y(n);
    gh_rhs(k1.data(), u, t, nthreads);
    for (long i = 0; i < n; ++i) y[i] = u[i] + (dt / 2) * k1[i];
    stage_limit(y.data(), t + dt / 2);
    gh_rhs(k2.data(), y.data(), t + dt / 2, nthreads);
    for (long i = 0; i < n; ++i) y[i] = u[i] + (dt / 2) * k2[i];
    stage_limit(y.data(), t + dt / 2);
    gh_rhs(k3.data(), y.data(), t + dt / 2, nthreads);
    for (long i = 0; i < n; ++i) y[i] = u[i] + dt * k3[i];
    stage_limit(y.data(), t + dt);
    gh_rhs(k4.data(), y.data(), t + dt, nthreads);
    for (long i = 0; i < n; ++i) u[i] = u[i] + (dt / 6) * (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
    stage_limit(u, t + dt);
  }
  // settle! (evolution.jl:1370-1374): fill, then gather back.
  void settle(double* u, double t) {
    prepare(u, t);
    const Mesh& m = mesh;
    for (long b = 0; b < m.nblocks(); ++b)
      for (int v = 0; v < 20; ++v)
        for (long k = 0; k <= m.N; ++k)
          for (long j = 0; j <= m.N; ++j)
            for (long i = 0; i <= m.N; ++i) u[m.sidx(b, v, i, j, k)] = work[m.widx(b, v, i, j, k)];
  }
};

}  // namespace orc
