// A frozen hierarchy of closed vertex blocks (TreeAMR 0.3's forest as
// gh_forest(; refined) and hole_forest build it, src/initialdata.jl:1370-1515):
// the host-side checks of a leaf list and the plan of its ghost fill, which
// k_fill_plan executes.
//
// A leaf (l, i, j, k) is block (i, j, k) of the brick of level l: nb_d 2^l
// blocks of N cells of spacing h / 2^l along axis d. The rules of the fill:
//  * every point has an owner, the finest leaf that holds it as a vertex;
//    among leaves of one level, the one where the point's local index is
//    below N on the most significant axes (the uniform rule: an axis where
//    it is N ranks after one where it is not, x lowest);
//  * a work point that is not its block's primary takes the owner's value:
//    a same-level copy, an injection from a finer leaf, a coincident coarse
//    vertex; an owned face point of a coarse leaf included;
//  * a point that no leaf holds as a vertex (a coarse midpoint on one or
//    more axes) is the order-p Lagrange interpolation of the level-(l-1)
//    leaf that holds it, p points centred on the midpoint along each such
//    axis, tensor product;
//  * outside a non-periodic axis, the Dirichlet hook.
// With half-open ownership (TreeAMR 0.1) a leaf owns [0, N) per axis
// instead: a point's owner is the one leaf whose half-open extent holds it,
// when the point is a vertex of that leaf (else it is prolongated as
// above), and the upper plane of a non-periodic axis is a Dirichlet ghost.
// The checks: the leaves tile the box exactly, neighbours (faces, edges and
// corners, across periodic seams) differ by at most one level, N is even and
// at least 2G, and the prolongation order is even and fits the ghosts.
#pragma once
#include <algorithm>
#include <cstdint>
#include <exception>
#include <string>
#include <thread>
#include <vector>

namespace tgh {

// Runs fn(chunk, b0, b1) over contiguous block ranges [b0, b1) on up to 16
// host threads; chunk c covers the c-th range, so per-chunk outputs joined in
// chunk order are the serial loop's. An exception in any chunk (or in
// starting a thread) is rethrown on the calling thread after every thread
// started has been joined.
template <class F>
inline int for_block_chunks(int nl, F fn) {
  unsigned hw = std::thread::hardware_concurrency();
  const int nt = (int)std::max(1u, std::min(16u, hw ? hw : 1u));
  const int nchunk = std::max(1, std::min(nt, nl / 8));
  std::vector<std::exception_ptr> err(nchunk);
  auto run = [&](int c) {
    try {
      fn(c, (int)((long)nl * c / nchunk), (int)((long)nl * (c + 1) / nchunk));
    } catch (...) {
      err[c] = std::current_exception();
    }
  };
  std::vector<std::thread> pool;
  std::exception_ptr start_err;
  try {
    pool.reserve(nchunk);
    for (int c = 1; c < nchunk; ++c) pool.emplace_back(run, c);
  } catch (...) {
    start_err = std::current_exception();
  }
  if (!start_err) run(0);
  for (auto& t : pool) t.join();
  if (start_err) std::rethrow_exception(start_err);
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
  return nchunk;
}

struct ForestPlan {
  std::vector<int> ops;            // 4 ints per entry: dst, src, mode, 0
  std::vector<int> pass_off;       // entry offsets of the passes (Lmax + 2 values)
  std::vector<unsigned char> prim; // per owned point: 1 if its block's primary
  int Lmax = 0;
  int dir_begin = 0, dir_end = 0;  // pass 0's Dirichlet entries (its tail)
};

// A checked leaf list by position: the cell map of its finest level and the
// rules above that name where a point takes its value from.
struct ForestIndex {
  int N = 0, Lmax = 0;
  bool half_open = false;
  int nb[3] = {0, 0, 0}, periodic[3] = {0, 0, 0};
  long nc[3] = {0, 0, 0};
  std::vector<int> lev, bc, cell;

  // Checks that the leaves tile the box with 2:1 balance; returns an error
  // message or "".
  std::string build(int N_, const int nb_[3], const int periodic_[3], const std::vector<int>& lev_,
                    const std::vector<int>& bc_) {
    N = N_;
    lev = lev_;
    bc = bc_;
    for (int d = 0; d < 3; ++d) {
      nb[d] = nb_[d];
      periodic[d] = periodic_[d];
    }
    const int nl = (int)lev.size();
    if (nl < 1) return "a forest needs at least one leaf";
    Lmax = 0;
    for (int b = 0; b < nl; ++b) {
      if (lev[b] < 0 || lev[b] > 10) return "leaf " + std::to_string(b) + " has level " + std::to_string(lev[b]) +
                                            ", outside 0 .. 10";
      Lmax = lev[b] > Lmax ? lev[b] : Lmax;
      for (int d = 0; d < 3; ++d)
        if (bc[3 * b + d] < 0 || bc[3 * b + d] >= (nb[d] << lev[b]))
          return "leaf " + std::to_string(b) + " lies outside the brick of its level";
    }
    // The cell map: the leaf of every block of the finest level.
    for (int d = 0; d < 3; ++d) nc[d] = (long)nb[d] << Lmax;
    const long ncell = nc[0] * nc[1] * nc[2];
    if (ncell > (1L << 26)) return "the forest's finest brick has too many blocks for the host's cell map";
    cell.assign(ncell, -1);
    for (int b = 0; b < nl; ++b) {
      const long f = 1L << (Lmax - lev[b]);
      const long c0[3] = {bc[3 * b] * f, bc[3 * b + 1] * f, bc[3 * b + 2] * f};
      for (long z = c0[2]; z < c0[2] + f; ++z)
        for (long y = c0[1]; y < c0[1] + f; ++y)
          for (long x = c0[0]; x < c0[0] + f; ++x) {
            int& c = cell[x + nc[0] * (y + nc[1] * z)];
            if (c >= 0) return "leaves " + std::to_string(c) + " and " + std::to_string(b) + " overlap";
            c = b;
          }
    }
    for (long c = 0; c < ncell; ++c)
      if (cell[c] < 0) return "the leaves do not cover the box";
    // 2:1 balance over the 13 forward neighbour directions.
    for (long z = 0; z < nc[2]; ++z)
      for (long y = 0; y < nc[1]; ++y)
        for (long x = 0; x < nc[0]; ++x) {
          const int a = cell[x + nc[0] * (y + nc[1] * z)];
          for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dx = -1; dx <= 1; ++dx) {
                if (dz * 9 + dy * 3 + dx <= 0) continue;
                long q[3] = {x + dx, y + dy, z + dz};
                bool in = true;
                for (int d = 0; d < 3; ++d) {
                  if (q[d] < 0 || q[d] >= nc[d]) {
                    if (!periodic[d]) in = false;
                    q[d] = (q[d] + nc[d]) % nc[d];
                  }
                }
                if (!in) continue;
                const int o = cell[q[0] + nc[0] * (q[1] + nc[1] * q[2])];
                if (lev[a] - lev[o] > 1 || lev[o] - lev[a] > 1)
                  return "the forest is not 2:1 balanced: leaves " + std::to_string(a) + " (level " +
                         std::to_string(lev[a]) + ") and " + std::to_string(o) + " (level " +
                         std::to_string(lev[o]) + ") touch";
              }
        }
    return "";
  }

  // The leaves whose closed extent holds point X of level Lw >= Lmax: those
  // of the one or two finest blocks along each axis around it.
  int around(int Lw, const long X[3], int* cand) const {
    long lo[3], n[3];
    for (int d = 0; d < 3; ++d) {
      const long csz = (long)N << (Lw - Lmax), tot = nc[d] * csz;
      lo[d] = X[d] / csz;
      n[d] = 1;
      if (X[d] % csz == 0) {
        if (X[d] == tot) lo[d] = nc[d] - 1;  // the top of a non-periodic axis
        else if (X[d] > 0 || periodic[d]) { lo[d] -= 1; n[d] = 2; }
      }
    }
    int k = 0;
    for (long z = 0; z < n[2]; ++z)
      for (long y = 0; y < n[1]; ++y)
        for (long x = 0; x < n[0]; ++x) {
          const long q[3] = {(lo[0] + x + nc[0]) % nc[0], (lo[1] + y + nc[1]) % nc[1], (lo[2] + z + nc[2]) % nc[2]};
          const int c = cell[q[0] + nc[0] * (q[1] + nc[1] * q[2])];
          bool seen = false;
          for (int i = 0; i < k; ++i) seen = seen || cand[i] == c;
          if (!seen) cand[k++] = c;
        }
    return k;
  }

  // Local index of point gv of leaf c's level in leaf c (across a periodic
  // seam too).
  long local(int c, int d, long gv) const {
    const long tot = ((long)nb[d] << lev[c]) * N;
    long loc = gv - (long)bc[3 * c + d] * N;
    if (loc < 0) loc += tot;
    if (loc > N) loc -= tot;
    return loc;
  }

  // Where point gl of level L (wrapped on periodic axes, inside the box)
  // takes its value from: 0, the owner -- the finest leaf holding it as a
  // vertex, then the smallest mask of axes with local index N -- and its
  // local index in loc; 1, the order-np prolongation from the level-(L-1)
  // leaf holding it (the smallest such mask), the stencil's first point in
  // loc and the midpoint axes in half; -1, no leaf of level L - 1 holds it.
  // L may exceed Lmax by one (a point of a forest refined from this one).
  int source(int L, const long gl[3], int np, int& leaf, long loc[3], int& half) const {
    const int Lw = L > Lmax ? L : Lmax;
    long X[3];
    for (int d = 0; d < 3; ++d) X[d] = gl[d] << (Lw - L);
    int cand[8];
    const int ncand = around(Lw, X, cand);
    // the axes where the point is the upper plane of a non-periodic box:
    // half-open ownership leaves it to the Dirichlet hook, and only a
    // transfer asks for it, which takes the closed rule there
    int top = 0;
    for (int d = 0; d < 3; ++d)
      if (!periodic[d] && X[d] == nc[d] * ((long)N << (Lw - Lmax))) top |= 1 << d;
    int own = -1, own_mask = 8;
    for (int q = 0; q < ncand; ++q) {
      const int c = cand[q];
      const int s = Lw - lev[c];
      if ((X[0] | X[1] | X[2]) & ((1L << s) - 1)) continue;
      long l[3];
      int mask = 0;
      for (int d = 0; d < 3; ++d) {
        l[d] = local(c, d, X[d] >> s);
        if (l[d] == N) mask |= 1 << d;
      }
      if (half_open && (mask & ~top) != 0) continue;
      if (own < 0 || lev[c] > lev[own] || (lev[c] == lev[own] && mask < own_mask)) {
        own = c;
        own_mask = mask;
        for (int d = 0; d < 3; ++d) loc[d] = l[d];
      }
    }
    half = 0;
    if (own >= 0) {
      leaf = own;
      return 0;
    }
    if (L == 0) return -1;
    for (int d = 0; d < 3; ++d) half |= (int)(gl[d] & 1) << d;
    int cb = -1, cb_mask = 8;
    for (int q = 0; q < ncand; ++q) {
      const int c = cand[q];
      if (lev[c] != L - 1) continue;
      long st[3];
      int mask = 0;
      for (int d = 0; d < 3; ++d) {
        if (half >> d & 1) {
          st[d] = local(c, d, (gl[d] - 1) / 2) - np / 2 + 1;
        } else {
          st[d] = local(c, d, gl[d] / 2);
          if (st[d] == N) mask |= 1 << d;
        }
      }
      if (cb < 0 || mask < cb_mask) {
        cb = c;
        cb_mask = mask;
        for (int d = 0; d < 3; ++d) loc[d] = st[d];
      }
    }
    if (cb < 0) return -1;
    leaf = cb;
    return 1;
  }
};

inline std::string check_forest_shape(int N, int G, int np) {
  if (N % 2 != 0 || N < 2 * G)
    return "a forest with levels needs an even N of at least 2G = " + std::to_string(2 * G) +
           " (a coarse vertex is a fine one, and a ghost reaches no further than the next leaf), got N = " +
           std::to_string(N);
  if (np < 2 || np % 2 != 0 || np / 2 > G)
    return "the prolongation order must be even, at least 2 and at most 2G = " + std::to_string(2 * G) +
           " so that its stencil stays inside the coarse block's ghosts, got " + std::to_string(np);
  return "";
}

// Builds the plan, or returns an error message.
// refl[d][0] / [d][1]: the reflecting low / high face of axis d; rot: the
// rotating seam (the low x and y faces glued by a quarter turn about z). A
// ghost beyond them takes the value of the point inside it is the image of,
// turned and mirrored: the entry's mode carries the mirrored axes in bits
// 4-6 and the number of quarter turns in bits 7-8.
inline std::string forest_plan(int N, int G, const int nb[3], const int periodic[3], int np,
                               const std::vector<int>& lev, const std::vector<int>& bc, ForestPlan& out,
                               bool half_open = false, const int (*refl)[2] = nullptr, int rot = 0) {
  std::string msg = check_forest_shape(N, G, np);
  if (!msg.empty()) return msg;
  ForestIndex ix;
  ix.half_open = half_open;
  msg = ix.build(N, nb, periodic, lev, bc);
  if (!msg.empty()) return msg;
  const int nl = (int)lev.size(), Lmax = ix.Lmax;
  const int S = N + 1 + 2 * G, P = N + 1;
  const long S3 = (long)S * S * S;
  if ((long)nl * S3 >= (1L << 31)) return "the forest has too many points for the fill plan's 32-bit indices";
  out.Lmax = Lmax;
  out.prim.assign((size_t)nl * P * P * P, 0);
  // per chunk of blocks: its passes' entries and its first error
  std::vector<std::vector<std::vector<int>>> cpass(16, std::vector<std::vector<int>>(Lmax + 1));
  std::vector<std::string> cerr(16);
  const int nchunk = for_block_chunks(nl, [&](int c, int b0, int b1) {
    std::vector<std::vector<int>>& pass = cpass[c];
    for (int b = b0; b < b1; ++b) {
      const int L = lev[b];
      for (int k = -G; k <= N + G; ++k)
        for (int j = -G; j <= N + G; ++j)
          for (int i = -G; i <= N + G; ++i) {
            const int dst = (int)((long)b * S3 + (i + G) + (long)S * ((j + G) + (long)S * (k + G)));
            // a point inside the block's extent lies in no other leaf: its primary
            if (i > 0 && i < N && j > 0 && j < N && k > 0 && k < N) {
              out.prim[(size_t)b * P * P * P + i + P * (j + (size_t)P * k)] = 1;
              continue;
            }
            const int l[3] = {i, j, k};
            long gl[3];
            bool outside = false;
            int mm = 0, tn = 0;
            for (int d = 0; d < 3; ++d) gl[d] = (long)bc[3 * b + d] * N + l[d];
            if (rot) {
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
                if (refl && refl[d][0]) {
                  gl[d] = -gl[d];
                  mm |= 1 << d;
                } else {
                  outside = true;
                }
              } else if (gl[d] > tot) {
                if (refl && refl[d][1]) {
                  gl[d] = 2 * tot - gl[d];
                  mm |= 1 << d;
                } else {
                  outside = true;
                }
              } else if (half_open && gl[d] == tot) {
                outside = true;
              }
            }
            if (outside) {
              pass[0].insert(pass[0].end(), {dst, 0, 8, 0});
              continue;
            }
            int leaf, half;
            long loc[3];
            const int kind = ix.source(L, gl, np, leaf, loc, half);
            if (kind < 0) {
              cerr[c] = L == 0 ? std::string("a level-0 point that no leaf holds as a vertex")
                               : "a ghost of a level-" + std::to_string(L) + " leaf lies in no leaf of level " +
                                     std::to_string(L - 1);
              return;
            }
            if (kind == 0 && mm == 0 && tn == 0 && leaf == b && loc[0] == i && loc[1] == j && loc[2] == k) {
              if (i >= 0 && i <= N && j >= 0 && j <= N && k >= 0 && k <= N)
                out.prim[(size_t)b * P * P * P + i + P * (j + (size_t)P * k)] = 1;
              continue;
            }
            if (kind == 1)
              for (int d = 0; d < 3; ++d) {
                const long hi = loc[d] + ((half >> d & 1) ? np - 1 : 0);
                if (loc[d] < -G || hi > N + G) {
                  cerr[c] = "a prolongation stencil leaves its coarse block's ghosts";
                  return;
                }
              }
            const int src =
                (int)((long)leaf * S3 + (loc[0] + G) + (long)S * ((loc[1] + G) + (long)S * (loc[2] + G)));
            const int xf = (mm << 4) | (tn << 7);
            if (kind == 0) pass[0].insert(pass[0].end(), {dst, src, xf, 0});
            else pass[L].insert(pass[L].end(), {dst, src, half | xf, 0});
          }
    }
  });
  for (int c = 0; c < nchunk; ++c)
    if (!cerr[c].empty()) return cerr[c];
  // Within a pass the entries are independent: order them by mode (stable,
  // so each mode's entries stay in destination order) for wave-uniform
  // branches -- a counting sort over the chunks in block order.
  size_t total = 0;
  for (int c = 0; c < nchunk; ++c)
    for (int l = 0; l <= Lmax; ++l) total += cpass[c][l].size();
  out.ops.assign(total, 0);
  out.pass_off.assign(1, 0);
  size_t pos = 0;
  for (int l = 0; l <= Lmax; ++l) {
    // bucket: the interpolation axes and the Dirichlet bit (mode & 15, 0..8)
    size_t off[9] = {0};
    for (int c = 0; c < nchunk; ++c)
      for (size_t e = 0; e < cpass[c][l].size(); e += 4) off[cpass[c][l][e + 2] & 15] += 4;
    size_t run = pos;
    for (int m = 0; m <= 8; ++m) {
      const size_t n = off[m];
      off[m] = run;
      run += n;
    }
    if (l == 0) out.dir_begin = (int)(off[8] / 4);
    for (int c = 0; c < nchunk; ++c) {
      const std::vector<int>& v = cpass[c][l];
      for (size_t e = 0; e < v.size(); e += 4) {
        size_t& o = off[v[e + 2] & 15];
        for (int q = 0; q < 4; ++q) out.ops[o + q] = v[e + q];
        o += 4;
      }
    }
    pos = run;
    out.pass_off.push_back((int)(pos / 4));
  }
  out.dir_end = out.pass_off[1];
  return "";
}

// The regrid's transfer (refine!/coarsen! and the state carried over,
// src/refinement.jl:973-1021): every owned point of the new forest (leaves
// nlev/nbc, same brick and N) takes its value as the old forest's fill would
// give it -- a copy of the old owner (an unchanged leaf, an injection from a
// finer one) or the order-np prolongation from the old level-(L-1) leaf --
// read from the old forest's prepared work array (ghost width G). Entries
// are 4 ints: dst (new block * P^3 + point), src (old block * S^3 + point),
// mode (0 copy, else the midpoint axes), 0; sorted by mode.
inline std::string transfer_plan(const ForestIndex& old, int G, int np, const std::vector<int>& nlev,
                                 const std::vector<int>& nbc, std::vector<int>& ops) {
  const int N = old.N, S = N + 1 + 2 * G, P = N + 1;
  const long S3 = (long)S * S * S, P3 = (long)P * P * P;
  const int nl = (int)nlev.size();
  if ((long)old.lev.size() * S3 >= (1L << 31) || (long)nl * P3 >= (1L << 31))
    return "the forests have too many points for the transfer plan's 32-bit indices";
  std::vector<std::vector<int>> cmode(16 * 8);
  std::vector<std::string> cerr(16);
  const int nchunk = for_block_chunks(nl, [&](int c, int b0, int b1) {
    std::vector<int>* by_mode = &cmode[8 * c];
    for (int b = b0; b < b1; ++b) {
      const int L = nlev[b];
      // a leaf of both forests: after the old fill its owned points hold their
      // owners' values, so they copy across in place
      int same = -1;
      if (L <= old.Lmax) {
        const int f = old.Lmax - L;
        const long x = (long)nbc[3 * b] << f, y = (long)nbc[3 * b + 1] << f, z = (long)nbc[3 * b + 2] << f;
        const int o = old.cell[x + old.nc[0] * (y + old.nc[1] * z)];
        if (old.lev[o] == L && old.bc[3 * o] == nbc[3 * b] && old.bc[3 * o + 1] == nbc[3 * b + 1] &&
            old.bc[3 * o + 2] == nbc[3 * b + 2])
          same = o;
      }
      for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N; ++j)
          for (int i = 0; i <= N; ++i) {
            const int dst = (int)((long)b * P3 + i + (long)P * (j + (long)P * k));
            if (same >= 0) {
              const int src = (int)((long)same * S3 + (i + G) + (long)S * ((j + G) + (long)S * (k + G)));
              by_mode[0].insert(by_mode[0].end(), {dst, src, 0, 0});
              continue;
            }
            const int l[3] = {i, j, k};
            long gl[3];
            for (int d = 0; d < 3; ++d) {
              const long tot = ((long)old.nb[d] << L) * N;
              gl[d] = (long)nbc[3 * b + d] * N + l[d];
              if (old.periodic[d]) gl[d] %= tot;
            }
            int leaf, half;
            long loc[3];
            const int kind = old.source(L, gl, np, leaf, loc, half);
            if (kind < 0) {
              cerr[c] = "new leaf " + std::to_string(b) + " (level " + std::to_string(L) +
                        ") is more than one level finer than the old forest where it lies";
              return;
            }
            for (int d = 0; d < 3; ++d) {
              const long hi = loc[d] + ((half >> d & 1) ? np - 1 : 0);
              if (loc[d] < -G || hi > N + G) {
                cerr[c] = "a prolongation stencil leaves its coarse block's ghosts";
                return;
              }
            }
            const int src =
                (int)((long)leaf * S3 + (loc[0] + G) + (long)S * ((loc[1] + G) + (long)S * (loc[2] + G)));
            by_mode[half].insert(by_mode[half].end(), {dst, src, half, 0});
          }
    }
  });
  for (int c = 0; c < nchunk; ++c)
    if (!cerr[c].empty()) return cerr[c];
  ops.clear();
  for (int m = 0; m < 8; ++m)
    for (int c = 0; c < nchunk; ++c) ops.insert(ops.end(), cmode[8 * c + m].begin(), cmode[8 * c + m].end());
  return "";
}

}  // namespace tgh
