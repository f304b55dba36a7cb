// Device building blocks shared by the fused multi-operator line sweeps
// (momentum.hip, projection.hip) and their host-side launch planning.
//
// y/z sweeps: a block holds TL lines (consecutive i, so every line point is
// a row of TL words) split into P segments of L points, one thread per
// segment.  x sweeps: one wavefront per line, V contiguous points per lane,
// halos through cross-lane reads and the recurrences' carries through a
// scan across the 64 lanes.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "lines_impl.h"

namespace x3d {

// ---------------------------------------------------------------- y / z segment solve
// Block: TL consecutive i (one row of TL words per line point) x P segments
// of L points; a thread owns one segment of one line.  The two recurrences
// of each solve run locally per segment with a zero carry-in; the carries
// are then chained through LDS.  The chain stops K segments back: the factor
// a carry picks up across a segment is the product of that segment's
// elimination factors, and K is chosen on the host so that the product over
// any K consecutive segments is below the precision's resolution (the terms
// dropped are smaller than the rounding of the full chain).
template <typename T, int L, int NR>
__device__ __forceinline__ void seg_solve(T (&v)[NR][L], const int e0, const T *__restrict__ cs, const T alfa,
                                          const T den, T *cf, T *cb, T *gf, T *gb, T *x0s, T *xns, const int line,
                                          const int seg, const int TL, const int P, const int K) {
  // NR right-hand sides share the matrix: each coefficient is read once.
  const int PT = P * TL;
  // forward elimination, zero carry-in
  T prev[NR], g = T(1);
#pragma unroll
  for (int q = 0; q < NR; ++q) prev[q] = T(0);
#pragma unroll
  for (int m = 0; m < L; ++m) {
    const T f = cs[4 * (e0 + m)];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      prev[q] = v[q][m] - prev[q] * f;
      v[q][m] = prev[q];
    }
    g = -f * g;
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) cf[q * PT + seg * TL + line] = prev[q];
  if (line == 0) gf[seg] = g;
  __syncthreads();
  if (seg > 0) {
    T cin[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) cin[q] = T(0);
    for (int s = seg - K > 0 ? seg - K : 0; s < seg; ++s) {
      const T gs = gf[s];
#pragma unroll
      for (int q = 0; q < NR; ++q) cin[q] = cf[q * PT + s * TL + line] + gs * cin[q];
    }
    T gg = T(1);
#pragma unroll
    for (int m = 0; m < L; ++m) {
      gg = -cs[4 * (e0 + m)] * gg;
#pragma unroll
      for (int q = 0; q < NR; ++q) v[q][m] += gg * cin[q];
    }
  }
  // back substitution, zero carry-in
  T nxt[NR], h = T(1);
#pragma unroll
  for (int q = 0; q < NR; ++q) nxt[q] = T(0);
#pragma unroll
  for (int m = L - 1; m >= 0; --m) {
    const T w = cs[4 * (e0 + m) + 1], ffw = cs[4 * (e0 + m) + 2];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      nxt[q] = w * v[q][m] - ffw * nxt[q];
      v[q][m] = nxt[q];
    }
    h = -ffw * h;
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) cb[q * PT + seg * TL + line] = nxt[q];
  if (line == 0) gb[seg] = h;
  __syncthreads();
  if (seg < P - 1) {
    T c[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) c[q] = T(0);
    for (int s = seg + K < P - 1 ? seg + K : P - 1; s > seg; --s) {
      const T gs = gb[s];
#pragma unroll
      for (int q = 0; q < NR; ++q) c[q] = cb[q * PT + s * TL + line] + gs * c[q];
    }
    T hh = T(1);
#pragma unroll
    for (int m = L - 1; m >= 0; --m) {
      hh = -cs[4 * (e0 + m) + 2] * hh;
#pragma unroll
      for (int q = 0; q < NR; ++q) v[q][m] += hh * c[q];
    }
  }
  // periodic closure (Sherman-Morrison with the precomputed vector sm)
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    if (seg == 0) x0s[q * TL + line] = v[q][0];
    if (seg == P - 1) xns[q * TL + line] = v[q][L - 1];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const T sx = (x0s[q * TL + line] - alfa * xns[q * TL + line]) / den;
#pragma unroll
    for (int m = 0; m < L; ++m) v[q][m] = v[q][m] - sx * cs[4 * (e0 + m) + 3];
  }
  // No barrier needed before the next solve: it writes cf/gf before its
  // first barrier (last read here before the second), cb/gb after its first
  // barrier (last read here before the third) and x0s/xns after its second.
}

template <typename T>
__device__ __forceinline__ T &at(T *p, unsigned byte_off) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + byte_off);
}
template <typename T>
__device__ __forceinline__ const T &at(const T *p, unsigned byte_off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + byte_off);
}

// ---------------------------------------------------------------- x wavefront solve
template <typename T, int V, int HL, int HR>
__device__ __forceinline__ void halo_window(const T (&own)[V], T (&v)[V + HL + HR], const int lane) {
#pragma unroll
  for (int k = 0; k < V + HL + HR; ++k) {
    const int d = k - HL;
    if (d >= 0 && d < V) {
      v[k] = own[d];
    } else {
      const int lo = floor_div(d, V);
      const int el = d - lo * V;
      v[k] = __shfl(own[el], (lane + lo) & 63);
    }
  }
}

template <typename T, int V>
__device__ __forceinline__ void wave_solve(T (&v)[V], const T (&fs)[V], const T (&fw)[V], const T (&ffw)[V],
                                           const T (&sm)[V], const T alfa, const T den, const int lane) {
  T prev = T(0), g = T(1);
#pragma unroll
  for (int m = 0; m < V; ++m) {
    prev = v[m] - prev * fs[m];
    v[m] = prev;
    g = -fs[m] * g;
  }
  {
    T a = g, bb = prev;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const T ap = __shfl_up(a, off), bp = __shfl_up(bb, off);
      if (lane >= off) {
        bb = a * bp + bb;
        a = a * ap;
      }
    }
    T cin = __shfl_up(bb, 1);
    if (lane == 0) cin = T(0);
    T gg = T(1);
#pragma unroll
    for (int m = 0; m < V; ++m) {
      gg = -fs[m] * gg;
      v[m] += gg * cin;
    }
  }
  T nxt = T(0), h = T(1);
#pragma unroll
  for (int m = V - 1; m >= 0; --m) {
    nxt = fw[m] * v[m] - ffw[m] * nxt;
    v[m] = nxt;
    h = -ffw[m] * h;
  }
  {
    T a = h, bb = nxt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const T an = __shfl_down(a, off), bn = __shfl_down(bb, off);
      if (lane + off < 64) {
        bb = a * bn + bb;
        a = a * an;
      }
    }
    T c = __shfl_down(bb, 1);
    if (lane == 63) c = T(0);
    T hh = T(1);
#pragma unroll
    for (int m = V - 1; m >= 0; --m) {
      hh = -ffw[m] * hh;
      v[m] += hh * c;
    }
  }
  const T x0 = __shfl(v[0], 0), xn = __shfl(v[V - 1], 63);
  const T sx = (x0 - alfa * xn) / den;
#pragma unroll
  for (int m = 0; m < V; ++m) v[m] = v[m] - sx * sm[m];
}

// ---------------------------------------------------------------- host planning
inline bool same_set(const TriSet &a, const TriSet &b) {
  return &a == &b || (a.ff == b.ff && a.fs == b.fs && a.fw == b.fw && a.sm == b.sm && a.alfa == b.alfa &&
                      a.sm_den == b.sm_den);
}
// Launch shape of the y/z sweep for a line of n points: segments of L = 8
// (FP64) or 16 (FP32) points, so that the three solves keep their windows,
// right-hand sides and results within 128 VGPRs; TL lines per block with
// TL*P <= 1024 threads and the LDS (two coefficient tables, the carries of
// two right-hand sides) within 64 KiB.
template <typename T>
struct YzShape {
  static constexpr int L = sizeof(T) == 8 ? 8 : 16;
  int P = 0, TL = 0;
  size_t shmem = 0;
  static size_t lds(int n, int P, int TL) {
    return sizeof(T) * (8 * (size_t)n + 4 * (size_t)P * TL + 2 * (size_t)P + 4 * (size_t)TL);
  }
  bool plan(int n) {
    if (n < 8 || n % L != 0) return false;
    P = n / L;
    // 128-byte rows (16 FP64 / 32 FP32 lines): full cache lines per point,
    // and at 256-point lines 512-thread blocks, two of which share a CU
    TL = 128 / (int)sizeof(T);
    while (TL >= 8 && (TL * P > 1024 || lds(n, P, TL) > 65536)) TL >>= 1;
    if (TL < 8) return false;
    shmem = lds(n, P, TL);
    return true;
  }
};

// Smallest K such that the elimination factors of any K consecutive segments
// of L points multiply to less than tol*(1-rho) in magnitude, rho being the
// largest single-segment product: the carries of segments further back then
// change no result by more than tol times the largest carry.  Forward factors
// are fs, backward factors ff*fw (the recurrences of prepare, schemes.f90:
// 413-439).  Returns P (the full chain) when no shorter depth qualifies.
inline int chain_depth(const TriSet &s, int L, double tol) {
  const int n = (int)s.fs.size();
  const int P = (n + L - 1) / L;
  std::vector<double> gfw(P, 1.0), gbw(P, 1.0);
  for (int e = 0; e < n; ++e) {
    gfw[e / L] *= std::fabs(s.fs[e]);
    gbw[e / L] *= std::fabs(s.ff[e] * s.fw[e]);
  }
  double rho = 0;
  for (int p = 0; p < P; ++p) rho = std::max(rho, std::max(gfw[p], gbw[p]));
  if (rho >= 1.0) return P;
  for (int K = 1; K < P; ++K) {
    double worst = 0;
    for (int p0 = 0; p0 + K <= P; ++p0) {
      double pf = 1, pb = 1;
      for (int q = p0; q < p0 + K; ++q) {
        pf *= gfw[q];
        pb *= gbw[q];
      }
      worst = std::max(worst, std::max(pf, pb));
    }
    if (worst < tol * (1.0 - rho)) return K;
  }
  return P;
}

}  // namespace x3d
