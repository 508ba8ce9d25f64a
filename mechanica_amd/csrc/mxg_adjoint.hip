// Differentiable overdamped Morse trajectory on gfx950: forward steps with
// checkpoints, and the discrete adjoint of the same map.
//
// Forward map, one step (engine_step's overdamped path,
// src/mdcore/src/engine_advance.cpp:126-239, with the pair terms of
// src/mdcore/src/potential_eval.hpp:106-161 and the bond term of bond_eval,
// src/mdcore/src/bond.cpp:99-295, evaluated analytically instead of from
// their tables, so that they have exact parameter derivatives):
//   F_i   = sum_j phi(r_ij) d_ij + sum_bonds phi_b(r_ib) d_ib,
//           d_ij = x_i - x_j (minimum image),
//           phi = -E'(s) / max(r, FLT_EPSILON), pairs with r^2 <= rc^2 and
//           s <= smax only, per type pair either
//             Morse (potential_create_morse): s = max(r, smin) - (R_i + R_j)
//             + r0 + 1, E(s) = D (u^2 - 2u), u = exp(-A (s - 1));
//             LJ 12-6 (potential_create_LJ126): s = max(r, smin),
//             E(s) = A s^-12 - B s^-6;
//           bonds (potential_create_harmonic): rb = clamp(r, bmin, bmax),
//           phi_b = -E'(rb) / rb, E(rb) = K (rb - r0)^2;
//   dx_i  = clamp((dt F_i) / m_i, +-maxx) per component;
//   x_i  <- x_i + dx_i, wrapped into the box.
// Adjoint, lambda_t = dL/dx_t:
//   mu_i      = (dt / m_i) [component not clamped] lambda_{t+1,i}
//   lambda_t,i = lambda_{t+1,i} + sum_j H_ij (mu_i - mu_j),
//   H_ij      = phi I + (phi'(r) / r) d d^T,
//   dL/dtheta += sum_{pairs, bonds} (dphi/dtheta) (mu_i - mu_j) . d_ij,
// with the bonds' H and parameter terms from phi_b.
// Every neighbour sum is a gather in a fixed order (cells in offset order,
// particles within a cell by id), accumulated in FP64 where the adjoint
// needs it; the forward recomputed from a checkpoint reproduces the first
// forward bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>

#include "mxg_adjoint.h"

namespace mxg {

namespace {

constexpr int kAB = 256;  // threads per block
static_assert(kAdjMaxTypes == 4, "k_backward accumulates four types by name");
constexpr double kFix = 1099511627776.0;  // 2^40: fixed-point scale of the histogram

enum : int { kPairMorse = 0, kPairLJ = 1 };

// One type pair's potential: Morse (d, a, r0 + 1, R_i + R_j) or LJ 12-6
// (A, B in d, a).
struct AdjPair {
  int kind;
  float d, a, r0p1, R;
  float smin, smax;
};

// The harmonic bond potential shared by all bonds: K (r - r0)^2 with r
// clamped to [bmin, bmax].
struct AdjBond {
  float K, r0, bmin, bmax;
};

struct AdjArgs {
  int N, ntypes, ncells;
  int cdim[3];
  float L[3], ih[3];
  float rc2;
  float maxx[3], maxx2[3];
  float dt;
  float imass[kAdjMaxTypes];
  float radius[kAdjMaxTypes];
  AdjPair p[kAdjMaxPairs];
  AdjBond bond;
  const int *boff;   // per-id bond list offsets (N + 1), null without bonds
  const int *bent;   // entries: partner id << 1 | (this end is the bond's i)
};

__host__ __device__ inline int pair_index(int ti, int tj, int nt) {
  const int a = ti < tj ? ti : tj, b = ti < tj ? tj : ti;
  return a * nt - a * (a - 1) / 2 + (b - a);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ int cell_of(const AdjArgs &A, float4 x, int &cx, int &cy, int &cz) {
  cx = min(max((int)(x.x * A.ih[0]), 0), A.cdim[0] - 1);
  cy = min(max((int)(x.y * A.ih[1]), 0), A.cdim[1] - 1);
  cz = min(max((int)(x.z * A.ih[2]), 0), A.cdim[2] - 1);
  return (cx * A.cdim[1] + cy) * A.cdim[2] + cz;
}

__device__ __forceinline__ int type_of(float4 x) { return __float_as_int(x.w); }

// Lanes per particle in k_forward and k_backward: lane `sub` of a particle's
// group walks candidates sub, sub + kLPP, ... of every neighbour cell, and the
// group's sums are combined by an xor butterfly (every lane ends with the
// same bits), so the result depends on kLPP but not on scheduling.
constexpr int kLPP = 8;

// Sum over a particle's kLPP lanes (aligned groups within a wave).
template <typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = 1; o < kLPP; o <<= 1) v += __shfl_xor(v, o);
  return v;
}

// Calls fn(j, dx, dy, dz, r2) for every particle j != i (sorted index) of
// the 27 cells around x_i with r^2 <= rc^2, cells in offset order, taking
// the candidates sub, sub + stride, ... of each cell.
template <typename Fn>
__device__ __forceinline__ void neighbours(const AdjArgs &A, const float4 *__restrict__ xs, const int *__restrict__ cstart, int i,
                                           float4 xi, int sub, int stride, Fn &&fn) {
  int cx, cy, cz;
  cell_of(A, xi, cx, cy, cz);
  for (int ox = -1; ox <= 1; ++ox) {
    int nx = cx + ox;
    float sx = 0.0f;
    if (nx < 0) {
      nx += A.cdim[0];
      sx = -A.L[0];
    } else if (nx >= A.cdim[0]) {
      nx -= A.cdim[0];
      sx = A.L[0];
    }
    for (int oy = -1; oy <= 1; ++oy) {
      int ny = cy + oy;
      float sy = 0.0f;
      if (ny < 0) {
        ny += A.cdim[1];
        sy = -A.L[1];
      } else if (ny >= A.cdim[1]) {
        ny -= A.cdim[1];
        sy = A.L[1];
      }
      for (int oz = -1; oz <= 1; ++oz) {
        int nz = cz + oz;
        float sz = 0.0f;
        if (nz < 0) {
          nz += A.cdim[2];
          sz = -A.L[2];
        } else if (nz >= A.cdim[2]) {
          nz -= A.cdim[2];
          sz = A.L[2];
        }
        const int c = (nx * A.cdim[1] + ny) * A.cdim[2] + nz;
        const int j1 = cstart[c + 1];
        for (int j = cstart[c] + sub; j < j1; j += stride) {
          if (j == i) continue;
          const float4 xj = xs[j];
          const float dx = xi.x - (xj.x + sx), dy = xi.y - (xj.y + sy), dz = xi.z - (xj.z + sz);
          const float r2 = dx * dx + dy * dy + dz * dz;
          if (r2 > A.rc2) continue;
          fn(j, xj, dx, dy, dz, r2);
        }
      }
    }
  }
}

// Terms of one pair: returns false when s lies beyond the domain.
struct PairEval {
  float r, ro, s, e, ep, epp;  // E(s), E'(s), E''(s)
  float u;                     // Morse: exp(-a (s - 1))
  float is, is6;               // LJ: 1 / s, s^-6
  bool inner;                  // r >= smin: s moves with r
};

__device__ __forceinline__ bool pair_eval(const AdjPair &P, float r2, PairEval &m) {
  m.r = sqrtf(r2);
  m.ro = fmaxf(m.r, FLT_EPSILON);
  m.inner = m.r >= P.smin;
  const float re = m.inner ? m.r : P.smin;
  if (P.kind == kPairLJ) {
    m.s = re;
    if (m.s > P.smax) return false;
    m.is = 1.0f / m.s;
    const float is2 = m.is * m.is;
    m.is6 = is2 * is2 * is2;
    const float a12 = P.d * m.is6 * m.is6, b6 = P.a * m.is6;
    m.e = a12 - b6;
    m.ep = (-12.0f * a12 + 6.0f * b6) * m.is;
    m.epp = (156.0f * a12 - 42.0f * b6) * is2;
    m.u = 0.0f;
    return true;
  }
  m.s = re - P.R + P.r0p1;
  if (m.s > P.smax) return false;
  m.u = expf(-P.a * (m.s - 1.0f));
  const float uu = m.u * m.u;
  m.e = P.d * (uu - 2.0f * m.u);
  m.ep = 2.0f * P.a * P.d * (m.u - uu);
  m.epp = 2.0f * P.a * P.a * P.d * (2.0f * uu - m.u);
  m.is = m.is6 = 0.0f;
  return true;
}

// dphi/dtheta of a pair, phi = -E'(s) / ro: Morse (d, a, r0), LJ (A, B, -).
__device__ __forceinline__ void pair_dphi(const AdjPair &P, const PairEval &m, float &g0, float &g1, float &g2) {
  if (P.kind == kPairLJ) {
    g0 = (12.0f * m.is6 * m.is6 * m.is) / m.ro;
    g1 = -(6.0f * m.is6 * m.is) / m.ro;
    g2 = 0.0f;
    return;
  }
  const float w = m.u - m.u * m.u;
  g0 = -(2.0f * P.a * w) / m.ro;
  g1 = -(2.0f * P.d * w - 2.0f * P.a * P.d * (m.s - 1.0f) * m.u * (1.0f - 2.0f * m.u)) / m.ro;
  g2 = -m.epp / m.ro;  // ds/dr0 = 1
}

// A bond seen from one end: d = x_self - x_other (minimum image), phi_b and
// psi_b = (dphi_b/dr) / r (zero where r is clamped).
struct BondEval {
  float dx, dy, dz, rb, phi, psi;
  bool in_range;
};

__device__ __forceinline__ float min_image(float d, float L) {
  if (d > 0.5f * L) return d - L;
  if (d < -0.5f * L) return d + L;
  return d;
}

__device__ __forceinline__ void bond_eval(const AdjArgs &A, float4 xi, float4 xj, BondEval &b) {
  b.dx = min_image(xi.x - xj.x, A.L[0]);
  b.dy = min_image(xi.y - xj.y, A.L[1]);
  b.dz = min_image(xi.z - xj.z, A.L[2]);
  const float r = sqrtf(b.dx * b.dx + b.dy * b.dy + b.dz * b.dz);
  b.in_range = r > A.bond.bmin && r < A.bond.bmax;
  b.rb = fminf(fmaxf(r, A.bond.bmin), A.bond.bmax);
  const float ep = 2.0f * A.bond.K * (b.rb - A.bond.r0);
  b.phi = -ep / b.rb;
  b.psi = b.in_range ? (-2.0f * A.bond.K + ep / b.rb) / (b.rb * b.rb) : 0.0f;
}

// ---------------------------------------------------------------- binning

__global__ __launch_bounds__(kAB) void k_init(AdjArgs A, const double *__restrict__ x0, const int *__restrict__ type, float4 *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.N) return;
  float v[3];
  for (int k = 0; k < 3; ++k) {
    const double L = A.L[k];
    double y = x0[3 * i + k];
    y -= L * floor(y / L);
    float f = (float)y;
    if (f >= A.L[k]) f -= A.L[k];
    if (f < 0.0f) f = 0.0f;
    v[k] = f;
  }
  x[i] = make_float4(v[0], v[1], v[2], __int_as_float(type[i]));
}

__global__ __launch_bounds__(kAB) void k_count(AdjArgs A, const float4 *__restrict__ x, int *__restrict__ cellid, int *__restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.N) return;
  int cx, cy, cz;
  const int c = cell_of(A, x[i], cx, cy, cz);
  cellid[i] = c;
  atomicAdd((int *)&cnt[c], 1);
}

// One block: exclusive scan of the counts into cstart, and the cursor of
// each cell (cnt[c] = cstart[c]) for the scatter.
__global__ __launch_bounds__(1024) void k_scan(int *__restrict__ cnt, int *__restrict__ cstart, int n) {
  __shared__ int wsum[16];
  __shared__ int carry_s;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) carry_s = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    const int idx = base + t;
    const int v = idx < n ? cnt[idx] : 0;
    int incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int wbase = 0;
    for (int k = 0; k < w; ++k) wbase += wsum[k];
    int total = 0;
    for (int k = 0; k < 16; ++k) total += wsum[k];
    const int excl = carry_s + wbase + incl - v;
    if (idx < n) {
      cstart[idx] = excl;
      cnt[idx] = excl;
    }
    __syncthreads();
    if (t == 0) carry_s += total;
    __syncthreads();
  }
  if (t == 0) cstart[n] = carry_s;
}

__global__ __launch_bounds__(kAB) void k_scatter(AdjArgs A, const int *__restrict__ cellid, int *__restrict__ cursor, int *__restrict__ pid_s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.N) return;
  const int slot = atomicAdd((int *)&cursor[cellid[i]], 1);
  pid_s[slot] = i;
}

// Per cell: order its slots by particle id, gather the positions, and clear
// the cell's counter for the next binning.  One wave per cell: each lane
// holds one slot's id and finds its rank among the cell's ids with wave
// shuffles (ids are distinct), so the order is the same as a sort's; a cell
// of more than 64 particles is sorted by its first lane.
constexpr int kCellsPerBlock = kAB / 64;

__global__ __launch_bounds__(kAB) void k_cellsort(AdjArgs A, const int *__restrict__ cstart, int *__restrict__ pid_s, const float4 *__restrict__ x,
                                                   float4 *__restrict__ xs, int *__restrict__ cnt) {
  const int c = blockIdx.x * kCellsPerBlock + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= A.ncells) return;
  const int a = cstart[c], b = cstart[c + 1], n = b - a;
  if (n <= 64) {
    const int mine = lane < n ? pid_s[a + lane] : INT_MAX;
    int rank = 0;
    for (int m = 0; m < n; ++m) rank += __shfl(mine, m) < mine ? 1 : 0;
    if (lane < n) {
      pid_s[a + rank] = mine;
      xs[a + rank] = x[mine];
    }
  } else if (lane == 0) {
    for (int k = a + 1; k < b; ++k) {
      const int v = pid_s[k];
      int m = k - 1;
      while (m >= a && pid_s[m] > v) {
        pid_s[m + 1] = pid_s[m];
        --m;
      }
      pid_s[m + 1] = v;
    }
    for (int k = a; k < b; ++k) xs[k] = x[pid_s[k]];
  }
  if (lane == 0) cnt[c] = 0;
}

// Sorted index of each particle id (the inverse of pid_s).
__global__ __launch_bounds__(kAB) void k_slot_of(int N, const int *__restrict__ pid_s, int *__restrict__ slot_of) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < N) slot_of[pid_s[k]] = k;
}

// ---------------------------------------------------------------- forward

// One overdamped step: forces, clamp, advance; kLPP lanes per particle (the
// grid covers N * kLPP threads).  Optional outputs: the force by id (f), the
// clamp bits by id (mask), and per-block FP64 energy sums.
__global__ __launch_bounds__(kAB) void k_forward(AdjArgs A, const float4 *__restrict__ xs, const int *__restrict__ pid_s,
                                                 const int *__restrict__ cstart, const float4 *__restrict__ xid,
                                                 float4 *__restrict__ x_next, unsigned char *__restrict__ mask,
                                                 float *__restrict__ f, double *__restrict__ epart) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / kLPP, sub = t % kLPP;
  double e = 0.0;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  float4 xi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int pid = 0;
  if (i < A.N) {
    xi = xs[i];
    const int ti = type_of(xi);
    neighbours(A, xs, cstart, i, xi, sub, kLPP, [&](int, float4 xj, float dx, float dy, float dz, float r2) __attribute__((always_inline)) {
      const AdjPair &P = A.p[pair_index(ti, type_of(xj), A.ntypes)];
      PairEval m;
      if (!pair_eval(P, r2, m)) return;
      const float phi = -m.ep / m.ro;
      fx += phi * dx;
      fy += phi * dy;
      fz += phi * dz;
      if (epart) e += 0.5 * (double)m.e;
    });
    pid = pid_s[i];
    if (A.boff) {
      // Bonds of this particle, in list order (partners by id), split over
      // the group's lanes like the candidates.
      for (int k = A.boff[pid] + sub; k < A.boff[pid + 1]; k += kLPP) {
        const int ent = A.bent[k];
        BondEval b;
        bond_eval(A, xi, xid[ent >> 1], b);
        fx += b.phi * b.dx;
        fy += b.phi * b.dy;
        fz += b.phi * b.dz;
        if (epart && (ent & 1)) {
          const float q = b.rb - A.bond.r0;
          e += (double)(A.bond.K * q * q);
        }
      }
    }
  }
  // Groups are aligned and share i, so whole groups are in or out of range.
  fx = group_sum(fx);
  fy = group_sum(fy);
  fz = group_sum(fz);
  if (i < A.N && sub == 0) {
    const int ti = type_of(xi);
    const float im = A.imass[ti];
    const float fv[3] = {fx, fy, fz};
    const float xv[3] = {xi.x, xi.y, xi.z};
    float out[3];
    unsigned bits = 0;
    for (int k = 0; k < 3; ++k) {
      float dx = (A.dt * fv[k]) * im;
      if (!(dx * dx <= A.maxx2[k])) {
        dx = copysignf(A.maxx[k], dx);
        bits |= 1u << k;
      }
      float y = xv[k] + dx;
      if (y >= A.L[k])
        y -= A.L[k];
      else if (y < 0.0f)
        y += A.L[k];
      out[k] = y;
    }
    x_next[pid] = make_float4(out[0], out[1], out[2], xi.w);
    if (mask) mask[pid] = (unsigned char)bits;
    if (f) {
      f[3 * pid + 0] = fx;
      f[3 * pid + 1] = fy;
      f[3 * pid + 2] = fz;
    }
  }
  if (epart) {
    __shared__ double red[kAB / 64];
    const double w = wave_sum(e);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int k = 0; k < kAB / 64; ++k) s += red[k];
      epart[blockIdx.x] = s;
    }
  }
}

__global__ __launch_bounds__(kAB) void k_out(int N, const float4 *__restrict__ x, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float4 v = x[i];
  out[3 * i + 0] = v.x;
  out[3 * i + 1] = v.y;
  out[3 * i + 2] = v.z;
}

__global__ __launch_bounds__(kAB) void k_copy4(int N, const float4 *__restrict__ a, float4 *__restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) b[i] = a[i];
}

// ---------------------------------------------------------------- adjoint

__global__ __launch_bounds__(kAB) void k_lam_init(int N, const double *__restrict__ g, double4 *__restrict__ lam) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  lam[i] = make_double4(g[3 * i], g[3 * i + 1], g[3 * i + 2], 0.0);
}

__global__ __launch_bounds__(kAB) void k_lam_out(int N, const double4 *__restrict__ lam, double *__restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double4 v = lam[i];
  g[3 * i] = v.x;
  g[3 * i + 1] = v.y;
  g[3 * i + 2] = v.z;
}

// mu in sorted order: (dt / m) lambda on the components the step did not
// clamp.
__global__ __launch_bounds__(kAB) void k_mu(AdjArgs A, const float4 *__restrict__ xs, const int *__restrict__ pid_s,
                                            const unsigned char *__restrict__ mask, const double4 *__restrict__ lam, double4 *__restrict__ mu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.N) return;
  const int pid = pid_s[i];
  const double c = (double)(A.dt * A.imass[type_of(xs[i])]);
  const unsigned m = mask[pid];
  const double4 l = lam[pid];
  mu[i] = make_double4((m & 1) ? 0.0 : c * l.x, (m & 2) ? 0.0 : c * l.y, (m & 4) ? 0.0 : c * l.z, 0.0);
}

__global__ __launch_bounds__(kAB) void k_zero(double *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// lambda_t = lambda_{t+1} + sum_j H_ij (mu_i - mu_j), and per-block sums of
// the parameter gradient (each pair once: sorted index i < j); kLPP lanes
// per particle, as in k_forward.
__global__ __launch_bounds__(kAB) void k_backward(AdjArgs A, const float4 *__restrict__ xs, const int *__restrict__ pid_s,
                                                  const int *__restrict__ cstart, const int *__restrict__ slot_of,
                                                  const double4 *__restrict__ mu,
                                                  double4 *__restrict__ lam, double *__restrict__ gpart) {
  const int th = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = th / kLPP, sub = th % kLPP;
  double gb0 = 0.0, gb1 = 0.0;  // bond K, r0
  double g[kAdjMaxTypes][3];
#pragma unroll
  for (int t = 0; t < kAdjMaxTypes; ++t) g[t][0] = g[t][1] = g[t][2] = 0.0;
  int ti = 0, pid = 0;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (i < A.N) {
    const float4 xi = xs[i];
    ti = type_of(xi);
    const double4 mi = mu[i];
    neighbours(A, xs, cstart, i, xi, sub, kLPP, [&](int j, float4 xj, float dx, float dy, float dz, float r2) __attribute__((always_inline)) {
      const int tj = type_of(xj);
      const AdjPair &P = A.p[pair_index(ti, tj, A.ntypes)];
      PairEval m;
      if (!pair_eval(P, r2, m)) return;
      const double4 mj = mu[j];
      const double ux = mi.x - mj.x, uy = mi.y - mj.y, uz = mi.z - mj.z;
      const double dd = (double)dx * ux + (double)dy * uy + (double)dz * uz;
      const float phi = -m.ep / m.ro;
      const float psi = ((m.inner ? -m.epp : 0.0f) + m.ep / m.ro) / (m.ro * m.ro);
      const double pp = (double)psi * dd;
      ax += (double)phi * ux + pp * dx;
      ay += (double)phi * uy + pp * dy;
      az += (double)phi * uz + pp * dz;
      if (i < j) {
        float gd, ga, gr;
        pair_dphi(P, m, gd, ga, gr);
        const double c0 = dd * gd, c1 = dd * ga, c2 = dd * gr;
        // Constant indices keep g in registers.
#define MXG_ADJ_ACC(t)  \
  if (tj == t) {        \
    g[t][0] += c0;      \
    g[t][1] += c1;      \
    g[t][2] += c2;      \
  }
        MXG_ADJ_ACC(0)
        MXG_ADJ_ACC(1)
        MXG_ADJ_ACC(2)
        MXG_ADJ_ACC(3)
#undef MXG_ADJ_ACC
      }
    });
    pid = pid_s[i];
    if (A.boff) {
      for (int k = A.boff[pid] + sub; k < A.boff[pid + 1]; k += kLPP) {
        const int ent = A.bent[k];
        const int sj = slot_of[ent >> 1];
        BondEval b;
        bond_eval(A, xi, xs[sj], b);
        const double4 mj = mu[sj];
        const double ux = mi.x - mj.x, uy = mi.y - mj.y, uz = mi.z - mj.z;
        const double dd = (double)b.dx * ux + (double)b.dy * uy + (double)b.dz * uz;
        const double pp = (double)b.psi * dd;
        ax += (double)b.phi * ux + pp * b.dx;
        ay += (double)b.phi * uy + pp * b.dy;
        az += (double)b.phi * uz + pp * b.dz;
        if (ent & 1) {
          // dphi_b/dK = -2 (rb - r0) / rb, dphi_b/dr0 = 2 K / rb
          gb0 += dd * (double)(-2.0f * (b.rb - A.bond.r0) / b.rb);
          gb1 += dd * (double)(2.0f * A.bond.K / b.rb);
        }
      }
    }
  }
  ax = group_sum(ax);
  ay = group_sum(ay);
  az = group_sum(az);
  if (i < A.N && sub == 0) {
    double4 l = lam[pid];
    l.x += ax;
    l.y += ay;
    l.z += az;
    lam[pid] = l;
  }
  // Block sums of the parameter gradient, slot = 3 * pair + component, then
  // the bond's K and r0.
  __shared__ double red[kAdjSlots][kAB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int np = A.ntypes * (A.ntypes + 1) / 2;
  for (int p = 0; p < np; ++p) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double v = 0.0;
      if (i < A.N)
#pragma unroll
        for (int t = 0; t < kAdjMaxTypes; ++t)
          if (t < A.ntypes && pair_index(ti, t, A.ntypes) == p) v += g[t][k];
      v = wave_sum(v);
      if (lane == 0) red[3 * p + k][wv] = v;
    }
  }
  {
    const double v0 = wave_sum(gb0), v1 = wave_sum(gb1);
    if (lane == 0) {
      red[3 * np][wv] = v0;
      red[3 * np + 1][wv] = v1;
      red[3 * np + 2][wv] = 0.0;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < 3 * np + 3) {
    double s = 0.0;
    for (int k = 0; k < kAB / 64; ++k) s += red[threadIdx.x][k];
    gpart[(size_t)blockIdx.x * kAdjSlots + threadIdx.x] = s;
  }
}

// gsum[k] += sum over blocks of gpart[b][k]: one block per slot, each
// thread summing a fixed stride of blocks, then a fixed-order block sum.
__global__ __launch_bounds__(kAB) void k_gsum(const double *__restrict__ gpart, int nblocks, int nslots, double *__restrict__ gsum) {
  __shared__ double red[kAB / 64];
  const int k = blockIdx.x;
  if (k >= nslots) return;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kAB) s += gpart[(size_t)b * kAdjSlots + k];
  const double w = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kAB / 64; ++q) t += red[q];
    gsum[k] += t;
  }
}

// ---------------------------------------------------------------- histogram

// Biweight bin kernel K(t) = (1 - t^2)^2 on |t| < 1, t = r / dr - (k + 1/2).
__device__ __forceinline__ void bins_of(float r, double inv_dr, int nbins, int &k0, double &t0) {
  const double q = (double)r * inv_dr - 0.5;
  k0 = (int)floor(q);
  t0 = q - k0;
}

__global__ __launch_bounds__(kAB) void k_zero_u64(unsigned long long *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Soft pair-distance histogram per type pair, each pair once; sums in
// 2^-40 fixed point so the result does not depend on the order of adds.
__global__ __launch_bounds__(kAB) void k_rdf(AdjArgs A, const float4 *__restrict__ xs, const int *__restrict__ cstart, int nbins,
                                             float rmax2, double inv_dr, unsigned long long *__restrict__ hist) {
  __shared__ unsigned long long h[kAdjMaxPairs * kAdjMaxBins];
  const int np = A.ntypes * (A.ntypes + 1) / 2;
  for (int k = threadIdx.x; k < np * nbins; k += blockDim.x) h[k] = 0ull;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < A.N) {
    const float4 xi = xs[i];
    const int ti = type_of(xi);
    neighbours(A, xs, cstart, i, xi, 0, 1, [&](int j, float4 xj, float, float, float, float r2) __attribute__((always_inline)) {
      if (j < i || r2 >= rmax2) return;
      const int p = pair_index(ti, type_of(xj), A.ntypes);
      int k0;
      double t0;
      bins_of(sqrtf(r2), inv_dr, nbins, k0, t0);
      for (int q = 0; q < 2; ++q) {
        const int k = k0 + q;
        const double t = t0 - q;
        if (k < 0 || k >= nbins || !(t > -1.0 && t < 1.0)) continue;
        const double w = (1.0 - t * t) * (1.0 - t * t);
        atomicAdd(&h[p * nbins + k], (unsigned long long)llrint(w * kFix));
      }
    });
  }
  __syncthreads();
  for (int k = threadIdx.x; k < np * nbins; k += blockDim.x)
    if (h[k]) atomicAdd((unsigned long long *)&hist[k], h[k]);
}

__global__ __launch_bounds__(kAB) void k_rdf_out(int n, const unsigned long long *__restrict__ hist, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)hist[i] / kFix;
}

// dL/dx_i = sum_j sum_k g[p][k] K'(t) / dr * d_ij / r.
__global__ __launch_bounds__(kAB) void k_rdf_backward(AdjArgs A, const float4 *__restrict__ xs, const int *__restrict__ pid_s,
                                                      const int *__restrict__ cstart, int nbins, float rmax2, double inv_dr,
                                                      const double *__restrict__ ghist, double *__restrict__ gx) {
  __shared__ double gh[kAdjMaxPairs * kAdjMaxBins];
  const int np = A.ntypes * (A.ntypes + 1) / 2;
  for (int k = threadIdx.x; k < np * nbins; k += blockDim.x) gh[k] = ghist[k];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.N) return;
  const float4 xi = xs[i];
  const int ti = type_of(xi);
  double ax = 0.0, ay = 0.0, az = 0.0;
  neighbours(A, xs, cstart, i, xi, 0, 1, [&](int, float4 xj, float dx, float dy, float dz, float r2) __attribute__((always_inline)) {
    if (r2 >= rmax2 || r2 <= 0.0f) return;
    const int p = pair_index(ti, type_of(xj), A.ntypes);
    const float r = sqrtf(r2);
    int k0;
    double t0;
    bins_of(r, inv_dr, nbins, k0, t0);
    double gr = 0.0;
    for (int q = 0; q < 2; ++q) {
      const int k = k0 + q;
      const double t = t0 - q;
      if (k < 0 || k >= nbins || !(t > -1.0 && t < 1.0)) continue;
      gr += gh[p * nbins + k] * (-4.0 * t * (1.0 - t * t)) * inv_dr;
    }
    const double s = gr / (double)r;
    ax += s * dx;
    ay += s * dy;
    az += s * dz;
  });
  const int pid = pid_s[i];
  gx[3 * pid + 0] = ax;
  gx[3 * pid + 1] = ay;
  gx[3 * pid + 2] = az;
}

inline int blocks(int n) { return (n + kAB - 1) / kAB; }

}  // namespace

// ---------------------------------------------------------------- host

struct AdjointDev {
  AdjArgs args;
  int *type = nullptr;
  float4 *x[2] = {nullptr, nullptr};
  int *cellid = nullptr, *cnt = nullptr;
  // Per-step slots of a checkpoint segment.
  float4 *xs = nullptr;
  int *pid_s = nullptr, *cstart = nullptr;
  unsigned char *mask = nullptr;
  double4 *lam = nullptr, *mu = nullptr;
  int *slot_of = nullptr;  // sorted index of each id (backward sweep)
  double *part = nullptr;  // per-block partial sums
  double *gsum = nullptr;
  unsigned long long *hist = nullptr;
  int *boff = nullptr, *bent = nullptr;  // bond lists by id
};

Adjoint::~Adjoint() {
  if (!d_) return;
  void *ptrs[] = {d_->type, d_->x[0], d_->x[1], d_->cellid, d_->cnt, d_->xs, d_->pid_s, d_->cstart,
                  d_->mask, d_->lam, d_->mu, d_->slot_of, d_->part, d_->gsum, d_->hist, d_->boff, d_->bent};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete d_;
}

int Adjoint::device(int st, const char *what) {
  if (st == hipSuccess) return 0;
  msg_ = std::string(what) + ": " + hipGetErrorString((hipError_t)st);
  return -21;
}

int Adjoint::create(int N, const double dim[3], const int cells[3], double cutoff, double dt, float max_dist_fraction,
                    int ntypes, const double *mass, const double *radius, const int *type) {
  if (d_) return fail(-1, "already created");
  if (N <= 0 || !dim || !cells || !mass || !radius || !type) return fail(-1, "missing argument");
  if (N > INT_MAX / kLPP) return fail(-6, "too many particles for one thread group each");
  if (ntypes < 1 || ntypes > kAdjMaxTypes) return fail(-6, "between 1 and 4 particle types");
  for (int k = 0; k < 3; ++k) {
    if (cells[k] < 3) return fail(-6, "a periodic grid needs at least 3 cells per axis");
    if (dim[k] / cells[k] < cutoff) return fail(-27, "cells narrower than the cutoff");
  }
  for (int i = 0; i < N; ++i)
    if (type[i] < 0 || type[i] >= ntypes) return fail(-6, "particle type out of range");
  N_ = N;
  ntypes_ = ntypes;
  cutoff_ = cutoff;
  dt_ = dt;
  maxfrac_ = max_dist_fraction;
  mass_.assign(mass, mass + ntypes);
  radius_.assign(radius, radius + ntypes);
  ncells_ = cells[0] * cells[1] * cells[2];
  nblocks_ = blocks(N);
  nlanes_ = blocks(N * kLPP);
  d_ = new AdjointDev();
  AdjArgs &A = d_->args;
  std::memset(&A, 0, sizeof(A));
  A.N = N;
  A.ntypes = ntypes;
  A.ncells = ncells_;
  for (int k = 0; k < 3; ++k) {
    cdim_[k] = cells[k];
    dim_[k] = dim[k];
    A.cdim[k] = cells[k];
    A.L[k] = (float)dim[k];
    A.ih[k] = (float)(cells[k] / dim[k]);
    const double h = dim[k] / cells[k];
    A.maxx[k] = (float)(h * max_dist_fraction);
    A.maxx2[k] = A.maxx[k] * A.maxx[k];
  }
  A.rc2 = (float)(cutoff * cutoff);
  A.dt = (float)dt;
  for (int t = 0; t < ntypes; ++t) {
    A.imass[t] = (float)(1.0 / mass[t]);
    A.radius[t] = (float)radius[t];
  }
  int st = 0;
  st |= hipMalloc(&d_->type, N * sizeof(int));
  st |= hipMalloc(&d_->x[0], N * sizeof(float4));
  st |= hipMalloc(&d_->x[1], N * sizeof(float4));
  st |= hipMalloc(&d_->cellid, N * sizeof(int));
  st |= hipMalloc(&d_->cnt, ncells_ * sizeof(int));
  st |= hipMalloc(&d_->lam, N * sizeof(double4));
  st |= hipMalloc(&d_->mu, N * sizeof(double4));
  st |= hipMalloc(&d_->slot_of, N * sizeof(int));
  st |= hipMalloc(&d_->part, (size_t)nlanes_ * kAdjSlots * sizeof(double));
  st |= hipMalloc(&d_->gsum, kAdjSlots * sizeof(double));
  st |= hipMalloc(&d_->hist, kAdjMaxPairs * kAdjMaxBins * sizeof(unsigned long long));
  if (st) return device(st, "hipMalloc");
  st |= hipMemcpy(d_->type, type, N * sizeof(int), hipMemcpyHostToDevice);
  st |= hipMemset(d_->cnt, 0, ncells_ * sizeof(int));
  if (st) return device(st, "upload");
  return ensure_workspace(1);
}

int Adjoint::ensure_workspace(int every) {
  if (every <= ws_cap_) return 0;
  if (d_->xs) (void)hipFree(d_->xs);
  if (d_->pid_s) (void)hipFree(d_->pid_s);
  if (d_->cstart) (void)hipFree(d_->cstart);
  if (d_->mask) (void)hipFree(d_->mask);
  d_->xs = nullptr;
  d_->pid_s = d_->cstart = nullptr;
  d_->mask = nullptr;
  ws_cap_ = 0;
  int st = 0;
  st |= hipMalloc(&d_->xs, (size_t)every * N_ * sizeof(float4));
  st |= hipMalloc(&d_->pid_s, (size_t)every * N_ * sizeof(int));
  st |= hipMalloc(&d_->cstart, (size_t)every * (ncells_ + 1) * sizeof(int));
  st |= hipMalloc(&d_->mask, (size_t)every * N_);
  if (st) return device(st, "hipMalloc workspace");
  ws_cap_ = every;
  return 0;
}

int Adjoint::set_args(const double *theta, double smin, double smax) {
  if (!d_) return fail(-1, "not created");
  if (!theta) return fail(-1, "missing parameters");
  if (!(smax > smin)) return fail(-6, "empty potential domain");
  AdjArgs &A = d_->args;
  for (int ti = 0; ti < ntypes_; ++ti)
    for (int tj = ti; tj < ntypes_; ++tj) {
      const int p = pair_index(ti, tj, ntypes_);
      AdjPair &P = A.p[p];
      P.kind = kind_[p];
      P.d = (float)theta[3 * p + 0];
      P.a = (float)theta[3 * p + 1];
      P.r0p1 = (float)(theta[3 * p + 2] + 1.0);
      P.R = (float)(radius_[ti] + radius_[tj]);
      P.smin = (float)(dom_[p] ? pmin_[p] : smin);
      P.smax = (float)(dom_[p] ? pmax_[p] : smax);
    }
  if (M_ > 0) {
    const int b = 3 * npairs();
    A.bond.K = (float)theta[b];
    A.bond.r0 = (float)theta[b + 1];
    A.bond.bmin = (float)bmin_;
    A.bond.bmax = (float)bmax_;
    A.boff = d_->boff;
    A.bent = d_->bent;
  } else {
    A.boff = A.bent = nullptr;
  }
  return 0;
}

int Adjoint::set_pair_kind(int ti, int tj, int kind, double smin, double smax) {
  if (!d_) return fail(-1, "not created");
  if (ti < 0 || tj < 0 || ti >= ntypes_ || tj >= ntypes_) return fail(-6, "type out of range");
  if (kind != kPairMorse && kind != kPairLJ) return fail(-6, "pair kind is 0 (Morse) or 1 (LJ 12-6)");
  const int p = pair_index(ti, tj, ntypes_);
  kind_[p] = kind;
  dom_[p] = smax > smin;
  pmin_[p] = smin;
  pmax_[p] = smax;
  return 0;
}

int Adjoint::set_bonds(int M, const int *bi, const int *bj, double rmin, double rmax) {
  if (!d_) return fail(-1, "not created");
  if (M < 0 || (M > 0 && (!bi || !bj))) return fail(-1, "bond arrays missing");
  if (M > 0 && !(rmax > rmin && rmin > 0)) return fail(-6, "bond domain must satisfy 0 < rmin < rmax");
  std::vector<int> off(N_ + 1, 0);
  for (int b = 0; b < M; ++b) {
    if (bi[b] < 0 || bi[b] >= N_ || bj[b] < 0 || bj[b] >= N_ || bi[b] == bj[b])
      return fail(-11, "bond " + std::to_string(b) + " has an invalid particle id");
    off[bi[b] + 1]++;
    off[bj[b] + 1]++;
  }
  for (int i = 0; i < N_; ++i) off[i + 1] += off[i];
  std::vector<int> ent(2 * (size_t)M + 1), fill(off.begin(), off.end() - 1);
  for (int b = 0; b < M; ++b) {
    ent[fill[bi[b]]++] = (bj[b] << 1) | 1;
    ent[fill[bj[b]]++] = bi[b] << 1;
  }
  if (d_->boff) (void)hipFree(d_->boff);
  if (d_->bent) (void)hipFree(d_->bent);
  d_->boff = d_->bent = nullptr;
  M_ = 0;
  if (M == 0) return 0;
  int st = hipMalloc(&d_->boff, (N_ + 1) * sizeof(int));
  st |= hipMalloc(&d_->bent, ent.size() * sizeof(int));
  if (st) return device(st, "hipMalloc bonds");
  st |= hipMemcpy(d_->boff, off.data(), (N_ + 1) * sizeof(int), hipMemcpyHostToDevice);
  st |= hipMemcpy(d_->bent, ent.data(), ent.size() * sizeof(int), hipMemcpyHostToDevice);
  if (st) return device(st, "bond upload");
  M_ = M;
  bmin_ = rmin;
  bmax_ = rmax;
  return 0;
}

// Bins positions x (by id) into workspace slot `slot`: sorted ids, cell
// offsets and the gathered positions.
void Adjoint::bin(const void *xv, int slot, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  const AdjArgs &A = d_->args;
  const float4 *x = (const float4 *)xv;
  int *pid_s = d_->pid_s + (size_t)slot * N_;
  int *cstart = d_->cstart + (size_t)slot * (ncells_ + 1);
  float4 *xs = d_->xs + (size_t)slot * N_;
  hipLaunchKernelGGL(k_count, dim3(nblocks_), dim3(kAB), 0, s, A, x, d_->cellid, d_->cnt);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, d_->cnt, cstart, ncells_);
  hipLaunchKernelGGL(k_scatter, dim3(nblocks_), dim3(kAB), 0, s, A, d_->cellid, d_->cnt, pid_s);
  hipLaunchKernelGGL(k_cellsort, dim3((ncells_ + kCellsPerBlock - 1) / kCellsPerBlock), dim3(kAB), 0, s, A, cstart,
                     pid_s, x, xs, d_->cnt);
}

int Adjoint::forward(const double *theta, double smin, double smax, const double *x0, double *xT, float *ckpt,
                     int nsteps, int every, void *stream) {
  if (int rc = set_args(theta, smin, smax)) return rc;
  if (!x0 || !xT || !ckpt) return fail(-1, "missing buffer");
  if (nsteps < 1 || every < 1) return fail(-6, "nsteps and every must be positive");
  hipStream_t s = (hipStream_t)stream;
  const AdjArgs &A = d_->args;
  float4 *xa = d_->x[0], *xb = d_->x[1];
  float4 *ck = (float4 *)ckpt;
  hipLaunchKernelGGL(k_init, dim3(nblocks_), dim3(kAB), 0, s, A, x0, d_->type, xa);
  for (int t = 0; t < nsteps; ++t) {
    if (t % every == 0) hipLaunchKernelGGL(k_copy4, dim3(nblocks_), dim3(kAB), 0, s, N_, xa, ck + (size_t)(t / every) * N_);
    bin(xa, 0, stream);
    hipLaunchKernelGGL(k_forward, dim3(nlanes_), dim3(kAB), 0, s, A, d_->xs, d_->pid_s, d_->cstart, xa, xb,
                       (unsigned char *)nullptr, (float *)nullptr, (double *)nullptr);
    std::swap(xa, xb);
  }
  hipLaunchKernelGGL(k_out, dim3(nblocks_), dim3(kAB), 0, s, N_, xa, xT);
  return device(hipGetLastError(), "forward launch");
}

int Adjoint::backward(const double *theta, double smin, double smax, const float *ckpt, int nsteps, int every,
                      const double *gxT, double *gx0, double *gtheta, void *stream) {
  if (int rc = set_args(theta, smin, smax)) return rc;
  if (!ckpt || !gxT || !gx0 || !gtheta) return fail(-1, "missing buffer");
  if (nsteps < 1 || every < 1) return fail(-6, "nsteps and every must be positive");
  if (int rc = ensure_workspace(std::min(every, nsteps))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const AdjArgs &A = d_->args;
  const float4 *ck = (const float4 *)ckpt;
  const int nslots = 3 * nrows();
  hipLaunchKernelGGL(k_lam_init, dim3(nblocks_), dim3(kAB), 0, s, N_, gxT, d_->lam);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(kAB), 0, s, d_->gsum, kAdjSlots);
  const int nseg = (nsteps + every - 1) / every;
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int t0 = seg * every, n = std::min(every, nsteps - t0);
    float4 *xa = d_->x[0], *xb = d_->x[1];
    hipLaunchKernelGGL(k_copy4, dim3(nblocks_), dim3(kAB), 0, s, N_, ck + (size_t)seg * N_, xa);
    // Recompute the segment, keeping each step's binning and clamp bits.
    for (int q = 0; q < n; ++q) {
      bin(xa, q, stream);
      hipLaunchKernelGGL(k_forward, dim3(nlanes_), dim3(kAB), 0, s, A, d_->xs + (size_t)q * N_,
                         d_->pid_s + (size_t)q * N_, d_->cstart + (size_t)q * (ncells_ + 1), xa, xb,
                         d_->mask + (size_t)q * N_, (float *)nullptr, (double *)nullptr);
      std::swap(xa, xb);
    }
    for (int q = n - 1; q >= 0; --q) {
      const float4 *xs = d_->xs + (size_t)q * N_;
      const int *pid_s = d_->pid_s + (size_t)q * N_;
      const int *cstart = d_->cstart + (size_t)q * (ncells_ + 1);
      hipLaunchKernelGGL(k_mu, dim3(nblocks_), dim3(kAB), 0, s, A, xs, pid_s, d_->mask + (size_t)q * N_, d_->lam,
                         d_->mu);
      if (M_ > 0) hipLaunchKernelGGL(k_slot_of, dim3(nblocks_), dim3(kAB), 0, s, N_, pid_s, d_->slot_of);
      hipLaunchKernelGGL(k_backward, dim3(nlanes_), dim3(kAB), 0, s, A, xs, pid_s, cstart, d_->slot_of, d_->mu,
                         d_->lam, d_->part);
      hipLaunchKernelGGL(k_gsum, dim3(nslots), dim3(kAB), 0, s, d_->part, nlanes_, nslots, d_->gsum);
    }
  }
  hipLaunchKernelGGL(k_lam_out, dim3(nblocks_), dim3(kAB), 0, s, N_, d_->lam, gx0);
  int st = hipGetLastError();
  if (st) return device(st, "backward launch");
  st = hipMemcpyAsync(gtheta, d_->gsum, nslots * sizeof(double), hipMemcpyDeviceToHost, s);
  st |= hipStreamSynchronize(s);
  return device(st, "backward");
}

int Adjoint::forces(const double *theta, double smin, double smax, const double *x, float *f, double *epot,
                    void *stream) {
  if (int rc = set_args(theta, smin, smax)) return rc;
  if (!x || !f || !epot) return fail(-1, "missing buffer");
  hipStream_t s = (hipStream_t)stream;
  const AdjArgs &A = d_->args;
  hipLaunchKernelGGL(k_init, dim3(nblocks_), dim3(kAB), 0, s, A, x, d_->type, d_->x[0]);
  bin(d_->x[0], 0, stream);
  hipLaunchKernelGGL(k_forward, dim3(nlanes_), dim3(kAB), 0, s, A, d_->xs, d_->pid_s, d_->cstart, d_->x[0],
                     d_->x[1], (unsigned char *)nullptr, f, d_->part);
  std::vector<double> part(nlanes_);
  int st = hipMemcpyAsync(part.data(), d_->part, nlanes_ * sizeof(double), hipMemcpyDeviceToHost, s);
  st |= hipStreamSynchronize(s);
  if (st) return device(st, "forces");
  double e = 0.0;
  for (double v : part) e += v;
  *epot = e;
  return 0;
}

int Adjoint::rdf(const double *x, int nbins, double rmax, double *hist, void *stream) {
  if (!d_) return fail(-1, "not created");
  if (!x || !hist) return fail(-1, "missing buffer");
  if (nbins < 1 || nbins > kAdjMaxBins) return fail(-6, "between 1 and 256 bins");
  if (!(rmax > 0) || rmax > cutoff_) return fail(-6, "rmax must lie in (0, cutoff]");
  hipStream_t s = (hipStream_t)stream;
  const AdjArgs &A = d_->args;
  const int n = npairs() * nbins;
  hipLaunchKernelGGL(k_init, dim3(nblocks_), dim3(kAB), 0, s, A, x, d_->type, d_->x[0]);
  bin(d_->x[0], 0, stream);
  hipLaunchKernelGGL(k_zero_u64, dim3(blocks(n)), dim3(kAB), 0, s, d_->hist, n);
  hipLaunchKernelGGL(k_rdf, dim3(nblocks_), dim3(kAB), 0, s, A, d_->xs, d_->cstart, nbins, (float)(rmax * rmax),
                     nbins / rmax, d_->hist);
  hipLaunchKernelGGL(k_rdf_out, dim3(blocks(n)), dim3(kAB), 0, s, n, d_->hist, hist);
  return device(hipGetLastError(), "rdf launch");
}

int Adjoint::rdf_backward(const double *x, int nbins, double rmax, const double *ghist, double *gx, void *stream) {
  if (!d_) return fail(-1, "not created");
  if (!x || !ghist || !gx) return fail(-1, "missing buffer");
  if (nbins < 1 || nbins > kAdjMaxBins) return fail(-6, "between 1 and 256 bins");
  if (!(rmax > 0) || rmax > cutoff_) return fail(-6, "rmax must lie in (0, cutoff]");
  hipStream_t s = (hipStream_t)stream;
  const AdjArgs &A = d_->args;
  hipLaunchKernelGGL(k_init, dim3(nblocks_), dim3(kAB), 0, s, A, x, d_->type, d_->x[0]);
  bin(d_->x[0], 0, stream);
  hipLaunchKernelGGL(k_rdf_backward, dim3(nblocks_), dim3(kAB), 0, s, A, d_->xs, d_->pid_s, d_->cstart, nbins,
                     (float)(rmax * rmax), nbins / rmax, ghist, gx);
  return device(hipGetLastError(), "rdf backward launch");
}

// ---------------------------------------------------------------- no-alloc

namespace {

// Scratch of morse_forces_static: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code never passes the array's name
// to a kernel (on the host it is a stand-in, not a device address): the
// address comes from hipGetSymbolAddress, once per device.
constexpr size_t kStaticPartBytes = (size_t)kStaticMaxParticles * (3 * sizeof(float4) + 2 * sizeof(int));
constexpr size_t kStaticCellBytes = (size_t)(2 * kStaticMaxCells + 1) * sizeof(int);
constexpr size_t kStaticPartials = ((size_t)kStaticMaxParticles * kLPP + kAB - 1) / kAB;  // k_forward's blocks
constexpr size_t kStaticBytes = kStaticPartBytes + (kStaticCellBytes + 255) / 256 * 256 + kStaticPartials * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch (looked up on the first call on the
// device, with the device current).
int scratch_of_current_device(unsigned char **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -21;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -21;
    g_scratch_ptr[dev] = (unsigned char *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

__global__ __launch_bounds__(kAB) void k_fill0(int *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Fixed-order sum of the per-block energies into one double.
__global__ __launch_bounds__(kAB) void k_esum(const double *__restrict__ part, int n, double *__restrict__ out) {
  __shared__ double red[kAB / 64];
  double s = 0.0;
  for (int k = threadIdx.x; k < n; k += kAB) s += part[k];
  const double w = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kAB / 64; ++k) t += red[k];
    *out = t;
  }
}

}  // namespace

int morse_forces_static(int N, const double dim[3], const int cells[3], double cutoff, int ntypes,
                        const double *radius, const double *theta, double smin, double smax, const int *type,
                        const double *x, float *f, double *epot, void *stream) {
  if (N <= 0 || !dim || !cells || !radius || !theta || !type || !x || !f || !epot) return -1;
  if (N > kStaticMaxParticles) return -6;
  if (ntypes < 1 || ntypes > kAdjMaxTypes) return -6;
  if (!(smax > smin)) return -6;
  long long nc = 1;
  for (int k = 0; k < 3; ++k) {
    if (cells[k] < 3 || !(dim[k] > 0)) return -6;
    if (dim[k] / cells[k] < cutoff) return -27;
    nc *= cells[k];
  }
  if (nc > kStaticMaxCells) return -6;
  unsigned char *base = nullptr;
  if (int rc = scratch_of_current_device(&base)) return rc;
  AdjArgs A;
  std::memset(&A, 0, sizeof(A));
  A.N = N;
  A.ntypes = ntypes;
  A.ncells = (int)nc;
  for (int k = 0; k < 3; ++k) {
    A.cdim[k] = cells[k];
    A.L[k] = (float)dim[k];
    A.ih[k] = (float)(cells[k] / dim[k]);
    A.maxx[k] = 0.0f;
    A.maxx2[k] = 0.0f;
  }
  A.rc2 = (float)(cutoff * cutoff);
  for (int t = 0; t < ntypes; ++t) {
    A.imass[t] = 1.0f;
    A.radius[t] = (float)radius[t];
  }
  for (int ti = 0; ti < ntypes; ++ti)
    for (int tj = ti; tj < ntypes; ++tj) {
      const int p = pair_index(ti, tj, ntypes);
      AdjPair &P = A.p[p];
      P.kind = kPairMorse;
      P.d = (float)theta[3 * p + 0];
      P.a = (float)theta[3 * p + 1];
      P.r0p1 = (float)(theta[3 * p + 2] + 1.0);
      P.R = (float)(radius[ti] + radius[tj]);
      P.smin = (float)smin;
      P.smax = (float)smax;
    }
  // Scratch layout: positions in, sorted, next; cell ids, sorted ids; cell
  // counts and offsets; per-block energies.
  float4 *xin = (float4 *)base;
  float4 *xs = xin + kStaticMaxParticles;
  float4 *xnext = xs + kStaticMaxParticles;
  int *cellid = (int *)(xnext + kStaticMaxParticles);
  int *pid_s = cellid + kStaticMaxParticles;
  int *cnt = pid_s + kStaticMaxParticles;
  int *cstart = cnt + kStaticMaxCells;
  double *epart = (double *)(base + kStaticPartBytes + (kStaticCellBytes + 255) / 256 * 256);
  const int nb = (N + kAB - 1) / kAB;
  hipStream_t s = (hipStream_t)stream;
  // The scratch keeps the previous call's contents: the counters start at
  // zero by a fill kernel (no memset node under graph capture).
  hipLaunchKernelGGL(k_fill0, dim3((A.ncells + kAB - 1) / kAB), dim3(kAB), 0, s, cnt, A.ncells);
  hipLaunchKernelGGL(k_init, dim3(nb), dim3(kAB), 0, s, A, x, type, xin);
  hipLaunchKernelGGL(k_count, dim3(nb), dim3(kAB), 0, s, A, xin, cellid, cnt);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, cnt, cstart, A.ncells);
  hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kAB), 0, s, A, cellid, cnt, pid_s);
  hipLaunchKernelGGL(k_cellsort, dim3((A.ncells + kCellsPerBlock - 1) / kCellsPerBlock), dim3(kAB), 0, s, A, cstart,
                     pid_s, xin, xs, cnt);
  const int nlb = (N * kLPP + kAB - 1) / kAB;
  hipLaunchKernelGGL(k_forward, dim3(nlb), dim3(kAB), 0, s, A, xs, pid_s, cstart, xin, xnext, (unsigned char *)nullptr,
                     f, epart);
  hipLaunchKernelGGL(k_esum, dim3(1), dim3(kAB), 0, s, epart, nlb, epot);
  return hipGetLastError() == hipSuccess ? 0 : -21;
}

}  // namespace mxg
