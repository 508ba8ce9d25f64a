// MI355X (gfx950) engine for mdcore's force-and-advance step.
//
// One step is six launches on the engine's stream (capturable as a HIP
// graph):
//   k_brick         per brick of cells: stage the brick and its halo in LDS,
//                   then per particle the non-bonded gather over the
//                   neighbour cells (candidate scan into a per-thread list,
//                   then the pair phase), bonds, the single-body force, the
//                   Euler/overdamped advance and the cell-migration rule;
//                   integer per-cell counts and FP64 per-brick partial sums
//   k_brick_global  the same for bricks whose halo overflowed LDS, with
//                   neighbours read from global memory
//   k_scan_block, k_scan_add  exclusive scan of the counts into the new cell
//                   offsets
//   k_place         stayers keep their rank in the cell, movers take a cursor;
//                   its first blocks sum each replica's brick partials in a
//                   fixed order
//   k_order_incoming  movers of each cell ordered by particle id
//
// Every pair is evaluated from both sides (no Newton-3 write-back), so each
// particle's force is written once by the thread that owns it and the result
// is independent of scheduling.  Each side computes the pair in the
// reference's orientation (space_getsid, src/mdcore/src/space.cpp:95-130:
// the cell whose shift to the other points "forward" is i, and
// dx = (x_i - shift) - x_j), so r^2, the cutoff test and the pair force are
// bit-identical to the reference's runner_dopair/runner_doself
// (src/mdcore/src/runner_dopair.cpp:91-340, 446-682).  Compile with
// -ffp-contract=off: the reference never fuses multiply-adds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "mxg_engine.h"
#include "mxg_philox.h"

namespace mxg {

// Pointers into device memory carry the global address space in device code,
// so loads through argument structs held in memory are global (or scalar)
// loads rather than flat ones; host code sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define G __attribute__((address_space(1)))
#else
#define G
#endif

namespace {

constexpr int kNRed = 12;  // epot_nonbond, epot_bond, ekin, virial[9]
constexpr int kTile = 256;

struct DevReplica {
  int cdim[3];
  int cell_base;
  int off_base, n_off;
  int shift_base[3];
  int span[3];
  float h[3], h2[3];
  float maxv[3], maxv2[3], maxx[3], maxx2[3];
  double hd[3];
  float bias, dscale;
  int dmaxdist;
  // Bits 0-2: axis x, y, z periodic.  Bits 3-8: the left, right, front,
  // back, bottom, top walls are active (free-slip, velocity or potential).
  int wall;
  // A particle can take part in dropping a cell pair (see DevOffset::flags)
  // only if each coordinate along the pair's axes lies within near_eps of
  // the cell's bound; drop_check is set when any offset can drop.
  float near_eps;
  int drop_check;
  // Random streams of the replica: its seed, and the gid of its first
  // particle (a particle's stream id is its gid minus id_base).
  uint32_t seed_lo, seed_hi;
  int id_base;
  int pad;
};

enum : int { kWallPeriodicShift = 0, kWallActiveShift = 3 };

// One side of the box (MxBoundaryCondition, MxBoundaryConditions.hpp:25-55):
// sides 0 left (x low), 1 right, 2 front (y low), 3 back, 4 bottom (z low),
// 5 top.
struct DevWall {
  int kind;
  float vel[3];
  float restore;
  float radius;
  float normal[3];
  int pad[3];
};

// A neighbour-cell offset.  side: 0 same cell, +1 this cell is the
// reference's i cell, -1 the neighbour is.  n: unit direction of the
// forward shift (the sort axis runner_dosort uses for this pair).
// flags, bit 31: the sort keys of this offset can drop the whole cell pair
// (runner_dopair's outer-loop break on a negative key distance); bits 0-5
// are then the near-bound bits a witness particle of this cell must show and
// bits 8-13 those of the neighbour cell (bit 2k: coordinate k near 0, bit
// 2k+1: near the cell width).  Bit 30: the sort-key window can reject a pair
// inside the cutoff; without it, every in-cutoff pair of a cell pair that is
// not dropped passes the window.
struct DevOffset {
  int ox, oy, oz, side;
  float nx, ny, nz;
  int flags;
};

constexpr int kOffDrop = (int)0x80000000u, kOffWindow = 0x40000000;

// Unit direction n = f / |f| of the forward shift f = side * sign(ox, oy, oz)
// as the offset table holds it: components 0 or +-1/sqrt(|f|^2), with
// 1/sqrt(2) and 1/sqrt(3) rounded to float as (float)(1.0 / sqrt(k)) on the
// host.
__host__ __device__ inline void offset_dir(int ox, int oy, int oz, int side, float &nx, float &ny, float &nz) {
  const int gx = (ox > 0) - (ox < 0), gy = (oy > 0) - (oy < 0), gz = (oz > 0) - (oz < 0);
  const int l2 = gx * gx + gy * gy + gz * gz;
  const float il = l2 == 1 ? 1.0f : (l2 == 2 ? 0.70710677f : 0.57735026f);
  nx = (float)(side * gx) * il;
  ny = (float)(side * gy) * il;
  nz = (float)(side * gz) * il;
}

struct DevPot {
  uint32_t kind, flags;
  float alpha0, alpha1, alpha2;
  float a, b, r0_plusone;
  int n, coff;
  float dpd_alpha, dpd_gamma, dpd_sigma;
  float a2;
  int pad[2];
};

// Single-body force of a type (engine's psbs[type], one per type):
// MxForce.cpp:330-385.
enum : int { kSbNone = 0, kSbGaussian = 1, kSbBerendsen = 2, kSbConstant = 3, kSbFriction = 4 };

struct DevType {
  double mass, imass;
  double target_energy;  // MxParticleType::target_energy (Berendsen)
  float radius;
  int dynamics;
  int sb;                // kSb*
  float rnd_mean, rnd_std;
  unsigned rnd_steps;
  float sb_coef;         // Berendsen 1/tau, friction coefficient
  float cf[3];           // constant force
  int spec_const;        // bit k: species k is constant (not integrated)
  int pad;
};

// One flux of a type pair (MxFlux entries, Flux.hpp; MxFluxes_AddFlux,
// Flux.cpp:288-316): kind 0 Fick, 1 secrete, 2 uptake; species ia of type
// ta and ib of type tb.
struct DevFlux {
  int kind, ta, tb, ia, ib;
  float coef, decay, target;
};

}  // namespace

// One angle, dihedral or exclusion.  Each role's force goes to its own
// entry of the per-particle term list (pos), which the particle sums in its
// force gather, so every term is evaluated once and the sums stay in a fixed
// order.
enum : int { kTermExclusion = 0, kTermAngle = 1, kTermDihedral = 2 };

struct DevTerm {
  int id[4];   // particle ids: exclusion (i, j), angle (i, j, k), dihedral (i, j, k, l)
  int pos[4];  // list entry of each role
  int kind;    // kTerm*
  int pot;     // angle or dihedral potential
  int tid;     // index of the term within its kind (random plane of an angle)
  int pad;
};

struct DeviceState {
  hipStream_t stream = nullptr;
  // particle state, ping-pong
  float4 *x[2] = {nullptr, nullptr};
  float4 *v[2] = {nullptr, nullptr};
  float4 *pf[2] = {nullptr, nullptr};
  // Runge-Kutta 4 state: start position, slope sums, previous velocity slope
  float4 *rk_p0[2] = {nullptr, nullptr};
  float4 *rk_ax[2] = {nullptr, nullptr};
  float4 *rk_av[2] = {nullptr, nullptr};
  float4 *rk_vk[2] = {nullptr, nullptr};
  int *pid[2] = {nullptr, nullptr};
  int *cell[2] = {nullptr, nullptr};
  int *cstart[2] = {nullptr, nullptr};
  float4 *f = nullptr;
  int *slot_of = nullptr;
  int *ncount = nullptr;
  int *cnt = nullptr;
  int *inc = nullptr;
  int *cursor = nullptr;
  int *bsum = nullptr;
  int *cell_rep = nullptr;
  DevReplica *rep = nullptr;
  DevOffset *offs = nullptr;
  float *shifts = nullptr;
  int4 *tiles = nullptr;
  int2 *rep_tiles = nullptr;
  int4 *bricks = nullptr;
  int2 *rep_bricks = nullptr;
  int *ovf_list = nullptr;
  int *ovf_count = nullptr;
  unsigned char *dargs = nullptr;  // two StepArgs, one per step parity
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // two captured steps per parity
  DevPot *pots = nullptr;
  DevPot *bpots = nullptr;
  float *coef = nullptr;
  int *potmat = nullptr;
  DevType *types = nullptr;
  int *boff = nullptr, *bent = nullptr, *bi = nullptr, *bj = nullptr, *bpot = nullptr;
  double *bdis = nullptr;
  unsigned char *bact[2] = {nullptr, nullptr};
  DevTerm *terms = nullptr;   // grouped by replica, then by kind
  int4 *tblk = nullptr;       // energy blocks: replica, kind, first and end term
  int *trep_blk = nullptr;    // first energy block of each replica, nrep + 1
  double *tpart = nullptr;    // energy of each block
  int *toff = nullptr;        // per-particle term-list offsets, N + 1
  float4 *tf = nullptr;       // term-list entries: force on the particle
  float *te = nullptr;        // energy of each term
  DevPot *tpots = nullptr;    // angle then dihedral potentials
  double *obs_terms = nullptr;  // per replica: angle, dihedral, exclusion energy
  float4 *spec[2] = {nullptr, nullptr};   // species values, up to 4 per particle
  float4 *specq[2] = {nullptr, nullptr};  // species flux accumulators
  DevFlux *fluxes = nullptr;              // by ordered type pair (CSR)
  int *flux_off = nullptr;                // ntypes * ntypes + 1
  DevWall *walls = nullptr;     // the six sides
  int *wallpot = nullptr;       // 6 x ntypes: index into wpots, -1 none
  DevPot *wpots = nullptr;      // wall potentials
  std::vector<DevReplica> drep; // host copy of rep
  double *partials = nullptr;
  double *type_ke = nullptr;       // kinetic energy per type (Berendsen)
  double *type_ke_part = nullptr;  // per-block partial sums of it
  double *obs = nullptr;
  unsigned *time = nullptr;
  int *err = nullptr;  // [0] error code, [1] pid, [2] error state at step start
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // Per-step events of the force kernel (start, end) while timing is on.
  std::vector<hipEvent_t> kev;
  std::vector<void *> owned;
};

namespace {

struct StepArgs {
  const float4 G *__restrict__ x;
  const float4 G *__restrict__ v;
  const float4 G *__restrict__ pf;
  const int G *__restrict__ pid;
  const int G *__restrict__ cell;
  const int G *__restrict__ cstart;
  float4 G *__restrict__ x_out;
  float4 G *__restrict__ v_out;
  float4 G *__restrict__ pf_out;
  float4 G *__restrict__ f_out;
  // Runge-Kutta 4 state per slot (MxParticle p0, xk, vk): the stage's start
  // position, the running sums of position and velocity slopes, and the
  // previous stage's velocity slope; read from buffer 0, written to 1.
  const float4 G *__restrict__ rk_p0;
  const float4 G *__restrict__ rk_ax;
  const float4 G *__restrict__ rk_av;
  const float4 G *__restrict__ rk_vk;
  float4 G *__restrict__ rk_p0_out;
  float4 G *__restrict__ rk_ax_out;
  float4 G *__restrict__ rk_av_out;
  float4 G *__restrict__ rk_vk_out;
  int G *__restrict__ cell_out;
  const DevReplica G *__restrict__ rep;
  const DevOffset G *__restrict__ offs;
  const float G *__restrict__ shifts;
  const int4 G *__restrict__ tiles;
  int ntiles;
  const DevPot G *__restrict__ pots;
  const DevPot G *__restrict__ bpots;
  const float G *__restrict__ coef;
  const int G *__restrict__ potmat;
  int ntypes;
  const DevType G *__restrict__ types;
  const int G *__restrict__ boff;
  const int G *__restrict__ bent;
  const int G *__restrict__ bi;
  const int G *__restrict__ bj;
  const int G *__restrict__ bpot;
  const double G *__restrict__ bdis;
  const unsigned char G *__restrict__ bact_in;
  unsigned char G *__restrict__ bact_out;
  const int G *__restrict__ slot_of;
  const DevWall G *__restrict__ walls;
  const int G *__restrict__ wallpot;
  const DevPot G *__restrict__ wpots;
  // species transport (flux mode): values and accumulators in (buffer 0)
  // and out (buffer 1), the flux CSR by ordered type pair
  const float4 G *__restrict__ spec;
  const float4 G *__restrict__ specq;
  float4 G *__restrict__ spec_out;
  float4 G *__restrict__ specq_out;
  const DevFlux G *__restrict__ fluxes;
  const int G *__restrict__ flux_off;
  double dt_d, cutoff_d;
  const int G *__restrict__ toff;   // term-list offsets (null without terms)
  const float4 G *__restrict__ tf;  // term-list entries
  float dt;
  float cut2_pair, cut2_self;
  float dpd_delta;
  uint32_t seed_lo, seed_hi;
  int dpd_noise;
  const double G *type_ke;  // kinetic energy per type at the start of the step (Berendsen)
  int with_pf;               // a type carries a persistent random force
  const unsigned G *time;
  int G *err;
  double G *partials;
  int G *ncount;  // neighbour-count mode only
  int random;   // some type carries a persistent random force
  int npots;    // pair potentials
  float cutoff;       // number-density kernel width
  float nd_norm;      // 1 / (pi cutoff^3)
  int G *cnt;     // particles staying in each cell (written by its brick; the scan adds inc)
  int G *inc;     // particles entering each cell
};

// Loop-invariant scalars of the step, read once per block into registers:
// reading them through the argument pointer inside the neighbour loops makes
// each read a scalar load whose wait also drains the outstanding LDS loads.
struct LoopConsts {
  const float G *coef;
  int ntypes;
  float cut2_pair, cut2_self;
  float dpd_delta;
  uint32_t seed_lo, seed_hi;  // of the replica being computed
  int id_base;                // its first gid (stream id = gid - id_base)
  int dpd_noise;
  float cutoff, nd_norm;
};

__device__ __forceinline__ LoopConsts loop_consts(const StepArgs &A) {
  LoopConsts C;
  C.coef = A.coef;
  C.ntypes = A.ntypes;
  C.cut2_pair = A.cut2_pair;
  C.cut2_self = A.cut2_self;
  C.dpd_delta = A.dpd_delta;
  C.seed_lo = A.seed_lo;
  C.seed_hi = A.seed_hi;
  C.id_base = 0;
  C.dpd_noise = A.dpd_noise;
  C.cutoff = A.cutoff;
  C.nd_norm = A.nd_norm;
  return C;
}

enum : int {
  kModeTable = 1,
  kModeDPD = 2,
  kModeVirial = 4,
  kModeBonds = 8,
  kModeFlux = 16,
};

// Uniform noise of zero mean and unit variance on [-sqrt(3), sqrt(3)).
__device__ __forceinline__ float unit_uniform(uint32_t a) {
  return 1.7320508f * (2.0f * u01_closed_open(a) - 1.0f);
}

__device__ __forceinline__ float gauss_from(uint32_t a, uint32_t b) {
  float u1 = u01_open_closed(a);
  float u2 = u01_closed_open(b);
  return sqrtf(-2.0f * logf(u1)) * cosf(6.2831853f * u2);
}

// Random direction of length len for a pair closer than the potential's
// lower bound (the reference draws it from a per-cell Mersenne twister,
// dpd_eval.hpp:122-131; here it is keyed by the pair and the time).
__device__ __forceinline__ float3 overlap_dir(uint32_t pi, uint32_t pj, uint32_t t, uint32_t klo,
                                              uint32_t khi, float len) {
  u32x4 ctr = {{pi, pj, t, kStreamOverlap}};
  u32x4 r = philox4x32_10(ctr, klo, khi);
  u32x4 r2 = philox4x32_10(r, klo ^ 0x9E3779B9u, khi);
  float g0 = gauss_from(r.v[0], r.v[1]);
  float g1 = gauss_from(r.v[2], r.v[3]);
  float g2 = gauss_from(r2.v[0], r2.v[1]);
  float l = sqrtf(g0 * g0 + g1 * g1 + g2 * g2);
  return make_float3(g0 * len / l, g1 * len / l, g2 * len / l);
}

// x / y to about one ulp from the hardware reciprocal (v_rcp_f32).  This
// build passes -fhip-fp32-correctly-rounded-divide-sqrt, which makes both `/`
// and __fdividef the IEEE division sequence (about ten instructions); the
// pair forces and the number density only have to agree to 1e-5 and 1e-3,
// while r^2, the cutoff tests and the advance stay exact.
__device__ __forceinline__ float div_1ulp(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }

// potential_eval_ex (potential_eval.hpp:106-161).
__device__ __forceinline__ bool table_eval_ex(const DevPot &P, const float *__restrict__ coef, float ri,
                                              float rj, float r_in, float &e, float &f) {
  const float eps = 1.1920928955078125e-07f;
  float r = r_in;  // sqrt(r^2), correctly rounded
  float ro = r < eps ? eps : r;
  r = r < P.a ? P.a : r;
  if (P.flags & kPotFlagScaled)
    r = r / (ri + rj);
  else if (P.flags & kPotFlagShifted)
    r = r - (ri + rj) + P.r0_plusone;
  float indf = fmaxf(0.0f, P.alpha0 + r * (P.alpha1 + r * P.alpha2));
  unsigned ind = (unsigned)indf;
  if (r > P.b || ind > (unsigned)P.n) return false;
  const float4 *c4 = reinterpret_cast<const float4 *>(coef + P.coff + (size_t)ind * kPotChunk);
  float4 c0 = c4[0], c1 = c4[1];
  // The interval's quintic and its derivative by Horner's rule in fused
  // multiply-adds: half the instructions and fewer roundings than the
  // reference's separate products and sums (the build keeps contraction off
  // for the pair-set test and the advance, which must round as the
  // reference's do).
  const float x = (r - c0.x) * c0.y;
  float ee = __builtin_fmaf(c0.z, x, c0.w);
  float eff = c0.z;
  eff = __builtin_fmaf(eff, x, ee);
  ee = __builtin_fmaf(ee, x, c1.x);
  eff = __builtin_fmaf(eff, x, ee);
  ee = __builtin_fmaf(ee, x, c1.y);
  eff = __builtin_fmaf(eff, x, ee);
  ee = __builtin_fmaf(ee, x, c1.z);
  eff = __builtin_fmaf(eff, x, ee);
  ee = __builtin_fmaf(ee, x, c1.w);
  e = ee;
  f = div_1ulp(eff * c0.y, ro);
  return true;
}

// potential_eval (potential_eval.hpp:67-104), used for bonds.
__device__ __forceinline__ void table_eval(const DevPot &P, const float *__restrict__ coef, float r2,
                                           float &e, float &f) {
  float r = sqrtf(r2);
  int ind = (int)fmaxf(0.0f, P.alpha0 + r * (P.alpha1 + r * P.alpha2));
  // r^2 is clamped to [a^2, b^2] by the caller, so ind <= n; the clamp only
  // keeps non-finite input inside the table.
  ind = min(max(ind, 0), P.n);
  const float4 *c4 = reinterpret_cast<const float4 *>(coef + P.coff + (size_t)ind * kPotChunk);
  float4 c0 = c4[0], c1 = c4[1];
  float x = (r - c0.x) * c0.y;
  float ee = c0.z * x + c0.w;
  float eff = c0.z;
  eff = eff * x + ee;
  ee = ee * x + c1.x;
  eff = eff * x + ee;
  ee = ee * x + c1.y;
  eff = eff * x + ee;
  ee = ee * x + c1.z;
  eff = eff * x + ee;
  ee = ee * x + c1.w;
  e = ee;
  f = eff * c0.y / r;
}

__device__ __forceinline__ unsigned sort_key(float bias, float dscale, float x, float y, float z, float nx,
                                             float ny, float nz) {
  return ((unsigned)(dscale * (bias + x * nx + y * ny + z * nz))) & 0xffffu;
}

// True when runner_dopair yields no pair for the cell pair with i-cell
// particles x[ib, ie) and j-cell particles x[jb, je) (its loop over the i
// cell stops at once, runner_dopair.cpp:164-166): the smallest j key plus
// dnshift minus the largest i key is negative.  (sx, sy, sz): the pair's
// shift in either orientation (only its norm enters).
__device__ __forceinline__ bool pair_dropped(const DevReplica &R, float nx, float ny, float nz, float sx, float sy,
                                             float sz, const float4 *x, int ib, int ie, int jb, int je) {
  if (ib >= ie || jb >= je) return false;
  unsigned kimax = 0u, kjmin = 0xffffu;
  for (int q = ib; q < ie; ++q) {
    const float4 p = x[q];
    kimax = max(kimax, sort_key(R.bias, R.dscale, p.x, p.y, p.z, nx, ny, nz));
  }
  for (int q = jb; q < je; ++q) {
    const float4 p = x[q];
    kjmin = min(kjmin, sort_key(R.bias, R.dscale, p.x, p.y, p.z, nx, ny, nz));
  }
  const int dn = (int)(R.dscale * sqrtf(sx * sx + sy * sy + sz * sz));
  return (int)kjmin + dn - (int)kimax < 0;
}

// Near-bound bits of a cell-relative position (bit 2k: coordinate k below
// eps, bit 2k+1: above the cell width minus eps).
__device__ __forceinline__ int near_bits(const DevReplica &R, const float4 &p) {
  const float e = R.near_eps;
  return (int)(p.x < e) | ((int)(p.x > R.h[0] - e) << 1) | ((int)(p.y < e) << 2) | ((int)(p.y > R.h[1] - e) << 3) |
         ((int)(p.z < e) << 4) | ((int)(p.z > R.h[2] - e) << 5);
}

// Cubic-spline number-density kernel W(r^2, h) (smoothing_kernel.hpp:28-50)
// with the exact square root (the reference uses an approximate x86 rsqrt).
__device__ __forceinline__ float w_cubic(float r, float h, float inv_norm) {
  const float x = div_1ulp(r, h);
  float y;
  if (x < 1.f) {
    const float x2 = x * x;
    y = 1.f - 1.5f * x2 + 0.75f * x2 * x;
  } else if (x < 2.f) {
    const float a = 2.f - x;
    y = 0.25f * a * a * a;
  } else {
    y = 0.f;
  }
  return y * inv_norm;
}

__device__ __forceinline__ void cell_coords(int lc, const DevReplica &R, int &lx, int &ly, int &lz) {
  lz = lc % R.cdim[2];
  int t = lc / R.cdim[2];
  ly = t % R.cdim[1];
  lx = t / R.cdim[1];
}

__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Neighbour cell coordinate along an axis: wrapped on a periodic axis, -1
// beyond a wall (space_init creates no cell pair across it,
// space.cpp:827-888).
__device__ __forceinline__ int nbr_coord(int a, int n, int periodic) {
  return periodic ? wrap(a, n) : ((a < 0 || a >= n) ? -1 : a);
}

// Walls of the cell (lx, ly, lz) as space_init flags them
// (cell_active_*, space.cpp:700-723): per axis the low wall for the first
// cell, else the high wall for the last.  Bit i is side i.
__device__ __forceinline__ int cell_walls(const DevReplica &R, int lx, int ly, int lz) {
  const int act = R.wall >> kWallActiveShift;
  int m = 0;
  if (lx == 0 && (act & 1))
    m |= 1;
  else if (lx + 1 == R.cdim[0] && (act & 2))
    m |= 2;
  if (ly == 0 && (act & 4))
    m |= 4;
  else if (ly + 1 == R.cdim[1] && (act & 8))
    m |= 8;
  if (lz == 0 && (act & 16))
    m |= 16;
  else if (lz + 1 == R.cdim[2] && (act & 32))
    m |= 32;
  return m;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Neighbour data read from global memory (slot q) or from LDS (index q).
struct GlobalNbr {
  const float4 *__restrict__ vv;
  const int *__restrict__ pp;
  __device__ __forceinline__ float4 v(int q) const { return vv[q]; }
  __device__ __forceinline__ int pid(int q) const { return pp[q]; }
};

// In LDS the velocity's w holds the particle id (DPD bricks stage
// velocities); without staged velocities the id comes from global memory
// through the halo map (only pairs closer than the potential's lower bound
// need it).
struct LdsNbr {
  const float4 *vv;
  bool staged_v;
  const int *hstart;
  const int *hgb;
  int nhalo;
  const int *__restrict__ gpid;
  __device__ __forceinline__ float4 v(int q) const { return vv[q]; }
  __device__ int pid(int q) const;
};

// Pair distance from r^2, the same in the LDS kernel and its global-memory
// fallback so a particle's force does not depend on which one ran its brick:
// table potentials index by r and keep the correctly rounded root; a mode
// with DPD potentials only takes the one-ulp hardware square root.
template <int MODE>
__device__ __forceinline__ float pair_root(float r2) {
  return (MODE & kModeTable) ? sqrtf(r2) : __builtin_amdgcn_sqrtf(r2);
}

// Accumulates one pair into the owner's force.  dx is in the reference
// orientation (x_i - x_j); p_is_i says which end the owner is.
// Returns whether the pair adds to the number density (no potential, or the
// potential was evaluated in range: potential_eval_super_ex adds it only on
// success, dpd_eval.hpp:135-160).
// Per-type-pair tables read by the pair code (staged in LDS by the kernels).
struct PairTabs {
  const DevPot *pots;
  const int *potmat;
  const float *radius;
};

template <int MODE, typename NB>
__device__ __forceinline__ bool pair_force(const LoopConsts &A, const PairTabs &tb, const NB &nb, int ti, int tj,
                                           bool p_is_i,
                                           int mypid, int q,
                                           bool self_pair, float dxx, float dxy, float dxz, float r2, float r,
                                           const float4 &vs, float &fx, float &fy, float &fz, double *acc,
                                           unsigned t1, uint32_t tk) {
  int pidx = tb.potmat[ti * A.ntypes + tj];
  if (pidx < 0) return true;
  const DevPot P = tb.pots[pidx];
  if (r2 < P.a2) {
    // Stream ids within the replica (gid - id_base).
    const int mine = mypid - A.id_base, other = nb.pid(q) - A.id_base;
    uint32_t ki, kj;
    float sgn = 1.0f;
    if (self_pair) {
      ki = (uint32_t)min(mine, other);
      kj = (uint32_t)max(mine, other);
      sgn = (mine == (int)ki) ? 1.0f : -1.0f;
    } else {
      ki = (uint32_t)(p_is_i ? mine : other);
      kj = (uint32_t)(p_is_i ? other : mine);
    }
    float3 d = overlap_dir(ki, kj, t1, A.seed_lo, A.seed_hi, P.a);
    dxx = sgn * d.x;
    dxy = sgn * d.y;
    dxz = sgn * d.z;
    r2 = P.a * P.a;
    r = sqrtf(r2);
  }
  const float ri = tb.radius[ti], rj = tb.radius[tj];
  // Without table potentials in the mode every potential is DPD.
  if ((MODE & kModeDPD) && (!(MODE & kModeTable) || P.kind == kPotKindDPD)) {
    bool shifted = P.flags & kPotFlagShifted;
    float cutoff = shifted ? (P.b + ri + rj) : P.b;
    // Beyond the DPD cutoff the pair adds an exact zero (a mask, not a
    // branch; the flux mode keeps the branch, where it measures faster).
    const bool in_range = r2 <= cutoff * cutoff;
    if ((MODE & kModeFlux) && !in_range) return false;
    const float ir = __builtin_amdgcn_rcpf(r);
    float ex = dxx * ir, ey = dxy * ir, ez = dxz * ir;
    float4 vq = nb.v(q);
    // v_i - v_j: the owner's difference, negated when the owner is j (the
    // negation is exact, b - a = -(a - b)), so no branch on the end.
    const float sg = p_is_i ? 1.0f : -1.0f;
    const float vx = sg * (vs.x - vq.x), vy = sg * (vs.y - vq.y), vz = sg * (vs.z - vq.z);
    float shifted_r = shifted ? r - ri - rj : r;
    float omega_c = shifted_r < 0.f ? 1.f : (1 - div_1ulp(shifted_r, cutoff));
    float fc = P.dpd_alpha * omega_c;
    float omega_d = omega_c * omega_c;
    float ev = ex * vx + ey * vy + ez * vz;
    float fd = -P.dpd_gamma * omega_d * ev;
    float fr;
    if (A.dpd_noise) {
      const int qpid = nb.pid(q);
      fr = P.dpd_sigma * omega_c *
           pair_uniform((uint32_t)(min(mypid, qpid) - A.id_base), (uint32_t)(max(mypid, qpid) - A.id_base), tk) *
           A.dpd_delta;
    } else {
      fr = P.dpd_sigma * omega_c * A.dpd_delta;
    }
    float f = in_range ? fc + fd + fr : 0.0f;
    float wx = f * ex, wy = f * ey, wz = f * ez;  // force on i
    fx += sg * wx;
    fy += sg * wy;
    fz += sg * wz;
    if (MODE & kModeVirial) {
      acc[3] += 0.5 * dxx * wx;
      acc[4] += 0.5 * dxx * wy;
      acc[5] += 0.5 * dxx * wz;
      acc[6] += 0.5 * dxy * wx;
      acc[7] += 0.5 * dxy * wy;
      acc[8] += 0.5 * dxy * wz;
      acc[9] += 0.5 * dxz * wx;
      acc[10] += 0.5 * dxz * wy;
      acc[11] += 0.5 * dxz * wz;
    }
    return in_range;
  }
  if (MODE & kModeTable) {
    float e, f;
    if (!table_eval_ex(P, A.coef, ri, rj, r, e, f)) return false;
    float wx = f * dxx, wy = f * dxy, wz = f * dxz;
    if (p_is_i) {
      fx -= wx;
      fy -= wy;
      fz -= wz;
    } else {
      fx += wx;
      fy += wy;
      fz += wz;
    }
    acc[0] += 0.5 * (double)e;
    if (MODE & kModeVirial) {
      acc[3] -= 0.5 * dxx * wx;
      acc[4] -= 0.5 * dxx * wy;
      acc[5] -= 0.5 * dxx * wz;
      acc[6] -= 0.5 * dxy * wx;
      acc[7] -= 0.5 * dxy * wy;
      acc[8] -= 0.5 * dxy * wz;
      acc[9] -= 0.5 * dxz * wx;
      acc[10] -= 0.5 * dxz * wy;
      acc[11] -= 0.5 * dxz * wz;
    }
  }
  return true;
}

// Visits every in-cutoff neighbour of slot s in the reference orientation,
// applying the sorted-axis window of runner_dopair (keys of runner_dosort,
// runner_dosort.cpp:126-137; window test runner_dopair.cpp:170,182) so the
// pair set is exactly the reference's.  fn(q, p_is_i, self, dx, r2).
template <typename Fn>
__device__ __forceinline__ void for_each_neighbour(const StepArgs &A, const LoopConsts &C, const DevReplica &R,
                                                   int s, int lx, int ly, int lz, float4 xs, Fn &&fn) {
  const int cd0 = R.cdim[0], cd1 = R.cdim[1], cd2 = R.cdim[2];
  const int w0 = 2 * R.span[0] + 1, w1 = 2 * R.span[1] + 1, w2 = 2 * R.span[2] + 1;
  for (int o = 0; o < R.n_off; ++o) {
    const DevOffset off = A.offs[R.off_base + o];
    const int nx = nbr_coord(lx + off.ox, cd0, R.wall & 1), ny = nbr_coord(ly + off.oy, cd1, R.wall & 2),
              nz = nbr_coord(lz + off.oz, cd2, R.wall & 4);
    if ((nx | ny | nz) < 0) continue;
    int c2 = R.cell_base + (nx * cd1 + ny) * cd2 + nz;
    int qb = A.cstart[c2], qe = A.cstart[c2 + 1];
    if (off.side == 0) {
      for (int q = qb; q < qe; ++q) {
        if (q == s) continue;
        float4 xq = A.x[q];
        float dx = xs.x - xq.x, dy = xs.y - xq.y, dz = xs.z - xq.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > C.cut2_self) continue;
        fn(q, true, true, dx, dy, dz, r2, xq);
      }
      continue;
    }
    // Shift from the reference i cell to the j cell.
    float sx, sy, sz;
    if (off.side > 0) {
      sx = A.shifts[R.shift_base[0] + lx * w0 + (off.ox + R.span[0])];
      sy = A.shifts[R.shift_base[1] + ly * w1 + (off.oy + R.span[1])];
      sz = A.shifts[R.shift_base[2] + lz * w2 + (off.oz + R.span[2])];
    } else {
      sx = A.shifts[R.shift_base[0] + nx * w0 + (-off.ox + R.span[0])];
      sy = A.shifts[R.shift_base[1] + ny * w1 + (-off.oy + R.span[1])];
      sz = A.shifts[R.shift_base[2] + nz * w2 + (-off.oz + R.span[2])];
    }
    if (off.flags & kOffDrop) {
      const int c = R.cell_base + (lx * cd1 + ly) * cd2 + lz;
      const int hb = A.cstart[c], he = A.cstart[c + 1];
      if (off.side > 0 ? pair_dropped(R, off.nx, off.ny, off.nz, sx, sy, sz, A.x, hb, he, qb, qe)
                       : pair_dropped(R, off.nx, off.ny, off.nz, sx, sy, sz, A.x, qb, qe, hb, he))
        continue;
    }
    float nshift = sqrtf(sx * sx + sy * sy + sz * sz);
    unsigned dnshift = (unsigned)(int)(R.dscale * nshift);
    unsigned dmax = (unsigned)R.dmaxdist;
    unsigned kp = sort_key(R.bias, R.dscale, xs.x, xs.y, xs.z, off.nx, off.ny, off.nz);
    if (off.side > 0) {
      float pix = xs.x - sx, piy = xs.y - sy, piz = xs.z - sz;
      for (int q = qb; q < qe; ++q) {
        float4 xq = A.x[q];
        float dx = pix - xq.x, dy = piy - xq.y, dz = piz - xq.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > C.cut2_pair) continue;
        unsigned kq = sort_key(R.bias, R.dscale, xq.x, xq.y, xq.z, off.nx, off.ny, off.nz);
        if (!(kq + dnshift - kp < dmax)) continue;
        fn(q, true, false, dx, dy, dz, r2, xq);
      }
    } else {
      for (int q = qb; q < qe; ++q) {
        float4 xq = A.x[q];
        float qix = xq.x - sx, qiy = xq.y - sy, qiz = xq.z - sz;
        float dx = qix - xs.x, dy = qiy - xs.y, dz = qiz - xs.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > C.cut2_pair) continue;
        unsigned kq = sort_key(R.bias, R.dscale, xq.x, xq.y, xq.z, off.nx, off.ny, off.nz);
        if (!(kp + dnshift - kq < dmax)) continue;
        fn(q, false, false, dx, dy, dz, r2, xq);
      }
    }
  }
}

constexpr int kScanBlock = 256, kScanItems = 4, kScanChunk = kScanBlock * kScanItems;

__device__ __forceinline__ int block_exclusive_scan(int v, int *lds, int &total) {
  // Inclusive scan over the block with wave shuffles, then exclusive.
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  int wsum = 0, tot = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    if (w < wave) wsum += lds[w];
    tot += lds[w];
  }
  __syncthreads();
  total = tot;
  return wsum + x - v;
}

// The type's single-body force, added in runner_doself before the pair
// loop (runner_dopair.cpp:519-523), MxForce.cpp:330-385:
//   gaussian  f += pf, pf = MxRandomVector(mean, std) refreshed when
//             (time + id) % steps == 0;
//   berendsen f += itau (E_target / E_kin(type) - 1) v;
//   constant  f += F;
//   friction  f += -coef |v|^2 v + pf (pf refreshed as gaussian's).
// refresh: INTEGRATOR_UPDATE_PERSISTENTFORCE (set for the Euler step and the
// first Runge-Kutta stage only).  The random vector comes from Philox keyed
// by (id, time).
__device__ __forceinline__ void single_body(const StepArgs &A, const LoopConsts &C, int s, int mypid, int ti,
                                            const DevType &T, unsigned t1, bool refresh, const float4 &vs, float &fx,
                                            float &fy, float &fz, float4 &pfs) {
  pfs = A.with_pf ? float4(A.pf[s]) : make_float4(0.f, 0.f, 0.f, 0.f);
  if (T.sb == kSbNone) return;
  if (T.sb == kSbGaussian || T.sb == kSbFriction) {
    // Stream id within the replica (MxParticle::id of a standalone engine).
    const uint32_t sid = (uint32_t)(mypid - C.id_base);
    if (refresh && ((uint64_t)t1 + (uint64_t)sid) % T.rnd_steps == 0) {
      u32x4 ctr = {{sid, t1, kStreamSingle, 0u}};
      u32x4 r = philox4x32_10(ctr, C.seed_lo, C.seed_hi);
      float g = gauss_from(r.v[0], r.v[1]);
      float mag = T.rnd_mean + T.rnd_std * g;
      float theta = 6.2831853f * u01_closed_open(r.v[2]);
      float phi = acosf(1.0f - 2.0f * u01_closed_open(r.v[3]));
      pfs.x = mag * sinf(phi) * cosf(theta);
      pfs.y = mag * sinf(phi) * sinf(theta);
      pfs.z = mag * cosf(phi);
    }
  }
  if (T.sb == kSbGaussian) {
    fx += pfs.x;
    fy += pfs.y;
    fz += pfs.z;
  } else if (T.sb == kSbFriction) {
    const float v2 = vs.x * vs.x + vs.y * vs.y + vs.z * vs.z;
    const float scale = (float)(-1. * (double)T.sb_coef * (double)v2);
    fx += scale * vs.x + pfs.x;
    fy += scale * vs.y + pfs.y;
    fz += scale * vs.z + pfs.z;
  } else if (T.sb == kSbBerendsen) {
    const float scale = (float)((double)T.sb_coef * ((T.target_energy / A.type_ke[ti]) - 1.0));
    fx += scale * vs.x;
    fy += scale * vs.y;
    fz += scale * vs.z;
  } else if (T.sb == kSbConstant) {
    fx += T.cf[0];
    fy += T.cf[1];
    fz += T.cf[2];
  }
}

// Bonds of one particle (bond_eval, bond.cpp:99-295), gathered per particle.
// Both ends evaluate the bond in its (i, j) orientation; the i end keeps the
// energy and writes the dissociation flag.
template <int MODE>
__device__ __forceinline__ void bond_forces(const StepArgs &A, const DevReplica &R, int s, int mypid, float &fx,
                                            float &fy, float &fz, double *acc) {
  if (!A.boff) return;
  const int b0 = A.boff[mypid], b1 = A.boff[mypid + 1];
  for (int k = b0; k < b1; ++k) {
    const int ent = A.bent[k];
    const int bid = ent >> 1;
    const bool p_is_i = (ent & 1) == 0;
    if (!A.bact_in[bid]) {
      if (p_is_i) A.bact_out[bid] = 0;
      continue;
    }
    const int si = p_is_i ? s : A.slot_of[A.bi[bid]];
    const int sj = p_is_i ? A.slot_of[A.bj[bid]] : s;
    const float4 xi = A.x[si], xj = A.x[sj];
    int lix, liy, liz, ljx, ljy, ljz;
    cell_coords(A.cell[si] - R.cell_base, R, lix, liy, liz);
    cell_coords(A.cell[sj] - R.cell_base, R, ljx, ljy, ljz);
    int sh[3] = {lix - ljx, liy - ljy, liz - ljz};
    for (int d = 0; d < 3; ++d) sh[d] = sh[d] > 1 ? -1 : (sh[d] < -1 ? 1 : sh[d]);
    const float pix = (float)((double)xi.x + R.hd[0] * sh[0]);
    const float piy = (float)((double)xi.y + R.hd[1] * sh[1]);
    const float piz = (float)((double)xi.z + R.hd[2] * sh[2]);
    const float dx = pix - xj.x, dy = piy - xj.y, dz = piz - xj.z;
    float r2 = dx * dx + dy * dy + dz * dz;
    const DevPot P = A.bpots[A.bpot[bid]];
    const float lo = P.a * P.a, hi = P.b * P.b;
    if (r2 < lo || r2 > hi) r2 = fmaxf(lo, fminf(hi, r2));
    float ee, eff;
    table_eval(P, A.coef, r2, ee, eff);
    if ((double)eff >= A.bdis[bid]) {
      if (p_is_i) A.bact_out[bid] = 0;
      continue;
    }
    if (p_is_i) A.bact_out[bid] = 1;
    const float wx = eff * dx, wy = eff * dy, wz = eff * dz;
    if (p_is_i) {
      fx -= wx;
      fy -= wy;
      fz -= wz;
      acc[1] += (double)ee;
      if (MODE & kModeVirial) {
        acc[3] -= dx * wx;
        acc[4] -= dx * wy;
        acc[5] -= dx * wz;
        acc[6] -= dy * wx;
        acc[7] -= dy * wy;
        acc[8] -= dy * wz;
        acc[9] -= dz * wx;
        acc[10] -= dz * wy;
        acc[11] -= dz * wz;
      }
    } else {
      fx += wx;
      fy += wy;
      fz += wz;
    }
  }
}

// boundary_eval (boundary_eval.hpp:162-221) for a particle in a cell
// touching walls m: each active side with a potential for the particle's
// type acts within the potential's range b, along the axis, from the wall
// (distance x on a low side, h - x on a high side).  As the reference does,
// the offset vector is kept from side to side (a later side overwrites only
// its own axis) and is replaced by normal * a when r^2 < a
// (boundary_potential_eval_ex, :124-158); DPD walls take the reference's
// noise-free dpd_boundary_eval (dpd_eval.hpp:71-106) against the wall's
// velocity.
__device__ void wall_forces(const StepArgs &A, const DevReplica &R, int m, int ti, float radius, const float4 &xs,
                            const float4 &vs, float dpd_delta, float &fx, float &fy, float &fz, double &epot) {
  float dx[3] = {0.0f, 0.0f, 0.0f};
  const float X[3] = {xs.x, xs.y, xs.z};
  for (int side = 0; side < 6; ++side) {
    if (!(m & (1 << side))) continue;
    const int pidx = A.wallpot[side * A.ntypes + ti];
    if (pidx < 0) continue;
    const DevPot P = A.wpots[pidx];
    const int axis = side >> 1;
    const bool low = (side & 1) == 0;
    const float r = low ? X[axis] : (float)(R.hd[axis] - (double)X[axis]);
    if (!(r <= P.b)) continue;
    dx[axis] = low ? r : -r;
    float r2 = r * r;
    const DevWall W = A.walls[side];
    if (r2 < P.a) {
      r2 = P.a * P.a;
      dx[0] = W.normal[0] * P.a;
      dx[1] = W.normal[1] * P.a;
      dx[2] = W.normal[2] * P.a;
    }
    if (P.kind == kPotKindDPD) {
      const float cutoff = P.b;
      if (r2 > cutoff * cutoff) continue;
      const float rr = sqrtf(r2);
      const float ex = dx[0] / rr, ey = dx[1] / rr, ez = dx[2] / rr;
      const float vx = vs.x - W.vel[0], vy = vs.y - W.vel[1], vz = vs.z - W.vel[2];
      const float omega_c = 1 - rr / cutoff;
      const float fc = P.dpd_alpha * omega_c;
      const float omega_d = omega_c * omega_c;
      const float fd = -P.dpd_gamma * omega_d * (ex * vx + ey * vy + ez * vz);
      const float fr = P.dpd_sigma * omega_c * dpd_delta;
      const float f = fc + fd + fr;
      fx = fx + f * ex;
      fy = fy + f * ey;
      fz = fz + f * ez;
    } else {
      float e, f;
      if (!table_eval_ex(P, A.coef, radius, W.radius, sqrtf(r2), e, f)) continue;
      fx -= f * dx[0];
      fy -= f * dx[1];
      fz -= f * dx[2];
      epot += (double)e;
    }
  }
}

// Forces of the angles, dihedrals and exclusions on one particle: the
// entries k_terms wrote into its term list.
__device__ __forceinline__ void term_forces(const StepArgs &A, int mypid, float &fx, float &fy, float &fz) {
  if (!A.toff) return;
  const int t0 = A.toff[mypid], t1 = A.toff[mypid + 1];
  for (int k = t0; k < t1; ++k) {
    const float4 t = A.tf[k];
    fx += t.x;
    fy += t.y;
    fz += t.z;
  }
}

// potential_eval_r (potential_eval.hpp:180-208): the table at r itself, no
// square root; f is dE/dr.  r has been clamped to [a, b] by the caller.
__device__ __forceinline__ void table_eval_r(const DevPot &P, const float *__restrict__ coef, float r, float &e,
                                             float &f) {
  int ind = (int)fmaxf(0.0f, P.alpha0 + r * (P.alpha1 + r * P.alpha2));
  ind = min(max(ind, 0), P.n);
  const float4 *c4 = reinterpret_cast<const float4 *>(coef + P.coff + (size_t)ind * kPotChunk);
  float4 c0 = c4[0], c1 = c4[1];
  float x = (r - c0.x) * c0.y;
  float ee = c0.z * x + c0.w;
  float eff = c0.z;
  eff = eff * x + ee;
  ee = ee * x + c1.x;
  eff = eff * x + ee;
  ee = ee * x + c1.y;
  eff = eff * x + ee;
  ee = ee * x + c1.z;
  eff = eff * x + ee;
  ee = ee * x + c1.w;
  e = ee;
  f = eff * c0.y;
}

__device__ __forceinline__ void cross3(const float a[3], const float b[3], float out[3]) {
  out[0] = a[1] * b[2] - a[2] * b[1];
  out[1] = a[2] * b[0] - a[0] * b[2];
  out[2] = a[0] * b[1] - a[1] * b[0];
}

// Cosine of the dihedral angle of (xi, xj, xk, xl) and its gradient with
// respect to xi, xj and xl (the gradient for xk is minus their sum).
__device__ __forceinline__ void dihedral_geometry(float3 xi, float3 xj, float3 xk, float3 xl, float &c, float gi[3],
                                                  float gj[3], float gl[3]) {
  const float b1[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
  const float b2[3] = {xk.x - xj.x, xk.y - xj.y, xk.z - xj.z};
  const float b3[3] = {xl.x - xk.x, xl.y - xk.y, xl.z - xk.z};
  float m[3], n[3];
  cross3(b1, b2, m);
  cross3(b2, b3, n);
  const float mm = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
  const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  const float mn = m[0] * n[0] + m[1] * n[1] + m[2] * n[2];
  const float inv = 1.0f / sqrtf(mm * nn);
  c = mn * inv;
  const float cm = c / mm, cn = c / nn;
  float gm[3], gn[3];
  for (int d = 0; d < 3; ++d) {
    gm[d] = n[d] * inv - cm * m[d];
    gn[d] = m[d] * inv - cn * n[d];
  }
  float d1[3], d3[3], t1[3], t2[3];
  cross3(b2, gm, d1);
  cross3(gn, b2, d3);
  cross3(gm, b1, t1);
  cross3(b3, gn, t2);
  for (int d = 0; d < 3; ++d) {
    gi[d] = -d1[d];
    gj[d] = d1[d] - (t1[d] + t2[d]);
    gl[d] = d3[d];
  }
}

struct TermArgs {
  const DevTerm *__restrict__ terms;
  int nterms;
  const float4 *__restrict__ x;
  const int *__restrict__ cell;
  const int *__restrict__ cell_rep;
  const int *__restrict__ slot_of;
  const DevReplica *__restrict__ rep;
  const DevPot *__restrict__ pots;   // pair potentials (exclusions)
  const DevPot *__restrict__ tpots;  // angle and dihedral potentials
  const float *__restrict__ coef;
  const int *__restrict__ potmat;
  int ntypes;
  float cut2;
  const unsigned *__restrict__ time;
  uint32_t seed_lo, seed_hi;
  float4 *__restrict__ tf;
  float *__restrict__ te;
};

// Position of the particle in slot s relative to the cell (jx, jy, jz): its
// cell offset folded to -1, 0, +1 as the bonded evaluators do
// (angle.cpp:148-163), the shift added in double and rounded once.
__device__ __forceinline__ float3 rel_pos(const TermArgs &T, const DevReplica &R, int s, int jx, int jy, int jz) {
  int lx, ly, lz;
  cell_coords(T.cell[s] - R.cell_base, R, lx, ly, lz);
  int sh[3] = {lx - jx, ly - jy, lz - jz};
  for (int d = 0; d < 3; ++d) sh[d] = sh[d] > 1 ? -1 : (sh[d] < -1 ? 1 : sh[d]);
  const float4 p = T.x[s];
  return make_float3((float)((double)p.x + R.hd[0] * sh[0]), (float)((double)p.y + R.hd[1] * sh[1]),
                     (float)((double)p.z + R.hd[2] * sh[2]));
}

__device__ __forceinline__ void put(float4 *tf, int pos, float x, float y, float z) {
  tf[pos] = make_float4(x, y, z, 0.0f);
}

// Angles (angle_eval, angle.cpp:92-354), dihedrals (dihedral_eval,
// dihedral.cpp:90-330) and exclusions (exclusion_eval, exclusion.cpp:83-260),
// one thread per term, from the positions at the start of the force phase.
__global__ void __launch_bounds__(256) k_terms(TermArgs T) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T.nterms) return;
  const DevTerm D = T.terms[t];
  const int sj = T.slot_of[D.id[1]];
  const DevReplica R = T.rep[T.cell_rep[T.cell[sj]]];
  int jx, jy, jz;
  cell_coords(T.cell[sj] - R.cell_base, R, jx, jy, jz);
  const float4 pj = T.x[sj];
  float ee = 0.0f;
  if (D.kind == kTermExclusion) {
    // dx = xi - xj + h shift (float difference, double shift), the pair
    // potential of the two types at r^2, forces and energy subtracted.
    const int si = T.slot_of[D.id[0]];
    int ix, iy, iz;
    cell_coords(T.cell[si] - R.cell_base, R, ix, iy, iz);
    int sh[3] = {ix - jx, iy - jy, iz - jz};
    for (int d = 0; d < 3; ++d) sh[d] = sh[d] > 1 ? -1 : (sh[d] < -1 ? 1 : sh[d]);
    const float4 pi = T.x[si];
    const float dx = (float)((double)(pi.x - pj.x) + R.hd[0] * sh[0]);
    const float dy = (float)((double)(pi.y - pj.y) + R.hd[1] * sh[1]);
    const float dz = (float)((double)(pi.z - pj.z) + R.hd[2] * sh[2]);
    const float r2 = dx * dx + dy * dy + dz * dz;
    const int ti = __float_as_int(pi.w) & 0xffff, tj = __float_as_int(pj.w) & 0xffff;
    const int pidx = T.potmat[tj * T.ntypes + ti];
    float wx = 0, wy = 0, wz = 0;
    if (r2 <= T.cut2 && pidx >= 0) {
      const DevPot P = T.pots[pidx];
      if (P.kind == kPotKindTable) {
        const float r = sqrtf(r2);
        const int ind = (int)fmaxf(0.0f, P.alpha0 + r * (P.alpha1 + r * P.alpha2));
        if (ind <= P.n) {
          float eff;
          table_eval(P, T.coef, r2, ee, eff);
          wx = eff * dx;
          wy = eff * dy;
          wz = eff * dz;
        }
      }
    }
    put(T.tf, D.pos[0], wx, wy, wz);
    put(T.tf, D.pos[1], -wx, -wy, -wz);
  } else if (D.kind == kTermAngle) {
    const float3 xi = rel_pos(T, R, T.slot_of[D.id[0]], jx, jy, jz);
    const float3 xk = rel_pos(T, R, T.slot_of[D.id[2]], jx, jy, jz);
    const float rji[3] = {xi.x - pj.x, xi.y - pj.y, xi.z - pj.z};
    const float rjk[3] = {xk.x - pj.x, xk.y - pj.y, xk.z - pj.z};
    const float dprod = rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2];
    const float inji = 1.0f / sqrtf(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
    const float injk = 1.0f / sqrtf(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
    float ct = fmaxf(-1.0f, fminf(1.0f, dprod * inji * injk));
    float dxi[3], dxk[3];
    if (ct == 0.0f || ct == -1.0f) {
      // Right or straight angle: the reference draws a random direction
      // (uniform components on [-1, 1]), removes its component along
      // xi - xk and uses the unit result for both outer particles; here the
      // draw is keyed by the angle and the step's time (engine_step counts
      // the step before its forces, so the counter plus one).
      mxg::u32x4 ctr = {{(uint32_t)D.tid, *T.time + 1u, mxg::kStreamAngle, 0u}};
      mxg::u32x4 r = mxg::philox4x32_10(ctr, T.seed_lo, T.seed_hi);
      float v[3], vik[3] = {xi.x - xk.x, xi.y - xk.y, xi.z - xk.z};
      for (int d = 0; d < 3; ++d) v[d] = 2.0f * mxg::u01_closed_open(r.v[d]) - 1.0f;
      const float dv = v[0] * vik[0] + v[1] * vik[1] + v[2] * vik[2];
      for (int d = 0; d < 3; ++d) v[d] = v[d] - dv * vik[d];
      const float il = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int d = 0; d < 3; ++d) dxi[d] = dxk[d] = v[d] * il;
    } else {
      for (int d = 0; d < 3; ++d) {
        dxi[d] = (rjk[d] * injk - ct * rji[d] * inji) * inji;
        dxk[d] = (rji[d] * inji - ct * rjk[d] * injk) * injk;
      }
    }
    const DevPot P = T.tpots[D.pot];
    if (ct < P.a || ct > P.b) ct = fmaxf(P.a, fminf(P.b, ct));
    float eff;
    table_eval_r(P, T.coef, ct, ee, eff);
    float wi[3], wk[3];
    for (int d = 0; d < 3; ++d) {
      wi[d] = eff * dxi[d];
      wk[d] = eff * dxk[d];
    }
    put(T.tf, D.pos[0], -wi[0], -wi[1], -wi[2]);
    put(T.tf, D.pos[2], -wk[0], -wk[1], -wk[2]);
    put(T.tf, D.pos[1], wi[0] + wk[0], wi[1] + wk[1], wi[2] + wk[2]);
  } else {
    // cos(phi) of the planes (i, j, k) and (j, k, l) from the normals
    // m = b1 x b2 and n = b2 x b3 (b1 = xj - xi, b2 = xk - xj,
    // b3 = xl - xk), and its gradient: with g_m = dc/dm and g_n = dc/dn,
    // dc/db1 = b2 x g_m, dc/db2 = g_m x b1 + b3 x g_n, dc/db3 = g_n x b2.
    const float3 xi = rel_pos(T, R, T.slot_of[D.id[0]], jx, jy, jz);
    const float3 xk = rel_pos(T, R, T.slot_of[D.id[2]], jx, jy, jz);
    const float3 xl = rel_pos(T, R, T.slot_of[D.id[3]], jx, jy, jz);
    float c, gi[3], gj[3], gl[3];
    dihedral_geometry(make_float3(xi.x, xi.y, xi.z), make_float3(pj.x, pj.y, pj.z), xk, xl, c, gi, gj, gl);
    const DevPot P = T.tpots[D.pot];
    float cp = fmaxf(-1.0f, fminf(1.0f, c));
    if (cp < P.a || cp > P.b) cp = fmaxf(P.a, fminf(P.b, cp));
    float eff;
    table_eval_r(P, T.coef, cp, ee, eff);
    float wi[3], wj[3], wl[3];
    for (int d = 0; d < 3; ++d) {
      wi[d] = eff * gi[d];
      wj[d] = eff * gj[d];
      wl[d] = eff * gl[d];
    }
    put(T.tf, D.pos[0], -wi[0], -wi[1], -wi[2]);
    put(T.tf, D.pos[1], -wj[0], -wj[1], -wj[2]);
    put(T.tf, D.pos[3], -wl[0], -wl[1], -wl[2]);
    put(T.tf, D.pos[2], wi[0] + wj[0] + wl[0], wi[1] + wj[1] + wl[1], wi[2] + wj[2] + wl[2]);
  }
  T.te[t] = ee;
}

// Energies of the terms in FP64 and a fixed order.  The terms are grouped
// by replica and, within it, by kind, and cut into blocks of at most
// kTermChunk terms of one replica and kind: each block sums its terms
// (k_term_partial), then one block per replica adds its blocks
// (k_term_energy), both in a fixed order.
constexpr int kTermChunk = 4096;

__global__ void __launch_bounds__(256) k_term_partial(const float *__restrict__ te, const int4 *__restrict__ tblk,
                                                      double *__restrict__ part) {
  __shared__ double red[256];
  const int4 B = tblk[blockIdx.x];
  double acc = 0.0;
  for (int t = B.z + threadIdx.x; t < B.w; t += 256) acc += (double)te[t];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_term_energy(const int4 *__restrict__ tblk, const double *__restrict__ part,
                                                     const int *__restrict__ trep_blk, double *__restrict__ out) {
  __shared__ double red[3][256];
  const int r = blockIdx.x;
  double e[3] = {0.0, 0.0, 0.0};
  for (int b = trep_blk[r] + threadIdx.x; b < trep_blk[r + 1]; b += 256) {
    const int kind = tblk[b].y;
    const double v = part[b];
    e[0] += kind == kTermAngle ? v : 0.0;
    e[1] += kind == kTermDihedral ? v : 0.0;
    e[2] += kind == kTermExclusion ? v : 0.0;
  }
  for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = e[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  // exclusion_eval subtracts its energies (exclusion.cpp:255).
  if (threadIdx.x < 3) out[3 * r + threadIdx.x] = threadIdx.x == 2 ? -red[2][0] : red[threadIdx.x][0];
}

// Component k of a float4 by selects (a runtime index into a register
// vector would otherwise send it to scratch memory).
__device__ __forceinline__ float comp(const float4 &v, int k) {
  return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

__device__ __forceinline__ void add_comp(float4 &v, int k, float d, float e) {
  // v[k] = v[k] + d - e, as the reference orders it (flux_eval.hpp:79-80)
  const float n = comp(v, k) + d - e;
  v.x = k == 0 ? n : v.x;
  v.y = k == 1 ? n : v.y;
  v.z = k == 2 ? n : v.z;
  v.w = k == 3 ? n : v.w;
}

// flux_eval_ex (flux_eval.hpp:33-81) for one in-cutoff pair of the
// reference's pair set, this particle's share: every flux of the pair's
// types adds -q - decay/2 s_i to the flux's a particle and q - decay/2 s_j
// to its b particle, q = (1 - r/cutoff)^2 times the kind's rate.  The roles
// follow the types; for equal types the pair's i particle is a (me_i: the
// reference orientation, or the lower slot within a cell).
__device__ __forceinline__ void flux_pair(const DevFlux G *fluxes, const int G *flux_off, int nt, double cutoff,
                                          int ti, int tj, bool me_i, float r, const float4 &me,
                                          const float4 &other, float4 &q) {
  const int f0 = flux_off[ti * nt + tj], f1 = flux_off[ti * nt + tj + 1];
  if (f0 == f1) return;
  float term = (float)(1. - (double)r / cutoff);
  term = term * term;
  const int type_i = me_i ? ti : tj, type_j = me_i ? tj : ti;
  for (int e = f0; e < f1; ++e) {
    const DevFlux X = fluxes[e];
    const bool pi_is_i = type_i == X.ta;  // pi = part_i if its type is a
    const bool pj_is_j = type_j == X.tb;  // pj = part_j if its type is b
    const bool me_pi = pi_is_i ? me_i : !me_i;
    const bool me_pj = pj_is_j ? !me_i : me_i;
    const float ssi = me_pi ? comp(me, X.ia) : comp(other, X.ia);
    const float ssj = me_pj ? comp(me, X.ib) : comp(other, X.ib);
    float rate;
    if (X.kind == 0) {
      rate = X.coef * (ssi - ssj);
    } else if (X.kind == 1) {
      const float qq = X.coef * (ssi - X.target);
      rate = (float)(qq > 0.f) * qq;
    } else {
      const float qq = X.coef * (X.target - ssj) * ssi;
      rate = (float)(qq > 0.f) * qq;
    }
    const float qf = term * rate;
    const float hd = X.decay / 2.f;
    if (me_pi) add_comp(q, X.ia, -qf, hd * ssi);
    if (me_pj) add_comp(q, X.ib, qf, hd * ssj);
  }
}

// MxFluxes_Integrate (Flux.cpp:261-284) for one particle after its advance:
// the reference integrates a cell's particles after that cell's advance, so
// a particle that changed cells keeps its accumulated flux for the next
// step; stayers add dt q (not for constant species) and clear it.  Under
// RK4 nothing is integrated (engine_advance_runge_kutta_4 calls no
// MxFluxes_Integrate).
__device__ __forceinline__ void flux_store(const StepArgs &A, int s, bool integrate, int spec_const, float4 v,
                                           float4 q) {
  if (integrate) {
    auto integ = [&](float s0, float q0, int k) {
      const float konst = (spec_const >> k) & 1 ? 0.f : 1.f;
      return (float)((double)s0 + A.dt_d * (double)q0 * (double)konst);
    };
    v = make_float4(integ(v.x, q.x, 0), integ(v.y, q.y, 1), integ(v.z, q.z, 2), integ(v.w, q.w, 3));
    q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  A.spec_out[s] = v;
  A.specq_out[s] = q;
}

// apply_update_pos_vel (boundary_eval.hpp:30-119): a particle that left a
// wall cell through an active free-slip or velocity wall is mirrored back
// (x -> -x on a low side, 2h - x on a high side, scaled by the restitution
// or the wall's restore factor) and its velocity reflected (free slip:
// that component negated; velocity wall: v -> 2 U - restore v).  Sides in
// the reference's order; returns whether any side acted.
__device__ __forceinline__ bool reflect(const StepArgs &A, const DevReplica &R, int m, float X[3], float V[3]) {
  bool enforced = false;
  for (int side = 0; side < 6; ++side) {
    if (!(m & (1 << side))) continue;
    const int axis = side >> 1;
    const bool low = (side & 1) == 0;
    if (low ? !(X[axis] <= 0.0f) : !((double)X[axis] >= R.hd[axis])) continue;
    const DevWall W = A.walls[side];
    if (W.kind & kBoundaryFreeslip) {
      X[axis] = low ? -X[axis] * 1.0f : (float)(R.hd[axis] - ((double)X[axis] - R.hd[axis]) * 1.0);
      V[axis] *= -1.0f;
      enforced = true;
    } else if (W.kind & kBoundaryVelocity) {
      X[axis] = low ? -X[axis] * W.restore : (float)(2.0 * R.hd[axis] - (double)(X[axis] * W.restore));
      for (int k = 0; k < 3; ++k) V[k] = 2.0f * W.vel[k] - V[k] * W.restore;
      enforced = true;
    }
  }
  return enforced;
}

// space_shuffle_local (space.cpp:333-393), with which the reference's step
// begins: a particle whose cell-local coordinate lies outside [0, h) moves
// to the neighbouring cell (x -= delta h in float, cell index modulo the
// grid).  The load already places particles so, and a step's own move
// leaves a coordinate outside [0, h) only one way: x + h rounding to h (x a
// tiny negative; x - h of an x in [h, 2h) is exact and non-negative).  Such
// a particle starts the next advance from the upper neighbour at 0, as the
// reference's does.  Returns the cell the particle starts in.
__device__ __forceinline__ int shuffle_local(const DevReplica &R, int c, float X[3], float P[3], int &lx, int &ly,
                                             int &lz) {
  const bool ux = X[0] >= R.h[0], uy = X[1] >= R.h[1], uz = X[2] >= R.h[2];
  if (__builtin_expect(!(ux | uy | uz), 1)) return c;
  const int d[3] = {ux ? 1 : 0, uy ? 1 : 0, uz ? 1 : 0};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    X[k] -= d[k] * R.h[k];
    if (P) P[k] -= d[k] * R.h[k];
  }
  lx = wrap(lx + d[0], R.cdim[0]);
  ly = wrap(ly + d[1], R.cdim[1]);
  lz = wrap(lz + d[2], R.cdim[2]);
  return R.cell_base + (lx * R.cdim[1] + ly) * R.cdim[2] + lz;
}

// Advance of one particle (cell_advance_forward_euler,
// engine_advance.cpp:126-239) and the store of its next-step state, the
// migration flag and the per-cell counts.
__device__ __forceinline__ bool advance_store(const StepArgs &A, const DevReplica &R, int s, int c, int lx, int ly,
                                              int lz, const float4 &xs, const float4 &vs, const DevType &T,
                                              int passthrough, float fx, float fy, float fz, const float4 &pfs,
                                              int mypid, float nd) {
  const int pflags = (unsigned)__float_as_int(xs.w) >> 16;
  float X[3] = {xs.x, xs.y, xs.z};
  float V[3] = {vs.x, vs.y, vs.z};
  const float F[3] = {fx, fy, fz};
  // c: the slot's cell; the step moves the particle from cb.
  const int cb = (passthrough || (pflags & kShuffledToFace)) ? c : shuffle_local(R, c, X, nullptr, lx, ly, lz);
  int newc = cb;
  const bool all_frozen = (pflags & kFrozenX) && (pflags & kFrozenY) && (pflags & kFrozenZ);
  if (!passthrough && !all_frozen) {
    const float mask[3] = {(pflags & kFrozenX) ? 0.0f : 1.0f, (pflags & kFrozenY) ? 0.0f : 1.0f,
                           (pflags & kFrozenZ) ? 0.0f : 1.0f};
    int delta[3];
    bool toofast = false;
    if (T.dynamics == kDynNewtonian) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float vn = (float)((double)mask[k] * ((double)V[k] + (double)(A.dt * F[k]) * T.imass));
        if (vn * vn <= R.maxv2[k]) {
          V[k] = vn;
        } else {
          float neg = vn / fabsf(vn);
          V[k] = neg * R.maxv[k];
        }
        X[k] += A.dt * V[k];
        delta[k] = (X[k] >= R.h[k] ? 1 : 0) - (X[k] < 0.0f ? 1 : 0);
        toofast = toofast || (X[k] >= R.h2[k] || X[k] <= -R.h[k]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float d = (float)((double)mask[k] * ((double)(A.dt * F[k]) * T.imass));
        float st;
        if (d * d <= R.maxx2[k]) {
          st = d;
        } else {
          float neg = d / fabsf(d);
          st = neg * R.maxx[k];
        }
        X[k] += st;
        delta[k] = (X[k] >= R.h[k] ? 1 : 0) - (X[k] < 0.0f ? 1 : 0);
        toofast = toofast || (X[k] >= R.h2[k] || X[k] <= -R.h[k]);
      }
    }
    if (toofast) {
      if (atomicCAS(&A.err[0], 0, (int)kErrToofast) == 0) A.err[1] = mypid;
    }
    if ((delta[0] | delta[1] | delta[2]) && (R.wall >> kWallActiveShift)) {
      const int wm = cell_walls(R, lx, ly, lz);
      if (wm && reflect(A, R, wm, X, V)) {
#pragma unroll
        for (int k = 0; k < 3; ++k) delta[k] = (X[k] >= R.h[k] ? 1 : 0) - (X[k] < 0.0f ? 1 : 0);
      }
    }
    if (delta[0] | delta[1] | delta[2]) {
#pragma unroll
      for (int k = 0; k < 3; ++k) X[k] -= delta[k] * R.h[k];
      // delta is -1, 0 or 1: a compare wraps the index (no division).
      const int nx = wrap(lx + delta[0], R.cdim[0]);
      const int ny = wrap(ly + delta[1], R.cdim[1]);
      const int nz = wrap(lz + delta[2], R.cdim[2]);
      newc = R.cell_base + (nx * R.cdim[1] + ny) * R.cdim[2] + nz;
    }
  }
  const float xw = passthrough ? xs.w : __uint_as_float(__float_as_uint(xs.w) & ~((unsigned)kShuffledToFace << 16));
  A.x_out[s] = make_float4(X[0], X[1], X[2], xw);
  A.v_out[s] = make_float4(V[0], V[1], V[2], vs.w);  // w carries the particle id
  A.cell_out[s] = newc != c ? (int)((unsigned)newc | 0x80000000u) : newc;
  // Movers are counted into their new cell here; the caller counts stayers
  // per home cell in LDS and writes the cell's count once.
  if (newc != c) atomicAdd(&A.inc[newc], 1);
  A.f_out[s] = make_float4(fx, fy, fz, nd);
  if (A.with_pf) A.pf_out[s] = pfs;
  return newc == c;
}

// One Runge-Kutta 4 stage of engine_advance_runge_kutta_4
// (engine_advance.cpp:443-698) for slot s, with this stage's force.  Stages
// 0-2 set the position for the next force evaluation, p0 + c dt xk (c = 1/2,
// 1/2, 1), check it (CHECK_TOOFAST, :424-431) and move it with p0 to the
// neighbour cell it entered, as the next engine_force's
// space_shuffle_local does (space.cpp:333-393); stage 3 forms the update
// p0 + dt/6 (xk0 + 2 xk1 + 2 xk2 + xk3) (likewise v) and migrates.  Float
// arithmetic as the reference's Magnum Vector3 (scalars converted to float:
// 0.5 dt, dt / 6., imass); no clamps, no frozen mask.  Returns whether the
// particle stayed in its cell.
__device__ __forceinline__ bool rk_store(const StepArgs &A, const DevReplica &R, int stage, int s, int c, int lx,
                                         int ly, int lz, const float4 &xs, const float4 &vs, const DevType &T,
                                         int passthrough, float fx, float fy, float fz, const float4 &pfs,
                                         int mypid, float nd) {
  const float dt = A.dt;
  const float half = (float)(0.5 * (double)dt);
  const float sixth = (float)((double)dt / 6.);
  const float imf = (float)T.imass;
  const bool newton = T.dynamics == kDynNewtonian;
  const float F[3] = {fx, fy, fz};
  float X[3] = {xs.x, xs.y, xs.z};
  const float V[3] = {vs.x, vs.y, vs.z};  // v0: velocities change only at the last stage
  float P0[3], AX[3], AV[3], VK[3], Xn[3], Vn[3] = {V[0], V[1], V[2]};
  if (stage == 0) {
    P0[0] = X[0];
    P0[1] = X[1];
    P0[2] = X[2];
  } else {
    const float4 p0 = float4(A.rk_p0[s]);
    P0[0] = p0.x;
    P0[1] = p0.y;
    P0[2] = p0.z;
  }
  // c: the slot's cell; the stage moves the particle (and p0) from cb.  The
  // first stage follows the previous step's update, which the reference's
  // next engine_force shuffles; each later stage follows this kernel's own
  // move at the end of the stage before, which is that force's shuffle.
  const bool at_face = (((unsigned)__float_as_int(xs.w) >> 16) & kShuffledToFace) != 0;
  const int cb = (passthrough || stage != 0 || at_face) ? c : shuffle_local(R, c, X, P0, lx, ly, lz);
  const float4 axv = stage ? float4(A.rk_ax[s]) : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 avv = stage ? float4(A.rk_av[s]) : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 vkv = stage ? float4(A.rk_vk[s]) : make_float4(0.f, 0.f, 0.f, 0.f);
  const float ax0[3] = {axv.x, axv.y, axv.z}, av0[3] = {avv.x, avv.y, avv.z}, vkp[3] = {vkv.x, vkv.y, vkv.z};
  int newc = cb;
  if (!passthrough) {
    int delta[3];
    bool toofast = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float vk = newton ? (stage == 3 ? imf * F[k] : F[k] * imf) : 0.0f;
      float xk;
      if (newton)
        xk = stage == 0 ? V[k] : V[k] + (stage == 3 ? dt : half) * vkp[k];
      else
        xk = stage == 3 ? imf * F[k] : F[k] * imf;
      if (stage < 3) {
        AX[k] = stage == 0 ? xk : ax0[k] + 2.0f * xk;
        AV[k] = stage == 0 ? vk : av0[k] + 2.0f * vk;
        VK[k] = vk;
        Xn[k] = P0[k] + (stage == 2 ? dt : half) * xk;
      } else {
        AX[k] = AV[k] = VK[k] = 0.0f;
        if (newton) Vn[k] = V[k] + sixth * (av0[k] + vk);
        Xn[k] = P0[k] + sixth * (ax0[k] + xk);
      }
      delta[k] = (Xn[k] >= R.h[k] ? 1 : 0) - (Xn[k] < 0.0f ? 1 : 0);
      toofast = toofast || (Xn[k] >= R.h2[k] || Xn[k] <= -R.h[k]);
    }
    if (toofast) {
      if (atomicCAS(&A.err[0], 0, (int)kErrToofast) == 0) A.err[1] = mypid;
    }
    if (delta[0] | delta[1] | delta[2]) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        Xn[k] -= delta[k] * R.h[k];
        P0[k] -= delta[k] * R.h[k];
      }
      // delta is -1, 0 or 1: a compare wraps the index (no division).
      const int nx = wrap(lx + delta[0], R.cdim[0]);
      const int ny = wrap(ly + delta[1], R.cdim[1]);
      const int nz = wrap(lz + delta[2], R.cdim[2]);
      newc = R.cell_base + (nx * R.cdim[1] + ny) * R.cdim[2] + nz;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Xn[k] = X[k];
      AX[k] = ax0[k];
      AV[k] = av0[k];
      VK[k] = vkp[k];
    }
  }
  const float xw = passthrough ? xs.w : __uint_as_float(__float_as_uint(xs.w) & ~((unsigned)kShuffledToFace << 16));
  A.x_out[s] = make_float4(Xn[0], Xn[1], Xn[2], xw);
  A.v_out[s] = make_float4(Vn[0], Vn[1], Vn[2], vs.w);
  A.cell_out[s] = newc != c ? (int)((unsigned)newc | 0x80000000u) : newc;
  if (newc != c) atomicAdd(&A.inc[newc], 1);
  A.f_out[s] = make_float4(fx, fy, fz, nd);
  if (A.with_pf) A.pf_out[s] = pfs;
  if (stage < 3) {
    A.rk_p0_out[s] = make_float4(P0[0], P0[1], P0[2], 0.f);
    A.rk_ax_out[s] = make_float4(AX[0], AX[1], AX[2], 0.f);
    A.rk_av_out[s] = make_float4(AV[0], AV[1], AV[2], 0.f);
    A.rk_vk_out[s] = make_float4(VK[0], VK[1], VK[2], 0.f);
  }
  return newc == c;
}

// Bricks of cells, the work unit of the force kernel: a block stages the
// brick and a halo of span cells around it (positions, velocities, ids) in
// LDS with coalesced row loads, then each thread owns one particle of the
// brick.  A brick whose halo holds more than pmax particles is computed from
// global memory instead.
struct BrickArgs {
  const int4 *__restrict__ bricks;  // (replica, x0, y0, z0)
  int nbricks;
  int bdim[3];  // brick size in cells
  int pmax;     // LDS particle capacity
  int cap;      // neighbour-list entries per thread
  int drain;    // entries a full list keeps when the wave drains (clamped to [0, cap - chunk])
  int hmax;     // halo-cell capacity
  int hdmax;    // halo cells per axis (capacity)
  int noff;     // neighbour offsets (capacity)
  int *ovf_list;   // bricks whose halo overflowed the LDS capacity
  int *ovf_count;
  int rk_stage;    // -1: forward Euler step; 0-3: Runge-Kutta 4 stage
};

constexpr int kBT = 256;        // threads per brick block
// Candidates loaded together in the scan: two in modes with DPD potentials
// (three per chunk costs the DPD fluid about 1%), three otherwise (the dense
// LJ fluid gains about 4%).  Neighbour lists hold at least kChunkMax entries.
template <int MODE>
constexpr int scan_chunk() { return (MODE & kModeDPD) ? 2 : 3; }
constexpr int kChunkMax = 3;
// Blocks of the global-memory fallback: bricks whose halo overflows LDS
// (dense clusters, small replicas) are spread over this many workgroups.
constexpr int kGlobalBlocks = 1024;
constexpr int kHaloAxis = 64;   // max halo cells per axis
constexpr int kSpanMax = 2;     // max cell span
constexpr int kOffMax = 125;    // (2 * kSpanMax + 1)^3
constexpr int kMaxTypes = 16;   // particle types per engine
constexpr int kMaxPots = 32;    // pair potentials per engine
// LDS bytes of the per-type-pair tables: the potentials in use, the type-pair
// matrix and the radii (sized to what the engine holds, so small systems
// keep LDS for particles).
__host__ __device__ inline size_t pair_tab_bytes(int npots, int ntypes) {
  const size_t b = (size_t)npots * sizeof(DevPot) + (size_t)ntypes * ntypes * 4 + (size_t)ntypes * 4;
  return (b + 15) / 16 * 16;
}

// Copies the per-type-pair tables into LDS at p (pair_tab_bytes).
__device__ __forceinline__ PairTabs stage_pair_tabs(const StepArgs &A, unsigned char *p, int npots) {
  const int nt = A.ntypes;
  DevPot *pots = reinterpret_cast<DevPot *>(p);
  int *potmat = reinterpret_cast<int *>(pots + npots);
  float *radius = reinterpret_cast<float *>(potmat + nt * nt);
  for (int i = threadIdx.x; i < npots; i += blockDim.x) pots[i] = A.pots[i];
  for (int i = threadIdx.x; i < nt * nt; i += blockDim.x) potmat[i] = A.potmat[i];
  for (int i = threadIdx.x; i < nt; i += blockDim.x) radius[i] = A.types[i].radius;
  return PairTabs{pots, potmat, radius};
}

__host__ __device__ inline size_t brick_lds_bytes(int pmax, int hmax, int hdmax, int noff, int cap, bool dpd,
                                                  bool flux, size_t tabs) {
  size_t b = 0;
  b += (size_t)pmax * 16;              // positions (w: type and flags)
  if (dpd) b += (size_t)pmax * 16;     // velocities (w: particle id)
  if (flux) b += (size_t)pmax * 16;    // species values
  b += (size_t)noff * 16;              // offsets: halo-index delta, list geometry, flags
  b += (size_t)(hmax + 1) * 4 * 5;     // halo starts, global starts, home starts, home stayers, near bits
  b += (size_t)3 * hdmax * (2 * kSpanMax + 1) * 4;  // shifts
  b += (size_t)cap * kBT * 4;          // neighbour lists
  b += tabs;                           // potentials, type pairs, radii
  return b;
}

__device__ __forceinline__ int upper_index(const int *a, int n, int f) {
  // largest i in [0, n) with a[i] <= f (a ascending, a[0] = 0)
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a[mid] <= f)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ int LdsNbr::pid(int q) const {
  if (staged_v) return __float_as_int(vv[q].w);
  const int h = upper_index(hstart, nhalo + 1, q);
  return gpid[hgb[h] + (q - hstart[h])];
}

// Workgroups go to the eight XCDs round-robin (workgroup g to XCD g % 8),
// and each XCD has its own L2.  Giving XCD k a contiguous run of the brick
// list (neighbouring bricks share most of their halo) keeps halo re-reads in
// that XCD's L2.  A bijection on [0, n); the tail past a multiple of 8 maps
// to itself.
__device__ __forceinline__ int xcd_brick(int g, int n) {
  constexpr int kXcd = 8;
  const int full = n - n % kXcd;
  if (g >= full) return g;
  return (g % kXcd) * (full / kXcd) + g / kXcd;
}

// Three waves per SIMD: the register budget (168 VGPRs) holds without
// spills, and bricks whose LDS footprint allows three blocks per CU keep
// more candidate loads in flight (up to 30% faster than two waves).  The
// bonded modes also come with four waves (128 VGPRs), which the autotuner
// may pick: their term gathers gain more from the extra wave than the pair
// loops lose to the tighter budget.
template <int MODE, int WAVES>
__global__ __launch_bounds__(kBT, WAVES) void k_brick(const StepArgs *__restrict__ Ap, BrickArgs B) {
  const StepArgs &A = *Ap;
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr bool kDPD = (MODE & kModeDPD) != 0;
  constexpr bool kFlux = (MODE & kModeFlux) != 0;
  constexpr int kChunk = scan_chunk<MODE>();
  float4 *hx = reinterpret_cast<float4 *>(smem);
  float4 *hv = hx + B.pmax;
  float4 *hsp = hv + (kDPD ? B.pmax : 0);  // species values (flux mode)
  // Per offset: halo-index delta, the low 12 bits of its list entries
  // (window << 11 | (side + 1) << 9 | (ox + 2) << 6 | (oy + 2) << 3 |
  // (oz + 2)) and DevOffset::flags.
  int4 *hoff = reinterpret_cast<int4 *>(hsp + (kFlux ? B.pmax : 0));
  int *hstart = reinterpret_cast<int *>(hoff + B.noff);
  int *hgb = hstart + (B.hmax + 1);
  int *homestart = hgb + (B.hmax + 1);
  int *home_stay = homestart + (B.hmax + 1);  // stayers per home cell
  int *hmask = home_stay + (B.hmax + 1);      // near-bound bits per halo cell
  float *hs = reinterpret_cast<float *>(hmask + (B.hmax + 1));
  unsigned *nlist = reinterpret_cast<unsigned *>(hs + 3 * B.hdmax * (2 * kSpanMax + 1));
  __shared__ int scan_lds[kBT / 64];
  __shared__ double red[kNRed][kBT / 64];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int passthrough = A.err[2];
  const unsigned t1 = *A.time + 1u;
  const PairTabs tb = stage_pair_tabs(A, reinterpret_cast<unsigned char *>(nlist + B.cap * kBT), A.npots);
  LoopConsts C = loop_consts(A);
  (void)kOffMax;
  const DevFlux G *const fluxes = kFlux ? A.fluxes : nullptr;
  const int G *const flux_off = kFlux ? A.flux_off : nullptr;
  const double flux_cut = kFlux ? A.cutoff_d : 0.0;

  for (int g = blockIdx.x; g < B.nbricks; g += gridDim.x) {
    const int bi = xcd_brick(g, B.nbricks);
    const int4 br = B.bricks[bi];
    const DevReplica R = A.rep[br.x];  // by value: the fields stay in registers
    C.seed_lo = R.seed_lo;
    C.seed_hi = R.seed_hi;
    C.id_base = R.id_base;
    const uint32_t tk = step_key(R.seed_lo, t1);
    const int sp0 = R.span[0], sp1 = R.span[1], sp2 = R.span[2];
    const int nb0 = min(B.bdim[0], R.cdim[0] - br.y);
    const int nb1 = min(B.bdim[1], R.cdim[1] - br.z);
    const int nb2 = min(B.bdim[2], R.cdim[2] - br.w);
    const int hd0 = nb0 + 2 * sp0, hd1 = nb1 + 2 * sp1, hd2 = nb2 + 2 * sp2;
    const int nhalo = hd0 * hd1 * hd2;
    const int nhome = nb0 * nb1 * nb2;

    // 1. Halo cell ranges and their exclusive scan (kScanItems per thread).
    {
      int cnt[kScanItems];
      int local = 0;
#pragma unroll
      for (int i = 0; i < kScanItems; ++i) {
        const int h = tid * kScanItems + i;
        cnt[i] = 0;
        if (h < nhalo) {
          const int hi = h / (hd1 * hd2), hj = (h / hd2) % hd1, hk = h % hd2;
          const int gx = nbr_coord(br.y - sp0 + hi, R.cdim[0], R.wall & 1);
          const int gy = nbr_coord(br.z - sp1 + hj, R.cdim[1], R.wall & 2);
          const int gz = nbr_coord(br.w - sp2 + hk, R.cdim[2], R.wall & 4);
          if ((gx | gy | gz) >= 0) {
            const int c = R.cell_base + (gx * R.cdim[1] + gy) * R.cdim[2] + gz;
            const int gb = A.cstart[c];
            cnt[i] = A.cstart[c + 1] - gb;
            hgb[h] = gb;
          } else {
            hgb[h] = 0;  // beyond a wall: an empty cell
          }
          hmask[h] = 0;
        }
        local += cnt[i];
      }
      int total;
      int off = block_exclusive_scan(local, scan_lds, total);
#pragma unroll
      for (int i = 0; i < kScanItems; ++i) {
        const int h = tid * kScanItems + i;
        if (h < nhalo) hstart[h] = off;
        off += cnt[i];
      }
      if (tid == 0) hstart[nhalo] = total;
    }
    // Shift table of the brick and the offset list.
    for (int i = tid; i < hd0 * (2 * sp0 + 1); i += kBT) {
      const int hl = i / (2 * sp0 + 1), d = i % (2 * sp0 + 1);
      hs[(0 * B.hdmax + hl) * (2 * kSpanMax + 1) + d] =
          A.shifts[R.shift_base[0] + wrap(br.y - sp0 + hl, R.cdim[0]) * (2 * sp0 + 1) + d];
    }
    for (int i = tid; i < hd1 * (2 * sp1 + 1); i += kBT) {
      const int hl = i / (2 * sp1 + 1), d = i % (2 * sp1 + 1);
      hs[(1 * B.hdmax + hl) * (2 * kSpanMax + 1) + d] =
          A.shifts[R.shift_base[1] + wrap(br.z - sp1 + hl, R.cdim[1]) * (2 * sp1 + 1) + d];
    }
    for (int i = tid; i < hd2 * (2 * sp2 + 1); i += kBT) {
      const int hl = i / (2 * sp2 + 1), d = i % (2 * sp2 + 1);
      hs[(2 * B.hdmax + hl) * (2 * kSpanMax + 1) + d] =
          A.shifts[R.shift_base[2] + wrap(br.w - sp2 + hl, R.cdim[2]) * (2 * sp2 + 1) + d];
    }
    for (int i = tid; i < R.n_off; i += kBT) {
      const DevOffset d = A.offs[R.off_base + i];
      const int geo = ((d.flags & kOffWindow) ? 0x800 : 0) | ((d.side + 1) << 9) | ((d.ox + 2) << 6) |
                      ((d.oy + 2) << 3) | (d.oz + 2);
      hoff[i] = make_int4((d.ox * hd1 + d.oy) * hd2 + d.oz, geo, d.flags, 0);
    }
    __syncthreads();

    // 2. Home cells: exclusive scan of their particle counts.
    {
      int cnt[kScanItems];
      int local = 0;
#pragma unroll
      for (int i = 0; i < kScanItems; ++i) {
        const int hc = tid * kScanItems + i;
        cnt[i] = 0;
        if (hc < nhome) {
          const int a = hc / (nb1 * nb2), b = (hc / nb2) % nb1, cc = hc % nb2;
          const int h = ((a + sp0) * hd1 + (b + sp1)) * hd2 + (cc + sp2);
          cnt[i] = hstart[h + 1] - hstart[h];
          home_stay[hc] = 0;
        }
        local += cnt[i];
      }
      int total;
      int off = block_exclusive_scan(local, scan_lds, total);
#pragma unroll
      for (int i = 0; i < kScanItems; ++i) {
        const int hc = tid * kScanItems + i;
        if (hc < nhome) homestart[hc] = off;
        off += cnt[i];
      }
      if (tid == 0) homestart[nhome] = total;
    }
    const int P = hstart[nhalo];
    if (P > B.pmax) {
      // The halo does not fit: k_brick_global computes this brick.
      if (tid == 0) B.ovf_list[atomicAdd(B.ovf_count, 1)] = bi;
      __syncthreads();
      continue;
    }
    // 3. Stage the halo particles (consecutive indices within a row of cells
    // are consecutive slots, so the loads coalesce).
    for (int f = tid; f < P; f += kBT) {
      const int h = upper_index(hstart, nhalo + 1, f);
      const int gs = hgb[h] + (f - hstart[h]);
      const float4 xf = A.x[gs];
      hx[f] = xf;
      if (kDPD) hv[f] = A.v[gs];
      if (kFlux) hsp[f] = A.spec[gs];
      if (R.drop_check) {
        const int m = near_bits(R, xf);
        if (m) atomicOr(&hmask[h], m);
      }
    }
    __syncthreads();
    const int nhp = homestart[nhome];

    double acc[kNRed];
#pragma unroll
    for (int k = 0; k < kNRed; ++k) acc[k] = 0.0;

    for (int f = tid; f < nhp; f += kBT) {
      const int hc = upper_index(homestart, nhome + 1, f);
      const int kk = f - homestart[hc];
      const int a = hc / (nb1 * nb2), b = (hc / nb2) % nb1, cc = hc % nb2;
      const int ha = a + sp0, hb = b + sp1, hcz = cc + sp2;
      const int h = (ha * hd1 + hb) * hd2 + hcz;
      const int s = hgb[h] + kk;       // global slot
      const int p = hstart[h] + kk;    // LDS index
      const int lx = br.y + a, ly = br.z + b, lz = br.w + cc;
      const int c = R.cell_base + (lx * R.cdim[1] + ly) * R.cdim[2] + lz;
      const float4 xs = hx[p];
      const float4 vs = kDPD ? float4(hv[p]) : float4(A.v[s]);
      const float4 sme = kFlux ? float4(hsp[p]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      float4 sq = kFlux ? float4(A.specq[s]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const int mypid = __float_as_int(vs.w);
      const int ti = __float_as_int(xs.w) & 0xffff;
      const DevType T = A.types[ti];
      {
        float e = 0.5 * T.mass * (vs.x * vs.x + vs.y * vs.y + vs.z * vs.z);
        acc[2] += (double)e;
      }
      float fx = 0.0f, fy = 0.0f, fz = 0.0f;
      // space_prepare sets f[3] to the self density W(0, cutoff).
      float nd = A.nd_norm;
      float4 pfs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (!passthrough) {
        if (A.random) single_body(A, C, s, mypid, ti, T, t1, B.rk_stage <= 0, vs, fx, fy, fz, pfs);
        if (R.wall >> kWallActiveShift) {
          const int wm = cell_walls(R, lx, ly, lz);
          if (wm) wall_forces(A, R, wm, ti, T.radius, xs, vs, C.dpd_delta, fx, fy, fz, acc[0]);
        }
        {
          const LdsNbr nb{hv, kDPD, hstart, hgb, nhalo, A.pid};
          const unsigned dmax = (unsigned)R.dmaxdist;
          // Shift of the pair for offset (ox, oy, oz) in the reference
          // orientation, from this particle's cell's rows of the shift table;
          // the reverse shift (side -1: the neighbour is the reference's i
          // cell) is the exact negation (checked when the tables are built),
          // and side 0 (the same cell) reads the table's exact +0.
          constexpr int W = 2 * kSpanMax + 1;
          const float *hsx = hs + (0 * B.hdmax + ha) * W + sp0, *hsy = hs + (1 * B.hdmax + hb) * W + sp1,
                      *hsz = hs + (2 * B.hdmax + hcz) * W + sp2;
          auto shift_of = [&](int ox, int oy, int oz, int side, float &sx, float &sy, float &sz) {
            const float x = hsx[ox], y = hsy[oy], z = hsz[oz];
            sx = side >= 0 ? x : -x;
            sy = side >= 0 ? y : -y;
            sz = side >= 0 ? z : -z;
          };
          // Each thread's neighbour list: entry k of thread t at
          // nlist[k * kBT + t], nn entries held.
          int nn = 0;
          // Evaluates the first k entries of every list, k the wave's
          // longest list less `target`, so that no list keeps more than
          // `target`, then moves what is left to the lists' fronts.  A wave
          // whose fullest list triggers the drain stops while most lists
          // still have entries, so most lanes work in each of the k passes;
          // each list is evaluated in the order it was appended, so each
          // particle's pair order, and its force bits, do not depend on when
          // the wave drains.
          auto process = [&](int target) {
            int longest = 0;  // the wave's longest list (nn <= cap <= 64), bit by bit
            for (int b = 64; b >= 1; b >>= 1)
              if (__any(nn >= longest + b)) longest += b;
            const int k = max(longest - target, 0);
            const int n = min(nn, k);
            for (int e = 0; e < n; ++e) {
              // Entry: LDS index << 12 | window << 11 | (side + 1) << 9 |
              // (ox + 2) << 6 | (oy + 2) << 3 | (oz + 2).
              const unsigned ent = nlist[e * kBT + tid];
              const int q = (int)(ent >> 12);
              const int side = (int)((ent >> 9) & 3u) - 1;
              const float4 xq = hx[q];
              // (a - shift) - b in the reference orientation; a self entry's
              // offset is (0, 0, 0), whose table shift is exactly +0, and
              // x - 0 is exact, so self pairs (x_i - x_j) take the same form
              // and no entry branches on its side.
              const int ox = (int)((ent >> 6) & 7u) - 2, oy = (int)((ent >> 3) & 7u) - 2, oz = (int)(ent & 7u) - 2;
              float sx, sy, sz;
              shift_of(ox, oy, oz, side, sx, sy, sz);
              // runner_dopair's sort-key window (runner_dopair.cpp:170,182) along the
              // pair's forward direction f = side * sign(ox, oy, oz), n = f / |f| as the
              // offset table holds it (components 0 or +-1/sqrt(|f|^2)); only on
              // offsets where it can reject an in-cutoff pair (kOffWindow).
              if (side != 0 && (ent & 0x800u)) {
                float nx, ny, nz;
                offset_dir(ox, oy, oz, side, nx, ny, nz);
                const unsigned dnshift = (unsigned)(int)(R.dscale * sqrtf(sx * sx + sy * sy + sz * sz));
                const unsigned kp = sort_key(R.bias, R.dscale, xs.x, xs.y, xs.z, nx, ny, nz);
                const unsigned kq = sort_key(R.bias, R.dscale, xq.x, xq.y, xq.z, nx, ny, nz);
                if (!(side > 0 ? (kq + dnshift - kp < dmax) : (kp + dnshift - kq < dmax))) continue;
              }
              const bool fwd = side >= 0;
              const float ax = fwd ? xs.x : xq.x, ay = fwd ? xs.y : xq.y, az = fwd ? xs.z : xq.z;
              const float bx = fwd ? xq.x : xs.x, by = fwd ? xq.y : xs.y, bz = fwd ? xq.z : xs.z;
              const float dx = (ax - sx) - bx, dy = (ay - sy) - by, dz = (az - sz) - bz;
              const float r2 = dx * dx + dy * dy + dz * dz;
              const float r = pair_root<MODE>(r2);
              const int tj = __float_as_int(xq.w) & 0xffff;
              if (pair_force<MODE>(C, tb, nb, ti, tj, side >= 0, mypid, q, side == 0, dx, dy, dz, r2, r, vs, fx, fy,
                                   fz, acc, t1, tk))
                nd += w_cubic(r, C.cutoff, C.nd_norm);
              if (kFlux)
                flux_pair(fluxes, flux_off, C.ntypes, flux_cut, ti, tj, side == 0 ? p < q : side > 0, r, sme,
                          hsp[q], sq);
            }
            for (int j = k; j < nn; ++j) nlist[(j - k) * kBT + tid] = nlist[j * kBT + tid];
            nn -= n;
          };
          // Gather in-cutoff neighbours into the per-thread list, process
          // it when it is full or the scan is done (one call site, so the
          // pair code is instantiated once).  The control flow is the
          // wave's: every lane visits the offsets in the same order, a chunk
          // loop runs while any lane has candidates left in its cell (a lane
          // that is done keeps its position and appends nothing), and the
          // wave drains its lists when any lane's list might not hold the
          // next chunk.  Draining early does not change a lane's list order,
          // so each particle's pair order is the same as a per-lane loop's.
          int o = 0, q = -1;
          const int hm = R.drop_check ? hmask[h] : 0;  // this cell's near-bound bits
          int4 offn = hoff[0];  // descriptor of offset o, read one offset ahead
          for (;;) {
            bool full = false;
            while (o < R.n_off) {
              // The offset is the same for the whole wave: its descriptor
              // goes to scalar registers.  The next one is read now, off
              // the path of this offset's cell bounds.
              const int4 off = make_int4(__builtin_amdgcn_readfirstlane(offn.x), __builtin_amdgcn_readfirstlane(offn.y),
                                         __builtin_amdgcn_readfirstlane(offn.z), 0);
              offn = hoff[min(o + 1, R.n_off - 1)];
              const int h2 = h + off.x;
              const unsigned geo = (unsigned)off.y;
              const int oside = (int)((geo >> 9) & 3u) - 1;
              const int qe = hstart[h2 + 1];
              if (q < 0) {
                const int qb = hstart[h2];
                q = qb;
                // A cell pair that runner_dopair drops whole: only when both
                // cells hold particles near the bounds the offset's keys
                // favour (rare), the exact key test; a dropped pair leaves
                // this lane nothing to scan in the cell.
                if ((off.z & kOffDrop) && (hm & off.z & 0x3f) == (off.z & 0x3f)) {
                  const int need = (off.z >> 8) & 0x3f;
                  if ((hmask[h2] & need) == need) {
                    const int ox = (int)((geo >> 6) & 7u) - 2, oy = (int)((geo >> 3) & 7u) - 2,
                              oz = (int)(geo & 7u) - 2;
                    float sx, sy, sz, nx, ny, nz;
                    shift_of(ox, oy, oz, oside, sx, sy, sz);
                    offset_dir(ox, oy, oz, oside, nx, ny, nz);
                    const int hb = hstart[h], he = hstart[h + 1];
                    if (oside > 0 ? pair_dropped(R, nx, ny, nz, sx, sy, sz, hx, hb, he, qb, qe)
                                  : pair_dropped(R, nx, ny, nz, sx, sy, sz, hx, qb, qe, hb, he))
                      q = qe;
                  }
                }
              }
              // Candidates go in chunks of kChunk: their LDS loads are issued
              // together, and the list is flushed before a chunk that might
              // not fit.  A chunk's tail past the cell reads the next staged
              // particles (or, past the last, the next LDS arrays of this
              // block: a lane that is done stays at most kChunk - 1 entries
              // past its cell, so reads end 2 kChunk - 2 entries beyond hx);
              // those entries are never kept (q + u < qe), and unclamped
              // indices let the chunk share one address.
              if (oside == 0) {
                while (__any(q < qe)) {
                  if (__any(nn > B.cap - kChunk)) {
                    full = true;
                    break;
                  }
                  float4 xq[kChunk];
#pragma unroll
                  for (int u = 0; u < kChunk; ++u) xq[u] = hx[q + u];
#pragma unroll
                  for (int u = 0; u < kChunk; ++u) {
                    const float dx = xs.x - xq[u].x, dy = xs.y - xq[u].y, dz = xs.z - xq[u].z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    // Written unconditionally (the flush check above leaves room
                    // for a whole chunk), kept by advancing nn: no divergent
                    // branch, so no exec-mask bookkeeping per candidate.
                    nlist[nn * kBT + tid] = ((unsigned)(q + u) << 12) | geo;
                    nn += (q + u < qe && q + u != p && r2 <= C.cut2_self) ? 1 : 0;
                  }
                  q += q < qe ? kChunk : 0;
                }
              } else {
                float sx, sy, sz;
                shift_of((int)((geo >> 6) & 7u) - 2, (int)((geo >> 3) & 7u) - 2, (int)(geo & 7u) - 2, oside, sx, sy,
                         sz);
                // (a - shift) - b with (a, b) = (i, j) in the reference
                // orientation.  This cell is i (side +1): a - shift is the
                // reference's pix, the same for every candidate.  The
                // neighbour is i (side -1): each candidate is shifted.
                // In-cutoff candidates go to the list; the sort-key window
                // of runner_dopair is applied when the list is processed.
                if (oside > 0) {
                  const float pix = xs.x - sx, piy = xs.y - sy, piz = xs.z - sz;
                  while (__any(q < qe)) {
                    if (__any(nn > B.cap - kChunk)) {
                      full = true;
                      break;
                    }
                    float4 xq[kChunk];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) xq[u] = hx[q + u];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) {
                      const float dx = pix - xq[u].x, dy = piy - xq[u].y, dz = piz - xq[u].z;
                      const float r2 = dx * dx + dy * dy + dz * dz;
                      nlist[nn * kBT + tid] = ((unsigned)(q + u) << 12) | geo;
                      nn += (q + u < qe && r2 <= C.cut2_pair) ? 1 : 0;
                    }
                    q += q < qe ? kChunk : 0;
                  }
                } else {
                  while (__any(q < qe)) {
                    if (__any(nn > B.cap - kChunk)) {
                      full = true;
                      break;
                    }
                    float4 xq[kChunk];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) xq[u] = hx[q + u];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) {
                      const float dx = (xq[u].x - sx) - xs.x, dy = (xq[u].y - sy) - xs.y, dz = (xq[u].z - sz) - xs.z;
                      const float r2 = dx * dx + dy * dy + dz * dz;
                      nlist[nn * kBT + tid] = ((unsigned)(q + u) << 12) | geo;
                      nn += (q + u < qe && r2 <= C.cut2_pair) ? 1 : 0;
                    }
                    q += q < qe ? kChunk : 0;
                  }
                }
              }
              if (full) break;
              ++o;
              q = -1;
            }
            // A full list drains to B.drain entries (at most cap - kChunk, so
            // that the next chunk fits); at the end of the scan every list
            // drains.
            process(full ? max(0, min(B.drain, B.cap - kChunk)) : 0);
            if (!full) break;
            offn = hoff[o];  // resume at the same offset
          }
        }
        if (MODE & kModeBonds) {
          bond_forces<MODE>(A, R, s, mypid, fx, fy, fz, acc);
          term_forces(A, mypid, fx, fy, fz);
        }
      }
      const bool stayed =
          B.rk_stage < 0 ? advance_store(A, R, s, c, lx, ly, lz, xs, vs, T, passthrough, fx, fy, fz, pfs, mypid, nd)
                         : rk_store(A, R, B.rk_stage, s, c, lx, ly, lz, xs, vs, T, passthrough, fx, fy, fz, pfs,
                                    mypid, nd);
      if (stayed) atomicAdd(&home_stay[hc], 1);
      if (kFlux) flux_store(A, s, stayed && B.rk_stage < 0, T.spec_const, sme, sq);
    }

    // Deterministic block reduction of the FP64 partials for this brick.
#pragma unroll
    for (int k = 0; k < kNRed; ++k) {
      if (!(MODE & kModeVirial) && k >= 3) break;
      double w = wave_sum(acc[k]);
      if (lane == 0) red[k][wave] = w;
    }
    __syncthreads();
    // kNRed partials per brick with the virial, 3 without (finalize_replica
    // reads the same stride).
    constexpr int kPart = (MODE & kModeVirial) ? kNRed : 3;
    if (tid < kPart) {
      double sum = 0.0;
      for (int w = 0; w < kBT / 64; ++w) sum += red[tid][w];
      A.partials[(size_t)bi * kPart + tid] = sum;
    }
    // Stayer counts of the home cells (this brick is their only writer).
    for (int hc = tid; hc < nhome; hc += kBT) {
      const int a = hc / (nb1 * nb2), b = (hc / nb2) % nb1, cc = hc % nb2;
      A.cnt[R.cell_base + ((br.y + a) * R.cdim[1] + (br.z + b)) * R.cdim[2] + (br.w + cc)] = home_stay[hc];
    }
    __syncthreads();
  }
}

// Per-particle count of in-cutoff neighbours (pairs the reference would
// evaluate), for bit-exact comparison with the CPU loops.
__global__ __launch_bounds__(kTile) void k_ncount(StepArgs A) {
  for (int t = blockIdx.x; t < A.ntiles; t += gridDim.x) {
    const int4 tile = A.tiles[t];
    const int s = tile.x + threadIdx.x;
    if (s >= tile.y) continue;
    const DevReplica R = A.rep[tile.z];
    const float4 xs = A.x[s];
    int lx, ly, lz;
    cell_coords(A.cell[s] - R.cell_base, R, lx, ly, lz);
    int n = 0;
    const LoopConsts C = loop_consts(A);
    for_each_neighbour(A, C, R, s, lx, ly, lz, xs,
                       [&](int, bool, bool, float, float, float, float, const float4 &) { ++n; });
    A.ncount[A.pid[s]] = n;
  }
}

struct RebuildArgs {
  const int *__restrict__ cstart;      // old cell offsets
  const int *__restrict__ cstart_new;  // new cell offsets
  const int *__restrict__ cell_out;    // new cell of each old slot, bit 31 set if it moved
  int *__restrict__ inc;               // particles entering each cell
  int *__restrict__ cursor;            // per-cell cursor for entering particles
  int nslots;
  int ncells;
  const float4 *__restrict__ x_in;
  const float4 *__restrict__ v_in;
  const float4 *__restrict__ pf_in;
  const int *__restrict__ pid_in;
  float4 *__restrict__ x;
  float4 *__restrict__ v;
  float4 *__restrict__ pf;
  int *__restrict__ pid;
  int *__restrict__ cell;
  int *__restrict__ slot_of;
  int with_slots;  // bonds or terms look partners up by id: keep slot_of current
  int with_pf;
  // Runge-Kutta 4 state, carried along between stages (buffer 1 -> 0)
  const float4 *__restrict__ rk_in[4];
  float4 *__restrict__ rk[4];
  int with_rk;
  // species values and flux accumulators (buffer 1 -> 0)
  const float4 *__restrict__ spec_in[2];
  float4 *__restrict__ spec[2];
  int with_spec;
};

// Exclusive scan of the per-cell counts (first pass); the counts are reset
// to zero for the next step once read.
__global__ __launch_bounds__(kScanBlock) void k_scan_block(int *__restrict__ cnt, const int *__restrict__ inc,
                                                          int *__restrict__ out, int *__restrict__ bsum, int n) {
  __shared__ int lds[kScanBlock / 64];
  const int base = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
  int vals[kScanItems];
  int local = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    vals[i] = (base + i < n) ? cnt[base + i] + inc[base + i] : 0;  // stayers + arrivals
    local += vals[i];
  }
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) cnt[base + i] = 0;
  int total;
  int off = block_exclusive_scan(local, lds, total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = off;
    off += vals[i];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Second pass: each block adds the total of the chunks before its own (it
// sums those chunk totals itself; a block's 256 cells lie in one chunk).
// Block 0 also writes the grand total, advances the time unless the step was
// skipped because of an earlier error, and records the error state:
// err[3] = this step was skipped, err[2] = error state for the next step.
static_assert(kScanChunk % 256 == 0, "a k_scan_add block must lie in one chunk");

__global__ __launch_bounds__(256) void k_scan_add(int *__restrict__ out, const int *__restrict__ bsum, int nb,
                                                 int n, unsigned *time, int *err, int *ovf_count, int tick) {
  __shared__ int lds[256 / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = (blockIdx.x * 256) / kScanChunk;
  const int upto = blockIdx.x == 0 ? nb : chunk;  // block 0 also needs the grand total
  int v = 0;
  for (int j = threadIdx.x; j < upto; j += 256) v += bsum[j];
  int total;
  block_exclusive_scan(v, lds, total);
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      out[n] = total;
      *ovf_count = 0;
      const int skipped = err[2];
      err[3] = skipped;
      if (!skipped && tick) *time += 1u;  // once per step (the last Runge-Kutta stage)
      err[2] = err[0];
    }
    return;  // chunk 0 has nothing before it
  }
  if (i < n) out[i] += total;
}

// Per-replica sums of the force kernels' per-brick partials: nred doubles per
// brick, brick-major (3 without the virial, kNRed with it).
struct FinalArgs {
  const double *__restrict__ partials;
  const int2 *__restrict__ rep_tiles;  // brick range of each replica
  int nrep, nred;
  double *__restrict__ obs;
  const int *__restrict__ err;
};

// One block (256 threads) sums replica r's partials in a fixed order: thread
// t takes bricks t, t + 256, ... in order, then waves in order, so the
// result is the same every run.  The slots past nred are stored as zero.
__device__ __forceinline__ void finalize_replica(const FinalArgs &F, int r) {
  __shared__ double red[kNRed][4];
  if (F.err[3]) return;  // the step was skipped: keep the last observables
  const int2 tr = F.rep_tiles[r];
  double acc[kNRed];
#pragma unroll
  for (int k = 0; k < kNRed; ++k) acc[k] = 0.0;
  int t = tr.x + (int)threadIdx.x;
  if (F.nred > 3) {
    for (; t < tr.y; t += 256)
#pragma unroll
      for (int k = 0; k < kNRed; ++k) acc[k] += F.partials[(size_t)t * kNRed + k];
  } else {
    // four bricks' loads in flight, added in brick order
    for (; t + 3 * 256 < tr.y; t += 4 * 256) {
      double p[4][3];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int k = 0; k < 3; ++k) p[u][k] = F.partials[(size_t)(t + u * 256) * 3 + k];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += p[u][k];
    }
    for (; t < tr.y; t += 256)
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[k] += F.partials[(size_t)t * 3 + k];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kNRed; ++k) {
    if (k >= F.nred) break;
    double w = wave_sum(acc[k]);
    if (lane == 0) red[k][wave] = w;
  }
  __syncthreads();
  if (threadIdx.x < kNRed) {
    double sum = 0.0;
    if ((int)threadIdx.x < F.nred)
      for (int w = 0; w < 4; ++w) sum += red[threadIdx.x][w];
    F.obs[(size_t)r * kNRed + threadIdx.x] = sum;
  }
}

// Moves each particle to its slot of the next step.  A particle that stays in
// its cell keeps its rank among the cell's stayers; particles entering a cell
// follow its stayers (their order is made deterministic by k_order_incoming).
// The first F.nrep blocks sum the replicas' partials instead (independent of
// the move; they run alongside it).
__global__ __launch_bounds__(256) void k_place(RebuildArgs A, FinalArgs F) {
  if ((int)blockIdx.x < F.nrep) {
    finalize_replica(F, blockIdx.x);
    return;
  }
  const int q = (blockIdx.x - F.nrep) * blockDim.x + threadIdx.x;
  if (q >= A.nslots) return;
  const int raw = A.cell_out[q];
  const int c1 = raw & 0x7fffffff;
  int dst;
  if (raw >= 0) {
    // Stayer: rank among the stayers of its cell (old slot order).
    int rank = 0;
    for (int q2 = A.cstart[c1]; q2 < q; ++q2) rank += (A.cell_out[q2] == c1);
    dst = A.cstart_new[c1] + rank;
  } else {
    const int stay = (A.cstart_new[c1 + 1] - A.cstart_new[c1]) - A.inc[c1];
    dst = A.cstart_new[c1] + stay + atomicAdd(&A.cursor[c1], 1);
  }
  A.x[dst] = A.x_in[q];
  A.v[dst] = A.v_in[q];
  if (A.with_pf) A.pf[dst] = A.pf_in[q];
  if (A.with_rk) {
#pragma unroll
    for (int k = 0; k < 4; ++k) A.rk[k][dst] = A.rk_in[k][q];
  }
  if (A.with_spec) {
    A.spec[0][dst] = A.spec_in[0][q];
    A.spec[1][dst] = A.spec_in[1][q];
  }
  const int p = A.pid_in[q];
  A.pid[dst] = p;
  A.cell[dst] = c1;
  if (A.with_slots) A.slot_of[p] = dst;
}

// slot_of from the slot order (after steps that did not keep it current).
__global__ __launch_bounds__(256) void k_slots(const int *__restrict__ pid, int n, int *__restrict__ slot_of) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) slot_of[pid[k]] = k;
}

// Sorts the particles that entered each cell by id (their arrival order in
// k_place depends on scheduling) and resets the per-cell counters.
__global__ __launch_bounds__(256) void k_order_incoming(RebuildArgs A) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= A.ncells) return;
  const int n_in = A.inc[c];
  if (n_in >= 2) {
    const int base = A.cstart_new[c + 1] - n_in;
    for (int i = 1; i < n_in; ++i) {
      for (int j = i; j > 0 && A.pid[base + j - 1] > A.pid[base + j]; --j) {
        const int a = base + j - 1, b = base + j;
        const int tp = A.pid[a];
        A.pid[a] = A.pid[b];
        A.pid[b] = tp;
        const float4 tx = A.x[a];
        A.x[a] = A.x[b];
        A.x[b] = tx;
        const float4 tv = A.v[a];
        A.v[a] = A.v[b];
        A.v[b] = tv;
        if (A.with_pf) {
          const float4 tf = A.pf[a];
          A.pf[a] = A.pf[b];
          A.pf[b] = tf;
        }
        if (A.with_rk) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float4 t = A.rk[r][a];
            A.rk[r][a] = A.rk[r][b];
            A.rk[r][b] = t;
          }
        }
        if (A.with_spec) {
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            const float4 t = A.spec[r][a];
            A.spec[r][a] = A.spec[r][b];
            A.spec[r][b] = t;
          }
        }
      }
    }
    if (A.with_slots)
      for (int k = base; k < base + n_in; ++k) A.slot_of[A.pid[k]] = k;
  }
  if (n_in) A.inc[c] = 0;
  A.cursor[c] = 0;
}


// engine_kinetic_energy's per-type sums (engine.cpp:1698-1719) for the
// Berendsen force: 1/2 m v^2 per particle in FP32, summed in FP64 per block
// (a fixed strided subset per thread), then over blocks in block order.
constexpr int kKeBlocks = 256;

__global__ __launch_bounds__(256) void k_type_ke(const float4 *__restrict__ x, const float4 *__restrict__ v,
                                                 const DevType *__restrict__ types, int n,
                                                 double *__restrict__ part) {
  double acc[kMaxTypes];
#pragma unroll
  for (int t = 0; t < kMaxTypes; ++t) acc[t] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int ti = __float_as_int(x[i].w) & 0xffff;
    const float4 vv = v[i];
    const float e = 0.5 * types[ti].mass * (vv.x * vv.x + vv.y * vv.y + vv.z * vv.z);
#pragma unroll
    for (int t = 0; t < kMaxTypes; ++t) acc[t] += ti == t ? (double)e : 0.0;
  }
  __shared__ double red[kMaxTypes][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < kMaxTypes; ++t) {
    const double w = wave_sum(acc[t]);
    if (lane == 0) red[t][wave] = w;
  }
  __syncthreads();
  if (threadIdx.x < kMaxTypes)
    part[blockIdx.x * kMaxTypes + threadIdx.x] = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] +
                                                 red[threadIdx.x][3];
}

__global__ __launch_bounds__(64) void k_type_ke_final(const double *__restrict__ part, double *__restrict__ ke) {
  const int t = threadIdx.x;
  if (t >= kMaxTypes) return;
  double s = 0.0;
  for (int b = 0; b < kKeBlocks; ++b) s += part[b * kMaxTypes + t];
  ke[t] = s;
}

// Bricks whose halo overflowed LDS: the same computation with neighbours
// read from global memory, one thread per particle of the brick.
template <int MODE>
__global__ __launch_bounds__(kBT) void k_brick_global(const StepArgs *__restrict__ Ap, BrickArgs B) {
  const StepArgs &A = *Ap;
  __shared__ int scan_lds[kBT / 64];
  __shared__ int homestart[kScanItems * kBT + 1];
  __shared__ int homegb[kScanItems * kBT];
  __shared__ int home_stay[kScanItems * kBT];  // stayers per home cell
  __shared__ double red[kNRed][kBT / 64];
  extern __shared__ __align__(16) unsigned char gsmem[];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int nov = *B.ovf_count;
  if ((int)blockIdx.x >= nov) return;  // usually no brick overflowed
  const int passthrough = A.err[2];
  const unsigned t1 = *A.time + 1u;
  const PairTabs tb = stage_pair_tabs(A, gsmem, A.npots);
  LoopConsts C = loop_consts(A);
  constexpr bool kFlux = (MODE & kModeFlux) != 0;
  const DevFlux G *const fluxes = kFlux ? A.fluxes : nullptr;
  const int G *const flux_off = kFlux ? A.flux_off : nullptr;
  const double flux_cut = kFlux ? A.cutoff_d : 0.0;
  __syncthreads();
  for (int i = blockIdx.x; i < nov; i += gridDim.x) {
    const int bi = B.ovf_list[i];
    const int4 br = B.bricks[bi];
    const DevReplica R = A.rep[br.x];  // by value: the fields stay in registers
    C.seed_lo = R.seed_lo;
    C.seed_hi = R.seed_hi;
    C.id_base = R.id_base;
    const uint32_t tk = step_key(R.seed_lo, t1);
    const int nb0 = min(B.bdim[0], R.cdim[0] - br.y);
    const int nb1 = min(B.bdim[1], R.cdim[1] - br.z);
    const int nb2 = min(B.bdim[2], R.cdim[2] - br.w);
    const int nhome = nb0 * nb1 * nb2;
    {
      int cnt[kScanItems];
      int local = 0;
#pragma unroll
      for (int k = 0; k < kScanItems; ++k) {
        const int hc = tid * kScanItems + k;
        cnt[k] = 0;
        if (hc < nhome) {
          const int a = hc / (nb1 * nb2), b = (hc / nb2) % nb1, cc = hc % nb2;
          const int c = R.cell_base + ((br.y + a) * R.cdim[1] + (br.z + b)) * R.cdim[2] + (br.w + cc);
          homegb[hc] = A.cstart[c];
          cnt[k] = A.cstart[c + 1] - A.cstart[c];
          home_stay[hc] = 0;
        }
        local += cnt[k];
      }
      int total;
      int off = block_exclusive_scan(local, scan_lds, total);
#pragma unroll
      for (int k = 0; k < kScanItems; ++k) {
        const int hc = tid * kScanItems + k;
        if (hc < nhome) homestart[hc] = off;
        off += cnt[k];
      }
      if (tid == 0) homestart[nhome] = total;
    }
    __syncthreads();
    const int nhp = homestart[nhome];
    double acc[kNRed];
#pragma unroll
    for (int k = 0; k < kNRed; ++k) acc[k] = 0.0;
    for (int f = tid; f < nhp; f += kBT) {
      const int hc = upper_index(homestart, nhome + 1, f);
      const int s = homegb[hc] + (f - homestart[hc]);
      const int a = hc / (nb1 * nb2), b = (hc / nb2) % nb1, cc = hc % nb2;
      const int lx = br.y + a, ly = br.z + b, lz = br.w + cc;
      const int c = R.cell_base + (lx * R.cdim[1] + ly) * R.cdim[2] + lz;
      const float4 xs = A.x[s];
      const float4 vs = A.v[s];
      const int mypid = A.pid[s];
      const float4 sme = kFlux ? float4(A.spec[s]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      float4 sq = kFlux ? float4(A.specq[s]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const int ti = __float_as_int(xs.w) & 0xffff;
      const DevType T = A.types[ti];
      {
        float e = 0.5 * T.mass * (vs.x * vs.x + vs.y * vs.y + vs.z * vs.z);
        acc[2] += (double)e;
      }
      float fx = 0.0f, fy = 0.0f, fz = 0.0f;
      // space_prepare sets f[3] to the self density W(0, cutoff).
      float nd = A.nd_norm;
      float4 pfs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (!passthrough) {
        if (A.random) single_body(A, C, s, mypid, ti, T, t1, B.rk_stage <= 0, vs, fx, fy, fz, pfs);
        if (R.wall >> kWallActiveShift) {
          const int wm = cell_walls(R, lx, ly, lz);
          if (wm) wall_forces(A, R, wm, ti, T.radius, xs, vs, C.dpd_delta, fx, fy, fz, acc[0]);
        }
        const GlobalNbr nb{A.v, A.pid};
        for_each_neighbour(A, C, R, s, lx, ly, lz, xs,
                           [&](int q, bool p_is_i, bool self, float dx, float dy, float dz, float r2,
                               const float4 &xq) {
                             const int tj = __float_as_int(xq.w) & 0xffff;
                             const float r = pair_root<MODE>(r2);
                             if (pair_force<MODE>(C, tb, nb, ti, tj, p_is_i, mypid, q, self, dx, dy, dz, r2, r, vs,
                                                  fx, fy, fz, acc, t1, tk))
                               nd += w_cubic(r, C.cutoff, C.nd_norm);
                             if (kFlux)
                               flux_pair(fluxes, flux_off, C.ntypes, flux_cut, ti, tj, self ? s < q : p_is_i, r, sme,
                                         float4(A.spec[q]), sq);
                           });
        if (MODE & kModeBonds) {
          bond_forces<MODE>(A, R, s, mypid, fx, fy, fz, acc);
          term_forces(A, mypid, fx, fy, fz);
        }
      }
      const bool stayed =
          B.rk_stage < 0 ? advance_store(A, R, s, c, lx, ly, lz, xs, vs, T, passthrough, fx, fy, fz, pfs, mypid, nd)
                         : rk_store(A, R, B.rk_stage, s, c, lx, ly, lz, xs, vs, T, passthrough, fx, fy, fz, pfs,
                                    mypid, nd);
      if (stayed) atomicAdd(&home_stay[hc], 1);
      if (kFlux) flux_store(A, s, stayed && B.rk_stage < 0, T.spec_const, sme, sq);
    }
#pragma unroll
    for (int k = 0; k < kNRed; ++k) {
      if (!(MODE & kModeVirial) && k >= 3) break;
      double w = wave_sum(acc[k]);
      if (lane == 0) red[k][wave] = w;
    }
    __syncthreads();
    // kNRed partials per brick with the virial, 3 without (finalize_replica
    // reads the same stride).
    constexpr int kPart = (MODE & kModeVirial) ? kNRed : 3;
    if (tid < kPart) {
      double sum = 0.0;
      for (int w = 0; w < kBT / 64; ++w) sum += red[tid][w];
      A.partials[(size_t)bi * kPart + tid] = sum;
    }
    for (int hc = tid; hc < nhome; hc += kBT) {
      const int a = hc / (nb1 * nb2), b = (hc / nb2) % nb1, cc = hc % nb2;
      A.cnt[R.cell_base + ((br.y + a) * R.cdim[1] + (br.z + b)) * R.cdim[2] + (br.w + cc)] = home_stay[hc];
    }
    __syncthreads();
  }
}

template <int MODE, int WAVES = 3>
void launch_step_mode(const StepArgs *A, const BrickArgs &B, size_t lds, size_t tabs, hipStream_t st) {
  hipLaunchKernelGGL((k_brick<MODE, WAVES>), dim3(B.nbricks), dim3(kBT), lds, st, A, B);
  hipLaunchKernelGGL(k_brick_global<MODE>, dim3(std::min(B.nbricks, kGlobalBlocks)), dim3(kBT), tabs, st, A, B);
}

using StepLauncher = void (*)(const StepArgs *, const BrickArgs &, size_t, size_t, hipStream_t);

// One instantiation per combination of table / DPD / virial / bonds / flux.
const StepLauncher kStepLaunchers[32] = {
    launch_step_mode<0>,  launch_step_mode<1>,  launch_step_mode<2>,  launch_step_mode<3>,
    launch_step_mode<4>,  launch_step_mode<5>,  launch_step_mode<6>,  launch_step_mode<7>,
    launch_step_mode<8>,  launch_step_mode<9>,  launch_step_mode<10>, launch_step_mode<11>,
    launch_step_mode<12>, launch_step_mode<13>, launch_step_mode<14>, launch_step_mode<15>,
    launch_step_mode<16>, launch_step_mode<17>, launch_step_mode<18>, launch_step_mode<19>,
    launch_step_mode<20>, launch_step_mode<21>, launch_step_mode<22>, launch_step_mode<23>,
    launch_step_mode<24>, launch_step_mode<25>, launch_step_mode<26>, launch_step_mode<27>,
    launch_step_mode<28>, launch_step_mode<29>, launch_step_mode<30>, launch_step_mode<31>,
};

// Four waves per SIMD, bonded modes only (nullptr: three waves).
const StepLauncher kStepLaunchers4[32] = {
    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
    launch_step_mode<8, 4>,  launch_step_mode<9, 4>,  launch_step_mode<10, 4>, launch_step_mode<11, 4>,
    launch_step_mode<12, 4>, launch_step_mode<13, 4>, launch_step_mode<14, 4>, launch_step_mode<15, 4>,
    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
    launch_step_mode<24, 4>, launch_step_mode<25, 4>, launch_step_mode<26, 4>, launch_step_mode<27, 4>,
    launch_step_mode<28, 4>, launch_step_mode<29, 4>, launch_step_mode<30, 4>, launch_step_mode<31, 4>,
};

}  // namespace

// ---------------------------------------------------------------------------
// Host side.

Engine::Engine() {}

Engine::~Engine() {
  if (d_) {
    if (d_->stream) hipStreamSynchronize(d_->stream);
    for (void *p : d_->owned) hipFree(p);
    for (auto &e : d_->ev)
      if (e) hipEventDestroy(e);
    for (auto &e : d_->kev)
      if (e) hipEventDestroy(e);
    for (auto &g : d_->gexec)
      if (g) hipGraphExecDestroy(g);
    if (d_->stream) hipStreamDestroy(d_->stream);
    delete d_;
  }
}

int Engine::fail(int code, const std::string &msg) {
  err_ = code;
  msg_ = msg;
  return code;
}

int Engine::reject(int code, const std::string &msg) {
  msg_ = msg;
  return code;
}

int Engine::check_device(int st, const char *what) {
  if (st != hipSuccess) return fail(kErrDevice, std::string(what) + ": " + hipGetErrorString((hipError_t)st));
  return kErrOk;
}

// Allocates n elements into *p, releasing what *p held before.
template <typename T>
static int dalloc(DeviceState *d, T **p, size_t n) {
  if (*p) {
    auto it = std::find(d->owned.begin(), d->owned.end(), (void *)*p);
    if (it != d->owned.end()) {
      d->owned.erase(it);
      hipFree(*p);
    }
    *p = nullptr;
  }
  void *q = nullptr;
  if (n == 0) n = 1;
  hipError_t st = hipMalloc(&q, n * sizeof(T));
  if (st != hipSuccess) return (int)st;
  d->owned.push_back(q);
  *p = (T *)q;
  return 0;
}

int Engine::init(const ReplicaSpec *specs, int nrep, double cutoff, double dt) {
  if (!specs || nrep < 1) return fail(kErrNull, "no replica given");
  if (!(cutoff > 0)) return fail(kErrCutoff, "cutoff must be positive");
  cutoff_ = cutoff;
  dt_ = dt;
  reps_.assign(nrep, HostReplica());
  int base = 0;
  for (int r = 0; r < nrep; ++r) {
    HostReplica &R = reps_[r];
    int cells[3] = {specs[r].cells[0], specs[r].cells[1], specs[r].cells[2]};
    // Periodic boundaries need at least three cells per axis
    // (MxBoundaryConditions.cpp:669-686).
    for (int k = 0; k < 3; ++k)
      if (cells[k] < 3) cells[k] = 3;
    double L[3];
    for (int k = 0; k < 3; ++k) L[k] = (specs[r].dim[k] - specs[r].origin[k]) / cells[k];
    // space_init (space.cpp:645-682): cdim = floor(dim / L), h = dim / cdim.
    for (int k = 0; k < 3; ++k) {
      R.origin[k] = specs[r].origin[k];
      R.dim[k] = specs[r].dim[k];
      R.cdim[k] = (int)std::floor(specs[r].dim[k] / L[k]);
      if (R.cdim[k] < 3) return fail(kErrSpace, "periodic space needs at least 3 cells per axis");
      R.h[k] = R.dim[k] / R.cdim[k];
      R.ih[k] = 1.0 / R.h[k];
      R.span[k] = (int)std::ceil(cutoff * R.ih[k]);
      if (R.cdim[k] < 2 * R.span[k] + 1)
        return fail(kErrSpace, "cell grid too coarse for the cutoff (needs cdim >= 2*span+1)");
      R.corg[k].resize(R.cdim[k]);
      for (int l = 0; l < R.cdim[k]; ++l) R.corg[k][l] = R.origin[k] + l * R.h[k];
    }
    R.cell_base = base;
    R.ncells = R.cdim[0] * R.cdim[1] * R.cdim[2];
    base += R.ncells;
  }
  ncells_ = base;
  rep_seed_.resize(nrep);
  for (int r = 0; r < nrep; ++r) rep_seed_[r] = replica_seed_default(seed_, r);
  d_ = new DeviceState();
  if (int rc = check_device(hipStreamCreateWithFlags(&d_->stream, hipStreamNonBlocking), "hipStreamCreate"))
    return rc;
  for (auto &e : d_->ev)
    if (int rc = check_device(hipEventCreate(&e), "hipEventCreate")) return rc;
  if (int rc = build_geometry()) return rc;
  return upload_walls();
}

// sort_key in host float arithmetic (same operations and order).
static unsigned host_sort_key(const DevReplica &D, const float x[3], const float n[3]) {
  const float v = D.dscale * (D.bias + x[0] * n[0] + x[1] * n[1] + x[2] * n[2]);
  return v <= 0.0f ? 0u : ((unsigned)v & 0xffffu);
}

// runner_dopair (runner_dopair.cpp:164-166) leaves its loop over the i cell
// when (smallest j key) + dnshift - (i key) exceeds dmaxdist as an unsigned
// value, and a negative distance wraps to a huge one: if the i particle with
// the largest key already sees a negative distance, the cell pair yields no
// pair at all, in the cutoff or not.  Keys are monotone in the position, so
// bounding them at the cell corners (positions in [-delta, h + delta]) and
// the shift norm from below over the shift table finds every offset where
// this can happen (kOffDrop); the others never drop.
//
// A dropping pair (i, j) has kj + dn < ki, hence
//   sum_k |n_k| (g_ik + g_jk) < sum_k |n_k| h_k - |shift| + 2.05 / dscale,
// with g the distance of each coordinate from the bound the key favours
// (g_ik = h_k - x_ik for n_k > 0, x_ik for n_k < 0, and the mirror for j).
// Every such gap is then below near_eps, which makes "both cells hold a
// particle that near their bounds" a filter that only rare cell pairs pass.
//
// Bit 30 (kOffWindow): in a cell pair that is not dropped every key distance
// kj + dn - ki is at least kjmin + dn - kimax >= 0, so the window can only
// reject a pair from above.  An in-cutoff pair has
//   kj + dn - ki < dscale (r + |shift| - shift.n) + 1 + (key rounding),
// and r <= cutoff: the window matters only where that bound reaches
// dmaxdist (non-cubic cells, offsets beyond the nearest cells, or a
// dscale * cutoff just below an integer).
static void flag_offsets(DevReplica &D, DevOffset *off, const std::vector<float> &shifts, float cut2) {
  float delta[3];
  for (int k = 0; k < 3; ++k) delta[k] = D.h[k] * 1e-6f;
  double slack = 0.0, dmax_delta = std::max(delta[0], std::max(delta[1], delta[2]));
  const double hmax = std::max(D.h[0], std::max(D.h[1], D.h[2]));
  // largest separation of a pair that passes r2 <= cut2 (float r2 and
  // float separations)
  const double rmax = std::sqrt((double)cut2) * (1.0 + 1e-6) + 2e-6 * hmax;
  D.drop_check = 0;
  for (int o = 0; o < D.n_off; ++o) {
    DevOffset &f = off[o];
    f.flags = 0;
    if (f.side == 0) continue;
    const int od[3] = {f.ox, f.oy, f.oz};
    const float n[3] = {f.nx, f.ny, f.nz};
    float xi[3], xj[3], smin[3], smax[3];
    int need_i = 0, need_j = 0;
    for (int k = 0; k < 3; ++k) {
      xi[k] = n[k] > 0 ? D.h[k] + delta[k] : (n[k] < 0 ? -delta[k] : 0.0f);
      xj[k] = n[k] > 0 ? -delta[k] : (n[k] < 0 ? D.h[k] + delta[k] : 0.0f);
      if (n[k] > 0) {
        need_i |= 2 << (2 * k);
        need_j |= 1 << (2 * k);
      } else if (n[k] < 0) {
        need_i |= 1 << (2 * k);
        need_j |= 2 << (2 * k);
      }
      // smallest |shift| of this offset's component over the cell rows
      const int w = 2 * D.span[k] + 1;
      float lo = INFINITY, hi = 0.0f;
      for (int l = 0; l < D.cdim[k]; ++l) {
        const float a = std::fabs(shifts[D.shift_base[k] + l * w + od[k] + D.span[k]]);
        lo = std::min(lo, a);
        hi = std::max(hi, a);
      }
      smin[k] = lo;
      smax[k] = hi;
    }
    {
      double ns = 0.0, dot = 0.0;
      for (int k = 0; k < 3; ++k) {
        ns += (double)smax[k] * smax[k];
        dot += std::fabs((double)n[k]) * smin[k];
      }
      const double bound = (double)D.dscale * (rmax + std::sqrt(ns) - dot) + 1.05;
      if (bound > (double)D.dmaxdist) f.flags |= kOffWindow;
    }
    const float nshift = std::sqrt(smin[0] * smin[0] + smin[1] * smin[1] + smin[2] * smin[2]);
    const int dn = (int)(D.dscale * nshift);
    const int ki = (int)host_sort_key(D, xi, n), kj = (int)host_sort_key(D, xj, n);
    if (kj + dn - ki >= 0) continue;
    int need_home = f.side > 0 ? need_i : need_j, need_nbr = f.side > 0 ? need_j : need_i;
    if (std::abs(f.ox) > 1 || std::abs(f.oy) > 1 || std::abs(f.oz) > 1) {
      // The near-bound bound above assumes adjacent cells (shift = cell
      // width per axis); a non-adjacent offset that can drop (none on any
      // grid seen so far) takes the exact test every time.
      need_home = 0;
      need_nbr = 0;
    }
    f.flags |= (int)(0x80000000u | (unsigned)need_home | ((unsigned)need_nbr << 8));
    D.drop_check = 1;
    if (need_home == 0) continue;  // no filter: near_eps does not apply
    double proj = 0.0;
    for (int k = 0; k < 3; ++k) proj += std::fabs((double)n[k]) * ((double)D.h[k] + 2.0 * delta[k]);
    slack = std::max(slack, proj - (double)nshift);
  }
  D.near_eps = (float)(1.5 * std::sqrt(3.0) * (slack + 2.05 / D.dscale + 6.0 * dmax_delta));
}

// engine_advance_forward_euler's clamps (engine_advance.cpp:262-281): the
// velocity and displacement limits of a step as a fraction of the cell.
static void set_clamps(DevReplica &D, float dt, float frac) {
  for (int k = 0; k < 3; ++k) {
    D.maxv[k] = (D.h[k] * frac) / dt;
    D.maxv2[k] = D.maxv[k] * D.maxv[k];
    D.maxx[k] = D.h[k] * frac;
    D.maxx2[k] = D.maxx[k] * D.maxx[k];
  }
}

// Offsets, shift tables and per-replica constants (once, at init: it
// allocates the geometry's device buffers).
int Engine::build_geometry() {
  std::vector<DevReplica> drep(reps_.size());
  std::vector<DevOffset> offs;
  std::vector<float> shifts;
  std::vector<int> cell_rep(ncells_);
  const float cutoff_f = (float)cutoff_;
  for (size_t r = 0; r < reps_.size(); ++r) {
    HostReplica &R = reps_[r];
    DevReplica &D = drep[r];
    std::memset(&D, 0, sizeof(D));
    for (int k = 0; k < 3; ++k) {
      D.cdim[k] = R.cdim[k];
      D.span[k] = R.span[k];
    }
    D.cell_base = R.cell_base;
    D.wall = 7 << kWallPeriodicShift;  // periodic until set_boundaries says otherwise
    D.seed_lo = (uint32_t)rep_seed_[r];
    D.seed_hi = (uint32_t)(rep_seed_[r] >> 32);
    D.id_base = R.gid_base;
    for (int c = 0; c < R.ncells; ++c) cell_rep[R.cell_base + c] = (int)r;
    for (int k = 0; k < 3; ++k) {
      D.h[k] = (float)R.h[k];
      D.h2[k] = (float)(2. * R.h[k]);
      D.hd[k] = R.h[k];
    }
    set_clamps(D, (float)dt_, max_frac_);
    // Sort-key scale of runner_dosort/runner_dopair.
    D.bias = (float)std::sqrt(R.h[0] * R.h[0] + R.h[1] * R.h[1] + R.h[2] * R.h[2]);
    D.dscale = (float)SHRT_MAX / (2 * D.bias);
    D.dmaxdist = (int)(2 + D.dscale * (cutoff_f + 2 * 0.0f));
    // Shift tables: shift[k][l][d] = min-image float(origin[l+d] - origin[l]).
    for (int k = 0; k < 3; ++k) {
      D.shift_base[k] = (int)shifts.size();
      const int w = 2 * R.span[k] + 1;
      for (int l = 0; l < R.cdim[k]; ++l)
        for (int d = -R.span[k]; d <= R.span[k]; ++d) {
          int l2 = ((l + d) % R.cdim[k] + R.cdim[k]) % R.cdim[k];
          float s = R.corg[k][l2] - R.corg[k][l];
          if (s * 2 > R.dim[k])
            s -= R.dim[k];
          else if (s * 2 < -R.dim[k])
            s += R.dim[k];
          shifts.push_back(s);
        }
      // The force kernel uses shift(l + d, -d) = -shift(l, d); check it.
      const int base = D.shift_base[k], sp = R.span[k];
      for (int l = 0; l < R.cdim[k]; ++l)
        for (int d = -sp; d <= sp; ++d) {
          const int l2 = ((l + d) % R.cdim[k] + R.cdim[k]) % R.cdim[k];
          if (shifts[base + l2 * w + (-d + sp)] != -shifts[base + l * w + (d + sp)])
            return fail(kErrSpace, "cell shift table is not antisymmetric");
        }
    }
    // Neighbour offsets within the span whose cells can hold a pair within
    // the cutoff (space.cpp:837-843), self first.
    D.off_base = (int)offs.size();
    DevOffset self = {0, 0, 0, 0, 0.f, 0.f, 0.f, 0};
    offs.push_back(self);
    for (int a = -R.span[0]; a <= R.span[0]; ++a)
      for (int b = -R.span[1]; b <= R.span[1]; ++b)
        for (int c = -R.span[2]; c <= R.span[2]; ++c) {
          if (a == 0 && b == 0 && c == 0) continue;
          double lh[3] = {R.h[0] * std::fmax(std::abs(a) - 1, 0), R.h[1] * std::fmax(std::abs(b) - 1, 0),
                          R.h[2] * std::fmax(std::abs(c) - 1, 0)};
          if (lh[0] * lh[0] + lh[1] * lh[1] + lh[2] * lh[2] > cutoff_ * cutoff_) continue;
          int sg[3] = {(a > 0) - (a < 0), (b > 0) - (b < 0), (c > 0) - (c < 0)};
          // forward = first non-zero sign component positive
          bool fwd = sg[0] > 0 || (sg[0] == 0 && (sg[1] > 0 || (sg[1] == 0 && sg[2] > 0)));
          int f[3] = {fwd ? sg[0] : -sg[0], fwd ? sg[1] : -sg[1], fwd ? sg[2] : -sg[2]};
          double len = std::sqrt((double)(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]));
          DevOffset o;
          o.ox = a;
          o.oy = b;
          o.oz = c;
          o.side = fwd ? 1 : -1;
          o.nx = (float)(f[0] / len);
          o.ny = (float)(f[1] / len);
          o.nz = (float)(f[2] / len);
          o.flags = 0;
          offs.push_back(o);
        }
    D.n_off = (int)offs.size() - D.off_base;
    flag_offsets(D, &offs[D.off_base], shifts, cutoff_f * cutoff_f);
  }
  DeviceState *d = d_;
  if (int st = dalloc(d, &d->rep, drep.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->offs, offs.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->shifts, shifts.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->cell_rep, cell_rep.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->cnt, (size_t)ncells_)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->inc, (size_t)ncells_)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->cursor, (size_t)ncells_)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->bsum, (size_t)ncells_ / kScanChunk + 2)) return check_device(st, "hipMalloc");
  for (int b = 0; b < 2; ++b)
    if (int st = dalloc(d, &d->cstart[b], (size_t)ncells_ + 1)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->time, 1)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->err, 4)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->obs, reps_.size() * kNRed)) return check_device(st, "hipMalloc");
  hipStream_t s = d->stream;
  int rc = 0;
  d->drep = drep;
  rc |= hipMemcpyAsync(d->rep, drep.data(), drep.size() * sizeof(DevReplica), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->offs, offs.data(), offs.size() * sizeof(DevOffset), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->shifts, shifts.data(), shifts.size() * sizeof(float), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->cell_rep, cell_rep.data(), cell_rep.size() * sizeof(int), hipMemcpyHostToDevice, s);
  rc |= hipMemsetAsync(d->cnt, 0, (size_t)ncells_ * sizeof(int), s);
  rc |= hipMemsetAsync(d->inc, 0, (size_t)ncells_ * sizeof(int), s);
  rc |= hipMemsetAsync(d->cursor, 0, (size_t)ncells_ * sizeof(int), s);
  rc |= hipMemsetAsync(d->time, 0, sizeof(unsigned), s);
  rc |= hipMemsetAsync(d->err, 0, 4 * sizeof(int), s);
  rc |= hipMemsetAsync(d->obs, 0, reps_.size() * kNRed * sizeof(double), s);
  rc |= hipStreamSynchronize(s);
  return check_device(rc, "geometry upload");
}

int Engine::add_type(double mass, float radius, int dynamics) {
  if (!(mass > 0)) return fail(kErrRange, "mass must be positive");
  if ((int)types_.size() >= kMaxTypes) return fail(kErrRange, "at most 16 particle types per engine");
  if (dynamics != kDynNewtonian && dynamics != kDynOverdamped) return fail(kErrRange, "unknown dynamics");
  TypeInfo t;
  t.mass = mass;
  t.imass = 1.0 / mass;
  t.radius = radius;
  t.dynamics = dynamics;
  types_.push_back(t);
  int nt = (int)types_.size();
  std::vector<int> m((size_t)nt * nt, -1);
  for (int i = 0; i < nt - 1; ++i)
    for (int j = 0; j < nt - 1; ++j) m[(size_t)i * nt + j] = potmat_[(size_t)i * (nt - 1) + j];
  potmat_.swap(m);
  potentials_dirty_ = true;
  return nt - 1;
}

// engine_addpot (engine.cpp:960-981): the same potential for (i, j), (j, i).
int Engine::add_pot(const Potential &p, int i, int j) {
  int nt = (int)types_.size();
  if (i < 0 || i >= nt || j < 0 || j >= nt) return fail(kErrRange, "type index out of range");
  if (p.flags & kPotFlagBound) return fail(kErrPotential, "cluster-bound potentials are not supported");
  if ((int)pots_.size() >= kMaxPots) return fail(kErrPotential, "at most 32 pair potentials per engine");
  if (p.kind == kPotKindTable && (p.n < 1 || p.c.size() < (size_t)(p.n + 1) * kPotChunk))
    return fail(kErrPotential, "potential has no table");
  pots_.push_back(p);
  int idx = (int)pots_.size() - 1;
  potmat_[(size_t)i * nt + j] = idx;
  potmat_[(size_t)j * nt + i] = idx;
  potentials_dirty_ = true;
  return kErrOk;
}

// engine_add_singlebody_force: one force per type, a later one replaces it.
int Engine::set_single_body(int type, int kind, const float *p) {
  if (type < 0 || type >= (int)types_.size()) return reject(kErrRange, "type index out of range");
  TypeInfo &t = types_[type];
  t.sb = kind;
  if (kind == kSbGaussian || kind == kSbFriction) {
    // random_create / friction_create: durration_steps = ceil(durration / dt)
    // (MxForce.cpp:397-420).
    const float mean = p[0], std = p[1], duration = p[2];
    t.rnd_mean = mean;
    t.rnd_std = std;
    double steps = std::ceil(duration / dt_);
    t.rnd_steps = steps < 1 ? 1u : (unsigned)steps;
    t.sb_coef = kind == kSbFriction ? p[3] : 0.f;
  } else if (kind == kSbBerendsen) {
    t.sb_coef = 1 / p[0];  // berenderson_create: itau = 1 / tau
  } else if (kind == kSbConstant) {
    for (int k = 0; k < 3; ++k) t.cf[k] = p[k];
  }
  any_random_ = false;
  any_sb_ = false;
  any_berendsen_ = false;
  for (const auto &u : types_) {
    any_random_ = any_random_ || u.sb == kSbGaussian || u.sb == kSbFriction;
    any_sb_ = any_sb_ || u.sb != kSbNone;
    any_berendsen_ = any_berendsen_ || u.sb == kSbBerendsen;
  }
  potentials_dirty_ = true;
  args_dirty_ = true;
  ++config_gen_;
  return kErrOk;
}

int Engine::set_random_force(int type, float mean, float std, float duration) {
  const float p[3] = {mean, std, duration};
  return set_single_body(type, kSbGaussian, p);
}

int Engine::set_target_energy(int type, double e) {
  if (type < 0 || type >= (int)types_.size()) return reject(kErrRange, "type index out of range");
  types_[type].target_energy = e;
  potentials_dirty_ = true;
  return kErrOk;
}

int Engine::set_seed(uint64_t seed) {
  seed_ = seed;
  for (size_t r = 0; r < rep_seed_.size(); ++r) rep_seed_[r] = replica_seed_default(seed, (int)r);
  args_dirty_ = true;
  return upload_replicas();
}

int Engine::set_replica_seed(int replica, uint64_t seed) {
  if (replica < 0 || replica >= (int)reps_.size()) return reject(kErrRange, "replica index out of range");
  rep_seed_[replica] = seed;
  return upload_replicas();
}

int Engine::replica_seed(int replica, uint64_t *seed) const {
  if (replica < 0 || replica >= (int)reps_.size() || !seed) return kErrRange;
  *seed = rep_seed_[replica];
  return kErrOk;
}

// Copies the per-replica constants that change after set-up (random-stream
// seeds and first gids) into the device table.
int Engine::upload_replicas() {
  if (!d_ || !d_->rep) return kErrOk;
  DeviceState *d = d_;
  for (size_t r = 0; r < reps_.size(); ++r) {
    DevReplica &D = d->drep[r];
    D.seed_lo = (uint32_t)rep_seed_[r];
    D.seed_hi = (uint32_t)(rep_seed_[r] >> 32);
    D.id_base = reps_[r].gid_base;
  }
  int rc = hipMemcpyAsync(d->rep, d->drep.data(), d->drep.size() * sizeof(DevReplica), hipMemcpyHostToDevice,
                          d->stream);
  rc |= hipStreamSynchronize(d->stream);
  return check_device(rc, "replica upload");
}

int Engine::set_dpd_counter_noise(int on) {
  dpd_noise_ = on ? 1 : 0;
  args_dirty_ = true;
  return kErrOk;
}

int Engine::set_max_dist_fraction(float f) {
  if (!(f > 0)) return reject(kErrRange, "max distance fraction must be positive");
  max_frac_ = f;
  args_dirty_ = true;
  // Only the advance's clamps change: the geometry and every device buffer
  // (cell offsets, counters, time) stay as they are.
  if (d_ && d_->rep) {
    for (size_t r = 0; r < reps_.size(); ++r) set_clamps(d_->drep[r], (float)dt_, max_frac_);
    return upload_replicas();
  }
  return kErrOk;
}

int Engine::set_block_size(int b) {
  if (b != kTile) return reject(kErrRange, "the force kernel is compiled for 256 threads per block");
  block_ = b;
  return kErrOk;
}

int Engine::upload_potentials() {
  DeviceState *d = d_;
  std::vector<float> coef;
  auto pack = [&](const Potential &p) {
    DevPot D;
    std::memset(&D, 0, sizeof(D));
    D.kind = p.kind;
    D.flags = p.flags;
    D.alpha0 = p.alpha[0];
    D.alpha1 = p.alpha[1];
    D.alpha2 = p.alpha[2];
    D.a = p.a;
    D.b = p.b;
    D.r0_plusone = p.r0_plusone;
    D.n = p.n;
    D.dpd_alpha = p.dpd_alpha;
    D.dpd_gamma = p.dpd_gamma;
    D.dpd_sigma = p.dpd_sigma;
    D.a2 = p.a * p.a;
    D.coff = (int)coef.size();
    if (p.kind == kPotKindTable) coef.insert(coef.end(), p.c.begin(), p.c.end());
    return D;
  };
  std::vector<DevPot> dp, dbp;
  any_dpd_ = any_table_ = false;
  for (auto &p : pots_) {
    dp.push_back(pack(p));
    if (p.kind == kPotKindDPD)
      any_dpd_ = true;
    else
      any_table_ = true;
  }
  for (auto &p : bpots_) dbp.push_back(pack(p));
  std::vector<DevPot> dtp;
  for (auto &p : apots_) dtp.push_back(pack(p));
  for (auto &p : dpots_) dtp.push_back(pack(p));
  std::vector<DevPot> dwp;
  for (auto &p : wpots_) dwp.push_back(pack(p));
  const int nt = (int)types_.size();
  std::vector<int> wallpot(6 * (size_t)nt, -1);
  for (size_t w = 0; w < wpot_key_.size(); ++w)
    if (wpot_key_[w].second < nt) wallpot[wpot_key_[w].first * nt + wpot_key_[w].second] = (int)w;
  std::vector<DevType> dt(types_.size());
  for (size_t i = 0; i < types_.size(); ++i) {
    dt[i].mass = types_[i].mass;
    dt[i].imass = types_[i].imass;
    dt[i].radius = types_[i].radius;
    dt[i].dynamics = types_[i].dynamics;
    dt[i].target_energy = types_[i].target_energy;
    dt[i].sb = types_[i].sb;
    dt[i].rnd_mean = types_[i].rnd_mean;
    dt[i].rnd_std = types_[i].rnd_std;
    dt[i].rnd_steps = types_[i].rnd_steps;
    dt[i].sb_coef = types_[i].sb_coef;
    for (int k = 0; k < 3; ++k) dt[i].cf[k] = types_[i].cf[k];
    dt[i].spec_const = types_[i].spec_const;
  }
  // Fluxes by ordered type pair; both orders of a pair list the same fluxes
  // in the order they were added (one MxFluxes per pair, Flux.cpp:197-208).
  std::vector<int> flux_off((size_t)nt * nt + 1, 0);
  std::vector<DevFlux> dflux;
  for (int t1 = 0; t1 < nt; ++t1)
    for (int t2 = 0; t2 < nt; ++t2) {
      for (const FluxInfo &F : fluxes_)
        if ((F.ta == t1 && F.tb == t2) || (F.ta == t2 && F.tb == t1))
          dflux.push_back(DevFlux{F.kind, F.ta, F.tb, F.ia, F.ib, F.k, F.decay, F.target});
      flux_off[t1 * nt + t2 + 1] = (int)dflux.size();
    }
  // Potentials change at set-up only, so reallocating here is fine.
  auto upload = [&](auto **dst, const auto &vec) -> int {
    using T = typename std::remove_reference<decltype(vec)>::type::value_type;
    if (int st = dalloc(d, dst, vec.size())) return st;
    if (!vec.empty())
      return (int)hipMemcpyAsync(*dst, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice, d->stream);
    return 0;
  };
  int rc = 0;
  rc |= upload(&d->pots, dp);
  rc |= upload(&d->bpots, dbp);
  rc |= upload(&d->tpots, dtp);
  rc |= upload(&d->wpots, dwp);
  rc |= upload(&d->wallpot, wallpot);
  rc |= upload(&d->fluxes, dflux);
  rc |= upload(&d->flux_off, flux_off);
  rc |= upload(&d->coef, coef);
  rc |= upload(&d->potmat, potmat_);
  rc |= upload(&d->types, dt);
  rc |= hipStreamSynchronize(d->stream);
  if (rc) return check_device(rc, "potential upload");
  potentials_dirty_ = false;
  args_dirty_ = true;
  return kErrOk;
}

// engine_load / space_addpart (space.cpp:397-477): cell index
// (x - origin) * ih in double, local coordinate x - cell origin, particles
// kept in insertion order inside each cell.
int Engine::load(int N, const double *x, const double *v, const int *type, const int *pid_in, const int *replica,
                 const uint16_t *flags) {
  if (!d_) return fail(kErrNull, "engine not initialised");
  if (N <= 0 || !x || !type) return fail(kErrNull, "empty particle set");
  const int nt = (int)types_.size();
  // Particle ids (MxParticle::id, engine_load's pid): unique, non-negative;
  // the id -> particle table is sized by the largest id, as partlist is.
  long long maxpid = N - 1;
  if (pid_in) {
    for (int i = 0; i < N; ++i) {
      if (pid_in[i] < 0) return fail(kErrRange, "particle ids must be non-negative");
      maxpid = std::max<long long>(maxpid, pid_in[i]);
    }
    if (maxpid >= 64LL * N + 65536) return fail(kErrRange, "particle ids must be below 64 N + 65536");
  }
  std::vector<int> rep_count(reps_.size(), 0), rank(N);
  for (int i = 0; i < N; ++i) {
    const int r = replica ? replica[i] : 0;
    if (r < 0 || r >= (int)reps_.size()) return fail(kErrRange, "replica index out of range");
    rank[i] = rep_count[r]++;
  }
  {
    int base = 0;
    for (size_t r = 0; r < reps_.size(); ++r) {
      reps_[r].gid_base = base;
      base += rep_count[r];
    }
  }
  std::vector<int> gid(N), gid_of_pid((size_t)maxpid + 1, -1), pid_of_gid(N);
  for (int i = 0; i < N; ++i) {
    gid[i] = reps_[replica ? replica[i] : 0].gid_base + rank[i];
    const int id = pid_in ? pid_in[i] : i;
    if (gid_of_pid[id] >= 0) return fail(kErrRange, "particle id " + std::to_string(id) + " given twice");
    gid_of_pid[id] = gid[i];
    pid_of_gid[gid[i]] = id;
  }
  std::vector<int> cellof(N);
  std::vector<float4> xl(N);
  for (int i = 0; i < N; ++i) {
    const int r = replica ? replica[i] : 0;
    if (r < 0 || r >= (int)reps_.size()) return fail(kErrRange, "replica index out of range");
    const int t = type[i];
    if (t < 0 || t >= nt) return fail(kErrRange, "type index out of range");
    if (types_[t].radius > cutoff_)
      return fail(kErrRange, "particles with radius larger than the cutoff (large particles) are not supported");
    const HostReplica &R = reps_[r];
    int ind[3];
    for (int k = 0; k < 3; ++k) ind[k] = (x[3 * (size_t)i + k] - R.origin[k]) * R.ih[k];
    for (int k = 0; k < 3; ++k)
      if (ind[k] < 0 || ind[k] >= R.cdim[k])
        return fail(kErrSpace, "particle " + std::to_string(i) + " lies outside the space");
    float loc[3];
    for (int k = 0; k < 3; ++k) loc[k] = x[3 * (size_t)i + k] - R.corg[k][ind[k]];
    // The first step's space_shuffle_local (space.cpp:333-393), applied
    // here: a local coordinate that rounded to the cell width (a particle
    // just below a face) or fell below zero moves the particle to the
    // neighbouring cell, x -= delta h in float, the index modulo the grid.
    bool at_face = false;
    for (int k = 0; k < 3; ++k) {
      const float hf = (float)R.h[k];
      const int d = (loc[k] >= hf ? 1 : 0) - (loc[k] < 0.0f ? 1 : 0);
      if (d) {
        loc[k] -= d * hf;
        ind[k] = (ind[k] + d + R.cdim[k]) % R.cdim[k];
        at_face = at_face || loc[k] >= hf;
      }
    }
    cellof[i] = R.cell_base + (ind[0] * R.cdim[1] + ind[1]) * R.cdim[2] + ind[2];
    const uint16_t fl = (uint16_t)(((flags ? flags[i] : 0) & ~kShuffledToFace) | (at_face ? kShuffledToFace : 0));
    const unsigned bits = (unsigned)t | ((unsigned)fl << 16);
    float wbits;
    std::memcpy(&wbits, &bits, sizeof(wbits));
    xl[i] = make_float4(loc[0], loc[1], loc[2], wbits);
  }
  // Stable counting sort by cell.
  std::vector<int> cstart(ncells_ + 1, 0);
  for (int i = 0; i < N; ++i) cstart[cellof[i] + 1]++;
  for (int c = 0; c < ncells_; ++c) {
    if (cstart[c + 1] > 65535) return fail(kErrMaxParts, "more than 65535 particles in one cell");
  }
  for (int c = 0; c < ncells_; ++c) cstart[c + 1] += cstart[c];
  std::vector<int> fill(cstart.begin(), cstart.end() - 1);
  std::vector<float4> xs(N), vs(N);
  std::vector<int> pid(N), cell(N), slot_of(N);
  for (int i = 0; i < N; ++i) {
    const int k = fill[cellof[i]]++;
    const int g = gid[i];
    xs[k] = xl[i];
    float idbits;
    std::memcpy(&idbits, &g, sizeof(idbits));  // v.w carries the gid
    vs[k] = v ? make_float4((float)v[3 * (size_t)i], (float)v[3 * (size_t)i + 1], (float)v[3 * (size_t)i + 2], idbits)
              : make_float4(0.f, 0.f, 0.f, idbits);
    pid[k] = g;
    cell[k] = cellof[i];
    slot_of[g] = k;
  }
  // Tiles: per replica, slot ranges of kTile.
  std::vector<int4> tiles;
  std::vector<int2> rep_tiles(reps_.size());
  for (size_t r = 0; r < reps_.size(); ++r) {
    HostReplica &R = reps_[r];
    R.slot_base = cstart[R.cell_base];
    R.nparts = cstart[R.cell_base + R.ncells] - R.slot_base;
    R.tile_begin = (int)tiles.size();
    for (int s = 0; s < R.nparts; s += kTile)
      tiles.push_back(make_int4(R.slot_base + s, R.slot_base + std::min(s + kTile, R.nparts), (int)r, 0));
    R.tile_end = (int)tiles.size();
    rep_tiles[r] = make_int2(R.tile_begin, R.tile_end);
  }
  N_ = N;
  prep_.assign(N, 0);
  gid_of_load_ = gid;
  load_of_gid_.assign(N, 0);
  for (int i = 0; i < N; ++i) {
    load_of_gid_[gid[i]] = i;
    prep_[gid[i]] = replica ? replica[i] : 0;
  }
  gid_of_pid_.swap(gid_of_pid);
  pid_of_gid_.swap(pid_of_gid);
  // Bonds name particles of the previous load.
  M_ = 0;
  // Species buffers are sized by the particle count.
  const bool had_species = species_on_;
  species_on_ = false;
  // Angles, dihedrals and exclusions name particle ids of the previous load.
  angles_.clear();
  dihedrals_.clear();
  exclusions_.clear();
  NT_ = 0;
  ntiles_ = (int)tiles.size();
  mean_cell_count_.assign(reps_.size(), 0.0);
  for (size_t r = 0; r < reps_.size(); ++r)
    mean_cell_count_[r] = reps_[r].ncells ? (double)reps_[r].nparts / reps_[r].ncells : 0.0;

  DeviceState *d = d_;
  // (Re)allocate particle buffers.
  for (int b = 0; b < 2; ++b) {
    if (int st = dalloc(d, &d->x[b], N)) return check_device(st, "hipMalloc");
    if (int st = dalloc(d, &d->v[b], N)) return check_device(st, "hipMalloc");
    if (int st = dalloc(d, &d->pf[b], N)) return check_device(st, "hipMalloc");
    if (b == 0) {
      if (int st = dalloc(d, &d->type_ke, kMaxTypes)) return check_device(st, "hipMalloc");
      if (int st = dalloc(d, &d->type_ke_part, (size_t)kKeBlocks * kMaxTypes)) return check_device(st, "hipMalloc");
    }
    if (int st = dalloc(d, &d->pid[b], N)) return check_device(st, "hipMalloc");
    if (int st = dalloc(d, &d->cell[b], N)) return check_device(st, "hipMalloc");
  }
  if (int st = dalloc(d, &d->f, N)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->slot_of, N)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->ncount, N)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->tiles, tiles.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->rep_tiles, rep_tiles.size())) return check_device(st, "hipMalloc");
  hipStream_t s = d->stream;
  int rc = 0;
  rc |= hipMemcpyAsync(d->x[0], xs.data(), N * sizeof(float4), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->v[0], vs.data(), N * sizeof(float4), hipMemcpyHostToDevice, s);
  rc |= hipMemsetAsync(d->pf[0], 0, N * sizeof(float4), s);
  rc |= hipMemsetAsync(d->pf[1], 0, N * sizeof(float4), s);
  rc |= hipMemsetAsync(d->f, 0, N * sizeof(float4), s);
  rc |= hipMemcpyAsync(d->pid[0], pid.data(), N * sizeof(int), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->pid[1], pid.data(), N * sizeof(int), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->cell[0], cell.data(), N * sizeof(int), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->cstart[0], cstart.data(), (ncells_ + 1) * sizeof(int), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->slot_of, slot_of.data(), N * sizeof(int), hipMemcpyHostToDevice, s);
  slots_valid_ = true;
  rc |= hipMemcpyAsync(d->tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->rep_tiles, rep_tiles.data(), rep_tiles.size() * sizeof(int2), hipMemcpyHostToDevice, s);
  rc |= hipStreamSynchronize(s);
  if (rc) return check_device(rc, "particle upload");
  pid_prev_valid_ = false;
  parity_ = 0;
  args_dirty_ = true;
  if (int rc = upload_replicas()) return rc;
  if (int rc = ensure_rk_buffers()) return rc;
  if (had_species || !fluxes_.empty())
    if (int rc = ensure_species()) return rc;
  return set_brick(0, 0, 0, 0, 0);
}

// Capacities of the LDS layout for a brick shape: halo cells, halo cells per
// axis, neighbour offsets, and the particle buffer (the expected halo
// population plus margin, within the LDS budget).
struct BrickFit {
  int pmax = 0, hmax = 0, hdmax = 0, noff = 0;
  double worst = 0;  // expected particles in the fullest halo
};

static BrickFit brick_fit(const std::vector<HostReplica> &reps, const std::vector<double> &mean, const int bd[3],
                          int cap, bool dpd, bool flux, size_t tabs, size_t budget) {
  BrickFit F;
  for (size_t r = 0; r < reps.size(); ++r) {
    int hc = 1;
    for (int k = 0; k < 3; ++k) {
      const int hd = std::min(bd[k], reps[r].cdim[k]) + 2 * reps[r].span[k];
      F.hdmax = std::max(F.hdmax, hd);
      hc *= hd;
    }
    const int sp = std::max(reps[r].span[0], std::max(reps[r].span[1], reps[r].span[2]));
    F.noff = std::max(F.noff, (2 * sp + 1) * (2 * sp + 1) * (2 * sp + 1));
    F.hmax = std::max(F.hmax, hc);
    F.worst = std::max(F.worst, mean[r] * hc);
  }
  int want = (int)std::ceil(1.2 * F.worst + 4.0 * std::sqrt(F.worst) + 32.0);
  want = (want + 31) / 32 * 32;
  int fit = 0;
  for (int pm = 32; pm <= 16384; pm += 32)
    if (brick_lds_bytes(pm, F.hmax, F.hdmax, F.noff, cap, dpd, flux, tabs) <= budget) fit = pm;
  F.pmax = std::min(want, fit);
  return F;
}

// Chooses (or sets) the brick shape of the force kernel and builds the brick
// list.  bx = 0 picks the largest shape whose expected halo fits in LDS.
int Engine::set_brick(int bx, int by, int bz, int cap, int pmax) {
  if (!d_ || N_ == 0) return fail(kErrNull, "no particles loaded");
  const size_t budget = 65536;
  const bool dpd = any_dpd_ || std::any_of(pots_.begin(), pots_.end(), [](const Potential &p) {
                     return p.kind == kPotKindDPD;
                   });
  int c = cap > 0 ? cap : cap_;
  if (c < kChunkMax || c > 64)
    return reject(kErrRange, "neighbour-list capacity must lie in [" + std::to_string(kChunkMax) + ", 64]");
  const size_t tabs = pair_tab_bytes((int)pots_.size(), (int)types_.size());
  int bd[3] = {bx, by, bz};
  int hmax = 0, pm = 0;
  BrickFit fitv;
  for (const auto &r : reps_)
    for (int k = 0; k < 3; ++k)
      if (r.span[k] > kSpanMax) return reject(kErrCutoff, "cutoff spans more than 2 cells");
  if (bx <= 0) {
    static const int shapes[][3] = {{4, 4, 4}, {2, 4, 4}, {2, 2, 4}, {2, 2, 2}, {1, 2, 2}, {1, 1, 2}, {1, 1, 1}};
    bool ok = false;
    for (const auto &sh : shapes) {
      for (int k = 0; k < 3; ++k) bd[k] = sh[k];
      fitv = brick_fit(reps_, mean_cell_count_, bd, c, dpd, species_on_, tabs, budget);
      if (fitv.pmax >= 1.1 * fitv.worst && fitv.hmax <= kScanItems * kBT) {
        ok = true;
        break;
      }
    }
    if (!ok) {
      for (int k = 0; k < 3; ++k) bd[k] = 1;
      fitv = brick_fit(reps_, mean_cell_count_, bd, c, dpd, species_on_, tabs, budget);
    }
  } else {
    if (bx < 1 || by < 1 || bz < 1) return reject(kErrRange, "brick dimensions must be positive");
    fitv = brick_fit(reps_, mean_cell_count_, bd, c, dpd, species_on_, tabs, budget);
  }
  hmax = fitv.hmax;
  pm = fitv.pmax;
  // An explicit capacity below the fit sends fuller bricks to the global path.
  if (pmax > 0) pm = std::min(pm, (pmax + 31) / 32 * 32);
  if (fitv.hdmax > kHaloAxis) return reject(kErrRange, "brick too large");
  if (hmax > kScanItems * kBT) return reject(kErrRange, "brick halo has more than 1024 cells");
  if (pm < 64) return reject(kErrRange, "brick does not fit in LDS");
  std::vector<int4> bricks;
  std::vector<int2> rep_bricks(reps_.size());
  for (size_t r = 0; r < reps_.size(); ++r) {
    const HostReplica &R = reps_[r];
    rep_bricks[r].x = (int)bricks.size();
    for (int x = 0; x < R.cdim[0]; x += bd[0])
      for (int y = 0; y < R.cdim[1]; y += bd[1])
        for (int z = 0; z < R.cdim[2]; z += bd[2]) bricks.push_back(make_int4((int)r, x, y, z));
    rep_bricks[r].y = (int)bricks.size();
  }
  // One universe: each XCD walks an eighth of the list in order (xcd_brick),
  // so the bricks cut short by the grid's upper edges go to the end of each
  // eighth, and the step's last workgroups are the cheap ones.
  if (reps_.size() == 1) {
    const HostReplica &R = reps_[0];
    const size_t part = (bricks.size() - bricks.size() % 8) / 8;
    auto whole = [&](const int4 &b) {
      return b.y + bd[0] <= R.cdim[0] && b.z + bd[1] <= R.cdim[1] && b.w + bd[2] <= R.cdim[2];
    };
    for (int k = 0; k < 8 && part > 0; ++k)
      std::stable_partition(bricks.begin() + k * part, bricks.begin() + (k + 1) * part, whole);
  }
  DeviceState *d = d_;
  if (int st = dalloc(d, &d->bricks, bricks.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->rep_bricks, rep_bricks.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->partials, bricks.size() * kNRed)) return check_device(st, "hipMalloc");
  int rc = 0;
  rc |= hipMemcpyAsync(d->bricks, bricks.data(), bricks.size() * sizeof(int4), hipMemcpyHostToDevice, d->stream);
  rc |= hipMemcpyAsync(d->rep_bricks, rep_bricks.data(), rep_bricks.size() * sizeof(int2), hipMemcpyHostToDevice,
                       d->stream);
  rc |= hipMemsetAsync(d->partials, 0, bricks.size() * kNRed * sizeof(double), d->stream);
  rc |= hipStreamSynchronize(d->stream);
  if (rc) return check_device(rc, "brick upload");
  for (int k = 0; k < 3; ++k) bdim_[k] = bd[k];
  cap_ = c;
  pmax_ = pm;
  pmax_req_ = pmax;
  hmax_ = hmax;
  hdmax_ = fitv.hdmax;
  noff_ = fitv.noff;
  nbricks_ = (int)bricks.size();
  brick_dpd_ = dpd;
  brick_flux_ = species_on_;
  brick_tabs_ = tabs;
  if (int st = dalloc(d, &d->ovf_list, bricks.size())) return check_device(st, "hipMalloc");
  if (!d->ovf_count) {
    if (int st = dalloc(d, &d->ovf_count, 1)) return check_device(st, "hipMalloc");
    if (int st = hipMemset(d->ovf_count, 0, sizeof(int))) return check_device(st, "hipMemset");
  }
  args_dirty_ = true;
  return kErrOk;
}

// engine_bond_alloc + MxBond fields (bond.h:47-70): bonds are kept in a CSR
// by particle id so each particle gathers its own bond forces.
int Engine::set_bonds(int M, const int *bi, const int *bj, const Potential *const *pots, const double *dissociation) {
  if (!d_ || N_ == 0) return fail(kErrNull, "load particles before bonds");
  if (M < 0) return fail(kErrRange, "negative bond count");
  std::vector<int> vi(M), vj(M), vp(M);
  std::vector<double> vd(M);
  bpots_.clear();
  std::vector<const Potential *> seen;
  for (int b = 0; b < M; ++b) {
    const int gi = gid_of(bi[b]), gj = gid_of(bj[b]);
    if (gi < 0 || gj < 0 || gi == gj)
      return fail(kErrBond, "bond " + std::to_string(b) + " has an invalid particle id");
    if (prep_[gi] != prep_[gj]) return fail(kErrBond, "bond " + std::to_string(b) + " joins two replicas");
    const Potential *p = pots[b];
    if (!p || p->kind != kPotKindTable || p->n < 1) return fail(kErrBond, "bond potential must be a table potential");
    auto it = std::find(seen.begin(), seen.end(), p);
    int idx;
    if (it == seen.end()) {
      seen.push_back(p);
      bpots_.push_back(*p);
      idx = (int)bpots_.size() - 1;
    } else {
      idx = (int)(it - seen.begin());
    }
    vi[b] = gi;
    vj[b] = gj;
    vp[b] = idx;
    vd[b] = dissociation ? dissociation[b] : std::numeric_limits<double>::infinity();
  }
  std::vector<int> off(N_ + 1, 0);
  for (int b = 0; b < M; ++b) {
    off[vi[b] + 1]++;
    off[vj[b] + 1]++;
  }
  for (int i = 0; i < N_; ++i) off[i + 1] += off[i];
  std::vector<int> ent(2 * (size_t)M), fill(off.begin(), off.end() - 1);
  for (int b = 0; b < M; ++b) {
    ent[fill[vi[b]]++] = 2 * b;
    ent[fill[vj[b]]++] = 2 * b + 1;
  }
  DeviceState *d = d_;
  if (int st = dalloc(d, &d->boff, N_ + 1)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->bent, ent.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->bi, M)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->bj, M)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->bpot, M)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->bdis, M)) return check_device(st, "hipMalloc");
  for (int b = 0; b < 2; ++b)
    if (int st = dalloc(d, &d->bact[b], M)) return check_device(st, "hipMalloc");
  hipStream_t s = d->stream;
  int rc = 0;
  rc |= hipMemcpyAsync(d->boff, off.data(), (N_ + 1) * sizeof(int), hipMemcpyHostToDevice, s);
  if (M > 0) {
    rc |= hipMemcpyAsync(d->bent, ent.data(), ent.size() * sizeof(int), hipMemcpyHostToDevice, s);
    rc |= hipMemcpyAsync(d->bi, vi.data(), M * sizeof(int), hipMemcpyHostToDevice, s);
    rc |= hipMemcpyAsync(d->bj, vj.data(), M * sizeof(int), hipMemcpyHostToDevice, s);
    rc |= hipMemcpyAsync(d->bpot, vp.data(), M * sizeof(int), hipMemcpyHostToDevice, s);
    rc |= hipMemcpyAsync(d->bdis, vd.data(), M * sizeof(double), hipMemcpyHostToDevice, s);
    rc |= hipMemsetAsync(d->bact[0], 1, M, s);
    rc |= hipMemsetAsync(d->bact[1], 1, M, s);
  }
  rc |= hipStreamSynchronize(s);
  if (rc) return check_device(rc, "bond upload");
  M_ = M;
  // Both flag buffers start active; the current one follows the parity.
  potentials_dirty_ = true;
  args_dirty_ = true;
  return kErrOk;
}

// MxBoundaryConditions_Init (MxBoundaryConditions.cpp:496-700): the six
// sides together.  An axis is periodic when both its sides are; a lone
// periodic side is refused (the reference warns and makes both periodic,
// check_periodicy).
int Engine::set_boundaries(const int kind_in[6], const float *vel_in, const float *restore, const float *radius) {
  if (!d_) return fail(kErrNull, "engine not initialised");
  if (!kind_in) return reject(kErrNull, "boundary kinds missing");
  // A no-slip side is a velocity wall at rest, as the reference's
  // configuration makes it (MxBoundaryConditions.cpp:415-418, 438-441).
  int kind[6];
  float vel[18];
  for (int side = 0; side < 6; ++side) {
    kind[side] = kind_in[side];
    for (int k = 0; k < 3; ++k) vel[3 * side + k] = vel_in ? vel_in[3 * side + k] : 0.0f;
    if (kind[side] & kBoundaryNoSlip) {
      kind[side] = (kind[side] & ~kBoundaryNoSlip) | kBoundaryVelocity;
      for (int k = 0; k < 3; ++k) vel[3 * side + k] = 0.0f;
    }
  }
  for (int side = 0; side < 6; ++side) {
    const int k = kind[side];
    if (k & ~(kBoundaryVelocity | kBoundaryPeriodic | kBoundaryFreeslip | kBoundaryPotential))
      return reject(kErrRange, "unknown boundary kind bits");
    if ((k & kBoundaryPeriodic) && (k & kBoundaryActive))
      return reject(kErrRange, "a periodic side cannot also be a wall");
  }
  for (int a = 0; a < 3; ++a)
    if (((kind[2 * a] & kBoundaryPeriodic) != 0) != ((kind[2 * a + 1] & kBoundaryPeriodic) != 0))
      return reject(kErrRange, "both sides of an axis must be periodic, or neither");
  for (int side = 0; side < 6; ++side) {
    WallInfo &W = walls_[side];
    W.kind = kind[side];
    for (int k = 0; k < 3; ++k) W.vel[k] = vel[3 * side + k];
    W.restore = restore ? restore[side] : 1.0f;
    W.radius = radius ? radius[side] : 0.0f;
  }
  return upload_walls();
}

int Engine::set_boundary_potential(int side, int type, const Potential *p) {
  if (side < 0 || side > 5) return reject(kErrRange, "boundary side must be 0..5");
  if (type < 0 || type >= (int)types_.size()) return reject(kErrRange, "type index out of range");
  if (p && p->kind == kPotKindTable && p->n < 1) return reject(kErrPotential, "empty potential table");
  for (size_t w = 0; w < wpot_key_.size(); ++w)
    if (wpot_key_[w] == std::make_pair(side, type)) {
      wpot_key_.erase(wpot_key_.begin() + w);
      wpots_.erase(wpots_.begin() + w);
      break;
    }
  if (p) {
    wpot_key_.push_back({side, type});
    wpots_.push_back(*p);
  }
  potentials_dirty_ = true;
  return kErrOk;
}

// Periodic and active-wall bits of every replica, and the side table.
int Engine::upload_walls() {
  DeviceState *d = d_;
  int bits = 0;
  for (int a = 0; a < 3; ++a)
    if ((walls_[2 * a].kind & kBoundaryPeriodic) && (walls_[2 * a + 1].kind & kBoundaryPeriodic))
      bits |= 1 << (kWallPeriodicShift + a);
  for (int side = 0; side < 6; ++side)
    if (walls_[side].kind & kBoundaryActive) bits |= 1 << (kWallActiveShift + side);
  static const float normals[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  std::vector<DevWall> dw(6);
  for (int side = 0; side < 6; ++side) {
    std::memset(&dw[side], 0, sizeof(DevWall));
    dw[side].kind = walls_[side].kind;
    for (int k = 0; k < 3; ++k) {
      dw[side].vel[k] = walls_[side].vel[k];
      dw[side].normal[k] = normals[side][k];
    }
    dw[side].restore = walls_[side].restore;
    dw[side].radius = walls_[side].radius;
  }
  for (auto &D : d->drep) D.wall = bits;
  if (!d->walls)
    if (int st = dalloc(d, &d->walls, 6)) return check_device(st, "hipMalloc");
  hipStream_t s = d->stream;
  int rc = 0;
  rc |= hipMemcpyAsync(d->walls, dw.data(), 6 * sizeof(DevWall), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->rep, d->drep.data(), d->drep.size() * sizeof(DevReplica), hipMemcpyHostToDevice, s);
  rc |= hipStreamSynchronize(s);
  if (rc) return check_device(rc, "boundary upload");
  args_dirty_ = true;
  return kErrOk;
}

int Engine::set_species(int type, int constant_mask) {
  if (type < 0 || type >= (int)types_.size()) return reject(kErrRange, "type index out of range");
  if (constant_mask & ~0xf) return reject(kErrRange, "at most four species per particle");
  types_[type].spec_const = constant_mask;
  potentials_dirty_ = true;
  return kErrOk;
}

int Engine::add_flux(int kind, int ta, int tb, int ia, int ib, float k, float decay, float target) {
  const int nt = (int)types_.size();
  if (kind < 0 || kind > 2) return reject(kErrRange, "flux kind must be 0 (Fick), 1 (secrete) or 2 (uptake)");
  if (ta < 0 || ta >= nt || tb < 0 || tb >= nt) return reject(kErrRange, "type index out of range");
  if (ia < 0 || ia > 3 || ib < 0 || ib > 3) return reject(kErrRange, "species index must be 0..3");
  fluxes_.push_back(FluxInfo{kind, ta, tb, ia, ib, k, decay, target});
  potentials_dirty_ = true;
  if (N_ > 0) return ensure_species();
  return kErrOk;
}

// Species buffers (values and accumulators, ping-pong), zero at first.
int Engine::ensure_species() {
  if (species_on_) return kErrOk;
  if (!d_ || N_ == 0) return fail(kErrNull, "load particles before species");
  DeviceState *d = d_;
  for (int b = 0; b < 2; ++b) {
    if (int st = dalloc(d, &d->spec[b], N_)) return check_device(st, "hipMalloc");
    if (int st = dalloc(d, &d->specq[b], N_)) return check_device(st, "hipMalloc");
  }
  int rc = 0;
  for (int b = 0; b < 2; ++b) {
    rc |= hipMemsetAsync(d->spec[b], 0, (size_t)N_ * sizeof(float4), d->stream);
    rc |= hipMemsetAsync(d->specq[b], 0, (size_t)N_ * sizeof(float4), d->stream);
  }
  rc |= hipStreamSynchronize(d->stream);
  if (rc) return check_device(rc, "species buffers");
  species_on_ = true;
  args_dirty_ = true;
  ++config_gen_;
  return kErrOk;
}

int Engine::set_species_values(const float *values) {
  if (!values) return reject(kErrNull, "species values missing");
  if (int rc = ensure_species()) return rc;
  if (int rc = sync()) return rc;
  DeviceState *d = d_;
  std::vector<int> pid(N_);
  std::vector<float4> v(N_), zero(N_, make_float4(0, 0, 0, 0));
  int rc = hipMemcpy(pid.data(), d->pid[parity_], N_ * sizeof(int), hipMemcpyDeviceToHost);
  if (rc) return check_device(rc, "species upload");
  for (int s = 0; s < N_; ++s) {
    const float *p = values + 4 * (size_t)load_of_gid_[pid[s]];
    v[s] = make_float4(p[0], p[1], p[2], p[3]);
  }
  rc |= hipMemcpy(d->spec[0], v.data(), N_ * sizeof(float4), hipMemcpyHostToDevice);
  rc |= hipMemcpy(d->specq[0], zero.data(), N_ * sizeof(float4), hipMemcpyHostToDevice);
  if (rc) return check_device(rc, "species upload");
  return kErrOk;
}

int Engine::get_species(float *values, float *q) {
  if (!d_ || N_ == 0) return reject(kErrNull, "no particles loaded");
  if (values) std::fill(values, values + 4 * (size_t)N_, 0.0f);
  if (q) std::fill(q, q + 4 * (size_t)N_, 0.0f);
  if (!species_on_) return kErrOk;
  if (int rc = sync()) return rc;
  DeviceState *d = d_;
  std::vector<int> pid(N_);
  std::vector<float4> v(N_), a(N_);
  int rc = hipMemcpy(pid.data(), d->pid[parity_], N_ * sizeof(int), hipMemcpyDeviceToHost);
  rc |= hipMemcpy(v.data(), d->spec[0], N_ * sizeof(float4), hipMemcpyDeviceToHost);
  rc |= hipMemcpy(a.data(), d->specq[0], N_ * sizeof(float4), hipMemcpyDeviceToHost);
  if (rc) return check_device(rc, "species read-back");
  for (int s = 0; s < N_; ++s) {
    const float vv[4] = {v[s].x, v[s].y, v[s].z, v[s].w}, qq[4] = {a[s].x, a[s].y, a[s].z, a[s].w};
    const size_t i = load_of_gid_[pid[s]];
    for (int k = 0; k < 4; ++k) {
      if (values) values[4 * i + k] = vv[k];
      if (q) q[4 * i + k] = qq[k];
    }
  }
  return kErrOk;
}

// Shared checks of a term's particles (gids): loaded, distinct, in one
// replica.
static bool term_ids_ok(const int *ids, int n, const std::vector<int> &prep) {
  for (int a = 0; a < n; ++a) {
    if (ids[a] < 0) return false;
    for (int b = 0; b < a; ++b)
      if (ids[a] == ids[b]) return false;
    if (prep[ids[a]] != prep[ids[0]]) return false;
  }
  return true;
}

static int dedup_pot(std::vector<Potential> &list, std::vector<const Potential *> &seen, const Potential *p) {
  auto it = std::find(seen.begin(), seen.end(), p);
  if (it != seen.end()) return (int)(it - seen.begin());
  seen.push_back(p);
  list.push_back(*p);
  return (int)list.size() - 1;
}

// engine_angle_add (engine.cpp): angle (i, j, k) with j the vertex and a
// potential tabulated in cos(theta).
int Engine::set_angles(int M, const int *i, const int *j, const int *k, const Potential *const *pots) {
  if (!d_ || N_ == 0) return fail(kErrNull, "load particles before angles");
  if (M < 0 || (M > 0 && (!i || !j || !k || !pots))) return reject(kErrNull, "angle arrays missing");
  std::vector<HostTerm> terms(M);
  std::vector<Potential> list;
  std::vector<const Potential *> seen;
  for (int a = 0; a < M; ++a) {
    const int ids[3] = {gid_of(i[a]), gid_of(j[a]), gid_of(k[a])};
    if (!term_ids_ok(ids, 3, prep_))
      return reject(kErrRange, "angle " + std::to_string(a) + " has invalid or repeated particle ids");
    const Potential *p = pots[a];
    if (!p || p->kind != kPotKindTable || p->n < 1)
      return reject(kErrPotential, "angle potential must be a table potential");
    terms[a] = HostTerm{{ids[0], ids[1], ids[2], -1}, dedup_pot(list, seen, p)};
  }
  angles_.swap(terms);
  apots_.swap(list);
  potentials_dirty_ = true;
  return upload_terms();
}

// engine_dihedral_add (engine.cpp): dihedral (i, j, k, l) about the bond
// j-k, with a potential tabulated in cos(phi).
int Engine::set_dihedrals(int M, const int *i, const int *j, const int *k, const int *l,
                          const Potential *const *pots) {
  if (!d_ || N_ == 0) return fail(kErrNull, "load particles before dihedrals");
  if (M < 0 || (M > 0 && (!i || !j || !k || !l || !pots))) return reject(kErrNull, "dihedral arrays missing");
  std::vector<HostTerm> terms(M);
  std::vector<Potential> list;
  std::vector<const Potential *> seen;
  for (int a = 0; a < M; ++a) {
    const int ids[4] = {gid_of(i[a]), gid_of(j[a]), gid_of(k[a]), gid_of(l[a])};
    if (!term_ids_ok(ids, 4, prep_))
      return reject(kErrRange, "dihedral " + std::to_string(a) + " has invalid or repeated particle ids");
    const Potential *p = pots[a];
    if (!p || p->kind != kPotKindTable || p->n < 1)
      return reject(kErrPotential, "dihedral potential must be a table potential");
    terms[a] = HostTerm{{ids[0], ids[1], ids[2], ids[3]}, dedup_pot(list, seen, p)};
  }
  dihedrals_.swap(terms);
  dpots_.swap(list);
  potentials_dirty_ = true;
  return upload_terms();
}

// engine_exclusion_add (engine.cpp): the pair's non-bonded interaction is
// evaluated again and subtracted.
int Engine::set_exclusions(int M, const int *i, const int *j) {
  if (!d_ || N_ == 0) return fail(kErrNull, "load particles before exclusions");
  if (M < 0 || (M > 0 && (!i || !j))) return reject(kErrNull, "exclusion arrays missing");
  std::vector<HostTerm> terms(M);
  for (int a = 0; a < M; ++a) {
    const int ids[2] = {gid_of(i[a]), gid_of(j[a])};
    if (!term_ids_ok(ids, 2, prep_))
      return reject(kErrRange, "exclusion " + std::to_string(a) + " has invalid or repeated particle ids");
    terms[a] = HostTerm{{ids[0], ids[1], -1, -1}, -1};
  }
  exclusions_.swap(terms);
  return upload_terms();
}

// Device lists of the terms: grouped by replica and, within a replica, in
// the reference's evaluation order (exclusions, angles, dihedrals;
// engine_bonded.cpp:1513-1530); the per-particle term lists hold one entry
// per role, in term order.
int Engine::upload_terms() {
  struct Ref {
    int rep, kind, idx;
  };
  std::vector<Ref> order;
  const std::vector<HostTerm> *lists[3] = {&exclusions_, &angles_, &dihedrals_};
  const int kinds[3] = {kTermExclusion, kTermAngle, kTermDihedral};
  const int roles[3] = {2, 3, 4};
  for (int g = 0; g < 3; ++g)
    for (int a = 0; a < (int)lists[g]->size(); ++a) order.push_back(Ref{prep_[(*lists[g])[a].id[1]], g, a});
  std::stable_sort(order.begin(), order.end(), [](const Ref &x, const Ref &y) { return x.rep < y.rep; });
  const int NT = (int)order.size();
  const int R = (int)reps_.size();
  std::vector<int> toff(N_ + 1, 0);
  for (const Ref &r : order) {
    const HostTerm &h = (*lists[r.kind])[r.idx];
    for (int q = 0; q < roles[r.kind]; ++q) toff[h.id[q] + 1]++;
  }
  for (int p = 0; p < N_; ++p) toff[p + 1] += toff[p];
  // Energy blocks: runs of one replica and kind, at most kTermChunk long.
  std::vector<int4> tblk;
  std::vector<int> trep_blk(R + 1, 0);
  for (int t = 0; t < NT;) {
    int e = t;
    while (e < NT && e - t < kTermChunk && order[e].rep == order[t].rep && order[e].kind == order[t].kind) ++e;
    tblk.push_back(make_int4(order[t].rep, kinds[order[t].kind], t, e));
    trep_blk[order[t].rep + 1]++;
    t = e;
  }
  for (int r = 0; r < R; ++r) trep_blk[r + 1] += trep_blk[r];
  std::vector<int> fill(toff.begin(), toff.end() - 1);
  std::vector<DevTerm> dev(NT);
  for (int t = 0; t < NT; ++t) {
    const Ref &r = order[t];
    const HostTerm &h = (*lists[r.kind])[r.idx];
    DevTerm &D = dev[t];
    std::memset(&D, 0, sizeof(D));
    for (int q = 0; q < 4; ++q) {
      D.id[q] = q < roles[r.kind] ? h.id[q] : -1;
      D.pos[q] = q < roles[r.kind] ? fill[h.id[q]]++ : -1;
    }
    D.kind = kinds[r.kind];
    D.pot = r.kind == 1 ? h.pot : (r.kind == 2 ? (int)apots_.size() + h.pot : -1);
    D.tid = r.idx;
  }
  DeviceState *d = d_;
  const size_t entries = toff[N_];
  if (int st = dalloc(d, &d->terms, dev.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->tblk, tblk.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->trep_blk, trep_blk.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->tpart, tblk.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->toff, toff.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->tf, entries)) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->te, dev.size())) return check_device(st, "hipMalloc");
  if (int st = dalloc(d, &d->obs_terms, 3 * (size_t)R)) return check_device(st, "hipMalloc");
  hipStream_t s = d->stream;
  int rc = 0;
  if (NT > 0) {
    rc |= hipMemcpyAsync(d->terms, dev.data(), dev.size() * sizeof(DevTerm), hipMemcpyHostToDevice, s);
    rc |= hipMemsetAsync(d->tf, 0, entries * sizeof(float4), s);
    rc |= hipMemsetAsync(d->te, 0, dev.size() * sizeof(float), s);
  }
  if (!tblk.empty())
    rc |= hipMemcpyAsync(d->tblk, tblk.data(), tblk.size() * sizeof(int4), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->trep_blk, trep_blk.data(), trep_blk.size() * sizeof(int), hipMemcpyHostToDevice, s);
  rc |= hipMemcpyAsync(d->toff, toff.data(), toff.size() * sizeof(int), hipMemcpyHostToDevice, s);
  rc |= hipMemsetAsync(d->obs_terms, 0, 3 * (size_t)R * sizeof(double), s);
  rc |= hipStreamSynchronize(s);
  if (rc) return check_device(rc, "term upload");
  NT_ = NT;
  NTB_ = (int)tblk.size();
  args_dirty_ = true;
  return kErrOk;
}

int Engine::term_energies(int replica, double *out) {
  if (!d_ || !out) return reject(kErrNull, "null argument");
  const int R = (int)reps_.size();
  if (replica < -1 || replica >= R) return reject(kErrRange, "replica index out of range");
  for (int k = 0; k < 3; ++k) out[k] = 0.0;
  if (NT_ == 0) return kErrOk;
  std::vector<double> h(3 * (size_t)R);
  int rc = hipMemcpyAsync(h.data(), d_->obs_terms, h.size() * sizeof(double), hipMemcpyDeviceToHost, d_->stream);
  rc |= hipStreamSynchronize(d_->stream);
  if (rc) return check_device(rc, "term energies");
  for (int r = 0; r < R; ++r)
    if (replica < 0 || r == replica)
      for (int k = 0; k < 3; ++k) out[k] += h[3 * (size_t)r + k];
  return kErrOk;
}

// Kernel arguments of the force kernel for the step parity par (the
// ping-pong buffers of ids, cell offsets and bond flags alternate).
void Engine::fill_step_args(void *out, int par) const {
  const DeviceState *d = d_;
  StepArgs &A = *reinterpret_cast<StepArgs *>(out);
  std::memset(&A, 0, sizeof(A));
  A.x = (decltype(A.x))(d->x[0]);
  A.v = (decltype(A.v))(d->v[0]);
  A.pf = (decltype(A.pf))(d->pf[0]);
  A.pid = (decltype(A.pid))(d->pid[par]);
  A.cell = (decltype(A.cell))(d->cell[0]);
  A.cstart = (decltype(A.cstart))(d->cstart[par]);
  A.x_out = (decltype(A.x_out))(d->x[1]);
  A.v_out = (decltype(A.v_out))(d->v[1]);
  A.pf_out = (decltype(A.pf_out))(d->pf[1]);
  A.f_out = (decltype(A.f_out))(d->f);
  A.cell_out = (decltype(A.cell_out))(d->cell[1]);
  A.rk_p0 = (decltype(A.rk_p0))(d->rk_p0[0]);
  A.rk_ax = (decltype(A.rk_ax))(d->rk_ax[0]);
  A.rk_av = (decltype(A.rk_av))(d->rk_av[0]);
  A.rk_vk = (decltype(A.rk_vk))(d->rk_vk[0]);
  A.rk_p0_out = (decltype(A.rk_p0_out))(d->rk_p0[1]);
  A.rk_ax_out = (decltype(A.rk_ax_out))(d->rk_ax[1]);
  A.rk_av_out = (decltype(A.rk_av_out))(d->rk_av[1]);
  A.rk_vk_out = (decltype(A.rk_vk_out))(d->rk_vk[1]);
  A.rep = (decltype(A.rep))(d->rep);
  A.offs = (decltype(A.offs))(d->offs);
  A.shifts = (decltype(A.shifts))(d->shifts);
  A.tiles = (decltype(A.tiles))(d->tiles);
  A.ntiles = (decltype(A.ntiles))(ntiles_);
  A.pots = (decltype(A.pots))(d->pots);
  A.bpots = (decltype(A.bpots))(d->bpots);
  A.coef = (decltype(A.coef))(d->coef);
  A.potmat = (decltype(A.potmat))(d->potmat);
  A.ntypes = (decltype(A.ntypes))((int)types_.size());
  A.npots = (decltype(A.npots))((int)pots_.size());
  A.types = (decltype(A.types))(d->types);
  A.boff = (decltype(A.boff))(d->boff);
  A.bent = (decltype(A.bent))(d->bent);
  A.bi = (decltype(A.bi))(d->bi);
  A.bj = (decltype(A.bj))(d->bj);
  A.bpot = (decltype(A.bpot))(d->bpot);
  A.bdis = (decltype(A.bdis))(d->bdis);
  A.bact_in = (decltype(A.bact_in))(M_ > 0 ? d->bact[par] : nullptr);
  A.bact_out = (decltype(A.bact_out))(M_ > 0 ? d->bact[par ^ 1] : nullptr);
  A.slot_of = (decltype(A.slot_of))(d->slot_of);
  A.spec = (decltype(A.spec))(d->spec[0]);
  A.specq = (decltype(A.specq))(d->specq[0]);
  A.spec_out = (decltype(A.spec_out))(d->spec[1]);
  A.specq_out = (decltype(A.specq_out))(d->specq[1]);
  A.fluxes = (decltype(A.fluxes))(d->fluxes);
  A.flux_off = (decltype(A.flux_off))(d->flux_off);
  A.dt_d = dt_;
  A.cutoff_d = cutoff_;
  A.walls = (decltype(A.walls))(d->walls);
  A.wallpot = (decltype(A.wallpot))(d->wallpot);
  A.wpots = (decltype(A.wpots))(d->wpots);
  A.toff = (decltype(A.toff))(NT_ > 0 ? d->toff : nullptr);
  A.tf = (decltype(A.tf))(NT_ > 0 ? d->tf : nullptr);
  A.dt = (decltype(A.dt))((float)dt_);
  const float cf = (float)cutoff_;
  A.cut2_pair = (decltype(A.cut2_pair))(cf * cf);
  A.cut2_self = (decltype(A.cut2_self))((float)(cutoff_ * cutoff_));
  // dpd_eval's delta = 1/sqrt(dt) (dpd_eval.hpp:23).
  A.dpd_delta = (decltype(A.dpd_delta))(1.f / std::sqrt(dt_));
  A.seed_lo = (decltype(A.seed_lo))((uint32_t)seed_);
  A.seed_hi = (decltype(A.seed_hi))((uint32_t)(seed_ >> 32));
  A.dpd_noise = (decltype(A.dpd_noise))(dpd_noise_);
  A.time = (decltype(A.time))(d->time);
  A.err = (decltype(A.err))(d->err);
  A.partials = (decltype(A.partials))(d->partials);
  A.cnt = (decltype(A.cnt))(d->cnt);
  A.inc = (decltype(A.inc))(d->inc);
  A.ncount = (decltype(A.ncount))(d->ncount);
  A.random = (decltype(A.random))(any_sb_ ? 1 : 0);
  A.with_pf = (decltype(A.with_pf))(any_random_ ? 1 : 0);
  A.type_ke = (decltype(A.type_ke))(d->type_ke);
  A.cutoff = (decltype(A.cutoff))((float)cutoff_);
  A.nd_norm = (decltype(A.nd_norm))((float)(1.0 / (M_PI * cutoff_ * cutoff_ * cutoff_)));
}

// Uploads both parities of the force-kernel arguments (device resident, so
// the kernel reads them through one pointer).
int Engine::refresh_args() {
  DeviceState *d = d_;
  if (!d->dargs)
    if (int st = dalloc(d, &d->dargs, 2 * sizeof(StepArgs))) return check_device(st, "hipMalloc");
  StepArgs h[2];
  fill_step_args(&h[0], 0);
  fill_step_args(&h[1], 1);
  int rc = hipMemcpyAsync(d->dargs, h, sizeof(h), hipMemcpyHostToDevice, d->stream);
  rc |= hipStreamSynchronize(d->stream);
  if (rc) return check_device(rc, "argument upload");
  args_dirty_ = false;
  ++config_gen_;  // captured graphs hold the old arguments
  return kErrOk;
}

int Engine::launch_step() {
  if (integrator_ == kIntegratorRK4) {
    // engine_advance_runge_kutta_4: four force evaluations, each followed
    // by the cell-list rebuild (the reference's space_shuffle_local in the
    // next engine_force); the time advances once.
    for (int stage = 0; stage < 4; ++stage)
      if (int st = launch_substep(stage, stage == 3)) return st;
    return 0;
  }
  return launch_substep(-1, 1);
}

int Engine::launch_substep(int rk_stage, int tick) {
  DeviceState *d = d_;
  const int par = parity_;
  int mode = 0;
  if (any_table_) mode |= kModeTable;
  if (any_dpd_) mode |= kModeDPD;
  if (virial_) mode |= kModeVirial;
  if (M_ > 0 || NT_ > 0) mode |= kModeBonds;
  if (species_on_) mode |= kModeFlux;
  hipStream_t s = d->stream;
  if (any_berendsen_) {
    hipLaunchKernelGGL(k_type_ke, dim3(kKeBlocks), dim3(256), 0, s, d->x[0], d->v[0], d->types, N_,
                       d->type_ke_part);
    hipLaunchKernelGGL(k_type_ke_final, dim3(1), dim3(64), 0, s, d->type_ke_part, d->type_ke);
  }
  if (NT_ > 0) {
    TermArgs T;
    T.terms = d->terms;
    T.nterms = NT_;
    T.x = d->x[0];
    T.cell = d->cell[0];
    T.cell_rep = d->cell_rep;
    T.slot_of = d->slot_of;
    T.rep = d->rep;
    T.pots = d->pots;
    T.tpots = d->tpots;
    T.coef = d->coef;
    T.potmat = d->potmat;
    T.ntypes = (int)types_.size();
    T.cut2 = (float)(cutoff_ * cutoff_);
    T.time = d->time;
    T.seed_lo = (uint32_t)seed_;
    T.seed_hi = (uint32_t)(seed_ >> 32);
    T.tf = d->tf;
    T.te = d->te;
    hipLaunchKernelGGL(k_terms, dim3((NT_ + 255) / 256), dim3(256), 0, s, T);
    hipLaunchKernelGGL(k_term_partial, dim3(NTB_), dim3(256), 0, s, d->te, d->tblk, d->tpart);
    hipLaunchKernelGGL(k_term_energy, dim3(reps_.size()), dim3(256), 0, s, d->tblk, d->tpart, d->trep_blk,
                       d->obs_terms);
  }
  const bool rec = timing_ && timed_steps_ < (int)d->kev.size() / 2;
  if (rec) hipEventRecord(d->kev[2 * timed_steps_], s);
  BrickArgs Bk;
  Bk.bricks = d->bricks;
  Bk.nbricks = nbricks_;
  for (int k = 0; k < 3; ++k) Bk.bdim[k] = bdim_[k];
  Bk.pmax = pmax_;
  Bk.cap = cap_;
  // A mid-scan drain leaves two thirds of the capacity in each list (of
  // 1/2, 2/3, 3/4 and 5/6, the fastest on the 1M DPD, LJ and lattice steps).
  Bk.drain = cap_ * 2 / 3;
  Bk.hmax = hmax_;
  Bk.hdmax = hdmax_;
  Bk.noff = noff_;
  Bk.ovf_list = d->ovf_list;
  Bk.ovf_count = d->ovf_count;
  Bk.rk_stage = rk_stage;
  const size_t tabs = pair_tab_bytes((int)pots_.size(), (int)types_.size());
  const size_t lds =
      brick_lds_bytes(pmax_, hmax_, hdmax_, noff_, cap_, (mode & kModeDPD) != 0, (mode & kModeFlux) != 0, tabs);
  const StepLauncher launch = (waves_ == 4 && kStepLaunchers4[mode]) ? kStepLaunchers4[mode] : kStepLaunchers[mode];
  launch(reinterpret_cast<const StepArgs *>(d->dargs) + par, Bk, lds, tabs, s);
  if (rec) hipEventRecord(d->kev[2 * timed_steps_ + 1], s);
  if (rec) ++timed_steps_;

  // Cell-list rebuild: x, v, pf, cell go from buffer 1 back to buffer 0;
  // ids and cell offsets go from parity par to par ^ 1.
  RebuildArgs B;
  B.cstart = d->cstart[par];
  B.cstart_new = d->cstart[par ^ 1];
  B.cell_out = d->cell[1];
  B.inc = d->inc;
  B.cursor = d->cursor;
  B.nslots = N_;
  B.ncells = ncells_;
  B.x_in = d->x[1];
  B.v_in = d->v[1];
  B.pf_in = d->pf[1];
  B.pid_in = d->pid[par];
  B.x = d->x[0];
  B.v = d->v[0];
  B.pf = d->pf[0];
  B.pid = d->pid[par ^ 1];
  B.cell = d->cell[0];
  B.slot_of = d->slot_of;
  B.with_slots = (M_ > 0 || NT_ > 0) ? 1 : 0;
  B.with_pf = any_random_ ? 1 : 0;
  B.with_rk = rk_stage >= 0 && rk_stage < 3 ? 1 : 0;
  B.with_spec = species_on_ ? 1 : 0;
  for (int k = 0; k < 2; ++k) {
    B.spec_in[k] = k == 0 ? d->spec[1] : d->specq[1];
    B.spec[k] = k == 0 ? d->spec[0] : d->specq[0];
  }
  for (int k = 0; k < 4; ++k) {
    float4 *const *buf = k == 0 ? d->rk_p0 : k == 1 ? d->rk_ax : k == 2 ? d->rk_av : d->rk_vk;
    B.rk_in[k] = buf[1];
    B.rk[k] = buf[0];
  }
  const int cg = (ncells_ + 255) / 256;
  const int pg = (N_ + 255) / 256;
  const int nb = (ncells_ + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(kScanBlock), 0, s, d->cnt, d->inc, d->cstart[par ^ 1], d->bsum,
                     ncells_);
  hipLaunchKernelGGL(k_scan_add, dim3(cg), dim3(256), 0, s, d->cstart[par ^ 1], d->bsum, nb, ncells_, d->time,
                     d->err, d->ovf_count, tick);
  FinalArgs F;
  F.partials = d->partials;
  F.rep_tiles = d->rep_bricks;
  F.nrep = (int)reps_.size();
  F.nred = virial_ ? kNRed : 3;
  F.obs = d->obs;
  F.err = d->err;
  hipLaunchKernelGGL(k_place, dim3(F.nrep + pg), dim3(256), 0, s, B, F);
  hipLaunchKernelGGL(k_order_incoming, dim3(cg), dim3(256), 0, s, B);
  parity_ ^= 1;
  pid_prev_valid_ = true;
  return (int)hipGetLastError();
}

// Captures two steps starting at parity par into a graph (kernel nodes only).
int Engine::ensure_graph(int par) {
  DeviceState *d = d_;
  if (d->gexec[par] && graph_gen_[par] == config_gen_) return kErrOk;
  if (d->gexec[par]) {
    hipGraphExecDestroy(d->gexec[par]);
    d->gexec[par] = nullptr;
  }
  const int saved = parity_;
  parity_ = par;
  hipGraph_t g = nullptr;
  if (int st = hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal))
    return check_device(st, "hipStreamBeginCapture");
  int st1 = launch_step();
  int st2 = launch_step();
  int st3 = hipStreamEndCapture(d->stream, &g);
  parity_ = saved;
  if (st1 || st2 || st3) {
    if (g) hipGraphDestroy(g);
    return check_device(st1 ? st1 : (st2 ? st2 : st3), "graph capture");
  }
  int st4 = hipGraphInstantiate(&d->gexec[par], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (st4) {
    d->gexec[par] = nullptr;
    return check_device(st4, "hipGraphInstantiate");
  }
  graph_gen_[par] = config_gen_;
  return kErrOk;
}

// Times the force kernel over a list of brick shapes and neighbour-list
// capacities, steps_per steps each, and keeps the fastest.  The steps are
// real steps of the simulation: every layout gives the same per-particle
// results (each particle's neighbours are visited in the same order), so
// tuning does not change the trajectory.
int Engine::autotune(int steps_per, int *taken) {
  if (!d_ || N_ == 0) return fail(kErrNull, "no particles loaded");
  if (steps_per < 1) return reject(kErrRange, "autotune needs at least one step per candidate");
  // Shapes whose home-particle count fills the 256-thread block at common
  // densities (the measured best for DPD at density 3, [5,5,3], and for LJ
  // at density 10, [6,4,1], among them), and the power-of-two series for
  // dense or small systems.
  static const int shapes[][3] = {{6, 4, 4}, {5, 5, 4}, {5, 4, 4}, {5, 5, 3}, {6, 4, 3}, {7, 5, 2}, {6, 6, 2},
                                  {4, 4, 4}, {4, 4, 3}, {4, 3, 3}, {3, 3, 3}, {4, 4, 2}, {6, 4, 1}, {6, 2, 2},
                                  {5, 2, 2}, {4, 2, 2}, {3, 3, 2}, {2, 2, 2}, {2, 2, 1}, {2, 1, 1}, {1, 1, 1}};
  static const int caps[] = {8, 10, 12, 14, 16, 20, 24};
  const bool was_timing = timing_;
  int best[4] = {bdim_[0], bdim_[1], bdim_[2], cap_};
  int best_waves = waves_;
  double best_ms = 1e30;
  int steps = 0;
  timing_ = true;
  tune_log_.clear();
  for (const auto &sh : shapes) {
    for (int c : caps) {
      if (set_brick(sh[0], sh[1], sh[2], c, pmax_req_) != kErrOk) continue;  // does not fit: skip
      // One untimed step after the change of layout, then the timed ones.
      timing_ = false;
      int rc = step(1);
      timing_ = true;
      if (rc == kErrOk) rc = step(steps_per);
      steps += steps_per + 1;
      if (rc != kErrOk) {
        timing_ = was_timing;
        if (taken) *taken = steps;
        return rc;
      }
      tune_log_.push_back({sh[0], sh[1], sh[2], c, waves_, force_ms_});
      if (force_ms_ > 0 && force_ms_ < best_ms) {
        best_ms = force_ms_;
        best[0] = sh[0];
        best[1] = sh[1];
        best[2] = sh[2];
        best[3] = c;
        best_waves = waves_;
      }
    }
  }
  // Bonded modes: the four fastest layouts again at four waves per SIMD.
  if (M_ > 0 || NT_ > 0) {
    std::vector<TuneEntry> top = tune_log_;
    std::sort(top.begin(), top.end(), [](const TuneEntry &a, const TuneEntry &b) {
      return (a.force_ms > 0 ? a.force_ms : 1e30) < (b.force_ms > 0 ? b.force_ms : 1e30);
    });
    top.resize(std::min<size_t>(top.size(), 4));
    const int saved = waves_;
    waves_ = 4;
    ++config_gen_;
    for (const TuneEntry &t : top) {
      if (set_brick(t.bx, t.by, t.bz, t.cap, pmax_req_) != kErrOk) continue;
      timing_ = false;
      int rc = step(1);
      timing_ = true;
      if (rc == kErrOk) rc = step(steps_per);
      steps += steps_per + 1;
      if (rc != kErrOk) {
        waves_ = saved;
        ++config_gen_;
        timing_ = was_timing;
        if (taken) *taken = steps;
        return rc;
      }
      tune_log_.push_back({t.bx, t.by, t.bz, t.cap, 4, force_ms_});
      if (force_ms_ > 0 && force_ms_ < best_ms) {
        best_ms = force_ms_;
        best[0] = t.bx;
        best[1] = t.by;
        best[2] = t.bz;
        best[3] = t.cap;
        best_waves = 4;
      }
    }
    waves_ = saved;
  }
  // A fresh load relaxes over the sweep's first steps, which favours the
  // layouts timed late, and five steps are short against the clock's
  // wanderings: the four fastest are timed again in two passes of four
  // times as many steps, and the layout with the best time in either pass
  // is kept.
  {
    std::vector<TuneEntry> top = tune_log_;
    std::sort(top.begin(), top.end(), [](const TuneEntry &a, const TuneEntry &b) {
      return (a.force_ms > 0 ? a.force_ms : 1e30) < (b.force_ms > 0 ? b.force_ms : 1e30);
    });
    top.resize(std::min<size_t>(top.size(), 4));
    const int saved = waves_;
    double round_ms = 1e30;
    for (int pass = 0; pass < 2; ++pass) {
      for (const TuneEntry &t : top) {
        if (!(t.force_ms > 0)) continue;
        if (t.waves != waves_) {
          waves_ = t.waves;
          ++config_gen_;
        }
        if (set_brick(t.bx, t.by, t.bz, t.cap, pmax_req_) != kErrOk) continue;
        timing_ = false;
        int rc = step(1);
        timing_ = true;
        if (rc == kErrOk) rc = step(4 * steps_per);
        steps += 4 * steps_per + 1;
        if (rc != kErrOk) {
          if (waves_ != saved) {
            waves_ = saved;
            ++config_gen_;
          }
          timing_ = was_timing;
          if (taken) *taken = steps;
          return rc;
        }
        tune_log_.push_back({t.bx, t.by, t.bz, t.cap, t.waves, force_ms_});
        if (force_ms_ > 0 && force_ms_ < round_ms) {
          round_ms = force_ms_;
          best[0] = t.bx;
          best[1] = t.by;
          best[2] = t.bz;
          best[3] = t.cap;
          best_waves = t.waves;
        }
      }
    }
    if (waves_ != saved) {
      waves_ = saved;
      ++config_gen_;
    }
  }
  timing_ = was_timing;
  if (taken) *taken = steps;
  msg_.clear();
  if (best_waves != waves_) {
    waves_ = best_waves;
    ++config_gen_;
  }
  return set_brick(best[0], best[1], best[2], best[3], pmax_req_);
}

// Waves per SIMD of the force kernel (3, or 4 for bonded modes; modes
// without bonds or terms always run at three).
int Engine::set_waves(int waves) {
  if (waves != 3 && waves != 4) return reject(kErrRange, "waves per SIMD must be 3 or 4");
  if (waves != waves_) {
    waves_ = waves;
    ++config_gen_;  // captured graphs hold the old kernel
  }
  return kErrOk;
}

// engine.integrator (EngineIntegrator FORWARD_EULER / RUNGE_KUTTA_4, chosen
// by the simulator's integrator argument, MxSimulator.cpp:218-231).
int Engine::set_integrator(int kind) {
  if (kind != kIntegratorEuler && kind != kIntegratorRK4) return reject(kErrRange, "unknown integrator");
  integrator_ = kind;
  ++config_gen_;
  if (d_ && N_ > 0) return ensure_rk_buffers();
  return kErrOk;
}

int Engine::ensure_rk_buffers() {
  if (integrator_ != kIntegratorRK4 || !d_ || N_ == 0) return kErrOk;
  DeviceState *d = d_;
  for (int b = 0; b < 2; ++b) {
    if (int st = dalloc(d, &d->rk_p0[b], N_)) return check_device(st, "hipMalloc");
    if (int st = dalloc(d, &d->rk_ax[b], N_)) return check_device(st, "hipMalloc");
    if (int st = dalloc(d, &d->rk_av[b], N_)) return check_device(st, "hipMalloc");
    if (int st = dalloc(d, &d->rk_vk[b], N_)) return check_device(st, "hipMalloc");
  }
  args_dirty_ = true;
  return kErrOk;
}

int Engine::set_graphs(int on) {
  use_graphs_ = on != 0;
  return kErrOk;
}

int Engine::step(int nsteps) {
  if (!d_ || N_ == 0) return fail(kErrNull, "no particles loaded");
  if (err_ != 0) return err_;
  if (potentials_dirty_)
    if (int rc = upload_potentials()) return rc;
  // Potentials bound after the load can change the LDS layout (velocities
  // are staged for DPD only).
  if (brick_dpd_ != any_dpd_ || brick_flux_ != species_on_ ||
      brick_tabs_ != pair_tab_bytes((int)pots_.size(), (int)types_.size()))
    if (int rc = set_brick(bdim_[0], bdim_[1], bdim_[2], cap_, pmax_req_)) return rc;
  if (args_dirty_)
    if (int rc = refresh_args()) return rc;
  // Without bonds or terms the rebuild skips slot_of (a scattered write per
  // particle); rebuild it before bonds or terms use it again.
  if (M_ > 0 || NT_ > 0) {
    if (!slots_valid_) {
      hipLaunchKernelGGL(k_slots, dim3((N_ + 255) / 256), dim3(256), 0, d_->stream, d_->pid[parity_], N_,
                         d_->slot_of);
      if (int st = hipGetLastError()) return check_device(st, "kernel launch");
      slots_valid_ = true;
    }
  } else if (nsteps > 0) {
    slots_valid_ = false;
  }
  if (timing_) {
    // Two events per force launch, for up to 4096 launches per call.
    const int per = integrator_ == kIntegratorRK4 ? 4 : 1;
    const size_t want = 2 * (size_t)std::min(nsteps * per, 4096);
    while (d_->kev.size() < want) {
      hipEvent_t e;
      if (int rc = check_device(hipEventCreate(&e), "hipEventCreate")) return rc;
      d_->kev.push_back(e);
    }
    hipEventRecord(d_->ev[0], d_->stream);
  }
  timed_steps_ = 0;
  int k = 0;
  if (use_graphs_ && !timing_) {
    // Pairs of steps replay a captured graph (the ping-pong buffers return
    // to the same parity after two steps).
    for (; k + 2 <= nsteps; k += 2) {
      if (int rc = ensure_graph(parity_)) return rc;
      if (int st = hipGraphLaunch(d_->gexec[parity_], d_->stream)) return check_device(st, "hipGraphLaunch");
      pid_prev_valid_ = true;
    }
  }
  for (; k < nsteps; ++k) {
    if (int st = launch_step()) return check_device(st, "kernel launch");
  }
  if (timing_) hipEventRecord(d_->ev[2], d_->stream);
  int herr[4];
  unsigned htime = 0;
  int rc = hipMemcpyAsync(herr, d_->err, sizeof(herr), hipMemcpyDeviceToHost, d_->stream);
  rc |= hipMemcpyAsync(&htime, d_->time, sizeof(htime), hipMemcpyDeviceToHost, d_->stream);
  rc |= hipStreamSynchronize(d_->stream);
  if (rc) return check_device(rc, "step");
  time_ = htime;
  if (timing_ && nsteps > 0) {
    double sum = 0;
    for (int k = 0; k < timed_steps_; ++k) {
      float ms = 0;
      hipEventElapsedTime(&ms, d_->kev[2 * k], d_->kev[2 * k + 1]);
      sum += ms;
    }
    force_ms_ = timed_steps_ ? sum / timed_steps_ : -1;
    float ms = 0;
    hipEventElapsedTime(&ms, d_->ev[0], d_->ev[2]);
    step_ms_ = ms / nsteps;
  }
  if (herr[0] != 0) {
    const int g = herr[1];
    const int id = (g >= 0 && g < (int)pid_of_gid_.size()) ? pid_of_gid_[g] : g;
    return fail(herr[0], "particle " + std::to_string(id) + " moved too fast (more than one cell in a step)");
  }
  return kErrOk;
}

int Engine::sync() {
  if (!d_) return kErrOk;
  return check_device(hipStreamSynchronize(d_->stream), "sync");
}

int Engine::set_timing(int on) {
  timing_ = on != 0;
  return kErrOk;
}

int Engine::set_virial(int on) {
  virial_ = on != 0;
  ++config_gen_;
  return kErrOk;
}

int Engine::get_state(float *xlocal, int *cell, float *v, float *f, float *pf, float *nd) {
  if (!d_ || N_ == 0) return fail(kErrNull, "no particles loaded");
  DeviceState *d = d_;
  std::vector<float4> hx(N_), hv(N_), hf(N_), hpf(N_);
  std::vector<int> hpid(N_), hcell(N_), hpidf(N_);
  hipStream_t s = d->stream;
  int rc = 0;
  rc |= hipMemcpyAsync(hx.data(), d->x[0], N_ * sizeof(float4), hipMemcpyDeviceToHost, s);
  rc |= hipMemcpyAsync(hv.data(), d->v[0], N_ * sizeof(float4), hipMemcpyDeviceToHost, s);
  rc |= hipMemcpyAsync(hf.data(), d->f, N_ * sizeof(float4), hipMemcpyDeviceToHost, s);
  rc |= hipMemcpyAsync(hpf.data(), d->pf[0], N_ * sizeof(float4), hipMemcpyDeviceToHost, s);
  rc |= hipMemcpyAsync(hpid.data(), d->pid[parity_], N_ * sizeof(int), hipMemcpyDeviceToHost, s);
  rc |= hipMemcpyAsync(hpidf.data(), d->pid[parity_ ^ 1], N_ * sizeof(int), hipMemcpyDeviceToHost, s);
  rc |= hipMemcpyAsync(hcell.data(), d->cell[0], N_ * sizeof(int), hipMemcpyDeviceToHost, s);
  rc |= hipStreamSynchronize(s);
  if (rc) return check_device(rc, "state download");
  for (int k = 0; k < N_; ++k) {
    const size_t i = load_of_gid_[hpid[k]];
    if (xlocal) {
      xlocal[3 * i] = hx[k].x;
      xlocal[3 * i + 1] = hx[k].y;
      xlocal[3 * i + 2] = hx[k].z;
    }
    if (v) {
      v[3 * i] = hv[k].x;
      v[3 * i + 1] = hv[k].y;
      v[3 * i + 2] = hv[k].z;
    }
    if (pf) {
      pf[3 * i] = hpf[k].x;
      pf[3 * i + 1] = hpf[k].y;
      pf[3 * i + 2] = hpf[k].z;
    }
    if (cell) cell[i] = hcell[k];
  }
  if (f || nd) {
    // Forces (and the number density in w) are stored in the slot order of
    // the step that computed them.
    const std::vector<int> &order = pid_prev_valid_ ? hpidf : hpid;
    for (int k = 0; k < N_; ++k) {
      const size_t i = load_of_gid_[order[k]];
      if (f) {
        f[3 * i] = hf[k].x;
        f[3 * i + 1] = hf[k].y;
        f[3 * i + 2] = hf[k].z;
      }
      if (nd) nd[i] = hf[k].w;
    }
  }
  return kErrOk;
}

int Engine::get_global(double *x, double *v) {
  std::vector<float> xl(3 * (size_t)N_), vv(3 * (size_t)N_);
  std::vector<int> cell(N_);
  if (int rc = get_state(xl.data(), cell.data(), vv.data(), nullptr, nullptr, nullptr)) return rc;
  for (int i = 0; i < N_; ++i) {
    const int c = cell[i];
    const HostReplica *R = nullptr;
    for (const auto &r : reps_)
      if (c >= r.cell_base && c < r.cell_base + r.ncells) R = &r;
    const int lc = c - R->cell_base;
    const int l[3] = {lc / (R->cdim[1] * R->cdim[2]), (lc / R->cdim[2]) % R->cdim[1], lc % R->cdim[2]};
    for (int k = 0; k < 3; ++k) {
      if (x) x[3 * (size_t)i + k] = R->corg[k][l[k]] + xl[3 * (size_t)i + k];
      if (v) v[3 * (size_t)i + k] = vv[3 * (size_t)i + k];
    }
  }
  return kErrOk;
}

int Engine::neighbour_counts(int *counts) {
  if (!d_ || N_ == 0) return fail(kErrNull, "no particles loaded");
  if (potentials_dirty_)
    if (int rc = upload_potentials()) return rc;
  DeviceState *d = d_;
  StepArgs A;
  fill_step_args(&A, parity_);
  hipLaunchKernelGGL(k_ncount, dim3(ntiles_), dim3(kTile), 0, d->stream, A);
  std::vector<int> by_gid(N_);
  int rc = hipMemcpyAsync(by_gid.data(), d->ncount, N_ * sizeof(int), hipMemcpyDeviceToHost, d->stream);
  rc |= hipStreamSynchronize(d->stream);
  if (rc) return check_device(rc, "neighbour counts");
  for (int g = 0; g < N_; ++g) counts[load_of_gid_[g]] = by_gid[g];
  return kErrOk;
}

int Engine::observables(int replica, double *out) {
  if (!d_) return fail(kErrNull, "engine not initialised");
  const int R = (int)reps_.size();
  std::vector<double> h((size_t)R * kNRed);
  int rc = hipMemcpyAsync(h.data(), d_->obs, h.size() * sizeof(double), hipMemcpyDeviceToHost, d_->stream);
  rc |= hipStreamSynchronize(d_->stream);
  if (rc) return check_device(rc, "observables");
  if (replica >= 0) {
    if (replica >= R) return fail(kErrRange, "replica index out of range");
    for (int k = 0; k < kNRed; ++k) out[k] = h[(size_t)replica * kNRed + k];
  } else {
    for (int k = 0; k < kNRed; ++k) {
      double s = 0;
      for (int r = 0; r < R; ++r) s += h[(size_t)r * kNRed + k];
      out[k] = s;
    }
  }
  return kErrOk;
}

int Engine::bond_active(uint8_t *active) {
  if (M_ == 0) return kErrOk;
  int rc = hipMemcpyAsync(active, d_->bact[parity_], M_, hipMemcpyDeviceToHost, d_->stream);
  rc |= hipStreamSynchronize(d_->stream);
  return check_device(rc, "bond flags");
}

int Engine::geometry(int replica, int cdim[3], double h[3]) {
  if (replica < 0 || replica >= (int)reps_.size()) return fail(kErrRange, "replica index out of range");
  for (int k = 0; k < 3; ++k) {
    cdim[k] = reps_[replica].cdim[k];
    h[k] = reps_[replica].h[k];
  }
  return kErrOk;
}

}  // namespace mxg
