// ORACLE — test infrastructure only.
//
// A CPU restatement of mdcore's force-and-advance step (engine_step) used as
// the checker for the HIP engine and as the CPU baseline.  Nothing in the
// product links or calls this file; only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load the library built from it.
//
// What it restates, with the reference lines it follows (paths relative to
// src/mdcore/src of the reference):
//   space geometry, cell origins, task graph ........ space.cpp:626-978
//   space_getsid (min-image shift, flip) ............ space.cpp:95-130
//   sort-direction tables ........................... space_cell.cpp:54-99
//   space_addpart (cell index, local coordinate) .... space.cpp:397-477
//   engine_kinetic_energy ........................... engine.cpp:1698-1719
//   space_prepare (force reset) ..................... space.cpp:202-242
//   space_shuffle_local / welcome ................... space.cpp:333-393,
//                                                     space_cell.cpp:224-392
//   single-runner task order ........................ queue.cpp:84-246,
//                                                     runner.cpp:244-418
//   runner_dosort + runner_sort_descending .......... runner_dosort.cpp:84-146,
//                                                     runner.cpp:102-165
//   runner_dopair (sorted, pruned cell pair) ........ runner_dopair.cpp:91-340
//   runner_doself (in-cell pairs, single body) ...... runner_dopair.cpp:446-682
//   potential_eval / potential_eval_ex .............. potential_eval.hpp:67-161
//   potential_eval_super_ex / dpd_eval .............. dpd_eval.hpp:21-169
//   get_potential ................................... potential_eval.hpp:1543
//   bond_eval (periodic shift, clamp, dissociation).. bond.cpp:99-295
//   exclusion_eval (pair potential subtracted) ...... exclusion.cpp:83-260
//   angle_eval (cos theta, potential_eval_r) ........ angle.cpp:92-354
//   dihedral_eval (cos phi, potential_eval_r) ....... dihedral.cpp:90-330
//   bonded evaluation order ......................... engine_bonded.cpp:1510-1538
//   boundary conditions: cell flags, pairs ........... space.cpp:700-723, 827-888
//   boundary_eval (wall potentials) ................. boundary_eval.hpp:124-221
//   dpd_boundary_eval ............................... dpd_eval.hpp:71-106
//   apply_update_pos_vel (wall reflection) .......... boundary_eval.hpp:30-119
//   flux_eval_ex (species transport) ................ flux_eval.hpp:17-81
//   MxFluxes_Integrate (after each cell's advance) .. Flux.cpp:261-284
//   cell_advance_forward_euler ...................... engine_advance.cpp:126-239
//   engine_advance_forward_euler .................... engine_advance.cpp:248-422
//   engine_advance_runge_kutta_4 .................... engine_advance.cpp:443-698
//   gaussian_force (persistent random force) ........ MxForce.cpp:370-385
//   berendsen_force, constant_force, friction_force . MxForce.cpp:330-367
//   per-type kinetic energy ......................... engine.cpp:1698-1719
//
// Deviations, each stated where it happens:
//   * gaussian_force draws from Philox4x32-10 instead of carbon's CRandom
//     (absent from the reference tree), keyed by (particle id, time);
//   * the optional counter-based DPD noise mode is this project's own; the
//     default DPD mode is the reference's deterministic sigma term;
//   * a right or straight angle (cos theta exactly 0 or -1) takes its random
//     plane from Philox4x32-10 keyed by (angle index, time) instead of
//     CRandom;
//   * dihedral_eval's gradient of cos(phi) is the reference's generated
//     expression; here it is derived from the plane normals (same
//     mathematics, different rounding);
//   * an exclusion whose distance lies beyond its table (the reference reads
//     past the table) or whose pair potential is DPD (no table) contributes
//     nothing.
//
// Two precisions: the faithful FP32 path (every float/double step as the
// reference), and an FP64 evaluation of the same pair set and tables
// (forces64) that states the FP32 path's rounding error.
//
// Build with -O3 -mavx -msse4.2 -ffp-contract=off (no FMA), as the reference
// is built (CMakeLists.txt:79-104).
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <random>
#include <set>
#include <thread>
#include <vector>

#include <immintrin.h>

#include "../mechanica_amd/csrc/mxg_philox.h"

namespace {

constexpr int kChunk = 8;

enum : uint32_t { kFlagScaled = 1u << 11, kFlagShifted = 1u << 12 };
enum : uint32_t { kKindTable = 0, kKindDPD = 1 };
// PARTICLE_FROZEN_X/Y/Z (MxParticle.h:58-60).
enum : uint16_t { kFrozenX = 1 << 3, kFrozenY = 1 << 4, kFrozenZ = 1 << 5 };
enum { kNewtonian = 0, kOverdamped = 1 };
enum { kTaskSort = 0, kTaskSelf = 1, kTaskPair = 2 };

// BoundaryConditionKind (MxBoundaryConditions.hpp:15-23).
enum { kBcVelocity = 1, kBcPeriodic = 2, kBcFreeslip = 4, kBcPotential = 8, kBcActive = 1 | 4 | 8 };
constexpr int kErrToofast = -29;
constexpr int kErrRange = -6;

// Sort-direction tables (space_cell.cpp:54-99): the ID of the sorted list
// for each of the 27 shift sign triples, the unit direction of each list, and
// whether the pair has to be flipped so the shift points "forward".
const char kSortlistID[27] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 0,
                              12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0};
const char kFlip[27] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0,
                        0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

struct SortDirs {
  float n[13 * 3];
  SortDirs() {
    // The 13 forward directions in list order: sign triples 26, 25, ..., 14.
    for (int id = 0; id < 13; ++id) {
      int s27 = 26 - id;
      int t[3] = {s27 / 9 - 1, (s27 / 3) % 3 - 1, s27 % 3 - 1};
      double len = std::sqrt((double)(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
      for (int k = 0; k < 3; ++k) n[3 * id + k] = (float)(t[k] / len);
    }
  }
};
const SortDirs &sortdirs() {
  static const SortDirs d;
  return d;
}

struct Pot {
  uint32_t kind = kKindTable, flags = 0;
  float alpha[4] = {0, 0, 0, 0};
  std::vector<float> c;
  float r0_plusone = 0, a = 0, b = 0;
  int n = 0;
  float dpd_alpha = 0, dpd_gamma = 0, dpd_sigma = 0;
};

struct Type {
  double mass = 1, imass = 1;
  float radius = 1;
  int dynamics = kNewtonian;
  // single-body persistent random force (gaussian_force)
  bool has_random = false;  // persistent random force (gaussian or friction)
  float rnd_mean = 0, rnd_std = 0;
  unsigned rnd_steps = 1;
  // single-body force (psbs[type]): 0 none, 1 gaussian, 2 berendsen,
  // 3 constant, 4 friction
  int sb = 0;
  float sb_coef = 0;  // berendsen itau, friction coef
  float cf[3] = {0, 0, 0};
  double target_energy = 0, kinetic_energy = 0;
  int spec_const = 0;  // bit k: species k is constant (not integrated)
};

// One flux (MxFlux entry): kind 0 Fick, 1 secrete, 2 uptake, between
// species ia of type ta and ib of type tb.
struct FluxEntry {
  int kind, ta, tb, ia, ib;
  float coef, decay, target;
};

struct Part {
  float f[4];
  float v[4];
  float x[4];
  float pf[4];  // persistent random force
  float sv[4];  // species values (CStateVector fvec), up to four
  float sq[4];  // species flux accumulators (CStateVector q)
  // Runge-Kutta stage state (MxParticle p0, v0, xk, vk: MxParticle.h:150-153)
  float p0[3], v0[3], xk[4][3], vk[4][3];
  double imass;
  float radius;
  double mass;
  int id;
  int16_t type;
  uint16_t flags;
  int ncount;  // in-cutoff pairs seen during the last force evaluation
};

struct Cell {
  int loc[3];
  int walls = 0;  // cell_active_* as bit i for side i (left, right, front, back, bottom, top)
  double origin[3];
  std::vector<Part> parts;
  std::vector<Part> incoming;
  std::vector<uint32_t> sortlist;
  double epot = 0;
  int sort_flags = 0;
  std::mt19937 gen;
  std::normal_distribution<float> gauss{0.f, 1.f};
  std::atomic<char> taboo{0};
  std::mutex mtx;
};

struct Task {
  int type, i, j, sid;
  std::vector<int> unlock;
  int wait = 0;
};

struct Bond {
  int i, j, pot;
  double dissociation;
  bool active;
};

struct Angle {
  int i, j, k, pot;
};

struct Dihedral {
  int i, j, k, l, pot;
};

struct Exclusion {
  int i, j;
};

struct Oracle {
  // geometry
  double origin[3], dim[3], h[3], ih[3];
  int cdim[3], span[3], ncells;
  double cutoff, cutoff2;
  float dt;
  double dt_d;
  float max_frac = 0.05f;
  int nthreads = 1;
  uint64_t time = 0;
  uint32_t seed_lo = 0, seed_hi = 0;
  int dpd_counter_noise = 0;
  int err = 0;
  int err_pid = -1;
  int integrator = 0;     // 0 forward Euler, 1 Runge-Kutta 4 (EngineIntegrator)
  bool update_pf = true;  // INTEGRATOR_UPDATE_PERSISTENTFORCE

  std::vector<Cell> cells;
  std::vector<Task> tasks;
  std::vector<int> order;  // single-runner execution order
  std::vector<Type> types;
  std::vector<Pot> pots;       // pair potentials
  std::vector<int> potmat;     // ntypes x ntypes -> index into pots, -1 none
  std::vector<Pot> bpots;      // bond potentials
  std::vector<Bond> bonds;
  std::vector<Pot> apots, dpots;  // angle and dihedral potentials
  std::vector<Angle> angles;
  std::vector<Dihedral> dihedrals;
  std::vector<Exclusion> exclusions;
  std::vector<FluxEntry> fluxes;
  std::vector<std::vector<int>> flux_pair;  // ntypes^2: fluxes of the unordered type pair
  int nparts = 0;

  double ekin = 0, epot_nb = 0, epot_bond = 0;
  double epot_angle = 0, epot_dihedral = 0, epot_exclusion = 0;

  // Boundary conditions (MxBoundaryConditions): per side the kind bits,
  // wall velocity, restore factor and radius; wall potentials by side and
  // type.  The space is rebuilt when they change.
  int bc_kind[6] = {2, 2, 2, 2, 2, 2};
  float bc_vel[6][3] = {};
  float bc_restore[6] = {1, 1, 1, 1, 1, 1};
  float bc_radius[6] = {};
  bool periodic[3] = {true, true, true};
  std::vector<Pot> wpots;
  std::vector<int> wpot_side, wpot_type;
  double init_origin[3] = {}, init_dim[3] = {};
  int init_cells[3] = {};
  float dpd_delta = 0;
  bool dpd_delta_set = false;

  int cellid(int i, int j, int k) const { return (i * cdim[1] + j) * cdim[2] + k; }
};

// ---------------------------------------------------------------------------
// Geometry and task graph (space.cpp:626-978, engine.cpp:1502-1560).

// space_getsid: min-image shift from ci to cj (float), flipping so it points
// forward.  Returns the sort-list ID.
int getsid(const Oracle &o, int &ci, int &cj, float shift[3]) {
  for (int k = 0; k < 3; ++k) {
    shift[k] = o.cells[cj].origin[k] - o.cells[ci].origin[k];
    if (shift[k] * 2 > o.dim[k])
      shift[k] -= o.dim[k];
    else if (shift[k] * 2 < -o.dim[k])
      shift[k] += o.dim[k];
  }
  int sid = 0;
  for (int k = 0; k < 3; ++k) sid = 3 * sid + ((shift[k] < 0) ? 0 : ((shift[k] > 0) ? 2 : 1));
  if (kFlip[sid]) {
    std::swap(ci, cj);
    shift[0] = -shift[0];
    shift[1] = -shift[1];
    shift[2] = -shift[2];
  }
  return kSortlistID[sid];
}

int init_space(Oracle &o, const double origin[3], const double dim[3], const int cells_in[3],
               double cutoff) {
  for (int k = 0; k < 3; ++k) {
    o.init_origin[k] = origin[k];
    o.init_dim[k] = dim[k];
    o.init_cells[k] = cells_in[k];
  }
  int cells[3] = {cells_in[0], cells_in[1], cells_in[2]};
  for (int k = 0; k < 3; ++k)
    if (cells[k] < 3) cells[k] = 3;  // periodic needs >= 3 cells
  double L[3];
  for (int k = 0; k < 3; ++k) L[k] = (dim[k] - origin[k]) / cells[k];
  for (int k = 0; k < 3; ++k) {
    o.origin[k] = origin[k];
    o.dim[k] = dim[k];
    o.cdim[k] = (int)std::floor(dim[k] / L[k]);
  }
  o.cutoff = cutoff;
  o.cutoff2 = cutoff * cutoff;
  o.ncells = o.cdim[0] * o.cdim[1] * o.cdim[2];
  for (int k = 0; k < 3; ++k) {
    o.h[k] = o.dim[k] / o.cdim[k];
    o.ih[k] = 1.0 / o.h[k];
  }
  o.cells = std::vector<Cell>(o.ncells);
  for (int i = 0; i < o.ncells; ++i) o.cells[i].gen = std::mt19937(i);
  int l[3];
  double org[3];
  for (l[0] = 0; l[0] < o.cdim[0]; l[0]++) {
    org[0] = origin[0] + l[0] * o.h[0];
    for (l[1] = 0; l[1] < o.cdim[1]; l[1]++) {
      org[1] = origin[1] + l[1] * o.h[1];
      for (l[2] = 0; l[2] < o.cdim[2]; l[2]++) {
        org[2] = origin[2] + l[2] * o.h[2];
        Cell &c = o.cells[o.cellid(l[0], l[1], l[2])];
        for (int k = 0; k < 3; ++k) {
          c.loc[k] = l[k];
          c.origin[k] = org[k];
        }
        // cell_active_* flags: the low side for the first cell of an axis,
        // else the high side for the last.
        c.walls = 0;
        for (int a = 0; a < 3; ++a) {
          if (l[a] == 0 && (o.bc_kind[2 * a] & kBcActive))
            c.walls |= 1 << (2 * a);
          else if (l[a] + 1 == o.cdim[a] && (o.bc_kind[2 * a + 1] & kBcActive))
            c.walls |= 1 << (2 * a + 1);
        }
      }
    }
  }
  for (int k = 0; k < 3; ++k) o.span[k] = (int)std::ceil(cutoff * o.ih[k]);
  for (int k = 0; k < 3; ++k)
    if (o.cdim[k] < 2 * o.span[k] + 1) return kErrRange;

  // Cell pairs and self tasks, in the reference's loop order.
  o.tasks.clear();
  for (int i = 0; i < o.cdim[0]; i++)
    for (int j = 0; j < o.cdim[1]; j++)
      for (int k = 0; k < o.cdim[2]; k++) {
        int id1 = o.cellid(i, j, k);
        // Neighbours across a non-periodic side are not paired.
        auto nbr = [&](int a, int v) { return o.periodic[a] ? (v + o.cdim[a]) % o.cdim[a] : v; };
        auto off = [&](int a, int v) { return !o.periodic[a] && (v < 0 || v >= o.cdim[a]); };
        for (l[0] = -o.span[0]; l[0] <= o.span[0]; l[0]++) {
          if (off(0, i + l[0])) continue;
          int ii = nbr(0, i + l[0]);
          for (l[1] = -o.span[1]; l[1] <= o.span[1]; l[1]++) {
            if (off(1, j + l[1])) continue;
            int jj = nbr(1, j + l[1]);
            for (l[2] = -o.span[2]; l[2] <= o.span[2]; l[2]++) {
              double lh[3];
              for (int q = 0; q < 3; ++q) lh[q] = o.h[q] * std::fmax(std::abs(l[q]) - 1, 0);
              if (lh[0] * lh[0] + lh[1] * lh[1] + lh[2] * lh[2] > o.cutoff2) continue;
              if (off(2, k + l[2])) continue;
              int kk = nbr(2, k + l[2]);
              int id2 = o.cellid(ii, jj, kk);
              int ci = id1, cj = id2;
              float sh[3];
              int sid = getsid(o, ci, cj, sh);
              if (id1 < id2 || (id1 == id2 && l[0] == 0 && l[1] == 0 && l[2] == 0)) {
                Task t;
                t.type = (id1 == id2) ? kTaskSelf : kTaskPair;
                t.i = ci;
                t.j = cj;
                t.sid = sid;
                o.tasks.push_back(t);
              }
            }
          }
        }
      }
  // Sort tasks, and pair -> sort dependencies.
  int first_sort = (int)o.tasks.size();
  for (int c = 0; c < o.ncells; ++c) {
    Task t;
    t.type = kTaskSort;
    t.i = c;
    t.j = -1;
    t.sid = 0;
    o.tasks.push_back(t);
  }
  for (int t = 0; t < first_sort; ++t)
    if (o.tasks[t].type == kTaskPair) {
      o.tasks[first_sort + o.tasks[t].i].unlock.push_back(t);
      o.tasks[first_sort + o.tasks[t].j].unlock.push_back(t);
      o.cells[o.tasks[t].i].sort_flags |= 1 << o.tasks[t].sid;
      o.cells[o.tasks[t].j].sort_flags |= 1 << o.tasks[t].sid;
    }

  // One runner, one queue: the first ready task in queue order runs next
  // (queue_get with a single runner always takes the first task whose wait
  // count is zero).  The order is fixed from the first step on.
  std::vector<int> wait(o.tasks.size(), 0);
  for (auto &t : o.tasks)
    for (int u : t.unlock) wait[u]++;
  std::set<int> ready;
  for (int t = 0; t < (int)o.tasks.size(); ++t)
    if (wait[t] == 0) ready.insert(t);
  o.order.clear();
  while (!ready.empty()) {
    int t = *ready.begin();
    ready.erase(ready.begin());
    o.order.push_back(t);
    for (int u : o.tasks[t].unlock)
      if (--wait[u] == 0) ready.insert(u);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Potential evaluation (potential_eval.hpp:67-161, dpd_eval.hpp:21-169).

inline void eval_plain(const Pot &p, float r2, float *e, float *f) {
  float r = std::sqrt(r2);
  int ind = (int)std::fmax(0.0f, p.alpha[0] + r * (p.alpha[1] + r * p.alpha[2]));
  const float *c = &p.c[(size_t)ind * kChunk];
  float x = (r - c[0]) * c[1];
  float ee = c[2] * x + c[3];
  float eff = c[2];
  for (int k = 4; k < kChunk; ++k) {
    eff = eff * x + ee;
    ee = ee * x + c[k];
  }
  *e = ee;
  *f = eff * c[1] / r;
}

inline bool eval_ex(const Pot &p, float ri, float rj, float r2, float *e, float *f) {
  const float eps = std::numeric_limits<float>::epsilon();
  float r = std::sqrt(r2);
  float ro = r < eps ? eps : r;
  r = r < p.a ? p.a : r;
  if (p.flags & kFlagScaled)
    r = r / (ri + rj);
  else if (p.flags & kFlagShifted)
    r = r - (ri + rj) + p.r0_plusone;
  unsigned ind = std::max(0.0f, p.alpha[0] + r * (p.alpha[1] + r * p.alpha[2]));
  if (r > p.b || ind > (unsigned)p.n) return false;
  const float *c = &p.c[(size_t)ind * kChunk];
  float x = (r - c[0]) * c[1];
  float ee = c[2] * x + c[3];
  float eff = c[2];
  for (int k = 4; k < kChunk; ++k) {
    eff = eff * x + ee;
    ee = ee * x + c[k];
  }
  *e = ee;
  *f = eff * c[1] / ro;
  return true;
}

// Pair noise of the counter-based DPD mode (zero mean, unit variance), keyed
// by the pair and the time step.
float pair_noise(uint32_t a, uint32_t b, uint64_t time, uint32_t klo, uint32_t) {
  return mxg::pair_uniform(a, b, mxg::step_key(klo, (uint32_t)time));
}

// dpd_eval: forces on the pair along e = dx / r.  Returns false out of range.
inline bool dpd_eval(Oracle &o, const Pot &p, Part *pi, Part *pj, const float dx[3], float r2) {
  float ri = pi->radius, rj = pj->radius;
  bool shifted = p.flags & kFlagShifted;
  float cutoff = shifted ? (p.b + ri + rj) : p.b;
  if (r2 > cutoff * cutoff) return false;
  float r = std::sqrt(r2);
  float e[3] = {dx[0] / r, dx[1] / r, dx[2] / r};
  float v[3] = {pi->v[0] - pj->v[0], pi->v[1] - pj->v[1], pi->v[2] - pj->v[2]};
  float shifted_r = shifted ? r - ri - rj : r;
  float omega_c = shifted_r < 0.f ? 1.f : (1 - shifted_r / cutoff);
  float fc = p.dpd_alpha * omega_c;
  float omega_d = omega_c * omega_c;
  float ev = e[0] * v[0] + e[1] * v[1] + e[2] * v[2];
  float fd = -p.dpd_gamma * omega_d * ev;
  float fr;
  if (o.dpd_counter_noise) {
    uint32_t a = (uint32_t)std::min(pi->id, pj->id), b = (uint32_t)std::max(pi->id, pj->id);
    fr = p.dpd_sigma * omega_c * pair_noise(a, b, o.time, o.seed_lo, o.seed_hi) * o.dpd_delta;
  } else {
    fr = p.dpd_sigma * omega_c * o.dpd_delta;
  }
  float f = fc + fd + fr;
  for (int k = 0; k < 3; ++k) {
    pj->f[k] = pj->f[k] - f * e[k];
    pi->f[k] = pi->f[k] + f * e[k];
  }
  return true;
}

// potential_eval_super_ex: overlap randomisation, then DPD or table.
inline void eval_pair(Oracle &o, Cell &cell, int pot, Part *pi, Part *pj, float dx[3], float r2, float nd,
                      float *epot) {
  const Pot &p = o.pots[pot];
  if (r2 < p.a * p.a) {
    dx[0] = cell.gauss(cell.gen);
    dx[1] = cell.gauss(cell.gen);
    dx[2] = cell.gauss(cell.gen);
    float len = std::sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    dx[0] = dx[0] * p.a / len;
    dx[1] = dx[1] * p.a / len;
    dx[2] = dx[2] * p.a / len;
    r2 = p.a * p.a;
  }
  if (p.kind == kKindDPD) {
    // the reference draws (and discards) one more gaussian here
    (void)cell.gauss(cell.gen);
    if (dpd_eval(o, p, pi, pj, dx, r2)) {
      pi->f[3] += nd;
      pj->f[3] += nd;
      *epot += 0.0f;
    }
  } else {
    float e, f;
    if (eval_ex(p, pi->radius, pj->radius, r2, &e, &f)) {
      for (int k = 0; k < 3; ++k) {
        float w = f * dx[k];
        pi->f[k] -= w;
        pj->f[k] += w;
      }
      pi->f[3] += nd;
      pj->f[3] += nd;
      *epot += e;
    }
  }
}

// Number-density kernel W(r^2, h): cubic spline of r / h with r from the
// approximate x86 reciprocal square root (smoothing_kernel.hpp:19-54).
inline float qsqrt(float f) {
  __m128 t = _mm_set_ss(f);
  t = _mm_rsqrt_ss(t);
  return 1.0 / _mm_cvtss_f32(t);
}

inline float w_cubic(float r2, float h) {
  float r = qsqrt(r2);
  float x = r / h;
  float y;
  if (x < 1.f) {
    float x2 = x * x;
    y = 1.f - (3.f / 2.f) * x2 + (3.f / 4.f) * x2 * x;
  } else if (x >= 1.f && x < 2.f) {
    float arg = 2.f - x;
    y = (1.f / 4.f) * arg * arg * arg;
  } else {
    y = 0.f;
  }
  return y / (M_PI * h * h * h);
}

inline float fr2(const float *x1, const float *x2, float *dx) {
  dx[0] = x1[0] - x2[0];
  dx[1] = x1[1] - x2[1];
  dx[2] = x1[2] - x2[2];
  return dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
}

// ---------------------------------------------------------------------------
// Sorting (runner_dosort.cpp:84-146, runner.cpp:102-165).

void sort_descending(uint32_t *parts, int N) {
  struct {
    short lo, hi;
  } qstack[10];
  int qpos, i, j, lo, hi, imax;
  unsigned pivot;
  uint32_t temp;
  qstack[0].lo = 0;
  qstack[0].hi = N - 1;
  qpos = 0;
  while (qpos >= 0) {
    lo = qstack[qpos].lo;
    hi = qstack[qpos].hi;
    qpos -= 1;
    if (hi - lo < 15) {
      for (i = lo; i < hi; i++) {
        imax = i;
        for (j = i + 1; j <= hi; j++)
          if ((parts[j] & 0xffff) > (parts[imax] & 0xffff)) imax = j;
        if (imax != i) {
          temp = parts[imax];
          parts[imax] = parts[i];
          parts[i] = temp;
        }
      }
    } else {
      pivot = parts[(lo + hi) / 2] & 0xffff;
      i = lo;
      j = hi;
      while (i <= j) {
        while ((parts[i] & 0xffff) > pivot) i++;
        while ((parts[j] & 0xffff) < pivot) j--;
        if (i <= j) {
          if (i < j) {
            temp = parts[i];
            parts[i] = parts[j];
            parts[j] = temp;
          }
          i += 1;
          j -= 1;
        }
      }
      if (j > (lo + hi) / 2) {
        if (lo < j) {
          qpos += 1;
          qstack[qpos].lo = lo;
          qstack[qpos].hi = j;
        }
        if (i < hi) {
          qpos += 1;
          qstack[qpos].lo = i;
          qstack[qpos].hi = hi;
        }
      } else {
        if (i < hi) {
          qpos += 1;
          qstack[qpos].lo = i;
          qstack[qpos].hi = hi;
        }
        if (lo < j) {
          qpos += 1;
          qstack[qpos].lo = lo;
          qstack[qpos].hi = j;
        }
      }
    }
  }
}

struct SortScale {
  float bias, dscale;
};
SortScale sort_scale(const Oracle &o) {
  SortScale s;
  s.bias = std::sqrt(o.h[0] * o.h[0] + o.h[1] * o.h[1] + o.h[2] * o.h[2]);
  s.dscale = (float)SHRT_MAX / (2 * s.bias);
  return s;
}

void do_sort(Oracle &o, Cell &c) {
  int count = (int)c.parts.size();
  c.sortlist.resize((size_t)13 * std::max(count, 1));
  if (count == 0) return;
  SortScale s = sort_scale(o);
  const SortDirs &D = sortdirs();
  for (int sid = 0; sid < 13; ++sid) {
    if (!(c.sort_flags & (1 << sid))) continue;
    const float *n = &D.n[3 * sid];
    uint32_t *ip = &c.sortlist[(size_t)count * sid];
    for (int i = 0; i < count; ++i) {
      const Part &p = c.parts[i];
      ip[i] = ((uint32_t)i << 16) |
              (unsigned int)(s.dscale * (s.bias + p.x[0] * n[0] + p.x[1] * n[1] + p.x[2] * n[2]));
    }
    sort_descending(ip, count);
  }
}

// ---------------------------------------------------------------------------
// Pair and self loops (runner_dopair.cpp:91-340, 446-682).

// flux_eval_ex: the fluxes of the pair's type pair, each between the
// particle of the flux's type a (the pair's i for equal types) and that of
// type b, scaled by (1 - r/cutoff)^2, with half the decay taken from each.
void flux_eval(const Oracle &o, const std::vector<int> &fl, float r2, Part *part_i, Part *part_j) {
  float r = std::sqrt(r2);
  float term = (1. - r / o.cutoff);
  term = term * term;
  for (int e : fl) {
    const FluxEntry &F = o.fluxes[e];
    Part *pi = part_i->type == F.ta ? part_i : part_j;
    Part *pj = part_j->type == F.tb ? part_j : part_i;
    const float ssi = pi->sv[F.ia], ssj = pj->sv[F.ib];
    float q = term;
    if (F.kind == 0) {
      q *= F.coef * (ssi - ssj);
    } else if (F.kind == 1) {
      const float qq = F.coef * (ssi - F.target);
      const float scale = qq > 0.f;
      q *= scale * qq;
    } else {
      const float qq = F.coef * (F.target - ssj) * ssi;
      const float scale = qq > 0.f;
      q *= scale * qq;
    }
    const float half_decay = F.decay / 2.f;
    pi->sq[F.ia] = pi->sq[F.ia] - q - half_decay * ssi;
    pj->sq[F.ib] = pj->sq[F.ib] + q - half_decay * ssj;
  }
}

void rebuild_flux_pairs(Oracle &o) {
  const int nt = (int)o.types.size();
  o.flux_pair.assign((size_t)nt * nt, {});
  for (size_t e = 0; e < o.fluxes.size(); ++e) {
    const FluxEntry &F = o.fluxes[e];
    if (F.ta >= nt || F.tb >= nt) continue;
    o.flux_pair[F.ta * nt + F.tb].push_back((int)e);
    if (F.ta != F.tb) o.flux_pair[F.tb * nt + F.ta].push_back((int)e);
  }
}

// MxFluxes_Integrate for one cell's particles (those still in it after its
// advance): s += dt q for species that are not constant, then q = 0.
void flux_integrate(Oracle &o, int cid) {
  for (Part &p : o.cells[cid].parts) {
    const int cm = o.types[p.type].spec_const;
    for (int k = 0; k < 4; ++k) {
      const float konst = (cm >> k) & 1 ? 0.f : 1.f;
      p.sv[k] += o.dt_d * p.sq[k] * konst;
      p.sq[k] = 0;
    }
  }
}

void do_pair(Oracle &o, int ci_, int cj_) {
  int ci = ci_, cj = cj_;
  Cell *cell_i = &o.cells[ci], *cell_j = &o.cells[cj];
  if (cell_i->parts.empty() || cell_j->parts.empty()) return;
  float cutoff = o.cutoff;
  float cutoff2 = cutoff * cutoff;
  SortScale s = sort_scale(o);
  float maxdx = 0.0f;
  int dmaxdist = 2 + s.dscale * (cutoff + 2 * maxdx);
  float shift[3];
  int sid = getsid(o, ci, cj, shift);
  cell_i = &o.cells[ci];
  cell_j = &o.cells[cj];
  int count_i = (int)cell_i->parts.size(), count_j = (int)cell_j->parts.size();
  float nshift = std::sqrt(shift[0] * shift[0] + shift[1] * shift[1] + shift[2] * shift[2]);
  int dnshift = s.dscale * nshift;
  const uint32_t *iparts = &cell_i->sortlist[(size_t)count_i * sid];
  const uint32_t *jparts = &cell_j->sortlist[(size_t)count_j * sid];
  float epot = 0.0f;
  float pix[3], dx[3];
  for (int i = 0; i < count_i; i++) {
    if ((jparts[count_j - 1] & 0xffff) + dnshift - (iparts[i] & 0xffff) > (unsigned)dmaxdist) break;
    Part *part_i = &cell_i->parts[iparts[i] >> 16];
    pix[0] = part_i->x[0] - shift[0];
    pix[1] = part_i->x[1] - shift[1];
    pix[2] = part_i->x[2] - shift[2];
    for (int j = count_j - 1;
         j >= 0 && (jparts[j] & 0xffff) + dnshift - (iparts[i] & 0xffff) < (unsigned)dmaxdist; j--) {
      Part *part_j = &cell_j->parts[jparts[j] >> 16];
      float r2 = fr2(pix, part_j->x, dx);
      if (r2 > cutoff2) continue;
      part_i->ncount++;
      part_j->ncount++;
      float nd = w_cubic(r2, cutoff);
      int pot = o.potmat[part_i->type * o.types.size() + part_j->type];
      const std::vector<int> *fl = o.fluxes.empty() ? nullptr : &o.flux_pair[part_i->type * o.types.size() + part_j->type];
      if (fl && fl->empty()) fl = nullptr;
      if (pot < 0 && !fl) {
        part_i->f[3] += nd;
        part_j->f[3] += nd;
        continue;
      }
      if (fl) flux_eval(o, *fl, r2, part_i, part_j);
      if (pot < 0) continue;
      eval_pair(o, *cell_i, pot, part_i, part_j, dx, r2, nd, &epot);
    }
  }
  cell_i->epot += epot;
}

void single_body(Oracle &o, Part *p) {
  const Type &t = o.types[p->type];
  if (t.sb == 0) return;
  if (t.has_random && o.update_pf && (o.time + (uint64_t)p->id) % t.rnd_steps == 0) {
    // MxRandomVector(mean, std) with Philox draws.
    mxg::u32x4 ctr = {{(uint32_t)p->id, (uint32_t)o.time, mxg::kStreamSingle, 0u}};
    mxg::u32x4 r = mxg::philox4x32_10(ctr, o.seed_lo, o.seed_hi);
    float u1 = mxg::u01_open_closed(r.v[0]);
    float u2 = mxg::u01_closed_open(r.v[1]);
    float ut = mxg::u01_closed_open(r.v[2]);
    float up = mxg::u01_closed_open(r.v[3]);
    float g = std::sqrt(-2.0f * std::log(u1)) * std::cos(6.2831853f * u2);
    float mag = t.rnd_mean + t.rnd_std * g;
    float theta = 6.2831853f * ut;
    float phi = std::acos(1.0f - 2.0f * up);
    p->pf[0] = mag * std::sin(phi) * std::cos(theta);
    p->pf[1] = mag * std::sin(phi) * std::sin(theta);
    p->pf[2] = mag * std::cos(phi);
  }
  if (t.sb == 1) {  // gaussian_force
    p->f[0] += p->pf[0];
    p->f[1] += p->pf[1];
    p->f[2] += p->pf[2];
  } else if (t.sb == 2) {  // berendsen_force
    float scale = t.sb_coef * ((t.target_energy / t.kinetic_energy) - 1.0);
    p->f[0] += scale * p->v[0];
    p->f[1] += scale * p->v[1];
    p->f[2] += scale * p->v[2];
  } else if (t.sb == 3) {  // constant_force, scaling constant 1
    p->f[0] += t.cf[0] * 1.f;
    p->f[1] += t.cf[1] * 1.f;
    p->f[2] += t.cf[2] * 1.f;
  } else if (t.sb == 4) {  // friction_force
    float v2 = p->v[0] * p->v[0] + p->v[1] * p->v[1] + p->v[2] * p->v[2];
    float scale = -1. * t.sb_coef * v2;
    p->f[0] += scale * p->v[0] + p->pf[0];
    p->f[1] += scale * p->v[1] + p->pf[1];
    p->f[2] += scale * p->v[2] + p->pf[2];
  }
}

// DPD's 1/sqrt(dt): a function-static in the reference, fixed by the dt of
// its first evaluation (dpd_eval.hpp:24, 74).  Before the first force phase
// it is the value that phase will latch.
float dpd_delta_of(const Oracle &o) { return o.dpd_delta_set ? o.dpd_delta : (float)(1.f / std::sqrt(o.dt_d)); }

// potential_eval_ex in FP64 (the FP64 statement of the table potentials).
bool table_ex64(const Pot &p, double ri, double rj, double r2, double *e, double *f) {
  double r = std::sqrt(r2);
  double ro = r < FLT_EPSILON ? FLT_EPSILON : r;
  r = r < p.a ? p.a : r;
  if (p.flags & kFlagScaled)
    r = r / (ri + rj);
  else if (p.flags & kFlagShifted)
    r = r - (ri + rj) + p.r0_plusone;
  double indf = std::max(0.0, p.alpha[0] + r * (p.alpha[1] + r * (double)p.alpha[2]));
  unsigned ind = (unsigned)indf;
  if (r > p.b || ind > (unsigned)p.n) return false;
  const float *c = &p.c[(size_t)ind * kChunk];
  double x = (r - c[0]) * (double)c[1];
  double ee = c[2] * x + c[3], eff = c[2];
  for (int k = 4; k < kChunk; ++k) {
    eff = eff * x + ee;
    ee = ee * x + c[k];
  }
  *e = ee;
  *f = eff * c[1] / ro;
  return true;
}

// boundary_eval / boundary_potential_eval_ex / dpd_boundary_eval: the
// offset vector dx persists from side to side, and r^2 < a (not a^2) moves
// the particle to the normal times a, as in the reference.  add(k, w) adds
// w to the particle's force component k; the energy goes to *epot.
template <typename T, typename Add>
void boundary_eval(const Oracle &o, const Cell &c, const Part &p, T *epot, Add add) {
  static const float normals[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  T dx[3] = {0, 0, 0};
  for (int side = 0; side < 6; ++side) {
    if (!(c.walls & (1 << side))) continue;
    int pidx = -1;
    for (size_t w = 0; w < o.wpots.size(); ++w)
      if (o.wpot_side[w] == side && o.wpot_type[w] == p.type) pidx = (int)w;
    if (pidx < 0) continue;
    const Pot &pot = o.wpots[pidx];
    const int axis = side >> 1;
    const bool low = (side & 1) == 0;
    const T r = low ? (T)p.x[axis] : (T)(o.h[axis] - p.x[axis]);
    if (!(r <= (T)pot.b)) continue;
    dx[axis] = low ? r : -r;
    T r2 = r * r;
    if (r2 < (T)pot.a) {
      r2 = (T)pot.a * (T)pot.a;
      for (int k = 0; k < 3; ++k) dx[k] = (T)normals[side][k] * (T)pot.a;
    }
    if (pot.kind == kKindDPD) {
      const T cutoff = pot.b;
      if (r2 > cutoff * cutoff) continue;
      const T rr = std::sqrt(r2);
      const T e[3] = {dx[0] / rr, dx[1] / rr, dx[2] / rr};
      const T v[3] = {(T)p.v[0] - (T)o.bc_vel[side][0], (T)p.v[1] - (T)o.bc_vel[side][1],
                      (T)p.v[2] - (T)o.bc_vel[side][2]};
      const T omega_c = 1 - rr / cutoff;
      const T fc = (T)pot.dpd_alpha * omega_c;
      const T omega_d = omega_c * omega_c;
      const T fd = -(T)pot.dpd_gamma * omega_d * (e[0] * v[0] + e[1] * v[1] + e[2] * v[2]);
      const T fr = (T)pot.dpd_sigma * omega_c * (T)dpd_delta_of(o);
      const T f = fc + fd + fr;
      for (int k = 0; k < 3; ++k) add(k, f * e[k]);
    } else {
      float ee, eff;
      if (sizeof(T) == 4) {
        if (!eval_ex(pot, p.radius, o.bc_radius[side], (float)r2, &ee, &eff)) continue;
        for (int k = 0; k < 3; ++k) add(k, -(T)eff * dx[k]);
        *epot += (T)ee;
      } else {
        double e64, f64;
        if (!table_ex64(pot, p.radius, o.bc_radius[side], (double)r2, &e64, &f64)) continue;
        for (int k = 0; k < 3; ++k) add(k, -(T)f64 * dx[k]);
        *epot += (T)e64;
      }
    }
  }
}

// apply_update_pos_vel for a particle that left its cell: reflection at
// active free-slip or velocity sides of its cell; true when it then stays.
bool apply_update_pos_vel(const Oracle &o, const Cell &c, Part *p, const float h[3], int delta[3]) {
  bool enforced = false;
  for (int side = 0; side < 6; ++side) {
    if (!(c.walls & (1 << side))) continue;
    const int i = side >> 1;
    const bool low = (side & 1) == 0;
    const double dim = o.h[i];
    if (low ? !(p->x[i] <= 0) : !(p->x[i] >= dim)) continue;
    const int kind = o.bc_kind[side];
    if (kind & kBcFreeslip) {
      const float restitution = 1.0f;
      if (low)
        p->x[i] = -p->x[i] * restitution;
      else
        p->x[i] = dim - (p->x[i] - dim) * restitution;
      p->v[i] *= -restitution;
      enforced = true;
    } else if (kind & kBcVelocity) {
      const float restore = o.bc_restore[side];
      if (low)
        p->x[i] = -p->x[i] * restore;
      else
        p->x[i] = 2.f * dim - (p->x[i] * restore);
      for (int k = 0; k < 3; ++k) p->v[k] = 2.f * o.bc_vel[side][k] - (p->v[k] * restore);
      enforced = true;
    }
  }
  if (enforced)
    for (int k = 0; k < 3; k++) delta[k] = std::isgreaterequal(p->x[k], h[k]) - std::isless(p->x[k], 0.0);
  return delta[0] == 0 && delta[1] == 0 && delta[2] == 0;
}

void do_self(Oracle &o, int cid) {
  Cell &c = o.cells[cid];
  int count = (int)c.parts.size();
  if (count == 0) return;
  float cutoff2 = o.cutoff2;
  float epot = 0.0f;
  float pix[3], dx[3];
  for (int i = 0; i < count; i++) {
    Part *part_i = &c.parts[i];
    pix[0] = part_i->x[0];
    pix[1] = part_i->x[1];
    pix[2] = part_i->x[2];
    single_body(o, part_i);
    if (c.walls) boundary_eval<float>(o, c, *part_i, &epot, [&](int k, float w) { part_i->f[k] += w; });
    for (int j = i + 1; j < count; j++) {
      Part *part_j = &c.parts[j];
      float r2 = fr2(pix, part_j->x, dx);
      if (r2 > cutoff2) continue;
      part_i->ncount++;
      part_j->ncount++;
      float nd = w_cubic(r2, (float)o.cutoff);
      int pot = o.potmat[part_i->type * o.types.size() + part_j->type];
      const std::vector<int> *fl = o.fluxes.empty() ? nullptr : &o.flux_pair[part_i->type * o.types.size() + part_j->type];
      if (fl && fl->empty()) fl = nullptr;
      if (pot < 0 && !fl) {
        part_i->f[3] += nd;
        part_j->f[3] += nd;
        continue;
      }
      if (fl) flux_eval(o, *fl, r2, part_i, part_j);
      if (pot < 0) continue;
      eval_pair(o, c, pot, part_i, part_j, dx, r2, nd, &epot);
    }
  }
  c.epot += epot;
}

void run_task(Oracle &o, const Task &t) {
  if (t.type == kTaskSort)
    do_sort(o, o.cells[t.i]);
  else if (t.type == kTaskSelf)
    do_self(o, t.i);
  else
    do_pair(o, t.i, t.j);
}

// Threaded runners: sorts in parallel, then self/pair tasks with CAS cell
// locks (the cells_taboo scheme of queue.cpp:128-182).
void run_tasks_threaded(Oracle &o) {
  int T = o.nthreads;
  std::vector<int> sorts, work;
  for (int t = 0; t < (int)o.tasks.size(); ++t)
    (o.tasks[t].type == kTaskSort ? sorts : work).push_back(t);
  {
    std::atomic<int> next{0};
    auto fn = [&]() {
      for (int k; (k = next.fetch_add(1)) < (int)sorts.size();) run_task(o, o.tasks[sorts[k]]);
    };
    std::vector<std::thread> th;
    for (int i = 0; i < T; ++i) th.emplace_back(fn);
    for (auto &x : th) x.join();
  }
  // Each runner owns a contiguous block of the task list (tasks are ordered
  // by cell, so blocks are spatially apart and rarely contend); a task whose
  // cells are held by another runner is retried after the block.
  auto try_run = [&](const Task &t) {
    char z = 0;
    if (!o.cells[t.i].taboo.compare_exchange_strong(z, 1)) return false;
    if (t.type == kTaskPair) {
      z = 0;
      if (!o.cells[t.j].taboo.compare_exchange_strong(z, 1)) {
        o.cells[t.i].taboo = 0;
        return false;
      }
    }
    run_task(o, t);
    o.cells[t.i].taboo = 0;
    if (t.type == kTaskPair) o.cells[t.j].taboo = 0;
    return true;
  };
  auto fn = [&](int rid) {
    size_t b = work.size() * rid / T, e = work.size() * (rid + 1) / T;
    std::vector<int> pending;
    for (size_t k = b; k < e; ++k)
      if (!try_run(o.tasks[work[k]])) pending.push_back(work[k]);
    while (!pending.empty()) {
      std::vector<int> again;
      for (int t : pending)
        if (!try_run(o.tasks[t])) again.push_back(t);
      if (again.size() == pending.size()) std::this_thread::yield();
      pending.swap(again);
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < T; ++i) th.emplace_back(fn, i);
  for (auto &x : th) x.join();
}

// ---------------------------------------------------------------------------
// Bonds (bond.cpp:99-295).

struct Loc {
  int cell, idx;
};

void bond_eval(Oracle &o, const std::vector<Loc> &where) {
  double epot = 0.0;
  float pix[3], dx[3];
  for (auto &b : o.bonds) {
    if (!b.active) continue;
    const Loc &li = where[b.i], &lj = where[b.j];
    if (li.cell < 0 || lj.cell < 0) continue;
    Part *pi = &o.cells[li.cell].parts[li.idx];
    Part *pj = &o.cells[lj.cell].parts[lj.idx];
    const Pot &pot = o.bpots[b.pot];
    const int *loci = o.cells[li.cell].loc, *locj = o.cells[lj.cell].loc;
    for (int k = 0; k < 3; ++k) {
      int shift = loci[k] - locj[k];
      if (shift > 1)
        shift = -1;
      else if (shift < -1)
        shift = 1;
      pix[k] = pi->x[k] + o.h[k] * shift;
    }
    float r2 = fr2(pix, pj->x, dx);
    if (r2 < pot.a * pot.a || r2 > pot.b * pot.b)
      r2 = std::fmax(pot.a * pot.a, std::fmin(pot.b * pot.b, r2));
    float ee, eff;
    eval_plain(pot, r2, &ee, &eff);
    if (eff >= b.dissociation) {
      b.active = false;
    } else {
      for (int k = 0; k < 3; ++k) {
        float w = eff * dx[k];
        pi->f[k] -= w;
        pj->f[k] += w;
      }
      epot += ee;
    }
  }
  o.epot_bond = epot;
}

// ---------------------------------------------------------------------------
// Exclusions, angles and dihedrals, in the working precision T: float
// restates the reference's arithmetic (positions shifted in double and
// rounded once, everything else in float), double is the FP64 statement of
// the same terms.

// potential_eval_r (potential_eval.hpp:180-208).
template <typename T>
T table_r(const Pot &p, T r, T *f) {
  int ind = (int)std::fmax((T)0, (T)p.alpha[0] + r * ((T)p.alpha[1] + r * (T)p.alpha[2]));
  ind = std::min(std::max(ind, 0), p.n);
  const float *c = &p.c[(size_t)ind * kChunk];
  T x = (r - (T)c[0]) * (T)c[1];
  T ee = (T)c[2] * x + (T)c[3];
  T eff = (T)c[2];
  for (int k = 4; k < kChunk; ++k) {
    eff = eff * x + ee;
    ee = ee * x + (T)c[k];
  }
  *f = eff * (T)c[1];
  return ee;
}

// Particle b's position relative to particle a's cell: the cell offset
// folded to -1, 0, +1 (angle.cpp:148-163).
template <typename T>
void rel_to(const Oracle &o, const Loc &lb, const Loc &la, T out[3]) {
  const Part &pb = o.cells[lb.cell].parts[lb.idx];
  const int *locb = o.cells[lb.cell].loc, *loca = o.cells[la.cell].loc;
  for (int k = 0; k < 3; ++k) {
    int shift = locb[k] - loca[k];
    if (shift > 1)
      shift = -1;
    else if (shift < -1)
      shift = 1;
    out[k] = (T)(pb.x[k] + shift * o.h[k]);
  }
}

inline Part &part_at(Oracle &o, const Loc &l) { return o.cells[l.cell].parts[l.idx]; }

// exclusion_eval: the pair's table potential at r^2 (potential_eval, with
// the float difference of the positions plus the double shift), its force
// and energy subtracted.  add(id, k, w) adds w to component k of the
// force on particle id.
template <typename T, typename Add>
double exclusions_eval(const Oracle &o, const std::vector<Loc> &where, Add add) {
  double epot = 0.0;
  const int nt = (int)o.types.size();
  const float cutoff2 = (float)o.cutoff2;
  for (const Exclusion &x : o.exclusions) {
    const Loc &li = where[x.i], &lj = where[x.j];
    const Part &pi = o.cells[li.cell].parts[li.idx];
    const Part &pj = o.cells[lj.cell].parts[lj.idx];
    const int pidx = o.potmat[pj.type * nt + pi.type];
    if (pidx < 0) continue;
    const Pot &pot = o.pots[pidx];
    if (pot.kind != kKindTable) continue;
    const int *loci = o.cells[li.cell].loc, *locj = o.cells[lj.cell].loc;
    T dx[3], r2 = 0;
    for (int k = 0; k < 3; ++k) {
      int shift = loci[k] - locj[k];
      if (shift > 1)
        shift = -1;
      else if (shift < -1)
        shift = 1;
      dx[k] = (T)((T)(pi.x[k] - pj.x[k]) + o.h[k] * shift);
      if (sizeof(T) == 8) dx[k] = ((T)pi.x[k] - (T)pj.x[k]) + o.h[k] * shift;
      r2 += dx[k] * dx[k];
    }
    if ((float)r2 > cutoff2) continue;
    const T r = std::sqrt(r2);
    const int ind = (int)std::fmax((T)0, (T)pot.alpha[0] + r * ((T)pot.alpha[1] + r * (T)pot.alpha[2]));
    if (ind > pot.n) continue;
    const float *c = &pot.c[(size_t)ind * kChunk];
    T xx = (r - (T)c[0]) * (T)c[1];
    T ee = (T)c[2] * xx + (T)c[3], eff = (T)c[2];
    for (int k = 4; k < kChunk; ++k) {
      eff = eff * xx + ee;
      ee = ee * xx + (T)c[k];
    }
    eff = eff * (T)c[1] / r;
    for (int k = 0; k < 3; ++k) {
      const T w = eff * dx[k];
      add(pi.id, k, w);
      add(pj.id, k, -w);
    }
    epot += ee;
  }
  return -epot;
}

// angle_eval: cos(theta) at the vertex j from the rays to i and k, its
// gradients dxi, dxk, clamped into the table's domain, forces
// -eff dxi on i, -eff dxk on k and their negative sum on j.
template <typename T, typename Add>
double angles_eval(const Oracle &o, const std::vector<Loc> &where, Add add) {
  double epot = 0.0;
  for (size_t a = 0; a < o.angles.size(); ++a) {
    const Angle &A = o.angles[a];
    const Loc &li = where[A.i], &lj = where[A.j], &lk = where[A.k];
    const Part &pj = o.cells[lj.cell].parts[lj.idx];
    T xi[3], xk[3], xj[3];
    rel_to(o, li, lj, xi);
    rel_to(o, lk, lj, xk);
    for (int k = 0; k < 3; ++k) xj[k] = pj.x[k];
    T rji[3], rjk[3];
    for (int k = 0; k < 3; ++k) {
      rji[k] = xi[k] - xj[k];
      rjk[k] = xk[k] - xj[k];
    }
    const T dprod = rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2];
    const T inji = (T)1 / std::sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
    const T injk = (T)1 / std::sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
    T ct = std::fmax((T)-1, std::fmin((T)1, dprod * inji * injk));
    T dxi[3], dxk[3];
    if (ct == 0 || ct == -1) {
      mxg::u32x4 ctr = {{(uint32_t)a, (uint32_t)o.time, mxg::kStreamAngle, 0u}};
      mxg::u32x4 r = mxg::philox4x32_10(ctr, o.seed_lo, o.seed_hi);
      T v[3], vik[3];
      for (int k = 0; k < 3; ++k) {
        v[k] = (T)(2.0f * mxg::u01_closed_open(r.v[k]) - 1.0f);
        vik[k] = xi[k] - xk[k];
      }
      const T dv = v[0] * vik[0] + v[1] * vik[1] + v[2] * vik[2];
      for (int k = 0; k < 3; ++k) v[k] = v[k] - dv * vik[k];
      const T il = (T)1 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int k = 0; k < 3; ++k) dxi[k] = dxk[k] = v[k] * il;
    } else {
      for (int k = 0; k < 3; ++k) {
        dxi[k] = (rjk[k] * injk - ct * rji[k] * inji) * inji;
        dxk[k] = (rji[k] * inji - ct * rjk[k] * injk) * injk;
      }
    }
    const Pot &pot = o.apots[A.pot];
    if (ct < (T)pot.a || ct > (T)pot.b) ct = std::fmax((T)pot.a, std::fmin((T)pot.b, ct));
    T eff;
    const T ee = table_r(pot, ct, &eff);
    for (int k = 0; k < 3; ++k) {
      const T wi = eff * dxi[k], wk = eff * dxk[k];
      add(A.i, k, -wi);
      add(A.k, k, -wk);
      add(A.j, k, wi + wk);
    }
    epot += ee;
  }
  return epot;
}

// dihedral_eval: cos(phi) between the planes (i, j, k) and (j, k, l),
// phi = 180 degrees for the trans form, and its gradient with respect to
// the four positions.  With b1 = xj - xi, b2 = xk - xj, b3 = xl - xk and
// the normals m = b1 x b2, n = b2 x b3, cos(phi) = m.n / (|m| |n|); its
// derivatives along m and n are g_m = n / (|m||n|) - cos(phi) m / |m|^2 and
// g_n likewise, so d/db1 = b2 x g_m, d/db2 = g_m x b1 + b3 x g_n,
// d/db3 = g_n x b2.  Forces -eff dxi, -eff dxj, -eff dxl and their negative
// sum on k.
template <typename T, typename Add>
double dihedrals_eval(const Oracle &o, const std::vector<Loc> &where, Add add) {
  auto cross = [](const T a[3], const T b[3], T c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
  };
  auto dot = [](const T a[3], const T b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  double epot = 0.0;
  for (const Dihedral &D : o.dihedrals) {
    const Loc &li = where[D.i], &lj = where[D.j], &lk = where[D.k], &ll = where[D.l];
    const Part &pj = o.cells[lj.cell].parts[lj.idx];
    T xi[3], xj[3], xk[3], xl[3];
    rel_to(o, li, lj, xi);
    rel_to(o, lk, lj, xk);
    rel_to(o, ll, lj, xl);
    for (int k = 0; k < 3; ++k) xj[k] = pj.x[k];
    T b1[3], b2[3], b3[3];
    for (int k = 0; k < 3; ++k) {
      b1[k] = xj[k] - xi[k];
      b2[k] = xk[k] - xj[k];
      b3[k] = xl[k] - xk[k];
    }
    T m[3], n[3];
    cross(b1, b2, m);
    cross(b2, b3, n);
    const T mm = dot(m, m), nn = dot(n, n), mn = dot(m, n);
    const T inv = (T)1 / std::sqrt(mm * nn);
    const T c = mn * inv;
    T gm[3], gn[3];
    for (int k = 0; k < 3; ++k) {
      gm[k] = n[k] * inv - (c / mm) * m[k];
      gn[k] = m[k] * inv - (c / nn) * n[k];
    }
    T d1[3], d3[3], e1[3], e2[3];
    cross(b2, gm, d1);
    cross(gn, b2, d3);
    cross(gm, b1, e1);
    cross(b3, gn, e2);
    T dxi[3], dxj[3], dxl[3];
    for (int k = 0; k < 3; ++k) {
      dxi[k] = -d1[k];
      dxj[k] = d1[k] - (e1[k] + e2[k]);
      dxl[k] = d3[k];
    }
    const Pot &pot = o.dpots[D.pot];
    T cp = std::fmax((T)-1, std::fmin((T)1, c));
    if (cp < (T)pot.a || cp > (T)pot.b) cp = std::fmax((T)pot.a, std::fmin((T)pot.b, cp));
    T eff;
    const T ee = table_r(pot, cp, &eff);
    for (int k = 0; k < 3; ++k) {
      const T wi = eff * dxi[k], wj = eff * dxj[k], wl = eff * dxl[k];
      add(D.i, k, -wi);
      add(D.j, k, -wj);
      add(D.l, k, -wl);
      add(D.k, k, wi + wj + wl);
    }
    epot += ee;
  }
  return epot;
}

// The reference's order: exclusions (stage 0) before the bonds, angles and
// dihedrals (stage 1) after them, each force added to the particle's float
// force as it is computed.
void terms_eval(Oracle &o, const std::vector<Loc> &where, int stage) {
  auto add = [&](int id, int k, float w) { part_at(o, where[id]).f[k] += w; };
  if (stage == 0) {
    if (!o.exclusions.empty()) o.epot_exclusion = exclusions_eval<float>(o, where, add);
  } else {
    if (!o.angles.empty()) o.epot_angle = angles_eval<float>(o, where, add);
    if (!o.dihedrals.empty()) o.epot_dihedral = dihedrals_eval<float>(o, where, add);
  }
}

// ---------------------------------------------------------------------------
// Advance (engine_advance.cpp:126-239, 248-422).

struct AdvanceConst {
  float dt, h[3], h2[3], maxv[3], maxv2[3], maxx[3], maxx2[3];
};

void advance_cell(Oracle &o, const AdvanceConst &K, int cid) {
  Cell &c = o.cells[cid];
  size_t pid = 0;
  while (pid < c.parts.size()) {
    Part *p = &c.parts[pid];
    bool toofast = false;
    int delta[3];
    if ((p->flags & kFrozenX) && (p->flags & kFrozenY) && (p->flags & kFrozenZ)) {
      pid++;
      continue;
    }
    float mask[3] = {(p->flags & kFrozenX) ? 0.0f : 1.0f, (p->flags & kFrozenY) ? 0.0f : 1.0f,
                     (p->flags & kFrozenZ) ? 0.0f : 1.0f};
    if (o.types[p->type].dynamics == kNewtonian) {
      for (int k = 0; k < 3; k++) {
        float v = mask[k] * (p->v[k] + K.dt * p->f[k] * p->imass);
        float neg = v / std::abs(v);
        p->v[k] = v * v <= K.maxv2[k] ? v : neg * K.maxv[k];
        p->x[k] += K.dt * p->v[k];
        delta[k] = std::isgreaterequal(p->x[k], K.h[k]) - std::isless(p->x[k], 0.0);
        toofast = toofast || (p->x[k] >= K.h2[k] || p->x[k] <= -K.h[k]);
      }
    } else {
      for (int k = 0; k < 3; k++) {
        float dx = mask[k] * (K.dt * p->f[k] * p->imass);
        float neg = dx / std::abs(dx);
        p->x[k] += dx * dx <= K.maxx2[k] ? dx : neg * K.maxx[k];
        delta[k] = std::isgreaterequal(p->x[k], K.h[k]) - std::isless(p->x[k], 0.0);
        toofast = toofast || (p->x[k] >= K.h2[k] || p->x[k] <= -K.h[k]);
      }
    }
    if (toofast && o.err == 0) {
      o.err = kErrToofast;
      o.err_pid = p->id;
    }
    if ((delta[0] != 0 || delta[1] != 0 || delta[2] != 0) && c.walls &&
        apply_update_pos_vel(o, c, p, K.h, delta)) {
      pid += 1;
    } else if (delta[0] != 0 || delta[1] != 0 || delta[2] != 0) {
      for (int k = 0; k < 3; k++) p->x[k] -= delta[k] * K.h[k];
      int dest = o.cellid((c.loc[0] + delta[0] + o.cdim[0]) % o.cdim[0],
                          (c.loc[1] + delta[1] + o.cdim[1]) % o.cdim[1],
                          (c.loc[2] + delta[2] + o.cdim[2]) % o.cdim[2]);
      {
        std::lock_guard<std::mutex> g(o.cells[dest].mtx);
        o.cells[dest].incoming.push_back(*p);
      }
      size_t last = c.parts.size() - 1;
      if (pid < last) c.parts[pid] = c.parts[last];
      c.parts.pop_back();
    } else {
      pid += 1;
    }
  }
}

void welcome_all(Oracle &o) {
  for (auto &c : o.cells) {
    for (auto &p : c.incoming) c.parts.push_back(p);
    c.incoming.clear();
  }
}

std::vector<Loc> locate(const Oracle &o) {
  std::vector<Loc> where(o.nparts, Loc{-1, -1});
  for (int c = 0; c < o.ncells; ++c)
    for (int k = 0; k < (int)o.cells[c].parts.size(); ++k) where[o.cells[c].parts[k].id] = Loc{c, k};
  return where;
}

// engine_force without the advance: kinetic energy, prepare, shuffle,
// non-bonded tasks, bonds.
void force(Oracle &o) {
  if (!o.dpd_delta_set) {
    o.dpd_delta = dpd_delta_of(o);
    o.dpd_delta_set = true;
  }
  double total = 0;
  for (auto &t : o.types) t.kinetic_energy = 0;
  for (auto &c : o.cells)
    for (auto &p : c.parts) {
      float e = 0.5 * o.types[p.type].mass * (p.v[0] * p.v[0] + p.v[1] * p.v[1] + p.v[2] * p.v[2]);
      o.types[p.type].kinetic_energy += e;
      total += e;
    }
  o.ekin = total;
  const float self_nd = w_cubic(0, (float)o.cutoff);
  for (auto &c : o.cells) {
    c.epot = 0.0;
    for (auto &p : c.parts) {
      p.f[0] = p.f[1] = p.f[2] = 0.0f;
      p.f[3] = self_nd;
      p.ncount = 0;
    }
  }
  // space_shuffle_local: particles outside [0, h) (only after a load).
  float h[3] = {(float)o.h[0], (float)o.h[1], (float)o.h[2]};
  for (int cid = 0; cid < o.ncells; ++cid) {
    Cell &c = o.cells[cid];
    size_t pid = 0;
    while (pid < c.parts.size()) {
      Part &p = c.parts[pid];
      int delta[3];
      for (int k = 0; k < 3; k++)
        delta[k] = std::isgreaterequal(p.x[k], h[k]) - std::isless(p.x[k], 0.0);
      if (delta[0] || delta[1] || delta[2]) {
        for (int k = 0; k < 3; k++) {
          p.x[k] -= delta[k] * h[k];
          p.p0[k] -= delta[k] * h[k];
        }
        int dest = o.cellid((c.loc[0] + delta[0] + o.cdim[0]) % o.cdim[0],
                            (c.loc[1] + delta[1] + o.cdim[1]) % o.cdim[1],
                            (c.loc[2] + delta[2] + o.cdim[2]) % o.cdim[2]);
        o.cells[dest].incoming.push_back(p);
        size_t last = c.parts.size() - 1;
        if (pid < last) c.parts[pid] = c.parts[last];
        c.parts.pop_back();
      } else {
        pid++;
      }
    }
  }
  welcome_all(o);

  if (o.nthreads <= 1) {
    for (int t : o.order) run_task(o, o.tasks[t]);
  } else {
    run_tasks_threaded(o);
  }
  o.epot_bond = 0;
  o.epot_angle = o.epot_dihedral = o.epot_exclusion = 0;
  const bool terms = !o.exclusions.empty() || !o.angles.empty() || !o.dihedrals.empty();
  if (o.bonds.empty() && !terms) return;
  const std::vector<Loc> where = locate(o);
  terms_eval(o, where, 0);
  if (!o.bonds.empty()) bond_eval(o, where);
  terms_eval(o, where, 1);
}

AdvanceConst advance_const(const Oracle &o) {
  AdvanceConst K;
  K.dt = o.dt;
  for (int k = 0; k < 3; ++k) {
    K.h[k] = o.h[k];
    K.h2[k] = 2. * o.h[k];
    K.maxv[k] = (K.h[k] * o.max_frac) / K.dt;
    K.maxv2[k] = K.maxv[k] * K.maxv[k];
    K.maxx[k] = K.h[k] * o.max_frac;
    K.maxx2[k] = K.maxx[k] * K.maxx[k];
  }
  return K;
}

void advance(Oracle &o);

int step_rk4(Oracle &o);

int step(Oracle &o) {
  if (o.err != 0) return o.err;
  if (o.integrator == 1) return step_rk4(o);
  o.time += 1;
  o.update_pf = true;
  force(o);
  double epot = 0;
  for (auto &c : o.cells) epot += c.epot;
  o.epot_nb = epot;
  advance(o);
  return o.err;
}

void advance(Oracle &o) {
  AdvanceConst K = advance_const(o);
  if (o.nthreads <= 1) {
    for (int cid = 0; cid < o.ncells; ++cid) {
      advance_cell(o, K, cid);
      flux_integrate(o, cid);
    }
  } else {
    std::atomic<int> next{0};
    auto fn = [&]() {
      for (int c; (c = next.fetch_add(1)) < o.ncells;) {
        advance_cell(o, K, c);
        flux_integrate(o, c);
      }
    };
    std::vector<std::thread> th;
    for (int i = 0; i < o.nthreads; ++i) th.emplace_back(fn);
    for (auto &x : th) x.join();
  }
  welcome_all(o);
}

// FP64 evaluation of the forces on the current configuration: the same pair
// set as the FP32 loops (in-cutoff in FP32, reference orientation) and the
// same tables, but every subsequent operation in double.  Returns forces by
// id, the potential energy and, when asked, the virial tensor W = sum over
// pairs and bonds of (x_i - x_j) (x) f_i, f_i the force on i (row-major).
// The reference's virial_pair never sets its force (metrics.cpp:183-246),
// so this definition is the one the engine's virial is checked against.
void forces64(const Oracle &o, double *F /* 3 * nparts, by id */, double *epot_out, double *virial = nullptr) {
  std::fill(F, F + 3 * (size_t)o.nparts, 0.0);
  double epot = 0;
  double W[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  auto vir = [&](const double dx[3], const double fi[3]) {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) W[3 * a + b] += dx[a] * fi[b];
  };
  auto table64 = [](const Pot &p, double ri, double rj, double r2, double *e, double *f) {
    return table_ex64(p, ri, rj, r2, e, f);
  };
  auto pair64 = [&](const Part &pi, const Part &pj, const double dx[3], double r2) {
    int pot = o.potmat[pi.type * o.types.size() + pj.type];
    if (pot < 0) return;
    const Pot &p = o.pots[pot];
    if (r2 < (double)p.a * p.a) return;  // overlap path is random; skipped
    if (p.kind == kKindDPD) {
      double ri = pi.radius, rj = pj.radius;
      bool shifted = p.flags & kFlagShifted;
      double cut = shifted ? (p.b + ri + rj) : p.b;
      if (r2 > cut * cut) return;
      double r = std::sqrt(r2);
      double e[3] = {dx[0] / r, dx[1] / r, dx[2] / r};
      double v[3] = {(double)pi.v[0] - pj.v[0], (double)pi.v[1] - pj.v[1], (double)pi.v[2] - pj.v[2]};
      double sr = shifted ? r - ri - rj : r;
      double wc = sr < 0 ? 1.0 : 1 - sr / cut;
      double f = p.dpd_alpha * wc - p.dpd_gamma * wc * wc * (e[0] * v[0] + e[1] * v[1] + e[2] * v[2]);
      if (!o.dpd_counter_noise) f += p.dpd_sigma * wc * (double)dpd_delta_of(o);
      const double fi[3] = {f * e[0], f * e[1], f * e[2]};
      for (int k = 0; k < 3; ++k) {
        F[3 * (size_t)pi.id + k] += fi[k];
        F[3 * (size_t)pj.id + k] -= fi[k];
      }
      vir(dx, fi);
    } else {
      double e, f;
      if (!table64(p, pi.radius, pj.radius, r2, &e, &f)) return;
      const double fi[3] = {-f * dx[0], -f * dx[1], -f * dx[2]};
      for (int k = 0; k < 3; ++k) {
        F[3 * (size_t)pi.id + k] += fi[k];
        F[3 * (size_t)pj.id + k] -= fi[k];
      }
      vir(dx, fi);
      epot += e;
    }
  };
  for (const Task &t : o.tasks) {
    if (t.type == kTaskSort) continue;
    if (t.type == kTaskSelf) {
      const Cell &c = o.cells[t.i];
      for (size_t i = 0; i < c.parts.size(); ++i)
        for (size_t j = i + 1; j < c.parts.size(); ++j) {
          float dxf[3];
          float r2f = fr2(c.parts[i].x, c.parts[j].x, dxf);
          if (r2f > (float)o.cutoff2) continue;
          double dx[3];
          for (int k = 0; k < 3; ++k) dx[k] = (double)c.parts[i].x[k] - c.parts[j].x[k];
          pair64(c.parts[i], c.parts[j], dx, dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        }
    } else {
      int ci = t.i, cj = t.j;
      float shift[3];
      getsid(o, ci, cj, shift);
      double shift64[3];
      for (int k = 0; k < 3; ++k) {
        shift64[k] = o.cells[cj].origin[k] - o.cells[ci].origin[k];
        if (shift64[k] * 2 > o.dim[k])
          shift64[k] -= o.dim[k];
        else if (shift64[k] * 2 < -o.dim[k])
          shift64[k] += o.dim[k];
      }
      float cutoff = o.cutoff;
      float cutoff2 = cutoff * cutoff;
      const Cell &A = o.cells[ci], &B = o.cells[cj];
      for (const Part &pi : A.parts) {
        float pix[3] = {pi.x[0] - shift[0], pi.x[1] - shift[1], pi.x[2] - shift[2]};
        for (const Part &pj : B.parts) {
          float dxf[3];
          float r2f = fr2(pix, pj.x, dxf);
          if (r2f > cutoff2) continue;
          double dx[3];
          for (int k = 0; k < 3; ++k) dx[k] = ((double)pi.x[k] - shift64[k]) - pj.x[k];
          pair64(pi, pj, dx, dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        }
      }
    }
  }
  // single-body forces (persistent parts as stored)
  for (const Cell &c : o.cells)
    for (const Part &p : c.parts) {
      const Type &t = o.types[p.type];
      double *Fp = F + 3 * (size_t)p.id;
      if (t.sb == 1) {
        for (int k = 0; k < 3; ++k) Fp[k] += p.pf[k];
      } else if (t.sb == 2) {
        const double scale = (double)t.sb_coef * ((t.target_energy / t.kinetic_energy) - 1.0);
        for (int k = 0; k < 3; ++k) Fp[k] += scale * p.v[k];
      } else if (t.sb == 3) {
        for (int k = 0; k < 3; ++k) Fp[k] += t.cf[k];
      } else if (t.sb == 4) {
        const double v2 = (double)p.v[0] * p.v[0] + (double)p.v[1] * p.v[1] + (double)p.v[2] * p.v[2];
        for (int k = 0; k < 3; ++k) Fp[k] += -(double)t.sb_coef * v2 * p.v[k] + p.pf[k];
      }
    }
  // walls
  for (const Cell &c : o.cells)
    if (c.walls)
      for (const Part &p : c.parts) {
        double *Fp = F + 3 * (size_t)p.id;
        boundary_eval<double>(o, c, p, &epot, [&](int k, double w) { Fp[k] += w; });
      }
  // bonds
  std::vector<Loc> where = locate(o);
  for (const Bond &b : o.bonds) {
    if (!b.active) continue;
    const Part &pi = o.cells[where[b.i].cell].parts[where[b.i].idx];
    const Part &pj = o.cells[where[b.j].cell].parts[where[b.j].idx];
    const int *loci = o.cells[where[b.i].cell].loc, *locj = o.cells[where[b.j].cell].loc;
    double dx[3];
    for (int k = 0; k < 3; ++k) {
      int s = loci[k] - locj[k];
      if (s > 1)
        s = -1;
      else if (s < -1)
        s = 1;
      dx[k] = ((double)pi.x[k] + o.h[k] * s) - pj.x[k];
    }
    const Pot &pot = o.bpots[b.pot];
    double r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    double lo = (double)pot.a * pot.a, hi = (double)pot.b * pot.b;
    r2 = std::fmax(lo, std::fmin(hi, r2));
    double r = std::sqrt(r2);
    int ind = (int)std::fmax(0.0, pot.alpha[0] + r * (pot.alpha[1] + r * (double)pot.alpha[2]));
    const float *c = &pot.c[(size_t)ind * kChunk];
    double x = (r - c[0]) * (double)c[1];
    double ee = c[2] * x + c[3], eff = c[2];
    for (int k = 4; k < kChunk; ++k) {
      eff = eff * x + ee;
      ee = ee * x + c[k];
    }
    eff = eff * c[1] / r;
    if (eff >= b.dissociation) continue;  // the bond breaks in this force phase (bond.cpp:230)
    const double fi[3] = {-eff * dx[0], -eff * dx[1], -eff * dx[2]};
    for (int k = 0; k < 3; ++k) {
      F[3 * (size_t)pi.id + k] += fi[k];
      F[3 * (size_t)pj.id + k] -= fi[k];
    }
    vir(dx, fi);
    epot += ee;
  }
  auto add64 = [&](int id, int k, double w) { F[3 * (size_t)id + k] += w; };
  epot += exclusions_eval<double>(o, where, add64);
  epot += angles_eval<double>(o, where, add64);
  epot += dihedrals_eval<double>(o, where, add64);
  if (epot_out) *epot_out = epot;
  if (virial)
    for (int k = 0; k < 9; ++k) virial[k] = W[k];
}

Pot make_pot(uint32_t kind, uint32_t flags, const float alpha[4], float a, float b, float r0_plusone,
             int n, const float *c, float dpd_alpha, float dpd_gamma, float dpd_sigma) {
  Pot p;
  p.kind = kind;
  p.flags = flags;
  for (int k = 0; k < 4; ++k) p.alpha[k] = alpha ? alpha[k] : 0.0f;
  p.a = a;
  p.b = b;
  p.r0_plusone = r0_plusone;
  p.n = n;
  if (kind == kKindTable && c) p.c.assign(c, c + (size_t)(n + 1) * kChunk);
  p.dpd_alpha = dpd_alpha;
  p.dpd_gamma = dpd_gamma;
  p.dpd_sigma = dpd_sigma;
  return p;
}

// engine_advance_runge_kutta_4 (engine_advance.cpp:443-698), the
// non-Verlet branch: four engine_force calls; between them each particle's
// position is set from its start position p0 and the stage slope, checked
// with CHECK_TOOFAST (:424-431), and the next engine_force's
// space_shuffle_local moves it (and p0) to a neighbour cell if it left its
// own (space.cpp:333-393).  No velocity or displacement clamp, no frozen
// mask.  The arithmetic is Magnum Vector3 (float) arithmetic: scalar factors
// are converted to float first (0.5 * dt, dt / 6., the double imass).
int step_rk4(Oracle &o) {
  o.time += 1;
  const float dt = o.dt;
  const float half = (float)(0.5 * dt);
  const float sixth = (float)(dt / 6.);
  float h[3], h2[3];
  for (int k = 0; k < 3; ++k) {
    h[k] = (float)o.h[k];
    h2[k] = (float)(2. * o.h[k]);
  }
  auto toofast = [&](Part &p) {
    for (int k = 0; k < 3; ++k)
      if (p.x[k] >= h2[k] || p.x[k] <= -h[k]) {
        if (o.err == 0) {
          o.err = kErrToofast;
          o.err_pid = p.id;
        }
        return true;
      }
    return false;
  };
  // K1 at the current positions (refreshes persistent forces).
  o.update_pf = true;
  force(o);
  o.update_pf = false;
  for (auto &c : o.cells)
    for (auto &p : c.parts) {
      const float imf = (float)p.imass;
      const bool newton = o.types[p.type].dynamics == kNewtonian;
      for (int k = 0; k < 3; ++k) {
        if (newton) {
          p.vk[0][k] = p.f[k] * imf;
          p.xk[0][k] = p.v[k];
        } else {
          p.xk[0][k] = p.f[k] * imf;
        }
        p.p0[k] = p.x[k];
        p.v0[k] = p.v[k];
        p.x[k] = p.p0[k] + half * p.xk[0][k];
      }
      if (toofast(p)) return o.err;
    }
  // K2 and K3.
  for (int st = 1; st <= 2; ++st) {
    force(o);
    const float scale = st == 1 ? half : dt;  // position for the next stage
    for (auto &c : o.cells)
      for (auto &p : c.parts) {
        const float imf = (float)p.imass;
        const bool newton = o.types[p.type].dynamics == kNewtonian;
        for (int k = 0; k < 3; ++k) {
          if (newton) {
            p.vk[st][k] = p.f[k] * imf;
            p.xk[st][k] = p.v0[k] + half * p.vk[st - 1][k];
          } else {
            p.xk[st][k] = p.f[k] * imf;
          }
          p.x[k] = p.p0[k] + scale * p.xk[st][k];
        }
        if (toofast(p)) return o.err;
      }
  }
  // K4 and the update.
  force(o);
  double epot = 0;
  for (auto &c : o.cells) epot += c.epot;
  o.epot_nb = epot;
  for (int cid = 0; cid < o.ncells; ++cid) {
    Cell &c = o.cells[cid];
    size_t pid = 0;
    while (pid < c.parts.size()) {
      Part *p = &c.parts[pid];
      const float imf = (float)p->imass;
      const bool newton = o.types[p->type].dynamics == kNewtonian;
      int delta[3];
      bool fast = false;
      for (int k = 0; k < 3; ++k) {
        if (newton) {
          p->vk[3][k] = imf * p->f[k];
          p->xk[3][k] = p->v0[k] + dt * p->vk[2][k];
          p->v[k] = p->v0[k] + sixth * (p->vk[0][k] + 2.0f * p->vk[1][k] + 2.0f * p->vk[2][k] + p->vk[3][k]);
        } else {
          p->xk[3][k] = imf * p->f[k];
        }
        p->x[k] = p->p0[k] + sixth * (p->xk[0][k] + 2.0f * p->xk[1][k] + 2.0f * p->xk[2][k] + p->xk[3][k]);
        delta[k] = std::isgreaterequal(p->x[k], h[k]) - std::isless(p->x[k], 0.0);
        fast = fast || (p->x[k] >= h2[k] || p->x[k] <= -h[k]);
      }
      if (fast && o.err == 0) {
        o.err = kErrToofast;
        o.err_pid = p->id;
      }
      if (delta[0] != 0 || delta[1] != 0 || delta[2] != 0) {
        for (int k = 0; k < 3; k++) p->x[k] -= delta[k] * h[k];
        int dest = o.cellid((c.loc[0] + delta[0] + o.cdim[0]) % o.cdim[0],
                            (c.loc[1] + delta[1] + o.cdim[1]) % o.cdim[1],
                            (c.loc[2] + delta[2] + o.cdim[2]) % o.cdim[2]);
        o.cells[dest].incoming.push_back(*p);
        size_t last = c.parts.size() - 1;
        if (pid < last) c.parts[pid] = c.parts[last];
        c.parts.pop_back();
      } else {
        pid += 1;
      }
    }
  }
  welcome_all(o);
  return o.err;
}


}  // namespace

// ---------------------------------------------------------------------------
// C interface for the tests and the CPU-baseline leg of bench.py.


extern "C" {

void *oracle_create(const double origin[3], const double dim[3], const int cells[3], double cutoff,
                    double dt, int nthreads) {
  Oracle *o = new Oracle();
  if (init_space(*o, origin, dim, cells, cutoff) != 0) {
    delete o;
    return nullptr;
  }
  o->dt = (float)dt;
  o->dt_d = dt;
  o->nthreads = nthreads < 1 ? 1 : nthreads;
  return o;
}

void oracle_free(void *h) { delete (Oracle *)h; }

int oracle_geometry(void *h, int cdim[3], double hh[3], int *ntasks) {
  Oracle *o = (Oracle *)h;
  for (int k = 0; k < 3; ++k) {
    cdim[k] = o->cdim[k];
    hh[k] = o->h[k];
  }
  *ntasks = (int)o->tasks.size();
  return 0;
}

void oracle_set_threads(void *h, int n) { ((Oracle *)h)->nthreads = n < 1 ? 1 : n; }
void oracle_set_max_dist_fraction(void *h, float f) { ((Oracle *)h)->max_frac = f; }

// Philox4x32-10 of the shared header, for known-answer checks.
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  mxg::u32x4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
  mxg::u32x4 r = mxg::philox4x32_10(c, key[0], key[1]);
  for (int k = 0; k < 4; ++k) out[k] = r.v[k];
}
void oracle_set_seed(void *h, uint64_t seed) {
  ((Oracle *)h)->seed_lo = (uint32_t)seed;
  ((Oracle *)h)->seed_hi = (uint32_t)(seed >> 32);
}
void oracle_set_dpd_counter_noise(void *h, int on) { ((Oracle *)h)->dpd_counter_noise = on; }
void oracle_set_time(void *h, uint64_t t) { ((Oracle *)h)->time = t; }

int oracle_add_type(void *h, double mass, float radius, int dynamics) {
  Oracle *o = (Oracle *)h;
  Type t;
  t.mass = mass;
  t.imass = 1.0 / mass;
  t.radius = radius;
  t.dynamics = dynamics;
  o->types.push_back(t);
  int nt = (int)o->types.size();
  std::vector<int> m(nt * nt, -1);
  for (int i = 0; i < nt - 1; ++i)
    for (int j = 0; j < nt - 1; ++j) m[i * nt + j] = o->potmat[i * (nt - 1) + j];
  o->potmat.swap(m);
  rebuild_flux_pairs(*o);
  return nt - 1;
}

// kind: 1 gaussian p = (mean, std, duration); 2 berendsen p = (tau);
// 3 constant p = F; 4 friction p = (mean, std, duration, coef).
int oracle_set_single_body(void *h, int type, int kind, const float *p) {
  Oracle *o = (Oracle *)h;
  Type &t = o->types[type];
  t.sb = kind;
  t.has_random = kind == 1 || kind == 4;
  if (t.has_random) {
    t.rnd_mean = p[0];
    t.rnd_std = p[1];
    t.rnd_steps = (unsigned)std::ceil(p[2] / o->dt_d);
    if (t.rnd_steps == 0) t.rnd_steps = 1;
    if (kind == 4) t.sb_coef = p[3];
  } else if (kind == 2) {
    t.sb_coef = 1 / p[0];
  } else if (kind == 3) {
    for (int k = 0; k < 3; ++k) t.cf[k] = p[k];
  }
  return 0;
}

int oracle_set_random_force(void *h, int type, float mean, float std, float duration) {
  const float p[3] = {mean, std, duration};
  return oracle_set_single_body(h, type, 1, p);
}

void oracle_set_target_energy(void *h, int type, double e) { ((Oracle *)h)->types[type].target_energy = e; }

int oracle_add_pot(void *h, int ti, int tj, uint32_t kind, uint32_t flags, const float alpha[4], float a,
                   float b, float r0_plusone, int n, const float *c, float dpd_alpha, float dpd_gamma,
                   float dpd_sigma) {
  Oracle *o = (Oracle *)h;
  o->pots.push_back(make_pot(kind, flags, alpha, a, b, r0_plusone, n, c, dpd_alpha, dpd_gamma, dpd_sigma));
  int nt = (int)o->types.size();
  o->potmat[ti * nt + tj] = (int)o->pots.size() - 1;
  o->potmat[tj * nt + ti] = (int)o->pots.size() - 1;
  return 0;
}

// The six sides (MxBoundaryConditions_Init); rebuilds the space, so it
// must come before the particles are loaded.
int oracle_set_boundaries(void *h, const int kind[6], const float *vel, const float *restore, const float *radius) {
  Oracle *o = (Oracle *)h;
  if (o->nparts > 0) return kErrRange;
  for (int side = 0; side < 6; ++side) {
    o->bc_kind[side] = kind[side];
    for (int k = 0; k < 3; ++k) o->bc_vel[side][k] = vel ? vel[3 * side + k] : 0.0f;
    o->bc_restore[side] = restore ? restore[side] : 1.0f;
    o->bc_radius[side] = radius ? radius[side] : 0.0f;
  }
  for (int a = 0; a < 3; ++a) o->periodic[a] = (o->bc_kind[2 * a] & kBcPeriodic) != 0;
  double origin[3], dim[3];
  int cells[3];
  for (int k = 0; k < 3; ++k) {
    origin[k] = o->init_origin[k];
    dim[k] = o->init_dim[k];
    cells[k] = o->init_cells[k];
  }
  return init_space(*o, origin, dim, cells, o->cutoff);
}

int oracle_set_boundary_pot(void *h, int side, int type, uint32_t kind, uint32_t flags, const float alpha[4], float a,
                            float b, float r0_plusone, int n, const float *c, float dpd_alpha, float dpd_gamma,
                            float dpd_sigma) {
  Oracle *o = (Oracle *)h;
  o->wpots.push_back(make_pot(kind, flags, alpha, a, b, r0_plusone, n, c, dpd_alpha, dpd_gamma, dpd_sigma));
  o->wpot_side.push_back(side);
  o->wpot_type.push_back(type);
  return 0;
}

int oracle_add_bond_pot(void *h, const float alpha[4], float a, float b, int n, const float *c) {
  Oracle *o = (Oracle *)h;
  o->bpots.push_back(make_pot(kKindTable, 0, alpha, a, b, 0.0f, n, c, 0, 0, 0));
  return (int)o->bpots.size() - 1;
}

// Adds particles as space_addpart does: cell index (x - origin) * ih in
// double, local coordinate x - cell origin.  Particle ids are 0..N-1 in
// argument order.
int oracle_load(void *h, int N, const double *x, const double *v, const int *type, const uint16_t *flags) {
  Oracle *o = (Oracle *)h;
  for (auto &c : o->cells) c.parts.clear();
  o->nparts = N;
  for (int i = 0; i < N; ++i) {
    Part p;
    std::memset(&p, 0, sizeof(p));
    const Type &t = o->types[type[i]];
    p.radius = t.radius;
    p.mass = t.mass;
    p.imass = t.imass;
    p.id = i;
    p.type = (int16_t)type[i];
    p.flags = flags ? flags[i] : 0;
    for (int k = 0; k < 3; ++k) p.v[k] = (float)v[3 * (size_t)i + k];
    int ind[3];
    for (int k = 0; k < 3; ++k) ind[k] = (x[3 * (size_t)i + k] - o->origin[k]) * o->ih[k];
    for (int k = 0; k < 3; ++k)
      if (ind[k] < 0 || ind[k] >= o->cdim[k]) return kErrRange;
    Cell &c = o->cells[o->cellid(ind[0], ind[1], ind[2])];
    for (int k = 0; k < 3; ++k) p.x[k] = x[3 * (size_t)i + k] - c.origin[k];
    c.parts.push_back(p);
  }
  return 0;
}

// Loads a cell-local state as the engine reports it: particle i (id i) in
// cell[i] at local position xl[i] (no rounding through global coordinates),
// velocity v[i] and persistent random force pf[i] (may be NULL).
int oracle_load_local(void *h, int N, const float *xl, const int *cell, const float *v, const int *type,
                      const uint16_t *flags, const float *pf) {
  Oracle *o = (Oracle *)h;
  for (auto &c : o->cells) c.parts.clear();
  o->nparts = N;
  for (int i = 0; i < N; ++i) {
    if (cell[i] < 0 || cell[i] >= (int)o->cells.size()) return kErrRange;
    Part p;
    std::memset(&p, 0, sizeof(p));
    const Type &t = o->types[type[i]];
    p.radius = t.radius;
    p.mass = t.mass;
    p.imass = t.imass;
    p.id = i;
    p.type = (int16_t)type[i];
    p.flags = flags ? flags[i] : 0;
    for (int k = 0; k < 3; ++k) {
      p.v[k] = v[3 * (size_t)i + k];
      p.x[k] = xl[3 * (size_t)i + k];
      if (pf) p.pf[k] = pf[3 * (size_t)i + k];
    }
    o->cells[cell[i]].parts.push_back(p);
  }
  return 0;
}

int oracle_set_bonds(void *h, int M, const int *bi, const int *bj, const int *pot, const double *dissoc) {
  Oracle *o = (Oracle *)h;
  o->bonds.clear();
  for (int b = 0; b < M; ++b) o->bonds.push_back(Bond{bi[b], bj[b], pot[b], dissoc[b], true});
  return 0;
}

int oracle_set_species(void *h, int type, int spec_const) {
  Oracle *o = (Oracle *)h;
  if (type < 0 || type >= (int)o->types.size()) return kErrRange;
  o->types[type].spec_const = spec_const;
  return 0;
}

int oracle_add_flux(void *h, int kind, int ta, int tb, int ia, int ib, float coef, float decay, float target) {
  Oracle *o = (Oracle *)h;
  o->fluxes.push_back(FluxEntry{kind, ta, tb, ia, ib, coef, decay, target});
  rebuild_flux_pairs(*o);
  return 0;
}

// Species values by id (4 per particle); the accumulators start at zero.
int oracle_set_species_values(void *h, const float *v) {
  Oracle *o = (Oracle *)h;
  for (auto &c : o->cells)
    for (auto &p : c.parts)
      for (int k = 0; k < 4; ++k) {
        p.sv[k] = v[4 * (size_t)p.id + k];
        p.sq[k] = 0;
      }
  return 0;
}

int oracle_get_species(void *h, float *v, float *q) {
  Oracle *o = (Oracle *)h;
  for (auto &c : o->cells)
    for (auto &p : c.parts)
      for (int k = 0; k < 4; ++k) {
        if (v) v[4 * (size_t)p.id + k] = p.sv[k];
        if (q) q[4 * (size_t)p.id + k] = p.sq[k];
      }
  return 0;
}

int oracle_add_angle_pot(void *h, const float alpha[4], float a, float b, int n, const float *c) {
  Oracle *o = (Oracle *)h;
  o->apots.push_back(make_pot(kKindTable, 0, alpha, a, b, 0.0f, n, c, 0, 0, 0));
  return (int)o->apots.size() - 1;
}

int oracle_add_dihedral_pot(void *h, const float alpha[4], float a, float b, int n, const float *c) {
  Oracle *o = (Oracle *)h;
  o->dpots.push_back(make_pot(kKindTable, 0, alpha, a, b, 0.0f, n, c, 0, 0, 0));
  return (int)o->dpots.size() - 1;
}

int oracle_set_angles(void *h, int M, const int *i, const int *j, const int *k, const int *pot) {
  Oracle *o = (Oracle *)h;
  o->angles.clear();
  for (int a = 0; a < M; ++a) o->angles.push_back(Angle{i[a], j[a], k[a], pot[a]});
  return 0;
}

int oracle_set_dihedrals(void *h, int M, const int *i, const int *j, const int *k, const int *l, const int *pot) {
  Oracle *o = (Oracle *)h;
  o->dihedrals.clear();
  for (int a = 0; a < M; ++a) o->dihedrals.push_back(Dihedral{i[a], j[a], k[a], l[a], pot[a]});
  return 0;
}

int oracle_set_exclusions(void *h, int M, const int *i, const int *j) {
  Oracle *o = (Oracle *)h;
  o->exclusions.clear();
  for (int a = 0; a < M; ++a) o->exclusions.push_back(Exclusion{i[a], j[a]});
  return 0;
}

// Energies of the last force evaluation: angles, dihedrals, exclusions.
void oracle_term_energies(void *h, double *out3) {
  Oracle *o = (Oracle *)h;
  out3[0] = o->epot_angle;
  out3[1] = o->epot_dihedral;
  out3[2] = o->epot_exclusion;
}

void oracle_set_integrator(void *h, int kind) { ((Oracle *)h)->integrator = kind; }

int oracle_step(void *h, int nsteps) {
  Oracle *o = (Oracle *)h;
  int rc = 0;
  for (int s = 0; s < nsteps; ++s) {
    rc = step(*o);
    if (rc != 0) break;
  }
  return rc;
}

// Force evaluation only (engine_force without the advance), so forces can be
// compared on a fixed configuration.
int oracle_force(void *h) {
  Oracle *o = (Oracle *)h;
  o->time += 1;
  o->update_pf = true;
  force(*o);
  double epot = 0;
  for (auto &c : o->cells) epot += c.epot;
  o->epot_nb = epot;
  return 0;
}

// One step whose forces are replaced by the given ones (3 floats per id)
// after the force evaluation: checks the advance and the migration rule on
// identical inputs.
int oracle_step_with_forces(void *h, const float *f) {
  Oracle *o = (Oracle *)h;
  if (o->err != 0) return o->err;
  o->time += 1;
  force(*o);
  double epot = 0;
  for (auto &c : o->cells) epot += c.epot;
  o->epot_nb = epot;
  for (auto &c : o->cells)
    for (auto &p : c.parts)
      for (int k = 0; k < 3; ++k) p.f[k] = f[3 * (size_t)p.id + k];
  advance(*o);
  return o->err;
}

// Per-particle state by id: local position, cell id, velocity, force,
// in-cutoff pair count of the last force evaluation, persistent force.
int oracle_get_state(void *h, float *xl, int *cell, float *v, float *f, int *ncount, float *pf) {
  Oracle *o = (Oracle *)h;
  for (int c = 0; c < o->ncells; ++c)
    for (const Part &p : o->cells[c].parts) {
      size_t i = p.id;
      for (int k = 0; k < 3; ++k) {
        if (xl) xl[3 * i + k] = p.x[k];
        if (v) v[3 * i + k] = p.v[k];
        if (f) f[3 * i + k] = p.f[k];
        if (pf) pf[3 * i + k] = p.pf[k];
      }
      if (cell) cell[i] = c;
      if (ncount) ncount[i] = p.ncount;
    }
  return 0;
}

int oracle_number_density(void *h, float *nd) {
  Oracle *o = (Oracle *)h;
  for (const Cell &c : o->cells)
    for (const Part &p : c.parts) nd[p.id] = p.f[3];
  return 0;
}

// Per-cell particle ids in storage order (count per cell, then ids).
int oracle_cell_counts(void *h, int *counts) {
  Oracle *o = (Oracle *)h;
  for (int c = 0; c < o->ncells; ++c) counts[c] = (int)o->cells[c].parts.size();
  return 0;
}

int oracle_global_positions(void *h, double *x) {
  Oracle *o = (Oracle *)h;
  for (const Cell &c : o->cells)
    for (const Part &p : c.parts)
      for (int k = 0; k < 3; ++k) x[3 * (size_t)p.id + k] = c.origin[k] + p.x[k];
  return 0;
}

int oracle_observables(void *h, double *epot_nb, double *epot_bond, double *ekin) {
  Oracle *o = (Oracle *)h;
  if (epot_nb) *epot_nb = o->epot_nb;
  if (epot_bond) *epot_bond = o->epot_bond;
  if (ekin) *ekin = o->ekin;
  return 0;
}

int oracle_bond_active(void *h, uint8_t *active) {
  Oracle *o = (Oracle *)h;
  for (size_t b = 0; b < o->bonds.size(); ++b) active[b] = o->bonds[b].active ? 1 : 0;
  return 0;
}

// Virial tensor of the FP64 force phase (pairs and bonds).
int oracle_virial64(void *h, double *W9) {
  Oracle *o = (Oracle *)h;
  std::vector<double> F(3 * (size_t)std::max(o->nparts, 1));
  forces64(*o, F.data(), nullptr, W9);
  return 0;
}

int oracle_forces64(void *h, double *F, double *epot) {
  forces64(*(Oracle *)h, F, epot);
  return 0;
}

int oracle_error(void *h, int *pid) {
  Oracle *o = (Oracle *)h;
  if (pid) *pid = o->err_pid;
  return o->err;
}

}  // extern "C"
