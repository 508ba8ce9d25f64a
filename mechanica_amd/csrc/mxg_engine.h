// Host side of the MI355X engine: geometry, device buffers and the step
// driver.  One Engine holds one or more independent replica universes that
// share particle types and potentials; a single universe is the case R = 1.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mx_potential.h"

namespace mxg {

// Error codes mirror mdcore's engine_err_* (src/mdcore/include/engine.h:33-63).
enum : int {
  kErrOk = 0,
  kErrNull = -1,
  kErrMalloc = -2,
  kErrSpace = -3,
  kErrRange = -6,
  kErrCell = -7,
  kErrBond = -11,
  kErrPotential = -17,
  kErrDevice = -21,  // engine_err_cuda: a device call failed
  kErrMaxParts = -24,
  kErrCutoff = -27,
  kErrToofast = -29,
};

enum : int {
  kIntegratorEuler = 0,
  kIntegratorRK4 = 1,
};

enum : int { kDynNewtonian = 0, kDynOverdamped = 1 };

// BoundaryConditionKind (MxBoundaryConditions.hpp:15-23).
enum : int {
  kBoundaryVelocity = 1 << 0,
  kBoundaryPeriodic = 1 << 1,
  kBoundaryFreeslip = 1 << 2,
  kBoundaryPotential = 1 << 3,
  kBoundaryActive = kBoundaryFreeslip | kBoundaryVelocity | kBoundaryPotential,
  // BOUNDARY_NO_SLIP: accepted by set_boundaries as a velocity wall at rest
  // (MxBoundaryConditions.cpp:415-418), never stored.
  kBoundaryNoSlip = 1 << 4,
};

// Particle flags understood by the advance (MxParticle.h:58-60,
// PARTICLE_FROZEN_X/Y/Z; bit 2 is PARTICLE_BOUND, which the advance ignores).
enum : uint16_t { kFrozenX = 1 << 3, kFrozenY = 1 << 4, kFrozenZ = 1 << 5 };
// Internal flag (the reference's flags use bits 0-6): the load's local
// shuffle, which stands for the first step's, left a coordinate at the cell
// width (a coordinate just below the cell's lower face, moved down a cell),
// so the first advance starts from there instead of shuffling again.
enum : uint16_t { kShuffledToFace = 1 << 15 };

struct ReplicaSpec {
  double origin[3];
  double dim[3];
  int cells[3];
};

struct HostReplica {
  double origin[3], dim[3], h[3], ih[3];
  int cdim[3], span[3];
  int cell_base = 0, ncells = 0;
  int slot_base = 0, nparts = 0;
  int gid_base = 0;  // first gid of the replica's particles
  int tile_begin = 0, tile_end = 0;
  std::vector<double> corg[3];  // cell origins per axis
};

struct DeviceState;  // device buffers (defined in mxg_engine.hip)

// Default seed of replica r under the engine seed: seed itself for replica 0
// (a single universe), then Weyl steps of the golden ratio.
inline uint64_t replica_seed_default(uint64_t seed, int r) { return seed + (uint64_t)r * 0x9E3779B97F4A7C15ull; }

class Engine {
 public:
  Engine();
  ~Engine();
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;

  int init(const ReplicaSpec *reps, int nrep, double cutoff, double dt);
  int add_type(double mass, float radius, int dynamics);
  int add_pot(const Potential &p, int i, int j);
  int set_random_force(int type, float mean, float std, float duration);
  // Single-body force of a type (kind: 1 gaussian p = mean, std, duration;
  // 2 berendsen p = tau; 3 constant p = F; 4 friction p = mean, std,
  // duration, coef).
  int set_single_body(int type, int kind, const float *p);
  int set_target_energy(int type, double e);
  // Seed of the counter-based generators.  Replica r draws from its own
  // seed (replica_seed_default(seed, r) unless set_replica_seed overrides
  // it); a particle's streams are keyed by its replica's seed and its rank
  // among that replica's particles in load order, so a replica evolves as a
  // standalone engine holding the same particles with that seed.
  int set_seed(uint64_t seed);
  int set_replica_seed(int replica, uint64_t seed);
  int replica_seed(int replica, uint64_t *seed) const;
  int set_dpd_counter_noise(int on);
  int set_max_dist_fraction(float f);
  // engine_load: particle k has id pid[k] (k when pid is null); bonds and
  // terms name particles by these ids, and per-particle read-backs are in
  // load order.
  int load(int N, const double *x, const double *v, const int *type, const int *pid, const int *replica,
           const uint16_t *flags);
  int set_bonds(int M, const int *bi, const int *bj, const Potential *const *pots,
                const double *dissociation);
  // Angles (engine_angle_add, angle.h:50-60), dihedrals (engine_dihedral_add,
  // dihedral.h:38-46) and exclusions (engine_exclusion_add, exclusion.h):
  // each call replaces the engine's whole set of that kind.
  int set_angles(int M, const int *i, const int *j, const int *k, const Potential *const *pots);
  int set_dihedrals(int M, const int *i, const int *j, const int *k, const int *l,
                    const Potential *const *pots);
  int set_exclusions(int M, const int *i, const int *j);
  // The six sides of the box (MxBoundaryConditions; sides 0 left, 1 right,
  // 2 front, 3 back, 4 bottom, 5 top): kind bits kBoundary*, the wall
  // velocities (3 per side) and restore factors of velocity walls, and the
  // radius each wall presents to potentials; null arrays take 0, 1 and 0.
  // An axis is periodic when both its sides are.
  int set_boundaries(const int kind[6], const float *vel, const float *restore, const float *radius);
  // Potential between a side and particles of a type (MxBoundaryCondition::
  // set_potential); null removes it.
  int set_boundary_potential(int side, int type, const Potential *p);
  // Species of a type (MxSpecies in a CStateVector): up to four per
  // particle; bit k of constant_mask keeps species k constant.
  int set_species(int type, int constant_mask);
  // Species values, four per particle by id; the flux accumulators restart
  // at zero.
  int set_species_values(const float *values);
  int get_species(float *values, float *q);
  // MxFluxes_FluxEx (Flux.cpp:162-211): kind 0 Fick, 1 secrete, 2 uptake
  // between species ia of type ta and ib of type tb.
  int add_flux(int kind, int ta, int tb, int ia, int ib, float k, float decay, float target);
  int step(int nsteps);
  int sync();

  // Per particle snapshots in load order; any pointer may be null.
  int get_state(float *xlocal, int *cell, float *v, float *f, float *pf, float *nd = nullptr);
  int get_global(double *x, double *v);
  int neighbour_counts(int *counts);
  int observables(int replica, double *out12);
  // Energies of the angles, dihedrals and exclusions of the last force
  // evaluation (space epot_angle, epot_dihedral, epot_exclusion); replica -1
  // sums the replicas.
  int term_energies(int replica, double *out3);
  int bond_active(uint8_t *active);
  int geometry(int replica, int cdim[3], double h[3]);

  int nparts() const { return N_; }
  int nreplicas() const { return (int)reps_.size(); }
  uint64_t time() const { return time_; }
  int error() const { return err_; }
  const std::string &message() const { return msg_; }

  int fail(int code, const std::string &msg);    // sticky: the engine stops
  int reject(int code, const std::string &msg);  // a refused request only

  // Replays captured HIP graphs of two steps (on by default).
  int set_graphs(int on);
  // 0: forward Euler (default), 1: Runge-Kutta 4.
  int set_integrator(int kind);
  int integrator() const { return integrator_; }

  // Adds the virial tensor to the per-step reductions.
  int set_virial(int on);

  // Average duration of the force kernel and of a whole step over the last
  // step() call (HIP events on the engine's stream); -1 when not recorded.
  int set_timing(int on);
  double force_kernel_ms() const { return force_ms_; }
  double step_ms() const { return step_ms_; }

  // Shape of the force kernel's work unit: a brick of bx*by*bz cells staged
  // with its halo in LDS, and the per-thread neighbour-list length.
  // set_brick(0, 0, 0, 0, 0) picks the largest brick whose halo fits; pmax
  // (> 0) lowers the LDS particle capacity.
  int set_brick(int bx, int by, int bz, int cap, int pmax);
  // Tries brick shapes and list capacities, steps_per steps each (with
  // bonds or terms, the four fastest again at four waves per SIMD), keeps
  // the fastest; *taken = steps advanced.
  int autotune(int steps_per, int *taken);
  // Waves per SIMD of the force kernel: 3, or 4 (used by the bonded modes).
  int set_waves(int waves);
  int waves() const { return waves_; }
  struct TuneEntry {
    int bx, by, bz, cap, waves;
    double force_ms;
  };
  const std::vector<TuneEntry> &tune_log() const { return tune_log_; }
  void brick(int dims[3], int *cap, int *pmax) const {
    for (int k = 0; k < 3; ++k) dims[k] = bdim_[k];
    *cap = cap_;
    *pmax = pmax_;
  }
  int block_size() const { return block_; }
  int set_block_size(int b);

 private:
  int build_geometry();
  int upload_potentials();
  int launch_step();
  int launch_substep(int rk_stage, int tick);
  int ensure_rk_buffers();
  int ensure_graph(int parity);
  int upload_terms();
  void fill_step_args(void *out, int parity) const;
  int refresh_args();
  int check_device(int hip_status, const char *what);

  std::vector<HostReplica> reps_;
  double cutoff_ = 1.0, dt_ = 0.01;
  float max_frac_ = 0.05f;
  uint64_t seed_ = 0x5eed5eedull;
  int dpd_noise_ = 0;

  struct TypeInfo {
    double mass, imass;
    float radius;
    int dynamics;
    int sb = 0;  // single-body force kind (0 none, 1 gaussian, 2 berendsen, 3 constant, 4 friction)
    float rnd_mean = 0, rnd_std = 0;
    unsigned rnd_steps = 1;
    float sb_coef = 0;
    float cf[3] = {0, 0, 0};
    double target_energy = 0;
    int spec_const = 0;
  };
  struct FluxInfo {
    int kind, ta, tb, ia, ib;
    float k, decay, target;
  };
  std::vector<FluxInfo> fluxes_;
  bool species_on_ = false;  // species buffers allocated: k_flux runs every step
  int ensure_species();
  std::vector<TypeInfo> types_;
  std::vector<Potential> pots_;
  std::vector<int> potmat_;  // ntypes * ntypes
  std::vector<Potential> bpots_;
  // Angles, dihedrals and exclusions: particle ids (4 per term, -1 unused)
  // and the index of the term's potential in apots_ or dpots_.
  struct HostTerm {
    int id[4];
    int pot;
  };
  std::vector<HostTerm> angles_, dihedrals_, exclusions_;
  std::vector<Potential> apots_, dpots_;
  // Particle ids.  Internally a particle is its gid: the replica's first
  // gid plus its rank among the replica's particles in load order (so each
  // replica holds a contiguous gid range, and gid - first is the key of its
  // random streams).  The caller's ids are pids (load indices without a pid
  // array).
  std::vector<int> prep_;           // replica of each gid
  std::vector<int> gid_of_load_;    // load index -> gid
  std::vector<int> load_of_gid_;    // gid -> load index
  std::vector<int> gid_of_pid_;     // pid -> gid, -1 for ids not loaded
  std::vector<int> pid_of_gid_;     // gid -> the caller's id
  int gid_of(int pid) const {
    return (pid >= 0 && pid < (int)gid_of_pid_.size()) ? gid_of_pid_[pid] : -1;
  }
  std::vector<uint64_t> rep_seed_;  // per replica
  int upload_replicas();
  struct WallInfo {
    int kind = kBoundaryPeriodic;
    float vel[3] = {0, 0, 0};
    float restore = 1.0f;
    float radius = 0.0f;
  };
  WallInfo walls_[6];
  std::vector<Potential> wpots_;                 // wall potentials
  std::vector<std::pair<int, int>> wpot_key_;   // (side, type) of each
  int upload_walls();
  int NT_ = 0;             // terms on the device
  int NTB_ = 0;            // blocks of their energy reduction

  int N_ = 0, M_ = 0, ncells_ = 0, ntiles_ = 0;
  int bdim_[3] = {4, 4, 4};
  size_t brick_tabs_ = 0;
  int integrator_ = 0;
  int cap_ = 12, pmax_ = 0, pmax_req_ = 0, hmax_ = 0, hdmax_ = 0, noff_ = 0, nbricks_ = 0;
  bool brick_dpd_ = false;
  bool brick_flux_ = false;
  std::vector<double> mean_cell_count_;
  int block_ = 256;
  bool any_dpd_ = false, any_table_ = false, any_random_ = false, any_sb_ = false, any_berendsen_ = false;
  bool potentials_dirty_ = true;
  uint64_t time_ = 0;
  int err_ = 0;
  std::vector<TuneEntry> tune_log_;
  std::string msg_;
  bool timing_ = false;
  int timed_steps_ = 0;
  double force_ms_ = -1, step_ms_ = -1;
  bool virial_ = false;
  bool pid_prev_valid_ = false;  // pid[1] holds the order forces were computed in
  int parity_ = 0;               // which ping-pong ids/offsets/bond flags are current
  bool args_dirty_ = true;        // device copy of the kernel arguments is stale
  bool slots_valid_ = true;       // device slot_of matches the current slot order
  int waves_ = 3;                 // force kernel waves per SIMD (4: bonded modes only)
  bool use_graphs_ = true;
  uint64_t config_gen_ = 1;       // bumped when the captured launches change
  uint64_t graph_gen_[2] = {0, 0};
  DeviceState *d_ = nullptr;
};

}  // namespace mxg
