/*
 * mxg.h — C ABI of the MI355X engine for mdcore's force-and-advance step.
 *
 * Plain C types only (pointers, sizes, scalars).  Every function returns 0
 * or a negative error code that mirrors mdcore's engine_err_* table
 * (src/mdcore/include/engine.h:33-63); the engine keeps the first error
 * sticky, as mdcore's engine_err does (MxUniverse.cpp:551-559), and refuses
 * further steps until it is destroyed.  Each declaration names the reference
 * interface it replaces.  Paths are relative to the reference repository.
 */
#ifndef MXG_H_
#define MXG_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes (engine.h:33-63). */
#define MXG_ERR_OK 0
#define MXG_ERR_NULL (-1)
#define MXG_ERR_MALLOC (-2)
#define MXG_ERR_SPACE (-3)
#define MXG_ERR_RANGE (-6)
#define MXG_ERR_BOND (-11)
#define MXG_ERR_POTENTIAL (-17)
#define MXG_ERR_DEVICE (-21)
#define MXG_ERR_MAXPARTS (-24)
#define MXG_ERR_CUTOFF (-27)
#define MXG_ERR_TOOFAST (-29)

/* Particle dynamics (MxParticle.h PARTICLE_NEWTONIAN / PARTICLE_OVERDAMPED). */
#define MXG_NEWTONIAN 0
#define MXG_OVERDAMPED 1

/* Particle flags honoured by the advance (MxParticle.h:58-60,
 * PARTICLE_FROZEN_X/Y/Z); the other MxParticleFlags bits are carried and
 * ignored. */
#define MXG_FROZEN_X (1 << 3)
#define MXG_FROZEN_Y (1 << 4)
#define MXG_FROZEN_Z (1 << 5)

/* Potential kinds and flags (MxPotential.h:55-110). */
#define MXG_POTENTIAL_KIND_TABLE 0
#define MXG_POTENTIAL_KIND_DPD 1
#define MXG_POTENTIAL_SCALED (1u << 11)
#define MXG_POTENTIAL_SHIFTED (1u << 12)

typedef struct mxg_potential mxg_potential;
typedef struct mxg_engine mxg_engine;

typedef struct mxg_potential_info {
  uint32_t kind;
  uint32_t flags;
  int32_t n;           /* number of intervals (table potentials) */
  float alpha[4];      /* interval map coefficients */
  float a, b;          /* domain */
  float r0_plusone;    /* shift of SHIFTED potentials */
  float dpd_alpha, dpd_gamma, dpd_sigma;
} mxg_potential_info;

/* ---- Potentials (host-side table construction) ------------------------ */

/* potential_create_LJ126 (src/mdcore/src/MxPotential.cpp:976-991), Python
 * m.Potential.lennard_jones_12_6 (:2459-2473). */
mxg_potential *mxg_potential_create_lj126(double min, double max, double A, double B, double tol);

/* potential_create_morse (MxPotential.cpp:3395-3424), m.Potential.morse
 * (:2652-2667). */
mxg_potential *mxg_potential_create_morse(double d, double a, double r0, double min, double max,
                                          double tol);

/* potential_create_harmonic (MxPotential.cpp:348-366), m.Potential.harmonic
 * (:2545-2559). */
mxg_potential *mxg_potential_create_harmonic(double min, double max, double K, double r0, double tol);

/* DPDPotential_New (DissapativeParticleDynamics.cpp:134-148). */
/* potential_create_harmonic_angle (MxPotential.cpp:555-597): K (theta -
 * theta0)^2 for angles in [min, max] (clamped to [0, pi]), tabulated in
 * cos(theta) for angle_eval. */
mxg_potential *mxg_potential_create_harmonic_angle(double min, double max, double K, double theta0,
                                                   double tol);

/* potential_create_harmonic_dihedral (MxPotential.cpp:497-517):
 * K (1 + cos(n phi - delta)), tabulated in cos(phi) for dihedral_eval. */
mxg_potential *mxg_potential_create_harmonic_dihedral(double K, int n, double delta, double tol);

mxg_potential *mxg_potential_create_dpd(float alpha, float gamma, float sigma, float cutoff, int shifted);

/* Py_DECREF of an MxPotential. */
void mxg_potential_free(mxg_potential *p);

/* Last error of the potential constructors in this thread (potential_err,
 * MxPotential.h:31-36). */
int mxg_potential_error(void);

/* Fields of struct MxPotential (MxPotential.h:121-155). */
int mxg_potential_get_info(const mxg_potential *p, mxg_potential_info *out);

/* The (n+1)*8 table coefficients (MxPotential.h:131); returns the count or a
 * negative code if capacity is too small. */
int mxg_potential_coefficients(const mxg_potential *p, float *c, int capacity);

/* potential_eval (potential_eval.hpp:67-104). */
int mxg_potential_eval(const mxg_potential *p, float r2, float *e, float *f);

/* potential_eval_ex (potential_eval.hpp:106-161); returns 1 in range, 0 not. */
/* potential_eval_r (potential_eval.hpp:180-208): the table at r itself (no
 * square root), f = dE/dr; r must lie in the potential's domain. */
int mxg_potential_eval_r(const mxg_potential *p, float r, float *e, float *f);

int mxg_potential_eval_ex(const mxg_potential *p, float ri, float rj, float r2, float *e, float *f);

/* ---- Engine ------------------------------------------------------------ */

/* engine_init (engine.h:600-601, engine.cpp:1502-1679) with the periodic
 * boundary default and universe_init's dt (MxSimulator.cpp:349-422). */
int mxg_engine_init(mxg_engine **e, const double origin[3], const double dim[3], const int cells[3],
                    double cutoff, double dt);

/* A batch of independent periodic universes stepped together (no reference
 * counterpart; each replica behaves as its own engine_init). origins, dims
 * and cells hold 3 values per replica. */
int mxg_ensemble_init(mxg_engine **e, int nrep, const double *origins, const double *dims, const int *cells,
                      double cutoff, double dt);

/* engine_finalize. */
void mxg_engine_free(mxg_engine *e);

/* engine_addtype (engine.h:573-574); returns the type id. */
int mxg_engine_addtype(mxg_engine *e, double mass, float radius, int dynamics);

/* engine_addpot (engine.h:440, engine.cpp:960-981). */
int mxg_engine_addpot(mxg_engine *e, const mxg_potential *p, int i, int j);

/* engine_add_singlebody_force with random_create (MxForce.cpp:370-407):
 * a persistent random force of magnitude N(mean, std) in a uniform
 * direction, refreshed every ceil(duration/dt) steps. */
int mxg_engine_add_random_force(mxg_engine *e, int type, float mean, float std, float duration);

/* The other single-body forces of MxForce.cpp (one per type; a later one
 * replaces the type's force, as engine_add_singlebody_force does):
 * berenderson_create(tau) (MxForce.cpp:386-393, force :330-337: f += v/tau
 * (E_target/E_kin(type) - 1) with E_kin the type's kinetic energy at the
 * start of the step and E_target set by mxg_engine_set_target_energy,
 * MxParticleType::target_energy), the constant force (:339-344, f += F) and
 * friction_create(coef, mean, std, durration) (:408-420, force :352-367:
 * f += -coef |v|^2 v + a persistent random vector). */
int mxg_engine_add_force_berendsen(mxg_engine *e, int type, float tau);
int mxg_engine_add_force_constant(mxg_engine *e, int type, const float force[3]);
int mxg_engine_add_force_friction(mxg_engine *e, int type, float coef, float mean, float std, float duration);
int mxg_engine_set_target_energy(mxg_engine *e, int type, double energy);

/* Seed of the counter-based generators (no reference counterpart: the
 * reference uses a global Mersenne twister).  Replica r draws from its own
 * seed, seed + r * 0x9E3779B97F4A7C15 (mod 2^64) unless set with
 * mxg_engine_set_replica_seed; a particle's random streams (persistent
 * random force, DPD counter noise, overlap direction) are keyed by its
 * replica's seed and its rank among the replica's particles in load order,
 * so a replica evolves as a standalone engine loaded with the same particles
 * and that seed.  Setting the engine seed resets every replica seed. */
int mxg_engine_set_seed(mxg_engine *e, uint64_t seed);
int mxg_engine_set_replica_seed(mxg_engine *e, int replica, uint64_t seed);
int mxg_engine_replica_seed(mxg_engine *e, int replica, uint64_t *seed);

/* 0: DPD random term as the reference computes it (sigma*omega/sqrt(dt),
 * dpd_eval.hpp:57); 1: a zero-mean, unit-variance uniform number per pair
 * and step from a counter-based generator. */
int mxg_engine_set_dpd_noise(mxg_engine *e, int mode);

/* engine::particle_max_dist_fraction (engine.h:383). */
int mxg_engine_set_max_dist_fraction(mxg_engine *e, float fraction);

/* Accumulate the virial tensor sum r_ij (x) f_ij (metrics.cpp:41-167 has no
 * working counterpart). */
int mxg_engine_set_virial(mxg_engine *e, int on);

/* Work unit of the force kernel: bricks of bx*by*bz cells staged with their
 * halo in LDS, with a per-thread neighbour list of cap entries; bx = 0 picks
 * the largest brick that fits; pmax > 0 caps the LDS particle capacity
 * (bricks above it run from global memory).  No reference counterpart. */
int mxg_engine_set_brick(mxg_engine *e, int bx, int by, int bz, int cap, int pmax);
int mxg_engine_get_brick(mxg_engine *e, int dims[3], int *cap, int *pmax);

/* Integrator of the step: MXG_INTEGRATOR_EULER (engine_advance_forward_euler,
 * engine_advance.cpp:248-422) or MXG_INTEGRATOR_RK4 (engine_advance_runge_kutta_4,
 * engine_advance.cpp:443-698); replaces setting engine.integrator
 * (EngineIntegrator, MxSimulator.cpp:218-231). */
#define MXG_INTEGRATOR_EULER 0
#define MXG_INTEGRATOR_RK4 1
int mxg_engine_set_integrator(mxg_engine *e, int kind);

/* Autotuner: times the force kernel for each brick shape and list capacity
 * that fits (steps_per real steps each: the layouts give identical
 * per-particle results, so the trajectory is unchanged) and keeps the
 * fastest; with bonds or terms it also times the four fastest at four waves
 * per SIMD.  *taken = steps advanced.  mxg_engine_tune_log returns the
 * number of entries and copies up to max of them: layout (bx, by, bz, cap,
 * waves), 5 ints each, and the mean force-kernel time in ms.  No reference
 * counterpart. */
int mxg_engine_autotune(mxg_engine *e, int steps_per, int *taken);
int mxg_engine_tune_log(mxg_engine *e, int max, int *layout, double *force_ms);

/* Waves per SIMD of the force kernel: 3 (default) or 4.  Four applies to
 * steps with bonds or terms only; other steps run at three.  get returns the
 * setting.  No reference counterpart. */
int mxg_engine_set_waves(mxg_engine *e, int waves);
int mxg_engine_get_waves(mxg_engine *e);

/* Replay steps from captured HIP graphs (two steps per graph); on by
 * default.  No reference counterpart (mdcore drives pthread runners). */
int mxg_engine_set_graphs(mxg_engine *e, int on);

/* engine_load (engine.h:612-613): x, v hold 3N doubles; pid[k] is the id of
 * particle k (MxParticle::id; NULL: ids 0..N-1 in argument order), unique
 * and below 64 N + 65536 (partlist is sized by the largest id); bonds,
 * angles, dihedrals and exclusions name particles by these ids.  replica
 * and flags may be NULL.  Per-particle read-backs (state, unload, number
 * density, neighbour counts, species) are in argument order.  A load
 * replaces the particles and drops the previous bonds and terms.  The local
 * shuffle that begins the reference's next step (space_shuffle_local,
 * space.cpp:333-393) is applied at the load: a particle whose float local
 * coordinate rounds to the cell width (just below a face) is reported in
 * the next cell at 0. */
int mxg_engine_load(mxg_engine *e, int N, const double *x, const double *v, const int *type, const int *pid,
                    const int *replica, const uint16_t *flags);

/* engine_bond_alloc + _bond_init (engine_bonded.cpp:1317-1360,
 * bond.cpp:497-529): M harmonic-table bonds between ids i[b], j[b]. */
int mxg_engine_set_bonds(mxg_engine *e, int M, const int32_t *i, const int32_t *j,
                         const mxg_potential *const *pots, const double *dissociation_energy);

/* engine_angle_add (engine.h:574, angle.h:50-60): M angles (i, j, k) with
 * vertex j, each with a potential in cos(theta) such as
 * mxg_potential_create_harmonic_angle; replaces the previous set.  Evaluated
 * as angle_eval (angle.cpp:92-354). */
int mxg_engine_set_angles(mxg_engine *e, int M, const int32_t *i, const int32_t *j, const int32_t *k,
                          const mxg_potential *const *pots);

/* engine_dihedral_add (engine.h:576, dihedral.h:38-46): M dihedrals
 * (i, j, k, l) about the axis j-k, each with a potential in cos(phi) such as
 * mxg_potential_create_harmonic_dihedral; replaces the previous set.
 * Evaluated as dihedral_eval (dihedral.cpp:90-330). */
int mxg_engine_set_dihedrals(mxg_engine *e, int M, const int32_t *i, const int32_t *j, const int32_t *k,
                             const int32_t *l, const mxg_potential *const *pots);

/* engine_exclusion_add (engine.h:578, exclusion.h): M excluded pairs whose
 * non-bonded table interaction is subtracted again (exclusion_eval,
 * exclusion.cpp:83-260); replaces the previous set. */
int mxg_engine_set_exclusions(mxg_engine *e, int M, const int32_t *i, const int32_t *j);

/* Boundary kinds (BoundaryConditionKind, MxBoundaryConditions.hpp:15-23) and
 * sides (left, right: x low, high; front, back: y; bottom, top: z). */
#define MXG_BOUNDARY_VELOCITY (1 << 0)
#define MXG_BOUNDARY_PERIODIC (1 << 1)
#define MXG_BOUNDARY_FREESLIP (1 << 2)
#define MXG_BOUNDARY_POTENTIAL (1 << 3)
/* BOUNDARY_NO_SLIP: a velocity wall at rest (MxBoundaryConditions.cpp:415-418);
 * BOUNDARY_RESETTING (1 << 5) is not supported and refused. */
#define MXG_BOUNDARY_NO_SLIP (1 << 4)
#define MXG_SIDE_LEFT 0
#define MXG_SIDE_RIGHT 1
#define MXG_SIDE_FRONT 2
#define MXG_SIDE_BACK 3
#define MXG_SIDE_BOTTOM 4
#define MXG_SIDE_TOP 5

/* MxBoundaryConditions_Init (MxBoundaryConditions.cpp:496-700): the kind of
 * each of the six sides, with velocity[3 * side] the velocity of a
 * velocity wall (no-slip is a velocity wall at rest), restore[side] its
 * restore factor and radius[side] the radius the wall presents to
 * potentials (NULL: 0, 1, 0).  Both sides of an axis are periodic or
 * neither.  Walls act in the advance (apply_update_pos_vel,
 * boundary_eval.hpp:30-119) and, with a potential, in the force phase
 * (boundary_eval, boundary_eval.hpp:162-221). */
int mxg_engine_set_boundaries(mxg_engine *e, const int kind[6], const float *velocity, const float *restore,
                              const float *radius);

/* MxBoundaryCondition::set_potential (MxBoundaryConditions.cpp:771-790): the
 * potential between a side and particles of a type (a table or DPD
 * potential); NULL removes it. */
int mxg_engine_set_boundary_potential(mxg_engine *e, int side, int type, const mxg_potential *p);

/* Species of a particle type (MxSpecies in its CStateVector, whose layout
 * lives in a dependency absent from the reference tree): up to four per
 * particle; bit k of constant_mask marks species k constant (the
 * SPECIES_KONSTANT flag MxFluxes_Integrate tests, Flux.cpp:261-267). */
int mxg_engine_set_species(mxg_engine *e, int type, int constant_mask);

/* CStateVector fvec of every particle: four floats per particle by id; the
 * flux accumulators (q) restart at zero.  After mxg_engine_load. */
int mxg_engine_set_species_values(mxg_engine *e, const float *values);

/* Species values and pending flux accumulators, four each per particle by
 * id (either pointer may be NULL). */
int mxg_engine_get_species(mxg_engine *e, float *values, float *q);

/* MxFluxes_FluxEx (Flux.cpp:162-211; MxFluxes_Fick / _Secrete / _Uptake,
 * :214-259): kind 0 Fick, 1 secrete, 2 uptake between species species_a of
 * type type_a and species_b of type type_b, rate k, decay and target.
 * Evaluated per in-cutoff pair as flux_eval_ex (flux_eval.hpp:33-81) and
 * integrated after the advance (MxFluxes_Integrate, Flux.cpp:261-284). */
int mxg_engine_add_flux(mxg_engine *e, int kind, int type_a, int type_b, int species_a, int species_b, float k,
                        float decay, float target);

/* engine_step (engine.h:628, engine.cpp:1198-1228), nsteps times. */
int mxg_engine_step(mxg_engine *e, int nsteps);

/* Waits for the engine's stream. */
int mxg_engine_sync(mxg_engine *e);

/* engine_unload (engine.h:634-635): global positions and velocities by id. */
int mxg_engine_unload(mxg_engine *e, double *x, double *v);

/* Per-particle state by id: cell-local position (MxParticle::x), cell index,
 * velocity, last forces, persistent random force.  NULL skips an array. */
int mxg_engine_state(mxg_engine *e, float *xlocal, int *cell, float *v, float *f, float *pf);

/* Number density of each particle from the last step (MxParticle::f[3]:
 * space_prepare's self term plus W(r^2, cutoff) of every in-cutoff pair,
 * space.cpp:202-242, runner_dopair.cpp:194-203), by id. */
int mxg_engine_number_density(mxg_engine *e, float *nd);

/* Number of in-cutoff neighbours of each particle in the current
 * configuration (pairs runner_dopair/runner_doself evaluate). */
int mxg_engine_neighbour_counts(mxg_engine *e, int *counts);

/* Observables of the last step, replica r or the sum over replicas (r < 0):
 * out[0] non-bonded energy, out[1] bond energy, out[2] kinetic energy at
 * the start of the step (engine_kinetic_energy, engine.cpp:1698-1719),
 * out[3..11] virial tensor (row-major), all accumulated in FP64. */
int mxg_engine_observables(mxg_engine *e, int replica, double out[12]);

/* Energies of the last force evaluation's angles, dihedrals and exclusions
 * (space epot_angle, epot_dihedral, epot_exclusion; engine_bonded.cpp:
 * 1533-1538), replica r or the sum over replicas (r < 0), in FP64. */
int mxg_engine_term_energies(mxg_engine *e, int replica, double out[3]);

/* BOND_ACTIVE of each bond (bond.h:40-43). */
int mxg_engine_bond_active(mxg_engine *e, uint8_t *active);

/* Cell grid of a replica (space.h cdim/h). */
int mxg_engine_geometry(mxg_engine *e, int replica, int cdim[3], double h[3]);

/* engine::time (engine.h). */
int mxg_engine_time(mxg_engine *e, uint64_t *time);

/* Sticky error code and message of the engine (engine_err / errs_getstring). */
int mxg_engine_error(const mxg_engine *e);
const char *mxg_engine_error_message(const mxg_engine *e);

/* HIP-event timing of the force kernel and of the whole step over the last
 * mxg_engine_step call (engine timers, engine.h:111-135). */
int mxg_engine_set_timing(mxg_engine *e, int on);
int mxg_engine_timing(mxg_engine *e, double *force_ms, double *step_ms);

/* ------------------------------------------------------------------------
 * Adjoint: a differentiable overdamped system (Morse or LJ 12-6 pairs,
 * harmonic bonds) and the discrete adjoint of its trajectory.  No reference
 * counterpart: the forward map restates engine_step's overdamped update
 * (engine_advance.cpp:126-239) with the pair term of potential_eval_ex
 * (potential_eval.hpp:106-161) and the bond term of bond_eval
 * (bond.cpp:99-295) evaluated analytically, so that their parameter
 * derivatives are exact.  Buffers marked (device) are
 * device pointers; work runs on `stream` (a hipStream_t, 0 = default).
 * ------------------------------------------------------------------------ */
typedef struct mxg_adjoint mxg_adjoint;

/* N particles of ntypes (<= 4) types in a periodic box dim with a cells grid
 * (>= 3 per axis, cells no narrower than cutoff); mass and radius per type;
 * type per particle (host arrays).  max_dist_fraction: the overdamped
 * displacement clamp as a fraction of the cell width. */
int mxg_adjoint_create(mxg_adjoint **out, int N, const double dim[3], const int cells[3], double cutoff, double dt,
                       float max_dist_fraction, int ntypes, const double *mass, const double *radius,
                       const int *type);
void mxg_adjoint_free(mxg_adjoint *a);
const char *mxg_adjoint_error_message(const mxg_adjoint *a);

/* Potential of a type pair (default Morse): 0 Morse, d (e^{-2a(s-1)} -
 * 2 e^{-a(s-1)}) with s = max(r, smin) - (R_i + R_j) + r0 + 1
 * (potential_create_morse, MxPotential.cpp:3395-3424); 1 Lennard-Jones 12-6,
 * A s^-12 - B s^-6 with s = max(r, smin) (potential_create_LJ126,
 * MxPotential.cpp:976-991); on its own domain [smin, smax] (the creator's
 * min and max) when smax > smin, else on the domain the forward and
 * backward calls pass. */
int mxg_adjoint_set_pair_kind(mxg_adjoint *a, int ti, int tj, int kind, double smin, double smax);

/* Harmonic bonds (potential_create_harmonic, MxPotential.cpp:348-366, as
 * bond_eval evaluates it, bond.cpp:99-295: r clamped to [rmin, rmax],
 * E = K (r - r0)^2, no dissociation) between particle ids i[b], j[b];
 * replaces the previous set (M = 0 removes them).  With bonds, theta and
 * dL/dtheta carry one more row (K, r0, unused) after the type pairs. */
int mxg_adjoint_set_bonds(mxg_adjoint *a, int M, const int32_t *i, const int32_t *j, double rmin, double rmax);

/* theta (host): ntypes*(ntypes+1)/2 rows, one per type pair ti <= tj in
 * row-major upper-triangle order, of (d, a, r0) for Morse pairs or (A, B,
 * unused) for LJ pairs, then the bond row (K, r0, unused) when bonds are
 * set; [smin, smax] is the pair potentials' domain (the creators' min and
 * max).
 * Forward: nsteps steps from x0 (device, N*3 FP64, rounded to FP32) to xT
 * (device, N*3 FP64); ckpt (device, ceil(nsteps/every)*N*4 FP32) receives
 * the state every `every` steps. */
int mxg_adjoint_forward(mxg_adjoint *a, const double *theta, double smin, double smax, const double *x0, double *xT,
                        float *ckpt, int nsteps, int every, void *stream);

/* Backward: from dL/dxT (device, N*3) and the forward's checkpoints, the
 * segments are recomputed and swept in reverse; writes dL/dx0 (device,
 * N*3) and dL/dtheta (host, 3 per theta row), accumulated in FP64.
 * Synchronises the stream. */
int mxg_adjoint_backward(mxg_adjoint *a, const double *theta, double smin, double smax, const float *ckpt, int nsteps,
                         int every, const double *gxT, double *gx0, double *gtheta, void *stream);

/* Forces (device, N*3 FP32) at x (device, N*3 FP64) and the total Morse
 * energy (host, FP64 sum).  Synchronises the stream. */
int mxg_adjoint_forces(mxg_adjoint *a, const double *theta, double smin, double smax, const double *x, float *f,
                       double *epot, void *stream);

/* Soft pair-distance histogram per type pair (device, npairs*nbins FP64):
 * each pair with r < rmax (<= cutoff) adds the biweight (1 - t^2)^2,
 * |t| < 1, t = r/dr - (k + 1/2), to bins k; and its gradient, dL/dx
 * (device, N*3) from dL/dhist (device). */
int mxg_adjoint_rdf(mxg_adjoint *a, const double *x, int nbins, double rmax, double *hist, void *stream);
int mxg_adjoint_rdf_backward(mxg_adjoint *a, const double *x, int nbins, double rmax, const double *ghist,
                             double *gx, void *stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  mxg_adjoint_forces without an mxg_adjoint and without
 * allocating: the cell binning, sorted positions and energy partials live in
 * a file-scope device array of about 8.4 MB per GPU, held until the process
 * ends.  N <= 131072 particles and cells[0]*cells[1]*cells[2] <= 131072
 * (larger sizes return MXG_ERR_RANGE); cells of at least 3 per axis, no
 * narrower than the cutoff (else MXG_ERR_CUTOFF).  type (N int32, by id), x
 * (N*3 FP64), f (N*3 FP32) and epot (one FP64, the total Morse energy) are
 * device pointers; radius (ntypes) and theta (3 per type pair, as
 * mxg_adjoint_forward) are host arrays.  The work is queued on `stream`
 * (a hipStream_t, 0 = default) and nothing synchronises.  Calls on one GPU
 * share the scratch and leave the last call's contents in it: calls that
 * overlap on one GPU (on other streams or threads) race. */
int mxg_morse_forces_noalloc(int N, const double dim[3], const int cells[3], double cutoff, int ntypes,
                             const double *radius, const double *theta, double smin, double smax, const int *type,
                             const double *x, float *f, double *epot, void *stream);

/* Number of visible HIP devices (0 when none). */
int mxg_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* MXG_H_ */
