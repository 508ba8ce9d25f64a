// C ABI over the engine and the potential builder (declarations in
// include/mxg.h).
#include <hip/hip_runtime_api.h>

#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mxg.h"
#include "mx_potential.h"
#include "mxg_adjoint.h"
#include "mxg_engine.h"

struct mxg_potential {
  mxg::Potential p;
};

struct mxg_engine {
  mxg::Engine e;
};

struct mxg_adjoint {
  mxg::Adjoint a;
};

namespace {
thread_local int g_pot_err = 0;

mxg_potential *finish(int rc, mxg_potential *h) {
  g_pot_err = rc;
  if (rc < 0) {
    delete h;
    return nullptr;
  }
  return h;
}
}  // namespace

extern "C" {

mxg_potential *mxg_potential_create_lj126(double min, double max, double A, double B, double tol) {
  auto *h = new (std::nothrow) mxg_potential();
  if (!h) return finish(mxg::kPotErrMalloc, nullptr);
  return finish(mxg::make_lj126(h->p, min, max, A, B, tol), h);
}

mxg_potential *mxg_potential_create_morse(double d, double a, double r0, double min, double max, double tol) {
  auto *h = new (std::nothrow) mxg_potential();
  if (!h) return finish(mxg::kPotErrMalloc, nullptr);
  return finish(mxg::make_morse(h->p, d, a, r0, min, max, tol), h);
}

mxg_potential *mxg_potential_create_harmonic(double min, double max, double K, double r0, double tol) {
  auto *h = new (std::nothrow) mxg_potential();
  if (!h) return finish(mxg::kPotErrMalloc, nullptr);
  return finish(mxg::make_harmonic(h->p, min, max, K, r0, tol), h);
}

mxg_potential *mxg_potential_create_harmonic_angle(double min, double max, double K, double theta0, double tol) {
  auto *h = new (std::nothrow) mxg_potential();
  if (!h) return finish(mxg::kPotErrMalloc, nullptr);
  return finish(mxg::make_harmonic_angle(h->p, min, max, K, theta0, tol), h);
}

mxg_potential *mxg_potential_create_harmonic_dihedral(double K, int n, double delta, double tol) {
  auto *h = new (std::nothrow) mxg_potential();
  if (!h) return finish(mxg::kPotErrMalloc, nullptr);
  return finish(mxg::make_harmonic_dihedral(h->p, K, n, delta, tol), h);
}

mxg_potential *mxg_potential_create_dpd(float alpha, float gamma, float sigma, float cutoff, int shifted) {
  auto *h = new (std::nothrow) mxg_potential();
  if (!h) return finish(mxg::kPotErrMalloc, nullptr);
  return finish(mxg::make_dpd(h->p, alpha, gamma, sigma, cutoff, shifted != 0), h);
}

void mxg_potential_free(mxg_potential *p) { delete p; }

int mxg_potential_error(void) { return g_pot_err; }

int mxg_potential_get_info(const mxg_potential *h, mxg_potential_info *out) {
  if (!h || !out) return MXG_ERR_NULL;
  const mxg::Potential &p = h->p;
  out->kind = p.kind;
  out->flags = p.flags;
  out->n = p.n;
  for (int k = 0; k < 4; ++k) out->alpha[k] = p.alpha[k];
  out->a = p.a;
  out->b = p.b;
  out->r0_plusone = p.r0_plusone;
  out->dpd_alpha = p.dpd_alpha;
  out->dpd_gamma = p.dpd_gamma;
  out->dpd_sigma = p.dpd_sigma;
  return MXG_ERR_OK;
}

int mxg_potential_coefficients(const mxg_potential *h, float *c, int capacity) {
  if (!h) return MXG_ERR_NULL;
  int n = (int)h->p.c.size();
  if (!c) return n;
  if (capacity < n) return MXG_ERR_RANGE;
  std::memcpy(c, h->p.c.data(), n * sizeof(float));
  return n;
}

int mxg_potential_eval(const mxg_potential *h, float r2, float *e, float *f) {
  if (!h || !e || !f) return MXG_ERR_NULL;
  if (h->p.kind != mxg::kPotKindTable || h->p.c.empty()) return MXG_ERR_POTENTIAL;
  mxg::eval_plain(h->p, r2, e, f);
  return MXG_ERR_OK;
}

int mxg_potential_eval_r(const mxg_potential *h, float r, float *e, float *f) {
  if (!h || !e || !f) return MXG_ERR_NULL;
  if (h->p.kind != mxg::kPotKindTable || h->p.c.empty()) return MXG_ERR_POTENTIAL;
  mxg::eval_r(h->p, r, e, f);
  return MXG_ERR_OK;
}

int mxg_potential_eval_ex(const mxg_potential *h, float ri, float rj, float r2, float *e, float *f) {
  if (!h || !e || !f) return MXG_ERR_NULL;
  if (h->p.kind != mxg::kPotKindTable || h->p.c.empty()) return MXG_ERR_POTENTIAL;
  return mxg::eval_ex(h->p, ri, rj, r2, e, f) ? 1 : 0;
}

int mxg_ensemble_init(mxg_engine **out, int nrep, const double *origins, const double *dims, const int *cells,
                      double cutoff, double dt) {
  if (!out || !origins || !dims || !cells || nrep < 1) return MXG_ERR_NULL;
  *out = nullptr;
  auto *h = new (std::nothrow) mxg_engine();
  if (!h) return MXG_ERR_MALLOC;
  std::vector<mxg::ReplicaSpec> specs(nrep);
  for (int r = 0; r < nrep; ++r)
    for (int k = 0; k < 3; ++k) {
      specs[r].origin[k] = origins[3 * r + k];
      specs[r].dim[k] = dims[3 * r + k];
      specs[r].cells[k] = cells[3 * r + k];
    }
  int rc = h->e.init(specs.data(), nrep, cutoff, dt);
  // The handle is returned even on failure so the caller can read the
  // message; it refuses every further call.
  *out = h;
  return rc;
}

int mxg_engine_init(mxg_engine **out, const double origin[3], const double dim[3], const int cells[3], double cutoff,
                    double dt) {
  return mxg_ensemble_init(out, 1, origin, dim, cells, cutoff, dt);
}

void mxg_engine_free(mxg_engine *e) { delete e; }

#define MXG_CHECK(e)                            \
  do {                                          \
    if (!(e)) return MXG_ERR_NULL;              \
    if ((e)->e.error()) return (e)->e.error();  \
  } while (0)

int mxg_engine_addtype(mxg_engine *e, double mass, float radius, int dynamics) {
  MXG_CHECK(e);
  return e->e.add_type(mass, radius, dynamics);
}

int mxg_engine_addpot(mxg_engine *e, const mxg_potential *p, int i, int j) {
  MXG_CHECK(e);
  if (!p) return MXG_ERR_NULL;
  return e->e.add_pot(p->p, i, j);
}

int mxg_engine_add_random_force(mxg_engine *e, int type, float mean, float std, float duration) {
  MXG_CHECK(e);
  return e->e.set_random_force(type, mean, std, duration);
}

int mxg_engine_add_force_berendsen(mxg_engine *e, int type, float tau) {
  MXG_CHECK(e);
  return e->e.set_single_body(type, 2, &tau);
}

int mxg_engine_add_force_constant(mxg_engine *e, int type, const float force[3]) {
  MXG_CHECK(e);
  if (!force) return MXG_ERR_NULL;
  return e->e.set_single_body(type, 3, force);
}

int mxg_engine_add_force_friction(mxg_engine *e, int type, float coef, float mean, float std, float duration) {
  MXG_CHECK(e);
  const float p[4] = {mean, std, duration, coef};
  return e->e.set_single_body(type, 4, p);
}

int mxg_engine_set_target_energy(mxg_engine *e, int type, double energy) {
  MXG_CHECK(e);
  return e->e.set_target_energy(type, energy);
}

int mxg_engine_set_seed(mxg_engine *e, uint64_t seed) {
  MXG_CHECK(e);
  return e->e.set_seed(seed);
}

int mxg_engine_set_replica_seed(mxg_engine *e, int replica, uint64_t seed) {
  MXG_CHECK(e);
  return e->e.set_replica_seed(replica, seed);
}

int mxg_engine_replica_seed(mxg_engine *e, int replica, uint64_t *seed) {
  MXG_CHECK(e);
  return e->e.replica_seed(replica, seed);
}

int mxg_engine_set_dpd_noise(mxg_engine *e, int mode) {
  MXG_CHECK(e);
  return e->e.set_dpd_counter_noise(mode);
}

int mxg_engine_set_max_dist_fraction(mxg_engine *e, float fraction) {
  MXG_CHECK(e);
  return e->e.set_max_dist_fraction(fraction);
}

int mxg_engine_set_brick(mxg_engine *e, int bx, int by, int bz, int cap, int pmax) {
  MXG_CHECK(e);
  return e->e.set_brick(bx, by, bz, cap, pmax);
}

int mxg_engine_get_brick(mxg_engine *e, int dims[3], int *cap, int *pmax) {
  if (!e || !dims || !cap || !pmax) return MXG_ERR_NULL;
  e->e.brick(dims, cap, pmax);
  return MXG_ERR_OK;
}

int mxg_engine_autotune(mxg_engine *e, int steps_per, int *taken) {
  MXG_CHECK(e);
  return e->e.autotune(steps_per, taken);
}

int mxg_engine_tune_log(mxg_engine *e, int max, int *layout, double *force_ms) {
  if (!e) return MXG_ERR_NULL;
  const auto &log = e->e.tune_log();
  const int n = (int)log.size();
  for (int k = 0; k < n && k < max; ++k) {
    if (layout) {
      layout[5 * k + 0] = log[k].bx;
      layout[5 * k + 1] = log[k].by;
      layout[5 * k + 2] = log[k].bz;
      layout[5 * k + 3] = log[k].cap;
      layout[5 * k + 4] = log[k].waves;
    }
    if (force_ms) force_ms[k] = log[k].force_ms;
  }
  return n;
}

int mxg_engine_set_waves(mxg_engine *e, int waves) {
  MXG_CHECK(e);
  return e->e.set_waves(waves);
}

int mxg_engine_get_waves(mxg_engine *e) {
  if (!e) return MXG_ERR_NULL;
  return e->e.waves();
}

int mxg_engine_set_integrator(mxg_engine *e, int kind) {
  MXG_CHECK(e);
  return e->e.set_integrator(kind);
}

int mxg_engine_set_graphs(mxg_engine *e, int on) {
  MXG_CHECK(e);
  return e->e.set_graphs(on);
}

int mxg_engine_set_virial(mxg_engine *e, int on) {
  MXG_CHECK(e);
  return e->e.set_virial(on);
}

int mxg_engine_load(mxg_engine *e, int N, const double *x, const double *v, const int *type, const int *pid,
                    const int *replica, const uint16_t *flags) {
  MXG_CHECK(e);
  return e->e.load(N, x, v, type, pid, replica, flags);
}

int mxg_engine_set_bonds(mxg_engine *e, int M, const int32_t *i, const int32_t *j, const mxg_potential *const *pots,
                         const double *dissociation_energy) {
  MXG_CHECK(e);
  if (M < 0) return MXG_ERR_RANGE;
  if (M > 0 && (!i || !j || !pots)) return MXG_ERR_NULL;
  std::vector<const mxg::Potential *> pp(M);
  for (int b = 0; b < M; ++b) pp[b] = pots[b] ? &pots[b]->p : nullptr;
  return e->e.set_bonds(M, i, j, pp.data(), dissociation_energy);
}

int mxg_engine_set_boundaries(mxg_engine *e, const int kind[6], const float *velocity, const float *restore,
                              const float *radius) {
  MXG_CHECK(e);
  if (!kind) return MXG_ERR_NULL;
  return e->e.set_boundaries(kind, velocity, restore, radius);
}

int mxg_engine_set_boundary_potential(mxg_engine *e, int side, int type, const mxg_potential *p) {
  MXG_CHECK(e);
  return e->e.set_boundary_potential(side, type, p ? &p->p : nullptr);
}

int mxg_engine_set_species(mxg_engine *e, int type, int constant_mask) {
  MXG_CHECK(e);
  return e->e.set_species(type, constant_mask);
}

int mxg_engine_set_species_values(mxg_engine *e, const float *values) {
  MXG_CHECK(e);
  if (!values) return MXG_ERR_NULL;
  return e->e.set_species_values(values);
}

int mxg_engine_get_species(mxg_engine *e, float *values, float *q) {
  if (!e) return MXG_ERR_NULL;
  return e->e.get_species(values, q);
}

int mxg_engine_add_flux(mxg_engine *e, int kind, int type_a, int type_b, int species_a, int species_b, float k,
                        float decay, float target) {
  MXG_CHECK(e);
  return e->e.add_flux(kind, type_a, type_b, species_a, species_b, k, decay, target);
}

int mxg_engine_set_angles(mxg_engine *e, int M, const int32_t *i, const int32_t *j, const int32_t *k,
                          const mxg_potential *const *pots) {
  MXG_CHECK(e);
  if (M > 0 && (!i || !j || !k || !pots)) return MXG_ERR_NULL;
  std::vector<const mxg::Potential *> pp(M > 0 ? M : 0);
  for (int a = 0; a < M; ++a) pp[a] = pots[a] ? &pots[a]->p : nullptr;
  return e->e.set_angles(M, i, j, k, pp.data());
}

int mxg_engine_set_dihedrals(mxg_engine *e, int M, const int32_t *i, const int32_t *j, const int32_t *k,
                             const int32_t *l, const mxg_potential *const *pots) {
  MXG_CHECK(e);
  if (M > 0 && (!i || !j || !k || !l || !pots)) return MXG_ERR_NULL;
  std::vector<const mxg::Potential *> pp(M > 0 ? M : 0);
  for (int a = 0; a < M; ++a) pp[a] = pots[a] ? &pots[a]->p : nullptr;
  return e->e.set_dihedrals(M, i, j, k, l, pp.data());
}

int mxg_engine_set_exclusions(mxg_engine *e, int M, const int32_t *i, const int32_t *j) {
  MXG_CHECK(e);
  if (M > 0 && (!i || !j)) return MXG_ERR_NULL;
  return e->e.set_exclusions(M, i, j);
}

int mxg_engine_term_energies(mxg_engine *e, int replica, double out[3]) {
  if (!e || !out) return MXG_ERR_NULL;
  return e->e.term_energies(replica, out);
}

int mxg_engine_step(mxg_engine *e, int nsteps) {
  MXG_CHECK(e);
  return e->e.step(nsteps);
}

int mxg_engine_sync(mxg_engine *e) {
  if (!e) return MXG_ERR_NULL;
  return e->e.sync();
}

int mxg_engine_unload(mxg_engine *e, double *x, double *v) {
  if (!e) return MXG_ERR_NULL;
  return e->e.get_global(x, v);
}

int mxg_engine_state(mxg_engine *e, float *xlocal, int *cell, float *v, float *f, float *pf) {
  if (!e) return MXG_ERR_NULL;
  return e->e.get_state(xlocal, cell, v, f, pf);
}

int mxg_engine_number_density(mxg_engine *e, float *nd) {
  if (!e || !nd) return MXG_ERR_NULL;
  return e->e.get_state(nullptr, nullptr, nullptr, nullptr, nullptr, nd);
}

int mxg_engine_neighbour_counts(mxg_engine *e, int *counts) {
  MXG_CHECK(e);
  if (!counts) return MXG_ERR_NULL;
  return e->e.neighbour_counts(counts);
}

int mxg_engine_observables(mxg_engine *e, int replica, double out[12]) {
  if (!e || !out) return MXG_ERR_NULL;
  return e->e.observables(replica, out);
}

int mxg_engine_bond_active(mxg_engine *e, uint8_t *active) {
  if (!e || !active) return MXG_ERR_NULL;
  return e->e.bond_active(active);
}

int mxg_engine_geometry(mxg_engine *e, int replica, int cdim[3], double h[3]) {
  if (!e || !cdim || !h) return MXG_ERR_NULL;
  return e->e.geometry(replica, cdim, h);
}

int mxg_engine_time(mxg_engine *e, uint64_t *time) {
  if (!e || !time) return MXG_ERR_NULL;
  *time = e->e.time();
  return MXG_ERR_OK;
}

int mxg_engine_error(const mxg_engine *e) { return e ? e->e.error() : MXG_ERR_NULL; }

const char *mxg_engine_error_message(const mxg_engine *e) { return e ? e->e.message().c_str() : "null engine"; }

int mxg_engine_set_timing(mxg_engine *e, int on) {
  if (!e) return MXG_ERR_NULL;
  return e->e.set_timing(on);
}

int mxg_engine_timing(mxg_engine *e, double *force_ms, double *step_ms) {
  if (!e) return MXG_ERR_NULL;
  if (force_ms) *force_ms = e->e.force_kernel_ms();
  if (step_ms) *step_ms = e->e.step_ms();
  return MXG_ERR_OK;
}

int mxg_adjoint_create(mxg_adjoint **out, int N, const double dim[3], const int cells[3], double cutoff, double dt,
                       float max_dist_fraction, int ntypes, const double *mass, const double *radius,
                       const int *type) {
  if (!out) return MXG_ERR_NULL;
  *out = nullptr;
  auto *h = new (std::nothrow) mxg_adjoint();
  if (!h) return MXG_ERR_MALLOC;
  int rc = h->a.create(N, dim, cells, cutoff, dt, max_dist_fraction, ntypes, mass, radius, type);
  *out = h;  // kept on failure so the message can be read; free it
  return rc;
}

void mxg_adjoint_free(mxg_adjoint *a) { delete a; }

const char *mxg_adjoint_error_message(const mxg_adjoint *a) { return a ? a->a.message() : "null adjoint"; }

int mxg_adjoint_forward(mxg_adjoint *a, const double *theta, double smin, double smax, const double *x0, double *xT,
                        float *ckpt, int nsteps, int every, void *stream) {
  if (!a) return MXG_ERR_NULL;
  return a->a.forward(theta, smin, smax, x0, xT, ckpt, nsteps, every, stream);
}

int mxg_adjoint_set_pair_kind(mxg_adjoint *a, int ti, int tj, int kind, double smin, double smax) {
  if (!a) return MXG_ERR_NULL;
  return a->a.set_pair_kind(ti, tj, kind, smin, smax);
}

int mxg_adjoint_set_bonds(mxg_adjoint *a, int M, const int32_t *i, const int32_t *j, double rmin, double rmax) {
  if (!a) return MXG_ERR_NULL;
  return a->a.set_bonds(M, i, j, rmin, rmax);
}

int mxg_adjoint_backward(mxg_adjoint *a, const double *theta, double smin, double smax, const float *ckpt, int nsteps,
                         int every, const double *gxT, double *gx0, double *gtheta, void *stream) {
  if (!a) return MXG_ERR_NULL;
  return a->a.backward(theta, smin, smax, ckpt, nsteps, every, gxT, gx0, gtheta, stream);
}

int mxg_adjoint_forces(mxg_adjoint *a, const double *theta, double smin, double smax, const double *x, float *f,
                       double *epot, void *stream) {
  if (!a) return MXG_ERR_NULL;
  return a->a.forces(theta, smin, smax, x, f, epot, stream);
}

int mxg_morse_forces_noalloc(int N, const double dim[3], const int cells[3], double cutoff, int ntypes,
                             const double *radius, const double *theta, double smin, double smax, const int *type,
                             const double *x, float *f, double *epot, void *stream) {
  return mxg::morse_forces_static(N, dim, cells, cutoff, ntypes, radius, theta, smin, smax, type, x, f, epot, stream);
}

int mxg_adjoint_rdf(mxg_adjoint *a, const double *x, int nbins, double rmax, double *hist, void *stream) {
  if (!a) return MXG_ERR_NULL;
  return a->a.rdf(x, nbins, rmax, hist, stream);
}

int mxg_adjoint_rdf_backward(mxg_adjoint *a, const double *x, int nbins, double rmax, const double *ghist,
                             double *gx, void *stream) {
  if (!a) return MXG_ERR_NULL;
  return a->a.rdf_backward(x, nbins, rmax, ghist, gx, stream);
}

int mxg_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

}  // extern "C"
