// Differentiable overdamped Morse system: the forward trajectory of
// engine_step's overdamped path with the Morse pair force evaluated
// analytically, and its discrete adjoint.
#pragma once

#include <string>
#include <vector>

namespace mxg {

constexpr int kAdjMaxTypes = 4;
constexpr int kAdjMaxPairs = kAdjMaxTypes * (kAdjMaxTypes + 1) / 2;
constexpr int kAdjMaxBins = 256;
// Gradient slots: (d, a, r0) or (A, B, -) per type pair, then the bond
// potential's (K, r0, -).
constexpr int kAdjSlots = 3 * kAdjMaxPairs + 3;
// Limits of morse_forces_static's file-scope scratch (about 8.4 MB per GPU).
constexpr int kStaticMaxParticles = 131072;
constexpr int kStaticMaxCells = 131072;

// Adjoint::forces without an Adjoint and without allocating: the binning,
// sorted positions and energy partials live in a file-scope device array;
// x, type (by id), f and epot (one double) are device pointers.  Returns 0,
// -6 beyond the limits above, -27 for cells narrower than the cutoff, -21
// on a device error.  Calls on one GPU share the scratch: calls that overlap
// on one GPU race.
int morse_forces_static(int N, const double dim[3], const int cells[3], double cutoff, int ntypes,
                        const double *radius, const double *theta, double smin, double smax, const int *type,
                        const double *x, float *f, double *epot, void *stream);

struct AdjointDev;

class Adjoint {
 public:
  ~Adjoint();
  int create(int N, const double dim[3], const int cells[3], double cutoff, double dt, float max_dist_fraction,
             int ntypes, const double *mass, const double *radius, const int *type);
  // Potential of a type pair: 0 Morse (theta row d, a, r0), 1 LJ 12-6
  // (theta row A, B, unused), on its own domain [smin, smax] when
  // smax > smin (else the domain the forward and backward calls pass).
  int set_pair_kind(int ti, int tj, int kind, double smin, double smax);
  // Harmonic bonds between ids i[b], j[b] (replacing any earlier set), r
  // clamped to [rmin, rmax]; theta then carries one more row (K, r0, -).
  int set_bonds(int M, const int *i, const int *j, double rmin, double rmax);
  // theta: one row per type pair (ti <= tj, row-major upper triangle) and,
  // with bonds, the bond row; [smin, smax] is the pair potentials' domain.
  int forward(const double *theta, double smin, double smax, const double *x0, double *xT, float *ckpt, int nsteps,
              int every, void *stream);
  int backward(const double *theta, double smin, double smax, const float *ckpt, int nsteps, int every,
               const double *gxT, double *gx0, double *gtheta, void *stream);
  int forces(const double *theta, double smin, double smax, const double *x, float *f, double *epot, void *stream);
  int rdf(const double *x, int nbins, double rmax, double *hist, void *stream);
  int rdf_backward(const double *x, int nbins, double rmax, const double *ghist, double *gx, void *stream);
  const char *message() const { return msg_.c_str(); }
  int npairs() const { return ntypes_ * (ntypes_ + 1) / 2; }
  // theta rows: the type pairs, then the bond row when bonds are set.
  int nrows() const { return npairs() + (M_ > 0 ? 1 : 0); }

 private:
  int fail(int code, const std::string &m) {
    msg_ = m;
    return code;
  }
  int device(int st, const char *what);
  int ensure_workspace(int every);
  int set_args(const double *theta, double smin, double smax);
  void bin(const void *x, int slot, void *stream);

  AdjointDev *d_ = nullptr;
  int N_ = 0, ntypes_ = 0, ncells_ = 0, ws_cap_ = 0, nblocks_ = 0, nlanes_ = 0, M_ = 0;
  int kind_[kAdjMaxPairs] = {};
  bool dom_[kAdjMaxPairs] = {};
  double pmin_[kAdjMaxPairs] = {}, pmax_[kAdjMaxPairs] = {};
  double bmin_ = 0, bmax_ = 0;
  int cdim_[3] = {0, 0, 0};
  double dim_[3] = {0, 0, 0}, cutoff_ = 0, dt_ = 0;
  float maxfrac_ = 0.05f;
  std::vector<double> mass_, radius_;
  std::string msg_;
};

}  // namespace mxg
