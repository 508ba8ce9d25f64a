// Test harness for the potential-table builder: compiles the builder's own
// source (mechanica_amd/csrc/mx_potential.cpp) into a small host library and
// exposes its internal stages (node exponent, edge placement, edge slopes,
// interval fits, the first-interval continuation) and a traced build through
// extern "C" functions, so the CPU tests can check each stage against an
// independent numpy restatement of the mathematics of potential_init
// (src/mdcore/src/MxPotential.cpp:1690-2340).  Test infrastructure only.
#include "mx_potential.cpp"

namespace {

// kind 0: Lennard-Jones 12-6 (A, B); 1: Morse (d, alpha); 2: harmonic
// (K, r0); 3: harmonic angle (K, theta0); 4: harmonic dihedral (K, n, delta).
mxg::PotentialFns fns_of(int kind, const double *par) {
  switch (kind) {
    case 0: return mxg::lj126_fns(par[0], par[1]);
    case 1: return mxg::morse_fns(par[0], par[1]);
    case 2: return mxg::harmonic_fns(par[0], par[1]);
    case 3: return mxg::harmonic_angle_fns(par[0], par[1]);
    default: return mxg::harmonic_dihedral_fns(par[0], (int)par[1], par[2]);
  }
}

}  // namespace

extern "C" {

// which 0: f, 1: f', 2: f^(6); returns -1 when the form has no such function.
int probe_eval(int kind, const double *par, int which, int n, const double *r, double *out) {
  const mxg::PotentialFns F = fns_of(kind, par);
  const mxg::Fn &g = which == 0 ? F.f : (which == 1 ? F.fp : F.f6p);
  if (!g) return -1;
  for (int i = 0; i < n; ++i) out[i] = g(r[i]);
  return 0;
}

double probe_node_exponent(int kind, const double *par, double a, double b) {
  return mxg::node_exponent(fns_of(kind, par).f6p, a, b);
}

int probe_place_edges(const float *al, float a, float b, int n, float *x) {
  return mxg::place_edges(al, a, b, n, x) ? 0 : -1;
}

void probe_edge_slopes(int kind, const double *par, double fpa, double fpb, int n, const float *x, double *slope,
                       int free_ends) {
  mxg::edge_slopes(fns_of(kind, par).f, fpa, fpb, n, x, slope, free_ends != 0);
}

// The n intervals' coefficients (8 floats each) and the builder's error.
void probe_fit(int kind, const double *par, const float *x, int n, float *out, float *err) {
  const mxg::PotentialFns F = fns_of(kind, par);
  mxg::fit_intervals(F.f, F.fp, x, n, out, err);
}

// c holds interval 0 (overwritten) and interval 1.
void probe_fix_first(float *c, float a) {
  std::vector<float> v(c, c + 2 * mxg::kPotChunk);
  mxg::fix_first_interval(v, a);
  for (int k = 0; k < 2 * mxg::kPotChunk; ++k) c[k] = v[k];
}

void probe_constants(double *out) {
  const double v[7] = {mxg::kWl0Wl1, mxg::kWl1Wl1, mxg::kWl0Wr1, mxg::kWl1Wr1,
                       mxg::kWr0Wr1, mxg::kWr1Wr1, mxg::kWl1Wr0};
  for (int k = 0; k < 7; ++k) out[k] = v[k];
}

int probe_sample_count() { return mxg::kPotN; }

// A traced build through the factory of the kind.  args: kind 0 (a, b, A, B),
// 1 (d, alpha, r0, min, max), 2 (a, b, K, r0), 3 (a, b, K, theta0),
// 4 (K, n, delta).  Returns the factory's code; on success *n, alpha4 and up
// to cap coefficients, the stretched domain ab, the exponent, the unscaled
// map al3 and up to maxtried (n, error) pairs of the search (*ntried).
int probe_build(int kind, const double *args, double tol, int *n, float *alpha4, float *coef, int cap, float *ab,
                double *exponent, float *al3, int *tried_n, float *tried_err, int maxtried, int *ntried) {
  mxg::Potential p;
  mxg::BuildTrace tr;
  int rc;
  switch (kind) {
    case 0: rc = mxg::make_lj126(p, args[0], args[1], args[2], args[3], tol, &tr); break;
    case 1: rc = mxg::make_morse(p, args[0], args[1], args[2], args[3], args[4], tol, &tr); break;
    case 2: rc = mxg::make_harmonic(p, args[0], args[1], args[2], args[3], tol, &tr); break;
    case 3: rc = mxg::make_harmonic_angle(p, args[0], args[1], args[2], args[3], tol, &tr); break;
    default: rc = mxg::make_harmonic_dihedral(p, args[0], (int)args[1], args[2], tol, &tr); break;
  }
  ab[0] = tr.a;
  ab[1] = tr.b;
  *exponent = tr.exponent;
  for (int k = 0; k < 3; ++k) al3[k] = tr.alpha[k];
  *ntried = (int)tr.tried_n.size();
  for (int k = 0; k < (int)tr.tried_n.size() && k < maxtried; ++k) {
    tried_n[k] = tr.tried_n[k];
    tried_err[k] = tr.tried_err[k];
  }
  if (rc < 0) return rc;
  *n = p.n;
  for (int k = 0; k < 4; ++k) alpha4[k] = p.alpha[k];
  for (int k = 0; k < (int)p.c.size() && k < cap; ++k) coef[k] = p.c[k];
  return rc;
}

}  // extern "C"
