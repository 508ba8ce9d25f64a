// Host-side construction of the tabulated potentials.
//
// The procedure is mdcore's potential_init (src/mdcore/src/MxPotential.cpp:
// 1690-1932) with its helpers potential_getalpha (:2277-2340),
// potential_getcoeffs (:2139-2260) and potential_getfp_fixend (:2032-2112).
// Every intermediate keeps the precision the reference gives it (float nodes,
// float coefficients, double sums) so that the interval count n and the
// stored coefficients come out the same; this file must be compiled with
// -ffp-contract=off.
#include "mx_potential.h"

#include <cfloat>
#include <cmath>
#include <limits>
#include <vector>

namespace mxg {
namespace {

// Weighted inner products of the cubic Hermite cardinal functions that the
// slope solve uses (MxPotential.cpp:2040-2047).  They are constants of the
// reference algorithm and are kept identical so the slopes match.
constexpr double kWl0Wl1 = 0.1125317885884428;
constexpr double kWl1Wl1 = 0.03215579530433858;
constexpr double kWl0Wr1 = -0.04823369227661384;
constexpr double kWl1Wr1 = -0.02143719641629633;
constexpr double kWr0Wr1 = -0.1125317885884429;
constexpr double kWr1Wr1 = 0.03215579530433859;
constexpr double kWl1Wr0 = 0.04823369227661384;

// Monomial coefficients (ascending) of the derivative cardinal functions
// wl1(x) = (1-x)^2 (1+x) / 4 and wr1(x) = -(1+x)^2 (1-x) / 4.
constexpr double kCardL1[4] = {0.25, -0.25, -0.25, 0.25};
constexpr double kCardR1[4] = {-0.25, -0.25, 0.25, 0.25};

// Chebyshev-Lobatto abscissae and weights 1/sqrt(1-x^2) for the projections
// in the slope solve.
struct LobattoTable {
  double x[kPotN];
  double w[kPotN];
  LobattoTable() {
    for (int k = 1; k < kPotN - 1; ++k) {
      x[k] = std::cos(k * M_PI / (kPotN - 1));
      w[k] = 1.0 / std::sqrt(1.0 - x[k] * x[k]);
    }
    x[0] = 1.0;
    x[kPotN - 1] = -1.0;
    w[0] = 0.0;
    w[kPotN - 1] = 0.0;
  }
};

// cos(j k pi / N) for k < 7, rounded to float as the reference stores them.
struct CosTable {
  float v[7 * kPotN];
  CosTable() {
    for (int k = 0; k < 7; ++k)
      for (int j = 0; j < kPotN; ++j)
        v[k * kPotN + j] = (float)std::cos(j * k * M_PI / kPotN);
  }
};

const LobattoTable &lobatto() {
  static const LobattoTable t;
  return t;
}
const CosTable &costab() {
  static const CosTable t;
  return t;
}

// Golden-section search for the exponent of the node distribution that
// equalises the sixth-derivative error bound across [a, b].
double node_exponent(const Fn &f6p, double a, double b) {
  double t[kPotN], g[kPotN];
  double gmax = 0.0;
  const double golden = 2.0 / (1 + std::sqrt(5.0));
  for (int i = 0; i < kPotN; ++i) {
    t[i] = ((double)i + 1) / (kPotN + 1);
    g[i] = f6p(a + (b - a) * t[i]);
    gmax = std::fmax(gmax, std::fabs(g[i]));
  }
  if (gmax == 0.0) return 1.0;

  auto bound = [&](double e) {
    double worst = 0.0;
    for (int j = 0; j < kPotN; ++j) {
      double v = std::fabs(std::pow(e + 2 * (1 - e) * t[j], -6) * g[j]);
      if (v > worst) worst = v;
    }
    return worst;
  };

  double lo = 0.0, hi = 2.0;
  double m1 = hi - 2 * golden, m2 = lo + 2 * golden;
  double b1 = bound(m1), b2 = bound(m2);
  while (hi - lo > 1.0e-4) {
    if (b1 < b2) {
      hi = m2;
      m2 = m1;
      b2 = b1;
      m1 = hi - (hi - lo) * golden;
      b1 = bound(m1);
    } else {
      lo = m1;
      m1 = m2;
      b1 = b2;
      m2 = lo + (hi - lo) * golden;
      b2 = bound(m2);
    }
  }
  return (lo + hi) / 2;
}

// Places the n+1 interval edges so that the quadratic map sends edge i to i/n
// (Newton iterations in float, as the reference does).
bool place_edges(const float al[3], float a, float b, int n, float *x) {
  const float mtol = 10 * FLT_EPSILON;
  x[0] = a;
  x[n] = b;
  for (int i = 1; i < n; ++i) {
    x[i] = a + (b - a) * i / n;
    int iter = 0;
    for (;;) {
      double e = i - n * (al[0] + x[i] * (al[1] + x[i] * al[2]));
      x[i] += e / (n * (al[1] + 2 * x[i] * al[2]));
      if (std::fabs(e) < n * mtol) break;
      if (++iter > 10000) return false;
    }
  }
  return true;
}

// Slopes at the edges from a tridiagonal least-squares system with the end
// slopes fixed to fpa and fpb (potential_getfp_fixend), or left free when
// free_ends is set (potential_getfp, MxPotential.cpp:1949-2029: the first and
// last rows keep their one-sided least-squares terms).
void edge_slopes(const Fn &f, double fpa, double fpb, int n, const float *x,
                 double *slope, bool free_ends = false) {
  const LobattoTable &L = lobatto();
  std::vector<double> fx(n + 1), hx(n), sl(n), sr(n);
  std::vector<double> lower(n + 1), diag(n + 1), upper(n + 1), rhs(n + 1);

  for (int i = 0; i <= n; ++i) fx[i] = f(x[i]);
  for (int i = 0; i < n; ++i) hx[i] = x[i + 1] - x[i];

  for (int i = 0; i < n; ++i) {
    sl[i] = 0.0;
    sr[i] = 0.0;
    double mid = 0.5 * (x[i] + x[i + 1]);
    double half = 0.5 * (x[i + 1] - x[i]);
    for (int k = 1; k < kPotN - 1; ++k) {
      const double s = L.x[k];
      double fv = f(mid + half * s);
      sl[i] += L.w[k] * (fv * (kCardL1[0] + s * (kCardL1[1] + s * (kCardL1[2] + s * kCardL1[3]))));
      sr[i] += L.w[k] * (fv * (kCardR1[0] + s * (kCardR1[1] + s * (kCardR1[2] + s * kCardR1[3]))));
    }
    sl[i] /= kPotN - 2;
    sr[i] /= kPotN - 2;
  }

  if (free_ends) {
    diag[0] = kWl1Wl1 * hx[0];
    upper[0] = kWl1Wr1 * hx[0];
    rhs[0] = 2 * (sl[0] - fx[0] * kWl0Wl1 - fx[1] * kWl1Wr0);
  } else {
    diag[0] = 1.0;
    upper[0] = 0.0;
    rhs[0] = fpa;
  }
  for (int i = 1; i < n; ++i) {
    lower[i] = kWl1Wr1 * hx[i - 1];
    diag[i] = kWr1Wr1 * hx[i - 1] + kWl1Wl1 * hx[i];
    upper[i] = kWl1Wr1 * hx[i];
    rhs[i] = 2 * (sr[i - 1] - fx[i - 1] * kWl0Wr1 - fx[i] * kWr0Wr1) +
             2 * (sl[i] - fx[i] * kWl0Wl1 - fx[i + 1] * kWl1Wr0);
  }
  if (free_ends) {
    lower[n] = kWl1Wr1 * hx[n - 1];
    diag[n] = kWr1Wr1 * hx[n - 1];
    rhs[n] = 2 * (sr[n - 1] - fx[n - 1] * kWl0Wr1 - fx[n] * kWr0Wr1);
  } else {
    lower[n] = 0.0;
    diag[n] = 1.0;
    rhs[n] = fpb;
  }

  // Thomas algorithm.
  for (int i = 1; i <= n; ++i) {
    double m = lower[i] / diag[i - 1];
    diag[i] = diag[i] - m * upper[i - 1];
    rhs[i] = rhs[i] - m * rhs[i - 1];
  }
  slope[n] = rhs[n] / diag[n];
  for (int i = n - 1; i >= 0; --i)
    slope[i] = (rhs[i] - upper[i] * slope[i + 1]) / diag[i];
}

// Fits one quintic per interval (Hermite cubic plus two Chebyshev
// corrections) into out[i*8 .. i*8+7] and returns the worst locally relative
// error at the Chebyshev nodes in *err.
void fit_intervals(const Fn &f, const Fn &fp, const float *x, int n, float *out,
                   float *err) {
  const CosTable &C = costab();
  std::vector<double> fedge(n + 1), slope(n + 1);
  double fs[kPotN], phi[7], cee[6];

  for (int k = 0; k <= n; ++k) fedge[k] = f(x[k]);
  if (fp)
    edge_slopes(f, fp(x[0]), fp(x[n]), n, x, slope.data());
  else
    edge_slopes(f, 0.0, 0.0, n, x, slope.data(), true);

  *err = 0.0;
  for (int i = 0; i < n; ++i) {
    float *c = out + i * kPotChunk;
    double mid = (x[i] + x[i + 1]) / 2;
    double half = (x[i + 1] - x[i]) / 2;
    double fa = fedge[i], fb = fedge[i + 1];
    double dfa = slope[i] * half, dfb = slope[i + 1] * half;

    for (int k = 0; k < kPotN; ++k) fs[k] = f(mid + half * std::cos(k * M_PI / kPotN));
    for (int j = 0; j < 7; ++j) {
      phi[j] = (fa + (1 - 2 * (j % 2)) * fb) / 2;
      for (int k = 1; k < kPotN; ++k) phi[j] += fs[k] * C.v[j * kPotN + k];
      phi[j] *= 2.0 / kPotN;
    }

    // Hermite cubic in Chebyshev form.
    cee[0] = (4 * (fa + fb) + dfa - dfb) / 4;
    cee[1] = -(9 * (fa - fb) + dfa + dfb) / 16;
    cee[2] = (dfb - dfa) / 8;
    cee[3] = (fa - fb + dfa + dfb) / 16;
    cee[4] = 0.0;
    cee[5] = 0.0;

    // Degree-4 and degree-5 corrections that keep the edge values and slopes.
    double w = (6 * (cee[0] - phi[0]) - 4 * (cee[2] - phi[2]) - phi[4]) / (36 + 16 + 1);
    cee[0] += -6 * w;
    cee[2] += 4 * w;
    cee[4] = -w;
    w = (2 * (cee[1] - phi[1]) - 3 * (cee[3] - phi[3]) - phi[5]) / (4 + 9 + 1);
    cee[1] += -2 * w;
    cee[3] += 3 * w;
    cee[5] = -w;

    // Chebyshev to monomials on [-1, 1], highest power first.
    c[7] = cee[0] / 2 - cee[2] + cee[4];
    c[6] = cee[1] - 3 * cee[3] + 5 * cee[5];
    c[5] = 2 * cee[2] - 8 * cee[4];
    c[4] = 4 * cee[3] - 20 * cee[5];
    c[3] = 8 * cee[4];
    c[2] = 16 * cee[5];
    c[1] = 1.0 / half;
    c[0] = mid;

    double fmaxabs = 0.0, worst = 0.0;
    for (int k = 1; k < kPotN; ++k) {
      fmaxabs = std::fmax(std::fabs(fs[k]), fmaxabs);
      double s = C.v[kPotN + k];
      double e = std::fabs(fs[k] - c[7] - s * (c[6] + s * (c[5] + s * (c[4] + s * (c[3] + s * c[2])))));
      worst = std::fmax(e, worst);
    }
    worst /= std::fmax(fmaxabs, 1.0);
    *err = std::fmax(worst, (double)*err);
  }
}

// A candidate table with n intervals (chunk 0 reserved for the continuation).
struct Candidate {
  int n = 0;
  float err = 0.0f;
  std::vector<float> c;
};

bool make_candidate(Candidate &cand, const Fn &f, const Fn &fp,
                    const float al[3], float a, float b, int n) {
  std::vector<float> x(n + 1);
  if (!place_edges(al, a, b, n, x.data())) return false;
  cand.n = n;
  cand.c.assign((size_t)(n + 1) * kPotChunk, 0.0f);
  fit_intervals(f, fp, x.data(), n, cand.c.data() + kPotChunk, &cand.err);
  return true;
}

// Interval 0 becomes the quadratic Taylor expansion of interval 1 at its left
// edge, so radii below a extrapolate smoothly.
void fix_first_interval(std::vector<float> &c, float a) {
  c[0] = a;
  c[1] = 1.0 / a;
  double coeffs[kPotDegree], d[kPotDegree];
  for (int k = 0; k < kPotDegree; ++k) {
    coeffs[k] = c[2 * kPotChunk - 1 - k];
    d[k] = 0.0;
  }
  for (int i = 0; i < kPotDegree; ++i)
    for (int k = kPotDegree - 1; k >= i; --k) {
      d[i] = coeffs[k] + (-1.0) * d[i];
      coeffs[k] *= (k - i) * c[kPotChunk + 1] * a;
    }
  c[kPotChunk - 1] = d[0];
  c[kPotChunk - 2] = d[1];
  c[kPotChunk - 3] = 0.5 * d[2];
  for (int k = 3; k <= kPotDegree; ++k) c[kPotChunk - 1 - k] = 0.0f;
}

}  // namespace

int build_table(Potential &p, const Fn &f, const Fn &fp, const Fn &f6p,
                float a, float b, float tol, BuildTrace *trace) {
  if (!f || !f6p) return kPotErrNull;
  p.a = a;
  p.b = b;

  // Stretch the domain slightly for index rounding.
  b += std::fabs(b) * std::sqrt(FLT_EPSILON);
  a -= std::fabs(a) * std::sqrt(FLT_EPSILON);

  double ex = node_exponent(f6p, a, b);

  float al[4];
  double w = 1.0 / (a - b);
  w *= w;
  al[0] = a * a * w - ex * b * a * w;
  al[1] = -2 * a * w + ex * (a + b) * w;
  al[2] = w - ex * w;
  al[3] = 0.0f;
  w = 2 * FLT_EPSILON * (std::fabs(al[0]) + std::fabs(al[1]) + std::fabs(al[2]));
  al[0] -= w * a / (a - b);
  al[1] += w / (a - b);
  if (trace) {
    trace->a = a;
    trace->b = b;
    trace->exponent = ex;
    for (int k = 0; k < 3; ++k) trace->alpha[k] = al[k];
    trace->tried_n.clear();
    trace->tried_err.clear();
  }
  auto tried = [trace](const Candidate &c) {
    if (!trace) return;
    trace->tried_n.push_back(c.n);
    trace->tried_err.push_back(c.err);
  };

  Candidate lo, hi, mid;
  if (!make_candidate(lo, f, fp, al, a, b, kPotIvalsA)) return kPotErrBounds;
  tried(lo);

  if (lo.err < tol) {
    p.n = lo.n;
    p.c.swap(lo.c);
    p.alpha[0] = al[0] * p.n;
    p.alpha[1] = al[1] * p.n;
    p.alpha[2] = al[2] * p.n;
    p.alpha[3] = al[3];
    p.alpha[0] += 1;
    fix_first_interval(p.c, a);
    return kPotErrOk;
  }

  // Double until the tolerance is met, then bisect on the interval count.
  int r = kPotIvalsB;
  for (;;) {
    if (!make_candidate(hi, f, fp, al, a, b, r)) return kPotErrBounds;
    tried(hi);
    if (hi.err < tol) break;
    if (2 * r > kPotIvalsMax) return kPotErrIvalsMax;
    lo = std::move(hi);
    hi = Candidate();
    r *= 2;
  }
  int l = lo.n;
  while (r - l > 1) {
    int m = 0.5 * (r + l);
    if (!make_candidate(mid, f, fp, al, a, b, m)) return kPotErrBounds;
    tried(mid);
    if (mid.err > tol) {
      l = m;
      lo = std::move(mid);
    } else {
      r = m;
      hi = std::move(mid);
    }
    mid = Candidate();
  }

  p.n = r;
  p.c.swap(hi.c);
  p.alpha[0] = al[0] * p.n;
  p.alpha[1] = al[1] * p.n;
  p.alpha[2] = al[2] * p.n;
  p.alpha[3] = al[3];
  p.alpha[0] += 1.0;
  fix_first_interval(p.c, a);
  return kPotErrOk;
}

PotentialFns lj126_fns(double A, double B) {
  PotentialFns F;
  F.f = [A, B](double r) {
    double ir = 1.0 / r, ir2 = ir * ir, ir6 = ir2 * ir2 * ir2, ir12 = ir6 * ir6;
    return A * ir12 - B * ir6;
  };
  F.fp = [A, B](double r) {
    double ir = 1.0 / r, ir2 = ir * ir, ir4 = ir2 * ir2, ir12 = ir4 * ir4 * ir4;
    return 6.0 * ir * (-2.0 * A * ir12 + B * ir4 * ir2);
  };
  F.f6p = [A, B](double r) {
    double r2 = r * r, ir2 = 1.0 / r2, ir6 = ir2 * ir2 * ir2, ir12 = ir6 * ir6;
    return 10080.0 * ir12 * (884.0 * A * ir6 - 33.0 * B);
  };
  return F;
}

// d (exp(-2 alpha (r-1)) - 2 exp(-alpha (r-1))); powers of e as the
// reference evaluates them (std::pow(M_E, .)).
PotentialFns morse_fns(double d, double alpha) {
  PotentialFns F;
  F.f = [d, alpha](double r) {
    return d * (std::pow(M_E, -2 * alpha * (-1 + r)) - 2 / std::pow(M_E, alpha * (-1 + r)));
  };
  F.fp = [d, alpha](double r) {
    return d * ((-2 * alpha) / std::pow(M_E, 2 * alpha * (-1 + r)) +
                (2 * alpha) / std::pow(M_E, alpha * (-1 + r)));
  };
  F.f6p = [d, alpha](double r) {
    return d * ((64 * std::pow(alpha, 6)) / std::pow(M_E, 2 * alpha * (-1 + r)) -
                (2 * std::pow(alpha, 6)) / std::pow(M_E, alpha * (-1 + r)));
  };
  return F;
}

PotentialFns harmonic_fns(double K, double r0) {
  PotentialFns F;
  F.f = [K, r0](double r) { return K * (r - r0) * (r - r0); };
  F.fp = [K, r0](double r) { return 2.0 * K * (r - r0); };
  F.f6p = [](double) { return 0.0; };
  return F;
}

// K (acos(c) - theta0)^2 in c = cos(theta); no derivative is supplied, so
// the slopes are the free-end least-squares ones.
PotentialFns harmonic_angle_fns(double K, double theta0) {
  PotentialFns F;
  F.f = [K, theta0](double r) {
    r = std::fmin(1.0, std::fmax(-1.0, r));
    const double theta = std::acos(r);
    return K * (theta - theta0) * (theta - theta0);
  };
  F.f6p = [](double) { return 0.0; };
  return F;
}

// K (1 + cos(n acos(c) - delta)) in c = cos(phi).  With theta = acos(c),
// cos(n theta) = T_n(c) and sin(n theta) = U_{n-1}(c) sin(theta), the
// Chebyshev polynomials of the first and second kind.
PotentialFns harmonic_dihedral_fns(double K, int n, double delta) {
  PotentialFns F;
  F.f = [K, n, delta](double r) {
    // T_k and U_k by their common three-term recurrence.
    double t0 = 1.0, t1 = r, u0 = 1.0, u1 = 2 * r;
    for (int k = 2; k <= n; ++k) {
      const double t2 = 2 * r * t1 - t0, u2 = 2 * r * u1 - u0;
      t0 = t1;
      t1 = t2;
      u0 = u1;
      u1 = u2;
    }
    const double Tn = t1, Un1 = u0;  // T_n and U_{n-1}
    if (delta == 0.0) return K * (1.0 + Tn);
    if (delta == M_PI) return K * (1.0 - Tn);
    if (std::fabs(r) < 1.0) return K * (1.0 + Tn * std::cos(delta) + Un1 * std::sqrt(1.0 - r * r) * std::sin(delta));
    return K * (1.0 + Tn * std::cos(delta));
  };
  F.f6p = [](double) { return 0.0; };
  return F;
}

int make_lj126(Potential &p, double a, double b, double A, double B,
               double tol, BuildTrace *trace) {
  Potential q;
  q.flags = kPotFlagR2 | kPotFlagLJ126;
  q.name = "Lennard-Jones";
  const PotentialFns F = lj126_fns(A, B);
  int rc = build_table(q, F.f, F.fp, F.f6p, (float)a, (float)b, (float)tol, trace);
  if (rc < 0) return rc;
  p = std::move(q);
  return kPotErrOk;
}

int make_morse(Potential &p, double d, double alpha, double r0, double min,
               double max, double tol, BuildTrace *trace) {
  Potential q;
  q.flags = kPotFlagR2 | kPotFlagShifted;
  q.name = "Morse";
  q.r0_plusone = r0 + 1;
  const PotentialFns F = morse_fns(d, alpha);
  int rc = build_table(q, F.f, F.fp, F.f6p, (float)min, (float)max, (float)tol, trace);
  if (rc < 0) return rc;
  p = std::move(q);
  return kPotErrOk;
}

int make_harmonic(Potential &p, double a, double b, double K, double r0,
                  double tol, BuildTrace *trace) {
  Potential q;
  // The reference sets POTENTIAL_HARMONIC & POTENTIAL_R2, which is 0.
  q.flags = kPotFlagHarmonic & kPotFlagR2;
  q.name = "Harmonic";
  const PotentialFns F = harmonic_fns(K, r0);
  int rc = build_table(q, F.f, F.fp, F.f6p, (float)a, (float)b, (float)tol, trace);
  if (rc < 0) return rc;
  p = std::move(q);
  return kPotErrOk;
}

// potential_create_harmonic_angle (MxPotential.cpp:555-597): tabulated on
// [cos(b), cos(a)], the angles clamped to [0, pi].
int make_harmonic_angle(Potential &p, double a, double b, double K, double theta0, double tol,
                        BuildTrace *trace) {
  Potential q;
  q.flags = kPotFlagAngle | kPotFlagHarmonic;
  q.name = "Harmonic Angle";
  if (a < 0.0) a = 0.0;
  if (b > M_PI) b = M_PI;
  const double left = std::cos(b), right = std::cos(a);
  const PotentialFns F = harmonic_angle_fns(K, theta0);
  int rc = build_table(q, F.f, F.fp, F.f6p, (float)left, (float)right, (float)tol, trace);
  if (rc < 0) return rc;
  p = std::move(q);
  return kPotErrOk;
}

// potential_create_harmonic_dihedral (MxPotential.cpp:428-517): the domain
// is [-1, 1], pulled in by a relative sqrt(eps) when delta is not a multiple
// of pi (the sin(theta) factor has unbounded slope at the ends).
int make_harmonic_dihedral(Potential &p, double K, int n, double delta, double tol, BuildTrace *trace) {
  if (n < 1) return kPotErrBounds;
  Potential q;
  q.flags = kPotFlagR | kPotFlagHarmonic | kPotFlagDihedral;
  q.name = "Harmonic Dihedral";
  double a = -1.0, b = 1.0;
  if (std::fmod(delta, M_PI) != 0) {
    const double e = std::sqrt((double)std::numeric_limits<float>::epsilon());
    a = -1.0 / (1.0 + e);
    b = 1.0 / (1.0 + e);
  }
  const PotentialFns F = harmonic_dihedral_fns(K, n, delta);
  int rc = build_table(q, F.f, F.fp, F.f6p, (float)a, (float)b, (float)tol, trace);
  if (rc < 0) return rc;
  p = std::move(q);
  return kPotErrOk;
}

int make_dpd(Potential &p, float alpha, float gamma, float sigma, float cutoff,
             bool shifted) {
  Potential q;
  q.kind = kPotKindDPD;
  q.dpd_alpha = alpha;
  q.dpd_gamma = gamma;
  q.dpd_sigma = sigma;
  q.a = std::sqrt(std::numeric_limits<float>::epsilon());
  q.b = cutoff;
  q.name = "Dissapative Particle Dynamics";
  if (shifted) q.flags |= kPotFlagShifted;
  p = std::move(q);
  return kPotErrOk;
}

void eval_plain(const Potential &p, float r2, float *e, float *f) {
  float r = std::sqrt(r2);
  int ind = (int)std::fmax(0.0f, p.alpha[0] + r * (p.alpha[1] + r * p.alpha[2]));
  const float *c = &p.c[(size_t)ind * kPotChunk];
  float x = (r - c[0]) * c[1];
  float ee = c[2] * x + c[3];
  float eff = c[2];
  for (int k = 4; k < kPotChunk; ++k) {
    eff = eff * x + ee;
    ee = ee * x + c[k];
  }
  *e = ee;
  *f = eff * c[1] / r;
}

void eval_r(const Potential &p, float r, float *e, float *f) {
  int ind = (int)std::fmax(0.0f, p.alpha[0] + r * (p.alpha[1] + r * p.alpha[2]));
  ind = std::min(std::max(ind, 0), p.n);
  const float *c = &p.c[(size_t)ind * kPotChunk];
  float x = (r - c[0]) * c[1];
  float ee = c[2] * x + c[3];
  float eff = c[2];
  for (int k = 4; k < kPotChunk; ++k) {
    eff = eff * x + ee;
    ee = ee * x + c[k];
  }
  *e = ee;
  *f = eff * c[1];
}

bool eval_ex(const Potential &p, float ri, float rj, float r2, float *e,
             float *f) {
  const float eps = std::numeric_limits<float>::epsilon();
  float r = std::sqrt(r2);
  float ro = r < eps ? eps : r;
  r = r < p.a ? p.a : r;
  if (p.flags & kPotFlagScaled)
    r = r / (ri + rj);
  else if (p.flags & kPotFlagShifted)
    r = r - (ri + rj) + p.r0_plusone;
  unsigned ind = std::max(0.0f, p.alpha[0] + r * (p.alpha[1] + r * p.alpha[2]));
  if (r > p.b || ind > (unsigned)p.n) return false;
  const float *c = &p.c[(size_t)ind * kPotChunk];
  float x = (r - c[0]) * c[1];
  float ee = c[2] * x + c[3];
  float eff = c[2];
  for (int k = 4; k < kPotChunk; ++k) {
    eff = eff * x + ee;
    ee = ee * x + c[k];
  }
  *e = ee;
  *f = eff * c[1] / ro;
  return true;
}

}  // namespace mxg
