// Tabulated pair potentials: piecewise quintic interpolants with a quadratic
// interval map, built on the host and consumed by the HIP pair kernels.
//
// The table layout and the construction procedure follow mdcore's
// MxPotential (src/mdcore/include/MxPotential.h:121-155,
// src/mdcore/src/MxPotential.cpp:1690-2340): every interval owns a chunk of
// 8 floats {centre, 1/half-width, c5, c4, c3, c2, c1, c0} (highest power
// first, evaluated by Horner on x = (r - centre) / half-width), and the
// interval of r is floor(alpha0 + r*(alpha1 + r*alpha2)).  Interval 0 is a
// quadratic continuation below the lower bound a.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace mxg {

constexpr int kPotDegree = 5;           // polynomial degree per interval
constexpr int kPotChunk = kPotDegree + 3;  // floats per interval
constexpr int kPotIvalsA = 1;           // first interval count tried
constexpr int kPotIvalsB = 10;          // second interval count tried
constexpr int kPotN = 100;              // Chebyshev sample count per interval
constexpr int kPotIvalsMax = 3048;      // hard cap on the interval count

// Potential flags: bit positions as MxPotential.h:55-103.
enum : uint32_t {
  kPotFlagLJ126 = 1u << 0,
  kPotFlagR2 = 1u << 4,
  kPotFlagR = 1u << 5,
  kPotFlagAngle = 1u << 6,
  kPotFlagHarmonic = 1u << 7,
  kPotFlagDihedral = 1u << 8,
  kPotFlagScaled = 1u << 11,
  kPotFlagShifted = 1u << 12,
  kPotFlagBound = 1u << 13,
};

enum : uint32_t { kPotKindTable = 0, kPotKindDPD = 1 };

// Error codes of the table builder (MxPotential.h:31-36).
enum : int {
  kPotErrOk = 0,
  kPotErrNull = -1,
  kPotErrMalloc = -2,
  kPotErrBounds = -3,
  kPotErrNyi = -4,
  kPotErrIvalsMax = -5,
};

struct Potential {
  uint32_t kind = kPotKindTable;
  uint32_t flags = 0;
  float alpha[4] = {0, 0, 0, 0};
  std::vector<float> c;  // (n + 1) * kPotChunk coefficients
  float r0_plusone = 0.0f;
  float a = 0.0f, b = 0.0f;
  float mu = 0.0f;
  int n = 0;
  // DPD parameters (kind == kPotKindDPD).
  float dpd_alpha = 0.0f, dpd_gamma = 0.0f, dpd_sigma = 0.0f;
  std::string name;
};

using Fn = std::function<double(double)>;

// The function a table interpolates, its first derivative (empty: the edge
// slopes are the free-end least-squares ones) and its sixth derivative
// (which places the interval edges).
struct PotentialFns {
  Fn f, fp, f6p;
};

// Closed forms of the reference's potentials (MxPotential.cpp:149-195
// Lennard-Jones 12-6, :305-346 harmonic, :3353-3394 Morse, :428-597 harmonic
// angle and dihedral in the cosine).
PotentialFns lj126_fns(double A, double B);
PotentialFns morse_fns(double d, double alpha);
PotentialFns harmonic_fns(double K, double r0);
PotentialFns harmonic_angle_fns(double K, double theta0);
PotentialFns harmonic_dihedral_fns(double K, int n, double delta);

// Intermediate results of build_table (potential_init's locals), recorded
// when a trace is passed: the stretched domain, the node exponent
// (potential_getalpha), the interval map before it is scaled by n, and the
// interval count and error of every candidate table in the order tried.
struct BuildTrace {
  float a = 0.0f, b = 0.0f;
  double exponent = 0.0;
  float alpha[3] = {0.0f, 0.0f, 0.0f};
  std::vector<int> tried_n;
  std::vector<float> tried_err;
};

// Builds the interpolation table of f on [a, b] to tolerance tol.
// Returns kPotErrOk or a negative kPotErr* code.
int build_table(Potential &p, const Fn &f, const Fn &fp, const Fn &f6p,
                float a, float b, float tol, BuildTrace *trace = nullptr);

// Factories (MxPotential.cpp:305-366, 932-991, 3353-3424 and
// DissapativeParticleDynamics.cpp:134-148).  They return a negative code on
// failure and leave p untouched.
int make_lj126(Potential &p, double a, double b, double A, double B,
               double tol, BuildTrace *trace = nullptr);
int make_morse(Potential &p, double d, double alpha, double r0, double min,
               double max, double tol, BuildTrace *trace = nullptr);
int make_harmonic(Potential &p, double a, double b, double K, double r0,
                  double tol, BuildTrace *trace = nullptr);
// potential_create_harmonic_angle (MxPotential.cpp:555-597): a table in
// cos(theta) on [cos(b), cos(a)] of K (theta - theta0)^2.
int make_harmonic_angle(Potential &p, double a, double b, double K,
                        double theta0, double tol, BuildTrace *trace = nullptr);
// potential_create_harmonic_dihedral (MxPotential.cpp:428-517): a table in
// cos(phi) of K (1 + cos(n phi - delta)).
int make_harmonic_dihedral(Potential &p, double K, int n, double delta,
                           double tol, BuildTrace *trace = nullptr);
int make_dpd(Potential &p, float alpha, float gamma, float sigma, float cutoff,
             bool shifted);

// Host evaluation with the kernels' arithmetic (potential_eval.hpp:67-161).
void eval_plain(const Potential &p, float r2, float *e, float *f);
// potential_eval_r (potential_eval.hpp:180-208): the table at r itself;
// f = dE/dr.
void eval_r(const Potential &p, float r, float *e, float *f);
bool eval_ex(const Potential &p, float ri, float rj, float r2, float *e,
             float *f);

}  // namespace mxg
