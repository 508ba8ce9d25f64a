// A C++ caller of the engine through include/mxg.hpp, written the way code
// against Mechanica's C++ API drives the step: universe_init from an
// MxUniverseConfig, a particle type, potential_create_LJ126 bound with
// MxUniverse_Bind, particles added by id (engine_addpart), then
// MxUniverse_Step.  Test infrastructure: the GPU test feeds it argon and
// checks every step against the oracle.
//
//   universe_argon IN OUT STEPS
//
// IN:  int32 N, float64 dim, int32 cells, then per particle float64 x[3],
//      float64 v[3], int32 id.
// OUT: per step, in load order: float32 xlocal[3N], int32 cell[N],
//      float32 v[3N], float32 f[3N]; then float64 observables[12] and the
//      uint64 engine time.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mxg.hpp"

using namespace mxg;

int main(int argc, char **argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s IN OUT STEPS\n", argv[0]);
    return 2;
  }
  FILE *in = std::fopen(argv[1], "rb");
  if (!in) return 3;
  int N = 0, cells = 0;
  double dim = 0;
  bool ok = std::fread(&N, sizeof(int), 1, in) == 1 && std::fread(&dim, sizeof(double), 1, in) == 1 &&
            std::fread(&cells, sizeof(int), 1, in) == 1;
  std::vector<double> x(3 * (size_t)N), v(3 * (size_t)N);
  std::vector<int> id(N);
  for (int k = 0; ok && k < N; ++k)
    ok = std::fread(&x[3 * k], sizeof(double), 3, in) == 3 && std::fread(&v[3 * k], sizeof(double), 3, in) == 3 &&
         std::fread(&id[k], sizeof(int), 1, in) == 1;
  std::fclose(in);
  if (!ok) return 4;
  const int steps = std::atoi(argv[3]);

  MxUniverseConfig conf;
  for (int k = 0; k < 3; ++k) {
    conf.dim[k] = dim;
    conf.spaceGridSize[k] = cells;
  }
  conf.cutoff = 1.0;
  conf.dt = 0.01;
  if (int rc = MxUniverse_Init(conf)) {
    std::fprintf(stderr, "init: %d %s\n", rc, Universe().error_message().c_str());
    return 5;
  }
  MxUniverse &u = Universe();
  const int argon = u.add_type(39.4, 0.1, PARTICLE_NEWTONIAN);
  MxPotentialPtr lj = potential_create_LJ126(0.275, 1.0, 9.5075e-06, 6.1545e-03, 1.0e-3);
  if (int rc = MxUniverse_Bind(lj, argon, argon)) {
    std::fprintf(stderr, "bind: %d\n", rc);
    return 6;
  }
  for (int k = 0; k < N; ++k) u.add_particle(argon, &x[3 * k], &v[3 * k], id[k]);

  FILE *out = std::fopen(argv[2], "wb");
  if (!out) return 7;
  std::vector<float> xl(3 * (size_t)N), vv(3 * (size_t)N), f(3 * (size_t)N);
  std::vector<int> cell(N);
  for (int s = 0; s < steps; ++s) {
    if (int rc = MxUniverse_Step(0.0, 0.0)) {
      std::fprintf(stderr, "step %d: %d %s\n", s, rc, u.error_message().c_str());
      return 8;
    }
    if (u.state(xl.data(), cell.data(), vv.data(), f.data())) return 9;
    std::fwrite(xl.data(), sizeof(float), xl.size(), out);
    std::fwrite(cell.data(), sizeof(int), cell.size(), out);
    std::fwrite(vv.data(), sizeof(float), vv.size(), out);
    std::fwrite(f.data(), sizeof(float), f.size(), out);
  }
  double obs[12];
  if (u.observables(obs)) return 10;
  const uint64_t t = u.time();
  std::fwrite(obs, sizeof(double), 12, out);
  std::fwrite(&t, sizeof(uint64_t), 1, out);
  std::fclose(out);
  std::printf("universe_argon: N=%d steps=%d time=%llu potential n=%d\n", N, steps, (unsigned long long)t, lj->n);
  return 0;
}
