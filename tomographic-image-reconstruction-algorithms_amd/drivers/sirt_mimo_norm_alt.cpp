// sirt-mimo-norm-alt <fo> <G> <Np> <target eps> <iters> <a1> <b1> <n1> <a2> <b2> <n2>
// Weighted SIRT with the alternating relaxation of report eq. 5
// (sirt-mimo-norm-alt.cu:122-158, 472-477): alpha, beta arguments are x1000.
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "driver_common.h"

namespace {
struct Alt {
  double p[6];
};
double alt_scale(int k, const void *u) { return tomo_schedule(TOMO_SCHED_ALTERNATING, k, static_cast<const Alt *>(u)->p); }
}  // namespace

int main(int argc, char **argv) {
  if (argc != 12) {
    std::cerr << "Usage: " << argv[0]
              << " <fo_filename> <group size> <parallel groups> <target epsilon> <max iterations> <alpha1> <beta1> "
                 "<n1> <alpha2> <beta2> <n2>\n\n";
    return 1;
  }
  const std::string fo_name = argv[1];
  const std::string base = drv::base_name(fo_name);
  const int G = std::stoi(argv[2]);
  const int Np = std::stoi(argv[3]);
  drv::Loop L;
  L.rule = drv::Rule::Weighted;
  L.target_epsilon = std::stof(argv[4]);
  L.max_iterations = std::stoi(argv[5]);
  Alt alt;
  alt.p[0] = std::stof(argv[6]) * 1000.f;
  alt.p[1] = std::stof(argv[7]) * 1000.f;
  alt.p[2] = std::stof(argv[8]);
  alt.p[3] = std::stof(argv[9]) * 1000.f;
  alt.p[4] = std::stof(argv[10]) * 1000.f;
  alt.p[5] = std::stof(argv[11]);
  L.prev_epsilon_init = 100.f;
  L.convergence_file = "sirt_convergence.txt";
  L.time_file = "sirt_time.txt";
  L.one_based = true;
  std::cout << std::setprecision(9);
  std::cerr << std::setprecision(9);
  int ng = 0;
  tomo_ctx *ctx = drv::reference_context(G, Np, &ng);
  std::vector<float> fo;
  if (!drv::load_inputs(ctx, fo_name, G, fo)) {
    tomo_destroy(ctx);
    return 0;
  }
  drv::run_loop(ctx, L, alt_scale, &alt);
  drv::write_field(ctx, "sirt-norm-" + std::to_string(G) + "-" + base + ".txt", 192, 192);
  tomo_destroy(ctx);
  return 0;
}
