// Shared main() of the file-driven SIRT / ART drivers (sirt-mimo-io.cu,
// sirt-mimo-norm.cu, art-mimo-io.cu, sirt-mimo-norm-mem.cu,
// sirt-mimo-norm-alt.cu): same arguments, file names and stop rules.
#pragma once

#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "driver_common.h"

namespace drv {

struct PowerSchedule {
  double p[4];  // omega, alpha, beta, gamma
};

inline double power_scale(int k, const void *u) {
  return tomo_schedule(TOMO_SCHED_POWER, k, static_cast<const PowerSchedule *>(u)->p);
}

// argv layout: <fo> <G> [<Np>] <target> <iters> <omega> <alpha> <beta> <gamma>
inline int sirt_io_main(int argc, char **argv, Rule rule, bool with_np, const std::string &prefix,
                        const std::string &log_prefix) {
  const int want = with_np ? 10 : 9;
  if (argc != want) {
    std::cerr << "Usage: " << argv[0] << " <fo_filename> <sensor group size>"
              << (with_np ? " <parallel groups>" : "")
              << " <target epsilon> <max iterations> <omega> <alpha> <beta> <gamma>\n\n";
    return 1;
  }
  int a = 1;
  const std::string fo_name = argv[a++];
  const std::string base = base_name(fo_name);
  const int G = std::stoi(argv[a++]);
  const int Np = with_np ? std::stoi(argv[a++]) : 0;
  Loop L;
  L.rule = rule;
  L.target_epsilon = std::stof(argv[a++]);
  L.max_iterations = std::stoi(argv[a++]);
  PowerSchedule ps;
  for (int i = 0; i < 4; ++i) ps.p[i] = std::stof(argv[a++]);
  L.prev_epsilon_init = 100.f;
  L.convergence_file = log_prefix + "_convergence.txt";
  L.time_file = log_prefix + "_time.txt";
  L.one_based = with_np;
  std::cout << std::setprecision(9);
  std::cerr << std::setprecision(9);
  int ng = 0;
  tomo_ctx *ctx = reference_context(G, Np, &ng);
  std::vector<float> fo;
  if (!load_inputs(ctx, fo_name, G, fo)) {
    tomo_destroy(ctx);
    return 0;
  }
  std::cerr << "group size:     " << G << "\n"
            << "target epsilon: " << L.target_epsilon << "\n"
            << "omega:          " << ps.p[0] << "\n"
            << "alpha:          " << ps.p[1] << "\n"
            << "beta:           " << ps.p[2] << "\n"
            << "gamma:          " << ps.p[3] << "\n\n";
  run_loop(ctx, L, power_scale, &ps);
  const std::string f_name = prefix + std::to_string(G) + "-" + base + ".txt";
  write_field(ctx, f_name, 192, 192);
  tomo_destroy(ctx);
  return 0;
}

}  // namespace drv
