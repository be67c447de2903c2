// Shared main() of the intermediate-field averaging and group-serial SIRT
// variants (sirt-mimo-avg.cu, sirt-mimo-avg-weight.cu, sirt-mimo-weight.cu,
// sirt-iter.cu, sirt-iter-norm.cu).  Each is a file-driven loop like
// sirt-mimo-io whose update is one of the engine's:
//  - averaging the per-group intermediate fields f + s mask df_g over the
//    groups equals f + s mask mean_g(df_g), the uniform update in mean mode;
//  - their distance-weighted sum is tomo_update_avg_weight;
//  - group-serial gradients (one group at a time) are passes of one group;
//  - the schedules are omega (alpha + beta / (k+1)^gamma) for fixed gamma.
// The FP32 atomic sums of the reference become FP64 sums in a fixed order.
#pragma once

#include <iomanip>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "driver_common.h"

namespace drv {

struct Variant {
  const char *usage;          // argument list after <fo_filename> <sensor group size>
  Rule rule;                  // Uniform (mean over groups), Weighted or AvgWeight
  int group_serial;           // 1: passes of one group
  float prev_epsilon_init;    // initial "previous epsilon" of the stop rule
  bool table;                 // iter/epsilon/difference/scale table, else "Iter: / epsilon ="
  const char *convergence_file, *time_file, *field_prefix;
  int nparams;                // numbers after <max iterations>
  // schedule parameters {omega, alpha, beta, gamma} from those numbers
  void (*schedule)(const float *params, double out[4]);
};

inline double variant_scale(int k, const void *u) {
  return tomo_schedule(TOMO_SCHED_POWER, k, static_cast<const double *>(u));
}

inline int variant_main(int argc, char **argv, const Variant &V) {
  if (argc != 5 + V.nparams) {
    std::cerr << "Usage: " << argv[0] << " <fo_filename> <sensor group size> <target epsilon> <max iterations>"
              << V.usage << "\n\n";
    return 1;
  }
  const std::string fo_name = argv[1];
  const std::string base = base_name(fo_name);
  const int G = std::stoi(argv[2]);
  Loop L;
  L.rule = V.rule;
  L.mode = TOMO_UPDATE_MEAN;
  L.target_epsilon = std::stof(argv[3]);
  L.max_iterations = std::stoi(argv[4]);
  float params[4] = {0.f, 0.f, 0.f, 0.f};
  for (int i = 0; i < V.nparams; ++i) params[i] = std::stof(argv[5 + i]);
  double sched[4];
  V.schedule(params, sched);
  L.prev_epsilon_init = V.prev_epsilon_init;
  L.convergence_file = V.convergence_file;
  L.time_file = V.time_file;
  L.table = V.table;
  std::cout << std::setprecision(9);
  std::cerr << std::setprecision(9);
  int ng = 0;
  tomo_ctx *ctx = reference_context(G, V.group_serial ? 1 : 0, &ng);
  std::vector<float> fo;
  if (!load_inputs(ctx, fo_name, G, fo)) {
    tomo_destroy(ctx);
    return 0;
  }
  std::cerr << "group size:     " << G << "\n"
            << "target epsilon: " << L.target_epsilon << "\n"
            << "omega:          " << sched[0] << "\n"
            << "alpha:          " << sched[1] << "\n"
            << "beta:           " << sched[2] << "\n\n";
  run_loop(ctx, L, variant_scale, sched);
  write_field(ctx, std::string(V.field_prefix) + std::to_string(G) + "-" + base + ".txt", 192, 192);
  tomo_destroy(ctx);
  return 0;
}

}  // namespace drv
