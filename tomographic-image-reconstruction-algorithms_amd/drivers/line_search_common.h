// Shared main() of the ground-truth line-search study drivers
// (sirt-test-weight.cu, sirt-norm-test-weight.cu): every iteration takes the
// relaxation from the grid 0, 100, .., 1e6 that minimises the relative error
// against the truth, as the producer of the data_dynamic logs did.
#pragma once

#include <chrono>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "driver_common.h"

namespace drv {

inline int line_search_main(int argc, char **argv, bool weighted, const std::string &prefix) {
  if (argc != 6) {
    std::cerr << "Usage: " << argv[0] << " <fo_filename> <group size> <parallel groups> <target epsilon> "
              << "<max iterations>\n\n";
    return 1;
  }
  const std::string fo_name = argv[1];
  const std::string base = base_name(fo_name);
  const int G = std::stoi(argv[2]);
  const int Np = std::stoi(argv[3]);
  const float target = std::stof(argv[4]);
  const int max_iterations = std::stoi(argv[5]);
  std::cout << std::setprecision(9);
  std::cerr << std::setprecision(9);
  int ng = 0;
  tomo_ctx *ctx = reference_context(G, Np, &ng);
  std::vector<float> fo;
  if (!load_inputs(ctx, fo_name, G, fo)) {
    tomo_destroy(ctx);
    return 0;
  }
  std::ofstream conv("sirt_convergence.txt"), tfile("sirt_time.txt");
  const int w_iter = 6, w_eps = 12, w_diff = 15;
  std::cout << std::setw(w_iter) << "iter" << " " << std::setw(w_eps) << "epsilon" << " " << std::setw(w_diff)
            << "difference" << " \n"
            << std::string(w_iter, '-') << " " << std::string(w_eps, '-') << " " << std::string(w_diff, '-')
            << " \n";
  float prev = 100.f, curr = 0.f, file_eps = std::numeric_limits<float>::infinity();
  bool reached = false, diverged = false;
  int iter = 0;
  const auto t0 = std::chrono::steady_clock::now();
  while (!reached && !diverged && iter < max_iterations) {
    ++iter;
    check(tomo_gradient(ctx, 0, ng, nullptr), ctx, "tomo_gradient");
    double scale = 0.0, eps_pred = 0.0;
    check(tomo_line_search(ctx, weighted ? 1 : 0, 0.0, 100.0, 10001, &scale, &eps_pred), ctx, "tomo_line_search");
    if (scale == 0.0) break;
    if (weighted) check(tomo_update_weighted(ctx, (float)scale, nullptr), ctx, "tomo_update_weighted");
    else check(tomo_update_uniform(ctx, (float)scale, TOMO_UPDATE_MEAN), ctx, "tomo_update_uniform");
    double eps = 0.0;
    check(tomo_error(ctx, &eps), ctx, "tomo_error");
    curr = (float)eps;
    const float t = std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
    if (std::fabs(file_eps - curr) > 0.2f) {
      conv << curr << " ";
      tfile << t << " ";
      file_eps = curr;
    }
    std::cout << std::setw(w_iter) << iter << " " << std::setw(w_eps) << curr << " " << std::setw(w_diff)
              << prev - curr << " " << (float)scale << " \n";
    reached = curr <= target;
    diverged = curr > prev || std::isnan(curr);
    prev = curr;
  }
  if (reached)
    std::cerr << "reached target epsilon: " << target << ", at iter: " << iter << ", epsilon: " << curr << "\n\n";
  else if (diverged)
    std::cerr << "diverged at iter: " << iter << ", epsilon: " << curr << "\n\n";
  else
    std::cerr << "max iterations reached: " << iter << ", epsilon: " << curr << "\n\n";
  std::cout << std::endl;
  std::cerr << "time (s): " << std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count() << "\n";
  write_field(ctx, prefix + std::to_string(G) + "-" + base + ".txt", 192, 192);
  tomo_destroy(ctx);
  return 0;
}

}  // namespace drv
