// driver_common.cpp -- see driver_common.h.
#include "driver_common.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>

namespace drv {

std::string base_name(const std::string &fo) {
  if (std::count(fo.begin(), fo.end(), '.') != 1) {
    std::cerr << "Error: '" << fo << "' should have only one period.\n"
              << "       It should be in the current directory and have only one filetype extension.\n\n";
    std::exit(1);
  }
  return fo.substr(0, fo.find('.'));
}

void check(int status, tomo_ctx *ctx, const char *what) {
  if (status >= 0) return;
  std::cerr << "Error: " << what << ": " << tomo_strerror(status);
  const char *m = tomo_last_error(ctx);
  if (m && *m) std::cerr << " (" << m << ")";
  std::cerr << "\n";
  std::exit(1);
}

tomo_ctx *reference_context(int G, int P, int *ng) {
  tomo_config cfg;
  tomo_config_reference(&cfg);
  cfg.shots_per_pass = P;
  tomo_ctx *ctx = nullptr;
  check(tomo_create(&cfg, -1, &ctx), nullptr, "tomo_create");
  std::vector<int> ii(640), jj(640);
  check(tomo_transducers_reference(ii.data(), jj.data(), 640), ctx, "transducers");
  check(tomo_set_transducers(ctx, ii.data(), jj.data(), 640, 1), ctx, "tomo_set_transducers");
  const int n = tomo_set_groups(ctx, G);
  check(n, ctx, "tomo_set_groups");
  *ng = n;
  return ctx;
}

bool load_inputs(tomo_ctx *ctx, const std::string &fo_filename, int G, std::vector<float> &fo) {
  const std::string base = base_name(fo_filename);
  tomo_config cfg;
  tomo_config_reference(&cfg);
  const int nx = cfg.nx, ny = cfg.ny, nt = cfg.nt, ng = tomo_num_shots(ctx);
  if (!std::ifstream(fo_filename)) {
    std::cerr << "Error: '" << fo_filename << "' file not found in current directory.\n\n";
    return false;
  }
  const char *sides[4] = {"bottom", "right", "top", "left"};
  std::vector<std::vector<float>> s(4, std::vector<float>((size_t)ng * nt * nx));
  for (int q = 0; q < 4; ++q) {
    const std::string name = base + "-data-" + sides[q] + "-" + std::to_string(G) + ".txt";
    if (!std::ifstream(name)) {
      std::cerr << "Error: '" << name << "' file not found in current directory.\n\n";
      return false;
    }
  }
  fo.assign((size_t)nx * ny, 0.f);
  check(tomo_read_text(fo_filename.c_str(), fo.data(), nx, ny, 1, 1), ctx, "read fo");
  for (int q = 0; q < 4; ++q) {
    const std::string name = base + "-data-" + sides[q] + "-" + std::to_string(G) + ".txt";
    check(tomo_read_text(name.c_str(), s[q].data(), nx, nt, ng, 1), ctx, "read data");
  }
  std::vector<float> traces((size_t)ng * nt * tomo_num_receivers(ctx));
  check(tomo_sides_to_traces(ctx, s[0].data(), s[1].data(), s[2].data(), s[3].data(), traces.data()), ctx,
        "sides_to_traces");
  check(tomo_set_measured(ctx, traces.data()), ctx, "tomo_set_measured");
  check(tomo_set_truth(ctx, fo.data()), ctx, "tomo_set_truth");
  return true;
}

int run_loop(tomo_ctx *ctx, const Loop &L, double (*scale_fn)(int, const void *), const void *user) {
  std::ofstream conv(L.convergence_file), tfile(L.time_file);
  const int w_iter = 6, w_eps = 12, w_diff = 15;
  if (L.table)
    std::cout << std::setw(w_iter) << "iter" << " " << std::setw(w_eps) << "epsilon" << " " << std::setw(w_diff)
              << "difference" << " \n"
              << std::string(w_iter, '-') << " " << std::string(w_eps, '-') << " " << std::string(w_diff, '-')
              << " \n";
  float prev = L.prev_epsilon_init;
  float file_eps = std::numeric_limits<float>::infinity();
  float curr = -std::numeric_limits<float>::infinity();
  const auto t0 = std::chrono::steady_clock::now();
  const int maxit = L.max_iterations < 0 ? std::numeric_limits<int>::max() : L.max_iterations;
  int iter = 0;
  for (; iter < maxit; ++iter) {
    const float scale = (float)scale_fn(iter, user);
    double misfit = 0, eps = 0;
    int rc;
    const bool art = L.rule == Rule::Art || (L.rule == Rule::ArtThenWeighted && iter < L.art_iterations);
    if (art) {
      rc = tomo_art_iteration(ctx, scale, &misfit, &eps);
    } else if (L.rule == Rule::AvgWeight) {
      rc = tomo_gradient(ctx, 0, tomo_num_shots(ctx), &misfit);
      if (!rc) rc = tomo_update_avg_weight(ctx, scale, nullptr);
      if (!rc) rc = tomo_error(ctx, &eps);
    } else {
      rc = tomo_sirt_iteration(ctx, scale, L.rule != Rule::Uniform, L.mode, &misfit, &eps);
    }
    check(rc, ctx, "iteration");
    curr = (float)eps;
    const float t = std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
    if (std::fabs(file_eps - curr) > 0.2f) {
      conv << curr << " ";
      tfile << t << " ";
      file_eps = curr;
    }
    const int shown = L.one_based ? iter + 1 : iter;
    if (L.table)
      std::cout << std::setw(w_iter) << shown << " " << std::setw(w_eps) << curr << " " << std::setw(w_diff)
                << prev - curr << " " << scale << " \n";
    else
      std::cout << "\nIter: " << iter << "\nepsilon = " << std::setprecision(8) << curr << "\n";
    if (curr <= L.target_epsilon) {
      std::cerr << "reached target epsilon = " << L.target_epsilon << ", at iter = " << shown
                << ", epsilon = " << curr << "\n\n";
      // the ART warm start only ends its own phase (sirt-mimo-seq.cu:410-414, 508)
      if (L.rule == Rule::ArtThenWeighted && iter < L.art_iterations) {
        prev = curr;
        iter = L.art_iterations - 1;
        continue;
      }
      break;
    }
    if (curr > prev || std::isnan(curr)) {
      std::cerr << "diverged at iter = " << shown << ", epsilon = " << curr << "\n\n";
      break;
    }
    prev = curr;
  }
  const float secs = std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
  std::cout << std::endl;
  std::cerr << "time (s): " << secs << "\n";
  return iter;
}

void write_field(tomo_ctx *ctx, const std::string &name, int nx, int ny) {
  std::vector<float> f((size_t)nx * ny);
  check(tomo_get_field(ctx, f.data()), ctx, "tomo_get_field");
  check(tomo_write_text(name.c_str(), f.data(), nx, ny, 1, 1, 0, 1), ctx, "write field");
}

}  // namespace drv
