// sirt-mimo <sensor group size> <target epsilon> <max iterations> <omega>
// Self-contained SIRT (sirt-mimo.cu:66-387): builds the phantom, makes the
// data with the same solver, then iterates f += omega mask sum_g G_g / (1 + f);
// writes dev_f0.txt and dev_f.txt in the IO_Files layout (sirt-mimo.cu:389-425).
#include <iomanip>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "driver_common.h"

namespace {
double const_scale(int, const void *u) { return *static_cast<const float *>(u); }
}  // namespace

int main(int argc, char **argv) {
  if (argc != 5) {
    std::cerr << "Usage: " << argv[0] << " <sensor group size> <target epsilon> <max iterations> <omega>\n\n";
    return 1;
  }
  const int G = std::stoi(argv[1]);
  drv::Loop L;
  L.rule = drv::Rule::Uniform;
  L.mode = TOMO_UPDATE_SUM;
  L.target_epsilon = std::stof(argv[2]);
  L.max_iterations = std::stoi(argv[3]);
  const float omega = std::stof(argv[4]);
  L.prev_epsilon_init = std::numeric_limits<float>::infinity();
  L.convergence_file = "sirt_convergence.txt";
  L.time_file = "sirt_time.txt";
  L.table = false;
  std::cout << std::fixed << std::setprecision(10);
  std::cerr << std::fixed << std::setprecision(10);
  int ng = 0;
  tomo_ctx *ctx = drv::reference_context(G, 0, &ng);
  tomo_config cfg;
  tomo_config_reference(&cfg);
  std::vector<float> fo((size_t)cfg.nx * cfg.ny), zero(fo.size(), 0.f);
  drv::check(tomo_phantom_reference(fo.data(), cfg.nx, cfg.ny, cfg.h), ctx, "phantom");
  drv::check(tomo_set_field(ctx, fo.data()), ctx, "tomo_set_field");
  std::vector<float> tr((size_t)ng * cfg.nt * tomo_num_receivers(ctx));
  drv::check(tomo_forward(ctx, 0, ng, 0, tr.data()), ctx, "tomo_forward");
  drv::check(tomo_set_measured(ctx, tr.data()), ctx, "tomo_set_measured");
  drv::check(tomo_set_truth(ctx, fo.data()), ctx, "tomo_set_truth");
  drv::check(tomo_set_field(ctx, zero.data()), ctx, "tomo_set_field");
  std::cerr << "group size:     " << G << "\n"
            << "target epsilon: " << std::setprecision(2) << L.target_epsilon << "\n"
            << "omega:          " << std::setprecision(0) << omega << "\n\n";
  drv::run_loop(ctx, L, const_scale, &omega);
  std::vector<float> f(fo.size());
  drv::check(tomo_get_field(ctx, f.data()), ctx, "tomo_get_field");
  std::cerr << "writing f0 to 'dev_f0.txt'...\nwriting f to 'dev_f.txt'...\n\n";
  drv::check(tomo_write_text("dev_f0.txt", fo.data(), cfg.nx, cfg.ny, 1, 1, 0, 0), ctx, "write dev_f0.txt");
  drv::check(tomo_write_text("dev_f.txt", f.data(), cfg.nx, cfg.ny, 1, 1, 0, 0), ctx, "write dev_f.txt");
  tomo_destroy(ctx);
  return 0;
}
