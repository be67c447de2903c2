// generate-data <fo filename> <sensor group size> [--fp64]
// Forward-propagates every source group on fo and writes the four sensor-line
// trace files <base>-data-{bottom,right,top,left}-<G>.txt
// (generate-data.cu:57-224).  --fp64 uses the FP64 solver.
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "driver_common.h"

int main(int argc, char **argv) {
  if (argc != 3 && !(argc == 4 && std::strcmp(argv[3], "--fp64") == 0)) {
    std::cerr << "Usage: " << argv[0] << " <fo filename> <sensor group size> [--fp64]\n\n";
    return 1;
  }
  const std::string fo_name = argv[1];
  const int G = std::stoi(argv[2]);
  const std::string base = drv::base_name(fo_name);
  if (!std::ifstream(fo_name)) {
    std::cerr << "Error: '" + fo_name + "' file not found in current directory.\n\n";
    return 1;
  }
  int ng = 0;
  tomo_ctx *ctx = drv::reference_context(G, 0, &ng);
  tomo_config cfg;
  tomo_config_reference(&cfg);
  std::vector<float> fo((size_t)cfg.nx * cfg.ny);
  drv::check(tomo_read_text(fo_name.c_str(), fo.data(), cfg.nx, cfg.ny, 1, 1), ctx, "read fo");
  drv::check(tomo_set_field(ctx, fo.data()), ctx, "tomo_set_field");
  const int nr = tomo_num_receivers(ctx);
  std::vector<float> tr((size_t)ng * cfg.nt * nr);
  drv::check(tomo_forward(ctx, 0, ng, argc == 4, tr.data()), ctx, "tomo_forward");
  std::vector<std::vector<float>> s(4, std::vector<float>((size_t)ng * cfg.nt * cfg.nx));
  drv::check(tomo_traces_to_sides(ctx, tr.data(), s[0].data(), s[1].data(), s[2].data(), s[3].data()), ctx,
             "traces_to_sides");
  const char *sides[4] = {"bottom", "right", "top", "left"};
  for (int q = 0; q < 4; ++q) {
    const std::string name = base + "-data-" + sides[q] + "-" + std::to_string(G) + ".txt";
    std::cerr << "writing to '" << name << "'...\n\n";
    drv::check(tomo_write_text(name.c_str(), s[q].data(), cfg.nx, cfg.nt, ng, 1, 0, 1), ctx, "write data");
  }
  tomo_destroy(ctx);
  return 0;
}
