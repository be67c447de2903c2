// generate-data <fo filename> <sensor group size> [--fp64] [--noise <rel> <seed>]
// Forward-propagates every source group on fo and writes the four sensor-line
// trace files <base>-data-{bottom,right,top,left}-<G>.txt
// (generate-data.cu:57-224).  --fp64 uses the FP64 solver.  --noise adds
// seeded Gaussian noise of standard deviation rel * RMS(trace) to every
// receiver trace (tomo_add_noise) before the files are written; it stands
// in for the reference's noisy generators (generate-data-art.cu:393-402,
// generate-data-noise.cu:199), whose unseeded streams cannot be reproduced.
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "driver_common.h"

int main(int argc, char **argv) {
  bool fp64 = false, noise = false, ok = argc >= 3;
  double rel = 0.0;
  unsigned long long seed = 0;
  for (int a = 3; ok && a < argc; ++a) {
    if (std::strcmp(argv[a], "--fp64") == 0) {
      fp64 = true;
    } else if (std::strcmp(argv[a], "--noise") == 0 && a + 2 < argc) {
      char *end1 = nullptr, *end2 = nullptr;
      rel = std::strtod(argv[a + 1], &end1);
      seed = std::strtoull(argv[a + 2], &end2, 10);
      ok = *end1 == '\0' && *end2 == '\0' && rel >= 0.0;
      noise = true;
      a += 2;
    } else {
      ok = false;
    }
  }
  if (!ok) {
    std::cerr << "Usage: " << argv[0] << " <fo filename> <sensor group size> [--fp64] [--noise <rel> <seed>]\n\n";
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
  drv::check(tomo_forward(ctx, 0, ng, fp64 ? 1 : 0, tr.data()), ctx, "tomo_forward");
  if (noise) drv::check(tomo_add_noise(tr.data(), ng, cfg.nt, nr, rel, seed), ctx, "tomo_add_noise");
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
