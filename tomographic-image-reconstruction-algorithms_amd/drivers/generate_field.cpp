// generate-field: writes fo.txt, the reference phantom on the 192 x 192 grid
// in the mimo-io text layout (generate-field.cu:74-113).
#include <iostream>
#include <vector>

#include "driver_common.h"

int main() {
  tomo_config cfg;
  tomo_config_reference(&cfg);
  std::vector<float> fo((size_t)cfg.nx * cfg.ny);
  drv::check(tomo_phantom_reference(fo.data(), cfg.nx, cfg.ny, cfg.h), nullptr, "phantom");
  std::cerr << "writing to 'fo.txt'...\n\n";
  drv::check(tomo_write_text("fo.txt", fo.data(), cfg.nx, cfg.ny, 1, 1, 0, 1), nullptr, "write fo.txt");
  return 0;
}
