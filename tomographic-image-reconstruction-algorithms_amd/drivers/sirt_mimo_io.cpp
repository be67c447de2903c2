// sirt-mimo-io <fo> <G> <target eps> <iters> <omega> <alpha> <beta> <gamma>
// Uniform SIRT from reference data files (sirt-mimo-io.cu:58-417):
// f += s_k mask (sum_g G_g) / (1 + f) / Ng, s_k = omega (alpha + beta / (k+1)^gamma).
#include "sirt_io_common.h"

int main(int argc, char **argv) {
  return drv::sirt_io_main(argc, argv, drv::Rule::Uniform, false, "sirt-", "sirt");
}
