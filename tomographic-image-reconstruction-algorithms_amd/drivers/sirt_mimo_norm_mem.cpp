// sirt-mimo-norm-mem <fo> <G> <Np> <target eps> <iters> <omega> <alpha> <beta> <gamma>
// Weighted SIRT computing Np groups per pass (sirt-mimo-norm-mem.cu:122-530).
#include "sirt_io_common.h"

int main(int argc, char **argv) {
  return drv::sirt_io_main(argc, argv, drv::Rule::Weighted, true, "sirt-norm-", "sirt");
}
