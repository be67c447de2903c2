// sirt-mimo-norm <fo> <G> <target eps> <iters> <omega> <alpha> <beta> <gamma>
// Non-uniform weighted SIRT (sirt-mimo-norm.cu:122-500).
#include "sirt_io_common.h"

int main(int argc, char **argv) {
  return drv::sirt_io_main(argc, argv, drv::Rule::Weighted, false, "sirt-norm-", "sirt");
}
