// sirt-mimo-seq <fo> <G> <target eps> <iters> <omega> <alpha> <beta> <gamma>
// ART warm start (5 iterations) then non-uniform weighted SIRT, one power
// schedule and one set of stop rules across both phases
// (sirt-mimo-seq.cu:75-620); output sirt-norm-<G>-<base>.txt.
#include "sirt_io_common.h"

int main(int argc, char **argv) {
  return drv::sirt_io_main(argc, argv, drv::Rule::ArtThenWeighted, false, "sirt-norm-", "sirt");
}
