// sirt-test-weight <fo> <G> <Np> <target eps> <iters>
// Uniform SIRT with the ground-truth line search (sirt-test-weight.cu:59-460);
// output sirt-<G>-<base>.txt.
#include "line_search_common.h"

int main(int argc, char **argv) { return drv::line_search_main(argc, argv, false, "sirt-"); }
