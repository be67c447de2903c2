// sirt-norm-test-weight <fo> <G> <Np> <target eps> <iters>
// Non-uniform weighted SIRT with the ground-truth line search
// (sirt-norm-test-weight.cu:123-565); output sirt-norm-<G>-<base>.txt.
#include "line_search_common.h"

int main(int argc, char **argv) { return drv::line_search_main(argc, argv, true, "sirt-norm-"); }
