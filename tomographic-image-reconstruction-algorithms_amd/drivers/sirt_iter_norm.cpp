// sirt-iter-norm <fo> <G> <target eps> <iters> <omega> <alpha>
// Group-serial weighted SIRT (sirt-iter-norm.cu:390-460, 686-784): per-group
// gradients one group at a time, weights w_g = 1 / |df_g - mean| normalised
// to sum 1, f += s mask sum_g w_g df_g with s = omega (alpha + (1 - alpha) /
// (k+1)^1.5).
#include "sirt_variant_common.h"

namespace {
void sched(const float *p, double o[4]) {  // p = omega, alpha
  o[0] = p[0];
  o[1] = p[1];
  o[2] = 1.0 - (double)p[1];
  o[3] = 1.5;
}
}  // namespace

int main(int argc, char **argv) {
  const drv::Variant V{" <omega> <alpha>", drv::Rule::Weighted, 1, 100.f, true, "art_convergence.txt",
                       "art_time.txt", "art-", 2, sched};
  return drv::variant_main(argc, argv, V);
}
