// sirt-iter <fo> <G> <target eps> <iters> <omega> <alpha>
// Group-serial SIRT (sirt-iter.cu:407-438, 622-710): the gradient of one
// group at a time summed over the groups, then f += s mask (sum_g df_g) / Ng
// with the growing relaxation s = omega (alpha + (1 - alpha) (k+1)).
#include "sirt_variant_common.h"

namespace {
void sched(const float *p, double o[4]) {  // p = omega, alpha
  o[0] = p[0];
  o[1] = p[1];
  o[2] = 1.0 - (double)p[1];
  o[3] = -1.0;  // beta (k+1)
}
}  // namespace

int main(int argc, char **argv) {
  const drv::Variant V{" <omega> <alpha>", drv::Rule::Uniform, 1, 100.f, true, "art_convergence.txt",
                       "art_time.txt", "art-", 2, sched};
  return drv::variant_main(argc, argv, V);
}
