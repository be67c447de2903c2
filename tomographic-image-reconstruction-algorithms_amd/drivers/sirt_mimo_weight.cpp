// sirt-mimo-weight <fo> <G> <target eps> <iters> <alpha> <beta> <omega>
// Averaged intermediate fields with a decaying relaxation
// (sirt-mimo-weight.cu:65-90, 786-805): f_g = f + omega (alpha + beta / k)
// mask df_g at 1-based iteration k, f = mean_g f_g.
#include "sirt_variant_common.h"

namespace {
void sched(const float *p, double o[4]) {  // p = alpha, beta, omega
  o[0] = p[2];
  o[1] = p[0];
  o[2] = p[1];
  o[3] = 1.0;
}
}  // namespace

int main(int argc, char **argv) {
  const drv::Variant V{" <alpha> <beta> <omega>", drv::Rule::Uniform, 0, std::numeric_limits<float>::infinity(),
                       false, "sirt_weight_convergence.txt", "sirt_time.txt", "sirt-", 3, sched};
  return drv::variant_main(argc, argv, V);
}
