// sirt-mimo-avg-weight <fo> <G> <target eps> <iters>
// Distance-weighted sum of intermediate fields (sirt-mimo-avg-weight.cu:
// 80-203, 386-590): f_g = f + 20000 mask df_g per group, w_g = ||f_g -
// mean_g f_g||, f += sum_g f_g / w_g (tomo_update_avg_weight at the
// reference's constant 20000).
#include "sirt_variant_common.h"

namespace {
void sched(const float *, double o[4]) {
  o[0] = 20000.0;
  o[1] = 1.0;
  o[2] = 0.0;
  o[3] = 1.0;
}
}  // namespace

int main(int argc, char **argv) {
  const drv::Variant V{"", drv::Rule::AvgWeight, 0, std::numeric_limits<float>::infinity(), false,
                       "sirt_convergence.txt", "sirt_time.txt", "sirt-", 0, sched};
  return drv::variant_main(argc, argv, V);
}
