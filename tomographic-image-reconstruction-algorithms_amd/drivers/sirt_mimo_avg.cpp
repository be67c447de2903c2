// sirt-mimo-avg <fo> <G> <target eps> <iters>
// Averaged intermediate fields (sirt-mimo-avg.cu:55-107, 386-460):
// f_g = f + 20000 mask df_g per group, f = mean_g f_g, i.e. the uniform
// update in mean mode at the constant relaxation 20000.
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
  const drv::Variant V{"", drv::Rule::Uniform, 0, std::numeric_limits<float>::infinity(), false,
                       "sirt_convergence.txt", "sirt_time.txt", "sirt-", 0, sched};
  return drv::variant_main(argc, argv, V);
}
