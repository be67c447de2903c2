// art-mimo-io <fo> <G> <target eps> <iters> <omega> <alpha> <beta> <gamma>
// Sequential ART over the source groups (art-mimo-io.cu:405-740).
#include "sirt_io_common.h"

int main(int argc, char **argv) {
  return drv::sirt_io_main(argc, argv, drv::Rule::Art, false, "art-", "art");
}
