// driver_common.h -- shared pieces of the command-line drivers that mirror
// the reference executables (argument checks, file names, the iteration
// loop with its stop rules and logs).
#pragma once

#include <string>
#include <vector>

#include "tomo.h"

namespace drv {

// "<base>.<ext>" with exactly one period (sirt-mimo-io.cu:68-73); exits 1 otherwise
std::string base_name(const std::string &fo_filename);

// Print "Error: ..." for a failed ABI call and exit 1.
void check(int status, tomo_ctx *ctx, const char *what);

// reference preset context with the square ring and groups of G; returns Ng
tomo_ctx *reference_context(int group_size, int shots_per_pass, int *ng);

// Read fo (192 x 192) and the four side files <base>-data-<side>-<G>.txt
// into the context (sirt-mimo-io.cu:126-181).  Returns false (after printing
// the reference's message) when a file is missing.
bool load_inputs(tomo_ctx *ctx, const std::string &fo_filename, int group_size, std::vector<float> &fo);

// Art then Weighted: the first art_iterations iterations are ART, the rest
// weighted SIRT (sirt-mimo-seq.cu:337-596); AvgWeight: the weighted average
// of intermediate fields (sirt-mimo-avg-weight.cu:526-544)
enum class Rule { Uniform, Weighted, Art, ArtThenWeighted, AvgWeight };

struct Loop {
  Rule rule;
  int mode = TOMO_UPDATE_MEAN;
  float target_epsilon;
  int max_iterations;          // -1: unbounded
  float prev_epsilon_init;     // 100 (io drivers) or +inf (sirt-mimo.cu:250)
  std::string convergence_file, time_file;
  bool table = true;           // the io drivers' iter/epsilon/difference table
  bool one_based = false;      // iteration numbers printed from 1 (norm-mem drivers)
  int art_iterations = 5;      // ArtThenWeighted: length of the ART phase
};

// Run the loop; scale(k, user) gives the relaxation of 0-based iteration k.
// Stops at epsilon <= target or epsilon > previous / NaN, as the reference
// (sirt-mimo-io.cu:379-393).  Returns the number of iterations run.
int run_loop(tomo_ctx *ctx, const Loop &L, double (*scale)(int, const void *), const void *user);

void write_field(tomo_ctx *ctx, const std::string &name, int nx, int ny);

}  // namespace drv
