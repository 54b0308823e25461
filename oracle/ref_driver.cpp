// Test-only driver for the reference's own CPU path (oracle checker).
//
// The reference's `fem/src/main.cpp:10-36` hard-codes its config path, task
// name and device and does not compile without CUDA (`main.cpp:20`), so this
// file stands in for that `main` only: it builds the reference's
// `dataKeeper(config, task)` (`DataKeeper/src/datakeeper.cpp:246-252`), calls
// the reference's CPU solver entry points (`femfunc/src/femfunc.cpp:3-23`,
// `:130-381`) exactly as `main.cpp:25-28` does, and prints the displacements
// with 9 significant digits (enough to round-trip FP32).
//
// Usage: ref_driver <config.json> <task_name> <out.txt>
// The output file holds one line "n_nodes dim" and then n_nodes lines of
// dim displacements.  Nothing in the product links or runs this program.
#include <cstdio>
#include <string>

#include <femfunc.h>
#include <datakeeper.h>

int main(int argc, char *argv[]) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s <config.json> <task> <out.txt>\n", argv[0]);
    return 2;
  }
  dataKeeper dk(argv[1], argv[2]);
  if (dk.getTaskType() == "dynamic")
    gpuCalculateFEM_DYN(CPU, dk, false);
  else
    gpuCalculateFEM_EbE_vec(CPU, dk, false);

  Matrix *u = dk.get_displacements();
  std::FILE *f = std::fopen(argv[3], "w");
  if (!f) return 3;
  std::fprintf(f, "%zu %zu\n", u->get_numRows(), u->get_numCols());
  for (size_t i = 0; i < u->get_numRows(); ++i) {
    for (size_t j = 0; j < u->get_numCols(); ++j)
      std::fprintf(f, j ? " %.9g" : "%.9g", (*u)(i, j));
    std::fprintf(f, "\n");
  }
  std::fclose(f);
  return 0;
}
