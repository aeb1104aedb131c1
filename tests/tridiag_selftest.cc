// CPU checks of the tridiagonal routines of the one-vector Lanczos.  With one argument:
// TridiagonalEigenvector (inverse iteration, for the residual test): reads "n theta" then n
// diagonal and n - 1 off-diagonal values from argv[1]; prints the eigenvector, one component per
// line.  With a second argument "lowest": TridiagonalLowest (the convergence test's eigenvalues):
// reads "n count nh", n diagonal and n - 1 off-diagonal values and nh hint values; prints the
// count lowest eigenvalues, one per line.
#include <cstdio>
#include <string>
#include <vector>

#include "block_krylov.hpp"
#include "diagham_host.hpp"

static int lowest(FILE* f) {
  int n, count, nh;
  if (std::fscanf(f, "%d %d %d", &n, &count, &nh) != 3) return 2;
  std::vector<double> d(n), e(n > 0 ? n - 1 : 0), hint(nh), ev;
  for (double& v : d)
    if (std::fscanf(f, "%lf", &v) != 1) return 2;
  for (double& v : e)
    if (std::fscanf(f, "%lf", &v) != 1) return 2;
  for (double& v : hint)
    if (std::fscanf(f, "%lf", &v) != 1) return 2;
  std::fclose(f);
  dhh::TridiagonalLowest(d, e, count, ev, nh > 0 ? &hint : nullptr);
  for (double v : ev) std::printf("%.17g\n", v);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = std::fopen(argv[1], "r");
  if (!f) return 2;
  if (argc > 2 && std::string(argv[2]) == "lowest") return lowest(f);
  int n;
  double theta;
  if (std::fscanf(f, "%d %lf", &n, &theta) != 2) return 2;
  std::vector<double> d(n), e(n > 0 ? n - 1 : 0);
  for (double& v : d)
    if (std::fscanf(f, "%lf", &v) != 1) return 2;
  for (double& v : e)
    if (std::fscanf(f, "%lf", &v) != 1) return 2;
  std::fclose(f);
  for (double v : dhh::TridiagonalEigenvector(d, e, theta)) std::printf("%.17g\n", v);
  return 0;
}
