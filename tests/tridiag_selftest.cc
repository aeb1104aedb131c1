// CPU check of TridiagonalEigenvector (inverse iteration on a symmetric tridiagonal matrix, used by
// the residual test of the one-vector Lanczos): reads "n theta" then n diagonal and n - 1
// off-diagonal values from argv[1]; prints the eigenvector, one component per line.
#include <cstdio>
#include <vector>

#include "diagham_host.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = std::fopen(argv[1], "r");
  if (!f) return 2;
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
