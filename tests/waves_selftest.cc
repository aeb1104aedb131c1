// CPU check of the wave planner of the batched sector solves (RunSectorWaves): a scripted device
// of `free` bytes runs each wave with Krylov storage for the largest number of blocks that fits
// all of the wave's rows; a sector converges when it needs no more blocks than that, otherwise it
// stops deferred at the storage's limit.  Reads "k kr nev maxIter free nbrSectors" then per sector
// "dim need" from argv[1]; prints "waves W" and per sector "converged attempts blocks".
#include <cstdio>
#include <exception>
#include <vector>

#include "diagham_host.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = std::fopen(argv[1], "r");
  if (!f) return 2;
  int k, kr, nev, maxIter, ns;
  unsigned long long freeBytes;
  if (std::fscanf(f, "%d %d %d %d %llu %d", &k, &kr, &nev, &maxIter, &freeBytes, &ns) != 6) return 2;
  std::vector<int64_t> dims(ns);
  std::vector<int> need(ns), attempts(ns, 0);
  for (int s = 0; s < ns; ++s) {
    long long d;
    if (std::fscanf(f, "%lld %d", &d, &need[s]) != 2) return 2;
    dims[s] = d;
  }
  std::fclose(f);
  std::vector<dhh::SectorResult> out(ns);
  try {
    const int waves = dhh::RunSectorWaves(
        dims, k, kr, nev, maxIter, (size_t)-1, out,
        [&](const std::vector<int>& wave, int) {
          int64_t rows = 0;
          for (int s : wave) rows += dims[s];
          const int64_t fit = (int64_t)(freeBytes / (8 * (uint64_t)rows) - 1) / kr - 2;
          for (int s : wave) {
            ++attempts[s];
            dhh::SectorResult& r = out[s];
            r.products += k * (int)std::min<int64_t>(need[s], fit);
            if (need[s] <= fit) {
              r.converged = true;
              r.deferred = false;
              r.blocks = need[s];
            } else {
              r.converged = false;
              r.deferred = true;
              r.blocks = (int)fit;
            }
          }
        },
        [&] { return (size_t)freeBytes + ((size_t)2 << 30); });
    std::printf("waves %d\n", waves);
  } catch (const std::exception& e) {
    std::printf("error %s\n", e.what());
    return 0;
  }
  for (int s = 0; s < ns; ++s) std::printf("%d %d %d\n", out[s].converged ? 1 : 0, attempts[s], out[s].blocks);
  return 0;
}
