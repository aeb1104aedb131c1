// Eigenvector files in DiagHam's binary vector format, read and written from device memory.
//
// RealVector::WriteVector (src/Vector/RealVector.cc:3669) and ComplexVector::WriteVector
// (src/Vector/ComplexVector.cc:2606): a little-endian int32 dimension, or -1 followed by an int64
// dimension when it does not fit in 31 bits, then the components as little-endian doubles (complex
// vectors: interleaved real and imaginary parts, the device layout of complex vectors here).
// RealVector::ReadVector (:3798) reads the same layout back.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "diagham_host.hpp"

namespace dhh {

namespace {
const int64_t kChunk = 1 << 22;  // doubles per host staging chunk

struct File {
  FILE* f;
  File(const char* name, const char* mode) : f(std::fopen(name, mode)) {
    if (!f) throw std::invalid_argument(std::string("cannot open the file: ") + name);
  }
  ~File() {
    if (f) std::fclose(f);
  }
};
}  // namespace

void WriteDeviceVector(GPUArchitecture* arch, const char* fileName, const double* d_vec, int64_t dim, bool complex) {
  if (dim < 0) throw std::invalid_argument("negative dimension");
  File file(fileName, "wb");
  if (dim < (int64_t(1) << 31)) {
    int32_t d = (int32_t)dim;
    if (std::fwrite(&d, sizeof(d), 1, file.f) != 1) throw std::runtime_error("write failed");
  } else {
    int32_t flag = -1;
    if (std::fwrite(&flag, sizeof(flag), 1, file.f) != 1 || std::fwrite(&dim, sizeof(dim), 1, file.f) != 1)
      throw std::runtime_error("write failed");
  }
  const int64_t n = complex ? 2 * dim : dim;
  std::vector<double> buffer((size_t)std::min(n, kChunk));
  for (int64_t i = 0; i < n; i += kChunk) {
    const int64_t m = std::min(kChunk, n - i);
    arch->Check(dh_memcpy_d2h(arch->Context(), buffer.data(), d_vec + i, sizeof(double) * m), "vector download");
    if (std::fwrite(buffer.data(), sizeof(double), (size_t)m, file.f) != (size_t)m)
      throw std::runtime_error("write failed");
  }
}

int64_t VectorFileDimension(const char* fileName, bool complex, int64_t* headerBytes) {
  File file(fileName, "rb");
  int32_t d;
  int64_t dim;
  if (std::fread(&d, sizeof(d), 1, file.f) != 1) throw std::invalid_argument("not a vector file");
  int64_t header = sizeof(d);
  if (d >= 0) {
    dim = d;
  } else {
    if (std::fread(&dim, sizeof(dim), 1, file.f) != 1) throw std::invalid_argument("not a vector file");
    header += sizeof(dim);
  }
  std::fseek(file.f, 0, SEEK_END);
  const int64_t payload = (int64_t)std::ftell(file.f) - header;
  const int64_t expect = (complex ? 2 : 1) * dim * (int64_t)sizeof(double);
  if (payload != expect)
    throw std::invalid_argument(payload == 2 * dim * (int64_t)sizeof(double)
                                    ? "the file holds a complex vector"
                                    : "the file size does not match its dimension");
  if (headerBytes) *headerBytes = header;
  return dim;
}

void ReadDeviceVector(GPUArchitecture* arch, const char* fileName, double* d_vec, int64_t dim, bool complex) {
  int64_t header = 0;
  const int64_t fileDim = VectorFileDimension(fileName, complex, &header);
  if (fileDim != dim) throw std::invalid_argument("the vector in the file has another dimension");
  File file(fileName, "rb");
  // the components start after the header the file actually holds: a writer may use the -1 and
  // int64 form for a dimension that fits in 31 bits
  std::fseek(file.f, (long)header, SEEK_SET);
  const int64_t n = complex ? 2 * dim : dim;
  std::vector<double> buffer((size_t)std::min(n, kChunk));
  for (int64_t i = 0; i < n; i += kChunk) {
    const int64_t m = std::min(kChunk, n - i);
    if (std::fread(buffer.data(), sizeof(double), (size_t)m, file.f) != (size_t)m)
      throw std::runtime_error("read failed");
    arch->Check(dh_memcpy_h2d(arch->Context(), d_vec + i, buffer.data(), sizeof(double) * m), "vector upload");
  }
  arch->Check(dh_synchronize(arch->Context()), "sync");
}

}  // namespace dhh
