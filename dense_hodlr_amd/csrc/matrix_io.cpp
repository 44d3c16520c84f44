/*
 * Dense matrix files of the reference's matrixIO (src/matrixIO.cpp), read and
 * written on the host:
 *
 *  binary  (saveMatrixXdToBinary / readBinaryIntoMatrixXd, :165-208): a
 *          32-bit int row count, a 32-bit int column count, then every entry
 *          as an 8-byte double in row-major order, all in the machine's byte
 *          order (little-endian here, as on the x86 hosts the reference
 *          targets);
 *  text    (readTxtIntoMatrix, :11-50): a first line "rows cols", then one
 *          line "i j value" per entry with 0-based indices.
 *
 * Matrices cross the C ABI column-major with a leading dimension, as every
 * other matrix does; the row-major file order is transposed while streaming
 * (one file row at a time, so a file of any size needs one row of buffer).
 */
#include "matrix_io.h"

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hodlr {

namespace {

struct File {
  std::FILE* f = nullptr;
  File(const char* path, const char* mode) {
    f = std::fopen(path, mode);
    if (!f)
      throw std::invalid_argument(std::string("file ") + path + " could not be opened: " +
                                  std::strerror(errno));
  }
  ~File() {
    if (f) std::fclose(f);
  }
};

void check_dims(long long rows, long long cols, const char* path) {
  if (rows < 0 || cols < 0 || rows > 0x7fffffffLL || cols > 0x7fffffffLL)
    throw std::invalid_argument(std::string("file ") + path + " has an invalid matrix size");
}

}  // namespace

void binary_matrix_dims(const char* path, int* rows, int* cols) {
  File fh(path, "rb");
  int32_t hdr[2];
  if (std::fread(hdr, sizeof(int32_t), 2, fh.f) != 2)
    throw std::invalid_argument(std::string("file ") + path + " is shorter than its header");
  check_dims(hdr[0], hdr[1], path);
  *rows = hdr[0];
  *cols = hdr[1];
}

void read_binary_matrix(const char* path, int rows, int cols, double* out, long long ld) {
  File fh(path, "rb");
  int32_t hdr[2];
  if (std::fread(hdr, sizeof(int32_t), 2, fh.f) != 2)
    throw std::invalid_argument(std::string("file ") + path + " is shorter than its header");
  if (hdr[0] != rows || hdr[1] != cols)
    throw std::invalid_argument(std::string("file ") + path + " holds a matrix of another size");
  if (ld < rows) throw std::invalid_argument("leading dimension < rows");
  std::vector<double> row((size_t)cols);
  for (int i = 0; i < rows; i++) {
    if (std::fread(row.data(), sizeof(double), (size_t)cols, fh.f) != (size_t)cols)
      throw std::invalid_argument(std::string("file ") + path + " ends before its last entry");
    for (int j = 0; j < cols; j++) out[(size_t)i + (size_t)j * ld] = row[j];
  }
}

void write_binary_matrix(const char* path, int rows, int cols, const double* A, long long lda) {
  if (rows < 0 || cols < 0 || lda < rows) throw std::invalid_argument("bad matrix size");
  File fh(path, "wb");
  const int32_t hdr[2] = {rows, cols};
  if (std::fwrite(hdr, sizeof(int32_t), 2, fh.f) != 2)
    throw std::runtime_error(std::string("write to ") + path + " failed");
  std::vector<double> row((size_t)cols);
  for (int i = 0; i < rows; i++) {
    for (int j = 0; j < cols; j++) row[j] = A[(size_t)i + (size_t)j * lda];
    if (std::fwrite(row.data(), sizeof(double), (size_t)cols, fh.f) != (size_t)cols)
      throw std::runtime_error(std::string("write to ") + path + " failed");
  }
  if (std::fflush(fh.f) != 0) throw std::runtime_error(std::string("write to ") + path + " failed");
}

void text_matrix_dims(const char* path, int* rows, int* cols) {
  File fh(path, "r");
  long long r = -1, c = -1;
  if (std::fscanf(fh.f, "%lld %lld", &r, &c) != 2)
    throw std::invalid_argument(std::string("file ") + path + ": bad format (first line)");
  check_dims(r, c, path);
  *rows = (int)r;
  *cols = (int)c;
}

/* Entries not listed in the file are zero (the reference leaves them
 * uninitialized); at most rows * cols entry lines are read, as in the
 * reference's loop; an index outside the matrix is an error. */
void read_text_matrix(const char* path, int rows, int cols, double* out, long long ld) {
  File fh(path, "r");
  long long r = -1, c = -1;
  if (std::fscanf(fh.f, "%lld %lld", &r, &c) != 2)
    throw std::invalid_argument(std::string("file ") + path + ": bad format (first line)");
  if (r != rows || c != cols)
    throw std::invalid_argument(std::string("file ") + path + " holds a matrix of another size");
  if (ld < rows) throw std::invalid_argument("leading dimension < rows");
  for (int j = 0; j < cols; j++)
    for (int i = 0; i < rows; i++) out[(size_t)i + (size_t)j * ld] = 0.0;
  const long long maxe = (long long)rows * cols;
  for (long long e = 0; e < maxe; e++) {
    long long i, j;
    double v;
    const int got = std::fscanf(fh.f, "%lld %lld %lf", &i, &j, &v);
    if (got == EOF) break;
    if (got != 3)
      throw std::invalid_argument(std::string("file ") + path + ": bad format (entry line " +
                                  std::to_string(e + 2) + ")");
    if (i < 0 || i >= rows || j < 0 || j >= cols)
      throw std::invalid_argument(std::string("file ") + path + ": entry index out of range");
    out[(size_t)i + (size_t)j * ld] = v;
  }
}

}  // namespace hodlr
