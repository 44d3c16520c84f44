/*
 * Dense matrix files in the formats of the reference's matrixIO (see
 * matrix_io.cpp).  Output matrices are column-major with leading dimension
 * ld; errors throw std::invalid_argument (bad file or arguments) or
 * std::runtime_error (a failed write).
 */
#ifndef DENSE_HODLR_MATRIX_IO_H
#define DENSE_HODLR_MATRIX_IO_H

namespace hodlr {

void binary_matrix_dims(const char* path, int* rows, int* cols);
void read_binary_matrix(const char* path, int rows, int cols, double* out, long long ld);
void write_binary_matrix(const char* path, int rows, int cols, const double* A, long long lda);
void text_matrix_dims(const char* path, int* rows, int* cols);
void read_text_matrix(const char* path, int rows, int cols, double* out, long long ld);

}  // namespace hodlr

#endif
