// The reference's integrals file (and.integrals.bin): Fortran sequential unformatted records with
// 4-byte little-endian length markers before and after each record (the gfortran / amdflang
// default). Stokes: WriteIntegralsToDiskStokes / ReadIntegralsFileStokes (src/lib_Stokes.f90:
// 2742-2872); Laplace: WriteIntegralsToDisk / ReadIntegralsFile (src/lib_int.f90:185-221, 136-180).
// Every matrix is written column by column (WrSMat, src/lib_util.f90:201-215), each column as
// records of 8192 values and one last record with the rest (wrvec, :264-287), empty when the
// column length is a multiple of 8192. Host code only: no device work.
#include "andr_internal.h"

#include <cstdio>
#include <cstring>
#include <vector>

namespace {

constexpr int32_t ID_STOKES = 1407;  // parStokes (src/mod_par.f90:109)
constexpr int64_t CHUNK = 8192;      // maxsize of wrvec / rdvec

struct File {
  FILE *f = nullptr;
  explicit File(const char *path, const char *mode) { f = std::fopen(path, mode); }
  ~File() {
    if (f) std::fclose(f);
  }
  bool record(const void *data, int64_t bytes) {  // one unformatted record
    const int32_t mark = (int32_t)bytes;
    return std::fwrite(&mark, 4, 1, f) == 1 && (bytes == 0 || std::fwrite(data, 1, bytes, f) == (size_t)bytes) &&
           std::fwrite(&mark, 4, 1, f) == 1;
  }
  bool read_record(void *data, int64_t bytes) {  // a record of exactly `bytes`
    int32_t a = -1, b = -1;
    if (std::fread(&a, 4, 1, f) != 1 || a != bytes) return false;
    if (bytes > 0 && std::fread(data, 1, bytes, f) != (size_t)bytes) return false;
    return std::fread(&b, 4, 1, f) == 1 && b == a;
  }
};

// a row-major nrow x ncol matrix, written as the reference's column-major one
bool write_matrix(File &F, const double *A, int64_t nrow, int64_t ncol) {
  std::vector<double> col(nrow);
  for (int64_t j = 0; j < ncol; ++j) {
    for (int64_t i = 0; i < nrow; ++i) col[i] = A[i * ncol + j];
    const int64_t nblock = nrow / CHUNK;
    for (int64_t b = 0; b < nblock; ++b)
      if (!F.record(col.data() + b * CHUNK, CHUNK * 8)) return false;
    if (!F.record(col.data() + nblock * CHUNK, (nrow - nblock * CHUNK) * 8)) return false;
  }
  return true;
}

bool read_matrix(File &F, double *A, int64_t nrow, int64_t ncol) {
  std::vector<double> col(nrow);
  for (int64_t j = 0; j < ncol; ++j) {
    const int64_t nblock = nrow / CHUNK;
    for (int64_t b = 0; b < nblock; ++b)
      if (!F.read_record(col.data() + b * CHUNK, CHUNK * 8)) return false;
    if (!F.read_record(col.data() + nblock * CHUNK, (nrow - nblock * CHUNK) * 8)) return false;
    if (A)
      for (int64_t i = 0; i < nrow; ++i) A[i * ncol + j] = col[i];
  }
  return true;
}

bool write_header(File &F, bool stokes, int32_t nsp, int32_t nn, int32_t nq, const double err[2]) {
  const int32_t sizes[3] = {nsp, nn, nq};
  const double e[2] = {err ? err[0] : 0.0, err ? err[1] : 0.0};
  return (!stokes || F.record(&ID_STOKES, 4)) && F.record(sizes, 12) && F.record(e, 16);
}

int read_header(File &F, bool stokes, int32_t nsp, int32_t nn, int32_t nq, double err[2]) {
  int32_t id = 0, sizes[3] = {0, 0, 0};
  double e[2] = {0.0, 0.0};
  if (stokes && (!F.read_record(&id, 4) || id != ID_STOKES)) return ANDR_ERR_ARG;
  if (!F.read_record(sizes, 12) || sizes[0] != nsp || sizes[1] != nn || sizes[2] != nq) return ANDR_ERR_ARG;
  if (!F.read_record(e, 16)) return ANDR_ERR_ARG;
  if (err) {
    err[0] = e[0];
    err[1] = e[1];
  }
  return ANDR_OK;
}

bool sizes_ok(const char *path, int32_t nsp, int32_t nn, int32_t nq) { return path && nsp > 0 && nn > 0 && nq > 0; }

}  // namespace

extern "C" int andr_integrals_write_stokes(const char *path, int32_t nsp, int32_t nnodes, int32_t nqnodes,
                                           const double err[2], double *const T[6], double *const G[6],
                                           double *const prT[3], double *const prG[3]) {
  if (!sizes_ok(path, nsp, nnodes, nqnodes) || !T || !G || !prT || !prG) return ANDR_ERR_ARG;
  for (int k = 0; k < 6; ++k)
    if (!T[k] || !G[k]) return ANDR_ERR_ARG;
  for (int k = 0; k < 3; ++k)
    if (!prT[k] || !prG[k]) return ANDR_ERR_ARG;
  File F(path, "wb");
  if (!F.f) return ANDR_ERR_ARG;
  // the reference's order: xx, yy, zz, xy, xz, yz; the arguments' order: xx, xy, xz, yy, yz, zz
  static const int ord[6] = {0, 3, 5, 1, 2, 4};
  bool ok = write_header(F, true, nsp, nnodes, nqnodes, err);
  for (int k = 0; ok && k < 6; ++k) ok = write_matrix(F, T[ord[k]], nsp, nnodes);
  for (int k = 0; ok && k < 6; ++k) ok = write_matrix(F, G[ord[k]], nsp, nqnodes);
  for (int k = 0; ok && k < 3; ++k) ok = write_matrix(F, prT[k], nsp, nnodes);
  for (int k = 0; ok && k < 3; ++k) ok = write_matrix(F, prG[k], nsp, nqnodes);
  return ok ? ANDR_OK : ANDR_ERR_ARG;
}

extern "C" int andr_integrals_read_stokes(const char *path, int32_t nsp, int32_t nnodes, int32_t nqnodes, double err[2],
                                          double *T[6], double *G[6], double *prT[3], double *prG[3]) {
  if (!sizes_ok(path, nsp, nnodes, nqnodes)) return ANDR_ERR_ARG;
  File F(path, "rb");
  if (!F.f) return ANDR_ERR_ARG;
  static const int ord[6] = {0, 3, 5, 1, 2, 4};
  int rc = read_header(F, true, nsp, nnodes, nqnodes, err);
  bool ok = rc == ANDR_OK;
  for (int k = 0; ok && k < 6; ++k) ok = read_matrix(F, T ? T[ord[k]] : nullptr, nsp, nnodes);
  for (int k = 0; ok && k < 6; ++k) ok = read_matrix(F, G ? G[ord[k]] : nullptr, nsp, nqnodes);
  for (int k = 0; ok && k < 3; ++k) ok = read_matrix(F, prT ? prT[k] : nullptr, nsp, nnodes);
  for (int k = 0; ok && k < 3; ++k) ok = read_matrix(F, prG ? prG[k] : nullptr, nsp, nqnodes);
  return ok ? ANDR_OK : ANDR_ERR_ARG;
}

extern "C" int andr_integrals_write_laplace(const char *path, int32_t nsp, int32_t nnodes, int32_t nqnodes,
                                            const double err[2], const double *H, const double *G) {
  if (!sizes_ok(path, nsp, nnodes, nqnodes) || !H || !G) return ANDR_ERR_ARG;
  File F(path, "wb");
  if (!F.f) return ANDR_ERR_ARG;
  const bool ok = write_header(F, false, nsp, nnodes, nqnodes, err) && write_matrix(F, H, nsp, nnodes) &&
                  write_matrix(F, G, nsp, nqnodes);
  return ok ? ANDR_OK : ANDR_ERR_ARG;
}

extern "C" int andr_integrals_read_laplace(const char *path, int32_t nsp, int32_t nnodes, int32_t nqnodes, double err[2],
                                           double *H, double *G) {
  if (!sizes_ok(path, nsp, nnodes, nqnodes)) return ANDR_ERR_ARG;
  File F(path, "rb");
  if (!F.f) return ANDR_ERR_ARG;
  const bool ok = read_header(F, false, nsp, nnodes, nqnodes, err) == ANDR_OK && read_matrix(F, H, nsp, nnodes) &&
                  read_matrix(F, G, nsp, nqnodes);
  return ok ? ANDR_OK : ANDR_ERR_ARG;
}
