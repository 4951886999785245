// netcdf3.cpp -- netCDF-3 writer for the step's output files.
//
// The files are what the reference's grid_nc_create / grid_nc_write write
// through netCDF-Fortran (nompi/mod_mppparam.F90:121-330): one double
// record variable over (KTAU, [KZ,] IY, JX) in the classic format.  The
// format is the published netCDF classic file layout: a header of
// big-endian 32-bit integers ("CDF", version byte, record count, dimension
// list, global attributes, variable list with each variable's record size
// and data offset) followed by the records, each a big-endian copy of the
// variable's values.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mlc_internal.h"
#include "moloch_output.h"

struct mlc_ncfile {
  FILE *fp = nullptr;
  long long begin = 0;    // offset of the first record
  long long recsize = 0;  // bytes per record
  long long nvalues = 0;  // doubles per record
  int numrecs = 0;
  std::vector<unsigned char> buf;  // big-endian staging
};

namespace {

void put32(std::vector<unsigned char> &h, uint32_t v) {
  h.push_back((unsigned char)(v >> 24));
  h.push_back((unsigned char)(v >> 16));
  h.push_back((unsigned char)(v >> 8));
  h.push_back((unsigned char)v);
}
void put64(std::vector<unsigned char> &h, uint64_t v) {
  put32(h, (uint32_t)(v >> 32));
  put32(h, (uint32_t)v);
}
void put_name(std::vector<unsigned char> &h, const std::string &s) {
  put32(h, (uint32_t)s.size());
  for (char c : s) h.push_back((unsigned char)c);
  while (h.size() % 4) h.push_back(0);
}

constexpr uint32_t kDimension = 0x0A, kVariable = 0x0B, kDouble = 6;

}  // namespace

extern "C" int mlc_nc_create(const char *path, const char *varname, int nx, int ny, int nz, mlc_ncfile **out) {
  if (!path || !varname || !out || nx <= 0 || ny <= 0 || nz < 0 || !*varname)
    return mlc_fail(MLC_EINVAL, "mlc_nc_create: bad arguments");
  *out = nullptr;
  const long long nvalues = (long long)nx * ny * (nz > 0 ? nz : 1);
  const long long recsize = nvalues * 8;
  // classic format unless a record does not fit its 32-bit sizes
  const bool offset64 = recsize > 2147483644LL;
  std::vector<unsigned char> h = {'C', 'D', 'F', (unsigned char)(offset64 ? 2 : 1)};
  put32(h, 0);  // numrecs
  // dimensions, in the order of the reference's nf90_def_dim calls
  put32(h, kDimension);
  put32(h, nz > 0 ? 4 : 3);
  put_name(h, "JX");
  put32(h, (uint32_t)nx);
  put_name(h, "IY");
  put32(h, (uint32_t)ny);
  if (nz > 0) {
    put_name(h, "KZ");
    put32(h, (uint32_t)nz);
  }
  put_name(h, "KTAU");
  put32(h, 0);  // the unlimited dimension
  // no global attributes
  put32(h, 0);
  put32(h, 0);
  // one variable; dimension ids slowest first (Fortran order reversed)
  put32(h, kVariable);
  put32(h, 1);
  put_name(h, varname);
  if (nz > 0) {
    put32(h, 4);
    put32(h, 3);
    put32(h, 2);
    put32(h, 1);
    put32(h, 0);
  } else {
    put32(h, 3);
    put32(h, 2);
    put32(h, 1);
    put32(h, 0);
  }
  put32(h, 0);  // no variable attributes
  put32(h, 0);
  put32(h, kDouble);
  put32(h, (uint32_t)(recsize > 4294967295LL ? 4294967295LL : recsize));  // vsize
  const size_t begin_at = h.size();
  if (offset64)
    put64(h, 0);
  else
    put32(h, 0);
  const long long begin = (long long)h.size();
  // patch the data offset
  std::vector<unsigned char> b;
  if (offset64)
    put64(b, (uint64_t)begin);
  else
    put32(b, (uint32_t)begin);
  std::memcpy(h.data() + begin_at, b.data(), b.size());

  FILE *fp = std::fopen(path, "wb");
  if (!fp) return mlc_fail(MLC_EINVAL, "mlc_nc_create: cannot create %s: %s", path, std::strerror(errno));
  if (std::fwrite(h.data(), 1, h.size(), fp) != h.size() || std::fflush(fp) != 0) {
    std::fclose(fp);
    return mlc_fail(MLC_EINVAL, "mlc_nc_create: cannot write %s", path);
  }
  auto *f = new mlc_ncfile();
  f->fp = fp;
  f->begin = begin;
  f->recsize = recsize;
  f->nvalues = nvalues;
  *out = f;
  return MLC_OK;
}

extern "C" int mlc_nc_append(mlc_ncfile *f, const double *values) {
  if (!f || !f->fp || !values) return mlc_fail(MLC_EINVAL, "mlc_nc_append: bad arguments");
  const long long off = f->begin + (long long)f->numrecs * f->recsize;
  if (fseeko(f->fp, (off_t)off, SEEK_SET) != 0) return mlc_fail(MLC_EINVAL, "mlc_nc_append: seek failed");
  const long long chunk = 1 << 20;  // doubles per write
  f->buf.resize((size_t)(chunk * 8));
  for (long long s = 0; s < f->nvalues; s += chunk) {
    const long long n = f->nvalues - s < chunk ? f->nvalues - s : chunk;
    unsigned char *o = f->buf.data();
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; i++) {
      uint64_t u;
      std::memcpy(&u, values + s + i, 8);
      for (int bte = 0; bte < 8; bte++) o[i * 8 + bte] = (unsigned char)(u >> (56 - 8 * bte));
    }
    if (std::fwrite(o, 1, (size_t)(n * 8), f->fp) != (size_t)(n * 8))
      return mlc_fail(MLC_EINVAL, "mlc_nc_append: write failed: %s", std::strerror(errno));
  }
  f->numrecs++;
  // record count, then flush (nf90_sync after every record)
  unsigned char nr[4] = {(unsigned char)(f->numrecs >> 24), (unsigned char)(f->numrecs >> 16),
                         (unsigned char)(f->numrecs >> 8), (unsigned char)f->numrecs};
  if (fseeko(f->fp, 4, SEEK_SET) != 0 || std::fwrite(nr, 1, 4, f->fp) != 4 || std::fflush(f->fp) != 0)
    return mlc_fail(MLC_EINVAL, "mlc_nc_append: cannot update the record count");
  return MLC_OK;
}

extern "C" int mlc_nc_records(const mlc_ncfile *f) { return f ? f->numrecs : -1; }

extern "C" int mlc_nc_close(mlc_ncfile *f) {
  if (!f) return MLC_OK;
  int rc = MLC_OK;
  if (f->fp && std::fclose(f->fp) != 0) rc = mlc_fail(MLC_EINVAL, "mlc_nc_close: close failed");
  delete f;
  return rc;
}
