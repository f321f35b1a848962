// rt_io.hip -- host-side reader of the k-coefficient data files that LAGINTERP
// opens at start-up (code/laginterp.f90:59-69): CO2H2O_V_12_95_INTEL and
// CO2H2O_IR_12_95_INTEL.  Each is a Fortran sequential unformatted file with
// two records, each framed by its byte length as a 4-byte little-endian
// integer before and after the payload:
//   record 1: CO2(7,11,10,nb,17) FP64, the k-coefficients on the 11 reference
//             pressures (first index fastest);
//   record 2: FZERO(nb) FP64.
// As Fortran's READ does, a record longer than the list it is read into is
// accepted and its tail skipped; a shorter one is an error.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/mgcm_rt.h"

namespace {

constexpr int NT = 7, NPREF = 11, NH = 10, NG = 17;

// Reads one record of at least `need` bytes into dst; returns false on a short,
// unframed or truncated record.
bool read_record(FILE *f, void *dst, size_t need) {
  uint32_t head = 0, tail = 0;
  if (fread(&head, 4, 1, f) != 1) return false;
  if (head < need) return false;
  if (need > 0 && fread(dst, 1, need, f) != need) return false;
  if (head > need && fseek(f, (long)(head - need), SEEK_CUR) != 0) return false;
  if (fread(&tail, 4, 1, f) != 1) return false;
  return tail == head;
}

}  // namespace

extern "C" int mgcm_rt_read_ktable(const char *path, int nbands, double *raw, double *fzero) {
  if (!path || !raw || !fzero || nbands < 1 || nbands > MGCM_RT_NV) return MGCM_RT_ERR_ARG;
  FILE *f = fopen(path, "rb");
  if (!f) return MGCM_RT_ERR_IO;
  const size_t n = (size_t)NT * NPREF * NH * nbands * NG;
  std::vector<double> k(n), z(nbands);
  const bool ok = read_record(f, k.data(), n * sizeof(double)) &&
                  read_record(f, z.data(), (size_t)nbands * sizeof(double));
  fclose(f);
  if (!ok) return MGCM_RT_ERR_IO;
  memcpy(raw, k.data(), n * sizeof(double));
  memcpy(fzero, z.data(), (size_t)nbands * sizeof(double));
  return MGCM_RT_OK;
}
