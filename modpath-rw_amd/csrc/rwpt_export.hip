// gpkde density rows in the files MODPATH-RW writes (host code).
//
// GridProjectedKDE%ExportDensityUnit and %ExportDensityUnitBinary
// (GridProjectedKDE.F90:6504-7021), in the branch MPathRW.F90:1626-1642
// reaches (time index, time and particle group present): one row per bin
// with density > 0, x fastest; the text form uses the Fortran edit
// descriptors I8 and ES18.9E3, the binary form is a stream of int32 and
// float64 fields without record markers (the unit is opened with
// access='stream', MPathRW.F90:583-586).
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "rwpt.h"

namespace {

// Fortran ESw.dEe with w = 18, d = 9, e = 3: a sign (or blank), one
// non-zero digit (0 for zero), the point, nine digits, E, the exponent sign
// and three digits, right-aligned in 18 columns.
void put_es18_9e3(std::string &out, double v) {
  char m[48];
  std::snprintf(m, sizeof m, "%.9E", v);  // d.dddddddddE+xx, correctly rounded
  char *e = std::strchr(m, 'E');
  char field[40];
  if (!e || !std::isfinite(v)) {
    std::snprintf(field, sizeof field, "%18s", std::isnan(v) ? "NaN" : (v > 0 ? "Infinity" : "-Infinity"));
    out += field;
    return;
  }
  const int ex = std::atoi(e + 1);
  *e = '\0';
  std::snprintf(field, sizeof field, "%sE%c%03d", m, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
  char padded[48];
  std::snprintf(padded, sizeof padded, "%18s", field);
  out += padded;
}

void put_i8(std::string &out, int32_t v) {
  char f[24];
  std::snprintf(f, sizeof f, "%8d", (int)v);
  out += f;
}

template <typename T>
void put_bin(std::string &out, T v) {
  out.append(reinterpret_cast<const char *>(&v), sizeof v);
}

}  // namespace

extern "C" int rwpt_kde_write(const char *path, int32_t append, int32_t format,
                              int32_t column_format, int32_t time_id, double time,
                              int32_t group_id, const rwpt_kde_options *opt,
                              const double *density, const double *histogram,
                              double histogram_scale) {
  if (!path || !opt || !density || !histogram) return RWPT_E_ARG;
  if (format != 0 && format != 1) return RWPT_E_ARG;
  if (column_format < 0 || column_format > 2) return RWPT_E_ARG;
  for (int d = 0; d < 3; ++d)
    if (opt->nbins[d] < 1) return RWPT_E_ARG;
  const int nx = opt->nbins[0], ny = opt->nbins[1], nz = opt->nbins[2];
  FILE *f = std::fopen(path, append ? "ab" : "wb");
  if (!f) return RWPT_E_IO;
  std::string buf;
  int rc = RWPT_OK;
  for (int iz = 1; iz <= nz && rc == RWPT_OK; ++iz)
    for (int iy = 1; iy <= ny; ++iy) {
      for (int ix = 1; ix <= nx; ++ix) {
        const int64_t q = (int64_t)(ix - 1) + (int64_t)nx * ((iy - 1) + (int64_t)ny * (iz - 1));
        const double dv = density[q];
        if (dv <= 0.0) continue;  // `.le. 0 cycle`: a NaN density is written
        const double hv = histogram[q] * histogram_scale;
        // bin centre coordinates, (real(idbin) - 0.5)*binSize + origin
        double xc[3];
        const int id[3] = {ix, iy, iz};
        for (int d = 0; d < 3; ++d)
          xc[d] = ((double)id[d] - 0.5) * opt->bin_size[d] + opt->origin[d];
        if (format == 0) {
          put_i8(buf, time_id);
          put_es18_9e3(buf, time);
          put_i8(buf, group_id);
          if (column_format != 2)
            for (int d = 0; d < 3; ++d) put_i8(buf, id[d]);
          if (column_format != 0)
            for (int d = 0; d < 3; ++d) put_es18_9e3(buf, xc[d]);
          put_es18_9e3(buf, dv);
          put_es18_9e3(buf, hv);
          buf += '\n';
        } else {
          put_bin<int32_t>(buf, time_id);
          put_bin<double>(buf, time);
          put_bin<int32_t>(buf, group_id);
          if (column_format != 2)
            for (int d = 0; d < 3; ++d) put_bin<int32_t>(buf, id[d]);
          if (column_format != 0)
            for (int d = 0; d < 3; ++d) put_bin<double>(buf, xc[d]);
          put_bin<double>(buf, dv);
          put_bin<double>(buf, hv);
        }
      }
      if (buf.size() > (1u << 20)) {
        if (std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) rc = RWPT_E_IO;
        buf.clear();
      }
    }
  if (rc == RWPT_OK && !buf.empty() &&
      std::fwrite(buf.data(), 1, buf.size(), f) != buf.size())
    rc = RWPT_E_IO;
  if (std::fclose(f) != 0 && rc == RWPT_OK) rc = RWPT_E_IO;
  return rc;
}
