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

// ---------------------------------------------------------------------------
// Timeseries and endpoint files (ParticleManager.F90:51-250), written from the
// engine's particle state.  Cell numbers and layers are those of the particle's
// own ensemble member (the same numbers for one member).

#include <vector>

#include "rwpt_internal.h"

namespace {

// Fortran Ew.dEe with w = 18, d = 10, e = 2 (the `e18.10` of the headers):
// 0.dddddddddd mantissa, E, sign, two exponent digits, right-aligned.
void put_e18_10(std::string &out, double v) {
  char f[64];
  if (v == 0.0) {
    std::snprintf(f, sizeof f, "%18s", "0.0000000000E+00");
  } else {
    char m[48];
    std::snprintf(m, sizeof m, "%.9E", v);  // d.dddddddddE+xx: 10 digits
    // 0.dddddddddd x 10^(x+1): the digits of d.ddddddddd with the point moved
    const char *p = m;
    const bool neg = *p == '-';
    if (neg) ++p;
    char digits[16];
    int nd = 0;
    for (; *p && *p != 'E'; ++p)
      if (*p != '.') digits[nd++] = *p;
    digits[nd] = '\0';
    const int ex = std::atoi(p + 1) + 1;
    std::snprintf(f, sizeof f, "%s0.%sE%c%02d", neg ? "-" : "", digits, ex < 0 ? '-' : '+',
                  ex < 0 ? -ex : ex);
    char g[64];
    std::snprintf(g, sizeof g, "%18s", f);
    out += g;
    return;
  }
  out += f;
}

void put_i(std::string &out, int64_t v, int w) {
  char f[32];
  std::snprintf(f, sizeof f, "%*lld", w, (long long)v);
  out += f;
}

struct HostGrid {
  int ncol = 0, nrow = 0;
  int64_t ncm = 0;
  std::vector<double> xf, yf, delr, delc, top, bot;
  std::vector<int32_t> zone;
  // unstructured grids: per-cell centres, sizes and layers
  std::vector<double> ucx, ucy, udx, udy;
  std::vector<int32_t> ulayer;
  bool load(rwpt_engine *e) {
    rwpt::DeviceProblem &P = e->prob;
    if (e->usg) {
      const rwpt::UsgData &U = *e->usg;
      const size_t n = (size_t)U.ncell;
      ncm = n;
      ucx.resize(n); ucy.resize(n); udx.resize(n); udy.resize(n); ulayer.resize(n);
      top.resize(n);
      bot.resize(n);
      hipError_t err = hipMemcpy(ucx.data(), U.cx, n * 8, hipMemcpyDeviceToHost);
      if (err == hipSuccess) err = hipMemcpy(ucy.data(), U.cy, n * 8, hipMemcpyDeviceToHost);
      if (err == hipSuccess) err = hipMemcpy(udx.data(), U.dx, n * 8, hipMemcpyDeviceToHost);
      if (err == hipSuccess) err = hipMemcpy(udy.data(), U.dy, n * 8, hipMemcpyDeviceToHost);
      if (err == hipSuccess) err = hipMemcpy(ulayer.data(), U.layer, n * 4, hipMemcpyDeviceToHost);
      if (err == hipSuccess) err = hipMemcpy(top.data(), model_top(P), n * 8, hipMemcpyDeviceToHost);
      if (err == hipSuccess) err = hipMemcpy(bot.data(), P.bot, n * 8, hipMemcpyDeviceToHost);
      if (err == hipSuccess && P.zone) {
        zone.resize(n);
        err = hipMemcpy(zone.data(), P.zone, n * 4, hipMemcpyDeviceToHost);
      }
      return err == hipSuccess;
    }
    ncol = P.ncol;
    nrow = P.nrow;
    ncm = P.ncell_member;
    xf.resize(ncol + 1);
    yf.resize(nrow + 1);
    delr.resize(ncol);
    delc.resize(nrow);
    top.resize(P.ncell);
    bot.resize(P.ncell);
    hipError_t err = hipMemcpy(xf.data(), e->xface, xf.size() * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(yf.data(), e->yface, yf.size() * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(delr.data(), P.delr, ncol * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(delc.data(), P.delc, nrow * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(top.data(), model_top(P), P.ncell * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(bot.data(), P.bot, P.ncell * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess && P.zone) {
      zone.resize(P.ncell);
      err = hipMemcpy(zone.data(), P.zone, P.ncell * 4, hipMemcpyDeviceToHost);
    }
    return err == hipSuccess;
  }
  // ConvertToModelXYZ (ModflowRectangularGrid.f90:524-549); c is 1-based
  void xyz(int32_t c, double x, double y, double z, double out[3]) const {
    const int64_t c0 = c - 1;
    double left, right, front, back;
    if (!ucx.empty()) {
      left = ucx[c0] - (udx[c0] / 2.0);
      right = ucx[c0] + (udx[c0] / 2.0);
      front = ucy[c0] - (udy[c0] / 2.0);
      back = ucy[c0] + (udy[c0] / 2.0);
    } else {
      const int64_t lc = (int64_t)nrow * ncol, r = c0 % lc;
      const int i = (int)(r / ncol), j = (int)(r % ncol);
      const double cx = (xf[j] + xf[j + 1]) / 2.0, cy = (yf[i] + yf[i + 1]) / 2.0;
      left = cx - (delr[j] / 2.0);
      right = cx + (delr[j] / 2.0);
      front = cy - (delc[i] / 2.0);
      back = cy + (delc[i] / 2.0);
    }
    const double b = bot[c0], t = top[c0];
    out[0] = std::fmin(std::fmax(((1.0 - x) * left) + (x * right), left), right);
    out[1] = std::fmin(std::fmax(((1.0 - y) * front) + (y * back), front), back);
    out[2] = std::fmin(std::fmax(((1.0 - z) * b) + (z * t), b), t);
  }
  int32_t local_cell(int32_t c) const { return (int32_t)((c - 1) % ncm + 1); }
  int32_t layer(int32_t c) const {
    if (!ulayer.empty()) return ulayer[c - 1];
    return (int32_t)(((c - 1) % ncm) / ((int64_t)nrow * ncol) + 1);
  }
  int32_t zone_of(int32_t c) const { return zone.empty() ? 1 : zone[c - 1]; }
};

struct HostParticles {
  std::vector<int32_t> cell, status;
  std::vector<double> lx, ly, lz, t;
  bool load(rwpt_engine *e) {
    rwpt::Particles &S = e->part;
    const size_t n = (size_t)S.n;
    cell.resize(n);
    status.resize(n);
    lx.resize(n);
    ly.resize(n);
    lz.resize(n);
    t.resize(n);
    if (n == 0) return true;
    hipError_t err = hipStreamSynchronize(e->stream);
    if (err == hipSuccess) err = hipMemcpy(cell.data(), S.cell, n * 4, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(status.data(), S.status, n * 4, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(lx.data(), S.lx, n * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(ly.data(), S.ly, n * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(lz.data(), S.lz, n * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(t.data(), S.t, n * 8, hipMemcpyDeviceToHost);
    return err == hipSuccess;
  }
};

int write_all(const char *path, bool append, const std::string &buf) {
  FILE *f = std::fopen(path, append ? "ab" : "wb");
  if (!f) return RWPT_E_IO;
  int rc = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size() ? RWPT_OK : RWPT_E_IO;
  if (std::fclose(f) != 0) rc = RWPT_E_IO;
  return rc;
}

// FindFace (ParticleManager.F90:665-690)
int find_face(double x, double y, double z) {
  if (x == 0.0) return 1;
  if (x == 1.0) return 2;
  if (y == 0.0) return 3;
  if (y == 1.0) return 4;
  if (z == 0.0) return 5;
  if (z == 1.0) return 6;
  return 0;
}

}  // namespace

extern "C" int rwpt_write_timeseries_header(const char *path, int32_t tracking_direction,
                                            double reference_time, double origin_x,
                                            double origin_y, double rotation) {
  if (!path) return RWPT_E_ARG;
  // WriteTimeseriesHeader (ParticleManager.F90:208-226)
  std::string b = "MODPATH_TIMESERIES_FILE";
  put_i(b, 7, 10);
  put_i(b, 2, 10);
  b += '\n';
  put_i(b, tracking_direction, 5);
  b += ' ';
  put_e18_10(b, reference_time);
  put_e18_10(b, origin_x);
  put_e18_10(b, origin_y);
  put_e18_10(b, rotation);
  b += "\nEND HEADER\n";
  return write_all(path, false, b);
}

extern "C" int rwpt_write_timeseries(rwpt_engine *e, const char *path, int32_t time_point,
                                     int32_t time_step, double time,
                                     const int32_t *sequence, const int32_t *id,
                                     const int32_t *group, const uint8_t *select) {
  if (!e || !path || !sequence || !id || !group) return RWPT_E_ARG;
  HostGrid G;
  HostParticles H;
  if (!G.load(e) || !H.load(e)) return RWPT_E_DEVICE;
  std::string b;
  for (size_t p = 0; p < H.cell.size(); ++p) {
    if (select ? !select[p] : H.status[p] != 1) continue;
    double g[3];
    G.xyz(H.cell[p], H.lx[p], H.ly[p], H.lz[p], g);
    // WriteTimeseriesRecord (ParticleManager.F90:229-250):
    // (2I8,es18.9e3,i10,i5,2i10,6es18.9e3,i10)
    put_i(b, time_point, 8);
    put_i(b, time_step, 8);
    put_es18_9e3(b, time);
    put_i(b, sequence[p], 10);
    put_i(b, group[p], 5);
    put_i(b, id[p], 10);
    put_i(b, G.local_cell(H.cell[p]), 10);
    put_es18_9e3(b, H.lx[p]);
    put_es18_9e3(b, H.ly[p]);
    put_es18_9e3(b, H.lz[p]);
    put_es18_9e3(b, g[0]);
    put_es18_9e3(b, g[1]);
    put_es18_9e3(b, g[2]);
    put_i(b, G.layer(H.cell[p]), 10);
    b += '\n';
  }
  return write_all(path, true, b);
}

extern "C" int rwpt_write_endpoints(
    rwpt_engine *e, const char *path, int32_t format, int32_t tracking_direction,
    double reference_time, double origin_x, double origin_y, double rotation,
    int32_t n_groups, const char *const *group_names, const int32_t *group_solute,
    const int32_t *group, const int32_t *sequence, const int32_t *id,
    const int32_t *initial_cell, const double *initial_lx, const double *initial_ly,
    const double *initial_lz, const double *initial_time, const int32_t *initial_face,
    const double *mass) {
  if (!e || !path || (format != 0 && format != 1) || n_groups < 0 || !group ||
      !sequence || !id || !initial_cell || !initial_lx || !initial_ly || !initial_lz ||
      !initial_time || (n_groups > 0 && !group_names))
    return RWPT_E_ARG;
  if (format == 0 && !initial_face) return RWPT_E_ARG;
  if (format == 1 && (!group_solute || !mass)) return RWPT_E_ARG;
  HostGrid G;
  HostParticles H;
  if (!G.load(e) || !H.load(e)) return RWPT_E_DEVICE;
  const size_t n = H.cell.size();
  // status sums, release count, largest released ID (:67-86)
  int64_t sums[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  int64_t maxId = 0;
  for (size_t p = 0; p < n; ++p) {
    const int s = H.status[p];
    if (s >= 0 && s <= 9) {
      sums[s] += 1;
      if (s != 0 && s != 8 && id[p] > maxId) maxId = id[p];
    }
  }
  const int64_t total = (int64_t)n, released = total - sums[0] - sums[8];
  std::string b = format == 0 ? "MODPATH_ENDPOINT_FILE" : "MODPATHRW_ENDPOINT_FILE";
  put_i(b, format == 0 ? 7 : 1, 10);
  put_i(b, format == 0 ? 2 : 0, 10);
  b += '\n';
  put_i(b, tracking_direction, 5);
  put_i(b, total, 10);
  put_i(b, released, 10);
  put_i(b, maxId, 10);
  b += ' ';
  put_e18_10(b, reference_time);
  put_e18_10(b, origin_x);
  put_e18_10(b, origin_y);
  put_e18_10(b, rotation);
  b += '\n';
  for (int s = 0; s < 10; ++s) put_i(b, sums[s], 8);
  b += '\n';
  put_i(b, n_groups, 10);
  b += '\n';
  for (int g = 0; g < n_groups; ++g) {
    // ParticleGroupType%Name is character(len=20): padded or cut to 20
    std::string nm = group_names[g] ? group_names[g] : "";
    nm.resize(20, ' ');
    b += nm;
    b += '\n';
  }
  b += "END HEADER\n";
  for (size_t p = 0; p < n; ++p) {
    if (H.status[p] == 0 || H.status[p] == 8) continue;
    double gi[3], gf[3];
    G.xyz(initial_cell[p], initial_lx[p], initial_ly[p], initial_lz[p], gi);
    G.xyz(H.cell[p], H.lx[p], H.ly[p], H.lz[p], gf);
    const int32_t c0 = G.local_cell(initial_cell[p]), c1 = G.local_cell(H.cell[p]);
    if (format == 0) {
      // WriteEndpoints (:100-121):
      // (3i10,i5,2es18.9e3,i10,i5,6es18.9e3,2i5,i10,i5,6es18.9e3,2i5)
      put_i(b, sequence[p], 10);
      put_i(b, group[p], 10);
      put_i(b, id[p], 10);
      put_i(b, H.status[p], 5);
      put_es18_9e3(b, initial_time[p]);
      put_es18_9e3(b, H.t[p]);
      put_i(b, c0, 10);
      put_i(b, G.layer(initial_cell[p]), 5);
      put_es18_9e3(b, initial_lx[p]);
      put_es18_9e3(b, initial_ly[p]);
      put_es18_9e3(b, initial_lz[p]);
      for (int d = 0; d < 3; ++d) put_es18_9e3(b, gi[d]);
      put_i(b, G.zone_of(initial_cell[p]), 5);
      put_i(b, initial_face[p], 5);
      put_i(b, c1, 10);
      put_i(b, G.layer(H.cell[p]), 5);
      put_es18_9e3(b, H.lx[p]);
      put_es18_9e3(b, H.ly[p]);
      put_es18_9e3(b, H.lz[p]);
      for (int d = 0; d < 3; ++d) put_es18_9e3(b, gf[d]);
      put_i(b, G.zone_of(H.cell[p]), 5);
      put_i(b, find_face(H.lx[p], H.ly[p], H.lz[p]), 5);
    } else {
      // WriteEndpointsMassParticles (:180-200):
      // (4i10,i5,3es18.9e3,i10,i5,3es18.9e3,1i5,i10,i5,3es18.9e3,1i5)
      put_i(b, sequence[p], 10);
      put_i(b, group[p], 10);
      const int32_t gidx = group[p] - 1;
      put_i(b, gidx >= 0 && gidx < n_groups ? group_solute[gidx] : 0, 10);
      put_i(b, id[p], 10);
      put_i(b, H.status[p], 5);
      put_es18_9e3(b, mass[p]);
      put_es18_9e3(b, initial_time[p]);
      put_es18_9e3(b, H.t[p]);
      put_i(b, c0, 10);
      put_i(b, G.layer(initial_cell[p]), 5);
      for (int d = 0; d < 3; ++d) put_es18_9e3(b, gi[d]);
      put_i(b, G.zone_of(initial_cell[p]), 5);
      put_i(b, c1, 10);
      put_i(b, G.layer(H.cell[p]), 5);
      for (int d = 0; d < 3; ++d) put_es18_9e3(b, gf[d]);
      put_i(b, G.zone_of(H.cell[p]), 5);
    }
    b += '\n';
  }
  return write_all(path, false, b);
}

// Observation records (Observation.f90:114-221) of the particles in one
// observation's cells, as MPathRW.F90 writes them: resident records at a
// time point for the active particles (:1449-1530), sink records for the
// particles stopped at a weak sink (weak source in backward tracking,
// :1344-1392).  Text: (2I8,es18.9e3,i10,es18.9e3,2i5,2i10,5es18.9e3) and
// (...,es18.9e3); binary: the same fields as a stream of int32 and float64.
extern "C" int rwpt_write_observation(rwpt_engine *e, const char *path, int32_t append,
                                      int32_t style, int32_t format, int32_t n_cells,
                                      const int32_t *cells, int32_t time_point,
                                      int32_t time_step, const int32_t *id,
                                      const int32_t *group, const int32_t *group_solute,
                                      const double *mass, double flow_rate,
                                      const uint8_t *select) {
  if (!e || !path || (style != 1 && style != 2) || (format != 0 && format != 1) ||
      n_cells < 0 || (n_cells > 0 && !cells) || !id || !group || !group_solute)
    return RWPT_E_ARG;
  rwpt::DeviceProblem &P = e->prob;
  for (int32_t k = 0; k < n_cells; ++k)
    if (cells[k] < 1 || cells[k] > P.ncell) return RWPT_E_ARG;
  HostGrid G;
  HostParticles H;
  if (!G.load(e) || !H.load(e)) return RWPT_E_DEVICE;
  // the observation's cells and, for resident records, their retardation
  // and water volume (GetWaterVolume, ModpathCellData.f90:3133-3157:
  // porosity DX DY times the saturated thickness of GetDZRW, 0 when dry)
  std::vector<int32_t> obs_index(P.ncell, -1);
  std::vector<double> rfac(n_cells), wvol(n_cells);
  for (int32_t k = 0; k < n_cells; ++k) {
    // the record of the cell, or on an unstructured grid of its first unit,
    // whose porosity, retardation, thickness and dry flag are the cell's;
    // DX and DY are then the whole cell's (CellData of TrackCell)
    int64_t rec_index = cells[k] - 1;
    if (e->usg) {
      int32_t u = 0;
      if (hipMemcpy(&u, e->usg->cell_unit + (cells[k] - 1), sizeof u,
                    hipMemcpyDeviceToHost) != hipSuccess)
        return RWPT_E_DEVICE;
      rec_index = u;
    }
    double rec[rwpt::CELL_DOUBLES];
    if (hipMemcpy(rec, P.cells + rec_index * rwpt::CELL_DOUBLES, sizeof rec,
                  hipMemcpyDeviceToHost) != hipSuccess)
      return RWPT_E_DEVICE;
    const int32_t *ints = reinterpret_cast<const int32_t *>(rec);
    const bool dry = (ints[2 * rwpt::F_FLAGS + 1] & rwpt::CF_DRY) != 0;
    const double dx = e->usg ? G.udx[cells[k] - 1] : rec[rwpt::F_DX];
    const double dy = e->usg ? G.udy[cells[k] - 1] : rec[rwpt::F_DY];
    rfac[k] = rec[rwpt::F_R];
    wvol[k] = dry ? 0.0 : rec[rwpt::F_PHI] * dx * dy * rec[rwpt::F_DZ];
    obs_index[cells[k] - 1] = k;
  }
  const int stop_status = e->opt.backward_tracking ? 4 : 3;
  std::string b;
  auto i4 = [&](int32_t v) { b.append(reinterpret_cast<const char *>(&v), 4); };
  auto f8 = [&](double v) { b.append(reinterpret_cast<const char *>(&v), 8); };
  for (size_t p = 0; p < H.cell.size(); ++p) {
    const int32_t c = H.cell[p];
    if (c < 1 || c > P.ncell || obs_index[c - 1] < 0) continue;
    if (style == 1 && select) {
      // the records MPathRW writes at a time point: the particles chosen
      if (!select[p]) continue;
    } else {
      if (H.status[p] != (style == 1 ? 1 : stop_status)) continue;
      if (select && !select[p]) continue;
    }
    const int k = obs_index[c - 1];
    const double m = mass ? mass[p] : 1.0;
    const int32_t sol = group_solute[group[p] - 1];
    double g[3] = {0.0, 0.0, 0.0};
    if (style == 1) G.xyz(c, H.lx[p], H.ly[p], H.lz[p], g);
    if (format == 0) {
      put_i(b, time_point, 8);
      put_i(b, time_step, 8);
      put_es18_9e3(b, H.t[p]);
      put_i(b, id[p], 10);
      put_es18_9e3(b, m);
      put_i(b, group[p], 5);
      put_i(b, sol, 5);
      put_i(b, G.local_cell(c), 10);
      put_i(b, G.layer(c), 10);
      if (style == 1) {
        put_es18_9e3(b, rfac[k]);
        put_es18_9e3(b, wvol[k]);
        put_es18_9e3(b, g[0]);
        put_es18_9e3(b, g[1]);
        put_es18_9e3(b, g[2]);
      } else {
        put_es18_9e3(b, flow_rate);
      }
      b += '\n';
    } else {
      i4(time_point);
      i4(time_step);
      f8(H.t[p]);
      i4(id[p]);
      f8(m);
      i4(group[p]);
      i4(sol);
      i4(G.local_cell(c));
      i4(G.layer(c));
      if (style == 1) {
        f8(rfac[k]);
        f8(wvol[k]);
        f8(g[0]);
        f8(g[1]);
        f8(g[2]);
      } else {
        f8(flow_rate);
      }
    }
  }
  return write_all(path, append != 0, b);
}
