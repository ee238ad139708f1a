// On-disk formats of the path: the VTK image-data mesh output of
// pmlib_output_mesh (source/output/pmlib_output_mesh.f:5-188) and the
// particle checkpoint of the vortex clients, written by output_particles
// (vortex_codes/vortex_structures/source/output/output_particles.f:50-69) and
// read by input_particles (.../input/input_particles.f:81-135, 200-230).
//
// Both are written from device-resident (or host) arrays: the fields are
// copied to the host once and serialised there.  The byte layout follows
// the reference's Fortran I/O:
//  * the .pvti index is formatted sequential output (one record per line),
//    the .vti piece is stream output (strings back to back, no newlines),
//    with Fortran's Iw / Iw.m / Ew.d edit descriptors restated below;
//  * the particle file is unformatted sequential: each record framed by its
//    byte count as a 4-byte integer before and after.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pmgpu.h"
#include "pm_common.h"

namespace pmgpu {
namespace {

// Fortran Iw: right-justified in w columns (asterisks if it does not fit)
std::string fmt_i(long long v, int w) {
  std::string s = std::to_string(v);
  if ((int)s.size() > w) return std::string(w, '*');
  return std::string(w - s.size(), ' ') + s;
}
// Fortran Iw.m: at least m digits, zero-padded
std::string fmt_i0(long long v, int w, int m) {
  std::string d = std::to_string(v < 0 ? -v : v);
  if ((int)d.size() < m) d = std::string(m - d.size(), '0') + d;
  if (v < 0) d = "-" + d;
  if ((int)d.size() > w) return std::string(w, '*');
  return std::string(w - d.size(), ' ') + d;
}
// Fortran Ew.d with the default scale factor: [-]0.d1...dd E+ee, the
// mantissa in [0.1, 1) rounded to d digits, right-justified in w columns
std::string fmt_e(double v, int w, int d) {
  char buf[64];
  std::string body;
  if (v == 0.0) {
    body = "0." + std::string(d, '0') + "E+00";
    if (std::signbit(v)) body = "-" + body;
  } else {
    // %.(d-1)e gives x.xxxxE+yy with the first significant digit before the
    // point and correct rounding; shift it behind the point
    std::snprintf(buf, sizeof(buf), "%.*e", d - 1, std::fabs(v));
    std::string s(buf);
    const size_t epos = s.find('e');
    std::string mant = s.substr(0, epos);
    int ex = std::atoi(s.c_str() + epos + 1) + 1;
    std::string digits = mant.substr(0, 1) + (mant.size() > 2 ? mant.substr(2) : std::string());
    char eb[16];
    if (ex >= -99 && ex <= 99)
      std::snprintf(eb, sizeof(eb), "E%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    else
      std::snprintf(eb, sizeof(eb), "%c%03d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    body = "0." + digits + eb;
    if (v < 0) body = "-" + body;
  }
  if ((int)body.size() > w) {
    // a leading zero is optional when the field is one column short
    if ((int)body.size() == w + 1 && body.compare(body[0] == '-' ? 1 : 0, 2, "0.") == 0)
      body.erase(body[0] == '-' ? 1 : 0, 1);
    else
      return std::string(w, '*');
  }
  return std::string(w - body.size(), ' ') + body;
}

std::string rtrim(const std::string &s) {
  size_t e = s.find_last_not_of(' ');
  return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}

// the interior cells of a (3, E...) mesh as FP32, component fastest, i
// fastest then j then k
int interior_f32(const pmgpu_geom *g, const void *mesh, int32_t precision, int32_t on_device,
                 std::vector<float> &out) {
  int32_t E[3];
  if (pmgpu_mesh_extent(g, E)) return -1;
  const size_t n = (size_t)3 * E[0] * E[1] * E[2];
  const size_t es = precision == PMGPU_FP32 ? 4 : 8;
  std::vector<unsigned char> host;
  const unsigned char *src = static_cast<const unsigned char *>(mesh);
  if (on_device) {
    host.resize(n * es);
    PM_HIP(hipMemcpy(host.data(), mesh, n * es, hipMemcpyDeviceToHost));
    src = host.data();
  }
  int ng[3];
  for (int d = 0; d < 3; ++d) ng[d] = (E[d] - g->ncell[d]) / 2;
  out.resize((size_t)3 * g->ncell[0] * g->ncell[1] * g->ncell[2]);
  size_t q = 0;
  for (int k = 0; k < g->ncell[2]; ++k)
    for (int j = 0; j < g->ncell[1]; ++j)
      for (int i = 0; i < g->ncell[0]; ++i) {
        const size_t o = 3 * ((size_t)(i + ng[0]) + (size_t)E[0] * ((j + ng[1]) + (size_t)E[1] * (k + ng[2])));
        for (int c = 0; c < 3; ++c) {
          if (precision == PMGPU_FP32) {
            float v;
            std::memcpy(&v, src + (o + c) * 4, 4);
            out[q++] = v;
          } else {
            double v;
            std::memcpy(&v, src + (o + c) * 8, 8);
            out[q++] = (float)v;
          }
        }
      }
  return 0;
}

bool put(FILE *f, const std::string &s) { return std::fwrite(s.data(), 1, s.size(), f) == s.size(); }
bool put_raw(FILE *f, const void *p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
// one unformatted sequential record
bool put_record(FILE *f, const void *p, uint32_t n) {
  return put_raw(f, &n, 4) && put_raw(f, p, n) && put_raw(f, &n, 4);
}
// reads one record of exactly n bytes (false at end of file or on a size mismatch)
bool get_record(FILE *f, void *p, uint32_t n) {
  uint32_t a = 0, b = 0;
  if (std::fread(&a, 4, 1, f) != 1 || a != n) return false;
  if (std::fread(p, 1, n, f) != n) return false;
  return std::fread(&b, 4, 1, f) == 1 && b == n;
}

}  // namespace
}  // namespace pmgpu

using namespace pmgpu;

extern "C" {

int pmgpu_output_mesh(const char *outdir, const char *tag, const pmgpu_geom *g, int32_t precision,
                      const void *vort, const void *vel, int32_t on_device) {
  PM_CHECK(outdir && tag && g && vort && vel, "pmgpu_output_mesh: null argument");
  PM_CHECK(precision == PMGPU_FP64 || precision == PMGPU_FP32, "pmgpu_output_mesh: bad precision");
  std::vector<float> w, u;
  if (interior_f32(g, vort, precision, on_device, w) || interior_f32(g, vel, precision, on_device, u)) return -1;
  const std::string dir = rtrim(outdir), tg = rtrim(tag);
  const int *n = g->ncell;
  // one rank: the piece is the whole patch (icell = 1)
  std::string ext;
  for (int d = 0; d < 3; ++d) ext += fmt_i(0, 8) + fmt_i(n[d], 8);
  std::string org, spc;
  for (int d = 0; d < 3; ++d) {
    org += fmt_e(g->xmin[d], 20, 12);
    spc += fmt_e(g->dx[d], 20, 12);
  }
  {  // the .pvti index (formatted sequential: one line per record)
    const std::string path = dir + tg + ".pvti";
    FILE *f = std::fopen(path.c_str(), "wb");
    PM_CHECK(f, "pmgpu_output_mesh: cannot open " + path);
    bool ok = put(f, "<?xml version=\"1.0\"?>\n") &&
              put(f, "<VTKFile type=\"PImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n") &&
              put(f, "  <PImageData WholeExtent=\"" + ext + "\" Ghostlevel=\"0\" Origin=\"" + org + "\" Spacing=\"" +
                         spc + "\">\n") &&
              put(f, "    <PCellData Vectors=\"output\">\n") &&
              put(f, "      <PDataArray type=\"Float32\" Name=\"vort\" NumberOfComponents=\"3\" format=\"appended\" "
                     "offset=\"0\"/>\n") &&
              put(f, "      <PDataArray type=\"Float32\" Name=\"vel\" NumberOfComponents=\"3\" format=\"appended\" "
                     "offset=\"0\"/>\n") &&
              put(f, "    </PCellData>\n") &&
              put(f, "    <Piece  Extent=\"" + ext + "\" Source=\"" + tg + "_P" + fmt_i0(0, 5, 5) + ".vti\"/>\n") &&
              put(f, "  </PImageData>\n") && put(f, "</VTKFile>\n");
    ok = (std::fclose(f) == 0) && ok;
    PM_CHECK(ok, "pmgpu_output_mesh: write failed: " + path);
  }
  {  // the .vti piece (stream output)
    const std::string path = dir + tg + "_P" + fmt_i0(0, 5, 5) + ".vti";
    FILE *f = std::fopen(path.c_str(), "wb");
    PM_CHECK(f, "pmgpu_output_mesh: cannot open " + path);
    const long long ndata = (long long)n[0] * n[1] * n[2];
    const int32_t nb = (int32_t)(3 * ndata * 4);
    bool ok = put(f, "<?xml version=\"1.0\"?>") &&
              put(f, "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">") &&
              put(f, rtrim("  <ImageData WholeExtent=\"" + ext + "\" Ghostlevel=\"0\" Origin=\"" + org +
                           "\" Spacing=\"" + spc + "\">")) &&
              put(f, rtrim("<Piece Extent=\"" + ext + "\">")) && put(f, "<PointData>") && put(f, "</PointData>") &&
              put(f, "<CellData>") &&
              put(f, rtrim("<DataArray type=\"Float32\" Name=\"vort\" NumberOfComponents=\"3\" format=\"appended\" "
                           "offset=\"" + fmt_i(0, 10) + "\"/>")) &&
              put(f, rtrim("<DataArray type=\"Float32\" Name=\"vel\" NumberOfComponents=\"3\" format=\"appended\" "
                           "offset=\"" + fmt_i(nb + 4, 10) + "\"/>")) &&
              put(f, "</CellData>") && put(f, "</Piece>") && put(f, "</ImageData>") &&
              put(f, "<AppendedData encoding=\"raw\">") && put(f, "_") && put_raw(f, &nb, 4) &&
              put_raw(f, w.data(), w.size() * 4) && put_raw(f, &nb, 4) && put_raw(f, u.data(), u.size() * 4) &&
              put(f, "</AppendedData>") && put(f, "</VTKFile>");
    ok = (std::fclose(f) == 0) && ok;
    PM_CHECK(ok, "pmgpu_output_mesh: write failed: " + path);
  }
  return 0;
}

int pmgpu_output_particles(const char *path, int32_t itime, double time, double dtime, const double *pos,
                           const double *vort, int64_t np, int32_t on_device) {
  PM_CHECK(path && (np == 0 || (pos && vort)), "pmgpu_output_particles: null argument");
  PM_CHECK(np >= 0 && np < 2147483648LL, "pmgpu_output_particles: the particle count is a default INTEGER");
  std::vector<double> hp, hw;
  const double *p = pos, *w = vort;
  if (on_device && np > 0) {
    hp.resize(3 * (size_t)np);
    hw.resize(3 * (size_t)np);
    PM_HIP(hipMemcpy(hp.data(), pos, 24 * (size_t)np, hipMemcpyDeviceToHost));
    PM_HIP(hipMemcpy(hw.data(), vort, 24 * (size_t)np, hipMemcpyDeviceToHost));
    p = hp.data();
    w = hw.data();
  }
  FILE *f = std::fopen(path, "wb");
  PM_CHECK(f, std::string("pmgpu_output_particles: cannot open ") + path);
  const int32_t it = itime + 1, n = (int32_t)np;
  const double t = time + dtime;
  bool ok = put_record(f, &it, 4) && put_record(f, &t, 8) && put_record(f, &n, 4);
  double rec[6];
  for (int64_t j = 0; ok && j < np; ++j) {
    for (int k = 0; k < 3; ++k) {
      rec[k] = p[3 * j + k];
      rec[3 + k] = w[3 * j + k];
    }
    ok = put_record(f, rec, sizeof(rec));
  }
  ok = (std::fclose(f) == 0) && ok;
  PM_CHECK(ok, std::string("pmgpu_output_particles: write failed: ") + path);
  return 0;
}

int pmgpu_input_particles(const char *path, double trunc, int32_t *itime, double *time, int64_t *np,
                          double *vort_max, double *pos, double *vort, int64_t cap, int32_t on_device) {
  PM_CHECK(path && np, "pmgpu_input_particles: null argument");
  FILE *f = std::fopen(path, "rb");
  PM_CHECK(f, std::string("pmgpu_input_particles: cannot open ") + path);
  int32_t it = 0, nglob = 0;
  double t = 0.0;
  if (!(get_record(f, &it, 4) && get_record(f, &t, 8) && get_record(f, &nglob, 4))) {
    std::fclose(f);
    PM_CHECK(false, std::string("pmgpu_input_particles: bad header in ") + path);
  }
  std::vector<double> hp, hw;
  double vmax = 0.0;
  double rec[6];
  for (int32_t i = 0; i < nglob; ++i) {
    if (!get_record(f, rec, sizeof(rec))) {
      std::fclose(f);
      PM_CHECK(false, std::string("pmgpu_input_particles: truncated particle records in ") + path);
    }
    const double mag = std::sqrt(rec[3] * rec[3] + rec[4] * rec[4] + rec[5] * rec[5]);
    if (mag >= trunc) {
      for (int k = 0; k < 3; ++k) {
        hp.push_back(rec[k]);
        hw.push_back(rec[3 + k]);
      }
      if (mag > vmax) vmax = mag;
    }
  }
  std::fclose(f);
  const int64_t kept = (int64_t)(hp.size() / 3);
  *np = kept;
  if (itime) *itime = it;
  if (time) *time = t;
  if (vort_max) *vort_max = vmax;
  PM_CHECK(kept > 0, "pmgpu_input_particles: No particles above truncation value.");
  if (pos || vort) {
    PM_CHECK(pos && vort && cap >= kept, "pmgpu_input_particles: output arrays missing or too small");
    if (on_device) {
      PM_HIP(hipMemcpy(pos, hp.data(), hp.size() * 8, hipMemcpyHostToDevice));
      PM_HIP(hipMemcpy(vort, hw.data(), hw.size() * 8, hipMemcpyHostToDevice));
    } else {
      std::memcpy(pos, hp.data(), hp.size() * 8);
      std::memcpy(vort, hw.data(), hw.size() * 8);
    }
  }
  return 0;
}

}  // extern "C"
