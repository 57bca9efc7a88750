// ASCII VTU output of a design (Write2Vtu::WriteResults, Write2Vtu.cpp:5-173):
// points ("nodes"), connectivity with the hex8 corners reordered from the
// diagonal numbering to VTK's counterclockwise one (0,1,3,2,4,5,7,6,
// Write2Vtu.cpp:89-96), offsets, cell types (12 hex8, 10 tet4,
// Write2Vtu.cpp:112-116), point data ux/uy/uz and cell data xphys.  Numbers
// are written in scientific notation with `digits` significant digits after
// the point (the reference's setprecision(6); 17 round-trips FP64 exactly).
#include <cstdio>
#include <string>
#include <vector>

#include "uft_internal.h"

namespace {

struct Out {
  std::FILE *f;
  std::string buf;
  bool ok = true;
  void flush() {
    if (!buf.empty() && std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) ok = false;
    buf.clear();
  }
  void put(const char *s) {
    buf += s;
    if (buf.size() > (1u << 20)) flush();
  }
  void num(double v, int digits, char sep) {
    char t[64];
    std::snprintf(t, sizeof(t), "%.*e%c", digits, v, sep);
    put(t);
  }
  void inum(long long v, char sep) {
    char t[32];
    std::snprintf(t, sizeof(t), "%lld%c", v, sep);
    put(t);
  }
};

}  // namespace

extern "C" int uft_vtu_write_host(const char *path, int npe, int64_t nn, int64_t ne, const double *xyz,
                                  const int32_t *elem, const double *u, const double *xphys, int digits) {
  if (!path || (npe != 8 && npe != 4) || nn <= 0 || ne <= 0 || !xyz || !elem || !u || !xphys || digits < 1 ||
      digits > 17)
    return uft_set_error(nullptr, UFT_ERR_ARG, "uft_vtu_write_host: bad arguments");
  Out o;
  o.f = std::fopen(path, "w");
  if (!o.f) return uft_set_error(nullptr, UFT_ERR_IO, "cannot create %s", path);
  char t[256];
  o.put("<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\">\n<UnstructuredGrid>\n");
  std::snprintf(t, sizeof(t), "<Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", (long long)nn, (long long)ne);
  o.put(t);
  o.put("<Points>\n<DataArray type=\"Float64\" Name=\"nodes\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int64_t i = 0; i < nn; ++i) {
    o.num(xyz[3 * i], digits, ' ');
    o.num(xyz[3 * i + 1], digits, ' ');
    o.num(xyz[3 * i + 2], digits, '\n');
  }
  o.put("</DataArray>\n</Points>\n<Cells>\n");
  o.put("<DataArray type=\"Int32\" Name=\"connectivity\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  static const int hex_vtk[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < npe; ++a) o.inum(elem[e * npe + (npe == 8 ? hex_vtk[a] : a)], a + 1 < npe ? ' ' : '\n');
  o.put("</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int64_t e = 0; e < ne; ++e) o.inum((long long)(e + 1) * npe, ' ');
  o.put("\n</DataArray>\n<DataArray type=\"Int32\" Name=\"types\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int64_t e = 0; e < ne; ++e) o.inum(npe == 8 ? 12 : 10, ' ');
  o.put("\n</DataArray>\n</Cells>\n<PointData>\n");
  static const char *comp[3] = {"ux", "uy", "uz"};
  for (int c = 0; c < 3; ++c) {
    std::snprintf(t, sizeof(t), "<DataArray type=\"Float64\" NumberOfComponents=\"1\" Name=\"%s\" format=\"ascii\">\n",
                  comp[c]);
    o.put(t);
    for (int64_t i = 0; i < nn; ++i) o.num(u[3 * i + c], digits, ' ');
    o.put("\n</DataArray>\n");
  }
  o.put("</PointData>\n<CellData>\n");
  o.put("<DataArray type=\"Float64\" NumberOfComponents=\"1\" Name=\"xphys\" format=\"ascii\">\n");
  for (int64_t e = 0; e < ne; ++e) o.num(xphys[e], digits, ' ');
  o.put("\n</DataArray>\n</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  o.flush();
  const bool closed = std::fclose(o.f) == 0;
  if (!o.ok || !closed) return uft_set_error(nullptr, UFT_ERR_IO, "write error on %s", path);
  return UFT_OK;
}

extern "C" int uft_write_vtu(uft_ctx *C, const char *path, int xphys_id, int u_id, int digits) {
  if (!C) return UFT_ERR_ARG;
  const int64_t ne = uft_size(C, UFT_SIZE_NE), nn = uft_size(C, UFT_SIZE_NN);
  const int npe = (int)uft_size(C, UFT_SIZE_NPE);
  if (ne <= 0 || nn <= 0) return uft_set_error(C, UFT_ERR_STATE, "uft_write_vtu: no mesh");
  std::vector<int32_t> elem((size_t)ne * npe);
  std::vector<double> xyz((size_t)nn * 3), u((size_t)nn * 3), x((size_t)ne);
  int rc = uft_mesh_get(C, elem.data(), xyz.data());
  if (!rc) rc = uft_vec_get(C, u_id, u.data(), nn * 3);
  if (!rc) rc = uft_vec_get(C, xphys_id, x.data(), ne);
  if (rc) return rc;
  rc = uft_vtu_write_host(path, npe, nn, ne, xyz.data(), elem.data(), u.data(), x.data(), digits);
  if (rc) return uft_set_error(C, rc, "%s", uft_last_error(nullptr));
  return UFT_OK;
}
