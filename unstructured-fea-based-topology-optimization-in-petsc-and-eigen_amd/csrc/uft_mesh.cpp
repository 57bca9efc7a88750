// Mesh-side host code: the synthetic cantilever generator, the reference's
// text formats (InputMeshData.cpp:100-149,228-285; GetBcs/GetBCs.m) and the
// node -> (element, local node) incidence used by the atomic-free node phase.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>

#include "uft_internal.h"

namespace {
// SplitMix64 (Steele, Lea, Flood 2014)
uint64_t sm64_next(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double sm64_u01(uint64_t &s) { return (double)(sm64_next(s) >> 11) * 0x1.0p-53; }
}  // namespace

void uft_node_incidence(int64_t nn, int64_t ne, int npe, const int32_t *elem, std::vector<int32_t> &ptr,
                        std::vector<int32_t> &inc) {
  ptr.assign(nn + 1, 0);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < npe; ++a) ptr[elem[e * npe + a] + 1]++;
  for (int64_t i = 0; i < nn; ++i) ptr[i + 1] += ptr[i];
  inc.assign(ptr[nn], 0);
  std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
  // element-major fill keeps each node's list in ascending element order
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < npe; ++a) inc[pos[elem[e * npe + a]]++] = (int32_t)(e * npe + a);
}

bool uft_fine_ell_pattern(int64_t nn, int npe, const int32_t *elem, const std::vector<int32_t> &iptr,
                          const std::vector<int32_t> &inc, std::vector<int64_t> &off, std::vector<int32_t> &width,
                          std::vector<int32_t> &col, std::vector<int32_t> &cptr, std::vector<int32_t> &clist) {
  const int64_t nslice = (nn + UFT_ELL_SLICE - 1) / UFT_ELL_SLICE;
  std::vector<int64_t> rp(nn + 1, 0);
  std::vector<int32_t> nb;  // sorted neighbour lists (the node itself included)
  nb.reserve((size_t)nn * 27);
  std::vector<int32_t> tmp;
  for (int64_t n = 0; n < nn; ++n) {
    tmp.clear();
    for (int32_t k = iptr[n]; k < iptr[n + 1]; ++k) {
      const int32_t e = inc[k] / npe;
      for (int b = 0; b < npe; ++b) tmp.push_back(elem[(int64_t)e * npe + b]);
    }
    if (tmp.empty()) tmp.push_back((int32_t)n);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    if (tmp.size() > 255) return false;  // slots are stored as bytes
    nb.insert(nb.end(), tmp.begin(), tmp.end());
    rp[n + 1] = (int64_t)nb.size();
  }
  off.assign(nslice + 1, 0);
  width.assign(nslice, 0);
  for (int64_t w = 0; w < nslice; ++w) {
    int wd = 0;
    for (int64_t n = w * UFT_ELL_SLICE; n < std::min(nn, (w + 1) * UFT_ELL_SLICE); ++n)
      wd = std::max(wd, (int)(rp[n + 1] - rp[n]));
    width[w] = wd;
    off[w + 1] = off[w] + (int64_t)wd * UFT_ELL_SLICE;
  }
  // padding slots point at the row's own node (their blocks are zero)
  col.assign(off[nslice], 0);
  for (int64_t n = 0; n < nn; ++n) {
    const int64_t w = n / UFT_ELL_SLICE, lane = n % UFT_ELL_SLICE;
    for (int s = 0; s < width[w]; ++s) {
      const int64_t k = rp[n] + s;
      col[off[w] + s * UFT_ELL_SLICE + lane] = k < rp[n + 1] ? nb[k] : (int32_t)n;
    }
  }
  // contributions of each slot, in slot order: (e * npe + a) * npe + b for
  // the element e holding the row's node as local node a and the column's
  // node as local node b (incidence order of the row's node)
  const int64_t total = off[nslice];
  if ((int64_t)iptr[nn] * npe >= (int64_t)INT32_MAX) return false;
  std::vector<int32_t> cnt(total + 1, 0);
  std::vector<int32_t> slot((size_t)iptr[nn] * npe);
  for (int64_t n = 0; n < nn; ++n) {
    const int32_t *b0 = nb.data() + rp[n], *b1 = nb.data() + rp[n + 1];
    const int64_t w = n / UFT_ELL_SLICE, lane = n % UFT_ELL_SLICE;
    for (int32_t k = iptr[n]; k < iptr[n + 1]; ++k) {
      const int32_t e = inc[k] / npe;
      for (int b = 0; b < npe; ++b) {
        const int s = (int)(std::lower_bound(b0, b1, elem[(int64_t)e * npe + b]) - b0);
        const int64_t pos = off[w] + (int64_t)s * UFT_ELL_SLICE + lane;
        slot[(size_t)k * npe + b] = (int32_t)s;
        cnt[pos + 1]++;
      }
    }
  }
  for (int64_t i = 0; i < total; ++i) cnt[i + 1] += cnt[i];
  cptr = cnt;
  clist.assign(cnt[total], 0);
  for (int64_t n = 0; n < nn; ++n) {
    const int64_t w = n / UFT_ELL_SLICE, lane = n % UFT_ELL_SLICE;
    for (int32_t k = iptr[n]; k < iptr[n + 1]; ++k)
      for (int b = 0; b < npe; ++b) {
        const int64_t pos = off[w] + (int64_t)slot[(size_t)k * npe + b] * UFT_ELL_SLICE + lane;
        clist[cnt[pos]++] = inc[k] * npe + b;  // inc[k] = e * npe + a
      }
  }
  return true;
}

namespace {
// row length of node n in the sliced-ELL pattern: its neighbours ascend and
// include n, and the padding slots after them repeat n
int ell_rowlen(int64_t n, const std::vector<int64_t> &off, const std::vector<int32_t> &width,
               const std::vector<int32_t> &col) {
  const int64_t w = n / UFT_ELL_SLICE, lane = n % UFT_ELL_SLICE;
  int len = width[w] > 0 ? 1 : 0;
  for (int s = 1; s < width[w]; ++s) {
    if (col[off[w] + (int64_t)s * UFT_ELL_SLICE + lane] <= col[off[w] + (int64_t)(s - 1) * UFT_ELL_SLICE + lane]) break;
    len = s + 1;
  }
  return len;
}

struct TileBuilder {
  const double *xyz;
  const std::vector<int64_t> &off;
  const std::vector<int32_t> &width, &col;
  const std::vector<int> &rowlen;
  int tile, hcap;
  std::vector<std::vector<int32_t>> leaves;
  std::vector<int32_t> halo;
  bool ok = true;

  // the tile's halo: every column of its nodes' rows, ascending
  void halo_of(const std::vector<int32_t> &nodes) {
    halo.clear();
    for (int32_t n : nodes) {
      const int64_t w = n / UFT_ELL_SLICE, lane = n % UFT_ELL_SLICE;
      for (int s = 0; s < rowlen[n]; ++s) halo.push_back(col[off[w] + (int64_t)s * UFT_ELL_SLICE + lane]);
      halo.push_back(n);
    }
    std::sort(halo.begin(), halo.end());
    halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
  }
  void split(std::vector<int32_t> &idx) {
    if (!ok) return;
    const int64_t n = (int64_t)idx.size();
    if (n == 0) return;
    if (n <= tile) {
      std::sort(idx.begin(), idx.end());
      halo_of(idx);
      if ((int)halo.size() <= hcap) {
        leaves.push_back(idx);
        return;
      }
      if (n <= UFT_ELL_SLICE) {  // a tile of one slice still over the cap: no tiling
        ok = false;
        return;
      }
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int32_t v : idx)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], xyz[3 * (int64_t)v + c]);
        hi[c] = std::max(hi[c], xyz[3 * (int64_t)v + c]);
      }
    int ax = 0;
    for (int c = 1; c < 3; ++c)
      if (hi[c] - lo[c] > hi[ax] - lo[ax]) ax = c;
    // the lower part: a multiple of tile nodes near half (half of a tile over the cap)
    const int64_t nl = n > tile ? (int64_t)tile * ((n / tile + 1) / 2) : n / 2;
    auto less = [&](int32_t a, int32_t b) {
      const double xa = xyz[3 * (int64_t)a + ax], xb = xyz[3 * (int64_t)b + ax];
      return xa < xb || (xa == xb && a < b);
    };
    std::nth_element(idx.begin(), idx.begin() + nl, idx.end(), less);
    std::vector<int32_t> hi_part(idx.begin() + nl, idx.end());
    idx.resize(nl);
    split(idx);
    split(hi_part);
  }
};
}  // namespace

bool uft_node_tiles(int64_t nn, const double *xyz, const std::vector<int64_t> &off, const std::vector<int32_t> &width,
                    const std::vector<int32_t> &col, int tile, int hcap, NodeTiles &T) {
  if (nn <= 0 || tile % UFT_ELL_SLICE != 0 || hcap > 65535) return false;
  std::vector<int> rowlen(nn);
  for (int64_t n = 0; n < nn; ++n) rowlen[n] = ell_rowlen(n, off, width, col);
  TileBuilder B{xyz, off, width, col, rowlen, tile, hcap, {}, {}, true};
  {
    std::vector<int32_t> all(nn);
    for (int64_t n = 0; n < nn; ++n) all[n] = (int32_t)n;
    B.split(all);
  }
  if (!B.ok) return false;
  T = NodeTiles();
  T.ntile = (int64_t)B.leaves.size();
  const int64_t npos = T.ntile * tile, nslice = npos / UFT_ELL_SLICE;
  T.node.assign(npos, -1);
  std::vector<int64_t> pos(nn, -1);
  for (int64_t t = 0; t < T.ntile; ++t)
    for (size_t i = 0; i < B.leaves[t].size(); ++i) {
      T.node[t * tile + (int64_t)i] = B.leaves[t][i];
      pos[B.leaves[t][i]] = t * tile + (int64_t)i;
    }
  T.off.assign(nslice + 1, 0);
  T.width.assign(nslice, 0);
  for (int64_t w = 0; w < nslice; ++w) {
    int wd = 0;
    for (int64_t p = w * UFT_ELL_SLICE; p < (w + 1) * UFT_ELL_SLICE; ++p)
      if (T.node[p] >= 0) wd = std::max(wd, rowlen[T.node[p]]);
    T.width[w] = wd;
    T.off[w + 1] = T.off[w] + (int64_t)wd * UFT_ELL_SLICE;
  }
  T.nslot = T.off[nslice];
  T.lc.assign(T.nslot, 0);
  T.kmap.assign(off.back(), -1);
  T.hoff.assign(T.ntile + 1, 0);
  std::vector<int32_t> mark(nn, -1);
  for (int64_t t = 0; t < T.ntile; ++t) {
    B.halo_of(B.leaves[t]);
    T.hmax = std::max(T.hmax, (int)B.halo.size());
    for (size_t i = 0; i < B.halo.size(); ++i) mark[B.halo[i]] = (int32_t)i;
    T.hnode.insert(T.hnode.end(), B.halo.begin(), B.halo.end());
    T.hoff[t + 1] = (int32_t)T.hnode.size();
    for (int32_t n : B.leaves[t]) {
      const int64_t p = pos[n], w2 = p / UFT_ELL_SLICE, l2 = p % UFT_ELL_SLICE;
      const int64_t w = n / UFT_ELL_SLICE, lane = n % UFT_ELL_SLICE;
      for (int s = 0; s < T.width[w2]; ++s) {
        const int64_t k2 = T.off[w2] + (int64_t)s * UFT_ELL_SLICE + l2;
        if (s < rowlen[n]) {
          const int64_t k = off[w] + (int64_t)s * UFT_ELL_SLICE + lane;
          T.lc[k2] = (uint16_t)mark[col[k]];
          T.kmap[k] = (int32_t)k2;
        } else {
          T.lc[k2] = (uint16_t)mark[n];  // padding: the node itself, zero block
        }
      }
    }
    for (int32_t v : B.halo) mark[v] = -1;
  }
  if ((int64_t)T.hnode.size() >= (int64_t)INT32_MAX || T.nslot >= (int64_t)INT32_MAX) return false;
  return true;
}

extern "C" {

int64_t uft_box_nodes(int nx, int ny, int nz) { return (int64_t)(nx + 1) * (ny + 1) * (nz + 1); }
int64_t uft_box_elems(int npe, int nx, int ny, int nz) {
  const int64_t h = (int64_t)nx * ny * nz;
  return npe == 4 ? 6 * h : h;
}

int uft_gen_box(int npe, int nx, int ny, int nz, double lx, double ly, double lz, double jitter, uint64_t seed,
                double *xyz, int32_t *elem) {
  if (nx < 1 || ny < 1 || nz < 1 || (npe != 8 && npe != 4) || !xyz || !elem) return UFT_ERR_ARG;
  if (uft_box_elems(npe, nx, ny, nz) * npe >= (int64_t)INT32_MAX) return UFT_ERR_ARG;
  uint64_t s = seed;
  const double hx = lx / nx, hy = ly / ny, hz = lz / nz;
  int64_t n = 0;
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i, ++n) {
        double x = lx * (double)i / (double)nx;
        double y = ly * (double)j / (double)ny;
        double z = lz * (double)k / (double)nz;
        if (jitter != 0.0 && i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz) {
          x += (2.0 * sm64_u01(s) - 1.0) * jitter * hx;
          y += (2.0 * sm64_u01(s) - 1.0) * jitter * hy;
          z += (2.0 * sm64_u01(s) - 1.0) * jitter * hz;
        }
        xyz[3 * n] = x;
        xyz[3 * n + 1] = y;
        xyz[3 * n + 2] = z;
      }
  // six positively oriented tets around the local 0-7 diagonal
  static const int tets[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7}, {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};
  const int64_t sx = nx + 1, sxy = (int64_t)(nx + 1) * (ny + 1);
  int64_t e = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t b = i + sx * j + sxy * k;
        const int32_t hv[8] = {(int32_t)b,         (int32_t)(b + 1),         (int32_t)(b + sx),
                               (int32_t)(b + sx + 1), (int32_t)(b + sxy),       (int32_t)(b + sxy + 1),
                               (int32_t)(b + sxy + sx), (int32_t)(b + sxy + sx + 1)};
        if (npe == 8) {
          for (int a = 0; a < 8; ++a) elem[8 * e + a] = hv[a];
          ++e;
        } else {
          for (int t = 0; t < 6; ++t, ++e)
            for (int a = 0; a < 4; ++a) elem[4 * e + a] = hv[tets[t][a]];
        }
      }
  return UFT_OK;
}

int uft_box_bcs(int nx, int ny, int nz, int load_kind, int32_t *fixed, int *nfixed, int32_t *load_dofs,
                double *load_vals, int *nload) {
  if (nx < 1 || ny < 1 || nz < 1 || !fixed || !nfixed || !load_dofs || !load_vals || !nload) return UFT_ERR_ARG;
  const int64_t sx = nx + 1, sxy = (int64_t)(nx + 1) * (ny + 1);
  int nf = 0;
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j) {
      const int64_t n = sx * j + sxy * k;
      fixed[nf++] = (int32_t)(3 * n);
      fixed[nf++] = (int32_t)(3 * n + 1);
      fixed[nf++] = (int32_t)(3 * n + 2);
    }
  *nfixed = nf;
  int nl = 0;
  if (load_kind == 0) {
    for (int k = 0; k <= nz; ++k) {
      const int64_t n = nx + sxy * k;
      load_dofs[nl] = (int32_t)(3 * n + 1);
      load_vals[nl++] = -1.0 / (double)(nz + 1);
    }
  } else if (load_kind == 1) {
    if (ny < 2 || nz < 2) return UFT_ERR_ARG;
    const int jc = ny / 2, kc = nz / 2;
    for (int k = kc - 1; k <= kc + 1; ++k)
      for (int j = jc - 1; j <= jc + 1; ++j) {
        const int64_t n = nx + sx * j + sxy * k;
        load_dofs[nl] = (int32_t)(3 * n + 1);
        load_vals[nl++] = -1.0 / 9.0;
      }
  } else if (load_kind == 2) {
    // eight load cases: seeded 3x3 patches on x = Lx, unit total each
    if (ny < 2 || nz < 2) return UFT_ERR_ARG;
    uint64_t st = 8;
    for (int lc = 0; lc < 8; ++lc) {
      const int jc = 1 + (int)std::floor(sm64_u01(st) * (ny - 1));
      const int kc = 1 + (int)std::floor(sm64_u01(st) * (nz - 1));
      for (int k = kc - 1; k <= kc + 1; ++k)
        for (int j = jc - 1; j <= jc + 1; ++j) {
          const int64_t n = nx + sx * j + sxy * k;
          load_dofs[nl] = (int32_t)(3 * n + 1);
          load_vals[nl++] = -1.0 / 9.0;
        }
    }
  } else {
    return UFT_ERR_ARG;
  }
  *nload = nl;
  return UFT_OK;
}

}  // extern "C"

// ---- text readers: whitespace-separated tokens, no terminator required
template <typename T>
static int read_tokens(uft_ctx *C, const char *path, int64_t count, T *out) {
  std::ifstream in(path);
  if (!in.is_open()) return uft_set_error(C, UFT_ERR_IO, "cannot open %s", path);
  for (int64_t i = 0; i < count; ++i) {
    if (!(in >> out[i]))
      return uft_set_error(C, UFT_ERR_IO, "%s: expected %lld values, read %lld", path, (long long)count, (long long)i);
  }
  return UFT_OK;
}

// InputMeshData::GetNodeAndElementData's parse (InputMeshData.cpp:100-149)
// into host arrays; errors go to C (NULL: the thread's context-free message).
static int mesh_read_text(uft_ctx *C, int npe, int64_t nn, int64_t ne, const char *vertex_path,
                          const char *element_path, double *xyz, int32_t *elem) {
  if (nn <= 0 || ne <= 0 || (npe != 8 && npe != 4) || !vertex_path || !element_path || !xyz || !elem)
    return uft_set_error(C, UFT_ERR_ARG, "mesh text reader: bad arguments");
  int rc = read_tokens(C, vertex_path, nn * 3, xyz);
  if (rc) return rc;
  return read_tokens(C, element_path, ne * npe, elem);
}

// InputMeshData::GetBcs (InputMeshData.cpp:228-285): for each case i = 1..nlc
// dir/loaddof_<i>.txt and dir/load_<i>.txt (nload values each), then
// dir/fixeddofs.txt (nfixed values), into host arrays.
static int bc_read_text(uft_ctx *C, const char *dir, int nfixed, int nlc, int nload, int32_t *fixed,
                        int32_t *ld, double *lv) {
  if (!dir || nfixed < 0 || nlc < 1 || nload < 0 || (nfixed && !fixed) || (nload && (!ld || !lv)))
    return uft_set_error(C, UFT_ERR_ARG, "BC text reader: bad arguments");
  char path[4096];
  for (int i = 0; i < nlc; ++i) {
    std::snprintf(path, sizeof(path), "%s/loaddof_%d.txt", dir, i + 1);
    int rc = read_tokens(C, path, nload, ld + (size_t)i * nload);
    if (rc) return rc;
    std::snprintf(path, sizeof(path), "%s/load_%d.txt", dir, i + 1);
    rc = read_tokens(C, path, nload, lv + (size_t)i * nload);
    if (rc) return rc;
  }
  std::snprintf(path, sizeof(path), "%s/fixeddofs.txt", dir);
  return read_tokens(C, path, nfixed, fixed);
}

extern "C" int uft_node_tiles_host(int npe, int64_t nn, int64_t ne, const int32_t *elem, const double *xyz,
                                   int64_t *ntile, int64_t *nhalo, int *hmax, int32_t *node, int32_t *hoff,
                                   int32_t *hnode) {
  if ((npe != 8 && npe != 4) || nn <= 0 || ne <= 0 || !elem || !xyz || !ntile || !nhalo || !hmax) return UFT_ERR_ARG;
  for (int64_t i = 0; i < ne * npe; ++i)
    if (elem[i] < 0 || elem[i] >= nn) return UFT_ERR_ARG;
  std::vector<int32_t> iptr, inc, width, col, cptr, clist;
  std::vector<int64_t> off;
  uft_node_incidence(nn, ne, npe, elem, iptr, inc);
  if (!uft_fine_ell_pattern(nn, npe, elem, iptr, inc, off, width, col, cptr, clist)) return UFT_ERR_ARG;
  NodeTiles T;
  if (!uft_node_tiles(nn, xyz, off, width, col, UFT_TILE, UFT_TILE_HCAP, T)) return UFT_ERR_ARG;
  *ntile = T.ntile;
  *nhalo = (int64_t)T.hnode.size();
  *hmax = T.hmax;
  if (node) std::copy(T.node.begin(), T.node.end(), node);
  if (hoff) std::copy(T.hoff.begin(), T.hoff.end(), hoff);
  if (hnode) std::copy(T.hnode.begin(), T.hnode.end(), hnode);
  return UFT_OK;
}

extern "C" int uft_mesh_read_text(int npe, int64_t nn, int64_t ne, const char *vertex_path, const char *element_path,
                                  double *xyz, int32_t *elem) {
  return mesh_read_text(nullptr, npe, nn, ne, vertex_path, element_path, xyz, elem);
}

extern "C" int uft_bc_read_text(const char *dir, int nfixed, int nlc, int nload, int32_t *fixed, int32_t *load_dofs,
                                double *load_vals) {
  return bc_read_text(nullptr, dir, nfixed, nlc, nload, fixed, load_dofs, load_vals);
}

extern "C" int uft_mesh_load_text(uft_ctx *C, int npe, int64_t nn, int64_t ne, const char *vertex_path,
                                  const char *element_path) {
  if (!C) return UFT_ERR_ARG;
  if (nn <= 0 || ne <= 0 || (npe != 8 && npe != 4) || !vertex_path || !element_path)
    return uft_set_error(C, UFT_ERR_ARG, "uft_mesh_load_text: bad arguments");
  std::vector<double> xyz(nn * 3);
  std::vector<int32_t> elem(ne * npe);
  int rc = mesh_read_text(C, npe, nn, ne, vertex_path, element_path, xyz.data(), elem.data());
  if (rc) return rc;
  return uft_mesh_set(C, npe, nn, ne, elem.data(), xyz.data());
}

extern "C" int uft_bc_load_text(uft_ctx *C, const char *dir, int nfixed, int nlc, int nload) {
  if (!C) return UFT_ERR_ARG;
  if (!dir || nfixed < 0 || nlc < 1 || nload < 0) return uft_set_error(C, UFT_ERR_ARG, "uft_bc_load_text: bad arguments");
  std::vector<int32_t> fixed(nfixed), ld((size_t)nlc * nload);
  std::vector<double> lv((size_t)nlc * nload);
  int rc = bc_read_text(C, dir, nfixed, nlc, nload, fixed.data(), ld.data(), lv.data());
  if (rc) return rc;
  return uft_bc_set(C, nfixed, fixed.data(), nlc, nload, ld.data(), lv.data());
}

// ---- binary cache of a mesh and its boundary conditions
// Layout (little endian): "UFTMESH1", int32 version 1, int32 npe, int64 nn,
// int64 ne, int32 nfixed, int32 nlc, int32 nload, int32 0, then xyz (nn*3
// f64), elem (ne*npe i32), fixed (i32), load dofs (nlc*nload i32), load
// values (f64), and a trailing FNV-1a 64 checksum of everything before it.
namespace {
constexpr char kMagic[8] = {'U', 'F', 'T', 'M', 'E', 'S', 'H', '1'};
struct CacheHead {
  char magic[8];
  int32_t version, npe;
  int64_t nn, ne;
  int32_t nfixed, nlc, nload, pad;
};
uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
  const unsigned char *b = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}
}  // namespace

extern "C" int uft_cache_save(uft_ctx *C, const char *path) {
  if (!C || !path) return UFT_ERR_ARG;
  if (!C->have_mesh || !C->have_bc) return uft_set_error(C, UFT_ERR_STATE, "uft_cache_save: mesh and BCs not set");
  CacheHead h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = 1;
  h.npe = C->npe;
  h.nn = C->nn;
  h.ne = C->ne;
  h.nfixed = (int32_t)C->fixed.size();
  h.nlc = C->nlc;
  h.nload = C->nload;
  std::FILE *f = std::fopen(path, "wb");
  if (!f) return uft_set_error(C, UFT_ERR_IO, "cannot create %s", path);
  uint64_t ck = 1469598103934665603ull;
  bool ok = true;
  auto put = [&](const void *p, size_t n) {
    if (!ok || n == 0) return;
    ok = std::fwrite(p, 1, n, f) == n;
    ck = fnv1a(ck, p, n);
  };
  put(&h, sizeof(h));
  put(C->xyz.data(), C->xyz.size() * sizeof(double));
  put(C->elem.data(), C->elem.size() * sizeof(int32_t));
  put(C->fixed.data(), C->fixed.size() * sizeof(int32_t));
  put(C->ldofs.data(), C->ldofs.size() * sizeof(int32_t));
  put(C->lvals.data(), C->lvals.size() * sizeof(double));
  if (ok) ok = std::fwrite(&ck, 1, sizeof(ck), f) == sizeof(ck);
  if (std::fclose(f) != 0) ok = false;
  return ok ? UFT_OK : uft_set_error(C, UFT_ERR_IO, "write error on %s", path);
}

extern "C" int uft_cache_load(uft_ctx *C, const char *path) {
  if (!C || !path) return UFT_ERR_ARG;
  std::FILE *f = std::fopen(path, "rb");
  if (!f) return uft_set_error(C, UFT_ERR_IO, "cannot open %s", path);
  uint64_t ck = 1469598103934665603ull;
  bool ok = true;
  auto get = [&](void *p, size_t n) {
    if (!ok || n == 0) return;
    ok = std::fread(p, 1, n, f) == n;
    if (ok) ck = fnv1a(ck, p, n);
  };
  CacheHead h{};
  get(&h, sizeof(h));
  if (!ok || std::memcmp(h.magic, kMagic, 8) != 0 || h.version != 1 || (h.npe != 8 && h.npe != 4) || h.nn <= 0 ||
      h.ne <= 0 || h.nfixed < 0 || h.nlc < 1 || h.nload < 0 || h.nn > (int64_t)1 << 31 || h.ne > (int64_t)1 << 31) {
    std::fclose(f);
    return uft_set_error(C, UFT_ERR_IO, "%s: not a mesh cache (bad header)", path);
  }
  std::vector<double> xyz(h.nn * 3), lv((size_t)h.nlc * h.nload);
  std::vector<int32_t> elem(h.ne * h.npe), fixed(h.nfixed), ld((size_t)h.nlc * h.nload);
  get(xyz.data(), xyz.size() * sizeof(double));
  get(elem.data(), elem.size() * sizeof(int32_t));
  get(fixed.data(), fixed.size() * sizeof(int32_t));
  get(ld.data(), ld.size() * sizeof(int32_t));
  get(lv.data(), lv.size() * sizeof(double));
  uint64_t stored = 0;
  const uint64_t want = ck;
  if (ok) ok = std::fread(&stored, 1, sizeof(stored), f) == sizeof(stored);
  std::fclose(f);
  if (!ok) return uft_set_error(C, UFT_ERR_IO, "%s: truncated mesh cache", path);
  if (stored != want) return uft_set_error(C, UFT_ERR_IO, "%s: checksum mismatch", path);
  int rc = uft_mesh_set(C, h.npe, h.nn, h.ne, elem.data(), xyz.data());
  if (rc) return rc;
  return uft_bc_set(C, h.nfixed, fixed.data(), h.nlc, h.nload, ld.data(), lv.data());
}
