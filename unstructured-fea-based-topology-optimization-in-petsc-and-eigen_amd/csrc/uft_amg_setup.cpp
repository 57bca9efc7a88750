// Host construction of the aggregation multigrid hierarchy.
//
// Replaces the setup half of PETSc's PCGAMG as configured by the reference
// (SolveFem.cpp:129-136: 4 levels, near-null space = the six rigid-body modes
// from the node coordinates).  The aggregates and tentative prolongators
// depend only on the mesh and the Dirichlet set, so they are built once; the
// density-dependent coarse operators are recomputed on the device for every
// design iteration from ordered contribution lists (no atomics).
//
//   aggregation   greedy three-phase aggregation on the unit graph (nodes that
//                 share an element; then aggregates whose operator blocks couple)
//   tentative P   per aggregate, QR of the stacked near-null-space rows; rows of
//                 fixed DOFs are zero so the coarse spaces never see them
//   lists         for each coarse block, the contributions in element (or
//                 fine-block) order
//   smoothed      with sa, level 0's prolongator is smoothed at every refresh
//                 with the current operator, P = (I - omega D^-1 A) P_tent
//                 (smoothed aggregation, as PCGAMG's one smoothing step); the
//                 host builds the patterns of P, of A P and of the level-1
//                 operator P^T A P, and the supports of the aggregates
#include <algorithm>
#include <cmath>
#include <cstring>

#include "uft_internal.h"

namespace {

// graph without self loops, sorted neighbour lists
struct Graph {
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  int64_t n() const { return (int64_t)rp.size() - 1; }
};

Graph node_graph(int64_t nn, int64_t ne, int npe, const int32_t *elem) {
  std::vector<int64_t> cnt(nn + 1, 0);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < npe; ++a) cnt[elem[e * npe + a] + 1] += npe - 1;
  for (int64_t i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
  std::vector<int32_t> tmp(cnt[nn]);
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < npe; ++a) {
      const int32_t n = elem[e * npe + a];
      for (int b = 0; b < npe; ++b)
        if (b != a) tmp[pos[n]++] = elem[e * npe + b];
    }
  Graph g;
  g.rp.assign(nn + 1, 0);
  g.ci.reserve(cnt[nn] / 2 + 16);
  for (int64_t i = 0; i < nn; ++i) {
    auto b = tmp.begin() + cnt[i], en = tmp.begin() + cnt[i + 1];
    std::sort(b, en);
    auto last = std::unique(b, en);
    for (auto it = b; it != last; ++it)
      if (*it != i) g.ci.push_back(*it);
    g.rp[i + 1] = (int64_t)g.ci.size();
  }
  return g;
}

// Greedy aggregation (Vanek, Mandel, Brezina 1996, phases 1-3).
int64_t aggregate(const Graph &g, std::vector<int32_t> &agg) {
  const int64_t n = g.n();
  agg.assign(n, -1);
  int32_t na = 0;
  // phase 1: a unit whose whole neighbourhood is free roots a new aggregate
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free_nb = true;
    for (int64_t k = g.rp[i]; k < g.rp[i + 1]; ++k)
      if (agg[g.ci[k]] >= 0) {
        free_nb = false;
        break;
      }
    if (!free_nb) continue;
    agg[i] = na;
    for (int64_t k = g.rp[i]; k < g.rp[i + 1]; ++k) agg[g.ci[k]] = na;
    ++na;
  }
  // phase 2: free units join the neighbouring aggregate they touch most
  std::vector<int32_t> tentative(agg);
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    int32_t best = -1, bestc = 0;
    for (int64_t k = g.rp[i]; k < g.rp[i + 1]; ++k) {
      const int32_t a = agg[g.ci[k]];
      if (a < 0) continue;
      int32_t c = 0;
      for (int64_t k2 = g.rp[i]; k2 < g.rp[i + 1]; ++k2)
        if (agg[g.ci[k2]] == a) ++c;
      if (c > bestc || (c == bestc && a < best)) {
        best = a;
        bestc = c;
      }
    }
    tentative[i] = best;
  }
  agg.swap(tentative);
  // phase 3: what is left forms aggregates with its free neighbours
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    agg[i] = na;
    for (int64_t k = g.rp[i]; k < g.rp[i + 1]; ++k)
      if (agg[g.ci[k]] < 0) agg[g.ci[k]] = na;
    ++na;
  }
  return na;
}

// Modified Gram-Schmidt (two passes) on the 6 columns of B (rows x 6,
// row-major); B is overwritten by Q, R (6x6 row-major) returned.  Columns
// whose remaining norm falls below tol * (initial norm) are dropped (zero).
void qr6(std::vector<double> &B, int64_t rows, double *R) {
  std::memset(R, 0, 36 * sizeof(double));
  double nrm0[6];
  for (int j = 0; j < 6; ++j) {
    double s = 0;
    for (int64_t r = 0; r < rows; ++r) s += B[r * 6 + j] * B[r * 6 + j];
    nrm0[j] = std::sqrt(s);
  }
  for (int j = 0; j < 6; ++j) {
    for (int pass = 0; pass < 2; ++pass)
      for (int i = 0; i < j; ++i) {
        double d = 0;
        for (int64_t r = 0; r < rows; ++r) d += B[r * 6 + i] * B[r * 6 + j];
        for (int64_t r = 0; r < rows; ++r) B[r * 6 + j] -= d * B[r * 6 + i];
        R[i * 6 + j] += d;
      }
    double s = 0;
    for (int64_t r = 0; r < rows; ++r) s += B[r * 6 + j] * B[r * 6 + j];
    s = std::sqrt(s);
    if (s <= 1e-10 * nrm0[j] || s == 0.0) {
      for (int64_t r = 0; r < rows; ++r) B[r * 6 + j] = 0.0;
      R[j * 6 + j] = 0.0;
    } else {
      for (int64_t r = 0; r < rows; ++r) B[r * 6 + j] /= s;
      R[j * 6 + j] = s;
    }
  }
}

int32_t find_col(const std::vector<int32_t> &rp, const std::vector<int32_t> &ci, int32_t row, int32_t col) {
  auto b = ci.begin() + rp[row], e = ci.begin() + rp[row + 1];
  auto it = std::lower_bound(b, e, col);
  return (it != e && *it == col) ? (int32_t)(it - ci.begin()) : -1;
}

// Design batches: renumber aggregates in order of first appearance over the
// units, so that (units being grouped by design, and no aggregate spanning
// two designs) each design's aggregates are contiguous; returns the design of
// every aggregate and each design's first aggregate.
void batch_renumber(std::vector<int32_t> &agg, int64_t na, const std::vector<int32_t> &udes, int nseg,
                    std::vector<int32_t> &ades, std::vector<int64_t> &aoff) {
  std::vector<int32_t> map(na, -1);
  int32_t next = 0;
  ades.assign(na, 0);
  for (size_t u = 0; u < agg.size(); ++u) {
    int32_t &a = agg[u];
    if (map[a] < 0) {
      map[a] = next;
      ades[next] = udes[u];
      ++next;
    }
    a = map[a];
  }
  aoff.assign(nseg + 1, 0);
  for (int64_t a = 0; a < na; ++a) aoff[ades[a] + 1]++;
  for (int d = 0; d < nseg; ++d) aoff[d + 1] += aoff[d];
}
int64_t max_per_design(const std::vector<int64_t> &off) {
  int64_t m = 0;
  for (size_t d = 0; d + 1 < off.size(); ++d) m = std::max(m, off[d + 1] - off[d]);
  return m;
}

// children lists of an aggregation
void children(const std::vector<int32_t> &agg, int64_t na, std::vector<int32_t> &cptr, std::vector<int32_t> &child) {
  cptr.assign(na + 1, 0);
  for (size_t u = 0; u < agg.size(); ++u) cptr[agg[u] + 1]++;
  for (int64_t a = 0; a < na; ++a) cptr[a + 1] += cptr[a];
  child.assign(agg.size(), 0);
  std::vector<int32_t> pos(cptr.begin(), cptr.end() - 1);
  for (size_t u = 0; u < agg.size(); ++u) child[pos[agg[u]]++] = (int32_t)u;
}

}  // namespace

// Builds the host description of every transfer level.  H[l] describes the
// transfer from level l to level l+1 and the pattern/lists of level l+1.
int uft_amg_build_host(uft_ctx *C, std::vector<AmgHostLevel> &H, std::vector<int64_t> &nunits, bool sa) {
  H.clear();
  nunits.clear();
  const int npe = C->npe;
  const int64_t nn = C->nn, ne = C->ne;
  if (ne * npe * npe / 2 >= (int64_t)INT32_MAX)
    return uft_set_error(C, UFT_ERR_ARG, "mesh too large for 32-bit Galerkin lists (ne=%lld)", (long long)ne);
  std::vector<uint8_t> fixm(nn, 0);
  for (int32_t d : C->fixed) fixm[d / 3] |= (uint8_t)(1u << (d % 3));
  // design batches: per-unit designs and per-design unit offsets of each level
  const int nseg = C->bt.nseg;
  std::vector<int32_t> udes, ades;
  if (nseg > 0) {
    udes.assign(nn, 0);
    for (int d = 0; d < nseg; ++d)
      for (int64_t n = C->bt.noff[d]; n < C->bt.noff[d + 1]; ++n) udes[n] = d;
    C->bt.uoff[0] = C->bt.noff;
  }

  const int maxlev = std::max(2, std::min((int)UFT_MAXLEV, C->P.amg_levels));
  // one design (one right-hand side or several load cases): its coarsest
  // level is inverted by the tiled sweep (kernels_dense.hip) and solved
  // exactly, so coarsening stops at coarse_max; design batches keep coarsest
  // levels of at most 600 unknowns per design for their per-design dense
  // solves in one workgroup each
  const bool single = C->bt.nseg == 0;
  const int64_t cmax = single ? std::max(6, C->P.coarse_max) : std::min(std::max(6, C->P.coarse_max), 600);
  const int64_t dense_limit = single ? C->P.coarse_dense_max : std::min<int64_t>(C->P.coarse_dense_max, UFT_TAIL_MAX);

  // ---- level 0 -> 1
  Graph g = node_graph(nn, ne, npe, C->elem.data());
  nunits.push_back(nn);
  {
    AmgHostLevel h;
    std::vector<int32_t> agg;
    const int64_t na = aggregate(g, agg);
    if (nseg > 0) {
      batch_renumber(agg, na, udes, nseg, ades, C->bt.uoff[1]);
      udes.swap(ades);
    }
    h.nunit_next = na;
    h.parent = agg;
    children(agg, na, h.cptr, h.child);
    h.P.assign(nn * 18, 0.0);
    std::vector<double> Rn(na * 36, 0.0);
    std::vector<double> B;
    for (int64_t a = 0; a < na; ++a) {
      const int32_t c0 = h.cptr[a], c1 = h.cptr[a + 1];
      const int64_t k = c1 - c0;
      double cx = 0, cy = 0, cz = 0;
      for (int32_t q = c0; q < c1; ++q) {
        const int32_t n = h.child[q];
        cx += C->xyz[3 * n];
        cy += C->xyz[3 * n + 1];
        cz += C->xyz[3 * n + 2];
      }
      cx /= k; cy /= k; cz /= k;
      B.assign(3 * k * 6, 0.0);
      for (int64_t q = 0; q < k; ++q) {
        const int32_t n = h.child[c0 + q];
        const double x = C->xyz[3 * n] - cx, y = C->xyz[3 * n + 1] - cy, z = C->xyz[3 * n + 2] - cz;
        double *r0 = &B[(3 * q + 0) * 6], *r1 = &B[(3 * q + 1) * 6], *r2 = &B[(3 * q + 2) * 6];
        r0[0] = 1; r1[1] = 1; r2[2] = 1;
        r0[3] = -y; r1[3] = x;   // rotation about z
        r1[4] = -z; r2[4] = y;   // rotation about x
        r2[5] = -x; r0[5] = z;   // rotation about y
        for (int c = 0; c < 3; ++c)
          if ((fixm[n] >> c) & 1u)
            for (int j = 0; j < 6; ++j) B[(3 * q + c) * 6 + j] = 0.0;
      }
      qr6(B, 3 * k, &Rn[a * 36]);
      for (int64_t q = 0; q < k; ++q) {
        const int32_t n = h.child[c0 + q];
        std::memcpy(&h.P[(size_t)n * 18], &B[3 * q * 6], 18 * sizeof(double));
      }
    }
    if (sa) {
      // smoothed prolongator P = (I - omega D^-1 A) P_tent: the row of node i
      // names the aggregates of i and of its neighbours (the columns of its
      // row of A); the row of A P names the aggregates of those rows
      h.sa = true;
      h.sa_pptr.assign(nn + 1, 0);
      std::vector<int32_t> tmp;
      for (int64_t i = 0; i < nn; ++i) {
        tmp.clear();
        tmp.push_back(agg[i]);
        for (int64_t k = g.rp[i]; k < g.rp[i + 1]; ++k) tmp.push_back(agg[g.ci[k]]);
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        h.sa_pagg.insert(h.sa_pagg.end(), tmp.begin(), tmp.end());
        h.sa_pptr[i + 1] = (int32_t)h.sa_pagg.size();
      }
      h.sa_vptr.assign(nn + 1, 0);
      for (int64_t i = 0; i < nn; ++i) {
        tmp.clear();
        tmp.insert(tmp.end(), h.sa_pagg.begin() + h.sa_pptr[i], h.sa_pagg.begin() + h.sa_pptr[i + 1]);
        for (int64_t k = g.rp[i]; k < g.rp[i + 1]; ++k) {
          const int32_t j = g.ci[k];
          tmp.insert(tmp.end(), h.sa_pagg.begin() + h.sa_pptr[j], h.sa_pagg.begin() + h.sa_pptr[j + 1]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        h.sa_vagg.insert(h.sa_vagg.end(), tmp.begin(), tmp.end());
        h.sa_vptr[i + 1] = (int32_t)h.sa_vagg.size();
      }
      // supports in node order
      h.sa_sptr.assign(na + 1, 0);
      for (int32_t a : h.sa_pagg) h.sa_sptr[a + 1]++;
      for (int64_t a = 0; a < na; ++a) h.sa_sptr[a + 1] += h.sa_sptr[a];
      h.sa_snode.assign(h.sa_pagg.size(), 0);
      h.sa_sidx.assign(h.sa_pagg.size(), 0);
      {
        std::vector<int32_t> pos(h.sa_sptr.begin(), h.sa_sptr.end() - 1);
        for (int64_t i = 0; i < nn; ++i)
          for (int32_t k = h.sa_pptr[i]; k < h.sa_pptr[i + 1]; ++k) {
            const int32_t q = pos[h.sa_pagg[k]]++;
            h.sa_snode[q] = (int32_t)i;
            h.sa_sidx[q] = k;
          }
      }
      // level-1 pattern: I couples J when a node of I's support has J in its row of A P
      h.rp.assign(na + 1, 0);
      for (int64_t A = 0; A < na; ++A) {
        tmp.clear();
        for (int32_t q = h.sa_sptr[A]; q < h.sa_sptr[A + 1]; ++q) {
          const int32_t i = h.sa_snode[q];
          tmp.insert(tmp.end(), h.sa_vagg.begin() + h.sa_vptr[i], h.sa_vagg.begin() + h.sa_vptr[i + 1]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        h.ci.insert(h.ci.end(), tmp.begin(), tmp.end());
        h.rp[A + 1] = (int32_t)h.ci.size();
      }
      const int64_t nbk = h.rp[na];
      h.tr.assign(nbk, 0);
      for (int64_t A = 0; A < na; ++A)
        for (int32_t k = h.rp[A]; k < h.rp[A + 1]; ++k) h.tr[k] = find_col(h.rp, h.ci, h.ci[k], (int32_t)A);
    } else {
      // level-1 pattern from element aggregate sets
      std::vector<std::vector<int32_t>> rows(na);
      for (int64_t e = 0; e < ne; ++e)
        for (int la = 0; la < npe; ++la) {
          const int32_t A = agg[C->elem[e * npe + la]];
          for (int lb = 0; lb < npe; ++lb) rows[A].push_back(agg[C->elem[e * npe + lb]]);
        }
      h.rp.assign(na + 1, 0);
      for (int64_t A = 0; A < na; ++A) {
        auto &r = rows[A];
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        h.rp[A + 1] = h.rp[A] + (int32_t)r.size();
      }
      h.ci.reserve(h.rp[na]);
      for (int64_t A = 0; A < na; ++A) h.ci.insert(h.ci.end(), rows[A].begin(), rows[A].end());
      rows.clear();
      rows.shrink_to_fit();
      // Galerkin entries for level 1: one per (upper block (A,B), A <= B,
      // element touching both), grouped by block, in element order within a
      // block.  Entry k carries G_k = sum over the element's node pairs (la, lb)
      // with agg(la) = A, agg(lb) = B of Q_la^T KE_e[la,lb] Q_lb (density free,
      // built once on the device); a refresh is then sum_k E(x_e) G_k.
      const int64_t nbk = h.rp[na];
      h.glptr.assign(nbk + 1, 0);
      h.tr.assign(nbk, 0);
      for (int64_t A = 0; A < na; ++A)
        for (int32_t k = h.rp[A]; k < h.rp[A + 1]; ++k) h.tr[k] = find_col(h.rp, h.ci, h.ci[k], (int32_t)A);
      std::vector<int32_t> eblk;  // (element, upper block) pairs in element order
      std::vector<int32_t> eent;
      std::vector<int32_t> loc;
      for (int64_t e = 0; e < ne; ++e) {
        loc.clear();
        for (int la = 0; la < npe; ++la) {
          const int32_t A = agg[C->elem[e * npe + la]];
          for (int lb = 0; lb < npe; ++lb) {
            const int32_t Bc = agg[C->elem[e * npe + lb]];
            if (A <= Bc) loc.push_back(find_col(h.rp, h.ci, A, Bc));
          }
        }
        std::sort(loc.begin(), loc.end());
        loc.erase(std::unique(loc.begin(), loc.end()), loc.end());
        for (int32_t kb : loc) {
          eblk.push_back(kb);
          eent.push_back((int32_t)e);
          h.glptr[kb + 1]++;
        }
      }
      for (int64_t k = 0; k < nbk; ++k) h.glptr[k + 1] += h.glptr[k];
      h.gl.assign(eblk.size(), 0);
      h.gblk.assign(eblk.size(), 0);
      std::vector<int32_t> pos(h.glptr.begin(), h.glptr.end() - 1);
      for (size_t q = 0; q < eblk.size(); ++q) {
        const int32_t k = pos[eblk[q]]++;
        h.gl[k] = eent[q];
        h.gblk[k] = eblk[q];
      }
    }
    H.push_back(std::move(h));
    nunits.push_back(na);
    // ---- levels >= 1 (Rn is the coarse near-null space of level 1)
    std::vector<double> Rcur = std::move(Rn);
    int lev = 1;
    while (true) {
      const AmgHostLevel &hp = H.back();
      const int64_t nu = hp.nunit_next;
      // a batch coarsens until each design's block of the coarsest level fits
      const int64_t nu_c = nseg > 0 ? max_per_design(C->bt.uoff[lev]) : nu;
      // (a batch to at least three levels: its coarsest solve works on 6x6 blocks)
      const bool want_more =
          (6 * nu_c > cmax && lev + 1 < maxlev) || 6 * nu_c > dense_limit || (nseg > 0 && lev < 2);
      if (!want_more || lev + 1 >= UFT_MAXLEV) break;
      // unit graph from the operator pattern
      Graph gu;
      gu.rp.assign(nu + 1, 0);
      for (int64_t A = 0; A < nu; ++A) {
        for (int32_t k = hp.rp[A]; k < hp.rp[A + 1]; ++k)
          if (hp.ci[k] != A) gu.ci.push_back(hp.ci[k]);
        gu.rp[A + 1] = (int64_t)gu.ci.size();
      }
      std::vector<int32_t> ag;
      const int64_t nb2 = aggregate(gu, ag);
      if (nb2 >= nu || nb2 * 3 > nu * 2) break;  // not coarsening any more
      if (nseg > 0) {
        batch_renumber(ag, nb2, udes, nseg, ades, C->bt.uoff[lev + 1]);
        udes.swap(ades);
      }
      AmgHostLevel h2;
      h2.nunit_next = nb2;
      h2.parent = ag;
      children(ag, nb2, h2.cptr, h2.child);
      h2.P.assign(nu * 36, 0.0);
      std::vector<double> Rn2(nb2 * 36, 0.0);
      for (int64_t a = 0; a < nb2; ++a) {
        const int32_t c0 = h2.cptr[a], c1 = h2.cptr[a + 1];
        const int64_t k = c1 - c0;
        B.assign(6 * k * 6, 0.0);
        for (int64_t q = 0; q < k; ++q) std::memcpy(&B[6 * q * 6], &Rcur[(size_t)h2.child[c0 + q] * 36], 36 * sizeof(double));
        qr6(B, 6 * k, &Rn2[a * 36]);
        for (int64_t q = 0; q < k; ++q)
          std::memcpy(&h2.P[(size_t)h2.child[c0 + q] * 36], &B[6 * q * 6], 36 * sizeof(double));
      }
      // pattern of the next level
      std::vector<std::vector<int32_t>> rws(nb2);
      for (int64_t A = 0; A < nu; ++A)
        for (int32_t k = hp.rp[A]; k < hp.rp[A + 1]; ++k) rws[ag[A]].push_back(ag[hp.ci[k]]);
      h2.rp.assign(nb2 + 1, 0);
      for (int64_t A = 0; A < nb2; ++A) {
        auto &r = rws[A];
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        h2.rp[A + 1] = h2.rp[A] + (int32_t)r.size();
      }
      for (int64_t A = 0; A < nb2; ++A) h2.ci.insert(h2.ci.end(), rws[A].begin(), rws[A].end());
      // lists: level-l blocks in block order
      const int64_t nbk2 = h2.rp[nb2];
      h2.glptr.assign(nbk2 + 1, 0);
      const int64_t nbk1 = hp.rp[nu];
      std::vector<int32_t> bi(nbk1);
      for (int64_t A = 0; A < nu; ++A)
        for (int32_t k = hp.rp[A]; k < hp.rp[A + 1]; ++k) {
          const int32_t kb = find_col(h2.rp, h2.ci, ag[A], ag[hp.ci[k]]);
          bi[k] = kb;
          h2.glptr[kb + 1]++;
        }
      for (int64_t k = 0; k < nbk2; ++k) h2.glptr[k + 1] += h2.glptr[k];
      h2.gl.assign(nbk1, 0);
      std::vector<int32_t> ps(h2.glptr.begin(), h2.glptr.end() - 1);
      for (int64_t k = 0; k < nbk1; ++k) h2.gl[ps[bi[k]]++] = (int32_t)k;
      H.push_back(std::move(h2));
      nunits.push_back(nb2);
      Rcur = std::move(Rn2);
      ++lev;
    }
  }
  const int64_t nlast = nunits.back();
  if (nseg > 0) {
    if (6 * max_per_design(C->bt.uoff[nunits.size() - 1]) > UFT_TAIL_MAX)
      return uft_set_error(C, UFT_ERR_ARG, "a design's coarsest level exceeds %d unknowns", UFT_TAIL_MAX);
    if (nunits.size() < 3)
      return uft_set_error(C, UFT_ERR_ARG, "design batches need at least three multigrid levels");
  }
  if (6 * nlast > std::max<int64_t>(dense_limit * 2, 6144))
    return uft_set_error(C, UFT_ERR_ARG, "coarsest level too large for the dense solve (%lld units)", (long long)nlast);
  return UFT_OK;
}
