// Host mesh pre-processing: topology and FP64 geometry with the reference's
// conventions, then the tiled, renumbered device layout.
//
// Reference: CalculationMeshGeo (gpu_init.cu:100-228) builds edges by a
// linear search over all existing edges and cells (O(C*E + C^2)).  Here one
// hash map from the sorted node pair to the edge gives the same edges, the
// same edge numbering (first appearance scanning cells then local edges), the
// same left/right cells and the same far/wall ids in O(C).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "fvm_internal.h"

namespace fvm {

Topology build_topology(const fvm_mesh_desc& m) {
  Topology t;
  if (m.nnode < 3 || m.ncell < 1 || !m.x || !m.y || !m.tri)
    throw Error{FVM_ERR_ARGUMENT, "mesh: need nnode >= 3, ncell >= 1 and non-NULL arrays"};
  const int C = m.ncell;
  t.nnode = m.nnode;
  t.ncell = C;
  t.x.assign(m.x, m.x + m.nnode);
  t.y.assign(m.y, m.y + m.nnode);
  t.pt.resize(3 * (size_t)C);
  for (size_t i = 0; i < 3 * (size_t)C; i++) {
    int p = m.tri[i];
    if (p < 1 || p > m.nnode)
      throw Error{FVM_ERR_MESH, "mesh: cell " + std::to_string(i / 3 + 1) + " has node id " +
                                    std::to_string(p) + " outside 1.." + std::to_string(m.nnode)};
    t.pt[i] = p - 1;
  }
  // area and centroid, gpu_init.cu:95-98 and :107-116
  t.area.resize(C);
  t.cx.resize(C);
  t.cy.resize(C);
  for (int c = 0; c < C; c++) {
    const int p0 = t.pt[3 * c], p1 = t.pt[3 * c + 1], p2 = t.pt[3 * c + 2];
    const double x1 = t.x[p0], y1 = t.y[p0], x2 = t.x[p1], y2 = t.y[p1], x3 = t.x[p2], y3 = t.y[p2];
    t.area[c] = 0.5 * ((y3 - y1) * (x2 - x1) - (y2 - y1) * (x3 - x1));
    if (!(t.area[c] > 0.0))
      throw Error{FVM_ERR_AREA, "mesh: cell " + std::to_string(c + 1) + " has area " +
                                    std::to_string(t.area[c]) + " (must be counter-clockwise)"};
    t.cx[c] = (x1 + x2 + x3) / 3.0;
    t.cy[c] = (y1 + y2 + y3) / 3.0;
  }
  // Edges in order of first appearance; local edge i of a cell joins
  // Point[(i+1)%3] -> Point[(i+2)%3] (gpu_init.cu:105,127-130).
  std::unordered_map<uint64_t, int> key_to_edge;
  key_to_edge.reserve((size_t)C * 2);
  std::vector<int> e_left, e_right, e_n1, e_n2;
  e_left.reserve((size_t)C * 3 / 2 + 8);
  std::vector<int> cedge(3 * (size_t)C);
  for (int c = 0; c < C; c++) {
    for (int i = 0; i < 3; i++) {
      const int N1 = t.pt[3 * c + (i + 1) % 3], N2 = t.pt[3 * c + (i + 2) % 3];
      const uint64_t lo = (uint64_t)std::min(N1, N2), hi = (uint64_t)std::max(N1, N2);
      if (lo == hi) throw Error{FVM_ERR_MESH, "mesh: degenerate cell " + std::to_string(c + 1)};
      auto ins = key_to_edge.emplace((lo << 32) | hi, (int)e_left.size());
      int e = ins.first->second;
      if (ins.second) {
        e_left.push_back(c);
        e_right.push_back(-1);
        e_n1.push_back(N1);
        e_n2.push_back(N2);
      } else {
        if (e_right[e] != -1 || e_left[e] == c)
          throw Error{FVM_ERR_MESH, "mesh: edge (" + std::to_string(lo + 1) + "," +
                                        std::to_string(hi + 1) + ") is shared by more than two cells"};
        e_right[e] = c;
      }
      cedge[3 * c + i] = e;
    }
  }
  const int E = (int)e_left.size();
  // Boundary classification in creation order, gpu_init.cu:187-204.
  std::vector<int> elog(E), farid(E, 0), wallid(E, 0);
  std::vector<double> vx(E), vy(E), vn(E), mx(E), my(E);
  int nwall = 0, nfar = 0;
  for (int e = 0; e < E; e++) {
    const int N1 = e_n1[e], N2 = e_n2[e];
    vx[e] = t.x[N1] - t.x[N2];
    vy[e] = t.y[N2] - t.y[N1];
    mx[e] = 0.5 * (t.x[N1] + t.x[N2]);
    my[e] = 0.5 * (t.y[N1] + t.y[N2]);
    vn[e] = std::sqrt(vx[e] * vx[e] + vy[e] * vy[e]);
    if (e_right[e] >= 0) {
      elog[e] = 0;
    } else if (std::fabs(mx[e]) < 2.0 && std::fabs(my[e]) < 2.0) {
      elog[e] = 2;
      wallid[e] = ++nwall;
    } else {
      elog[e] = 1;
      farid[e] = ++nfar;
    }
  }
  // Stable interior / far / wall partition (changeEdge, gpu_solver.cu:1832-1906).
  std::vector<int> order;
  order.reserve(E);
  for (int pass = 0; pass < 3; pass++)
    for (int e = 0; e < E; e++)
      if (elog[e] == pass) order.push_back(e);
  std::vector<int> newid(E);
  for (int k = 0; k < E; k++) newid[order[k]] = k;
  t.nedge = E;
  t.nwall = nwall;
  t.nfar = nfar;
  auto permute_i = [&](const std::vector<int>& a) {
    std::vector<int> r(E);
    for (int k = 0; k < E; k++) r[k] = a[order[k]];
    return r;
  };
  auto permute_d = [&](const std::vector<double>& a) {
    std::vector<double> r(E);
    for (int k = 0; k < E; k++) r[k] = a[order[k]];
    return r;
  };
  t.elog = permute_i(elog);
  t.left = permute_i(e_left);
  t.right = permute_i(e_right);
  t.farid = permute_i(farid);
  t.wallid = permute_i(wallid);
  t.node1 = permute_i(e_n1);
  t.node2 = permute_i(e_n2);
  t.vx = permute_d(vx);
  t.vy = permute_d(vy);
  t.vn = permute_d(vn);
  t.midx = permute_d(mx);
  t.midy = permute_d(my);
  t.celledge.resize(3 * (size_t)C);
  for (size_t i = 0; i < 3 * (size_t)C; i++) t.celledge[i] = newid[cedge[i]];
  t.wall_edge.assign(nwall, -1);
  for (int e = 0; e < E; e++)
    if (t.elog[e] == 2) t.wall_edge[t.wallid[e] - 1] = e;
  // neighbor[] and CLog, gpu_init.cu:209-222
  t.neicell.resize(3 * (size_t)C);
  t.clog.resize(C);
  for (int c = 0; c < C; c++) {
    int nn = 3;
    for (int i = 0; i < 3; i++) {
      const int e = t.celledge[3 * c + i];
      int nc = (t.left[e] == c) ? t.right[e] : t.left[e];
      if (nc == -1) {
        nc = -t.elog[e];
        nn--;
      }
      t.neicell[3 * c + i] = nc;
    }
    t.clog[c] = nn;
  }
  return t;
}

// Recursive coordinate bisection: splits ids[b,e) at the median of the longer
// centroid extent into parts holding whole numbers of tiles, and orders the
// leaves so that consecutive ids are spatially close.
static void rcb(std::vector<int>& ids, size_t b, size_t e, int ntile, const Topology& t,
                std::vector<std::pair<size_t, size_t>>& tiles) {
  const size_t n = e - b;
  if (ntile <= 1) {
    tiles.emplace_back(b, e);
    // order inside the tile by further bisection down to small groups
    if (n > 16) {
      std::vector<std::pair<size_t, size_t>> sub;
      rcb(ids, b, e, (int)((n + 15) / 16), t, sub);
    }
    return;
  }
  double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
  for (size_t i = b; i < e; i++) {
    const int c = ids[i];
    x0 = std::min(x0, t.cx[c]);
    x1 = std::max(x1, t.cx[c]);
    y0 = std::min(y0, t.cy[c]);
    y1 = std::max(y1, t.cy[c]);
  }
  const bool ax = (x1 - x0) >= (y1 - y0);
  const int nl = ntile / 2;
  const size_t m = b + (size_t)((double)n * nl / ntile + 0.5);
  auto key = [&](int c) { return ax ? t.cx[c] : t.cy[c]; };
  std::nth_element(ids.begin() + b, ids.begin() + m, ids.begin() + e, [&](int p, int q) {
    const double kp = key(p), kq = key(q);
    return kp < kq || (kp == kq && p < q);
  });
  rcb(ids, b, m, nl, t, tiles);
  rcb(ids, m, e, ntile - nl, t, tiles);
}

TiledMesh build_tiled(const Topology& t, int tile_cells) {
  TiledMesh T;
  const int C = t.ncell;
  if (tile_cells < 8) tile_cells = 8;
  T.ncell = C;
  T.tile_cells = tile_cells;
  const int ntile = (C + tile_cells - 1) / tile_cells;
  std::vector<int> ids(C);
  std::iota(ids.begin(), ids.end(), 0);
  std::vector<std::pair<size_t, size_t>> tiles;
  rcb(ids, 0, C, ntile, t, tiles);
  T.ntile = ntile;
  T.new_to_old = ids;
  T.old_to_new.assign(C, -1);
  for (int i = 0; i < C; i++) T.old_to_new[ids[i]] = i;
  T.tile_cell0.resize(ntile + 1);
  for (int k = 0; k < ntile; k++) T.tile_cell0[k] = (int)tiles[k].first;
  T.tile_cell0[ntile] = C;
  std::vector<int> tile_of(C);
  for (int k = 0; k < ntile; k++)
    for (int i = T.tile_cell0[k]; i < T.tile_cell0[k + 1]; i++) tile_of[i] = k;

  // per-cell arrays in new numbering
  T.gvx.resize(3 * (size_t)C);
  T.gvy.resize(3 * (size_t)C);
  T.inv_area.resize(C);
  T.area.resize(C);
  T.nbr.resize(3 * (size_t)C);
  T.clog.resize(C);
  T.slot.assign(3 * (size_t)C, 0);
  T.nloc.assign(3 * (size_t)C, -1);
  for (int n = 0; n < C; n++) {
    const int o = ids[n];
    for (int i = 0; i < 3; i++) {
      // the gradient's per-cell edge vector, gpu_solver.cu:261-265
      const int N1 = t.pt[3 * o + (i + 1) % 3], N2 = t.pt[3 * o + (i + 2) % 3];
      T.gvx[3 * n + i] = t.x[N1] - t.x[N2];
      T.gvy[3 * n + i] = t.y[N2] - t.y[N1];
      const int nc = t.neicell[3 * o + i];
      T.nbr[3 * n + i] = nc >= 0 ? T.old_to_new[nc] : nc;
    }
    T.area[n] = t.area[o];
    T.inv_area[n] = 1.0 / t.area[o];
    T.clog[n] = t.clog[o];
  }

  // tiles: halo ring, edge records, gather slots
  T.tile_halo0.resize(ntile + 1);
  T.tile_edge0.resize(ntile + 1);
  std::vector<int> local_of(C, -1);
  std::vector<int> edge_local(t.nedge, -1);
  for (int k = 0; k < ntile; k++) {
    const int c0 = T.tile_cell0[k], c1 = T.tile_cell0[k + 1], nt = c1 - c0;
    T.tile_halo0[k] = (int)T.halo.size();
    T.tile_edge0[k] = (int)T.e_l.size();
    for (int i = c0; i < c1; i++) local_of[i] = i - c0;
    std::vector<int> h;
    for (int i = c0; i < c1; i++)
      for (int j = 0; j < 3; j++) {
        const int nb = T.nbr[3 * i + j];
        if (nb >= 0 && tile_of[nb] != k) h.push_back(nb);
      }
    std::sort(h.begin(), h.end());
    h.erase(std::unique(h.begin(), h.end()), h.end());
    for (size_t q = 0; q < h.size(); q++) local_of[h[q]] = nt + (int)q;
    T.halo.insert(T.halo.end(), h.begin(), h.end());
    std::vector<int> tile_edges, ekl, ekr;
    for (int i = c0; i < c1; i++) {
      const int o = ids[i];
      for (int j = 0; j < 3; j++) {
        const int e = t.celledge[3 * o + j];
        if (edge_local[e] < 0) {
          edge_local[e] = (int)tile_edges.size();
          tile_edges.push_back(e);
          ekl.push_back(0);
          ekr.push_back(0);
        }
        const int side = (t.left[e] == o) ? 0 : 1;
        T.slot[3 * i + j] = 2 * edge_local[e] + side;
        // the edge's local index in this tile cell, on its side
        (side == 0 ? ekl : ekr)[edge_local[e]] = j;
        const int nb = T.nbr[3 * i + j];
        T.nloc[3 * i + j] = nb >= 0 ? local_of[nb] : -1;
      }
    }
    // the edge's local index in a ring-1 cell on its other side too (FP32
    // addresses the face slots of both sides through it), then both checked
    for (size_t q = 0; q < tile_edges.size(); q++) {
      const int e = tile_edges[q];
      for (int side = 0; side < 2; side++) {
        const int o = side == 0 ? t.left[e] : t.right[e];
        if (o < 0 || tile_of[T.old_to_new[o]] == k) continue;
        for (int j = 0; j < 3; j++)
          if (t.celledge[3 * o + j] == e) (side == 0 ? ekl : ekr)[q] = j;
      }
      if (t.celledge[3 * t.left[e] + ekl[q]] != e ||
          (t.right[e] >= 0 && t.celledge[3 * t.right[e] + ekr[q]] != e))
        throw Error{FVM_ERR_MESH, "internal: edge local index check failed"};
    }
    for (size_t q = 0; q < tile_edges.size(); q++) {
      const int e = tile_edges[q];
      const int L = T.old_to_new[t.left[e]];
      T.e_l.push_back(local_of[L]);
      T.e_kl.push_back(ekl[q]);
      T.e_kr.push_back(ekr[q]);
      if (t.right[e] >= 0) {
        T.e_r.push_back(local_of[T.old_to_new[t.right[e]]]);
      } else {
        T.e_r.push_back(-t.elog[e]);
      }
      // n = (vy, vx)/vn as the flux call forms it, gpu_solver.cu:727; the
      // reconstruction offsets mid - centroid, gpu_solver.cu:514-515,535-536
      T.e_nx.push_back(t.vy[e] / t.vn[e]);
      T.e_ny.push_back(t.vx[e] / t.vn[e]);
      T.e_vn.push_back(t.vn[e]);
      T.e_vx.push_back(t.vx[e]);
      T.e_vy.push_back(t.vy[e]);
      const int lo = t.left[e];
      T.e_rlx.push_back(t.midx[e] - t.cx[lo]);
      T.e_rly.push_back(t.midy[e] - t.cy[lo]);
      if (t.right[e] >= 0) {
        const int ro = t.right[e];
        T.e_rrx.push_back(t.midx[e] - t.cx[ro]);
        T.e_rry.push_back(t.midy[e] - t.cy[ro]);
      } else {
        T.e_rrx.push_back(0.0);
        T.e_rry.push_back(0.0);
      }
    }
    T.max_local = std::max(T.max_local, nt + (int)h.size());
    T.max_tile = std::max(T.max_tile, nt);
    T.max_edges = std::max(T.max_edges, (int)tile_edges.size());
    for (int e : tile_edges) edge_local[e] = -1;
    for (int i = c0; i < c1; i++) local_of[i] = -1;
    for (int q : h) local_of[q] = -1;
  }
  T.tile_halo0[ntile] = (int)T.halo.size();
  T.tile_edge0[ntile] = (int)T.e_l.size();

  // rings 2..4 and the tile-local neighbour table for tile + rings 1..3
  T.tile_ring0.resize(ntile + 1);
  T.tile_lnb0.resize(ntile + 1);
  T.ring_n1.resize(ntile);
  T.ring_n2.resize(ntile);
  T.ring_n3.resize(ntile);
  T.ring_n4.resize(ntile);
  std::vector<int> mark(C, -1);
  for (int k = 0; k < ntile; k++) {
    const int c0 = T.tile_cell0[k], c1 = T.tile_cell0[k + 1], nt = c1 - c0;
    T.tile_ring0[k] = (int)T.rings.size();
    T.tile_lnb0[k] = (int)T.lnb.size() / 3;
    std::vector<int> local;  // tile-local index -> new id
    for (int i = c0; i < c1; i++) {
      mark[i] = (int)local.size();
      local.push_back(i);
    }
    const int h0 = T.tile_halo0[k], h1 = T.tile_halo0[k + 1];
    for (int q = h0; q < h1; q++) {
      mark[T.halo[q]] = (int)local.size();
      local.push_back(T.halo[q]);
    }
    int begin = nt, end = (int)local.size();
    int ring_sizes[4] = {end - begin, 0, 0, 0};
    for (int r = 1; r <= 3; r++) {  // build rings 2..4 from rings 1..3
      std::vector<int> nxt;
      for (int j = begin; j < end; j++)
        for (int q = 0; q < 3; q++) {
          const int nb = T.nbr[3 * local[j] + q];
          if (nb >= 0 && mark[nb] < 0) nxt.push_back(nb);
        }
      std::sort(nxt.begin(), nxt.end());
      nxt.erase(std::unique(nxt.begin(), nxt.end()), nxt.end());
      for (int id : nxt) {
        mark[id] = (int)local.size();
        local.push_back(id);
      }
      ring_sizes[r] = (int)nxt.size();
      begin = end;
      end = (int)local.size();
    }
    T.ring_n1[k] = ring_sizes[0];
    T.ring_n2[k] = ring_sizes[1];
    T.ring_n3[k] = ring_sizes[2];
    T.ring_n4[k] = ring_sizes[3];
    T.rings.insert(T.rings.end(), local.begin() + nt, local.end());
    const int nw = nt + ring_sizes[0] + ring_sizes[1];
    const int nr = nw + ring_sizes[2];
    for (int j = 0; j < nr; j++)
      for (int q = 0; q < 3; q++) {
        const int nb = T.nbr[3 * local[j] + q];
        T.lnb.push_back(nb >= 0 ? mark[nb] : nb);
      }
    T.max_r4_local = std::max(T.max_r4_local, (int)local.size());
    T.max_ring_local = std::max(T.max_ring_local, nr);
    T.max_w_local = std::max(T.max_w_local, nw);
    // every interior neighbour of tile + rings 1..3 is local, those of
    // tile + rings 1..2 (smoothing sweep 2) lie in tile + rings 1..3 and those
    // of tile + ring 1 (the gradient cells) in tile + rings 1..2; check it
    const int ng = nt + ring_sizes[0];
    for (int j = 0; j < nr; j++)
      for (int q = 0; q < 3; q++) {
        const int v = T.lnb[(size_t)3 * (T.tile_lnb0[k] + j) + q];
        if (v < -2 || v >= (int)local.size() || (T.nbr[3 * local[j] + q] >= 0 && v < 0) ||
            (j < nw && v >= nr) || (j < ng && v >= nw) || (j < nt && v >= ng))
          throw Error{FVM_ERR_MESH, "internal: ring table check failed at tile " + std::to_string(k)};
      }
    for (int id : local) mark[id] = -1;
  }
  T.tile_ring0[ntile] = (int)T.rings.size();
  T.tile_lnb0[ntile] = (int)T.lnb.size() / 3;

  // Bounds check of every index the kernels dereference (global ids, LDS
  // slots), so a layout error surfaces here and never as a device fault.
  for (int k = 0; k < ntile; k++) {
    const int c0 = T.tile_cell0[k], c1 = T.tile_cell0[k + 1], nt = c1 - c0;
    const int nh = T.tile_halo0[k + 1] - T.tile_halo0[k];
    const int ne = T.tile_edge0[k + 1] - T.tile_edge0[k];
    const int nl = nt + nh;
    bool ok = nt > 0 && nl <= T.max_local && ne <= T.max_edges;
    for (int q = T.tile_halo0[k]; ok && q < T.tile_halo0[k + 1]; q++)
      ok = T.halo[q] >= 0 && T.halo[q] < C && (T.halo[q] < c0 || T.halo[q] >= c1);
    for (int s = T.tile_edge0[k]; ok && s < T.tile_edge0[k + 1]; s++)
      ok = T.e_l[s] >= 0 && T.e_l[s] < nl && T.e_r[s] >= -2 && T.e_r[s] < nl &&
           (T.e_r[s] < 0 || T.e_r[s] != T.e_l[s]);
    for (int i = c0; ok && i < c1; i++)
      for (int j = 0; ok && j < 3; j++) {
        const int sl = T.slot[3 * i + j], nb = T.nbr[3 * i + j], lo = T.nloc[3 * i + j];
        ok = sl >= 0 && (sl >> 1) < ne && nb >= -2 && nb < C && lo >= -1 && lo < nl &&
             ((nb >= 0) == (lo >= 0));
        // the tile edge in that slot names this cell and this local index on
        // that side, so the kernels can rebuild the slot from the edge record
        if (ok) {
          const int q = T.tile_edge0[k] + (sl >> 1);
          ok = (sl & 1) == 0 ? (T.e_l[q] == i - c0 && T.e_kl[q] == j)
                             : (T.e_r[q] == i - c0 && T.e_kr[q] == j);
        }
      }
    if (!ok) throw Error{FVM_ERR_MESH, "internal: tile " + std::to_string(k) + " index check failed"};
  }
  for (int i = 0; i < 3 * C; i++)
    if (T.nbr[i] < -2 || T.nbr[i] >= C)
      throw Error{FVM_ERR_MESH, "internal: neighbour index check failed"};

  for (int w = 0; w < t.nwall; w++) {
    const int e = t.wall_edge[w];
    T.w_cell.push_back(T.old_to_new[t.left[e]]);
    T.w_vx.push_back(t.vx[e]);
    T.w_vy.push_back(t.vy[e]);
    T.w_vn.push_back(t.vn[e]);
  }
  return T;
}

}  // namespace fvm
