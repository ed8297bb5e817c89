// pairs.hip -- centres, linked-cell lists and the correction pair lists on gfx950.
//
//   k_group_centres / k_mol_com   appbou.f group centres, molecule centres of mass
//   k_cell_*                      linked-cell binning of group centres (lc_list.f:93-109)
//   k_list_count / k_list_offsets / k_list_write
//                                 stencil walk and pair test (lc_list.f:133-201), two
//                                 passes so the output layout is exact and compact
//   k_h2f_*                       transpose of a half list into a full list, each
//                                 group's transposed tail sorted so the order is fixed
//   k_inner_list                  next-inner-shell list of force.f:259-286
//   k_corr_*                      ferrf.f / fnbgrp.f pair lists: per-pair kernel plus a
//                                 per-atom gather in fixed order
//   k_cself                       cself.f self term
// The pair force kernel itself is in force_pairs.hip.  Every batched kernel
// takes its system from blockIdx.y (blockIdx.x for one-block-per-system kernels).
// The list and classification arithmetic is compiled without FMA contraction so
// the distances are the reference's bit for bit.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace orac {

// A&S 7.1.26 (force.f:123-125)
#define AS_A1 0.2548296
#define AS_A2 -0.28449674
#define AS_A3 1.4214137
#define AS_A4 -1.453152
#define AS_A5 1.0614054
#define AS_QP 0.3275911

__device__ inline int fmodp(int a, int n) { return ((a % n) + n) % n; }

// ---------------------------------------------------------------- centres

// group centres (appbou.f): mass-fraction weighted sum, no contraction
__global__ void __launch_bounds__(256) k_group_centres(const PairDesc *__restrict__ D) {
#pragma clang fp contract(off)
  const PairDesc &d = D[blockIdx.y];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= d.ngrp) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int m = d.gfirst[g]; m <= d.glast[g]; ++m) {
    sx = sx + d.gmass[m] * d.xp[m];
    sy = sy + d.gmass[m] * d.yp[m];
    sz = sz + d.gmass[m] * d.zp[m];
  }
  d.xpg[g] = sx;
  d.ypg[g] = sy;
  d.zpg[g] = sz;
}

// Molecule centres of mass.  Blocks [0, nbb) give each large molecule (more
// than MOL_SMALL atoms) a wavefront: lanes stride over its atoms, fixed-order
// butterfly sum.  The blocks after them give every small molecule (a water) a
// thread that sums its atoms in order.  Each molecule has one fixed path.
__global__ void __launch_bounds__(256) k_mol_com(const PairDesc *__restrict__ D, int nbb) {
  const PairDesc &d = D[blockIdx.y];
  if ((int)blockIdx.x < nbb) {
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (k >= d.nbig) return;
    const int m = gmem(d.mbig)[k];
    double sx = 0.0, sy = 0.0, sz = 0.0, sm = 0.0;
    for (int a = gmem(d.mfirst)[m] + lane; a <= gmem(d.mlast)[m]; a += 64) {
      const double w = gmem(d.mass)[a];
      sx += w * gmem(d.xp)[a];
      sy += w * gmem(d.yp)[a];
      sz += w * gmem(d.zp)[a];
      sm += w;
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    sm = wave_sum(sm);
    if (lane == 0) {
      gmem(d.xpcm)[m] = sx / sm;
      gmem(d.ypcm)[m] = sy / sm;
      gmem(d.zpcm)[m] = sz / sm;
    }
    return;
  }
  const int m = (blockIdx.x - nbb) * blockDim.x + threadIdx.x;
  if (m >= d.nmol) return;
  const int a0 = gmem(d.mfirst)[m], a1 = gmem(d.mlast)[m];
  if (a1 - a0 + 1 > MOL_SMALL) return;
  double sx = 0.0, sy = 0.0, sz = 0.0, sm = 0.0;
  for (int a = a0; a <= a1; ++a) {
    const double w = gmem(d.mass)[a];
    sx += w * gmem(d.xp)[a];
    sy += w * gmem(d.yp)[a];
    sz += w * gmem(d.zp)[a];
    sm += w;
  }
  gmem(d.xpcm)[m] = sx / sm;
  gmem(d.ypcm)[m] = sy / sm;
  gmem(d.zpcm)[m] = sz / sm;
}

void launch_centres(const PairDesc *d_desc, const PairDesc *h, int nsys, hipStream_t st) {
  int max_mol = 0, max_grp = 0, max_big = 0;
  for (int s = 0; s < nsys; ++s) {
    max_mol = std::max(max_mol, h[s].mfirst ? h[s].nmol : 0);
    max_big = std::max(max_big, h[s].mfirst ? h[s].nbig : 0);
    max_grp = std::max(max_grp, h[s].ngrp);
  }
  if (max_mol > 0) {
    const int nbb = (max_big + 3) / 4;
    hipLaunchKernelGGL(k_mol_com, dim3(nbb + (max_mol + 255) / 256, nsys), dim3(256), 0, st, d_desc, nbb);
  }
  if (max_grp > 0) hipLaunchKernelGGL(k_group_centres, dim3((max_grp + 255) / 256, nsys), dim3(256), 0, st, d_desc);
}

// ---------------------------------------------------------------- cells

// cell index along one axis (lc_list.f:94-103): floor toward -inf then wrap.
// A non-finite or absurdly large coordinate goes to cell 0 instead of an
// undefined integer conversion, so no index leaves the grid.
__device__ inline int cell_axis(double xr, double dcell, int nc) {
#pragma clang fp contract(off)
  double x1 = xr / dcell;
  if (!(fabs(x1) < 1.0e9)) return 0;
  int ix = (int)x1;
  double sg = (x1 - (double)ix) >= 0.0 ? 1.0 : -1.0;
  int n = (int)((double)ix + (sg - 1.0) / 2.0);
  return fmodp(n, nc);
}

// the build of a system with a step counter runs only on its update steps
__device__ inline bool list_due(const ListDesc &d) { return !d.step || (*d.step % d.every) == 0; }

__global__ void __launch_bounds__(256) k_list_zero(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.y];
  if (!list_due(d)) return;
  const int ncell = d.cg.ncx * d.cg.ncy * d.cg.ncz + 1;  // + the parking cell
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ncell; i += gridDim.x * blockDim.x) {
    d.cell_cnt[i] = 0;
    d.cell_cursor[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < 2) d.total[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(256) k_cell_assign(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.y];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= d.ngrp || !list_due(d)) return;
  const CellGrid &cg = d.cg;
  const double x = d.xpg[g], y = d.ypg[g], z = d.zpg[g];
  int c;
  if (fabs(x) < 1.0e8 && fabs(y) < 1.0e8 && fabs(z) < 1.0e8) {
    int nx = cell_axis(x, 2.0 / cg.ncx, cg.ncx);
    int ny = cell_axis(y, 2.0 / cg.ncy, cg.ncy);
    int nz = cell_axis(z, 2.0 / cg.ncz, cg.ncz);
    c = nz + cg.ncz * (ny + cg.ncy * nx);
  } else {
    // non-finite (or absurd) centre: the parking cell after the grid, which no
    // stencil visits; the group gets no neighbours and is nobody's neighbour
    c = cg.ncx * cg.ncy * cg.ncz;
  }
  d.cell_of[g] = c;
  atomicAdd(&d.cell_cnt[c], 1);
}

__global__ void __launch_bounds__(1024) k_cell_scan(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.x];
  if (!list_due(d)) return;
  block_exclusive_scan(d.cell_cnt, d.cell_start, d.cg.ncx * d.cg.ncy * d.cg.ncz + 1);
}

__global__ void __launch_bounds__(256) k_cell_fill(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.y];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= d.ngrp || !list_due(d)) return;
  const int c = d.cell_of[g];
  const int pos = atomicAdd(&d.cell_cursor[c], 1);
  d.cell_fill[d.cell_start[c] + pos] = g;
}

// each cell's members in descending group order: the order of the
// reference's head-inserted chains (lc_list.f:104-108).  One wavefront per
// cell ranks its members (group indices are distinct): a member's place is
// the number of members above it, so the cost per cell is n^2 / 64 compares
// and one cell holding every group (the single-cell fallback of a small box)
// is as cheap as many small ones.
__global__ void __launch_bounds__(256) k_cell_sort(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.y];
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (c > d.cg.ncx * d.cg.ncy * d.cg.ncz || !list_due(d)) return;
  const int s = gmem(d.cell_start)[c], n = gmem(d.cell_cnt)[c];
  const auto in = gmem(d.cell_fill + s);
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    const int v = i < n ? in[i] : -1;
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += in[j] > v ? 1 : 0;
    if (i < n) d.cell_items[s + rank] = v;
  }
}

// centres copied into cell order, with the group index, so a stencil walk
// reads each cell's candidates from contiguous memory
__global__ void __launch_bounds__(256) k_cell_pack(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.y];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.ngrp || !list_due(d)) return;
  const int g = d.cell_items[t];
  d.cg.cpos[t] = make_double4(gmem(d.xpg)[g], gmem(d.ypg)[g], gmem(d.zpg)[g], (double)g);
  if (d.soff && d.cellorder) d.soff[t] = gmem(d.gsize)[g];  // group sizes in cell order; scanned in place below
}

// first atom of each group in cell order (exclusive prefix of the sizes), one block per system
__global__ void __launch_bounds__(1024) k_soff_scan(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.x];
  if (!d.soff || !d.cellorder || !list_due(d)) return;
  block_exclusive_scan(d.soff, d.soff, d.ngrp);
}

__global__ void k_zero_i32(int *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// single-block exclusive scan (fixed order); total written to *total
__global__ void __launch_bounds__(1024) k_scan(const int *in, int *out, int n, int *total) {
  const int t = block_exclusive_scan(in, out, n);
  if (threadIdx.x == 0 && total) *total = t;
}

void launch_exclusive_scan(const int *in, int *out, int n, int *total, hipStream_t st) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, in, out, n, total);
}

// Pair test of lc_list.f:159-170 (full co, minimum image), no contraction.
// p = {xpg, ypg, zpg, group} of the candidate, from the cell-ordered copy.
__device__ inline bool list_hit(double x1, double y1, double z1, double4 p, const double *co, double sqcut) {
#pragma clang fp contract(off)
  double x2 = x1 - p.x, y2 = y1 - p.y, z2 = z1 - p.z;
  x2 = x2 - 2.0 * fnint(0.5 * x2);
  y2 = y2 - 2.0 * fnint(0.5 * y2);
  z2 = z2 - 2.0 * fnint(0.5 * z2);
  double xx = co[0] * x2 + co[3] * y2 + co[6] * z2;
  double yy = co[1] * x2 + co[4] * y2 + co[7] * z2;
  double zz = co[2] * x2 + co[5] * y2 + co[8] * z2;
  return xx * xx + yy * yy + zz * zz < sqcut;
}

__device__ inline int wave_sum_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline int wave_excl_scan_int(int v, int lane) {
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  return x - v;
}

// Stencil walk for group n (lc_list.f:133-201) by one wavefront: the lanes
// test 64 candidate groups at a time, in stencil order and then member order;
// write positions come from a wave prefix sum of the hits, so the output order
// is the reference's: stencil order, then descending group index inside a cell.  half: only l > n (the reference list); else every l != n.
// Entries: group number + 1 (entry 0), cell-order position (1), group number (2).
template <bool WRITE>
__device__ inline int list_walk_wave(int n, const ListDesc &d, int *out, int entry, int cap = 0x7fffffff) {
  const CellGrid &cg = d.cg;
  const int lane = threadIdx.x & 63;
  const double x1 = gmem(d.xpg)[n], y1 = gmem(d.ypg)[n], z1 = gmem(d.zpg)[n];
  const int c = gmem(d.cell_of)[n];
  if (c >= cg.ncx * cg.ncy * cg.ncz) return 0;  // parked (non-finite centre)
  const int i = c / (cg.ncy * cg.ncz), j = (c / cg.ncz) % cg.ncy, k = c % cg.ncz;
  int base = 0;
  if (cg.nsten < 64) {
    // few, crowded stencil cells (small grids, the single-cell fallback of a
    // small box): lanes take the members of one cell at a time, same order
    for (int m = 0; m < cg.nsten; ++m) {
      const int nx = fmodp(i + gmem(cg.sten)[3 * m], cg.ncx);
      const int ny = fmodp(j + gmem(cg.sten)[3 * m + 1], cg.ncy);
      const int nz = fmodp(k + gmem(cg.sten)[3 * m + 2], cg.ncz);
      const int cc = nz + cg.ncz * (ny + cg.ncy * nx);
      const int s = gmem(d.cell_start)[cc], e = s + gmem(d.cell_cnt)[cc];
      for (int t0 = s; t0 < e; t0 += 64) {
        const int t = t0 + lane;
        bool hit = false;
        int l = 0;
        if (t < e) {
          const double4 p = gload(cg.cpos + t);
          l = (int)p.w;
          hit = !(cg.half ? l <= n : l == n) && list_hit(x1, y1, z1, p, cg.co, cg.sqcut);
        }
        const unsigned long long mask = __ballot(hit);
        if (WRITE && hit) {
          const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
          if (pos < cap) out[pos] = entry == 1 ? t : (entry == 0 ? l + 1 : l);
        }
        base += __popcll(mask);
      }
    }
    return base;
  }
  // Many stencil cells: lane m reads the bounds of cell m of each chunk of
  // 64, then the chunk's members are tested 64 at a time as one flat sequence
  // (stencil order, then member order: the order above), so each round issues
  // one position load per lane in parallel rather than each lane walking its
  // own cell's members one dependent load after another.  Candidate q finds
  // its cell by binary search over the chunk's member offsets (an exclusive
  // prefix sum, read across lanes): the last cell whose offset is <= q, which
  // holds q since q < the chunk's total.
  for (int m0 = 0; m0 < cg.nsten; m0 += 64) {
    const int m = m0 + lane;
    int s = 0, cnt = 0;
    if (m < cg.nsten) {
      int nx = fmodp(i + gmem(cg.sten)[3 * m], cg.ncx);
      int ny = fmodp(j + gmem(cg.sten)[3 * m + 1], cg.ncy);
      int nz = fmodp(k + gmem(cg.sten)[3 * m + 2], cg.ncz);
      int cc = nz + cg.ncz * (ny + cg.ncy * nx);
      s = gmem(d.cell_start)[cc];
      cnt = gmem(d.cell_cnt)[cc];
    }
    const int ofs = wave_excl_scan_int(cnt, lane);
    const int total = __shfl(ofs + cnt, 63, 64);
    for (int q0 = 0; q0 < total; q0 += 64) {
      const int q = q0 + lane;
      int c = 0;  // c + step <= 63 at every step
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (__shfl(ofs, c + step, 64) <= q) c += step;
      const int t = __shfl(s, c, 64) + (q - __shfl(ofs, c, 64));
      bool hit = false;
      int l = 0;
      if (q < total) {
        const double4 p = gload(cg.cpos + t);
        l = (int)p.w;
        hit = !(cg.half ? l <= n : l == n) && list_hit(x1, y1, z1, p, cg.co, cg.sqcut);
      }
      const unsigned long long mask = __ballot(hit);
      if (WRITE && hit) {
        const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
        if (pos < cap) out[pos] = entry == 1 ? t : (entry == 0 ? l + 1 : l);
      }
      base += __popcll(mask);
    }
  }
  return base;
}

// owners: group order, or cell order in mode 0 with cellorder (owner o is group cell_items[o])
__global__ void __launch_bounds__(256) k_list_count(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.y];
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (o >= d.ngrp || !list_due(d)) return;
  const int n = (d.mode == 0 && d.cellorder) ? d.cell_items[o] : o;
  const int c = list_walk_wave<false>(n, d, nullptr, 0);
  if ((threadIdx.x & 63) == 0) d.count[o] = c;
}

// one block per system: offsets, total and the capacity check.  On overflow
// every offset is zero and total[1] = 1, so the write pass writes nothing and
// a pair kernel reading the list sees empty neighbour lists.
__global__ void __launch_bounds__(1024) k_list_offsets(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.x];
  if (!list_due(d)) return;
  const int n = d.ngrp;
  const int tot = block_exclusive_scan(d.count, d.off, n);
  const long long need = (long long)tot + (d.mode == 1 ? n : 0);
  const bool over = need > d.cap;
  if (over)
    for (int i = threadIdx.x; i < n; i += blockDim.x) d.off[i] = 0;
  if (threadIdx.x == 0) {
    d.off[n] = over ? 0 : tot;
    d.total[0] = tot;
    d.total[1] = over ? 1 : 0;
    if (over) d.total[3] = 1;  // sticky: cleared only by the host
  }
}

// One instantiation per layout (a batch uses one), so each holds one walk:
// LW_REF the reference's count-prefixed layout, LW_SLAB one slab per owner,
// LW_CSR offsets from the count pass.
enum { LW_REF = 0, LW_SLAB = 1, LW_CSR = 2 };
template <int LAYOUT>
__global__ void __launch_bounds__(256) k_list_write(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.y];
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (o >= d.ngrp || d.total[1] || !list_due(d)) return;
  if (LAYOUT == LW_REF) {
    const int w = d.off[o] + o;  // one count slot per preceding group
    const int nv = list_walk_wave<true>(o, d, d.out + w + 1, 0);
    if ((threadIdx.x & 63) == 0) d.out[w] = nv;
  } else if (LAYOUT == LW_SLAB) {
    if (d.split) return;  // k_list_write_split
    const int n = d.cellorder ? d.cell_items[o] : o;
    const int c = list_walk_wave<true>(n, d, d.out + (size_t)o * d.slab, d.cellorder ? 1 : 2, d.slab);
    if ((threadIdx.x & 63) == 0) d.count[o] = c;
  } else {
    const int n = d.cellorder ? d.cell_items[o] : o;
    list_walk_wave<true>(n, d, d.out + d.off[o], d.cellorder ? 1 : 2);
  }
}

// ---------------------------------------------------------------- split slab layout
//
// The stencil walk of a full list, written as two lists per owner n: the
// forward entries (group l > n) to out[o * slab] in walk order, and the
// reverse entries (l < n) to rout[o * slab] as group numbers, ascending.  The pair
// kernel's half-list form evaluates each group pair once, from the forward
// list of its lower group; the forces on the upper group's atoms go to
// per-pair slots that the upper group sums in the order of its sorted
// reverse list (mirror[] gives each forward entry its slot).
__device__ inline void split_put(int *outf, int *lrev, int cap, bool fwd, int posf, int posr, int v) {
  if (fwd) {
    if (posf < cap) outf[posf] = v;
  } else {
    if (posr < cap) lrev[posr] = v;
  }
}

// returns {forward, reverse} counts; lrev (LDS) receives the reverse entries in walk order
__device__ inline int2 list_walk_split(int n, const ListDesc &d, int *outf, int *lrev, int cap, int entry) {
  const CellGrid &cg = d.cg;
  const int lane = threadIdx.x & 63;
  const double x1 = gmem(d.xpg)[n], y1 = gmem(d.ypg)[n], z1 = gmem(d.zpg)[n];
  const int c = gmem(d.cell_of)[n];
  if (c >= cg.ncx * cg.ncy * cg.ncz) return make_int2(0, 0);  // parked (non-finite centre)
  const int i = c / (cg.ncy * cg.ncz), j = (c / cg.ncz) % cg.ncy, k = c % cg.ncz;
  int basef = 0, baser = 0;
  if (cg.nsten < 64) {
    for (int m = 0; m < cg.nsten; ++m) {
      const int nx = fmodp(i + gmem(cg.sten)[3 * m], cg.ncx);
      const int ny = fmodp(j + gmem(cg.sten)[3 * m + 1], cg.ncy);
      const int nz = fmodp(k + gmem(cg.sten)[3 * m + 2], cg.ncz);
      const int cc = nz + cg.ncz * (ny + cg.ncy * nx);
      const int s = gmem(d.cell_start)[cc], e = s + gmem(d.cell_cnt)[cc];
      for (int t0 = s; t0 < e; t0 += 64) {
        const int t = t0 + lane;
        bool hit = false;
        int l = 0;
        if (t < e) {
          const double4 p = gload(cg.cpos + t);
          l = (int)p.w;
          hit = l != n && list_hit(x1, y1, z1, p, cg.co, cg.sqcut);
        }
        const bool f = hit && l > n, r = hit && l < n;
        const unsigned long long mf = __ballot(f), mr = __ballot(r);
        const unsigned long long below = (1ull << lane) - 1ull;
        // only hits are written (a lane past the cell, t >= e, holds no entry)
        if (f || r)
          split_put(outf, lrev, cap, f, basef + __popcll(mf & below), baser + __popcll(mr & below),
                          entry == 1 ? t : l);
        basef += __popcll(mf);
        baser += __popcll(mr);
      }
    }
    return make_int2(basef, baser);
  }
  for (int m0 = 0; m0 < cg.nsten; m0 += 64) {
    const int m = m0 + lane;
    int s = 0, e = 0;
    if (m < cg.nsten) {
      int nx = fmodp(i + gmem(cg.sten)[3 * m], cg.ncx);
      int ny = fmodp(j + gmem(cg.sten)[3 * m + 1], cg.ncy);
      int nz = fmodp(k + gmem(cg.sten)[3 * m + 2], cg.ncz);
      int cc = nz + cg.ncz * (ny + cg.ncy * nx);
      s = gmem(d.cell_start)[cc];
      e = s + gmem(d.cell_cnt)[cc];
    }
    // this lane's cell: forward and reverse hits as bit masks over its (at most 64) first members
    int cf = 0, cr = 0;
    unsigned long long hf = 0ull, hr = 0ull;
    for (int t = s; t < e; ++t) {
      const double4 p = gload(cg.cpos + t);
      const int l = (int)p.w;
      if (l == n || !list_hit(x1, y1, z1, p, cg.co, cg.sqcut)) continue;
      if (l > n) {
        ++cf;
        if (t - s < 64) hf |= 1ull << (t - s);
      } else {
        ++cr;
        if (t - s < 64) hr |= 1ull << (t - s);
      }
    }
    int pf = basef + wave_excl_scan_int(cf, lane), pr = baser + wave_excl_scan_int(cr, lane);
    for (int t = s; t < e; ++t) {
      const int r = t - s;
      bool f, b;
      int l;
      if (r < 64) {
        f = (hf >> r) & 1ull;
        b = (hr >> r) & 1ull;
        l = ((f || b) && entry != 1) ? (int)gload(cg.cpos + t).w : 0;
      } else {
        const double4 p = gload(cg.cpos + t);
        l = (int)p.w;
        const bool hit = l != n && list_hit(x1, y1, z1, p, cg.co, cg.sqcut);
        f = hit && l > n;
        b = hit && l < n;
      }
      if (f || b) {
        split_put(outf, lrev, cap, f, pf, pr, entry == 1 ? t : l);
        if (f) ++pf;
        else ++pr;
      }
    }
    basef += wave_sum_int(cf);
    baser += wave_sum_int(cr);
  }
  return make_int2(basef, baser);
}

// Split slab layout: the walk, then the reverse entries ranked in LDS by group
// number (distinct) and stored in ascending order, as group numbers.
// Dynamic LDS: one slab of ints per wavefront.
__global__ void __launch_bounds__(256) k_list_write_split(const ListDesc *__restrict__ D) {
  extern __shared__ int lrev_all[];
  const ListDesc &d = D[blockIdx.y];
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (!d.split || o >= d.ngrp || d.total[1] || !list_due(d)) return;
  int *lrev = lrev_all + (threadIdx.x >> 6) * d.slab;
  const int n = d.cellorder ? d.cell_items[o] : o;
  const int2 c = list_walk_split(n, d, d.out + (size_t)o * d.slab, lrev, d.slab, d.cellorder ? 1 : 2);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // the reverse entries as group numbers, ranked: the order (and with it the
  // summation order of the neighbour-side forces) is the same in cell order and in group order
  const int nr = min(c.y, d.slab);
  if (d.cellorder)
    for (int i = lane; i < nr; i += 64) {
      const int v = lrev[i];  // a cell-order position, below ngrp
      lrev[i] = (v >= 0 && v < d.ngrp) ? gmem(d.cell_items)[v] : -1;
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int *dst = d.rout + (size_t)o * d.slab;
  for (int i = lane; i < nr; i += 64) {
    const int v = lrev[i];
    int rank = 0;
    for (int q = 0; q < nr; ++q) rank += lrev[q] < v ? 1 : 0;
    dst[rank] = v;
  }
  if (lane == 0) {
    d.count[o] = c.x;
    d.rcount[o] = c.y;
  }
}

// entry value -> group size (entries are cell-order positions or group numbers)
__device__ inline int entry_size(const ListDesc &d, int v) {
  return gmem(d.gsize)[d.cellorder ? gmem(d.cell_items)[v] : v];
}

// one block per system: first partial slot of each owner's reverse entries
// (owner-major, atom-minor, in reverse-list order); capacity check
__global__ void __launch_bounds__(1024) k_list_rbase(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.x];
  if (!d.split || !list_due(d) || d.total[1]) return;
  const int n = d.ngrp;
  // rbase[o] = rcount[o] * size(o) first, scanned in place
  for (int o = threadIdx.x; o < n; o += blockDim.x) d.rbase[o] = min(d.rcount[o], d.slab) * entry_size(d, o);
  __syncthreads();
  const int tot = block_exclusive_scan(d.rbase, d.rbase, n);
  if (threadIdx.x == 0 && (long long)tot > d.pcap) {
    d.total[1] = 1;
    d.total[3] = 1;  // sticky: cleared only by the host
  }
}

// one wavefront per owner: each forward entry's partial slot, by a binary
// search for the owner in the neighbour's sorted reverse list
__global__ void __launch_bounds__(256) k_list_mirror(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.y];
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (!d.split || o >= d.ngrp || !list_due(d)) return;
  const int nf = min(gmem(d.count)[o], d.slab);
  const bool over = d.total[1] != 0;
  for (int e = lane; e < nf; e += 64) {
    const int v = gmem(d.out)[(size_t)o * d.slab + e];
    int slot = -1;
    if (!over) {
      const int key = d.cellorder ? gmem(d.cell_items)[o] : o;  // the owner's group number
      const int *rl = d.rout + (size_t)v * d.slab;
      int lo = 0, hi = min(gmem(d.rcount)[v], d.slab);
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (gmem(rl)[mid] < key) lo = mid + 1;
        else hi = mid;
      }
      if (lo < min(gmem(d.rcount)[v], d.slab) && gmem(rl)[lo] == key) slot = gmem(d.rbase)[v] + lo * entry_size(d, v);
    }
    gmem(d.mirror)[(size_t)o * d.slab + e] = slot;
  }
}

// slab layout, one block per system: total and largest count (fixed order),
// overflow when a count exceeds the slab
__global__ void __launch_bounds__(1024) k_list_total(const ListDesc *__restrict__ D, int split_written) {
  __shared__ int wmax[16];
  const ListDesc &d = D[blockIdx.x];
  if (!list_due(d)) return;
  __shared__ int wsum[16];
  const int n = d.ngrp;
  // the slab layout reads counts only: sum and largest count, no offsets
  int tot = 0, mx = 0, s = 0;
  if (d.slab > 0) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      // after a split write: forward and reverse entries (both directions), the larger part against the slab
      const int c = d.count[i], r = (split_written && d.split) ? d.rcount[i] : 0;
      s += c + r;
      mx = max(mx, max(c, r));
    }
  } else {
    tot = block_exclusive_scan(d.count, d.off, n);
    for (int i = threadIdx.x; i < n; i += blockDim.x) mx = max(mx, d.count[i]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mx = max(mx, __shfl_xor(mx, off, 64));
    s += __shfl_xor(s, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    wmax[threadIdx.x >> 6] = mx;
    wsum[threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < 16; ++w) mx = max(mx, wmax[w]);
    if (d.slab > 0) {
      tot = 0;
      for (int w = 0; w < 16; ++w) tot += wsum[w];
    }
    d.off[n] = tot;
    d.total[0] = tot;
    d.total[1] = mx > d.slab ? 1 : 0;
    d.total[2] = mx;
    if (mx > d.slab) d.total[3] = 1;  // sticky: cleared only by the host
  }
}

void launch_list_build(const ListDesc *d_desc, const ListDesc *h, int nsys, int phases, hipStream_t st) {
  int max_cell = 1, max_grp = 1;
  for (int s = 0; s < nsys; ++s) {
    max_cell = std::max(max_cell, h[s].cg.ncx * h[s].cg.ncy * h[s].cg.ncz + 1);  // + the parking cell
    max_grp = std::max(max_grp, h[s].ngrp);
  }
  const dim3 b(256), gg((max_grp + 255) / 256, nsys), gw((max_grp + 3) / 4, nsys);
  // a batch uses one layout: all slab (resident systems) or none (plain lc_list);
  // slab systems may mix the split and the plain slab layouts
  const bool slab = h[0].slab > 0;
  bool any_split = false, any_plain = false;
  int max_slab = 0;
  for (int s = 0; s < nsys; ++s) {
    if ((h[s].slab > 0) != slab || h[s].mode != h[0].mode) throw Error(E_STATE, "list layouts differ within a batch");
    if (h[s].split) {
      any_split = true;
      max_slab = std::max(max_slab, h[s].slab);
    } else {
      any_plain = true;
    }
  }
  const int layout = h[0].mode == 1 ? LW_REF : (slab ? LW_SLAB : LW_CSR);
  if (phases & LIST_COUNT) {
    hipLaunchKernelGGL(k_list_zero, dim3(std::min(64, (max_cell + 255) / 256), nsys), b, 0, st, d_desc);
    hipLaunchKernelGGL(k_cell_assign, gg, b, 0, st, d_desc);
    hipLaunchKernelGGL(k_cell_scan, dim3(nsys), dim3(1024), 0, st, d_desc);
    hipLaunchKernelGGL(k_cell_fill, gg, b, 0, st, d_desc);
    hipLaunchKernelGGL(k_cell_sort, dim3((max_cell + 3) / 4, nsys), b, 0, st, d_desc);
    hipLaunchKernelGGL(k_cell_pack, gg, b, 0, st, d_desc);
    hipLaunchKernelGGL(k_soff_scan, dim3(nsys), dim3(1024), 0, st, d_desc);
    if (!slab || !(phases & LIST_WRITE)) {
      // counts only (or the compact layout's first pass)
      hipLaunchKernelGGL(k_list_count, gw, b, 0, st, d_desc);
      if (slab)
        hipLaunchKernelGGL(k_list_total, dim3(nsys), dim3(1024), 0, st, d_desc, 0);
      else
        hipLaunchKernelGGL(k_list_offsets, dim3(nsys), dim3(1024), 0, st, d_desc);
    }
  }
  if (phases & LIST_WRITE) {
    if (layout == LW_REF)
      hipLaunchKernelGGL(k_list_write<LW_REF>, gw, b, 0, st, d_desc);
    else if (layout == LW_SLAB) {
      if (any_plain) hipLaunchKernelGGL(k_list_write<LW_SLAB>, gw, b, 0, st, d_desc);
      if (any_split) {
        const size_t shm = (size_t)4 * max_slab * sizeof(int);
        if (shm > 65536) throw Error(E_CAPACITY, "neighbour-list slab too large for the split layout's LDS sort");
        hipLaunchKernelGGL(k_list_write_split, gw, b, shm, st, d_desc);
      }
    } else {
      hipLaunchKernelGGL(k_list_write<LW_CSR>, gw, b, 0, st, d_desc);
    }
    if (slab) hipLaunchKernelGGL(k_list_total, dim3(nsys), dim3(1024), 0, st, d_desc, any_split ? 1 : 0);
    if (any_split) {
      hipLaunchKernelGGL(k_list_rbase, dim3(nsys), dim3(1024), 0, st, d_desc);
      hipLaunchKernelGGL(k_list_mirror, gw, b, 0, st, d_desc);
    }
  }
}

// Inner list of owner o (one wavefront): the entries of its slab whose centre
// distance is below sqrt(r2m), in list order (wave ballot compaction), so the
// inner shell's pair pass walks only them.  Same schedule as the build.
__global__ void __launch_bounds__(256) k_list_inner(const ListDesc *__restrict__ D) {
  const ListDesc &d = D[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (o >= d.ngrp || !(d.r2m > 0.0) || !list_due(d)) return;  // uniform over the wave
  const int n = d.cellorder ? gmem(d.cell_items)[o] : o;
  const double x1 = gmem(d.xpg)[n], y1 = gmem(d.ypg)[n], z1 = gmem(d.zpg)[n];
  const int cnt = min(gmem(d.count)[o], d.slab);
  const int *in = d.out + (size_t)o * d.slab;
  int *out = d.out_m + (size_t)o * d.slab;
  int w = 0;
  for (int e0 = 0; e0 < cnt; e0 += 64) {
    const int e = e0 + lane;
    bool keep = false;
    int l = 0;
    if (e < cnt) {
      l = in[e];
      const double4 p = d.cellorder ? gload(d.cg.cpos + l)
                                    : make_double4(gmem(d.xpg)[l], gmem(d.ypg)[l], gmem(d.zpg)[l], 0.0);
      keep = list_hit(x1, y1, z1, p, d.cg.co, d.r2m);
    }
    const unsigned long long m = __ballot(keep);
    if (keep) out[w + __popcll(m & ((1ull << lane) - 1ull))] = l;
    w += __popcll(m);
  }
  if (lane == 0) d.count_m[o] = w;
}

void launch_list_inner(const ListDesc *d_desc, const ListDesc *h, int nsys, hipStream_t st) {
  int max_grp = 1;
  for (int s = 0; s < nsys; ++s) {
    if (h[s].r2m > 0.0 && (h[s].slab <= 0 || h[s].split)) throw Error(E_STATE, "inner lists need the plain slab layout");
    max_grp = std::max(max_grp, h[s].ngrp);
  }
  hipLaunchKernelGGL(k_list_inner, dim3((max_grp + 3) / 4, nsys), dim3(256), 0, st, d_desc);
}

// Dense LJ table of a system: entry (i, j) = {ecc12, ecc6} of the packed
// pair index type_table(i, j) (1-based, column-major), and its FP32 copy
// when asked (the mixed-precision pair kernel's).  The host has checked the
// indices.
__global__ void __launch_bounds__(256) k_lj_dense(int ma, const int *__restrict__ tt, const double *__restrict__ e12,
                                                  const double *__restrict__ e6, double2 *dense, float2 *dense32) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)ma * ma) return;
  const int i = (int)(k / ma), j = (int)(k - (size_t)i * ma);
  const int lij = tt[i + (size_t)ma * j] - 1;
  const double2 v = make_double2(e12[lij], e6[lij]);
  dense[k] = v;
  if (dense32) dense32[k] = make_float2((float)v.x, (float)v.y);
}

void launch_lj_dense(int ma, const int *tt, const double *e12, const double *e6, double2 *dense, float2 *dense32,
                     hipStream_t st) {
  const size_t n = (size_t)ma * ma;
  if (n > 0) hipLaunchKernelGGL(k_lj_dense, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ma, tt, e12, e6,
                                dense, dense32);
}

// ---------------------------------------------------------------- half -> full

__global__ void k_h2f_count(int ngrp, const int *h_off, const int *h_idx, int *incount) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngrp) return;
  for (int t = h_off[g]; t < h_off[g + 1]; ++t) atomicAdd(&incount[h_idx[t]], 1);
}

__global__ void k_h2f_total(int ngrp, const int *h_off, int *incount) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngrp) return;
  incount[g] += h_off[g + 1] - h_off[g];
}

__global__ void k_h2f_fill(int ngrp, const int *h_off, const int *h_idx, const int *f_off, int *f_idx, int *cursor) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngrp) return;
  int nh = h_off[g + 1] - h_off[g];
  for (int t = 0; t < nh; ++t) f_idx[f_off[g] + t] = h_idx[h_off[g] + t];
  for (int t = h_off[g]; t < h_off[g + 1]; ++t) {
    int j = h_idx[t];
    int nhj = h_off[j + 1] - h_off[j];
    int pos = atomicAdd(&cursor[j], 1);
    f_idx[f_off[j] + nhj + pos] = g;
  }
}

__global__ void k_h2f_sort(int ngrp, const int *h_off, const int *f_off, int *f_idx) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngrp) return;
  int s = f_off[g] + (h_off[g + 1] - h_off[g]), n = f_off[g + 1] - s;
  for (int i = 1; i < n; ++i) {
    int v = f_idx[s + i];
    int j = i - 1;
    while (j >= 0 && f_idx[s + j] > v) {
      f_idx[s + j + 1] = f_idx[s + j];
      --j;
    }
    f_idx[s + j + 1] = v;
  }
}

void launch_half_to_full(int ngrp, const int *h_off, const int *h_idx, int *incount, int *f_off, int *f_idx,
                         int *cursor, int *total, hipStream_t st) {
  dim3 b(256), g((ngrp + 255) / 256);
  if (ngrp <= 0) return;
  hipLaunchKernelGGL(k_zero_i32, g, b, 0, st, incount, ngrp + 1);
  hipLaunchKernelGGL(k_zero_i32, g, b, 0, st, cursor, ngrp);
  hipLaunchKernelGGL(k_h2f_count, g, b, 0, st, ngrp, h_off, h_idx, incount);
  hipLaunchKernelGGL(k_h2f_total, g, b, 0, st, ngrp, h_off, incount);
  launch_exclusive_scan(incount, f_off, ngrp + 1, total, st);
  hipLaunchKernelGGL(k_h2f_fill, g, b, 0, st, ngrp, h_off, h_idx, f_off, f_idx, cursor);
  hipLaunchKernelGGL(k_h2f_sort, g, b, 0, st, ngrp, h_off, f_off, f_idx);
}

// ---------------------------------------------------------------- inner-shell list

// force.f:265-286: for each group i, the neighbours j (in list order) whose
// centre distance is below r2neigh.  iz = 0 walks j = i+1..ngrp instead.
__device__ inline double centre_d2(const double *xpg, const double *ypg, const double *zpg, const double *co, int i,
                                   int j) {
#pragma clang fp contract(off)
  double xgg = xpg[i] - xpg[j], ygg = ypg[i] - ypg[j], zgg = zpg[i] - zpg[j];
  xgg = xgg - 2.0 * fnint(0.5 * xgg);
  ygg = ygg - 2.0 * fnint(0.5 * ygg);
  zgg = zgg - 2.0 * fnint(0.5 * zgg);
  double xc = co[0] * xgg + co[3] * ygg + co[6] * zgg;
  double yc = co[4] * ygg + co[7] * zgg;
  double zc = co[8] * zgg;
  return xc * xc + yc * yc + zc * zc;
}

struct Co9 {
  double m[9];
};

__global__ void k_inner_list(int ngrp, int iz, const int *h_off, const int *h_idx, const double *xpg,
                             const double *ypg, const double *zpg, Co9 co, double r2neigh, int *count, const int *off,
                             int *out, int pass) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ngrp) return;
  int nv = 0;
  int b = iz == 1 ? h_off[i] : 0, e = iz == 1 ? h_off[i + 1] : 0;
  int jlo = iz == 1 ? 0 : i + 1, jhi = iz == 1 ? 0 : ngrp;
  int o = pass ? off[i] + i + 1 : 0;
  for (int t = b; t < e; ++t) {
    int j = h_idx[t];
    if (centre_d2(xpg, ypg, zpg, co.m, i, j) < r2neigh) {
      if (pass) out[o + nv] = j + 1;
      ++nv;
    }
  }
  for (int j = jlo; j < jhi; ++j) {
    if (centre_d2(xpg, ypg, zpg, co.m, i, j) < r2neigh) {
      if (pass) out[o + nv] = j + 1;
      ++nv;
    }
  }
  if (pass) out[off[i] + i] = nv;
  else count[i] = nv;
}

void launch_inner_list(int ngrp, int iz, const int *h_off, const int *h_idx, const double *xpg, const double *ypg,
                       const double *zpg, const double *co, double r2neigh, int *count, const int *off, int *out,
                       int pass, hipStream_t st) {
  if (ngrp <= 0) return;
  Co9 c;
  for (int k = 0; k < 9; ++k) c.m[k] = co[k];
  hipLaunchKernelGGL(k_inner_list, dim3((ngrp + 127) / 128), dim3(128), 0, st, ngrp, iz, h_off, h_idx, xpg, ypg, zpg,
                     c, r2neigh, count, off, out, pass);
}

// ---------------------------------------------------------------- ferrf / fnbgrp

// The scaling factors of one alchemical 1-4 pair (fnb14_alchemy.f:131-155,
// 161-185, 214-255): normal pairs and pairs of equal nonzero lambda keep the full
// potential, a pair with one alchemical atom is scaled by 1 - |lambda| of that
// atom, a pair of an added and a removed atom is switched off.
__device__ __forceinline__ void alch14_fact(double li, double lj, double li0, double lj0, double &f, double &f0) {
  if (li == 0.0 && lj != 0.0) {
    f = 1.0 - fabs(lj);
    f0 = 1.0 - fabs(lj0);
  } else if (li != 0.0 && lj == 0.0) {
    f = 1.0 - fabs(li);
    f0 = 1.0 - fabs(li0);
  } else if (li == 0.0 && lj == 0.0) {
    f = 1.0;
    f0 = 1.0;
  } else if (li != lj) {
    f = 0.0;
    f0 = 0.0;
  } else {
    f = 1.0;
    f0 = 1.0;
  }
}

__device__ __forceinline__ double erfc_as_dev(double alphar, double &expcst) {
  const double qt = 1.0 / (1.0 + AS_QP * alphar);
  expcst = exp(-(alphar * alphar));
  return ((((AS_A5 * qt + AS_A4) * qt + AS_A3) * qt + AS_A2) * qt + AS_A1) * qt * expcst;
}

__global__ void __launch_bounds__(256) k_corr_pairs(const CorrDesc *__restrict__ D) {
  const CorrDesc &d = D[blockIdx.y];
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= d.npairs) return;
  const int kind = d.pkind ? d.pkind[p] : d.kind;
  int ia = d.ia[p], ib = d.ib[p];
  double xab, yab, zab;
  if (kind == 1 || kind == 2 || kind == 4) {
    // fnbgrp.f, fnb14.f, fnb14_alchemy.f: the coordinates as given, no PBC
    xab = d.xp[ia] - d.xp[ib];
    yab = d.yp[ia] - d.yp[ib];
    zab = d.zp[ia] - d.zp[ib];
  } else if (d.x) {
    xab = d.x[ia] - d.x[ib];
    yab = d.y[ia] - d.y[ib];
    zab = d.z[ia] - d.z[ib];
  } else {
    const double *co = d.co;
    const double xa = co[0] * d.xp[ia] + co[3] * d.yp[ia] + co[6] * d.zp[ia];
    const double ya = co[1] * d.xp[ia] + co[4] * d.yp[ia] + co[7] * d.zp[ia];
    const double za = co[2] * d.xp[ia] + co[5] * d.yp[ia] + co[8] * d.zp[ia];
    const double xb = co[0] * d.xp[ib] + co[3] * d.yp[ib] + co[6] * d.zp[ib];
    const double yb = co[1] * d.xp[ib] + co[4] * d.yp[ib] + co[7] * d.zp[ib];
    const double zb = co[2] * d.xp[ib] + co[5] * d.yp[ib] + co[8] * d.zp[ib];
    xab = xa - xb;
    yab = ya - yb;
    zab = za - zb;
  }
  const double rsq = xab * xab + yab * yab + zab * zab;
  const double rsp = sqrt(rsq);
  const int c = d.ss[ia] - 1;  // energies by the class of the first atom (ferrf.f:169, fnbgrp.f:130)
  double ev[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double qforce = 0.0;
  if (kind == 0) {
    // ferrf.f:162-188 (iz = 1) and 212-226 (iz = 0)
    double fud = d.fudge_p ? d.fudge_p[p] : d.fudge;
    double furpar = d.iz == 1 ? fud * d.q[ia] * d.q[ib] : d.q[ia] * d.q[ib];
    double alphar = d.alpha * rsp;
    double expcst;
    const double erfcst = erfc_as_dev(alphar, expcst);
    double erfst = 1.0 - erfcst;
    double ecoul;
    if (d.erf_tab) {
      // ERF_CORR (ferrf.f:68-105 and 170-193): the reciprocal-cutoff correction
      // from its cubic spline table, interval ic with offset h
      int ic = (int)(1.0 + (rsp - d.rlew) / d.delew);
      ic = min(max(ic, 1), d.erf_n);
      const double h = rsp - (d.rlew + (ic - 1) * d.delew);
      const double *cs = d.erf_tab + 4 * (size_t)(ic - 1);
      const double poly = cs[0] + h * (cs[1] + h * (cs[2] + h * cs[3] / 3) / 2);
      if (d.iz == 1) {
        const double qq = d.q[ia] * d.q[ib];
        furpar = fud * qq;
        const double corr = qq * poly;
        const double dcorr = qq * (cs[1] + h * (cs[2] + h * cs[3] / 2)) / rsp;
        qforce = -dcorr - furpar * (erfst - d.twrtpi * alphar * expcst) / (rsp * rsq);
        ecoul = -(furpar * erfst / rsp - corr);
      } else {
        ecoul = -(furpar * (erfst / rsp - poly));
      }
    } else {
      ecoul = -(furpar * erfst / rsp);
      if (d.iz == 1) qforce = -(furpar * (erfst - d.twrtpi * alphar * expcst) / (rsp * rsq));
    }
    ev[c] = ecoul;
    if (d.q0) {
      // ferrf_alchemy.f:144-145,190-191: the same term on the lambda0 charges (en0)
      double furpar0 = d.iz == 1 ? fud * d.q0[ia] * d.q0[ib] : d.q0[ia] * d.q0[ib];
      ev[2] = -(furpar0 * erfst / rsp);
    }
  } else if (kind == 3) {
    // ferrf_alchemy_add.f:94-141 (ERF_CORR) and 152-199: the intra-alchemical
    // pairs beyond 1-4, with factl = (1 - |lambdaq_a|)(1 - |lambdaq_b|) - facQ on
    // the created charges; with INTRAMOL only the added-removed pairs (:105)
    const int fq = d.facq[p];
    if (!(d.intramol && fq != 0)) {
      const double facq = (double)fq;
      const double factl = (1.0 - fabs(d.lq[ia])) * (1.0 - fabs(d.lq[ib])) - facq;
      const double factl0 = (1.0 - fabs(d.lq0[ia])) * (1.0 - fabs(d.lq0[ib])) - facq;
      const double qq = factl * d.qa[ia] * d.qa[ib];
      const double qq0 = factl0 * d.qa[ia] * d.qa[ib];
      const double alphar = d.alpha * rsp;
      double expcst;
      const double erfst = 1.0 - erfc_as_dev(alphar, expcst);
      if (d.erf_tab) {
        int ic = (int)(1.0 + (rsp - d.rlew) / d.delew);
        ic = min(max(ic, 1), d.erf_n);
        const double h = rsp - (d.rlew + (ic - 1) * d.delew);
        const double *cs = d.erf_tab + 4 * (size_t)(ic - 1);
        const double corr = cs[0] + h * (cs[1] + h * (cs[2] + h * cs[3] / 3) / 2);
        const double dcorr = (cs[1] + h * (cs[2] + h * cs[3] / 2)) / rsp;
        qforce = -qq * (dcorr + (erfst - d.twrtpi * alphar * expcst) / (rsp * rsq));
        ev[4 + c] = -(qq * (erfst / rsp - corr));
        ev[6] = -(qq0 * (erfst / rsp - corr));
      } else {
        qforce = -qq * (erfst - d.twrtpi * alphar * expcst) / (rsp * rsq);
        ev[4 + c] = -(qq * erfst / rsp);
        ev[6] = -(qq0 * erfst / rsp);
      }
    }
  } else if (kind == 2) {
    // fnb14.f:69-158: LJ by the pair's packed type scaled by lj_fudge; charges
    // by fudge; erfc only with ewald and cut
    const int lij = d.pair_type[p] - 1;
    const double rsqi = 1.0 / rsq;
    const double r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
    const double ssvir = 12.0 * d.ecc12[lij] * r12 - 6.0 * d.ecc6[lij] * r6;
    ev[2 + c] = d.lj_fudge * (d.ecc12[lij] * r12 - d.ecc6[lij] * r6);
    if (d.ewald) {
      const double alphar = d.alpha * rsp;
      double expcst;
      const double erfcst = erfc_as_dev(alphar, expcst);
      const double furpar = d.fudge * d.q[ia] * d.q[ib];
      ev[c] = furpar * erfcst / rsp;
      qforce = d.lj_fudge * ssvir * rsqi + furpar * (erfcst + d.twrtpi * alphar * expcst) * (rsqi / rsp);
    } else {
      const double ecoul = d.fudge * d.q[ia] * d.q[ib] / rsp;
      ev[c] = ecoul;
      qforce = d.lj_fudge * ssvir * rsqi + ecoul * rsqi;
    }
    if (d.remf && d.remf[p]) {  // fnb14_rem.f:146-149, 209-212
      ev[4] = ev[2 + c] + ev[c];
      qforce = qforce * d.remf2;
    }
  } else if (kind == 4) {
    // fnb14_alchemy.f:120-356, with the reference's own arithmetic: without
    // Ewald the forces are unscaled and the LJ lambda0 energy uses alch_fact
    // (:203-210); with Ewald only the repulsive LJ force term carries
    // alch_fact (:286), and pairs of two alchemical atoms add the erf term
    double af, af0, aq, aq0;
    alch14_fact(d.lam[ia], d.lam[ib], d.lam0[ia], d.lam0[ib], af, af0);
    alch14_fact(d.lq[ia], d.lq[ib], d.lq0[ia], d.lq0[ib], aq, aq0);
    const int lij = d.pair_type[p] - 1;
    const double rsqi = 1.0 / rsq;
    const double r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
    const double qi = d.q[ia], qj = d.q[ib];
    if (!d.ewald) {
      const double ssvir = 12.0 * d.ecc12[lij] * r12 - 6.0 * d.ecc6[lij] * r6;
      qforce = d.lj_fudge * ssvir * rsqi;
      const double elj = d.lj_fudge * (d.ecc12[lij] * r12 - d.ecc6[lij] * r6);
      ev[2 + c] = af * elj;
      ev[6 + c] = af * elj;
      const double sv = d.fudge * qi * qj / rsp;
      ev[c] = aq * sv;
      ev[4 + c] = aq0 * sv;
      qforce = qforce + sv * rsqi;
    } else {
      double aq2 = 0.0, aq20 = 0.0;
      const double lqa = d.lq[ia], lqb = d.lq[ib];
      if (lqa != 0.0 && lqb != 0.0) {
        const double pr = (1.0 - fabs(lqa)) * (1.0 - fabs(lqb));
        const double pr0 = (1.0 - fabs(d.lq0[ia])) * (1.0 - fabs(d.lq0[ib]));
        aq2 = lqa != lqb ? -pr : 1.0 - pr;
        aq20 = lqa != lqb ? -pr0 : 1.0 - pr0;
      }
      const double rspqi = rsqi / rsp;
      const double ssvir = af * 12.0 * d.ecc12[lij] * r12 - 6.0 * d.ecc6[lij] * r6;
      qforce = d.lj_fudge * ssvir * rsqi;
      const double aux = d.lj_fudge * (d.ecc12[lij] * r12 - d.ecc6[lij] * r6);
      ev[2 + c] = af * aux;
      ev[6 + c] = af0 * aux;
      const double alphar = d.alpha * rsp;
      double expcst;
      const double erfcst = erfc_as_dev(alphar, expcst);
      const double furpar = aq * d.fudge * qi * qj;
      const double furpar0 = aq0 * d.fudge * qi * qj;
      const double erff = 1.0 - erfcst;
      const double furpar2 = aq2 * d.fudge * qi * qj;
      const double furpar20 = aq20 * d.fudge * qi * qj;
      ev[c] = furpar * erfcst / rsp + furpar2 * erff / rsp;
      ev[4 + c] = furpar0 * erfcst / rsp + furpar20 * erff / rsp;
      const double a1 = (erfcst + d.twrtpi * alphar * expcst) * rspqi;
      const double a2 = (erff - d.twrtpi * alphar * expcst) * rspqi;
      qforce = qforce + furpar * a1 + furpar2 * a2;
    }
  } else {
    // fnbgrp.f:120-202 (no PBC)
    int ti = d.type[ia], tj = d.type[ib];
    int lj = ti > tj ? ti : tj, li = ti < tj ? ti : tj;
    int lij = lj * (lj - 1) / 2 + li - 1;
    double rsqi = 1.0 / rsq;
    double r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
    double ssvir = 12.0 * d.ecc12[lij] * r12 - 6.0 * d.ecc6[lij] * r6;
    qforce = ssvir * rsqi;
    ev[2 + c] = d.ecc12[lij] * r12 - d.ecc6[lij] * r6;
    double furpar = d.q[ia] * d.q[ib];
    if (d.ewald) {
      double alphar = d.alpha * rsp;
      double expcst;
      const double erfcst = erfc_as_dev(alphar, expcst);
      ev[c] = furpar * erfcst / rsp;
      qforce += furpar * (erfcst + d.twrtpi * alphar * expcst) * (rsqi / rsp);
    } else {
      const double ecoul = furpar / rsp;
      ev[c] = ecoul;
      qforce += ecoul * rsqi;
    }
  }
  d.pf[3 * (size_t)p] = qforce * xab;
  d.pf[3 * (size_t)p + 1] = qforce * yab;
  d.pf[3 * (size_t)p + 2] = qforce * zab;
  const int ncol = d.ncol ? d.ncol : 4;
  double *pe = d.pe + ncol * (size_t)p;
  for (int k = 0; k < ncol; ++k) pe[k] = ev[k];
}

__global__ void __launch_bounds__(256) k_corr_gather(const CorrDesc *__restrict__ D) {
  const CorrDesc &d = D[blockIdx.y];
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (d.iz != 1 || a >= d.natoms || d.npairs <= 0) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int t = d.g_off[a]; t < d.g_off[a + 1]; ++t) {
    int sp = d.g_pair[t];
    int p = (sp > 0 ? sp : -sp) - 1;
    double s = sp > 0 ? 1.0 : -1.0;
    sx += s * d.pf[3 * (size_t)p];
    sy += s * d.pf[3 * (size_t)p + 1];
    sz += s * d.pf[3 * (size_t)p + 2];
  }
  if (d.write) {
    d.fx[a] = sx;
    d.fy[a] = sy;
    d.fz[a] = sz;
  } else {
    d.fx[a] += sx;
    d.fy[a] += sy;
    d.fz[a] += sz;
  }
}

// d.out[ncol]: coul(slt), coul(slv), lj(slt), lj(slv) (and the alchemical
// columns of kinds 3 and 4); block k reduces column k in fixed order
__global__ void __launch_bounds__(256) k_corr_energy(const CorrDesc *__restrict__ D) {
  __shared__ double red[4];
  const CorrDesc &d = D[blockIdx.y];
  const int k = blockIdx.x;
  const int ncol = d.ncol ? d.ncol : 4;
  if (k >= ncol) return;
  double s = 0.0;
  for (int p = threadIdx.x; p < d.npairs; p += blockDim.x) s += d.pe[ncol * (size_t)p + k];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    d.out[k] = t;
  }
}

void launch_corr(const CorrDesc *d_desc, const CorrDesc *h, int nsys, hipStream_t st) {
  int max_pairs = 0, max_atoms = 0, max_col = 4;
  for (int s = 0; s < nsys; ++s) {
    max_pairs = std::max(max_pairs, h[s].npairs);
    if (h[s].iz == 1 && h[s].npairs > 0) max_atoms = std::max(max_atoms, h[s].natoms);
    max_col = std::max(max_col, h[s].ncol);
  }
  if (max_pairs > 0)
    hipLaunchKernelGGL(k_corr_pairs, dim3((max_pairs + 255) / 256, nsys), dim3(256), 0, st, d_desc);
  if (max_atoms > 0)
    hipLaunchKernelGGL(k_corr_gather, dim3((max_atoms + 255) / 256, nsys), dim3(256), 0, st, d_desc);
  hipLaunchKernelGGL(k_corr_energy, dim3(max_col, nsys), dim3(256), 0, st, d_desc);
}

// ---------------------------------------------------------------- cself

// cself.f:45-56 sums, fixed order: out[0] = sum q^2 (ss 1), out[1] = (ss 2)
__global__ void k_cself(int n, const double *q, const int *ss, double *out) {
  __shared__ double red[2][16];
  double s[2] = {0.0, 0.0};
  int per = (n + blockDim.x - 1) / blockDim.x;
  int a0 = threadIdx.x * per, a1 = min(n, a0 + per);
  for (int a = a0; a < a1; ++a) s[ss[a] - 1] += q[a] * q[a];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int k = 0; k < 2; ++k) {
    double v = wave_sum(s[k]);
    if (lane == 0) red[k][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[threadIdx.x][w];
    out[threadIdx.x] = t;
  }
}

void launch_cself(int n, const double *q, const int *ss, double *d_out2, hipStream_t st) {
  hipLaunchKernelGGL(k_cself, dim3(1), dim3(1024), 0, st, n, q, ss, d_out2);
}

// ---------------------------------------------------------------- small utilities

__global__ void k_zero_u64(unsigned long long *p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}
__global__ void k_zero_f64(double *p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
void launch_zero_u64(unsigned long long *p, size_t n, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}
void launch_zero_f64(double *p, size_t n, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}
}  // namespace orac
