// force_pairs.hip -- the direct-space pair kernel (force.f:235-870) on gfx950.
//
// One wavefront per owner group.  Lane l serves owner atom ai = l % nI for
// the neighbour-group slot l / nI, so a pass covers 64 / nI neighbour groups
// (21 for a water owner) and each lane evaluates one owner atom against the
// atoms of one neighbour group.  The neighbour list is full (both directions),
// so every atom's force is summed by its own wavefront: no atomics, and the
// order of every sum is fixed.  Energies are taken from the j > i direction
// only, so each pair counts once.
//
// Molecular virial.  force.f accumulates st += f_ab (x) (C_a - C_b - s) per pair
// (C = molecule centre, s = the group-pair image shift) and, for switched
// pairs, the S'V terms with C_a and -(C_b + s).  Summed over pairs this equals
//     sum_a F_a (x) C_a  -  sum_pairs f_ab (x) s  -  sum_switched x3 (x) s,
// with F_a the total direct force on atom a.  The kernel accumulates that
// form: the shift terms are nonzero only for pairs across the cell boundary,
// and the per-atom term is added once per atom after the force reduction.
#include <algorithm>
#include <cstdlib>

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

// Cell-ordered copies of the step's group and atom data: sorted position t
// holds group perm[t] (identity without a permutation) and its atoms from
// soff[t] on.  Neighbours listed next to each other are then next to each
// other in memory, and one group is one 32-byte and one 16-byte record.
__global__ void __launch_bounds__(256) k_pack_sorted(const PairDesc *__restrict__ D) {
  const PairDesc &d = D[blockIdx.y];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.ngrp) return;
  const int g = d.perm ? d.perm[t] : t;
  const int f0 = d.gfirst[g], n = d.glast[g] - f0 + 1;
  const int sf = d.soff ? d.soff[t] : f0;
  d.sgp[t] = make_double4(d.xpg[g], d.ypg[g], d.zpg[g], 0.0);
  d.sgi[t] = make_int4(sf, f0, n | (d.gss[g] << 8), g);
  for (int k = 0; k < n; ++k) {
    const int a = f0 + k;
    d.sap[sf + k] = make_double4(d.xp[a], d.yp[a], d.zp[a], d.q[a]);
    d.sai[sf + k] = make_int2(d.type[a], d.ex_inter[a] ? 1 : 0);
  }
}

__device__ inline bool has_partner_in(const int *ex_off, const int *ex_idx, int a, int lo, int hi) {
  for (int t = ex_off[a]; t < ex_off[a + 1]; ++t) {
    int p = ex_idx[t];
    if (p >= lo && p <= hi) return true;
  }
  return false;
}

__device__ inline bool is_excluded(const int *ex_off, const int *ex_idx, int a, int b) {
  for (int t = ex_off[a]; t < ex_off[a + 1]; ++t)
    if (ex_idx[t] == b) return true;
  return false;
}

// Pair arithmetic built on one reciprocal square root per pair.  FP64: exact
// divide and sqrt (results agree with the reference's expressions to a few
// ulp).  FP32 (mixed precision): the hardware reciprocal, reciprocal square
// root and exponential, 1-2 ulp each.
template <typename T>
__device__ inline T pair_rsqrt(T x);
template <>
__device__ inline double pair_rsqrt<double>(double x) { return 1.0 / sqrt(x); }
template <>
__device__ inline float pair_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <typename T>
__device__ inline T pair_rcp(T x);
template <>
__device__ inline double pair_rcp<double>(double x) { return 1.0 / x; }
template <>
__device__ inline float pair_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <typename T>
__device__ inline T pair_exp(T x);
template <>
__device__ inline double pair_exp<double>(double x) { return exp(x); }
template <>
__device__ inline float pair_exp<float>(float x) { return __expf(x); }

__device__ inline void lj_params(const PairDesc &d, int ti, int tj, double &c12, double &c6) {
  if (d.lj_dense) {
    double2 v = d.lj_dense[(size_t)ti * d.ntypes + tj];
    c12 = v.x;
    c6 = v.y;
  } else {
    int lij = d.type_table[ti + (size_t)d.ntypes * tj];
    c12 = d.ecc12[lij];
    c6 = d.ecc6[lij];
  }
}

struct PairShared {
  double red[PAIR_WAVES][16];
  double facc[PAIR_WAVES][64][3];
  double acc[PAIR_WAVES][15][64];  // per-lane energy and virial sums (lane-private slots)
};

// One work item: block blk of system d (PAIR_WAVES owner groups).
template <typename T>
__device__ __forceinline__ void pairs_item(const PairDesc &d, int blk, PairShared &sh) {
  if (blk >= d.nblk) return;
  double(&red)[PAIR_WAVES][16] = sh.red;
  double(&facc)[PAIR_WAVES][64][3] = sh.facc;
  double(&acc)[PAIR_WAVES][15][64] = sh.acc;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int t = blk * PAIR_WAVES + wid;  // owner group, sorted position
  double fxa = 0.0, fya = 0.0, fza = 0.0;
  double *en = &acc[wid][0][0];  // en[k * 64 + lane]: uconf[typeij], ucoul[typeij]
  double *st = &acc[wid][6][0];  // st[k * 64 + lane]
#pragma unroll
  for (int k = 0; k < 15; ++k) acc[wid][k][lane] = 0.0;
  int ia0 = 0, nI = 1, nslot = 64;  // ia0: owner's first atom, original numbering
  if (t < d.ngrp) {
    const int4 gi = d.sgi[t];
    const int g = gi.w;
    ia0 = gi.y;
    nI = gi.z & 255;
    nslot = 64 / nI;
    const int slot = lane / nI, ai = lane - slot * nI;
    const bool active = slot < nslot;
    const int a = ia0 + (active ? ai : 0);        // original index
    const int as = gi.x + (active ? ai : 0);      // sorted index
    const double4 pa4 = d.sap[as];
    const int2 ia4 = d.sai[as];
    const double ma = d.gmass[a];
    const int typei = gi.z >> 8;
    const double4 gpi = d.sgp[t];
    const double xpgi = gpi.x, ypgi = gpi.y, zpgi = gpi.z;
    const double co0 = d.co[0], co3 = d.co[3], co6 = d.co[6], co4 = d.co[4], co7 = d.co[7], co8 = d.co[8];
    int nb0, nb1;
    if (d.nb_slab > 0) {
      nb0 = t * d.nb_slab;
      nb1 = nb0 + min(d.nb_cnt[t], d.nb_slab);
    } else {
      nb0 = d.nb_off[t];
      nb1 = d.nb_off[t + 1];
    }
    int tjn = (active && nb0 + slot < nb1) ? d.nb_idx[nb0 + slot] : t;
    for (int base = nb0; base < nb1; base += nslot) {
      const int jj = base + slot;
      const bool have = active && jj < nb1;
      const int tj = tjn;
      // the next neighbour's index is fetched before this one is worked on
      tjn = (active && jj + nslot < nb1) ? d.nb_idx[jj + nslot] : t;
      const double4 gpj = d.sgp[tj];
      const int4 gij = d.sgi[tj];
      const int j = gij.w;
      // group classification (force.f:265-335), no contraction
      double sxm, sym, szm, xcc, ycc, zcc, drj;
      {
#pragma clang fp contract(off)
        double xgg = xpgi - gpj.x, ygg = ypgi - gpj.y, zgg = zpgi - gpj.z;
        sxm = 2.0 * fnint(0.5 * xgg);
        sym = 2.0 * fnint(0.5 * ygg);
        szm = 2.0 * fnint(0.5 * zgg);
        xgg = xgg - sxm;
        ygg = ygg - sym;
        zgg = zgg - szm;
        xcc = co0 * xgg + co3 * ygg + co6 * zgg;
        ycc = co4 * ygg + co7 * zgg;
        zcc = co8 * zgg;
        drj = xcc * xcc + ycc * ycc + zcc * zcc;
      }
      int cls = 0;  // 0 none, 1 normal, 2 switched
      double sw = 1.0, dsw = 0.0;
      if (have) {
        if (drj >= d.r2inn0 && drj < d.r2out) {
          cls = 1;
        } else if (drj >= d.r2out && drj < d.r2out0) {
          cls = 2;
          double rsp = sqrt(drj);
          double auxa = (d.arsout1 + 2.0 * rsp) / d.arsout2;
          double auxb = d.rout0 - rsp;
          sw = auxa * (auxb * auxb);
          dsw = (-(2.0 * auxa * auxb) + 2.0 * (auxb * auxb) / d.arsout2) / rsp;
        } else if (drj >= d.r2inn && drj < d.r2inn0) {
          cls = 2;
          double rsp = sqrt(drj);
          double auxa = (d.arsinn1 + 2.0 * rsp) / d.arsinn2;
          double auxb = d.rinn0 - rsp;
          sw = 1.0 - auxa * (auxb * auxb);
          dsw = (2.0 * auxa * auxb - 2.0 * (auxb * auxb) / d.arsinn2) / rsp;
        }
      }
      const bool fwd = j > g;
      const bool shifted = sxm != 0.0 || sym != 0.0 || szm != 0.0;
      double cm = 0.0;                      // this lane's part of V (switched pairs)
      double fgx = 0.0, fgy = 0.0, fgz = 0.0;  // this lane's pair forces for the group pair
      if (cls != 0) {
        const int nJ = gij.z & 255;
        const int jb0 = gij.y, jb1 = jb0 + nJ - 1, jbs = gij.x;  // original range, sorted start
        const int typeij = typei + (gij.z >> 8) - 2;  // 0..2
        const bool mask_a = fwd && (ia4.y & 1) && has_partner_in(d.ex_off, d.ex_idx, a, jb0, jb1);
        double ucoula = 0.0, uconfa = 0.0;
        for (int k = 0; k < nJ; ++k) {
          const int b = jb0 + k;
          const int2 ib4 = d.sai[jbs + k];
          const bool masked =
              fwd ? mask_a : ((ib4.y & 1) && has_partner_in(d.ex_off, d.ex_idx, b, ia0, ia0 + nI - 1));
          if (masked) {
            if (is_excluded(d.ex_off, d.ex_idx, a, b)) continue;
            if (d.any_heavy && !(d.mass[b] < d.massinfty || d.mass[a] < d.massinfty)) continue;
          }
          double c12, c6;
          lj_params(d, ia4.x, ib4.x, c12, c6);
          if (d.lskip_ewald && c6 < 1e-12 && d.ewald) continue;
          const double4 pb4 = d.sap[jbs + k];
          double xg, yg, zg;
          if (cls == 1) {
            xg = (pa4.x - sxm) - pb4.x;
            yg = (pa4.y - sym) - pb4.y;
            zg = (pa4.z - szm) - pb4.z;
          } else {
            xg = (pa4.x - pb4.x) - sxm;
            yg = (pa4.y - pb4.y) - sym;
            zg = (pa4.z - pb4.z) - szm;
          }
          const T xc = T(co0 * xg + co3 * yg + co6 * zg);
          const T yc = T(co4 * yg + co7 * zg);
          const T zc = T(co8 * zg);
          const T rsq = xc * xc + yc * yc + zc * zc;
          const T rinv = pair_rsqrt<T>(rsq);
          const T rsqi = rinv * rinv;
          T qforce = T(0);
          double vpair = 0.0;
          if (c6 > 1e-12) {
            T r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
            qforce = (T(12.0 * c12) * r12 - T(6.0 * c6) * r6) * rsqi;
            double conf = (double)(T(c12) * r12 - T(c6) * r6);
            vpair += conf;
            if (cls == 1) uconfa += conf;
            else if (fwd) en[typeij * 64 + lane] += sw * conf;
          }
          if (!d.lskip_ewald) {
            T rsp = rsq * rinv;
            T rspqi = rsqi * rinv;
            T alphar = T(d.alpha) * rsp;
            T qt = pair_rcp<T>(T(1) + T(AS_QP) * alphar);
            T expcst = pair_exp<T>(-(alphar * alphar));
            T erfcst = ((((T(AS_A5) * qt + T(AS_A4)) * qt + T(AS_A3)) * qt + T(AS_A2)) * qt + T(AS_A1)) * qt * expcst;
            T furpar = T(pa4.w * pb4.w);
            double vel = (double)(furpar * erfcst * rinv);
            vpair += vel;
            if (cls == 1) ucoula += vel;
            else if (fwd) en[(3 + typeij) * 64 + lane] += sw * vel;
            qforce += furpar * (erfcst + T(d.twrtpi) * alphar * expcst) * rspqi;
          } else if (!d.ewald) {
            T ssv = T(pa4.w * pb4.w) * rinv;
            vpair += (double)ssv;
            if (fwd) en[(3 + typeij) * 64 + lane] += (cls == 1 ? 1.0 : sw) * (double)ssv;
            qforce += ssv * rsqi;
          }
          double qf = (double)qforce;
          if (cls == 2) {
            qf *= sw;
            cm += vpair;
          }
          const double fxv = qf * (double)xc, fyv = qf * (double)yc, fzv = qf * (double)zc;
          fgx += fxv;
          fgy += fyv;
          fgz += fzv;
        }
        fxa += fgx;
        fya += fgy;
        fza += fgz;
        if (fwd && cls == 1) {
          en[typeij * 64 + lane] += uconfa;
          en[(3 + typeij) * 64 + lane] += ucoula;
        }
      }
      // total V of the group pair over the slot's lanes (every lane shuffles)
      double cmt = 0.0;
      for (int u = 0; u < nI; ++u) cmt += __shfl(cm, (slot * nI + u) & 63, 64);
      if (cls == 2) {
        // S'(r) V spread over group atoms by mass fraction (force.f:756-817)
        const double x3 = -(dsw * cmt * xcc), y3 = -(dsw * cmt * ycc), z3 = -(dsw * cmt * zcc);
        fxa += ma * x3;
        fya += ma * y3;
        fza += ma * z3;
        if (fwd && shifted && ai == 0) {
          fgx += x3;
          fgy += y3;
          fgz += z3;
        }
      }
      if (fwd && shifted) {
        st[0 * 64 + lane] -= fgx * sxm; st[1 * 64 + lane] -= fgx * sym; st[2 * 64 + lane] -= fgx * szm;
        st[3 * 64 + lane] -= fgy * sxm; st[4 * 64 + lane] -= fgy * sym; st[5 * 64 + lane] -= fgy * szm;
        st[6 * 64 + lane] -= fgz * sxm; st[7 * 64 + lane] -= fgz * sym; st[8 * 64 + lane] -= fgz * szm;
      }
    }
  }
  // owner-atom forces: lanes ai, ai + nI, ... summed in lane order
  facc[wid][lane][0] = fxa;
  facc[wid][lane][1] = fya;
  facc[wid][lane][2] = fza;
  __syncthreads();
  if (t < d.ngrp && lane < nI) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int s = 0; s < nslot; ++s) {
      sx += facc[wid][s * nI + lane][0];
      sy += facc[wid][s * nI + lane][1];
      sz += facc[wid][s * nI + lane][2];
    }
    const int a = ia0 + lane;
    d.fx[a] = sx;
    d.fy[a] = sy;
    d.fz[a] = sz;
    const int pa = d.atomp[a];
    const double cx = d.xpcm[pa], cy = d.ypcm[pa], cz = d.zpcm[pa];
    st[0 * 64 + lane] += sx * cx; st[1 * 64 + lane] += sx * cy; st[2 * 64 + lane] += sx * cz;
    st[3 * 64 + lane] += sy * cx; st[4 * 64 + lane] += sy * cy; st[5 * 64 + lane] += sy * cz;
    st[6 * 64 + lane] += sz * cx; st[7 * 64 + lane] += sz * cy; st[8 * 64 + lane] += sz * cz;
  }
  // energies and virial: wave then block, fixed order
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    double v = wave_sum(acc[wid][k][lane]);
    if (lane == 0) red[wid][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 15) {
    double s = 0.0;
    for (int w = 0; w < PAIR_WAVES; ++w) s += red[w][threadIdx.x];
    d.part[(size_t)blk * 16 + threadIdx.x] = s;
  }
}

// item l of nitems = nsys x nblkmax (system l / nblkmax), optionally XCD-grouped
__device__ __forceinline__ void pairs_item_index(int h, int nitems, int nblkmax, int remap, int &sys, int &blk) {
  const int l = remap ? xcd_remap(h, nitems) : h;
  sys = l / nblkmax;
  blk = l - sys * nblkmax;
}

// W = minimum waves per SIMD the register allocation must allow (occupancy target)
template <typename T, int W>
__global__ void __launch_bounds__(256, W) k_pairs(const PairDesc *__restrict__ D, int remap) {
  __shared__ PairShared sh;
  int sys, blk;
  pairs_item_index(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y, gridDim.x, remap, sys, blk);
  pairs_item<T>(D[sys], blk, sh);
}

// Persistent form: a fixed grid (a multiple of 8 workgroups, so each keeps to
// one XCD) walks all items of the batch with a grid stride.
template <typename T, int W>
__global__ void __launch_bounds__(256, W) k_pairs_persistent(const PairDesc *__restrict__ D, int remap, int nitems,
                                                             int nblkmax) {
  __shared__ PairShared sh;
  for (int h = blockIdx.x; h < nitems; h += gridDim.x) {
    int sys, blk;
    pairs_item_index(h, nitems, nblkmax, remap, sys, blk);
    pairs_item<T>(D[sys], blk, sh);
    __syncthreads();
  }
}

// d.out[15]: uconf[3], ucoul[3] (slt, slt-slv, slv order of typeij), st[9] row-major;
// block k reduces component k over the per-block partials in fixed order
__global__ void __launch_bounds__(256) k_pairs_finalize(const PairDesc *__restrict__ D) {
  __shared__ double red[4];
  const PairDesc &d = D[blockIdx.y];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < d.nblk; b += blockDim.x) s += d.part[(size_t)b * 16 + k];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    d.out[k] = t;
  }
}

// Occupancy target of k_pairs: measured best on MI355X for the 24.5k-atom
// system is 3 waves/SIMD for FP64 (4 spills) and 4 for FP32.  The tuning
// value, then ORAC_HIP_PAIR_WAVES=2|3|4, override it.
template <typename T>
static int pair_occupancy(int tuned) {
  if (tuned == 2 || tuned == 3 || tuned == 4) return tuned;
  static int w = [] {
    const char *e = getenv("ORAC_HIP_PAIR_WAVES");
    int v = e ? atoi(e) : (sizeof(T) == 8 ? 3 : 4);
    return (v == 2 || v == 3 || v == 4) ? v : 3;
  }();
  return w;
}

template <typename T>
void launch_pairs(const PairDesc *d_desc, const PairDesc *h, int nsys, hipStream_t st, const Tuning &tu,
                  hipEvent_t kbeg, hipEvent_t kend) {
  int max_grp = 1, max_blk = 1;
  for (int s = 0; s < nsys; ++s) {
    if (h[s].max_group > 64)
      throw Error(E_UNSUPPORTED, "groups larger than 64 atoms are not supported by the pair kernel");
    max_grp = std::max(max_grp, h[s].ngrp);
    max_blk = std::max(max_blk, h[s].nblk);
  }
  hipLaunchKernelGGL(k_pack_sorted, dim3((max_grp + 255) / 256, nsys), dim3(256), 0, st, d_desc);
  dim3 grd(max_blk, nsys), blk(64 * PAIR_WAVES);
  const int remap = tu.xcd_remap ? 1 : 0;
  const int waves = pair_occupancy<T>(tu.pair_waves);
  const int nitems = max_blk * nsys;
  if (kbeg) ORAC_HIP(hipEventRecord(kbeg, st));
  if (tu.persistent) {
    // CUs x workgroups per CU at the occupancy target (4 waves per workgroup)
    static int ncu = [] {
      int dev = 0, n = 0;
      if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                   hipSuccess)
        n = 256;
      return n > 0 ? n : 256;
    }();
    const int g = std::max(8, std::min(nitems, ncu * waves) / 8 * 8);
    switch (waves) {
      case 3: hipLaunchKernelGGL((k_pairs_persistent<T, 3>), dim3(g), blk, 0, st, d_desc, remap, nitems, max_blk); break;
      case 4: hipLaunchKernelGGL((k_pairs_persistent<T, 4>), dim3(g), blk, 0, st, d_desc, remap, nitems, max_blk); break;
      default:
        hipLaunchKernelGGL((k_pairs_persistent<T, 2>), dim3(g), blk, 0, st, d_desc, remap, nitems, max_blk);
        break;
    }
  } else {
    switch (waves) {
      case 3: hipLaunchKernelGGL((k_pairs<T, 3>), grd, blk, 0, st, d_desc, remap); break;
      case 4: hipLaunchKernelGGL((k_pairs<T, 4>), grd, blk, 0, st, d_desc, remap); break;
      default: hipLaunchKernelGGL((k_pairs<T, 2>), grd, blk, 0, st, d_desc, remap); break;
    }
  }
  if (kend) ORAC_HIP(hipEventRecord(kend, st));
  hipLaunchKernelGGL(k_pairs_finalize, dim3(15, nsys), dim3(256), 0, st, d_desc);
}

template void launch_pairs<double>(const PairDesc *, const PairDesc *, int, hipStream_t, const Tuning &, hipEvent_t,
                                   hipEvent_t);
template void launch_pairs<float>(const PairDesc *, const PairDesc *, int, hipStream_t, const Tuning &, hipEvent_t,
                                  hipEvent_t);

}  // namespace orac
