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
#include <string>

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
  const int g = d.perm ? gmem(d.perm)[t] : t;
  const int f0 = gmem(d.gfirst)[g], n = gmem(d.glast)[g] - f0 + 1;
  const int sf = d.soff ? gmem(d.soff)[t] : f0;
  gstore(d.sgp + t, make_double4(gmem(d.xpg)[g], gmem(d.ypg)[g], gmem(d.zpg)[g], 0.0));
  gstore(d.sgi + t, make_int4(sf, f0, n | (gmem(d.gss)[g] << 8), g));
  const double gx = gmem(d.xpg)[g], gy = gmem(d.ypg)[g], gz = gmem(d.zpg)[g];
  for (int k = 0; k < n; ++k) {
    const int a = f0 + k;
    const double x = gmem(d.xp)[a], y = gmem(d.yp)[a], z = gmem(d.zp)[a], q = gmem(d.q)[a];
    gstore(d.sap + sf + k, make_double4(x, y, z, q));
    gstore(d.sai + sf + k, make_int2(gmem(d.type)[a], gmem(d.ex_inter)[a] ? 1 : 0));
    if (d.sapc) {
      // FP64 Cartesian offset from the group centre (upper triangle of co, as force.f)
      const double dx = x - gx, dy = y - gy, dz = z - gz;
      gstore(d.sapc + sf + k, make_double4(d.co[0] * dx + d.co[3] * dy + d.co[6] * dz, d.co[4] * dy + d.co[7] * dz,
                                           d.co[8] * dz, q));
    }
    if (d.sapf) {
      // Cartesian offset from the group centre: a pair's separation is the
      // group separation (FP64) plus the difference of two small offsets
      const double dx = x - gx, dy = y - gy, dz = z - gz;
      gstore(d.sapf + sf + k, make_float4((float)(d.co[0] * dx + d.co[3] * dy + d.co[6] * dz),
                                          (float)(d.co[4] * dy + d.co[7] * dz), (float)(d.co[8] * dz), (float)q));
    }
  }
}

template <typename P>
__device__ inline bool has_partner_in(P ex_off, P ex_idx, int a, int lo, int hi) {
  for (int t = ex_off[a]; t < ex_off[a + 1]; ++t) {
    int p = ex_idx[t];
    if (p >= lo && p <= hi) return true;
  }
  return false;
}

template <typename P>
__device__ inline bool is_excluded(P ex_off, P ex_idx, int a, int b) {
  for (int t = ex_off[a]; t < ex_off[a + 1]; ++t)
    if (ex_idx[t] == b) return true;
  return false;
}

// Pair arithmetic built on one reciprocal square root per pair.  FP64: the
// hardware reciprocal square root and reciprocal seeds refined by one
// Newton-Raphson step each (~1e-14 relative; the reference's 1/sqrt and
// divides are correctly rounded, and the difference stays far below the
// force tolerance) for a fraction of the instructions of an IEEE divide and
// square root, and exp by a two-constant ln 2 reduction and a degree-10
// Taylor series (~3e-13 relative) in place of the libm routine.  FP32
// (mixed precision): the hardware reciprocal, reciprocal square root and
// exponential, 1-2 ulp each.  Pair distances are positive and finite, so no
// special cases arise.
template <typename T>
__device__ inline T pair_rsqrt(T x);
template <>
__device__ inline double pair_rsqrt<double>(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  return fma(r, fma(-0.5 * x, r * r, 0.5), r);
}
template <>
__device__ inline float pair_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <typename T>
__device__ inline T pair_rcp(T x);
template <>
__device__ inline double pair_rcp<double>(double x) {
  const double q = __builtin_amdgcn_rcp(x);
  return fma(q, fma(-x, q, 1.0), q);
}
template <>
__device__ inline float pair_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <typename T>
__device__ inline T pair_exp(T x);
// exp(x) for the Ewald factor exp(-(alpha r)^2) (x <= 0, modest): x = n ln2 +
// r, |r| <= ln2 / 2 (ln 2 split in two constants, the reduction by FMA), e^r
// by its Taylor series to degree 10 (truncation below 3e-13), scaled by 2^n
template <>
__device__ inline double pair_exp<double>(double x) {
  const double n = __builtin_rint(x * 1.4426950408889634);
  double r = fma(-n, 0.6931471805599453, x);
  r = fma(-n, 2.3190468138462996e-17, r);
  double p = 2.755731922398589e-07;
  p = fma(p, r, 2.7557319223985893e-06);
  p = fma(p, r, 2.48015873015873e-05);
  p = fma(p, r, 0.0001984126984126984);
  p = fma(p, r, 0.001388888888888889);
  p = fma(p, r, 0.008333333333333333);
  p = fma(p, r, 0.041666666666666664);
  p = fma(p, r, 0.16666666666666666);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return __builtin_amdgcn_ldexp(p, (int)n);
}
template <>
__device__ inline float pair_exp<float>(float x) { return __expf(x); }

// The pair (a, b) is masked: excluded (the symmetric exclusion list), or
// both atoms at or above massinfty (force.f:488-489).  For the reverse
// direction the owner group must first hold one of b's partners.
__device__ __noinline__ bool pair_masked(const PairDesc &d, int a, int b, int ia0, int nI, bool fwd) {
  if (!fwd && !has_partner_in(gmem(d.ex_off), gmem(d.ex_idx), b, ia0, ia0 + nI - 1)) return false;
  if (is_excluded(gmem(d.ex_off), gmem(d.ex_idx), a, b)) return true;
  return d.any_heavy && !(gmem(d.mass)[b] < d.massinfty || gmem(d.mass)[a] < d.massinfty);
}

__device__ inline void lj_params(const PairDesc &d, int ti, int tj, double &c12, double &c6) {
  if (d.lj_dense) {
    double2 v = gload_at(d.lj_dense, ti * d.ntypes + tj);
    c12 = v.x;
    c6 = v.y;
  } else {
    int lij = gmem(d.type_table)[ti + (size_t)d.ntypes * tj];
    c12 = gmem(d.ecc12)[lij];
    c6 = gmem(d.ecc6)[lij];
  }
}

// Group-pair geometry (force.f:265-292): the image shift s (even integers in
// reduced units) and the centre separation after it, without contraction.
struct GroupGeom {
  double sxm, sym, szm, xcc, ycc, zcc, drj;
};
__device__ __forceinline__ GroupGeom group_geom(double xpgi, double ypgi, double zpgi, const double4 &gpj,
                                                const double *co) {
#pragma clang fp contract(off)
  GroupGeom g;
  double xgg = xpgi - gpj.x, ygg = ypgi - gpj.y, zgg = zpgi - gpj.z;
  g.sxm = 2.0 * fnint(0.5 * xgg);
  g.sym = 2.0 * fnint(0.5 * ygg);
  g.szm = 2.0 * fnint(0.5 * zgg);
  xgg = xgg - g.sxm;
  ygg = ygg - g.sym;
  zgg = zgg - g.szm;
  g.xcc = co[0] * xgg + co[3] * ygg + co[6] * zgg;
  g.ycc = co[4] * ygg + co[7] * zgg;
  g.zcc = co[8] * zgg;
  g.drj = g.xcc * g.xcc + g.ycc * g.ycc + g.zcc * g.zcc;
  return g;
}

// the group pair is inside the shell: normal [rinn0, rout) or switched
// [rout, rout0) and [rinn, rinn0) (force.f:293-335)
__device__ __forceinline__ bool in_shell(const PairDesc &d, double drj) {
  return (drj >= d.r2inn0 && drj < d.r2out) || (drj >= d.r2out && drj < d.r2out0) ||
         (drj >= d.r2inn && drj < d.r2inn0);
}

// One atom pair of a group pair (force.f:420-560): the LJ term when the pair
// has one (c6 > 1e-12), the A&S erfc Ewald term or plain Coulomb, as the
// scalar force factor qf (force = qf * r), the separation and the two energies.
struct PairTerms {
  double qf, xc, yc, zc, conf, cou;
  __device__ __forceinline__ void add(double &uconf, double &ucoul, double &fx, double &fy, double &fz) const {
    uconf += conf;
    ucoul += cou;
    fx += qf * xc;
    fy += qf * yc;
    fz += qf * zc;
  }
};

// The pair at Cartesian separation (xc, yc, zc) with charge product qq.
template <typename T>
__device__ __forceinline__ PairTerms pair_terms_c(const PairDesc &d, T xc, T yc, T zc, T qq, double c12, double c6);

template <typename T>
__device__ __forceinline__ PairTerms pair_terms(const PairDesc &d, double pxs, double pys, double pzs, double qa,
                                                const double *co, const double4 &pb4, double c12, double c6) {
  const double xg = pxs - pb4.x, yg = pys - pb4.y, zg = pzs - pb4.z;
  return pair_terms_c<T>(d, T(co[0] * xg + co[3] * yg + co[6] * zg), T(co[4] * yg + co[7] * zg), T(co[8] * zg),
                         T(qa * pb4.w), c12, c6);
}

template <typename T>
__device__ __forceinline__ PairTerms pair_terms_c(const PairDesc &d, T xc, T yc, T zc, T qq, double c12, double c6) {
  const T rsq = xc * xc + yc * yc + zc * zc;
  const T rinv = pair_rsqrt<T>(rsq);
  const T rsqi = rinv * rinv;
  // LJ by predication: a pair without LJ takes zero parameters
  const bool lj = c6 > 1e-12;
  const T c12t = lj ? T(c12) : T(0), c6t = lj ? T(c6) : T(0);
  const T c12f = lj ? T(12.0 * c12) : T(0), c6f = lj ? T(6.0 * c6) : T(0);
  const T r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
  T qforce = (c12f * r12 - c6f * r6) * rsqi;
  PairTerms r;
  r.conf = lj ? (double)(c12t * r12 - c6t * r6) : 0.0;
  r.cou = 0.0;
  if (!d.lskip_ewald) {
    const T rsp = rsq * rinv;
    const T rspqi = rsqi * rinv;
    const T alphar = T(d.alpha) * rsp;
    const T qt = pair_rcp<T>(T(1) + T(AS_QP) * alphar);
    const T expcst = pair_exp<T>(-(alphar * alphar));
    const T erfcst = ((((T(AS_A5) * qt + T(AS_A4)) * qt + T(AS_A3)) * qt + T(AS_A2)) * qt + T(AS_A1)) * qt * expcst;
    const T furpar = qq;
    r.cou = (double)(furpar * erfcst * rinv);
    qforce += furpar * (erfcst + T(d.twrtpi) * alphar * expcst) * rspqi;
  } else if (!d.ewald) {
    const T ssv = qq * rinv;
    r.cou = (double)ssv;
    qforce += ssv * rsqi;
  }
  r.qf = (double)qforce;
  r.xc = (double)xc;
  r.yc = (double)yc;
  r.zc = (double)zc;
  return r;
}

// One atom pair in soft-core form (force_alchemy.f:439-523): F = the pair's
// row of alch_fac, ap = {1 / sigma^6, epsmin, typeaij + 4} of its classes.
// LJ eps (gp^-2 - gp^-1) with gp = ljsoft f1 + r^6 / sigma^6, scaled by f2;
// Coulomb q_a q_b fq2 erfc(alpha r) / sqrt(qsoft fq1^2 + r^2) (or without
// erfc, plain); the energies at lambda0 come from the same distances with
// the lambda0 factors.  dw = U(lambda) - U(lambda0) of the pair.
struct AlchTerms {
  double qf, xc, yc, zc, conf, cou, dw;
  __device__ __forceinline__ void add(double &uconf, double &ucoul, double &dwa, double &fx, double &fy,
                                      double &fz) const {
    uconf += conf;
    ucoul += cou;
    dwa += dw;
    fx += qf * xc;
    fy += qf * yc;
    fz += qf * zc;
  }
};

template <typename T>
__device__ __forceinline__ AlchTerms alch_terms(const PairDesc &d, double pxs, double pys, double pzs, double qa,
                                                const double *co, const double4 &pb4, const double4 &ap,
                                                const double *F) {
  const double xg = pxs - pb4.x, yg = pys - pb4.y, zg = pzs - pb4.z;
  const T xc = T(co[0] * xg + co[3] * yg + co[6] * zg);
  const T yc = T(co[4] * yg + co[7] * zg);
  const T zc = T(co[8] * zg);
  const T rsq = xc * xc + yc * yc + zc * zc;
  const T rinv = pair_rsqrt<T>(rsq);
  T qforce = T(0), conf = T(0), conf0 = T(0), cou = T(0), cou0 = T(0);
  const T qq = T(qa * pb4.w);
  if (d.alch_shifted) {
    // force_alchemy_shifted.f:469-509: LJ at r + ljshift 2^(1/6) sigma f1, Coulomb over
    // r + qshift fq1 (erfc of the real distance)
    const T r = rsq * rinv;
    if (ap.y > 1e-12) {
      const T sig = T(ap.w), eps = T(ap.y);
      const T irl = pair_rcp<T>(r + T(F[0]) * sig), irl0 = pair_rcp<T>(r + T(F[2]) * sig);
      const T x = sig * irl, x2 = x * x, r6 = x2 * x2 * x2;
      const T y = sig * irl0, y2 = y * y, r60 = y2 * y2 * y2;
      qforce = T(F[1]) * eps * (T(12) * r6 * r6 - T(6) * r6) * (rinv * irl);
      conf = T(F[1]) * (eps * (r6 * r6 - r6));
      conf0 = T(F[3]) * (eps * (r60 * r60 - r60));
    }
    const T irq = pair_rcp<T>(r + T(F[4])), irq0 = pair_rcp<T>(r + T(F[6]));
    if (!d.lskip_ewald) {
      const T alphar = T(d.alpha) * r;
      const T qt = pair_rcp<T>(T(1) + T(AS_QP) * alphar);
      const T expcst = pair_exp<T>(-(alphar * alphar));
      const T erfcst = ((((T(AS_A5) * qt + T(AS_A4)) * qt + T(AS_A3)) * qt + T(AS_A2)) * qt + T(AS_A1)) * qt * expcst;
      const T furpar = T(F[5]) * qq;
      cou = furpar * erfcst * irq;
      cou0 = T(F[7]) * qq * erfcst * irq0;
      qforce += furpar * (rinv * irq) * (erfcst * irq + T(d.twrtpi) * alphar * expcst * rinv);
    } else if (!d.ewald) {
      cou = T(F[5]) * qq * irq;
      cou0 = T(F[7]) * qq * irq0;
      qforce += cou * (irq * rinv);
    }
  } else {
    const T auxaa = pair_rsqrt<T>(T(F[4]) + rsq);
    const T auxa0 = pair_rsqrt<T>(T(F[6]) + rsq);
    if (ap.y > 1e-12) {
      const T r6 = rsq * rsq * rsq * T(ap.x);
      const T gp1 = pair_rcp<T>(T(F[0]) + r6), gp2 = gp1 * gp1;
      const T gp10 = pair_rcp<T>(T(F[2]) + r6), gp20 = gp10 * gp10;
      const T eps = T(ap.y);
      const T ssvir = T(F[1]) * eps * (T(2) * gp1 * gp2 - gp2);
      qforce = T(6) * r6 * ssvir * (rinv * rinv);
      conf = T(F[1]) * (eps * (gp2 - gp1));
      conf0 = T(F[3]) * (eps * (gp20 - gp10));
    }
    if (!d.lskip_ewald) {
      const T alphar = T(d.alpha) * (rsq * rinv);
      const T qt = pair_rcp<T>(T(1) + T(AS_QP) * alphar);
      const T expcst = pair_exp<T>(-(alphar * alphar));
      const T erfcst = ((((T(AS_A5) * qt + T(AS_A4)) * qt + T(AS_A3)) * qt + T(AS_A2)) * qt + T(AS_A1)) * qt * expcst;
      const T furpar = T(F[5]) * qq, furpar0 = T(F[7]) * qq;
      cou = furpar * erfcst * auxaa;
      cou0 = furpar0 * erfcst * auxa0;
      qforce += furpar * auxaa * (erfcst * (auxaa * auxaa) + T(d.twrtpi) * alphar * expcst * (rinv * rinv));
    } else if (!d.ewald) {
      cou = T(F[5]) * qq * auxaa;
      cou0 = T(F[7]) * qq * auxa0;
      qforce += cou * (auxaa * auxaa);
    }
  }
  AlchTerms r;
  r.qf = (double)qforce;
  r.xc = (double)xc;
  r.yc = (double)yc;
  r.zc = (double)zc;
  r.conf = (double)conf;
  r.cou = (double)cou;
  r.dw = (double)(conf - conf0) + (double)(cou - cou0);
  return r;
}

constexpr int PAIR_NACC = 16;      // per-lane sums: uconf[3], ucoul[3], st[9], dwrk

struct PairShared {
  double red[PAIR_WAVES][PAIR_NACC];
  double facc[PAIR_WAVES][64][3];
  double acc[PAIR_WAVES][PAIR_NACC][64];  // per-lane energy and virial sums (lane-private slots)
  double alch[9][8];                      // the system's soft-core factors (alchemical kernels)
};

// One work item: block blk of system d (PAIR_WAVES owner groups).  ALCH:
// the kernel also takes the soft-core form for systems that carry alch_pair.
template <typename T, bool ALCH>
__device__ __forceinline__ void pairs_item(const PairDesc &d, int blk, PairShared &sh) {
  if (blk >= d.nblk || d.half) return;
  double(&red)[PAIR_WAVES][PAIR_NACC] = sh.red;
  double(&facc)[PAIR_WAVES][64][3] = sh.facc;
  double(&acc)[PAIR_WAVES][PAIR_NACC][64] = sh.acc;
  const bool alch = ALCH && d.alch_pair != nullptr;
  if (alch) {
    if (threadIdx.x < 72) sh.alch[threadIdx.x >> 3][threadIdx.x & 7] = gmem(d.alch_fac)[threadIdx.x];
    __syncthreads();
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // S waves share one owner group, each taking a contiguous part of its list
  const int S = d.split, part = wid % S;
  const int t = blk * (PAIR_WAVES / S) + wid / S;  // owner group, sorted position
  double fxa = 0.0, fya = 0.0, fza = 0.0;
  double *en = &acc[wid][0][0];  // en[k * 64 + lane]: uconf[typeij], ucoul[typeij]
  double *st = &acc[wid][6][0];  // st[k * 64 + lane]
#pragma unroll
  for (int k = 0; k < PAIR_NACC; ++k) acc[wid][k][lane] = 0.0;
  int ia0 = 0, nI = 1, nslot = 64;  // ia0: owner's first atom, original numbering
  if (t < d.ngrp) {
    const int4 gi = gload(d.sgi + t);
    const int g = gi.w;
    ia0 = gi.y;
    nI = gi.z & 255;
    nslot = 64 / nI;
    const int slot = lane / nI, ai = lane - slot * nI;
    const bool active = slot < nslot;
    const int a = ia0 + (active ? ai : 0);        // original index
    const int as = gi.x + (active ? ai : 0);      // sorted index
    const double4 pa4 = gload(d.sap + as);
    const int2 ia4 = gload(d.sai + as);
    // mixed precision reads compact FP32 records (the soft-core form keeps the FP64 ones)
    constexpr bool FREC = sizeof(T) == 4;
    const float4 paf = (FREC && !alch) ? gload(d.sapf + as) : make_float4(0.f, 0.f, 0.f, 0.f);
    const double ma = gmem(d.gmass)[a];
    const int typei = gi.z >> 8;
    const double4 gpi = gload(d.sgp + t);
    const double xpgi = gpi.x, ypgi = gpi.y, zpgi = gpi.z;
    double co[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) co[k] = d.co[k];
    const double co0 = co[0], co3 = co[3], co6 = co[6], co4 = co[4], co7 = co[7], co8 = co[8];
    int nb0, nb1;
    if (d.nb_slab > 0) {
      nb0 = t * d.nb_slab;
      nb1 = nb0 + min(gmem(d.nb_cnt)[t], d.nb_slab);
    } else {
      nb0 = gmem(d.nb_off)[t];
      nb1 = gmem(d.nb_off)[t + 1];
    }
    if (S > 1) {
      const int len = nb1 - nb0;
      nb1 = nb0 + (len * (part + 1)) / S;
      nb0 = nb0 + (len * part) / S;
    }
    // The list reaches rc + rtol + rnei; only groups inside the shell do work.
    // Entries are classified 64 at a time and the survivors compacted, in list
    // order, into this wave's part of facc (unused until the force reduction),
    // so every slot of the pair loop is busy.
    int *cbuf = reinterpret_cast<int *>(&facc[wid][0][0]);
    constexpr int CBUF = (int)(sizeof(facc[0]) / sizeof(int));
    for (int c0 = nb0; c0 < nb1;) {
      int nc = 0;
      for (; c0 < nb1 && nc <= CBUF - 64; c0 += 64) {
        const int jj = c0 + lane;
        const bool in = jj < nb1;
        const int tj = in ? gmem(d.nb_idx)[jj] : t;
        const GroupGeom gg = group_geom(xpgi, ypgi, zpgi, gload_at(d.sgp, tj), co);
        const bool keep = in && in_shell(d, gg.drj);
        const unsigned long long m = __ballot(keep);
        if (keep) cbuf[nc + __popcll(m & ((1ull << lane) - 1ull))] = tj;
        nc += __popcll(m);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int base = 0; base < nc; base += nslot) {
        const int jj = base + slot;
        const bool have = active && jj < nc;
        const int tj = have ? cbuf[jj] : t;
        const double4 gpj = gload_at(d.sgp, tj);
        const int4 gij = gload_at(d.sgi, tj);
        const int j = gij.w;
        const GroupGeom gg = group_geom(xpgi, ypgi, zpgi, gpj, co);
        const double sxm = gg.sxm, sym = gg.sym, szm = gg.szm, xcc = gg.xcc, ycc = gg.ycc, zcc = gg.zcc,
                     drj = gg.drj;
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
          const bool mask_a = fwd && (ia4.y & 1) && has_partner_in(gmem(d.ex_off), gmem(d.ex_idx), a, jb0, jb1);
          // owner position minus the image shift: one association for every class
          const double pxs = pa4.x - sxm, pys = pa4.y - sym, pzs = pa4.z - szm;
          // unswitched sums over the neighbour's atoms; the switch (a function of
          // the group distance only) scales them once below
          double ucoula = 0.0, uconfa = 0.0, dwa = 0.0;
          double sega = 0.0, cmr = 0.0;  // REM: segment energy, V with the selected pairs scaled
          if (alch) {
            for (int k = 0; k < nJ; ++k) {
              const int2 ib0 = gload_at(d.sai, jbs + k);
              const double4 pb0 = gload_at(d.sap, jbs + k);
              const double4 ap = gload_at(d.alch_pair, ia4.x * d.ntypes + ib0.x);
              bool skip0 = d.lskip_ewald && ap.y < 1e-12 && d.ewald;
              if (fwd ? mask_a : (ib0.y & 1) != 0) skip0 = skip0 || pair_masked(d, a, jb0 + k, ia0, nI, fwd);
              const AlchTerms r0 = alch_terms<T>(d, pxs, pys, pzs, pa4.w, co, pb0, ap, sh.alch[(int)ap.z]);
              if (!skip0) r0.add(uconfa, ucoula, dwa, fgx, fgy, fgz);
            }
          } else if constexpr (FREC) {
            // compact records: separation = group separation (FP64) + offset difference
            for (int k = 0; k < nJ; ++k) {
              const int2 ib0 = gload_at(d.sai, jbs + k);
              const float4 pb0 = gload_at(d.sapf, jbs + k);
              double c12a, c6a;
              if (d.lj_dense32) {
                const float2 lj = gload_at(d.lj_dense32, ia4.x * d.ntypes + ib0.x);
                c12a = lj.x;
                c6a = lj.y;
              } else {
                lj_params(d, ia4.x, ib0.x, c12a, c6a);
              }
              bool skip0 = d.lskip_ewald && c6a < 1e-12 && d.ewald;
              if (fwd ? mask_a : (ib0.y & 1) != 0) skip0 = skip0 || pair_masked(d, a, jb0 + k, ia0, nI, fwd);
              const PairTerms r0 =
                  pair_terms_c<T>(d, T(gg.xcc + (double)(paf.x - pb0.x)), T(gg.ycc + (double)(paf.y - pb0.y)),
                                  T(gg.zcc + (double)(paf.z - pb0.z)), T(paf.w * pb0.w), c12a, c6a);
              if (!skip0) r0.add(uconfa, ucoula, fgx, fgy, fgz);
            }
          } else if (d.rematom) {
            // force_REM.f:491-494, 649-664, 679-682 (Ewald branch): a selected
            // pair's force is scaled by rem_f3 and its energy goes to the
            // segment energy; its share of V for the switch derivative is
            // scaled too (cmap2), the shell energies are not
            const bool ra = gmem(d.rematom)[a] != 0;
            for (int k = 0; k < nJ; ++k) {
              const int2 ib0 = gload_at(d.sai, jbs + k);
              const double4 pb0 = gload_at(d.sap, jbs + k);
              double c12a, c6a;
              lj_params(d, ia4.x, ib0.x, c12a, c6a);
              bool skip0 = false;
              if (fwd ? mask_a : (ib0.y & 1) != 0) skip0 = pair_masked(d, a, jb0 + k, ia0, nI, fwd);
              const PairTerms r0 = pair_terms<T>(d, pxs, pys, pzs, pa4.w, co, pb0, c12a, c6a);
              if (skip0) continue;
              const bool rb = gmem(d.rematom)[jb0 + k] != 0;
              const bool sel = d.rem_kind == 0 ? (ra || rb) : (d.rem_kind == 1 ? (ra && rb) : (ra != rb));
              const double f = sel ? d.rem_f3 : 1.0;
              uconfa += r0.conf;
              ucoula += r0.cou;
              fgx += (f * r0.qf) * r0.xc;
              fgy += (f * r0.qf) * r0.yc;
              fgz += (f * r0.qf) * r0.zc;
              if (sel) sega += r0.conf + r0.cou;
              cmr += f * r0.conf + f * r0.cou;
            }
          } else {
            for (int k = 0; k < nJ; ++k) {
              const int2 ib0 = gload_at(d.sai, jbs + k);
              const double4 pb0 = gload_at(d.sap, jbs + k);
              double c12a, c6a;
              lj_params(d, ia4.x, ib0.x, c12a, c6a);
              bool skip0 = d.lskip_ewald && c6a < 1e-12 && d.ewald;
              if (fwd ? mask_a : (ib0.y & 1) != 0) skip0 = skip0 || pair_masked(d, a, jb0 + k, ia0, nI, fwd);
              const PairTerms r0 = pair_terms<T>(d, pxs, pys, pzs, pa4.w, co, pb0, c12a, c6a);
              if (!skip0) r0.add(uconfa, ucoula, fgx, fgy, fgz);
            }
          }
          const double s = cls == 2 ? sw : 1.0;
          fgx *= s;
          fgy *= s;
          fgz *= s;
          fxa += fgx;
          fya += fgy;
          fza += fgz;
          if (fwd) {
            en[typeij * 64 + lane] += s * uconfa;
            en[(3 + typeij) * 64 + lane] += s * ucoula;
            if (ALCH) acc[wid][15][lane] += s * dwa;
            if (!ALCH && FREC == false && d.rematom) acc[wid][15][lane] += s * sega;
          }
          if (cls == 2) cm = (!ALCH && !FREC && d.rematom) ? cmr : uconfa + ucoula;
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
      // every lane is done with this chunk before the next fill
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  // owner-atom forces: the owner's waves in part order, in each the lanes
  // ai, ai + nI, ... in lane order
  facc[wid][lane][0] = fxa;
  facc[wid][lane][1] = fya;
  facc[wid][lane][2] = fza;
  __syncthreads();
  if (t < d.ngrp && part == 0 && lane < nI) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int p = 0; p < S; ++p)
      for (int s = 0; s < nslot; ++s) {
        sx += facc[wid + p][s * nI + lane][0];
        sy += facc[wid + p][s * nI + lane][1];
        sz += facc[wid + p][s * nI + lane][2];
      }
    const int a = ia0 + lane;
    gmem(d.fx)[a] = sx;
    gmem(d.fy)[a] = sy;
    gmem(d.fz)[a] = sz;
    const int pa = gmem(d.atomp)[a];
    const double cx = gmem(d.xpcm)[pa], cy = gmem(d.ypcm)[pa], cz = gmem(d.zpcm)[pa];
    st[0 * 64 + lane] += sx * cx; st[1 * 64 + lane] += sx * cy; st[2 * 64 + lane] += sx * cz;
    st[3 * 64 + lane] += sy * cx; st[4 * 64 + lane] += sy * cy; st[5 * 64 + lane] += sy * cz;
    st[6 * 64 + lane] += sz * cx; st[7 * 64 + lane] += sz * cy; st[8 * 64 + lane] += sz * cz;
  }
  // energies and virial: wave then block, fixed order
#pragma unroll
  for (int k = 0; k < PAIR_NACC; ++k) {
    double v = wave_sum(acc[wid][k][lane]);
    if (lane == 0) red[wid][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < PAIR_NACC) {
    double s = 0.0;
    for (int w = 0; w < PAIR_WAVES; ++w) s += red[w][threadIdx.x];
    gmem(d.part)[(size_t)blk * 16 + threadIdx.x] = s;
  }
}

// item l of nitems = nsys x nblkmax (system l / nblkmax), optionally XCD-grouped
__device__ __forceinline__ void pairs_item_index(int h, int nitems, int nblkmax, int remap, int &sys, int &blk) {
  const int l = remap ? xcd_remap(h, nitems) : h;
  sys = l / nblkmax;
  blk = l - sys * nblkmax;
}

// W = minimum waves per SIMD the register allocation must allow (occupancy target)
template <typename T, int W, bool ALCH>
__global__ void __launch_bounds__(256, W) k_pairs(const PairDesc *__restrict__ D, int remap) {
  __shared__ PairShared sh;
  int sys, blk;
  pairs_item_index(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y, gridDim.x, remap, sys, blk);
  pairs_item<T, ALCH>(D[sys], blk, sh);
}

// Persistent form: a fixed grid (a multiple of 8 workgroups, so each keeps to
// one XCD) walks all items of the batch with a grid stride.
template <typename T, int W, bool ALCH>
__global__ void __launch_bounds__(256, W) k_pairs_persistent(const PairDesc *__restrict__ D, int remap, int nitems,
                                                             int nblkmax) {
  __shared__ PairShared sh;
  for (int h = blockIdx.x; h < nitems; h += gridDim.x) {
    int sys, blk;
    pairs_item_index(h, nitems, nblkmax, remap, sys, blk);
    pairs_item<T, ALCH>(D[sys], blk, sh);
    __syncthreads();
  }
}

template <typename T, int W, bool ALCH>
static void launch_pair_kernel(bool persistent, dim3 grd, int g, hipStream_t st, const PairDesc *d_desc, int remap,
                               int nitems, int max_blk) {
  const dim3 blk(64 * PAIR_WAVES);
  if (persistent)
    hipLaunchKernelGGL((k_pairs_persistent<T, W, ALCH>), dim3(g), blk, 0, st, d_desc, remap, nitems, max_blk);
  else
    hipLaunchKernelGGL((k_pairs<T, W, ALCH>), grd, blk, 0, st, d_desc, remap);
}

// d.out[16]: uconf[3], ucoul[3] (slt, slt-slv, slv order of typeij), st[9] row-major,
// dwrk; block k reduces component k over the per-block partials in fixed order
__global__ void __launch_bounds__(256) k_pairs_finalize(const PairDesc *__restrict__ D) {
  __shared__ double red[4];
  const PairDesc &d = D[blockIdx.y];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < d.nblk; b += blockDim.x) s += gmem(d.part)[(size_t)b * 16 + k];
  if (d.half && k >= 6 && k < 15)  // the half form's molecular virial of the total forces
    for (int b = threadIdx.x; b < d.nblk; b += blockDim.x) s += gmem(d.part2)[(size_t)b * 9 + (k - 6)];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    d.out[k] = t;
  }
}

// ================================================================ owner-atom loop form
//
// k_pairs_own: one wavefront per owner group, as k_pairs, with the roles of
// the loop and the lanes exchanged.  The lanes take the atoms of the
// neighbour groups (one atom each; whole groups packed into a pass of 64
// lanes) and loop over the owner group's atoms.  The owner atoms are the same
// across the wavefront and come through scalar loads; each lane loads its
// neighbour atom's record and LJ entries once per pass, and the pair loop
// touches no memory.  A list entry's group geometry, shell class and switch
// are computed once, when the entries are classified, and kept in LDS for
// the passes.  Each lane sums the owner atoms' forces in registers across the
// whole list (owner groups of at most OWN_REG atoms; systems with larger
// groups use k_pairs) and the wavefront reduces them at the end in a fixed
// order.
//
// Pair separations are the Cartesian group separation (image applied, FP64)
// plus the difference of the two atoms' Cartesian offsets from their group
// centres, so each pair costs three additions before its distance.  FP64
// reciprocal, reciprocal square root and exponential are the hardware seeds
// refined by one Newton step (~1e-14 relative) and a degree-10 Taylor
// polynomial after a two-constant ln 2 reduction (~3e-13 relative).

// (OWN_REG, the owner atoms whose force sums stay in registers: water, single atoms, is in kernels.hpp)

struct OwnWave {
  int ent[64];                  // the round's in-shell entries: neighbour group (sorted position)
  int pb[64];                   // half form: their first neighbour-side force slot
  int aoff[65];                 // exclusive prefix of their atom counts (aoff[n] = total)
  int nj[64];
  double gx[64], gy[64], gz[64];  // Cartesian separation of the group centres, image applied
  double sx[64], sy[64], sz[64];  // the image shift (reduced units)
  double sw[64], dsw[64];         // switch S and S'(r) / r of the group distance (1 and 0 unswitched)
};

struct OwnShared {
  OwnWave w[PAIR_WAVES];
  double red[PAIR_WAVES][PAIR_NACC];
};

template <typename T>
__device__ __forceinline__ T own_rsqrt(T x);
template <>
__device__ __forceinline__ double own_rsqrt<double>(double x) { return pair_rsqrt<double>(x); }
template <>
__device__ __forceinline__ float own_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <typename T>
__device__ __forceinline__ T own_rcp(T x);
template <>
__device__ __forceinline__ double own_rcp<double>(double x) { return pair_rcp<double>(x); }
template <>
__device__ __forceinline__ float own_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <typename T>
__device__ __forceinline__ T own_exp(T x);
template <>
__device__ __forceinline__ double own_exp<double>(double x) { return pair_exp<double>(x); }
template <>
__device__ __forceinline__ float own_exp<float>(float x) { return __expf(x); }

// LJ {c12, c6} of the type pair (ti, tj), zero when c6 <= 1e-12 (the pair
// then has no LJ term, force.f:455)
template <typename T>
__device__ __forceinline__ void own_lj(const PairDesc &d, int ti, int tj, T &c12, T &c6) {
  double a, b;
  if (sizeof(T) == 4 && d.lj_dense32) {
    const float2 v = gload_at(d.lj_dense32, ti * d.ntypes + tj);
    a = v.x;
    b = v.y;
  } else {
    lj_params(d, ti, tj, a, b);
  }
  const bool lj = b > 1e-12;
  c12 = lj ? T(a) : T(0);
  c6 = lj ? T(b) : T(0);
}

// The owner-side and neighbour-side records of one atom: Cartesian offset
// from the group centre and the charge, in the pair precision.
template <typename T>
struct OwnRec {
  T x, y, z, q;
};
template <typename T>
__device__ __forceinline__ OwnRec<T> own_rec(const PairDesc &d, int s) {
  if constexpr (sizeof(T) == 4) {
    const float4 v = gload(d.sapf + s);
    return {v.x, v.y, v.z, v.w};
  } else {
    const double4 v = gload(d.sapc + s);
    return {v.x, v.y, v.z, v.w};
  }
}

template <typename T, bool HALF>
__device__ __forceinline__ void own_item(const PairDesc &d, int blk, OwnShared &sh) {
  if (blk >= d.nblk || (d.half != 0) != HALF) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  OwnWave &W = sh.w[wid];
  // the owner group: wave-uniform, so its records come through scalar loads
  const int t = __builtin_amdgcn_readfirstlane(blk * PAIR_WAVES + wid);
  double fr[OWN_REG][3];
#pragma unroll
  for (int u = 0; u < OWN_REG; ++u) fr[u][0] = fr[u][1] = fr[u][2] = 0.0;
  double x3x = 0.0, x3y = 0.0, x3z = 0.0;  // the switched group pairs' S'(r) V vectors (owner side)
  double ec[2] = {0.0, 0.0}, eq[2] = {0.0, 0.0};  // LJ and Coulomb energy by the neighbour group's class
  double st[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) st[k] = 0.0;
  int ia0 = 0, nI = 0, typei = 1, as0 = 0;
  if (t < d.ngrp) {
    const int4 gi = gload(d.sgi + t);
    const int g = gi.w;
    as0 = gi.x;
    ia0 = gi.y;
    nI = gi.z & 255;
    typei = gi.z >> 8;
    const double4 gpi = gload(d.sgp + t);
    int nb0, nb1;
    if (d.nb_slab > 0) {
      nb0 = t * d.nb_slab;
      nb1 = nb0 + min(gmem(d.nb_cnt)[t], d.nb_slab);
    } else {
      nb0 = gmem(d.nb_off)[t];
      nb1 = gmem(d.nb_off)[t + 1];
    }
    for (int c0 = nb0; c0 < nb1; c0 += 64) {
      // ---- classify 64 entries; the in-shell ones are kept in list order
      const int jj = c0 + lane;
      const bool in = jj < nb1;
      const int tj = in ? gmem(d.nb_idx)[jj] : t;
      const GroupGeom gg = group_geom(gpi.x, gpi.y, gpi.z, gload_at(d.sgp, tj), d.co);
      const bool keep = in && in_shell(d, gg.drj);
      const int nJin = in ? (gload_at(d.sgi, tj).z & 255) : 0;
      const int nJ = keep ? nJin : 0;
      int mslot = -1;
      if (HALF && in) {
        // every forward entry's slots are written: the pair forces, or zeros outside the shell
        mslot = gmem(d.nb_mirror)[jj];
        if (!keep && mslot >= 0)
          for (int k = 0; k < nJin; ++k) {
            gmem(d.jpx)[mslot + k] = 0.0;
            gmem(d.jpy)[mslot + k] = 0.0;
            gmem(d.jpz)[mslot + k] = 0.0;
          }
      }
      const unsigned long long m = __ballot(keep);
      const int pos = __popcll(m & ((1ull << lane) - 1ull)), nc = __popcll(m);
      int incl = nJ;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
      }
      if (keep) {
        double sw = 1.0, dsw = 0.0;
        if (gg.drj >= d.r2out && gg.drj < d.r2out0) {
          const double rsp = sqrt(gg.drj);
          const double auxa = (d.arsout1 + 2.0 * rsp) / d.arsout2;
          const double auxb = d.rout0 - rsp;
          sw = auxa * (auxb * auxb);
          dsw = (-(2.0 * auxa * auxb) + 2.0 * (auxb * auxb) / d.arsout2) / rsp;
        } else if (gg.drj >= d.r2inn && gg.drj < d.r2inn0) {
          const double rsp = sqrt(gg.drj);
          const double auxa = (d.arsinn1 + 2.0 * rsp) / d.arsinn2;
          const double auxb = d.rinn0 - rsp;
          sw = 1.0 - auxa * (auxb * auxb);
          dsw = (2.0 * auxa * auxb - 2.0 * (auxb * auxb) / d.arsinn2) / rsp;
        }
        W.ent[pos] = tj;
        if (HALF) W.pb[pos] = mslot;
        W.aoff[pos] = incl - nJ;
        W.nj[pos] = nJ;
        W.gx[pos] = gg.xcc;
        W.gy[pos] = gg.ycc;
        W.gz[pos] = gg.zcc;
        W.sx[pos] = gg.sxm;
        W.sy[pos] = gg.sym;
        W.sz[pos] = gg.szm;
        W.sw[pos] = sw;
        W.dsw[pos] = dsw;
      }
      if (lane == 63) W.aoff[nc] = incl;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // ---- passes: whole neighbour groups, one atom per lane
      for (int e0 = 0; e0 < nc;) {
        const int base = W.aoff[e0];
        const int ee = e0 + lane;
        const bool fits = ee < nc && W.aoff[ee] + W.nj[ee] - base <= 64;
        const int e1 = e0 + __popcll(__ballot(fits));
        // this lane's entry: the last e of the pass with aoff[e] - base <= lane
        int lo = e0, hi = e1;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (W.aoff[mid] - base <= lane) lo = mid;
          else hi = mid;
        }
        const int e = lo, k = lane - (W.aoff[e] - base);
        const bool have = k < W.nj[e];
        const double xcc = W.gx[e], ycc = W.gy[e], zcc = W.gz[e];
        const double sw = have ? W.sw[e] : 0.0, dsw = have ? W.dsw[e] : 0.0;
        double fsx = 0.0, fsy = 0.0, fsz = 0.0;  // this lane's pair forces on the owner group
        double fbx = 0.0, fby = 0.0, fbz = 0.0;  // half form: the forces on this lane's neighbour atom
        double vc = 0.0, vq = 0.0;               // this lane's unswitched LJ and Coulomb energy
        int typej = 1, bs = 0, bo = 0, nJl = 0;
        bool fwd = false;
        int2 ib = make_int2(0, 0);
        OwnRec<T> rb = {T(0), T(0), T(0), T(0)};
        if (have) {
          const int4 gij = gload_at(d.sgi, W.ent[e]);
          typej = gij.z >> 8;
          fwd = HALF || gij.w > g;
          bs = gij.x + k;
          bo = gij.y + k;
          nJl = gij.z & 255;
          rb = own_rec<T>(d, bs);
          ib = gload_at(d.sai, bs);
        }
        // separation = (owner centre - neighbour centre - image) + owner offset - neighbour offset
        const T bx = T(xcc - (double)rb.x), by = T(ycc - (double)rb.y), bz = T(zcc - (double)rb.z);
        const T swt = T(sw);
        if (have) {
          T c12[OWN_REG], c6[OWN_REG];
#pragma unroll
          for (int u = 0; u < OWN_REG; ++u) {
            c12[u] = T(0);
            c6[u] = T(0);
            if (u < nI) own_lj<T>(d, gload(d.sai + as0 + u).x, ib.x, c12[u], c6[u]);
          }
#pragma unroll
          for (int u = 0; u < OWN_REG; ++u) {
            if (u >= nI) break;
            const OwnRec<T> ra = own_rec<T>(d, as0 + u);
            const int2 ia = gload(d.sai + as0 + u);
            bool masked = false;
            if (ia.y & ib.y & 1) masked = is_excluded(gmem(d.ex_off), gmem(d.ex_idx), ia0 + u, bo);
            if (d.any_heavy && !masked && !(gmem(d.mass)[bo] < d.massinfty || gmem(d.mass)[ia0 + u] < d.massinfty)) {
              // force.f:486-490: two atoms at or above massinfty are skipped only where the
              // list owner's atom has an exclusion partner in the other group (the owner
              // is the lower group: this one when fwd, the neighbour's otherwise)
              if (fwd)
                masked = (ia.y & 1) && has_partner_in(gmem(d.ex_off), gmem(d.ex_idx), ia0 + u, bo - k, bo - k + nJl - 1);
              else
                masked = (ib.y & 1) && has_partner_in(gmem(d.ex_off), gmem(d.ex_idx), bo, ia0, ia0 + nI - 1);
            }
            const T xc = bx + ra.x, yc = by + ra.y, zc = bz + ra.z;
            const T rsq = xc * xc + yc * yc + zc * zc;
            const T rinv = own_rsqrt<T>(rsq);
            const T rsqi = rinv * rinv;
            const T r6 = rsqi * rsqi * rsqi;
            const T e6 = c6[u] * r6, e12 = c12[u] * r6 * r6;
            T conf = e12 - e6;
            T qforce = (T(12) * e12 - T(6) * e6) * rsqi;
            T cou = T(0);
            const T qq = ra.q * rb.q;
            if (!d.lskip_ewald) {
              const T rsp = rsq * rinv;
              const T alphar = T(d.alpha) * rsp;
              const T qt = own_rcp<T>(T(1) + T(AS_QP) * alphar);
              const T expcst = own_exp<T>(-(alphar * alphar));
              const T erfcst =
                  ((((T(AS_A5) * qt + T(AS_A4)) * qt + T(AS_A3)) * qt + T(AS_A2)) * qt + T(AS_A1)) * qt * expcst;
              cou = qq * erfcst * rinv;
              qforce += qq * (erfcst + T(d.twrtpi) * alphar * expcst) * (rsqi * rinv);
            } else if (!d.ewald) {
              cou = qq * rinv;
              qforce += cou * rsqi;
            }
            if (masked) {
              conf = T(0);
              cou = T(0);
              qforce = T(0);
            }
            const T qf = qforce * swt;
            const double fx = (double)(qf * xc), fy = (double)(qf * yc), fz = (double)(qf * zc);
            fr[u][0] += fx;
            fr[u][1] += fy;
            fr[u][2] += fz;
            fsx += fx;
            fsy += fy;
            fsz += fz;
            if (HALF) {
              fbx -= fx;
              fby -= fy;
              fbz -= fz;
            }
            vc += (double)conf;
            vq += (double)cou;
          }
        }
        // energies: each pair once (the j > i direction), scaled by the switch
        if (fwd) {
          ec[typej - 1] += sw * vc;
          eq[typej - 1] += sw * vq;
        }
        // S'(r) V of switched group pairs (force.f:756-817), V the group pair's
        // unswitched energy summed over the neighbour group's lanes
        double x3 = 0.0, y3 = 0.0, z3 = 0.0;
        if (__ballot(dsw != 0.0)) {
          const double v = vc + vq;
          const int first = lane - k;
          const int njl = have ? W.nj[e] : 0;
          double cmt = 0.0;
          for (int u = 0; u < d.max_group; ++u) {
            const double o = __shfl(v, (first + u) & 63, 64);
            if (u < njl) cmt += o;
          }
          x3 = -(dsw * cmt * xcc);
          y3 = -(dsw * cmt * ycc);
          z3 = -(dsw * cmt * zcc);
          if (have && k == 0) {
            x3x += x3;
            x3y += y3;
            x3z += z3;
          }
          if (HALF && have && dsw != 0.0) {
            // the neighbour group's share, -x3 by its atoms' mass fractions
            const double mb = gmem(d.gmass)[bo];
            fbx -= mb * x3;
            fby -= mb * y3;
            fbz -= mb * z3;
          }
        }
        if (HALF && have) {
          const int slot = W.pb[e];
          if (slot >= 0) {
            gmem(d.jpx)[slot + k] = fbx;
            gmem(d.jpy)[slot + k] = fby;
            gmem(d.jpz)[slot + k] = fbz;
          }
        }
        // molecular virial: the image-shift terms of pairs across the cell boundary
        if (fwd) {
          const double sxm = W.sx[e], sym = W.sy[e], szm = W.sz[e];
          if (sxm != 0.0 || sym != 0.0 || szm != 0.0) {
            if (k == 0) {
              fsx += x3;
              fsy += y3;
              fsz += z3;
            }
            st[0] -= fsx * sxm; st[1] -= fsx * sym; st[2] -= fsx * szm;
            st[3] -= fsy * sxm; st[4] -= fsy * sym; st[5] -= fsy * szm;
            st[6] -= fsz * sxm; st[7] -= fsz * sym; st[8] -= fsz * szm;
          }
        }
        e0 = e1;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  // ---- owner-atom forces: lane sums in lane order, then the S'V terms by mass fraction
  if (t < d.ngrp) {
    const double X = wave_sum(x3x), Y = wave_sum(x3y), Z = wave_sum(x3z);
    double fxa = 0.0, fya = 0.0, fza = 0.0;
#pragma unroll
    for (int u = 0; u < OWN_REG; ++u) {
      if (u < nI) {
        const double sx = wave_sum(fr[u][0]), sy = wave_sum(fr[u][1]), sz = wave_sum(fr[u][2]);
        if (lane == u) {
          fxa = sx;
          fya = sy;
          fza = sz;
        }
      }
    }
    if (lane < nI) {
      const int a = ia0 + lane;
      const double ma = gmem(d.gmass)[a];
      fxa += ma * X;
      fya += ma * Y;
      fza += ma * Z;
      gmem(d.fx)[a] = fxa;
      gmem(d.fy)[a] = fya;
      gmem(d.fz)[a] = fza;
      if (!HALF) {  // (the half form adds the molecular virial once the neighbour-side forces are in)
        const int pa = gmem(d.atomp)[a];
        const double cx = gmem(d.xpcm)[pa], cy = gmem(d.ypcm)[pa], cz = gmem(d.zpcm)[pa];
        st[0] += fxa * cx; st[1] += fxa * cy; st[2] += fxa * cz;
        st[3] += fya * cx; st[4] += fya * cy; st[5] += fya * cz;
        st[6] += fza * cx; st[7] += fza * cy; st[8] += fza * cz;
      }
    }
  }
  // ---- energies and virial: wave, then block, fixed order
  double vals[PAIR_NACC];
#pragma unroll
  for (int k = 0; k < PAIR_NACC; ++k) vals[k] = 0.0;
  // uconf / ucoul by typeij = typei + typej - 2 (0 slt-slt, 1 slt-slv, 2 slv-slv)
  vals[typei - 1] = ec[0];
  vals[typei] = ec[1];
  vals[3 + typei - 1] = eq[0];
  vals[3 + typei] = eq[1];
#pragma unroll
  for (int k = 0; k < 9; ++k) vals[6 + k] = st[k];
#pragma unroll
  for (int k = 0; k < PAIR_NACC; ++k) {
    const double v = wave_sum(vals[k]);
    if (lane == 0) sh.red[wid][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < PAIR_NACC) {
    double s = 0.0;
    for (int w = 0; w < PAIR_WAVES; ++w) s += sh.red[w][threadIdx.x];
    gmem(d.part)[(size_t)blk * 16 + threadIdx.x] = s;
  }
}

template <typename T, int W, bool HALF>
__global__ void __launch_bounds__(256, W) k_pairs_own(const PairDesc *__restrict__ D, int remap) {
  __shared__ OwnShared sh;
  int sys, blk;
  pairs_item_index(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y, gridDim.x, remap, sys, blk);
  own_item<T, HALF>(D[sys], blk, sh);
}

template <typename T, int W, bool HALF>
__global__ void __launch_bounds__(256, W) k_pairs_own_persistent(const PairDesc *__restrict__ D, int remap,
                                                                 int nitems, int nblkmax) {
  __shared__ OwnShared sh;
  for (int h = blockIdx.x; h < nitems; h += gridDim.x) {
    int sys, blk;
    pairs_item_index(h, nitems, nblkmax, remap, sys, blk);
    own_item<T, HALF>(D[sys], blk, sh);
    __syncthreads();
  }
}

template <typename T, int W, bool HALF>
static void launch_own_kernel(bool persistent, dim3 grd, int g, hipStream_t st, const PairDesc *d_desc, int remap,
                              int nitems, int max_blk) {
  const dim3 blk(64 * PAIR_WAVES);
  if (persistent)
    hipLaunchKernelGGL((k_pairs_own_persistent<T, W, HALF>), dim3(g), blk, 0, st, d_desc, remap, nitems, max_blk);
  else
    hipLaunchKernelGGL((k_pairs_own<T, W, HALF>), grd, blk, 0, st, d_desc, remap);
}

// Half form, after the pair kernel: each group's atoms add the neighbour-side
// forces of its reverse entries (one wavefront per group, lanes over
// (reverse entry, atom) in memory order, lane sums combined in lane order),
// then the molecular virial sum F_a (x) C_a of the total forces.
__global__ void __launch_bounds__(256) k_pairs_jsum(const PairDesc *__restrict__ D) {
  __shared__ double red[PAIR_WAVES][64][3];
  __shared__ double vir[PAIR_WAVES][9];
  const PairDesc &d = D[blockIdx.y];
  const int blk = blockIdx.x;
  if (blk >= d.nblk || !d.half) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int t = __builtin_amdgcn_readfirstlane(blk * PAIR_WAVES + wid);
  double st[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) st[k] = 0.0;
  if (t < d.ngrp) {
    const int4 gi = gload(d.sgi + t);
    const int ia0 = gi.y, nI = gi.z & 255;
    int nr = min(gmem(d.nb_rcount)[t], d.nb_slab);
    const int base = gmem(d.nb_rbase)[t];
    if (base < 0 || (long long)base + (long long)nr * nI > d.pcap) nr = 0;  // an overflowed build (flagged)
    const int nrc = 64 / nI, rc = lane / nI, k = lane - rc * nI;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (rc < nrc)
      for (int r = rc; r < nr; r += nrc) {
        const int idx = base + r * nI + k;
        sx += gmem(d.jpx)[idx];
        sy += gmem(d.jpy)[idx];
        sz += gmem(d.jpz)[idx];
      }
    red[wid][lane][0] = sx;
    red[wid][lane][1] = sy;
    red[wid][lane][2] = sz;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < nI) {
      const int a = ia0 + lane;
      double fxa = gmem(d.fx)[a], fya = gmem(d.fy)[a], fza = gmem(d.fz)[a];
      for (int q = 0; q < nrc; ++q) {
        fxa += red[wid][q * nI + lane][0];
        fya += red[wid][q * nI + lane][1];
        fza += red[wid][q * nI + lane][2];
      }
      gmem(d.fx)[a] = fxa;
      gmem(d.fy)[a] = fya;
      gmem(d.fz)[a] = fza;
      const int pa = gmem(d.atomp)[a];
      const double cx = gmem(d.xpcm)[pa], cy = gmem(d.ypcm)[pa], cz = gmem(d.zpcm)[pa];
      st[0] = fxa * cx; st[1] = fxa * cy; st[2] = fxa * cz;
      st[3] = fya * cx; st[4] = fya * cy; st[5] = fya * cz;
      st[6] = fza * cx; st[7] = fza * cy; st[8] = fza * cz;
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double v = wave_sum(st[k]);
    if (lane == 0) vir[wid][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 9) {
    double s = 0.0;
    for (int w = 0; w < PAIR_WAVES; ++w) s += vir[w][threadIdx.x];
    gmem(d.part2)[(size_t)blk * 9 + threadIdx.x] = s;
  }
}

int pair_split_for(int ngrp) {
  // ORAC_HIP_PAIR_SPLIT=1|2|4 overrides the choice for systems created after it is set
  const char *e = getenv("ORAC_HIP_PAIR_SPLIT");
  const int v = e ? atoi(e) : 0;
  if (v == 1 || v == 2 || v == 4) return v;
  (void)ngrp;
  return 1;  // measured: extra waves per owner cost more than the tail they trim
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

// The owner-atom loop form (and with it the half-list layout of resident
// systems) is selected by ORAC_HIP_PAIR_KERNEL=own: read when a resident
// system is created (its list layout) and at each launch of a system on the
// full list; by default every system takes k_pairs on the full list.
// Measured on MI355X on the 24,460-atom system (one evaluation with a list
// rebuild): the half-list form's per-pair slots and its list build (split,
// reverse ranking, mirror search) cost more than the halved arithmetic saves
// (0.454 vs 0.346 ms in FP64).
bool use_owner_kernel() {
  const char *e = getenv("ORAC_HIP_PAIR_KERNEL");
  return e && std::string(e) == "own";
}

template <typename T>
void launch_pairs(const PairDesc *d_desc, const PairDesc *h, int nsys, hipStream_t st, const Tuning &tu,
                  hipEvent_t kbeg, hipEvent_t kend) {
  int max_grp = 1, max_blk = 1;
  // half-list systems take the half form; the others the full owner-atom form
  // when every one of them qualifies, else k_pairs
  bool any_alch = false, own = use_owner_kernel(), any_half = false, any_full = false;
  for (int s = 0; s < nsys; ++s) {
    // every system's groups are packed and its blocks launched, whichever form serves it
    max_grp = std::max(max_grp, h[s].ngrp);
    max_blk = std::max(max_blk, h[s].nblk);
    const bool fits = h[s].alch_pair == nullptr && h[s].rematom == nullptr && h[s].split == 1 &&
                      h[s].max_group <= OWN_REG &&
                      (sizeof(T) == 4 ? h[s].sapf != nullptr : h[s].sapc != nullptr);
    if (h[s].half) {
      if (!fits) throw Error(E_STATE, "half-list pair form on a system it does not serve");
      any_half = true;
      continue;
    }
    any_full = true;
    any_alch = any_alch || h[s].alch_pair != nullptr;
    own = own && fits;
    if (h[s].max_group > 64)
      throw Error(E_UNSUPPORTED, "groups larger than 64 atoms are not supported by the pair kernel");
  }
  hipLaunchKernelGGL(k_pack_sorted, dim3((max_grp + 255) / 256, nsys), dim3(256), 0, st, d_desc);
  dim3 grd(max_blk, nsys);
  const int remap = tu.xcd_remap ? 1 : 0;
  const int waves = pair_occupancy<T>(tu.pair_waves);
  const int nitems = max_blk * nsys;
  if (kbeg) ORAC_HIP(hipEventRecord(kbeg, st));
  int g = 0;
  if (tu.persistent) {
    // CUs x workgroups per CU at the occupancy target (4 waves per workgroup)
    static int ncu = [] {
      int dev = 0, n = 0;
      if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                   hipSuccess)
        n = 256;
      return n > 0 ? n : 256;
    }();
    g = std::max(8, std::min(nitems, ncu * waves) / 8 * 8);
  }
  const bool pers = tu.persistent != 0;
  if (any_half) {
    switch (waves) {
      case 3: launch_own_kernel<T, 3, true>(pers, grd, g, st, d_desc, remap, nitems, max_blk); break;
      case 4: launch_own_kernel<T, 4, true>(pers, grd, g, st, d_desc, remap, nitems, max_blk); break;
      default: launch_own_kernel<T, 2, true>(pers, grd, g, st, d_desc, remap, nitems, max_blk); break;
    }
    hipLaunchKernelGGL(k_pairs_jsum, grd, dim3(64 * PAIR_WAVES), 0, st, d_desc);
  }
  if (!any_full) {
    // (every system in the half-list form)
  } else if (own) {
    switch (waves) {
      case 3: launch_own_kernel<T, 3, false>(pers, grd, g, st, d_desc, remap, nitems, max_blk); break;
      case 4: launch_own_kernel<T, 4, false>(pers, grd, g, st, d_desc, remap, nitems, max_blk); break;
      default: launch_own_kernel<T, 2, false>(pers, grd, g, st, d_desc, remap, nitems, max_blk); break;
    }
  } else if (any_alch) {
    // the soft-core form at the default occupancy target only
    launch_pair_kernel<T, sizeof(T) == 8 ? 3 : 4, true>(pers, grd, g, st, d_desc, remap, nitems, max_blk);
  } else {
    switch (waves) {
      case 3: launch_pair_kernel<T, 3, false>(pers, grd, g, st, d_desc, remap, nitems, max_blk); break;
      case 4: launch_pair_kernel<T, 4, false>(pers, grd, g, st, d_desc, remap, nitems, max_blk); break;
      default: launch_pair_kernel<T, 2, false>(pers, grd, g, st, d_desc, remap, nitems, max_blk); break;
    }
  }
  if (kend) ORAC_HIP(hipEventRecord(kend, st));
  hipLaunchKernelGGL(k_pairs_finalize, dim3(PAIR_NACC, nsys), dim3(256), 0, st, d_desc);
}

template void launch_pairs<double>(const PairDesc *, const PairDesc *, int, hipStream_t, const Tuning &, hipEvent_t,
                                   hipEvent_t);
template void launch_pairs<float>(const PairDesc *, const PairDesc *, int, hipStream_t, const Tuning &, hipEvent_t,
                                  hipEvent_t);

}  // namespace orac
