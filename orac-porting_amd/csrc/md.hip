// md.hip -- on-device r-RESPA integration of resident systems on gfx950.
//
// Two time levels of ORAC's multiple-time-step integrator (mtsmd.f), in its
// order of operations: the fast level (ORAC's n0 level: the solute's
// intramolecular harmonic distance terms) runs ninner substeps of
// h = dt/ninner inside one step dt of the slow level (ORAC's m level: the full
// nonbonded force -- direct space, PME, corrections -- with a single shell).
// Rigid units (3-site waters) are constrained by ORAC's matrix-inversion
// SHAKE after every drift and its RATTLE after every kick:
//
//   v += Fs*0.5*dt/m; RATTLE                                 (:2598-2604)
//   ninner times:  x += h*v + Ff*0.5*h^2/m; v += Ff*0.5*h/m;
//                  SHAKE against the positions before        (:2663-2683)
//                  Ff(x); v += Ff*0.5*h/m; RATTLE            (:2843-2884)
//   Fs(x)          (the force evaluation of the engine)
//   v += Fs*0.5*dt/m; RATTLE                                 (:3121-3146)
//
// Positions are Cartesian FP64 with the reduced copy xp = oc x refreshed after
// every drift; the force path reads the reduced copy.  A unit is one atom or
// one rigid 3-atom molecule; one thread advances a unit, so constraint
// iterations stay in registers.  Every reduction (kinetic and bond energy) is
// per system and in fixed order.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace orac {

__device__ inline void to_reduced(const double *oc, double x, double y, double z, double &xp, double &yp, double &zp) {
  xp = oc[0] * x + oc[3] * y + oc[6] * z;
  yp = oc[1] * x + oc[4] * y + oc[7] * z;
  zp = oc[2] * x + oc[5] * y + oc[8] * z;
}

// Constraints of one rigid triangle by ORAC's matrix-inversion method (MIM),
// the branch ORAC takes for molecules of at most mim_lim constraints
// (default 20, blkdta.f90:361): SHAKE after a drift as rattle_verlet.f:228-560
// does it, RATTLE after a kick as rattle_correc.f:176-385.  Constraint order
// (0,1), (0,2), (1,2); im = 1/m; dss the squared lengths.  mim_dir gives the
// direction in which atom `at` moves for constraint (ca, cb) with bond vector
// d (rattle_verlet.f:286-303): -d for its first atom, +d for its second.
__device__ inline double3 mim_dir(int at, int ca, int cb, double3 d) {
  if (at == ca) return make_double3(-d.x, -d.y, -d.z);
  if (at == cb) return d;
  return make_double3(0.0, 0.0, 0.0);
}

// 3x3 solve by cofactors (rattle_verlet.f:381-397); later solves reuse them
struct MimInv {
  double b11, b12, b13, b21, b22, b23, b31, b32, b33, det;
};

__device__ inline MimInv mim_inverse(const double (&m)[3][3]) {
  MimInv r;
  const double a11 = m[0][0], a12 = m[0][1], a13 = m[0][2], a21 = m[1][0], a22 = m[1][1], a23 = m[1][2],
               a31 = m[2][0], a32 = m[2][1], a33 = m[2][2];
  r.b11 = a22 * a33 - a32 * a23;
  r.b12 = -a21 * a33 + a31 * a23;
  r.b13 = a21 * a32 - a22 * a31;
  r.b21 = -a12 * a33 + a32 * a13;
  r.b22 = a11 * a33 - a31 * a13;
  r.b23 = -a11 * a32 + a31 * a12;
  r.b31 = a12 * a23 - a22 * a13;
  r.b32 = -a11 * a23 + a21 * a13;
  r.b33 = a11 * a22 - a21 * a12;
  r.det = 1.0 / (a11 * r.b11 + a21 * r.b21 + a31 * r.b31);
  return r;
}

__device__ inline void mim_solve(const MimInv &r, double (&g)[3]) {
  const double xd = r.b11 * g[0] + r.b21 * g[1] + r.b31 * g[2];
  const double yd = r.b12 * g[0] + r.b22 * g[1] + r.b32 * g[2];
  const double zd = r.b13 * g[0] + r.b23 * g[1] + r.b33 * g[2];
  g[0] = xd * r.det;
  g[1] = yd * r.det;
  g[2] = zd * r.det;
}

constexpr double MIM_TOL = 1.0e-7;  // relative change of the multipliers (rattle_verlet.f:77)
constexpr int MIM_MAXIT = 120;      // rattle_verlet.f:461

// SHAKE: q are the drifted positions, p the positions before the drift (whose
// bond vectors give the directions of the corrections); v += correction * dti.
// The multipliers solve dss = |q_ij + corrections|^2 by iterating on the
// quadratic term; that term weights every constraint with the masses of the
// last constraint's atoms, as rattle_verlet.f:406-407 does.  Returns false
// when MIM_MAXIT iterations do not converge (the reference then starts the
// molecule over; here the unit is counted as failed and keeps what it has).
__device__ inline bool shake_mim3(double3 (&q)[3], const double3 (&p)[3], const double (&im)[3], const double *dss,
                                  double dti, double3 (&v)[3]) {
  constexpr int ca[3] = {0, 0, 1}, cb[3] = {1, 2, 2};
  double3 x0[3], xc[3], da[3][3], db[3][3];
  double dd[3], mat[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x0[k] = make_double3(p[ca[k]].x - p[cb[k]].x, p[ca[k]].y - p[cb[k]].y, p[ca[k]].z - p[cb[k]].z);
    xc[k] = make_double3(q[ca[k]].x - q[cb[k]].x, q[ca[k]].y - q[cb[k]].y, q[ca[k]].z - q[cb[k]].z);
  }
#pragma unroll
  for (int ka = 0; ka < 3; ++ka) {
    const int la = ca[ka], lb = cb[ka];
    dd[ka] = dss[ka] - (xc[ka].x * xc[ka].x + xc[ka].y * xc[ka].y + xc[ka].z * xc[ka].z);
#pragma unroll
    for (int kb = 0; kb < 3; ++kb) {
      const double3 a = mim_dir(la, ca[kb], cb[kb], x0[kb]);
      const double3 b = mim_dir(lb, ca[kb], cb[kb], x0[kb]);
      da[ka][kb] = a;
      db[ka][kb] = b;
      double aux3 = (-a.x * xc[ka].x + -a.y * xc[ka].y + -a.z * xc[ka].z) * im[la];
      aux3 = aux3 + (b.x * xc[ka].x + b.y * xc[ka].y + b.z * xc[ka].z) * im[lb];
      mat[ka][kb] = -(aux3 * 4.0);
    }
  }
  const MimInv inv = mim_inverse(mat);
  const double gg1 = im[ca[2]] * 2.0, gg2 = im[cb[2]] * 2.0;  // the last constraint's atoms
  double gam[3] = {0.0, 0.0, 0.0};
  bool ok = true;
  for (int iter = 0;;) {
    const double gamo[3] = {gam[0], gam[1], gam[2]};
    double g[3];
    if (iter == 0) {
#pragma unroll
      for (int ka = 0; ka < 3; ++ka) g[ka] = dd[ka];
    } else {
#pragma unroll
      for (int ka = 0; ka < 3; ++ka) {
        double xd = 0.0, yd = 0.0, zd = 0.0;
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) {
          double aux = gam[kb] * gg1;
          xd = xd + da[ka][kb].x * aux;
          yd = yd + da[ka][kb].y * aux;
          zd = zd + da[ka][kb].z * aux;
          aux = gam[kb] * gg2;
          xd = xd + -db[ka][kb].x * aux;
          yd = yd + -db[ka][kb].y * aux;
          zd = zd + -db[ka][kb].z * aux;
        }
        g[ka] = dd[ka] - (xd * xd + yd * yd + zd * zd);
      }
    }
    mim_solve(inv, g);
#pragma unroll
    for (int ka = 0; ka < 3; ++ka) gam[ka] = g[ka];
    ++iter;
    if (iter > MIM_MAXIT) {
      ok = false;
      break;
    }
    bool more = false;
#pragma unroll
    for (int ka = 0; ka < 3; ++ka)
      if (gam[ka] != 0.0 && fabs((gam[ka] - gamo[ka]) / gam[ka]) > MIM_TOL) more = true;
    if (!more) break;
  }
  // corrections, each atom once: by the first constraint that holds it
  // (rattle_verlet.f:514-541)
  bool mask[3] = {true, true, true};
#pragma unroll
  for (int ka = 0; ka < 3; ++ka) {
    const int la = ca[ka], lb = cb[ka];
    const double xd = im[la] * 2.0, yd = im[lb] * 2.0;
#pragma unroll
    for (int kb = 0; kb < 3; ++kb) {
      if (mask[la]) {
        double aux = gam[kb] * xd;
        q[la].x = q[la].x + da[ka][kb].x * aux;
        q[la].y = q[la].y + da[ka][kb].y * aux;
        q[la].z = q[la].z + da[ka][kb].z * aux;
        aux = aux * dti;
        v[la].x = v[la].x + da[ka][kb].x * aux;
        v[la].y = v[la].y + da[ka][kb].y * aux;
        v[la].z = v[la].z + da[ka][kb].z * aux;
      }
      if (mask[lb]) {
        double aux = gam[kb] * yd;
        q[lb].x = q[lb].x + db[ka][kb].x * aux;
        q[lb].y = q[lb].y + db[ka][kb].y * aux;
        q[lb].z = q[lb].z + db[ka][kb].z * aux;
        aux = aux * dti;
        v[lb].x = v[lb].x + db[ka][kb].x * aux;
        v[lb].y = v[lb].y + db[ka][kb].y * aux;
        v[lb].z = v[lb].z + db[ka][kb].z * aux;
      }
    }
    mask[la] = false;
    mask[lb] = false;
  }
  return ok;
}

// RATTLE: velocities made orthogonal to the bonds at positions r by one
// linear solve (rattle_correc.f:176-385).
__device__ inline void rattle_mim3(const double3 (&r)[3], double3 (&v)[3], const double (&im)[3]) {
  constexpr int ca[3] = {0, 0, 1}, cb[3] = {1, 2, 2};
  double3 bv[3], da[3][3], db[3][3];
  double mat[3][3], gam[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    bv[k] = make_double3(r[ca[k]].x - r[cb[k]].x, r[ca[k]].y - r[cb[k]].y, r[ca[k]].z - r[cb[k]].z);
#pragma unroll
  for (int ka = 0; ka < 3; ++ka) {
    const int la = ca[ka], lb = cb[ka];
    const double3 xc = make_double3(v[la].x - v[lb].x, v[la].y - v[lb].y, v[la].z - v[lb].z);
    gam[ka] = -(xc.x * bv[ka].x + xc.y * bv[ka].y + xc.z * bv[ka].z);
#pragma unroll
    for (int kb = 0; kb < 3; ++kb) {
      const double3 a = mim_dir(la, ca[kb], cb[kb], bv[kb]);
      const double3 b = mim_dir(lb, ca[kb], cb[kb], bv[kb]);
      da[ka][kb] = a;
      db[ka][kb] = b;
      double aux3 = (bv[ka].x * -a.x + bv[ka].y * -a.y + bv[ka].z * -a.z) * im[la];
      aux3 = aux3 + (bv[ka].x * b.x + bv[ka].y * b.y + bv[ka].z * b.z) * im[lb];
      mat[ka][kb] = -(aux3 * 2.0);
    }
  }
  mim_solve(mim_inverse(mat), gam);
  bool mask[3] = {true, true, true};
#pragma unroll
  for (int ka = 0; ka < 3; ++ka) {
    const int la = ca[ka], lb = cb[ka];
    const double xd = im[la] * 2.0, yd = im[lb] * 2.0;
#pragma unroll
    for (int kb = 0; kb < 3; ++kb) {
      if (mask[la]) {
        v[la].x = v[la].x + (da[ka][kb].x * gam[kb]) * xd;
        v[la].y = v[la].y + (da[ka][kb].y * gam[kb]) * xd;
        v[la].z = v[la].z + (da[ka][kb].z * gam[kb]) * xd;
      }
      if (mask[lb]) {
        v[lb].x = v[lb].x + (db[ka][kb].x * gam[kb]) * yd;
        v[lb].y = v[lb].y + (db[ka][kb].y * gam[kb]) * yd;
        v[lb].z = v[lb].z + (db[ka][kb].z * gam[kb]) * yd;
      }
    }
    mask[la] = false;
    mask[lb] = false;
  }
}

// One unit's share of an integration stage, in ORAC's order (mtsmd.f):
//   MD_KICK_FAST  v += Ff*0.5*h/m, RATTLE        correc + rattle_correc (:2876-2884)
//   MD_KICK_MID   v += Fm*0.5*hm/m, RATTLE       the inner nonbonded shell (two shells)
//   MD_KICK_SLOW  v += Fs*0.5*dt/m, RATTLE       correc + rattle_correc (:2598-2604, :3121-3146)
//   MD_VERLET     x += h*v + Ff*0.5*h^2/m, v += Ff*0.5*h/m, then SHAKE against the
//                 positions before the drift      verlet + rattle_verlet (:2663-2683)
//   MD_RATTLE     RATTLE alone
// applied in that order when several are set (a stage that ends one level's
// step and starts the next).  Ff are the fast forces, Fs the slow ones; the
// expressions keep verlet.f's and correc.f's association.  One thread per
// unit; NA atoms per unit as a template parameter so the unit stays in
// registers.
template <int NA>
__device__ __forceinline__ void md_unit(const MdDesc &d, int u, int a0, double dt, double hm, double h, int flags) {
  double3 x[NA], v[NA], ff[NA];
  double m[NA], im[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const int a = a0 + k;
    m[k] = gmem(d.mass)[a];
    im[k] = 1.0 / m[k];
    x[k] = make_double3(gmem(d.x)[a], gmem(d.y)[a], gmem(d.z)[a]);
    v[k] = make_double3(gmem(d.vx)[a], gmem(d.vy)[a], gmem(d.vz)[a]);
    ff[k] = make_double3(gmem(d.bfx)[a], gmem(d.bfy)[a], gmem(d.bfz)[a]);
  }
  if (flags & MD_KICK_FAST) {
#pragma unroll
    for (int k = 0; k < NA; ++k) {
      v[k].x = v[k].x + ff[k].x * 0.5 * h / m[k];
      v[k].y = v[k].y + ff[k].y * 0.5 * h / m[k];
      v[k].z = v[k].z + ff[k].z * 0.5 * h / m[k];
    }
    if constexpr (NA == 3) rattle_mim3(x, v, im);
  }
  if (flags & MD_KICK_MID) {
#pragma unroll
    for (int k = 0; k < NA; ++k) {
      const int a = a0 + k;
      v[k].x = v[k].x + gmem(d.fmx)[a] * 0.5 * hm / m[k];
      v[k].y = v[k].y + gmem(d.fmy)[a] * 0.5 * hm / m[k];
      v[k].z = v[k].z + gmem(d.fmz)[a] * 0.5 * hm / m[k];
    }
    if constexpr (NA == 3) rattle_mim3(x, v, im);
  }
  if (flags & MD_KICK_SLOW) {
#pragma unroll
    for (int k = 0; k < NA; ++k) {
      const int a = a0 + k;
      v[k].x = v[k].x + gmem(d.fx)[a] * 0.5 * dt / m[k];
      v[k].y = v[k].y + gmem(d.fy)[a] * 0.5 * dt / m[k];
      v[k].z = v[k].z + gmem(d.fz)[a] * 0.5 * dt / m[k];
    }
    if constexpr (NA == 3) rattle_mim3(x, v, im);
  }
  if (flags & MD_VERLET) {
    double3 p[NA];
    const double ts2 = h * h;
#pragma unroll
    for (int k = 0; k < NA; ++k) {
      p[k] = x[k];
      x[k].x = x[k].x + h * v[k].x + ff[k].x * 0.5 * ts2 / m[k];
      x[k].y = x[k].y + h * v[k].y + ff[k].y * 0.5 * ts2 / m[k];
      x[k].z = x[k].z + h * v[k].z + ff[k].z * 0.5 * ts2 / m[k];
      v[k].x = v[k].x + ff[k].x * 0.5 * h / m[k];
      v[k].y = v[k].y + ff[k].y * 0.5 * h / m[k];
      v[k].z = v[k].z + ff[k].z * 0.5 * h / m[k];
    }
    if constexpr (NA == 3)
      if (!shake_mim3(x, p, im, d.ud2 + 3 * (size_t)u, h != 0.0 ? 1.0 / h : 0.0, v)) atomicAdd(d.fail, 1);
  }
  if constexpr (NA == 3)
    if (flags & MD_RATTLE) rattle_mim3(x, v, im);
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const int a = a0 + k;
    gmem(d.vx)[a] = v[k].x;
    gmem(d.vy)[a] = v[k].y;
    gmem(d.vz)[a] = v[k].z;
    if (flags & MD_VERLET) {
      gmem(d.x)[a] = x[k].x;
      gmem(d.y)[a] = x[k].y;
      gmem(d.z)[a] = x[k].z;
      double xp, yp, zp;
      to_reduced(d.oc, x[k].x, x[k].y, x[k].z, xp, yp, zp);
      gmem(d.xp)[a] = xp;
      gmem(d.yp)[a] = yp;
      gmem(d.zp)[a] = zp;
    }
  }
}

__global__ void __launch_bounds__(256) k_md_units(const MdDesc *__restrict__ D, double dt, double hm, double h,
                                                  int flags) {
  const MdDesc &d = D[blockIdx.y];
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= d.nunits) return;
  const int a0 = gmem(d.ufirst)[u];
  if (gmem(d.urigid)[u])
    md_unit<3>(d, u, a0, dt, hm, h, flags);
  else
    md_unit<1>(d, u, a0, dt, hm, h, flags);
}

__global__ void __launch_bounds__(256) k_md_bonds(const MdDesc *__restrict__ D) {
  const MdDesc &d = D[blockIdx.y];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.nbond) return;
  const int i = d.bi[t], j = d.bj[t];
  const double dx = d.x[i] - d.x[j], dy = d.y[i] - d.y[j], dz = d.z[i] - d.z[j];
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const double dr = r - d.br0[t];
  const double f = -2.0 * d.bk[t] * dr / r;  // force on i along (x_i - x_j)
  d.bpf[3 * (size_t)t] = f * dx;
  d.bpf[3 * (size_t)t + 1] = f * dy;
  d.bpf[3 * (size_t)t + 2] = f * dz;
  d.bpe[t] = d.bk[t] * dr * dr;
}

// fast forces per atom from its terms, in fixed order
__global__ void __launch_bounds__(256) k_md_bond_gather(const MdDesc *__restrict__ D) {
  const MdDesc &d = D[blockIdx.y];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= d.natoms) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int t = d.bg_off[a]; t < d.bg_off[a + 1]; ++t) {
    const int sp = d.bg_term[t];
    const int p = (sp > 0 ? sp : -sp) - 1;
    const double s = sp > 0 ? 1.0 : -1.0;
    sx += s * d.bpf[3 * (size_t)p];
    sy += s * d.bpf[3 * (size_t)p + 1];
    sz += s * d.bpf[3 * (size_t)p + 2];
  }
  d.bfx[a] = sx;
  d.bfy[a] = sy;
  d.bfz[a] = sz;
}

// Per-system energies into d.out: [0] kinetic, [1] bond energy, [2..4]
// momentum; one block per component, fixed order over atoms / terms.
__global__ void __launch_bounds__(256) k_md_energy(const MdDesc *__restrict__ D) {
  __shared__ double red[4];
  const MdDesc &d = D[blockIdx.y];
  const int k = blockIdx.x;
  double s = 0.0;
  if (k == 1) {
    for (int t = threadIdx.x; t < d.nbond; t += blockDim.x) s += d.bpe[t];
  } else {
    for (int a = threadIdx.x; a < d.natoms; a += blockDim.x) {
      const double m = d.mass[a];
      if (k == 0) s += 0.5 * m * (d.vx[a] * d.vx[a] + d.vy[a] * d.vy[a] + d.vz[a] * d.vz[a]);
      else if (k == 2) s += m * d.vx[a];
      else if (k == 3) s += m * d.vy[a];
      else s += m * d.vz[a];
    }
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    d.out[k] = t;
  }
}

// steepest-descent relaxation step: each unit moves along its force by at
// most dmax (per atom h F/m capped at dmax), then SHAKE; velocities are zeroed
template <int NA>
__device__ __forceinline__ void md_relax_unit(const MdDesc &d, int u, int a0, double hstep, double dmax) {
  double3 x[NA], p[NA], dv[NA];
  double im[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const int a = a0 + k;
    im[k] = 1.0 / gmem(d.mass)[a];
    p[k] = make_double3(gmem(d.x)[a], gmem(d.y)[a], gmem(d.z)[a]);
    double fx = gmem(d.fx)[a] + gmem(d.bfx)[a], fy = gmem(d.fy)[a] + gmem(d.bfy)[a],
           fz = gmem(d.fz)[a] + gmem(d.bfz)[a];
    if (d.fmx) {
      fx += gmem(d.fmx)[a];
      fy += gmem(d.fmy)[a];
      fz += gmem(d.fmz)[a];
    }
    double sx = hstep * im[k] * fx, sy = hstep * im[k] * fy, sz = hstep * im[k] * fz;
    const double s = sqrt(sx * sx + sy * sy + sz * sz);
    if (!(s <= dmax)) {  // also catches a non-finite step
      const double c = s > 0.0 && s < 1e300 ? dmax / s : 0.0;
      sx *= c;
      sy *= c;
      sz *= c;
    }
    x[k] = make_double3(p[k].x + sx, p[k].y + sy, p[k].z + sz);
    dv[k] = make_double3(0.0, 0.0, 0.0);
  }
  if constexpr (NA == 3)
    if (!shake_mim3(x, p, im, d.ud2 + 3 * (size_t)u, 0.0, dv)) atomicAdd(d.fail, 1);
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const int a = a0 + k;
    gmem(d.x)[a] = x[k].x;
    gmem(d.y)[a] = x[k].y;
    gmem(d.z)[a] = x[k].z;
    gmem(d.vx)[a] = 0.0;
    gmem(d.vy)[a] = 0.0;
    gmem(d.vz)[a] = 0.0;
    double xp, yp, zp;
    to_reduced(d.oc, x[k].x, x[k].y, x[k].z, xp, yp, zp);
    gmem(d.xp)[a] = xp;
    gmem(d.yp)[a] = yp;
    gmem(d.zp)[a] = zp;
  }
}

// one steepest-descent step of every unit: displacement hstep F/m capped at
// dmax per atom, rigid units put back on their constraints, velocities zeroed
__global__ void __launch_bounds__(256) k_md_relax(const MdDesc *__restrict__ D, double hstep, double dmax) {
  const MdDesc &d = D[blockIdx.y];
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= d.nunits) return;
  const int a0 = gmem(d.ufirst)[u];
  if (gmem(d.urigid)[u])
    md_relax_unit<3>(d, u, a0, hstep, dmax);
  else
    md_relax_unit<1>(d, u, a0, hstep, dmax);
}

// v *= scale (per system), e.g. for velocity rescaling during equilibration
__global__ void __launch_bounds__(256) k_md_scale_v(const MdDesc *__restrict__ D, const double *scale) {
  const MdDesc &d = D[blockIdx.y];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= d.natoms) return;
  const double s = scale[blockIdx.y];
  d.vx[a] *= s;
  d.vy[a] *= s;
  d.vz[a] *= s;
}

// one outer step done: advance every system's step counter
__global__ void k_md_tick(const MdDesc *__restrict__ D) {
  if (threadIdx.x == 0) *D[blockIdx.y].step += 1;
}

static void md_dims(const MdDesc *h, int nsys, int &max_units, int &max_atoms, int &max_bonds) {
  max_units = max_atoms = max_bonds = 0;
  for (int s = 0; s < nsys; ++s) {
    max_units = std::max(max_units, h[s].nunits);
    max_atoms = std::max(max_atoms, h[s].natoms);
    max_bonds = std::max(max_bonds, h[s].nbond);
  }
}

void launch_md_units(const MdDesc *d, const MdDesc *h, int nsys, double dt, double hm, double hin, int flags,
                     hipStream_t st) {
  int mu, ma, mb;
  md_dims(h, nsys, mu, ma, mb);
  if (mu > 0) hipLaunchKernelGGL(k_md_units, dim3((mu + 255) / 256, nsys), dim3(256), 0, st, d, dt, hm, hin, flags);
}

void launch_md_fast_forces(const MdDesc *d, const MdDesc *h, int nsys, hipStream_t st) {
  int mu, ma, mb;
  md_dims(h, nsys, mu, ma, mb);
  if (mb > 0) hipLaunchKernelGGL(k_md_bonds, dim3((mb + 255) / 256, nsys), dim3(256), 0, st, d);
  if (ma > 0) hipLaunchKernelGGL(k_md_bond_gather, dim3((ma + 255) / 256, nsys), dim3(256), 0, st, d);
}

void launch_md_tick(const MdDesc *d, int nsys, hipStream_t st) {
  hipLaunchKernelGGL(k_md_tick, dim3(1, nsys), dim3(64), 0, st, d);
}

void launch_md_energy(const MdDesc *d, int nsys, hipStream_t st) {
  hipLaunchKernelGGL(k_md_energy, dim3(5, nsys), dim3(256), 0, st, d);
}

void launch_md_relax(const MdDesc *d, const MdDesc *h, int nsys, double hstep, double dmax, hipStream_t st) {
  int mu, ma, mb;
  md_dims(h, nsys, mu, ma, mb);
  if (mu > 0) hipLaunchKernelGGL(k_md_relax, dim3((mu + 255) / 256, nsys), dim3(256), 0, st, d, hstep, dmax);
}

void launch_md_scale_v(const MdDesc *d, const MdDesc *h, int nsys, const double *scale, hipStream_t st) {
  int mu, ma, mb;
  md_dims(h, nsys, mu, ma, mb);
  if (ma > 0) hipLaunchKernelGGL(k_md_scale_v, dim3((ma + 255) / 256, nsys), dim3(256), 0, st, d, scale);
}

}  // namespace orac
