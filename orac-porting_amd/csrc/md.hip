// md.hip -- on-device r-RESPA integration of resident systems on gfx950.
//
// Two time levels, as ORAC's multiple-time-step integrator nests them
// (mtsmd.f): the fast level holds the solute's intramolecular harmonic
// distance terms and runs ninner substeps of dt/ninner; the slow level is the
// full nonbonded force (direct space, PME, corrections) once per outer step
// of dt.  Rigid units (3-site waters) keep their bond lengths by SHAKE after
// every drift and RATTLE after every kick.  Velocity Verlet at both levels:
//
//   v += dt/2 F_slow/m
//   ninner times:  v += h/2 F_fast/m;  x += h v (SHAKE);  F_fast(x);  v += h/2 F_fast/m (RATTLE)
//   F_slow(x)      (the force evaluation of the engine)
//   v += dt/2 F_slow/m (RATTLE)
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

// SHAKE on one rigid triangle: q[k] are the drifted positions, p[k] the
// positions before the drift (which satisfy the constraints).  Returns the
// positions moved onto the constraint surface; dv[k] += the velocity change
// (displacement / h).  Constraints (0,1), (0,2), (1,2) with squared lengths d2.
__device__ inline bool shake3(double3 *q, const double3 *p, const double *im, const double *d2, double h,
                              double3 *dv, double tol, int maxit) {
  const int ci[3] = {0, 0, 1}, cj[3] = {1, 2, 2};
  for (int it = 0; it < maxit; ++it) {
    bool done = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int i = ci[c], j = cj[c];
      const double sx = q[i].x - q[j].x, sy = q[i].y - q[j].y, sz = q[i].z - q[j].z;
      const double diff = d2[c] - (sx * sx + sy * sy + sz * sz);
      if (fabs(diff) > tol * d2[c]) {
        done = false;
        const double rx = p[i].x - p[j].x, ry = p[i].y - p[j].y, rz = p[i].z - p[j].z;
        const double g = diff / (2.0 * (sx * rx + sy * ry + sz * rz) * (im[i] + im[j]));
        q[i].x += g * im[i] * rx; q[i].y += g * im[i] * ry; q[i].z += g * im[i] * rz;
        q[j].x -= g * im[j] * rx; q[j].y -= g * im[j] * ry; q[j].z -= g * im[j] * rz;
        dv[i].x += g * im[i] * rx / h; dv[i].y += g * im[i] * ry / h; dv[i].z += g * im[i] * rz / h;
        dv[j].x -= g * im[j] * rx / h; dv[j].y -= g * im[j] * ry / h; dv[j].z -= g * im[j] * rz / h;
      }
    }
    if (done) return true;
  }
  return false;
}

// RATTLE velocity stage on one rigid triangle: removes the velocity
// components along the bonds, r_ij . v_ij = 0.
__device__ inline bool rattle3(const double3 *r, double3 *v, const double *im, double tol, int maxit) {
  const int ci[3] = {0, 0, 1}, cj[3] = {1, 2, 2};
  for (int it = 0; it < maxit; ++it) {
    bool done = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int i = ci[c], j = cj[c];
      const double rx = r[i].x - r[j].x, ry = r[i].y - r[j].y, rz = r[i].z - r[j].z;
      const double vx = v[i].x - v[j].x, vy = v[i].y - v[j].y, vz = v[i].z - v[j].z;
      const double rr = rx * rx + ry * ry + rz * rz;
      const double rv = rx * vx + ry * vy + rz * vz;
      if (fabs(rv) > tol * rr) {
        done = false;
        const double k = rv / (rr * (im[i] + im[j]));
        v[i].x -= k * im[i] * rx; v[i].y -= k * im[i] * ry; v[i].z -= k * im[i] * rz;
        v[j].x += k * im[j] * rx; v[j].y += k * im[j] * ry; v[j].z += k * im[j] * rz;
      }
    }
    if (done) return true;
  }
  return false;
}

constexpr double SHAKE_TOL = 1.0e-12;
constexpr double RATTLE_TOL = 1.0e-12;
constexpr int CONSTRAINT_MAXIT = 500;

// Kick (v += c F/m, F = ff (fast) or the slow forces), optionally a drift by h
// with SHAKE, optionally RATTLE.  flags: 1 kick slow, 2 kick fast, 4 drift,
// 8 rattle.  One thread per unit; NA atoms per unit as a template parameter
// so the unit's arrays stay in registers.
template <int NA>
__device__ __forceinline__ void md_unit(const MdDesc &d, int u, int a0, double cslow, double cfast, double h,
                                        int flags) {
  double3 x[NA], v[NA], p[NA], dv[NA];
  double im[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const int a = a0 + k;
    im[k] = gmem(d.invm)[a];
    x[k] = make_double3(gmem(d.x)[a], gmem(d.y)[a], gmem(d.z)[a]);
    v[k] = make_double3(gmem(d.vx)[a], gmem(d.vy)[a], gmem(d.vz)[a]);
    if (flags & 1) {
      v[k].x += cslow * im[k] * gmem(d.fx)[a];
      v[k].y += cslow * im[k] * gmem(d.fy)[a];
      v[k].z += cslow * im[k] * gmem(d.fz)[a];
    }
    if (flags & 2) {
      v[k].x += cfast * im[k] * gmem(d.bfx)[a];
      v[k].y += cfast * im[k] * gmem(d.bfy)[a];
      v[k].z += cfast * im[k] * gmem(d.bfz)[a];
    }
  }
  if (flags & 4) {
#pragma unroll
    for (int k = 0; k < NA; ++k) {
      p[k] = x[k];
      x[k].x += h * v[k].x;
      x[k].y += h * v[k].y;
      x[k].z += h * v[k].z;
      dv[k] = make_double3(0.0, 0.0, 0.0);
    }
    if constexpr (NA == 3) {
      if (!shake3(x, p, im, d.ud2 + 3 * (size_t)u, h, dv, SHAKE_TOL, CONSTRAINT_MAXIT)) atomicAdd(d.fail, 1);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        v[k].x += dv[k].x;
        v[k].y += dv[k].y;
        v[k].z += dv[k].z;
      }
    }
  }
  if constexpr (NA == 3)
    if (flags & 8)
      if (!rattle3(x, v, im, RATTLE_TOL, CONSTRAINT_MAXIT)) atomicAdd(d.fail, 1);
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const int a = a0 + k;
    gmem(d.vx)[a] = v[k].x;
    gmem(d.vy)[a] = v[k].y;
    gmem(d.vz)[a] = v[k].z;
    if (flags & 4) {
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

__global__ void __launch_bounds__(256) k_md_units(const MdDesc *__restrict__ D, double cslow, double cfast,
                                                  double h, int flags) {
  const MdDesc &d = D[blockIdx.y];
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= d.nunits) return;
  const int a0 = gmem(d.ufirst)[u];
  if (gmem(d.urigid)[u])
    md_unit<3>(d, u, a0, cslow, cfast, h, flags);
  else
    md_unit<1>(d, u, a0, cslow, cfast, h, flags);
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
      const double m = 1.0 / d.invm[a];
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
    im[k] = gmem(d.invm)[a];
    p[k] = make_double3(gmem(d.x)[a], gmem(d.y)[a], gmem(d.z)[a]);
    double fx = gmem(d.fx)[a] + gmem(d.bfx)[a], fy = gmem(d.fy)[a] + gmem(d.bfy)[a],
           fz = gmem(d.fz)[a] + gmem(d.bfz)[a];
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
    if (!shake3(x, p, im, d.ud2 + 3 * (size_t)u, 1.0, dv, SHAKE_TOL, CONSTRAINT_MAXIT)) atomicAdd(d.fail, 1);
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

void launch_md_units(const MdDesc *d, const MdDesc *h, int nsys, double cslow, double cfast, double hdt, int flags,
                     hipStream_t st) {
  int mu, ma, mb;
  md_dims(h, nsys, mu, ma, mb);
  if (mu > 0) hipLaunchKernelGGL(k_md_units, dim3((mu + 255) / 256, nsys), dim3(256), 0, st, d, cslow, cfast, hdt, flags);
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
