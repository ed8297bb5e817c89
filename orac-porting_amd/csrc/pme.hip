// pme.hip -- smooth PME reciprocal sum on gfx950.
//
// Stages (each one kernel, all systems of a batch in one launch via blockIdx.y):
//   k_spread      fractionals + order-p B-splines + charge spreading with
//                 64-bit fixed-point atomics (pme_fftw3.f90:132-163,234-298,487-576)
//   k_fft_x_r2c   real rows along K1 -> half spectrum (FFTW r2c, pme_fftw3.f90:114)
//   k_fft_pencil  complex pencils along K2 (forward / inverse)
//   k_fft_z_conv  pencils along K3: forward FFT, influence function with
//                 energy and virial (scalar_sum.f:1-135), inverse FFT, in LDS
//   k_fft_x_c2r   half spectrum -> real rows (FFTW c2r, pme_fftw3.f90:116)
//   k_gather      forces from the convolved grid (pme_fftw3.f90:311-372) and
//                 the molecular-virial correction terms (fft_pme.f:65-106)
//   k_pme_finalize fixed-order reduction of the per-block partials
// FFTs are LDS-staged mixed radix-2/3/4/5 Stockham passes over tiles of
// pencils; loads and stores walk the contiguous K1 direction across threads.
#include "common.hpp"
#include "kernels.hpp"

namespace orac {

template <typename T>
struct cx {
  T x, y;
};
template <typename T>
__device__ inline cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ inline cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ inline cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ inline cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
// i*s*a for s = +-1
template <typename T, int S>
__device__ inline cx<T> cmuli(cx<T> a) { return S > 0 ? cx<T>{-a.y, a.x} : cx<T>{a.y, -a.x}; }

template <typename T>
__device__ inline cx<T> twiddle(const FftAxis &ax, int m);
template <>
__device__ inline cx<double> twiddle<double>(const FftAxis &ax, int m) {
  double2 w = ax.tw64[m];
  return {w.x, w.y};
}
template <>
__device__ inline cx<float> twiddle<float>(const FftAxis &ax, int m) {
  float2 w = ax.tw32[m];
  return {w.x, w.y};
}

// R-point DFT in registers, S = -1 forward (e^{-i}), +1 backward.
template <typename T, int R, int S>
__device__ inline void dft(cx<T> *v) {
  if constexpr (R == 2) {
    cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const T s3 = T(0.8660254037844386467637231707529362);
    cx<T> t = cadd(v[1], v[2]), u = csub(v[1], v[2]);
    cx<T> a = csub(v[0], cscale(t, T(0.5)));
    cx<T> b = cmuli<T, S>(cscale(u, s3));
    v[0] = cadd(v[0], t);
    v[1] = cadd(a, b);
    v[2] = csub(a, b);
  } else if constexpr (R == 4) {
    cx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cx<T> t2 = cadd(v[1], v[3]), t3 = cmuli<T, S>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else {  // R == 5
    const T c1 = T(0.3090169943749474241022934171828191), c2 = T(-0.8090169943749474241022934171828191);
    const T s1 = T(0.9510565162951535721164393333793821), s2 = T(0.5877852522924731291687059546390728);
    cx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    cx<T> a1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
    cx<T> a2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
    cx<T> b1 = cmuli<T, S>(cadd(cscale(t3, s1), cscale(t4, s2)));
    cx<T> b2 = cmuli<T, S>(csub(cscale(t3, s2), cscale(t4, s1)));
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  }
}

template <typename T, int R, int S>
__device__ inline void stockham_pass(const cx<T> *a, cx<T> *b, int N, int TP, int Ns, const FftAxis &ax) {
  const int NR = N / R;
  const int twstep = N / (Ns * R);
  for (int bi = threadIdx.x; bi < TP * NR; bi += blockDim.x) {
    int t = bi / NR, j = bi - t * NR;
    const cx<T> *src = a + t * N;
    cx<T> *dst = b + t * N;
    int k = j % Ns;
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[j + r * NR];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx<T> w = twiddle<T>(ax, r * k * twstep);
        if (S > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    dft<T, R, S>(v);
    int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[idxD + r * Ns] = v[r];
  }
}

// TP pencils of length N held in LDS buffer a (b is scratch); returns the
// buffer that holds the result.
template <typename T, int S>
__device__ cx<T> *lds_fft(cx<T> *a, cx<T> *b, int N, int TP, const FftAxis &ax) {
  int Ns = 1;
  for (int p = 0; p < ax.nr; ++p) {
    int R = ax.radix[p];
    switch (R) {
      case 2: stockham_pass<T, 2, S>(a, b, N, TP, Ns, ax); break;
      case 3: stockham_pass<T, 3, S>(a, b, N, TP, Ns, ax); break;
      case 4: stockham_pass<T, 4, S>(a, b, N, TP, Ns, ax); break;
      default: stockham_pass<T, 5, S>(a, b, N, TP, Ns, ax); break;
    }
    __syncthreads();
    cx<T> *t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// ---------------------------------------------------------------- spread

// Fractional grid coordinate along one reciprocal axis (pme_fftw3.f90:151-160).
__device__ inline double frac_coord(double x, double y, double z, const double *recip, int col, int K) {
  double w = x * recip[3 * col] + y * recip[3 * col + 1] + z * recip[3 * col + 2];
  return K * (w - fnint(w) + 0.5);
}

__device__ inline int wrapK(int k0, int K) { return k0 < 0 ? k0 + K : (k0 >= K ? k0 - K : k0); }

__device__ inline void cartesian(const double *co, double xp, double yp, double zp, double &x, double &y, double &z) {
  x = co[0] * xp + co[3] * yp + co[6] * zp;
  y = co[1] * xp + co[4] * yp + co[7] * zp;
  z = co[2] * xp + co[5] * yp + co[8] * zp;
}

// one thread per (atom, z-offset): order^2 fixed-point atomic adds
__global__ void __launch_bounds__(256) k_grid_zero(const PmeDesc *__restrict__ D) {
  const PmeDesc &d = D[blockIdx.y];
  if (d.tiled) return;
  const size_t G = (size_t)d.K1 * d.K2 * d.K3;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < G; i += (size_t)gridDim.x * blockDim.x)
    d.Qfix[i] = 0ull;
}

template <typename T>
__global__ void __launch_bounds__(256) k_spread(const PmeDesc *__restrict__ D) {
  const PmeDesc &d = D[blockIdx.y];
  if (d.tiled) return;
  const int order = d.order;
  int tid = blockIdx.x * blockDim.x + threadIdx.x;
  int a = tid / order, i3 = tid - a * order;
  if (a >= d.natoms) return;
  double x, y, z;
  cartesian(d.co, d.xp[a], d.yp[a], d.zp[a], x, y, z);
  double fr1 = frac_coord(x, y, z, d.recip, 0, d.K1);
  double fr2 = frac_coord(x, y, z, d.recip, 1, d.K2);
  double fr3 = frac_coord(x, y, z, d.recip, 2, d.K3);
  // fr is in [0, K] for finite coordinates; anything else (NaN, inf) is skipped
  if (!(fr1 >= 0.0 && fr1 <= d.K1 && fr2 >= 0.0 && fr2 <= d.K2 && fr3 >= 0.0 && fr3 <= d.K3)) return;
  int b1 = (int)fr1, b2 = (int)fr2, b3 = (int)fr3;
  T th1[MAX_ORDER], th2[MAX_ORDER], th3[MAX_ORDER], dd[MAX_ORDER];
  fill_bspline<T>(T(fr1 - b1), order, th1, dd);
  fill_bspline<T>(T(fr2 - b2), order, th2, dd);
  fill_bspline<T>(T(fr3 - b3), order, th3, dd);
  // same product order as the tiled spread, th1 * (th2 * (th3 * q)), so both
  // paths put identical fixed-point values on the grid
  const T q3 = th3[i3] * T(d.q[a]);
  const int kk = wrapK(b3 - order + i3 + 1, d.K3);
  const double scale = d.fixscale;
  for (int i2 = 0; i2 < order; ++i2) {
    int jj = wrapK(b2 - order + i2 + 1, d.K2);
    T prod = th2[i2] * q3;
    size_t row = (size_t)d.K1 * (jj + (size_t)d.K2 * kk);
    for (int i1 = 0; i1 < order; ++i1) {
      int ii = wrapK(b1 - order + i1 + 1, d.K1);
      atomicAdd(&d.Qfix[row + ii], to_fix((double)(th1[i1] * prod), scale));
    }
  }
}

// ---------------------------------------------------------------- tiled spread
//
// Atoms are binned by the tile of their spline base point b = int(fr).  The
// tile (T1,T2,T3) of st1 x st2 x st3 grid points receives contributions from
// atoms with b in [t0, t0 + st + order - 2] per axis, i.e. from bins T and
// T + 1 (mod nb) when st >= order - 1 and st divides K.  One workgroup per
// tile accumulates those contributions in LDS with int64 atomics and stores
// its points: every grid point has one owner, no global atomics, no clearing.

__global__ void __launch_bounds__(256) k_bin_zero(const PmeDesc *__restrict__ D) {
  const PmeDesc &d = D[blockIdx.y];
  if (!d.tiled) return;
  int nb = d.nb1 * d.nb2 * d.nb3;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) d.bin_cnt[i] = 0;
}

// Neighbouring atoms fall in the same tile, so a wavefront's atoms share a
// handful of bins: lanes of one bin are counted with one atomic by their
// lowest lane, and each lane takes base + its rank among them.
__global__ void __launch_bounds__(256) k_bin_count(const PmeDesc *__restrict__ D) {
  const PmeDesc &d = D[blockIdx.y];
  if (!d.tiled) return;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int bin = -1;
  if (a < d.natoms) {
    double x, y, z;
    cartesian(d.co, d.xp[a], d.yp[a], d.zp[a], x, y, z);
    double fr1 = frac_coord(x, y, z, d.recip, 0, d.K1);
    double fr2 = frac_coord(x, y, z, d.recip, 1, d.K2);
    double fr3 = frac_coord(x, y, z, d.recip, 2, d.K3);
    if (!(fr1 >= 0.0 && fr1 <= d.K1 && fr2 >= 0.0 && fr2 <= d.K2 && fr3 >= 0.0 && fr3 <= d.K3)) {
      d.atom_bin[a] = make_int2(-1, 0);  // non-finite coordinates are not spread
    } else {
      // fr = K (w - anint(w) = 0.5 exactly) is the grid origin: same spline weights
      if (fr1 >= d.K1) fr1 -= d.K1;
      if (fr2 >= d.K2) fr2 -= d.K2;
      if (fr3 >= d.K3) fr3 -= d.K3;
      d.afr[a] = make_double4(fr1, fr2, fr3, d.q[a]);
      int b1 = (int)fr1, b2 = (int)fr2, b3 = (int)fr3;
      bin = b1 / d.st1 + d.nb1 * (b2 / d.st2 + d.nb2 * (b3 / d.st3));
    }
  }
  unsigned long long todo = __ballot(bin >= 0);
  int pos = 0;
  while (todo) {  // the same on every lane of the wavefront
    const int leader = __ffsll((long long)todo) - 1;
    const int lb = __shfl(bin, leader, 64);
    const unsigned long long same = __ballot(bin == lb);
    int base = 0;
    if (lane == leader) base = atomicAdd(&d.bin_cnt[lb], __popcll(same));
    base = __shfl(base, leader, 64);
    if (bin == lb) pos = base + __popcll(same & ((1ull << lane) - 1ull));
    todo &= ~same;
  }
  if (bin >= 0) d.atom_bin[a] = make_int2(bin, pos);
}

// one block per system: exclusive prefix of the bin counts, fixed order
__global__ void __launch_bounds__(1024) k_bin_scan(const PmeDesc *__restrict__ D) {
  const PmeDesc &d = D[blockIdx.x];
  if (!d.tiled) return;
  block_exclusive_scan(d.bin_cnt, d.bin_start, d.nb1 * d.nb2 * d.nb3);
}

// atoms into bin order, with their B-spline weights along the three axes
// (computed once here instead of once per tile and stencil plane)
template <typename T>
__global__ void __launch_bounds__(256) k_bin_fill(const PmeDesc *__restrict__ D) {
  const PmeDesc &d = D[blockIdx.y];
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (!d.tiled || a >= d.natoms) return;
  int2 bp = d.atom_bin[a];
  if (bp.x < 0) return;
  const int s = d.bin_start[bp.x] + bp.y, order = d.order;
  const double4 f = d.afr[a];
  d.sfr[s] = f;
  double *w = d.swt + (size_t)s * 3 * order;
  T th[MAX_ORDER], dd[MAX_ORDER];
  fill_bspline<T>(T(f.x - (int)f.x), order, th, dd);
  for (int i = 0; i < order; ++i) w[i] = (double)th[i];
  fill_bspline<T>(T(f.y - (int)f.y), order, th, dd);
  for (int i = 0; i < order; ++i) w[order + i] = (double)th[i];
  fill_bspline<T>(T(f.z - (int)f.z), order, th, dd);
  for (int i = 0; i < order; ++i) w[2 * order + i] = (double)th[i];
}

constexpr int SPREAD_TILE_MAX = 8;

// The tile-local range of an atom's stencil along one axis.  The stencil
// points are base + i, i = 0..order-1, base = b - order + 1 (+ K when the atom
// comes from the wrapped neighbour bin) minus the tile origin.  With two or
// more tiles along the axis (st >= order - 1) the points inside the tile are
// one range [ia, ib) and none of them wraps; with one tile (st = K) every
// point is inside after a wrap of the negative ones.
struct AxisRange {
  int l0, ia, ib, wrap;
};
__device__ inline AxisRange axis_range(int b, int order, int t0, int st, int K, int wrapped) {
  AxisRange r;
  r.l0 = b - order + 1 + (wrapped ? K : 0) - t0;
  if (st >= K) {
    r.ia = 0;
    r.ib = order;
    r.wrap = K;
  } else {
    r.ia = max(0, -r.l0);
    r.ib = min(order, st - r.l0);
    r.wrap = 0;
  }
  return r;
}

// The tile is held in LDS with padded strides (st1 + 1, st2 + 1): z-planes and
// y-rows of one column then fall on different banks, which the int64 atomics
// of neighbouring stencil points would otherwise serialise on.  Threads take
// (atom, z-plane) pairs, 8 planes per atom (order <= 8), 64 atoms per sweep;
// a pair whose plane lies outside the tile stops at once, the others add
// their in-tile rows.  Local indices need no modulo: the bin offset tells
// which atoms come across the periodic boundary.
constexpr int SPREAD_THREADS = 512;  // 64 atoms x 8 z-planes per sweep

template <typename T>
__global__ void __launch_bounds__(SPREAD_THREADS) k_spread_tile(const PmeDesc *__restrict__ D) {
  __shared__ unsigned long long tile[(SPREAD_TILE_MAX + 1) * (SPREAD_TILE_MAX + 1) * SPREAD_TILE_MAX];
  const PmeDesc &d = D[blockIdx.y];
  if (!d.tiled || (int)blockIdx.x >= d.nb1 * d.nb2 * d.nb3) return;
  const int order = d.order;
  const int T1 = blockIdx.x % d.nb1, T2 = (blockIdx.x / d.nb1) % d.nb2, T3 = blockIdx.x / (d.nb1 * d.nb2);
  const int st1 = d.st1, st2 = d.st2, st3 = d.st3, P1 = st1 + 1, P2 = st2 + 1;
  const int t01 = T1 * st1, t02 = T2 * st2, t03 = T3 * st3;
  const int npts = st1 * st2 * st3;
  for (int i = threadIdx.x; i < P1 * P2 * st3; i += blockDim.x) tile[i] = 0ull;
  __syncthreads();
  const double scale = d.fixscale;
  const int slot = threadIdx.x >> 3, i3 = threadIdx.x & 7;
  for (int sb = 0; sb < 8; ++sb) {
    const int o1 = sb & 1, o2 = (sb >> 1) & 1, o3 = sb >> 2;
    // with a single bin along an axis, the +1 neighbour is the same bin: visit it once
    if ((o1 && d.nb1 == 1) || (o2 && d.nb2 == 1) || (o3 && d.nb3 == 1)) continue;
    const int w1 = T1 + o1 >= d.nb1, w2 = T2 + o2 >= d.nb2, w3 = T3 + o3 >= d.nb3;
    const int B1 = T1 + o1 - (w1 ? d.nb1 : 0), B2 = T2 + o2 - (w2 ? d.nb2 : 0), B3 = T3 + o3 - (w3 ? d.nb3 : 0);
    const int bin = B1 + d.nb1 * (B2 + d.nb2 * B3);
    const int s = d.bin_start[bin], n = d.bin_cnt[bin];
    for (int at = slot; at < n; at += SPREAD_THREADS / 8) {
      if (i3 >= order) continue;
      const double4 f = d.sfr[s + at];
      const AxisRange z = axis_range((int)f.z, order, t03, st3, d.K3, w3);
      int lz = z.l0 + i3;
      if (lz < 0) lz += z.wrap;
      if (lz < 0 || lz >= st3) continue;
      const AxisRange x = axis_range((int)f.x, order, t01, st1, d.K1, w1);
      const AxisRange y = axis_range((int)f.y, order, t02, st2, d.K2, w2);
      const double *wt = d.swt + (size_t)(s + at) * 3 * order;
      const T q3 = T(wt[2 * order + i3]) * T(f.w);
      for (int i2 = y.ia; i2 < y.ib; ++i2) {
        int ly = y.l0 + i2;
        if (ly < 0) ly += y.wrap;
        const T prod = T(wt[order + i2]) * q3;
        unsigned long long *row = tile + P1 * (ly + P2 * lz);
        for (int i1 = x.ia; i1 < x.ib; ++i1) {
          int lx = x.l0 + i1;
          if (lx < 0) lx += x.wrap;
          atomicAdd(&row[lx], to_fix((double)(T(wt[i1]) * prod), scale));
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < npts; i += blockDim.x) {
    const int lx = i % st1, ly = (i / st1) % st2, lz = i / (st1 * st2);
    d.Qfix[(size_t)(t01 + lx) + (size_t)d.K1 * ((t02 + ly) + (size_t)d.K2 * (t03 + lz))] =
        tile[lx + P1 * (ly + P2 * lz)];
  }
}

// tile edge along one axis: a divisor of K in [max(order - 1, 2), 8], the
// largest or the smallest; 0 if none
int spread_tile_edge(int K, int order, bool smallest) {
  const int lo = std::max(2, order - 1);
  if (smallest) {
    for (int s = lo; s <= SPREAD_TILE_MAX; ++s)
      if (K % s == 0) return s;
    return 0;
  }
  for (int s = SPREAD_TILE_MAX; s >= lo; --s)
    if (K % s == 0) return s;
  return 0;
}

// ---------------------------------------------------------------- FFT passes

// Rows along K1: fixed-point grid -> half spectrum.  TP rows per block.
template <typename T>
__global__ void __launch_bounds__(256) k_fft_x_r2c(const PmeDesc *__restrict__ D) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PmeDesc &d = D[blockIdx.y];
  const int N = d.K1, H = N / 2 + 1, rows = d.K2 * d.K3, TP = d.tpx;
  if ((int)blockIdx.x >= d.gx) return;
  cx<T> *a = reinterpret_cast<cx<T> *>(smem);
  cx<T> *b = a + TP * N;
  const int row0 = blockIdx.x * TP;
  const double inv = d.fixinv;
  for (int i = threadIdx.x; i < TP * N; i += blockDim.x) {
    int t = i / N, e = i - t * N;
    int row = row0 + t;
    T v = T(0);
    if (row < rows) v = T((double)(long long)d.Qfix[(size_t)row * N + e] * inv);
    a[i] = {v, T(0)};
  }
  __syncthreads();
  cx<T> *r = lds_fft<T, -1>(a, b, N, TP, d.ax1);
  cx<T> *FQ = reinterpret_cast<cx<T> *>(d.FQ);
  for (int i = threadIdx.x; i < TP * H; i += blockDim.x) {
    int t = i / H, e = i - t * H;
    int row = row0 + t;
    if (row < rows) FQ[(size_t)row * H + e] = r[t * N + e];
  }
}

// Complex pencils along a strided axis of the half grid: pencil p has base
// (p % H) + (p / H) * S1 and element stride Se.
template <typename T, int S>
__global__ void __launch_bounds__(256) k_fft_pencil(const PmeDesc *__restrict__ D, int axis) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PmeDesc &d = D[blockIdx.y];
  const int H = d.K1 / 2 + 1, TP = d.tpy;
  if ((int)blockIdx.x >= d.gy || (S > 0 && d.energy_only)) return;
  const FftAxis &ax = axis == 2 ? d.ax2 : d.ax3;
  const int N = ax.n;
  const int npen = axis == 2 ? H * d.K3 : H * d.K2;
  const size_t S1 = axis == 2 ? (size_t)H * d.K2 : (size_t)H;
  const size_t Se = axis == 2 ? (size_t)H : (size_t)H * d.K2;
  const int p0 = blockIdx.x * TP;
  if (p0 >= npen) return;
  cx<T> *a = reinterpret_cast<cx<T> *>(smem);
  cx<T> *b = a + TP * N;
  cx<T> *FQ = reinterpret_cast<cx<T> *>(d.FQ);
  for (int i = threadIdx.x; i < TP * N; i += blockDim.x) {
    int e = i / TP, t = i - e * TP;
    int p = p0 + t;
    cx<T> v = {T(0), T(0)};
    if (p < npen) v = FQ[(size_t)(p % H) + (size_t)(p / H) * S1 + (size_t)e * Se];
    a[t * N + e] = v;
  }
  __syncthreads();
  cx<T> *r = lds_fft<T, S>(a, b, N, TP, ax);
  for (int i = threadIdx.x; i < TP * N; i += blockDim.x) {
    int e = i / TP, t = i - e * TP;
    int p = p0 + t;
    if (p < npen) FQ[(size_t)(p % H) + (size_t)(p / H) * S1 + (size_t)e * Se] = r[t * N + e];
  }
}

// Pencils along K3: forward FFT, influence function (scalar_sum.f:40-121,
// FFTW variant, with its fac12/fac13/fac23 virial masks and the unscaled
// m = 0 term), inverse FFT.  Per-block partials: energy, v11 v22 v33 v12 v13 v23.
template <typename T>
__global__ void __launch_bounds__(256) k_fft_z_conv(const PmeDesc *__restrict__ D) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[8][8];
  const PmeDesc &d = D[blockIdx.y];
  if ((int)blockIdx.x >= d.conv_nblk) return;
  const int H = d.K1 / 2 + 1, N = d.K3, TP = d.tpz;
  const int npen = H * d.K2;
  const size_t Se = (size_t)H * d.K2;
  const int p0 = blockIdx.x * TP;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  cx<T> *FQ = reinterpret_cast<cx<T> *>(d.FQ);
  cx<T> *a = reinterpret_cast<cx<T> *>(smem);
  cx<T> *b = a + TP * N;
  if (p0 < npen) {
    for (int i = threadIdx.x; i < TP * N; i += blockDim.x) {
      int e = i / TP, t = i - e * TP;
      int p = p0 + t;
      cx<T> v = {T(0), T(0)};
      if (p < npen) v = FQ[(size_t)p + (size_t)e * Se];
      a[t * N + e] = v;
    }
    __syncthreads();
    cx<T> *r = lds_fft<T, -1>(a, b, N, TP, d.ax3);
    cx<T> *s = (r == a) ? b : a;
    const int nfft1 = d.K1, nfft2 = d.K2, nfft3 = d.K3, ndim1 = H;
    int nf1 = nfft1 / 2; if (2 * nf1 < nfft1) nf1 += 1;
    int nf2 = nfft2 / 2; if (2 * nf2 < nfft2) nf2 += 1;
    int nf3 = nfft3 / 2; if (2 * nf3 < nfft3) nf3 += 1;
    const double pi = d.pi_f32, fac = d.fac;
    for (int i = threadIdx.x; i < TP * N; i += blockDim.x) {
      int e = i / TP, t = i - e * TP;
      int p = p0 + t;
      if (p >= npen) continue;
      int k1 = p % H + 1, k2 = p / H + 1, k3 = e + 1;  // 1-based as in scalar_sum
      cx<T> qv = r[t * N + e];
      if (k1 == 1 && k2 == 1 && k3 == 1) continue;     // ind = 0 is never scaled
      int m1 = k1 - 1; if (k1 > nf1) m1 = k1 - 1 - nfft1;
      int m2 = k2 - 1; if (k2 > nf2) m2 = k2 - 1 - nfft2;
      int m3 = k3 - 1; if (k3 > nf3) m3 = k3 - 1 - nfft3;
      const double *rc = d.recip;
      double mhat1 = rc[0] * m1 + rc[3] * m2 + rc[6] * m3;
      double mhat2 = rc[1] * m1 + rc[4] * m2 + rc[7] * m3;
      double mhat3 = rc[2] * m1 + rc[5] * m2 + rc[8] * m3;
      double msq = mhat1 * mhat1 + mhat2 * mhat2 + mhat3 * mhat3;
      double eterm = 0.0;
      if (!(msq > d.rkcut2)) {
        double denom = pi * d.volume * d.bsp1[k1 - 1] * d.bsp2[k2 - 1] * d.bsp3[k3 - 1] * msq;
        eterm = exp(-fac * msq) / denom;
        double vterm = 2.0 * (fac * msq + 1.0) / msq;
        double qr = (double)qv.x, qi = (double)qv.y;
        double struc2, f12 = 1.0, f13 = 1.0, f23 = 1.0;
        if (k1 == 1 || (k1 == ndim1 && (nfft1 & 1) == 0)) {
          struc2 = eterm * qr * qr + eterm * qi * qi;
        } else {
          double e2 = 2.0 * eterm;
          struc2 = e2 * qr * qr + e2 * qi * qi;
          f12 = (mhat1 * mhat2 < 0.0 && k2 <= ndim1) ? 0.0 : 1.0;
          f13 = (mhat1 * mhat3 < 0.0 && k3 <= ndim1) ? 0.0 : 1.0;
          bool neg = mhat2 * mhat3 < 0.0;
          bool lo2 = k2 <= ndim1, lo3 = k3 <= ndim1;
          f23 = (lo2 == lo3) ? (neg ? 0.0 : 1.0) : (neg ? 1.0 : 0.0);
        }
        double vvir = struc2 * vterm;
        acc[0] += struc2;
        acc[1] += struc2 * (vterm * mhat1 * mhat1 - 1.0);
        acc[2] += struc2 * (vterm * mhat2 * mhat2 - 1.0);
        acc[3] += struc2 * (vterm * mhat3 * mhat3 - 1.0);
        acc[4] += vvir * mhat1 * mhat2 * f12;
        acc[5] += vvir * mhat1 * mhat3 * f13;
        acc[6] += vvir * mhat2 * mhat3 * f23;
      }
      r[t * N + e] = {T((double)qv.x * eterm), T((double)qv.y * eterm)};
    }
    __syncthreads();
    if (!d.energy_only) {
      cx<T> *o = lds_fft<T, +1>(r, s, N, TP, d.ax3);
      for (int i = threadIdx.x; i < TP * N; i += blockDim.x) {
        int e = i / TP, t = i - e * TP;
        int p = p0 + t;
        if (p < npen) FQ[(size_t)p + (size_t)e * Se] = o[t * N + e];
      }
    }
  }
  // fixed-order block reduction of the 7 partial sums
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    double v = wave_sum(acc[k]);
    if (lane == 0) red[wid][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w][threadIdx.x];
    d.conv_part[(size_t)blockIdx.x * 8 + threadIdx.x] = s;
  }
}

// Half spectrum rows -> real rows (imaginary parts of k1 = 0 and Nyquist ignored).
template <typename T>
__global__ void __launch_bounds__(256) k_fft_x_c2r(const PmeDesc *__restrict__ D) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PmeDesc &d = D[blockIdx.y];
  const int N = d.K1, H = N / 2 + 1, rows = d.K2 * d.K3, TP = d.tpx;
  if ((int)blockIdx.x >= d.gx || d.energy_only) return;
  cx<T> *a = reinterpret_cast<cx<T> *>(smem);
  cx<T> *b = a + TP * N;
  const int row0 = blockIdx.x * TP;
  const cx<T> *FQ = reinterpret_cast<const cx<T> *>(d.FQ);
  for (int i = threadIdx.x; i < TP * H; i += blockDim.x) {
    int t = i / H, e = i - t * H;
    int row = row0 + t;
    cx<T> v = {T(0), T(0)};
    if (row < rows) v = FQ[(size_t)row * H + e];
    if (e == 0 || 2 * e == N) v.y = T(0);
    a[t * N + e] = v;
    if (e > 0 && N - e >= H) a[t * N + (N - e)] = {v.x, -v.y};
  }
  __syncthreads();
  cx<T> *r = lds_fft<T, +1>(a, b, N, TP, d.ax1);
  T *Qr = reinterpret_cast<T *>(d.Qr);
  for (int i = threadIdx.x; i < TP * N; i += blockDim.x) {
    int t = i / N, e = i - t * N;
    int row = row0 + t;
    if (row < rows) Qr[(size_t)row * N + e] = r[t * N + e].x;
  }
}

// ---------------------------------------------------------------- gather

// `order` lanes per atom, 64 / order atoms per wavefront: lane i3 sums the
// z-plane i3 of the atom's order^3 stencil, and the partial f1, f2, f3 are
// summed over the atom's lanes in plane order (fixed order).
template <typename T>
__global__ void __launch_bounds__(256) k_gather(const PmeDesc *__restrict__ D) {
  __shared__ double red[8][9];
  const PmeDesc &d = D[blockIdx.y];
  if ((int)blockIdx.x >= d.gather_nblk || d.energy_only) return;
  const int order = d.order;
  const int apw = 64 / order;
  const int lane = threadIdx.x & 63;
  const int slot = lane / order, sub = lane - slot * order;
  int a = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * apw + slot;
  if (slot >= apw) a = d.natoms;  // lanes past the last whole atom of the wave
  double stc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double xp = 0, yp = 0, zp = 0, fr1 = -1, fr2 = -1, fr3 = -1;
  if (a < d.natoms) {
    xp = d.xp[a];
    yp = d.yp[a];
    zp = d.zp[a];
    double x, y, z;
    cartesian(d.co, xp, yp, zp, x, y, z);
    fr1 = frac_coord(x, y, z, d.recip, 0, d.K1);
    fr2 = frac_coord(x, y, z, d.recip, 1, d.K2);
    fr3 = frac_coord(x, y, z, d.recip, 2, d.K3);
  }
  const bool valid =
      a < d.natoms && fr1 >= 0.0 && fr1 <= d.K1 && fr2 >= 0.0 && fr2 <= d.K2 && fr3 >= 0.0 && fr3 <= d.K3;
  double f1 = 0.0, f2 = 0.0, f3 = 0.0;
  int b1 = 0, b2 = 0, b3 = 0;
  if (valid) {
    b1 = (int)fr1;
    b2 = (int)fr2;
    b3 = (int)fr3;
    T th1[MAX_ORDER], th2[MAX_ORDER], th3[MAX_ORDER], dt1[MAX_ORDER], dt2[MAX_ORDER], dt3[MAX_ORDER];
    fill_bspline<T>(T(fr1 - b1), order, th1, dt1);
    fill_bspline<T>(T(fr2 - b2), order, th2, dt2);
    fill_bspline<T>(T(fr3 - b3), order, th3, dt3);
    const T *Qr = reinterpret_cast<const T *>(d.Qr);
    const int i3 = sub;
    int kk = wrapK(b3 - order + i3 + 1, d.K3);
    for (int i2 = 0; i2 < order; ++i2) {
      int jj = wrapK(b2 - order + i2 + 1, d.K2);
      const T *row = Qr + (size_t)d.K1 * (jj + (size_t)d.K2 * kk);
      T s1 = T(0), s23 = T(0);
      for (int i1 = 0; i1 < order; ++i1) {
        int ii = wrapK(b1 - order + i1 + 1, d.K1);
        T g = row[ii];
        s1 += dt1[i1] * g;
        s23 += th1[i1] * g;
      }
      f1 -= (double)(s1 * th2[i2] * th3[i3]);
      f2 -= (double)(s23 * dt2[i2] * th3[i3]);
      f3 -= (double)(s23 * th2[i2] * dt3[i3]);
    }
  }
  {
    // every lane takes part in the shuffles; the atom's lanes are slot*order + k
    const int src0 = (slot < apw ? slot : 0) * order;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int k = 0; k < order; ++k) {
      s1 += __shfl(f1, src0 + k, 64);
      s2 += __shfl(f2, src0 + k, 64);
      s3 += __shfl(f3, src0 + k, 64);
    }
    f1 = s1;
    f2 = s2;
    f3 = s3;
  }
  if (valid && sub == 0) {
    const double q = d.q[a];
    f1 *= d.K1 * q;
    f2 *= d.K2 * q;
    f3 *= d.K3 * q;
    const double *rc = d.recip;
    double fx = rc[0] * f1 + rc[3] * f2 + rc[6] * f3;
    double fy = rc[1] * f1 + rc[4] * f2 + rc[7] * f3;
    double fz = rc[2] * f1 + rc[5] * f2 + rc[8] * f3;
    if (d.accumulate) {
      d.fx[a] += fx;
      d.fy[a] += fy;
      d.fz[a] += fz;
    } else {
      d.fx[a] = fx;
      d.fy[a] = fy;
      d.fz[a] = fz;
    }
    if (d.atomp) {
      int j = d.atomp[a];
      double sx = xp - d.xpcm[j], sy = yp - d.ypcm[j], sz = zp - d.zpcm[j];
      const double *co = d.co;
      double gx = sx * co[0] + sy * co[3] + sz * co[6];
      double gy = sx * co[1] + sy * co[4] + sz * co[7];
      double gz = sx * co[2] + sy * co[5] + sz * co[8];
      stc[0] = -gx * fx; stc[1] = -gy * fx; stc[2] = -gz * fx;
      stc[3] = -gx * fy; stc[4] = -gy * fy; stc[5] = -gz * fy;
      stc[6] = -gx * fz; stc[7] = -gy * fz; stc[8] = -gz * fz;
    }
  }
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    double v = wave_sum(stc[k]);
    if (lane == 0) red[wid][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 9) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w][threadIdx.x];
    d.gather_part[(size_t)blockIdx.x * 9 + threadIdx.x] = s;
  }
}

// d.out[20]: eer, vir(3x3 col-major), stc(3x3 row-major as stc1..stc9).
// Block k reduces one component in fixed order: k < 7 the convolution
// partials (scalar_sum.f:125-133 halving and symmetrisation), k >= 7 the
// molecular-virial partials.
__global__ void __launch_bounds__(256) k_pme_finalize(const PmeDesc *__restrict__ D) {
  __shared__ double red[4];
  const PmeDesc &d = D[blockIdx.y];
  const int k = blockIdx.x;
  if (k >= 7 && d.energy_only) {  // no gather: no molecular-virial terms
    if (threadIdx.x == 0) d.out[10 + (k - 7)] = 0.0;
    return;
  }
  const double *src = k < 7 ? d.conv_part + k : d.gather_part + (k - 7);
  const int nrow = k < 7 ? d.conv_nblk : d.gather_nblk, stride = k < 7 ? 8 : 9;
  double s = 0.0;
  for (int r = threadIdx.x; r < nrow; r += blockDim.x) s += src[(size_t)r * stride];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x != 0) return;
  s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  double *o = d.out;
  switch (k) {
    case 0: o[0] = 0.5 * s; break;
    case 1: o[1] = 0.5 * s; break;
    case 2: o[5] = 0.5 * s; break;
    case 3: o[9] = 0.5 * s; break;
    case 4: o[4] = o[2] = 0.5 * s; break;  // vir(1,2) = vir(2,1)
    case 5: o[7] = o[3] = 0.5 * s; break;  // vir(1,3) = vir(3,1)
    case 6: o[8] = o[6] = 0.5 * s; break;  // vir(2,3) = vir(3,2)
    default: o[10 + (k - 7)] = s; break;
  }
}

// ---------------------------------------------------------------- launchers

static int tile_pencils(int n, int npen, int elems) {
  int tp = elems / n;
  if (tp < 1) tp = 1;
  if (tp > npen) tp = npen;
  return tp;
}

// FFT tiling of one system for LDS buffers of lds_elems complex elements
void pme_set_tiles(PmeDesc &d, int lds_elems) {
  const int H = d.K1 / 2 + 1;
  d.tpx = tile_pencils(d.K1, d.K2 * d.K3, lds_elems);
  d.gx = (d.K2 * d.K3 + d.tpx - 1) / d.tpx;
  d.tpy = tile_pencils(d.K2, H * d.K3, lds_elems);
  d.gy = (H * d.K3 + d.tpy - 1) / d.tpy;
  d.tpz = tile_pencils(d.K3, H * d.K2, lds_elems);
  d.conv_nblk = (H * d.K2 + d.tpz - 1) / d.tpz;
  const int apb = 4 * (64 / d.order);     // k_gather: order lanes per atom, 4 waves per block
  d.gather_nblk = (d.natoms + apb - 1) / apb;
}

template <typename T>
void launch_pme(const PmeDesc *d_desc, const PmeDesc *h, int nsys, hipStream_t st, const Tuning &tu,
                const hipEvent_t *ev) {
  (void)tu;
  if (nsys <= 0) return;
  int max_atoms = 1, max_bins = 1, any_tiled = 0, any_atomic = 0, max_gx = 1, max_gy = 1, max_gz = 1, max_gg = 1;
  size_t max_grid = 1, shx = 0, shy = 0, shz = 0, max_work = 1;
  for (int s = 0; s < nsys; ++s) {
    const PmeDesc &d = h[s];
    max_atoms = std::max(max_atoms, d.natoms);
    max_gx = std::max(max_gx, d.gx);
    max_gy = std::max(max_gy, d.gy);
    max_gz = std::max(max_gz, d.conv_nblk);
    max_gg = std::max(max_gg, d.gather_nblk);
    shx = std::max(shx, 2 * (size_t)d.tpx * d.K1 * sizeof(cx<T>));
    shy = std::max(shy, 2 * (size_t)d.tpy * d.K2 * sizeof(cx<T>));
    shz = std::max(shz, 2 * (size_t)d.tpz * d.K3 * sizeof(cx<T>));
    if (d.tiled) {
      any_tiled = 1;
      max_bins = std::max(max_bins, d.nb1 * d.nb2 * d.nb3);
    } else {
      any_atomic = 1;
      max_grid = std::max(max_grid, (size_t)d.K1 * d.K2 * d.K3);
      max_work = std::max(max_work, (size_t)d.natoms * d.order);
    }
  }
  dim3 blk(256);
  if (ev) ORAC_HIP(hipEventRecord(ev[0], st));
  if (any_tiled) {
    dim3 ga((max_atoms + 255) / 256, nsys);
    hipLaunchKernelGGL(k_bin_zero, dim3((max_bins + 255) / 256, nsys), blk, 0, st, d_desc);
    hipLaunchKernelGGL(k_bin_count, ga, blk, 0, st, d_desc);
    hipLaunchKernelGGL(k_bin_scan, dim3(nsys), dim3(1024), 0, st, d_desc);
    hipLaunchKernelGGL(k_bin_fill<T>, ga, blk, 0, st, d_desc);
    hipLaunchKernelGGL(k_spread_tile<T>, dim3(max_bins, nsys), dim3(SPREAD_THREADS), 0, st, d_desc);
  }
  if (any_atomic) {
    unsigned gzero = (unsigned)std::min<size_t>((max_grid + 255) / 256, 4096);
    hipLaunchKernelGGL(k_grid_zero, dim3(gzero, nsys), blk, 0, st, d_desc);
    hipLaunchKernelGGL(k_spread<T>, dim3((unsigned)((max_work + 255) / 256), nsys), blk, 0, st, d_desc);
  }
  if (ev) ORAC_HIP(hipEventRecord(ev[1], st));
  hipLaunchKernelGGL(k_fft_x_r2c<T>, dim3(max_gx, nsys), blk, shx, st, d_desc);
  hipLaunchKernelGGL((k_fft_pencil<T, -1>), dim3(max_gy, nsys), blk, shy, st, d_desc, 2);
  hipLaunchKernelGGL(k_fft_z_conv<T>, dim3(max_gz, nsys), blk, shz, st, d_desc);
  hipLaunchKernelGGL((k_fft_pencil<T, +1>), dim3(max_gy, nsys), blk, shy, st, d_desc, 2);
  hipLaunchKernelGGL(k_fft_x_c2r<T>, dim3(max_gx, nsys), blk, shx, st, d_desc);
  if (ev) ORAC_HIP(hipEventRecord(ev[2], st));
  hipLaunchKernelGGL(k_gather<T>, dim3(max_gg, nsys), blk, 0, st, d_desc);
  if (ev) ORAC_HIP(hipEventRecord(ev[3], st));
  hipLaunchKernelGGL(k_pme_finalize, dim3(16, nsys), dim3(256), 0, st, d_desc);
}

// The reciprocal forces of a system evaluated beside its direct sum are
// written to their own arrays; this adds them to the direct forces once both
// are done, so the sum is the same as accumulating in sequence.
__global__ void __launch_bounds__(256) k_pme_combine(const PmeDesc *__restrict__ D) {
  const PmeDesc &d = D[blockIdx.y];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (!d.fx_dst || a >= d.natoms) return;
  double x = d.fx_dst[a] + d.fx[a], y = d.fy_dst[a] + d.fy[a], z = d.fz_dst[a] + d.fz[a];
  if (d.cfx) {  // the corrections, in the order of a sequential evaluation
    x += d.cfx[a];
    y += d.cfy[a];
    z += d.cfz[a];
  }
  d.fx_dst[a] = x;
  d.fy_dst[a] = y;
  d.fz_dst[a] = z;
}

void launch_pme_combine(const PmeDesc *d_desc, const PmeDesc *h, int nsys, hipStream_t st) {
  int max_atoms = 1;
  for (int s = 0; s < nsys; ++s)
    if (h[s].fx_dst) max_atoms = std::max(max_atoms, h[s].natoms);
  hipLaunchKernelGGL(k_pme_combine, dim3((max_atoms + 255) / 256, nsys), dim3(256), 0, st, d_desc);
}

template void launch_pme<double>(const PmeDesc *, const PmeDesc *, int, hipStream_t, const Tuning &,
                                 const hipEvent_t *);
template void launch_pme<float>(const PmeDesc *, const PmeDesc *, int, hipStream_t, const Tuning &,
                                const hipEvent_t *);

}  // namespace orac
