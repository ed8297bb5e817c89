// kernels.hpp -- device descriptors and host launchers shared by the .hip files.
#pragma once
#include "common.hpp"

namespace orac {

struct PmeTiles {
  int lds_elems = 512;  // complex elements per LDS buffer in the FFT kernels
};

template <typename T>
void launch_pme(const PmeDesc *d_desc, const PmeDesc &h, int nsys, double *d_out, hipStream_t st,
                const PmeTiles &tiles);
int conv_blocks(int K1, int K2, int K3, int lds_elems);
// tile edge of the LDS-tiled spread along an axis of K points (0: use the atomic spread)
int spread_tile_edge(int K, int order);

// ---------------------------------------------------------------- direct space

// One system's direct-space state (force.f semantics).  All indices 0-based.
struct PairDesc {
  int natoms, ngrp;
  const double *xp, *yp, *zp;       // reduced coordinates
  const double *q;
  const int *type;                  // LJ type of each atom
  const double *mass, *gmass;
  const int *atomp;                 // molecule of each atom
  const double *xpcm, *ypcm, *zpcm; // molecule centres (reduced)
  const int *gfirst, *glast;        // atom range of each group
  const int *gss;                   // ss_index of the group's first atom (1 solute, 2 solvent)
  const double *xpg, *ypg, *zpg;    // group centres
  int ntypes;
  const int *type_table;            // ma x ma, 0-based packed index
  const double *ecc12, *ecc6;
  const double2 *lj_dense;          // ma x ma {ecc12, ecc6}, or null
  const int *ex_off, *ex_idx;       // symmetric exclusion CSR (sorted partners)
  const unsigned char *ex_inter;    // atom has an excluded partner outside its own group
  int any_heavy;
  double massinfty;
  const int *nb_off, *nb_idx;       // full neighbour list of each group (CSR)
  double co[9];
  double alpha, twrtpi;
  double r2inn, r2out, rinn0, r2inn0, rout0, r2out0, arsout1, arsout2, arsinn1, arsinn2;
  int lskip_ewald, ewald;
  double *fx, *fy, *fz;             // direct forces (written)
  double *part;                     // per-block partials [nblk][16]: uconf[3] ucoul[3] st[9]
  int nblk;
  double4 *apos;                    // packed {xp, yp, zp, q} per atom (k_pack_atoms)
  int2 *ainfo;                      // packed {type, flags} per atom; flags bit 0 = ex_inter
};

constexpr int PAIR_WAVES = 4;       // owner groups per 256-thread block

template <typename T>
void launch_pairs(const PairDesc *d_desc, const PairDesc &h, int nsys, int max_group, hipStream_t st);
void launch_pairs_finalize(const PairDesc *d_desc, int nsys, double *d_out, hipStream_t st);

// group centres (appbou.f) and molecule centres of mass
void launch_group_centres(int ngrp, const int *gfirst, const int *glast, const double *pmass, const double *xp,
                          const double *yp, const double *zp, double *xpg, double *ypg, double *zpg, hipStream_t st);
void launch_mol_com(int nmol, const int *mfirst, const int *mlast, const double *mass, const double *xp,
                    const double *yp, const double *zp, double *xpcm, double *ypcm, double *zpcm, hipStream_t st);

// linked cells (lc_list.f)
struct CellGrid {
  int ncx, ncy, ncz;
  int nsten;                 // stencil entries (full sphere, entry 0 = (0,0,0))
  const int *sten;           // [nsten][3]
  double co[9];
  double sqcut;
  int half;                  // 1: l > n only (reference lc_list); 0: every l != n
  double4 *cpos;             // [ngrp] {xpg, ypg, zpg, group} in cell order (filled by the cell build)
};
void launch_cell_build(int ngrp, const double *xpg, const double *ypg, const double *zpg, const CellGrid &cg,
                       int *cell_of, int *cell_cnt, int *cell_start, int *cell_cursor, int *cell_items,
                       hipStream_t st);
void launch_list_count(int ngrp, const double *xpg, const double *ypg, const double *zpg, const CellGrid &cg,
                       const int *cell_of, const int *cell_start, const int *cell_cnt, const int *cell_items,
                       int *count, hipStream_t st);
// mode 0: CSR 0-based (off[n], idx), mode 1: reference layout [nv, l..] 1-based at off[n]
void launch_list_write(int ngrp, const double *xpg, const double *ypg, const double *zpg, const CellGrid &cg,
                       const int *cell_of, const int *cell_start, const int *cell_cnt, const int *cell_items,
                       const int *off, int *out, int mode, hipStream_t st);
void launch_exclusive_scan(const int *in, int *out, int n, int *total, hipStream_t st);
// off[n] = total[0]; if total[0] > cap: total[1] = 1 and all offsets zero, so nothing is written or read
void launch_list_guard(int n, int *off, int *total, int cap, hipStream_t st);

// half list (CSR) -> full list (CSR) with deterministic order
void launch_half_to_full(int ngrp, const int *h_off, const int *h_idx, int *incount, int *f_off, int *f_idx,
                         int *cursor, int *total, hipStream_t st);

// next-inner-shell list written by force.f:259-286
void launch_inner_list(int ngrp, int iz, const int *h_off, const int *h_idx, const double *xpg, const double *ypg,
                       const double *zpg, const double *co, double r2neigh, int *count, const int *off, int *out,
                       int pass, hipStream_t st);

// ferrf / fnbgrp style pair lists: per pair kernel then per atom gather
struct CorrDesc {
  int npairs;
  const int *ia, *ib;                 // 0-based atoms
  const double *x, *y, *z;            // cartesian coordinates (or null: use reduced + co)
  const double *xp, *yp, *zp;
  double co[9];
  const double *q;
  const int *ss;                      // ss_index (1/2) of each atom
  double alpha, twrtpi, fudge;
  const double *fudge_p;              // per-pair fudge (overrides fudge when set)
  int iz;                             // 1: forces + energy, 0: energy only
  int kind;                           // 0: ferrf, 1: fnbgrp (LJ + Ewald/Coulomb)
  int ewald;
  const int *type;                    // fnbgrp: per-atom LJ type (1-based, as fnbgrp.f uses)
  const double *ecc12, *ecc6;
  double *pf;                         // per pair force vector (3 per pair)
  double *pe;                         // per pair energies (4 per pair: coul slt, coul slv, lj slt, lj slv)
  const int *g_off, *g_pair;          // per atom gather CSR: signed pair index (+/-(p+1))
  int natoms;
  double *fx, *fy, *fz;               // accumulated (+=)
};
void launch_corr(const CorrDesc &h, CorrDesc *d_desc, hipStream_t st);
void launch_corr_energy(const CorrDesc *d_desc, double *d_out, hipStream_t st);

void launch_cself(int n, const double *q, const int *ss, double *d_out2, hipStream_t st);
void launch_zero_u64(unsigned long long *p, size_t n, hipStream_t st);
void launch_zero_f64(double *p, size_t n, hipStream_t st);
void launch_add3(int n, double *fx, double *fy, double *fz, const double *ax, const double *ay, const double *az,
                 hipStream_t st);

}  // namespace orac
