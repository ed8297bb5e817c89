// kernels.hpp -- device descriptors and host launchers shared by the .hip files.
//
// Every launcher takes a device array of descriptors and the same array on
// the host, one entry per system.  Grids are sized by the largest system of
// the batch; blockIdx.y selects the system and each kernel bounds its block
// range by that system's own sizes, so systems of different sizes share a
// launch and each system's result does not depend on what it is batched with.
#pragma once
#include "common.hpp"

namespace orac {

// Launch choices that do not change any result (tile counts of the FFT
// partial sums aside, which change the energy only in its last bits).
struct Tuning {
  int lds_elems = 512;  // complex elements per LDS buffer in the FFT kernels
  int pair_waves = 0;   // occupancy target of the pair kernel (0: default for the precision)
  int spread_tiled = 1; // LDS-tiled spread where the grid allows it (1 largest tiles, 2 smallest),
                        // 0 global atomics
  int xcd_remap = 1;    // keep each system's pair blocks on one XCD (L2 locality)
  int cell_order = 0;   // pair kernel owners and neighbour records in cell order (1) or group order (0)
  int persistent = 0;   // pair kernel as a fixed grid looping over the batch's work items
};

// ---------------------------------------------------------------- reciprocal space

void pme_set_tiles(PmeDesc &d, int lds_elems);
template <typename T>
// ev (optional, 4 events): recorded before the spread, before the forward FFT,
// before the gather and after it (the stage timings of a profiled evaluation)
void launch_pme(const PmeDesc *d_desc, const PmeDesc *h, int nsys, hipStream_t st, const Tuning &tu,
                const hipEvent_t *ev = nullptr);
// fx_dst += fx (and y, z) for every descriptor that has a destination
void launch_pme_combine(const PmeDesc *d_desc, const PmeDesc *h, int nsys, hipStream_t st);
// tile edge of the LDS-tiled spread along an axis of K points (0: use the atomic spread)
int spread_tile_edge(int K, int order, bool smallest);
// Waves per owner group of the pair kernel for a system of ngrp groups (1
// unless ORAC_HIP_PAIR_SPLIT says 2 or 4 when the system is created).  A
// property of the system, never of a batch or a tuning: it fixes the
// summation order.
int pair_split_for(int ngrp);

// ---------------------------------------------------------------- direct space

// One system's direct-space state (force.f semantics).  All indices 0-based.
struct PairDesc {
  int natoms, ngrp;
  const double *xp, *yp, *zp;       // reduced coordinates
  const double *q;
  const int *type;                  // LJ type of each atom
  const double *mass, *gmass;       // atom mass, mass fraction within its group
  const int *atomp;                 // molecule of each atom
  double *xpcm, *ypcm, *zpcm;       // molecule centres (reduced), written by the centre kernels
  int nmol;
  const int *mfirst, *mlast;        // atom range of each molecule
  const int *mbig;                  // molecules of more than MOL_SMALL atoms (one wavefront each)
  int nbig;
  const int *gfirst, *glast;        // atom range of each group
  const int *gss;                   // ss_index of the group's first atom (1 solute, 2 solvent)
  double *xpg, *ypg, *zpg;          // group centres, written by the centre kernels
  int ntypes;
  const int *type_table;            // ma x ma, 0-based packed index
  const double *ecc12, *ecc6;
  const double2 *lj_dense;          // ma x ma {ecc12, ecc6}, or null
  const int *ex_off, *ex_idx;       // symmetric exclusion CSR (sorted partners)
  const unsigned char *ex_inter;    // atom has an excluded partner outside its own group
  int any_heavy;
  double massinfty;
  const int *nb_off, *nb_idx;       // full neighbour list (CSR), owners and entries in the list's order
  const int *nb_cnt;                // slab layout (nb_slab > 0): owner t's entries at nb_idx[t * nb_slab],
  int nb_slab;                      //   nb_cnt[t] of them (at most nb_slab are stored)
  double co[9];
  double alpha, twrtpi;
  double r2inn, r2out, rinn0, r2inn0, rout0, r2out0, arsout1, arsout2, arsinn1, arsinn2;
  int lskip_ewald, ewald;
  double *fx, *fy, *fz;             // direct forces (written)
  double *part;                     // per-block partials [nblk][16]: uconf[3] ucoul[3] st[9]
  int nblk;
  int split;                        // waves per owner group (1, 2 or 4), fixed per system
  // Cell-ordered copies written every evaluation by k_pack_sorted.  Sorted
  // position t holds group perm[t] (t itself without a permutation), whose
  // atoms start at sorted atom soff[t] (gfirst[t] without).
  const int *perm;
  const int *soff;
  double4 *sgp;                     // [ngrp] {xpg, ypg, zpg, 0}
  int4 *sgi;                        // [ngrp] {sorted first atom, first atom, natoms | gss << 8, group}
  double4 *sap;                     // [natoms] {xp, yp, zp, q}
  float4 *sapf;                     // [natoms] {Cartesian offset from the group centre, q}: mixed-precision pairs
  double4 *sapc;                    // [natoms] the same in FP64 (FP64 pairs of the owner-atom loop form)
  // Half-list form (half = 1, split list layout): the list holds each group
  // pair once (forward entries, upper group l > owner); the forces on the
  // neighbour's atoms go to per-pair slots, mirror[t * nb_slab + e] the first
  // slot of forward entry e, and each group sums its reverse entries' slots
  // (nb_rcount, nb_rbase) in the order of its sorted reverse list.
  int half;
  const int *nb_mirror;
  const int *nb_rcount, *nb_rbase;
  double *jpx, *jpy, *jpz;          // neighbour-side force slots
  long long pcap;                   //   (pcap of them)
  double *part2;                    // [nblk][9] molecular virial of the total forces (half form)
  // REM segment scaling (force_REM.f): rematom[a] != 0 marks the segment's
  // atoms (original numbering); the pairs rem_kind picks (0 either atom, 1
  // both, 2 exactly one, rem.f90:631-645) have their force scaled by rem_f3
  // and their energy summed in out[15]; null: no REM
  const int *rematom;
  int rem_kind;
  double rem_f3;
  const float2 *lj_dense32;         // the dense LJ table in FP32 (mixed-precision pairs), or null
  int2 *sai;                        // [natoms] {type, flags}; flags bit 0 = ex_inter
  double *out;                      // [16]: uconf[3], ucoul[3] by typeij, st[9] row-major, dwrk
  int max_group;                    // largest group (atoms)
  // Alchemy (force_alchemy.f): with alch_pair set, `type` is each atom's
  // alchemical class (distinct rmin, epsm, ss_alchemy) and the pairs take the
  // soft-core form.
  const double4 *alch_pair;         // [ntypes x ntypes] {1 / sigma^6, epsmin, typeaij + 4, sigma}
  const double *alch_fac;           // [9][8] by typeaij + 4: ljsoft f1, f2, ljsoft f1_0, f2_0,
                                    //   qsoft fq1^2, fq2, qsoft fq1_0^2, fq2_0 (soft core); shifted:
                                    //   ljshift 2^(1/6) f1, f2, (lambda0), qshift fq1, fq2, (lambda0)
  int alch_shifted;                 // 1: shifted potential (force_alchemy_shifted.f)
};

constexpr int PAIR_WAVES = 4;       // owner groups per 256-thread block
constexpr int OWN_REG = 3;          // largest owner group of the owner-atom loop form (k_pairs_own)
// whether the owner-atom loop form (and with it the half-list layout) is in use
bool use_owner_kernel();
constexpr int MOL_SMALL = 16;       // molecules up to this size get one thread in k_mol_com
// LJ parameters as a dense ntypes x ntypes table of {c12, c6} (one gather per
// pair) up to this many types (64 MB); beyond it the packed table and
// type_table of the reference (two dependent gathers per pair)
constexpr int LJ_DENSE_MAX = 2048;

// kbeg/kend (optional) are recorded right before and after the pair kernel itself
template <typename T>
void launch_pairs(const PairDesc *d_desc, const PairDesc *h, int nsys, hipStream_t st, const Tuning &tu,
                  hipEvent_t kbeg = nullptr, hipEvent_t kend = nullptr);

// group centres (appbou.f) and molecule centres of mass of every system
void launch_centres(const PairDesc *d_desc, const PairDesc *h, int nsys, hipStream_t st);

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

// One system's list build: cells of the group centres, then the stencil walk.
struct ListDesc {
  int ngrp;
  const double *xpg, *ypg, *zpg;
  CellGrid cg;
  int *cell_of, *cell_cnt, *cell_start, *cell_cursor, *cell_items;
  int *cell_fill;            // [ngrp] cell members in arrival order, before the sort
  int *count;                // [ngrp] entries per group
  int *off;                  // [ngrp+1] CSR offsets (mode 0) / prefix of count (mode 1)
  int *out;                  // list entries
  long long cap;             // capacity of out (ints, count slots included in mode 1)
  int *total;                // [0] entries (count slots excluded), [1] overflow flag, [2] largest count
  int mode;                  // 0: CSR, 0-based; 1: reference layout per group [nv, l1..lnv], 1-based
  const int *step;           // optional step counter of the system: the build runs only
  int every;                 //   when *step % every == 0 (device-side update schedule)
  // mode 0 only.  cellorder = 1: owners are taken in cell order (cell_items)
  // and entries are cell-order positions, soff[t] = first atom of the t-th
  // group in that order; cellorder = 0: owners and entries are group numbers.
  int cellorder;
  const int *gsize;          // [ngrp] atoms per group (original numbering)
  int *soff;
  // mode 0 only.  slab > 0: one pass writes owner o's entries at out[o * slab]
  // and its count at count[o]; a count above slab sets the overflow flag.
  int slab;
  // split slab layout (split = 1): out / count hold the forward entries (group
  // l > owner) in walk order; rout[o * slab] the reverse entries (l < owner)
  // as ascending group numbers, rcount[o] of them; rbase[o] the first per-pair slot of
  // o's reverse entries (owner-major, atom-minor); mirror[o * slab + e] the
  // slot of forward entry e; pcap the slots available
  int split;
  int *rout, *rcount, *rbase, *mirror;
  // inner list (two shells): the entries of owner o whose centre distance is
  // below sqrt(r2m), in list order, at out_m[o * slab], count_m[o] of them
  int *out_m, *count_m;
  double r2m;
  long long pcap;
};
enum ListPhase { LIST_COUNT = 1, LIST_WRITE = 2 };
// LIST_COUNT: cells, per-group counts, offsets and the capacity check;
// LIST_WRITE: the entries.  On overflow total[1] = 1 and nothing is written.
void launch_list_build(const ListDesc *d_desc, const ListDesc *h, int nsys, int phases, hipStream_t st);
// the inner lists of a batch's systems after a slab-layout build (systems with r2m > 0)
void launch_list_inner(const ListDesc *d_desc, const ListDesc *h, int nsys, hipStream_t st);
// dense {ecc12, ecc6} table (and its FP32 copy if dense32) from the packed
// type_table (ma x ma, 1-based, column-major) and ecc12/ecc6
void launch_lj_dense(int ma, const int *tt, const double *e12, const double *e6, double2 *dense, float2 *dense32,
                     hipStream_t st);
void launch_exclusive_scan(const int *in, int *out, int n, int *total, hipStream_t st);

// half list (CSR) -> full list (CSR) with deterministic order
void launch_half_to_full(int ngrp, const int *h_off, const int *h_idx, int *incount, int *f_off, int *f_idx,
                         int *cursor, int *total, hipStream_t st);

// next-inner-shell list written by force.f:259-286
void launch_inner_list(int ngrp, int iz, const int *h_off, const int *h_idx, const double *xpg, const double *ypg,
                       const double *zpg, const double *co, double r2neigh, int *count, const int *off, int *out,
                       int pass, hipStream_t st);

// ---------------------------------------------------------------- corrections

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
  int kind;                           // 0: ferrf, 1: fnbgrp (LJ + Ewald/Coulomb), 2: fnb14 (fudged 1-4 pairs)
  int ewald;
  const int *type;                    // fnbgrp: per-atom LJ type (1-based, as fnbgrp.f uses)
  const double *ecc12, *ecc6;
  double *pf;                         // per pair force vector (3 per pair)
  double *pe;                         // per pair energies (4 per pair: coul slt, coul slv, lj slt, lj slv)
  const int *g_off, *g_pair;          // per atom gather CSR: signed pair index (+/-(p+1))
  int natoms;
  double *fx, *fy, *fz;               // accumulated (+=), or written (=) when write is set
  int write;
  double *out;                        // [4]: coul slt, coul slv, lj slt, lj slv
  const double *q0;                   // ferrf_alchemy: lambda0 charges; en0 goes to the lj slt column
  const double *erf_tab;              // ferrf ERF_CORR: spline table (4, erf_n) over r = rlew + (i - 1) delew
  int erf_n;
  double delew, rlew;
  const int *pair_type;               // fnb14: packed LJ index of each pair (type14, 1-based)
  double lj_fudge;                    // fnb14: LJ-FUDGE (fudge is QQ-FUDGE)
  // kinds 3 (ferrf_alchemy_add) and 4 (fnb14_alchemy) write 8 energy columns per pair:
  //   kind 3: [4 + class] the term on lambdaq, [6] en0 on lambdaq0
  //   kind 4: coul, lj by class at [0..3] on lambda, at [4..7] on lambda0
  int ncol;                           // energy columns per pair and in out (4 when 0)
  const int *pkind;                   // per pair kind (overrides kind when set)
  const int *facq;                    // kind 3: facQ of each pair (1: same alchemical kind, 0: added-removed)
  const double *qa;                   // kind 3: the created charges (q holds the lambda charges of kind 0)
  const double *lq, *lq0;             // per atom lambdaq, lambdaq0
  const double *lam, *lam0;           // per atom lambda, lambda0 (kind 4)
  int intramol;                       // kind 3: INTRAMOL (skip the facQ = 1 pairs)
  // fnb14_rem.f (kind 2): remf[p] != 0 puts the pair's energy in column 4
  // (ncol = 5) and scales its force by remf2
  const int *remf;
  double remf2;
};
void launch_corr(const CorrDesc *d_desc, const CorrDesc *h, int nsys, hipStream_t st);

// Proper torsions (fptors.f): one thread per torsion writes its four atoms'
// forces and its energy; a per-atom gather sums them in a fixed order.
struct TorsDesc {
  int ntors, natoms;
  const int4 *atoms;                  // 0-based l1..l4 (the first arity used)
  const double *pota, *potb;          // torsions: signed multiplicity, barrier; bonds, bends: rest value, constant
  const double *potc, *potd, *pote, *potf, *potg;  // bends: Urey-Bradley and coupling parameters
  int arity;                          // atoms per term: 2 bonds, 3 bends, 4 torsions
  int improper;                       // torsions: fitors rules (potc > 0 harmonic; cosine n = 1-4, 6)
  const double *x, *y, *z;            // coordinates as given
  const int *ss;                      // ss_index of each atom
  double *tf;                         // [ntors][12] force components: x of l1..l4, y, z (unused slots zero)
  double *te;                         // [ntors][2] energy by the class of l1
  int *status;                        // [0]: 1 singular torsion; [1]: 2 unsupported multiplicity (sticky words)
  const int *g_off, *g_slot;          // per atom gather CSR: 4 * torsion + position
  double *fx, *fy, *fz;               // accumulated (+=)
  double *out;                        // [2]: utors slt, slv ([3] with the segment energy under REM)
  // REM twins (fptors_rem.f, fitors_rem.f): remf[i] != 0 puts term i's energy
  // in tseg[i] (summed to out[2]) and scales its force by remf2; rem_variant
  // selects fptors_rem's 60-degree-phase forms (no sign factor, dsoa guard)
  const int *remf;
  double remf2;
  int rem_variant;
  double *tseg;
};
void launch_tors(const TorsDesc *d_desc, const TorsDesc &h, hipStream_t st);

// ---------------------------------------------------------------- integration

// One system's molecular-dynamics state.  A unit is one atom or one rigid
// 3-atom molecule (consecutive atoms, SHAKE/RATTLE on its three distances).
struct MdDesc {
  int natoms, nunits, nbond;
  double *x, *y, *z;                 // Cartesian positions
  double *vx, *vy, *vz;              // velocities
  double *xp, *yp, *zp;              // reduced copy read by the force path
  const double *fx, *fy, *fz;        // slow forces (the nonbonded evaluation)
  double *bfx, *bfy, *bfz;           // fast forces (harmonic distance terms)
  const double *fmx, *fmy, *fmz;     // inner-shell nonbonded forces (two shells; else null)
  const double *mass;                // atom masses
  const int *ufirst;                 // first atom of each unit
  const unsigned char *urigid;       // 1: rigid 3-atom unit
  const double *ud2;                 // [nunits][3] squared lengths (0,1), (0,2), (1,2)
  const int *bi, *bj;                // harmonic terms E = k (r - r0)^2
  const double *bk, *br0;
  double *bpf, *bpe;                 // per-term force vector and energy
  const int *bg_off, *bg_term;       // per-atom gather CSR, signed term index +/-(t+1)
  double oc[9];
  double *out;                       // [5]: kinetic, bond energy, momentum x, y, z
  int *fail;                         // constraint iterations that did not converge
  int *step;                         // outer steps taken (device counter, drives the list schedule)
};
// stages of launch_md_units, applied in this order when combined: half kick
// by the fast forces over h + RATTLE; half kick by the inner-shell forces
// over hm + RATTLE; half kick by the slow forces over dt + RATTLE; drift over
// h (ORAC's verlet with the fast forces) + SHAKE; RATTLE
constexpr int MD_KICK_FAST = 1, MD_KICK_SLOW = 2, MD_VERLET = 4, MD_RATTLE = 8, MD_KICK_MID = 16;
void launch_md_units(const MdDesc *d, const MdDesc *h, int nsys, double dt, double hm, double hin, int flags,
                     hipStream_t st);
void launch_md_fast_forces(const MdDesc *d, const MdDesc *h, int nsys, hipStream_t st);
void launch_md_energy(const MdDesc *d, int nsys, hipStream_t st);
void launch_md_tick(const MdDesc *d, int nsys, hipStream_t st);
void launch_md_relax(const MdDesc *d, const MdDesc *h, int nsys, double hstep, double dmax, hipStream_t st);
void launch_md_scale_v(const MdDesc *d, const MdDesc *h, int nsys, const double *scale, hipStream_t st);

void launch_cself(int n, const double *q, const int *ss, double *d_out2, hipStream_t st);
void launch_zero_u64(unsigned long long *p, size_t n, hipStream_t st);
void launch_zero_f64(double *p, size_t n, hipStream_t st);

}  // namespace orac
