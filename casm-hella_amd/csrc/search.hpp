// search.hpp -- boxcar / hella-kernel search over dedispersed planes:
// per-width statistics, normalisation, threshold and ordered compaction.
#pragma once

#include "internal.hpp"

namespace hamd {

constexpr int SEARCH_BLOCK = 4096;  // canonical reduction block (samples)
constexpr int SEARCH_MAXW = 12;     // maximum number of width trials

enum SearchMode { SEARCH_BOXCAR = 0, SEARCH_HELLA = 1 };

struct Cand {
  float snr;
  int samp;    // output sample index within the gulp
  int width;   // width trial index
  int dm_idx;  // DM trial index
};

struct SScal {
  double mu[SEARCH_MAXW];
  double inv[SEARCH_MAXW];
  float sigf[SEARCH_MAXW];
  double mu0;      // mean of the normalised first width plane
  float sig0;      // its standard deviation
  float thr;       // snr * sig0
  int gated;       // 1 when sig0 is outside (0.1, 5): no candidates
  int pad;
  // BOXCAR: smallest integer box sum whose normalised value exceeds thr
  // (normalisation is monotone in the sum); 2^32 when none does
  uint64_t rthr[SEARCH_MAXW];
};

struct SearchBeamDev {
  const uint32_t *D;  // dedispersed sums [ndm][ds]
  int64_t ds;
  int rows;           // searched rows (boxcar: ndm; hella: ndm - 2)
  int row0;           // first D row searched
  float *planes;      // hella mode: raw smoothed planes [nw][rows][ps]
  int64_t ps;
  double *part;       // [nw][nblk][rows] block sums of x, rows innermost so a
                      // finisher thread per row reads them coalesced (BOXCAR
                      // statistics: exact u64 sums, [nw][groups][rows])
  double *part2;      // [nw][nblk][rows] block sums of x*x (BOXCAR statistics:
                      // exact u128 sums per group, lo/hi u64 pairs; 2x the size)
  SScal *scal;
  uint32_t *counts;   // [nw][rows][nblk]
  uint32_t *offs;     // exclusive scan of counts
  uint32_t *bmax;     // [nw][rows][nblk][4] largest raw value of each wave's
                      // samples of a block (u32 box sum, or the bits of a
                      // non-negative float in hella mode)
  uint32_t *tiles;    // scan scratch [ceil(n/SEARCH_BLOCK)]
  Cand *cands;        // ordered candidates (first cap stored)
  uint32_t *ncand;    // total crossings (uncapped)
  // clustering (cluster.hip)
  uint32_t *cl_grid;  // [cl_rows][cl_cols] lowest candidate index of each cell; all
                      // 0xffffffff between launches
  uint32_t *cl_work;  // [6][cl_cap] union-find parents, count, sample and DM bounds
  unsigned long long *cl_key;  // [cl_cap] best (S/N, -index) of each component root
  struct ClusterRec *clusters;  // [cl_cap] clusters by representative index
  uint32_t *cl_n;     // [0] clusters, [1] candidates taken
  int cl_rows, cl_cols, cl_cap;
};

// one cluster (the layout of hella_cluster; rep is the index in the beam's
// taken candidates, beam is filled in by the host)
struct ClusterRec {
  int rep, count, samp_min, samp_max, dm_min, dm_max, beam;
};

struct ClusterArgs {
  int cand_cap, max_giants;  // a beam keeps min(crossings, cand_cap, max_giants)
  int prefix;                // candidates already in the gulp's list before the first beam
  int cumulative;            // beams share the list (each counts the ones before it)
  int ct, cd;                // cell size: samples x DM trials
};

struct SearchCommon {
  int nout;      // searched samples per row
  int ntime_dd;  // dedispersed samples available per row
  int maxW, minW;
  int nw;
  int mode;
  float snr;
  int cap;       // candidate buffer capacity per beam
  int nblk;      // ceil(nout / SEARCH_BLOCK)
  int wp;        // statistics columns
  int64_t box_max;  // bound on every BOXCAR box sum: WIDTH_MAX x 255 x channels
};

int search_nwidth(int mode, int minW, int maxW);
void hella_kernel_host(int sm, std::vector<float> &k);

struct SearchRunner {
  // Full search for nbeam beams (device descriptors).  max_rows is the
  // largest row count.  kern: hella kernels on the device (mode HELLA),
  // concatenated per width; kofs: their offsets (host).
  static void run(const SearchBeamDev *d_beams, int nbeam, int max_rows, const SearchCommon &C,
                  const float *d_kern, const std::vector<int> &kofs, hipStream_t s);
  // the two halves of run(): width trials + per-width statistics (smooth.cpp),
  // then sigma gate, threshold and ordered compaction (peaks.cu)
  static void smooth(const SearchBeamDev *d_beams, int nbeam, int max_rows,
                     const SearchCommon &C, const float *d_kern, const std::vector<int> &kofs,
                     hipStream_t s);
  static void peaks(const SearchBeamDev *d_beams, int nbeam, int max_rows, const SearchCommon &C,
                    hipStream_t s);
  // clusters of every beam's taken candidates (one workgroup per beam)
  static void cluster(const SearchBeamDev *d_beams, int nbeam, const ClusterArgs &A, hipStream_t s);
  // one beam's results written by the device straight into pinned host
  // memory: its three counts (crossings, clusters, candidates taken) into
  // h_n[0..2], its first min(crossings, cap_c) candidates into h_c and its
  // first min(clusters, cap_k) clusters into h_k
  static void harvest(const SearchBeamDev *d_beam, uint32_t *h_n, Cand *h_c, int cap_c,
                      ClusterRec *h_k, int cap_k, hipStream_t s);
};

}  // namespace hamd
